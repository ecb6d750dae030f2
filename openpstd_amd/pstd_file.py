"""openPSTD's .pstd scene/results file over the library's C ABI (pstd_file_*).

PSTDFile keeps the method names of shared/PSTDFile.h (Open, New,
GetSceneConf, SetSceneConf, GetResultsSceneConf, GetResultsFrameCount,
GetResultsFrame, GetReceiverData, Commit); ``run`` is the CLI's ``run``
command (CLI/main.cpp:365-464) on the file.
"""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import lib
from .config import PSTDConfiguration
from .kernel import KernelCallbackLog, _callbacks, _check


class PSTDFile:
    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def Open(path):
        """shared/PSTDFile.cpp:165-178 (version 4 required)."""
        h = C.c_void_p()
        rc = lib().pstd_file_open(str(path).encode(), C.byref(h))
        f = PSTDFile(h)
        _check(rc, f._err)
        return f

    @staticmethod
    def New(path):
        """shared/PSTDFile.cpp:180-199: the default scene and an empty results scene, written at once."""
        h = C.c_void_p()
        rc = lib().pstd_file_new(str(path).encode(), C.byref(h))
        f = PSTDFile(h)
        _check(rc, f._err)
        return f

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().pstd_file_close(h)
            self._h = None

    def _err(self):
        return lib().pstd_file_last_error(self._h).decode()

    def _scene(self, results):
        c = T.pstd_config()
        _check(lib().pstd_file_get_scene(self._h, int(results), C.byref(c)), self._err)
        return PSTDConfiguration.from_c(c)

    def GetSceneConf(self):
        return self._scene(False)

    def GetResultsSceneConf(self):
        return self._scene(True)

    def SetSceneConf(self, conf):
        c, keep = conf.to_c()
        _check(lib().pstd_file_set_scene(self._h, C.byref(c)), self._err)

    def GetResultsDomainCount(self):
        return len(self.GetResultsSceneConf().Domains)

    def GetResultsReceiverCount(self):
        return len(self.GetResultsSceneConf().Receivers)

    def GetResultsFrameCount(self, domain):
        n = lib().pstd_file_frame_count(self._h, int(domain))
        if n < 0:
            _check(n, self._err)
        return n

    def GetResultsFrame(self, frame, domain):
        n = lib().pstd_file_get_frame(self._h, int(frame), int(domain), None, 0)
        if n < 0:
            _check(int(n), self._err)
        out = np.zeros(n, np.float32)
        lib().pstd_file_get_frame(self._h, int(frame), int(domain),
                                  out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
        return out

    def GetReceiverData(self, receiver):
        n = lib().pstd_file_get_receiver(self._h, int(receiver), None, 0)
        if n < 0:
            _check(int(n), self._err)
        out = np.zeros(n, np.float32)
        lib().pstd_file_get_receiver(self._h, int(receiver), out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
        return out

    # raw records (keys: bytes of native 32-bit ints, values: bytes)
    @staticmethod
    def CreateKey(prefix, args=()):
        """PSTDFile::CreateKey (shared/PSTDFile.cpp:344-366)."""
        return np.array([prefix, *args], dtype=np.uint32).tobytes()

    def GetRawValue(self, key):
        n = lib().pstd_file_get_raw(self._h, key, len(key), None, 0)
        if n < 0:
            _check(int(n), self._err)
        buf = C.create_string_buffer(max(1, n))
        lib().pstd_file_get_raw(self._h, key, len(key), buf, n)
        return buf.raw[:n]

    def SetRawValue(self, key, value, append=False):
        _check(lib().pstd_file_set_raw(self._h, key, len(key), value, len(value), int(append)), self._err)

    def DeleteValue(self, key):
        _check(lib().pstd_file_delete_raw(self._h, key, len(key)), self._err)

    def keys(self):
        out = []
        for i in range(lib().pstd_file_record_count(self._h)):
            n = lib().pstd_file_record_key(self._h, i, None, 0)
            buf = C.create_string_buffer(max(1, n))
            lib().pstd_file_record_key(self._h, i, buf, n)
            out.append(buf.raw[:n])
        return out

    def Commit(self):
        _check(lib().pstd_file_commit(self._h), self._err)

    def run(self, gpu=True, mcpu=False, receiver_mode=T.PSTD_RECEIVER_PHYSICAL, log=None):
        """RunCommand::execute: delete old results, initialise new ones, run
        PSTDKernel(gpu, mcpu) on the file's scene writing frames and receiver
        samples into the file, commit."""
        cb, keep = _callbacks(log if log is not None else KernelCallbackLog())
        _check(lib().pstd_file_run(self._h, int(bool(gpu)), int(bool(mcpu)), int(receiver_mode), C.byref(cb)),
               self._err)
