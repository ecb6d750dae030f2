"""openPSTD's kernel API (kernel/KernelInterface.h:243-356, kernel/PSTDKernel.h:52-150)
over the native library.

    kernel = PSTDKernel(GPU=True, MCPU=False)
    kernel.initialize_kernel(conf, KernelCallbackLog())
    kernel.run(callback)            # WriteFrame / WriteSample / Info ... on this thread
    meta = kernel.get_metadata()

run() drives the MI355X solver; there is no CPU path in this package.
"""
import ctypes as C

import numpy as np

from . import abi_types as T
from ._lib import PSTDError, lib


class PSTDKernelNotConfiguredException(RuntimeError):
    """kernel/KernelInterface.h:353-356."""

    def __init__(self):
        super().__init__("Kernel is not yet configured")


class KernelCallbackLog:
    """kernel/KernelInterface.h:243-271: non-blocking log sink (no-op defaults)."""

    def Fatal(self, message): pass
    def Error(self, message): pass
    def Warning(self, message): pass
    def Info(self, message): pass
    def Debug(self, message): pass


class KernelCallback(KernelCallbackLog):
    """kernel/KernelInterface.h:279-304."""

    def WriteFrame(self, frame, domain, data):
        raise NotImplementedError

    def WriteSample(self, startSample, receiver, data):
        raise NotImplementedError


class SimulationMetadata:
    """kernel/KernelInterface.h:309-327."""

    def __init__(self):
        self.DomainMetadata = []
        self.Framecount = 0
        self.DomainPositions = []


class KernelInterface:
    """kernel/KernelInterface.h:332-351."""

    def initialize_kernel(self, config, callbackLog): raise NotImplementedError
    def run(self, callback): raise NotImplementedError
    def get_metadata(self): raise NotImplementedError


def _check(status, handle_error):
    if status == T.PSTD_ERR_NOT_CONFIGURED:
        raise PSTDKernelNotConfiguredException()
    if status != T.PSTD_OK:
        raise PSTDError(status, handle_error())


def _callbacks(cb):
    """Builds a pstd_callbacks struct (and keeps the ctypes thunks alive)."""
    levels = ("Fatal", "Error", "Warning", "Info", "Debug")

    def log(user, level, msg):
        if cb is not None:
            getattr(cb, levels[level] if 0 <= level < 5 else "Debug")(msg.decode())

    def frame(user, f, d, data, n):
        cb.WriteFrame(f, d, np.ctypeslib.as_array(data, shape=(n,)).copy())

    def sample(user, s, r, data, n):
        cb.WriteSample(s, r, [float(data[i]) for i in range(n)])

    c = T.pstd_callbacks()
    keep = [T.LOG_CB(log)]
    c.log = keep[0]
    if isinstance(cb, KernelCallback):
        keep += [T.FRAME_CB(frame), T.SAMPLE_CB(sample)]
        c.write_frame, c.write_sample = keep[1], keep[2]
    return c, keep


class PSTDKernel(KernelInterface):
    """kernel/PSTDKernel.cpp:44-168 with the GPU solver slot filled."""

    def __init__(self, GPU=True, MCPU=False):
        self._h = C.c_void_p()
        lib().pstd_kernel_create(int(bool(GPU)), int(bool(MCPU)), C.byref(self._h))
        self._configured = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            lib().pstd_kernel_destroy(h)
            self._h = None

    def _err(self):
        return lib().pstd_kernel_last_error(self._h).decode()

    def initialize_kernel(self, config, callbackLog=None):
        cconf, keep = config.to_c()
        cbs, keep2 = _callbacks(callbackLog)
        _check(lib().pstd_kernel_initialize(self._h, C.byref(cconf), C.byref(cbs)), self._err)
        self._configured = True
        self._config = config

    def run(self, callback):
        cbs, keep = _callbacks(callback)
        _check(lib().pstd_kernel_run(self._h, C.byref(cbs)), self._err)

    def get_metadata(self):
        n = C.c_int32()
        fc = C.c_int32()
        _check(lib().pstd_kernel_get_metadata(self._h, C.byref(n), None, 0, C.byref(fc)), self._err)
        sizes = (C.c_int32 * (3 * max(1, n.value)))()
        _check(lib().pstd_kernel_get_metadata(self._h, C.byref(n), sizes, n.value, C.byref(fc)), self._err)
        m = SimulationMetadata()
        m.DomainMetadata = [[sizes[3 * i], sizes[3 * i + 1], sizes[3 * i + 2]] for i in range(n.value)]
        m.Framecount = fc.value
        return m

    def set_form(self, form):
        """Derivative form of the next run/prepare: "fft" (default), "gemm" / "gemm_f16" (dense
        derivative matrices on MFMA, FP16 hi+lo split, FP32 accumulate), "gemm_bf16", or "auto"
        (the faster of "fft" and "gemm_f16" for the plan, timed at upload)."""
        _check(lib().pstd_kernel_set_form(self._h, T.FORMS[form]), self._err)

    def prepare(self):
        """Device context with the scene uploaded, frame 0 next (see DeviceSolver)."""
        return DeviceSolver(self)

    # ---- scene inspection (the test-only PSTDKernel::get_scene analogue)
    def domain_count(self):
        return len(self.get_metadata().DomainMetadata)

    def domain_info(self, i):
        info = (C.c_int32 * 16)()
        f = (C.c_float * 3)()
        _check(lib().pstd_kernel_domain_info(self._h, i, info, f), self._err)
        keys = ["id", "tlx", "tly", "nx", "ny", "is_pml", "is_secondary_pml", "is_rigid", "local", "n_left",
                "n_right", "n_top", "n_bottom", "update_x", "update_y", "pml_for_count"]
        d = {k: int(info[j]) for j, k in enumerate(keys)}
        d.update(alpha=f[0], impedance=f[1], rho=f[2])
        return d

    def _shape(self, i, name):
        d = self.domain_info(i)
        nx, ny = d["nx"], d["ny"]
        return {"vx0": (ny, nx + 1), "vy0": (ny + 1, nx)}.get(name, (ny, nx))

    def field(self, i, name):
        shape = self._shape(i, name)
        out = np.zeros(shape, dtype=np.float32)
        _check(lib().pstd_kernel_read_field(self._h, i, T.FIELD_IDS[name], out.ctypes.data_as(C.POINTER(C.c_float)),
                                            out.size), self._err)
        return out

    def pml(self, i, which):
        shape = self._shape(i, ["px0", "py0", "vx0", "vy0"][which])
        out = np.zeros(shape, dtype=np.float32)
        _check(lib().pstd_kernel_read_pml(self._h, i, which, out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
               self._err)
        return out

    def segments(self):
        n = lib().pstd_kernel_segment_count(self._h)
        out = []
        for i in range(max(0, n)):
            s = T.pstd_plan_segment()
            lib().pstd_kernel_get_segment(self._h, i, C.byref(s))
            out.append({name: getattr(s, name) for name, _ in s._fields_})
        return out


class _Context:
    """Shared driver of a device context (pstd_ctx_*)."""

    n_receivers = 0

    def _err(self):
        return lib().pstd_ctx_last_error(self.ctx).decode()

    def advance(self, steps):
        """Runs `steps` time steps; returns the receiver samples (steps x receivers)."""
        out = np.zeros((steps, max(1, self.n_receivers)), dtype=np.float32)
        ptr = out.ctypes.data_as(C.POINTER(C.c_float)) if self.n_receivers else None
        _check(lib().pstd_ctx_advance(self.ctx, steps, ptr), self._err)
        return out[:, : self.n_receivers]

    def advance_async(self, steps):
        _check(lib().pstd_ctx_advance_async(self.ctx, steps), self._err)

    def advance_static(self, steps):
        """Like advance_async without any device allocation (pstd_ctx_advance_static)."""
        _check(lib().pstd_ctx_advance_static(self.ctx, steps), self._err)

    def synchronize(self):
        _check(lib().pstd_ctx_synchronize(self.ctx), self._err)

    @property
    def steps_done(self):
        return lib().pstd_ctx_steps_done(self.ctx)

    @property
    def cells_per_stage(self):
        """Cells of the non-rigid domains (PML included): the metric's cell count."""
        return int(lib().pstd_ctx_cells_per_stage(self.ctx))

    def profile_stages(self, steps):
        """Advances `steps` steps between two HIP events on the solver stream and
        waits; returns (GPU time per RK-stage launch in ms, launches)."""
        ms = C.c_double()
        n = C.c_int()
        _check(lib().pstd_ctx_profile_stages(self.ctx, steps, C.byref(ms), C.byref(n)), self._err)
        return ms.value, n.value

    KERNEL_CLASSES = ("fft_small", "fft_4096", "fft_8192", "gemm")

    def profile_kernels(self, steps):
        """Advances `steps` steps with HIP event pairs around every launch of each
        kernel class and waits; returns {class: (GPU ms summed, launches)} and
        "total": (GPU ms of the whole run, RK stages)."""
        ms = (C.c_double * 5)()
        n = (C.c_int * 5)()
        _check(lib().pstd_ctx_profile_kernels(self.ctx, steps, ms, n), self._err)
        out = {c: (ms[i], n[i]) for i, c in enumerate(self.KERNEL_CLASSES)}
        out["total"] = (ms[4], n[4])
        return out

    def gemm_work(self):
        """(MFMA flops, line outputs) of the GEMM-form segments per RK stage."""
        f, n = C.c_double(), C.c_longlong()
        _check(lib().pstd_ctx_gemm_work(self.ctx, C.byref(f), C.byref(n)), self._err)
        return f.value, n.value

    def set_graph_steps(self, steps):
        """Replay blocks of `steps` time steps as one captured hipGraph (default 20; 0: eager launches)."""
        _check(lib().pstd_ctx_set_graph_steps(self.ctx, steps), self._err)

    def set_stage_overlap(self, on):
        """GEMM-form stages: FFT-form tiles on a second stream beside the GEMM kernel, or in order (default)."""
        _check(lib().pstd_ctx_set_stage_overlap(self.ctx, int(bool(on))), self._err)

    def capture_graph(self):
        """Capture the step graph now (after a first time step), outside a timed loop."""
        _check(lib().pstd_ctx_capture_graph(self.ctx), self._err)

    def set_tile_order(self, xcd_contiguous):
        _check(lib().pstd_ctx_set_tile_order(self.ctx, int(bool(xcd_contiguous))), self._err)

    FLAVOURS = {"early": 0, "late": 1, "mid": 2}

    def set_kernel_flavour(self, flavour):
        """Stage-kernel flavour: "early", "late" or "mid" (or 0, 1, 2)."""
        f = self.FLAVOURS.get(flavour, flavour)
        _check(lib().pstd_ctx_set_kernel_flavour(self.ctx, int(f)), self._err)

    def set_tile_threads(self, threads):
        """Block size (128, 256 or 512; 128 needs F <= 1024) of the stage kernel for lines of FFT length <= 2048."""
        _check(lib().pstd_ctx_set_tile_threads(self.ctx, int(threads)), self._err)

    def form(self):
        """The form the uploaded plan runs in ("fft", "gemm_f16", "gemm_bf16"); with set_form("auto")
        the one the upload found faster."""
        f = lib().pstd_ctx_get_form(self.ctx)
        _check(min(f, 0), self._err)
        return {T.PSTD_FORM_FFT: "fft", T.PSTD_FORM_GEMM_F16: "gemm_f16", T.PSTD_FORM_GEMM_BF16: "gemm_bf16"}[f]

    def tuning(self):
        """The autotuner's choice: {"lean" (flavour 0 early, 1 late, 2 mid), "xcd_order", "tile_threads",
        "stage_ms", "trial_steps" (time steps its trials ran on this solver), "cached" (the choice came from
        the per-(device, plan signature) cache, without trials)}."""
        lean, xcd = C.c_int32(), C.c_int32()
        ms = C.c_double()
        _check(lib().pstd_ctx_get_tuning(self.ctx, C.byref(lean), C.byref(xcd), C.byref(ms)), self._err)
        nt = lib().pstd_ctx_get_tile_threads(self.ctx)
        _check(min(nt, 0), self._err)
        trials, cached = C.c_int32(), C.c_int32()
        _check(lib().pstd_ctx_get_autotune(self.ctx, C.byref(trials), C.byref(cached)), self._err)
        return {"lean": lean.value, "xcd_order": xcd.value, "tile_threads": nt, "stage_ms": ms.value,
                "trial_steps": trials.value, "cached": bool(cached.value)}

    def _get(self, i, name, shape):
        out = np.zeros(shape, dtype=np.float32)
        _check(lib().pstd_ctx_get_field(self.ctx, i, T.FIELD_IDS[name], out.ctypes.data_as(C.POINTER(C.c_float)),
                                        out.size), self._err)
        return out

    def _set(self, i, name, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        _check(lib().pstd_ctx_set_field(self.ctx, i, T.FIELD_IDS[name], v.ctypes.data_as(C.POINTER(C.c_float)),
                                        v.size), self._err)


class DeviceSolver(_Context):
    """The device context of a configured PSTDKernel, driven step by step.

    ``PSTDKernel.prepare()`` uploads the plan and the initial fields (frame 0
    next); ``advance(n)`` runs n time steps and returns the receiver samples
    (n x receivers), one per step as openPSTD's Receiver::received_values.
    """

    def __init__(self, kernel):
        self.kernel = kernel  # owns the context
        self.ctx = C.c_void_p()
        _check(lib().pstd_kernel_prepare(kernel._h, C.byref(self.ctx)), kernel._err)
        self.n_receivers = len(kernel._config.Receivers)

    def field(self, i, name):
        return self._get(i, name, self.kernel._shape(i, name))

    def set_field(self, i, name, values):
        self._set(i, name, values)


class BatchSolver(_Context):
    """Independent scenes advanced together: one launch per RK stage covers all
    of them (pstd_batch_prepare).  Every scene evolves exactly as it would alone.

        b = BatchSolver([conf0, conf1, ...])
        samples = b.advance(n)          # n x (receivers of scene 0, scene 1, ...)
        b.field(scene, domain, "p0")
    """

    def __init__(self, configs, device=0, form="fft"):
        self.kernels = []
        for conf in configs:
            k = PSTDKernel(GPU=True)
            k.initialize_kernel(conf)
            self.kernels.append(k)
        self.kernels[0].set_form(form)
        self.domain_base, self.receiver_base = [], []
        nd = nr = 0
        for k, conf in zip(self.kernels, configs):
            self.domain_base.append(nd)
            self.receiver_base.append(nr)
            nd += k.domain_count()
            nr += len(conf.Receivers)
        self.n_receivers = nr
        handles = (C.c_void_p * len(self.kernels))(*[k._h.value for k in self.kernels])
        self.ctx = C.c_void_p()
        _check(lib().pstd_batch_prepare(handles, len(self.kernels), device, C.byref(self.ctx)), self.kernels[0]._err)

    def __del__(self):
        ctx = getattr(self, "ctx", None)
        if ctx and ctx.value:
            lib().pstd_ctx_destroy(ctx)
            self.ctx = None

    def scene_samples(self, samples, scene):
        """Columns of `samples` that belong to `scene`."""
        b = self.receiver_base[scene]
        n = len(self.kernels[scene]._config.Receivers)
        return samples[:, b:b + n]

    def field(self, scene, domain, name):
        return self._get(self.domain_base[scene] + domain, name, self.kernels[scene]._shape(domain, name))

    def set_field(self, scene, domain, name, values):
        self._set(self.domain_base[scene] + domain, name, values)


class RecordingCallback(KernelCallback):
    """Collects frames and samples in memory (the CLI's PSTDFile output in
    CLI/output.cpp:30-67 writes the same calls to disk)."""

    def __init__(self):
        self.frames = []   # (frame, domain, ndarray)
        self.samples = {}  # receiver -> list of (startSample, value)
        self.messages = []

    def WriteFrame(self, frame, domain, data):
        self.frames.append((frame, domain, data))

    def WriteSample(self, startSample, receiver, data):
        self.samples.setdefault(receiver, []).append((startSample, data[0]))

    def Info(self, message):
        self.messages.append(("Info", message))

    def Error(self, message):
        self.messages.append(("Error", message))

    def Debug(self, message):
        self.messages.append(("Debug", message))
