"""The CLI's `run` on a .pstd file (pstd_file_run, RunCommand of
CLI/main.cpp:365-464) on the GPU: the frames and receiver series it writes
into the file are bitwise those the kernel API hands to a KernelCallback for
the same scene, and openPSTD's own unqlite (oracle/_ref, where built) reads
them back the same."""
import os

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import workloads as W

from test_pstd_file import REF, OPEN_READWRITE, RefDB

pytestmark = pytest.mark.gpu


def test_cli_run_writes_the_kernel_results(tmp_path):
    steps = 120
    conf = W.lshape(0, steps)
    conf.Settings.SetSaveNth(40)
    path = str(tmp_path / "lshape.pstd")
    f = P.PSTDFile.New(path)
    f.SetSceneConf(conf)
    f.Commit()
    del f
    from openpstd_amd.__main__ import main
    print("run -g", path, flush=True)
    assert main(["run", "-g", path]) == 0
    g = P.PSTDFile.Open(path)
    scene = g.GetSceneConf()
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(scene)
    rec = P.RecordingCallback()
    k.run(rec)
    frames = {}
    for _, d, data in rec.frames:
        frames.setdefault(d, []).append(np.asarray(data, np.float32))
    assert g.GetResultsDomainCount() == 3 and sorted(frames) == [0, 1, 2]
    for d, fl in frames.items():
        assert g.GetResultsFrameCount(d) == len(fl) == steps // 40  # frames 0, 40, 80
        for i, want in enumerate(fl):
            assert np.array_equal(g.GetResultsFrame(i, d), want)
    # one WriteSample per receiver per saved frame (the receiver's last value,
    # kernel/Solver.cpp:86-91), appended to the receiver's record
    assert sorted(rec.samples) == [0, 1, 2, 3]
    for r, series in rec.samples.items():
        want = np.array([v for _, v in series], np.float32)
        assert want.size == steps // 40
        assert np.array_equal(g.GetReceiverData(r), want)
    if os.path.exists(REF):
        db = RefDB(path, OPEN_READWRITE)
        for d, fl in frames.items():
            assert np.frombuffer(db.fetch(P.PSTDFile.CreateKey(102, (d,))), np.int32)[0] == len(fl)
            for i, want in enumerate(fl):
                assert db.fetch(P.PSTDFile.CreateKey(103, (d, i))) == want.tobytes()
        db.close()


def test_cli_run_spectral_receivers(tmp_path):
    """`run -g --receivers spectral`: the receiver records hold the spectral
    receivers' series, bitwise those of the kernel API in that mode."""
    from openpstd_amd import abi_types as T
    steps = 60
    conf = W.lshape_spectral(steps)
    conf.Settings.SetSaveNth(1)
    path = str(tmp_path / "spectral.pstd")
    f = P.PSTDFile.New(path)
    f.SetSceneConf(conf)
    f.Commit()
    del f
    from openpstd_amd.__main__ import main
    assert main(["run", "-g", path, "--receivers", "spectral"]) == 0
    g = P.PSTDFile.Open(path)
    scene = g.GetSceneConf()
    scene.receiver_mode = T.PSTD_RECEIVER_SPECTRAL
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(scene)
    rec = P.RecordingCallback()
    k.run(rec)
    assert sorted(rec.samples) == list(range(len(conf.Receivers)))
    for r, series in rec.samples.items():
        want = np.array([v for _, v in series], np.float32)
        assert want.size == steps
        assert np.array_equal(g.GetReceiverData(r), want)
