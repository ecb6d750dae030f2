"""openPSTD's C++ kernel API end to end, without the Python facade: the driver
tests/cpp/kernel_api_run (built by tests/cpp/Makefile from the restated
kernel_api.h, linked to libopenpstd_amd.so only) configures a scene, runs
PSTDKernel(GPU=true, MCPU=false) with a KernelCallback subclass and reports
every WriteFrame / WriteSample / Info call; they are compared with the FP64
oracle's frames and receiver series and with openPSTD's call contract
(kernel/Solver.cpp:62-93: frames of non-PML domains and one sample per receiver
when frame % SaveNth == 0; "Starting simulation", "Finished frame: N",
"Succesfully finished simulation")."""
import os
import subprocess

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import workloads as W
from oracle import binding as ob

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "kernel_api_run")


def write_config(conf, path):
    s = conf.Settings
    f32 = lambda v: repr(float(np.float32(v)))
    with open(path, "w") as f:
        f.write("settings " + " ".join(f32(v) for v in (s.calctime, s.c1, s.ampMax, s.freqMax, s.rho, s.patcherror,
                                                        s.tfactRK, s.gridSpacing, s.band_width, s.wave_length)) +
                f" {int(s.PMLCells)} {int(s.SaveNth)}\n")
        f.write("rk " + " ".join(f32(a) for a in s.rk_coefficients) + "\n")
        for d in conf.Domains:
            f.write("domain " + " ".join(f32(v) for v in (*d.TopLeft, *d.Size, d.T.Absorption, d.L.Absorption,
                                                           d.B.Absorption, d.R.Absorption)) + "\n")
        for p in conf.Speakers:
            f.write("speaker " + " ".join(f32(v) for v in p) + "\n")
        for p in conf.Receivers:
            f.write("receiver " + " ".join(f32(v) for v in p) + "\n")


def parse(path):
    frames, samples, logs, meta, domains, other = [], [], [], None, [], {}
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "frame":
            frames.append((int(t[1]), int(t[2]), np.array([float(v) for v in t[4:]])))
            assert len(frames[-1][2]) == int(t[3])
        elif t[0] == "sample":
            samples.append((int(t[1]), int(t[2]), float(t[3])))
        elif t[0] == "log":
            logs.append((t[1], line.split(" ", 2)[2].strip()))
        elif t[0] == "meta":
            meta = (int(t[1]), int(t[2]))
        elif t[0] == "domain":
            domains.append(tuple(int(v) for v in t[1:4]))
        else:
            other[t[0]] = t[1:]
    return frames, samples, logs, meta, domains, other


def test_cpp_kernel_api_room(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (tests/cpp/Makefile, run by __graft_entry__.build())")
    steps, nth = 120, 10
    conf = W.room(steps)
    conf.Settings.SetSaveNth(nth)
    cfg, out = str(tmp_path / "conf.txt"), str(tmp_path / "out.txt")
    write_config(conf, cfg)
    r = subprocess.run([DRIVER, cfg, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    frames, samples, logs, meta, domains, other = parse(out)
    assert other["not_configured_throws"] == ["1"]
    # the oracle of the same configuration (FP64), with its WriteFrame calls
    o = ob.OracleScene(conf, 64, 8)
    ofr = []
    series = o.step(steps, frames=ofr)
    # metadata: every domain incl. PML (kernel/PSTDKernel.cpp:154-168)
    assert meta == (o.n_domains, steps)
    assert [d[:2] for d in domains] == [(o.domain_info(i)["nx"], o.domain_info(i)["ny"]) for i in range(o.n_domains)]
    # frames: same calls in the same order, values within the FFT form's bar
    assert [(f, d) for f, d, _ in frames] == [(f, d) for f, d, _ in ofr]
    assert [f for f, _, _ in frames] == [f for f in range(steps) if f % nth == 0]
    for (f, d, got), (_, _, want) in zip(frames, ofr):
        assert np.abs(got - want).max() <= 2e-6, (f, d)
    # samples: one per receiver per saved frame, startSample = frame, the last value
    assert [(s, r) for s, r, _ in samples] == [(f, 0) for f in range(steps) if f % nth == 0]
    got = np.array([v for _, _, v in samples])
    want = series[[f for f in range(steps) if f % nth == 0], 0]
    assert np.abs(got - want).max() <= 2e-6
    infos = [m for lvl, m in logs if lvl == "info"]
    assert infos[0] == "Starting simulation" and infos[-1] == "Succesfully finished simulation"
    assert [m for m in infos if m.startswith("Finished frame")] == [f"Finished frame: {f}" for f in range(steps)]
