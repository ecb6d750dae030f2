"""HIP solver (libopenpstd_amd.so on the GPU) against the CPU oracle.

Tolerances (FP32 device arithmetic):
  * fields after a few steps: rel-L2 <= 2e-6 per domain field vs the FP64 oracle;
  * receiver series: rel-L2 <= 1e-4 vs the FP64 oracle (the metric's accuracy bar),
    and no worse than 2x the FP32 oracle's own error vs FP64 plus 1e-5.
"""
import math
import os

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import abi_types as T
from openpstd_amd import workloads as W
from oracle import binding as ob

from test_oracle_kat import rigid_box

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def device(conf):
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(conf)
    return k, k.prepare()


def compare_fields(k, dev, o, tol, names=("p0", "px0", "py0", "vx0", "vy0")):
    worst = 0.0
    for i in range(k.domain_count()):
        for f in names:
            a, b = dev.field(i, f), o.field(i, f)
            scale = max(np.linalg.norm(b), 1e-6 * math.sqrt(b.size))
            e = np.linalg.norm(a - b) / scale
            worst = max(worst, e)
            assert e <= tol, (i, f, e)
    return worst


@pytest.mark.parametrize("steps", [1, 3])
def test_c1_fields_after_steps(steps):
    conf = W.c1(steps)
    k, dev = device(conf)
    dev.advance(steps)
    o = ob.OracleScene(conf, 64)
    o.step(steps)
    compare_fields(k, dev, o, 2e-6)


def test_c1_receiver_series():
    steps = 300
    conf = W.c1(steps)
    k, dev = device(conf)
    got = dev.advance(steps)
    o64 = ob.OracleScene(conf, 64, THREADS).step(steps)
    o32 = ob.OracleScene(conf, 32, THREADS).step(steps)
    e_dev, e_cpu32 = rel(got, o64), rel(o32, o64)
    print(f"C1 {steps} steps: device vs FP64 {e_dev:.2e}, CPU FP32 vs FP64 {e_cpu32:.2e}")
    assert e_dev <= 1e-4
    assert e_dev <= 2 * e_cpu32 + 1e-5


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_c2_short(seed):
    steps = 60
    conf = W.c2(seed, steps)
    k, dev = device(conf)
    got = dev.advance(steps)
    o = ob.OracleScene(conf, 64, THREADS)
    want = o.step(steps)
    compare_fields(k, dev, o, 2e-5)
    assert rel(got, want) <= 1e-4


def test_c2_full_golden():
    """The bench workload at full length against the committed FP64 oracle series."""
    g = np.load(os.path.join(HERE, "golden", "c2_receivers.npz"))
    conf = W.c2(0, 2000)
    k, dev = device(conf)
    got = dev.advance(2000)
    e_dev, e_cpu32 = rel(got, g["fp64"]), rel(g["fp32"], g["fp64"])
    print(f"C2 2000 steps: device vs FP64 {e_dev:.2e}, CPU FP32 vs FP64 {e_cpu32:.2e}")
    assert e_dev <= 1e-4
    assert np.isfinite(got).all()


def test_c1_full_golden():
    g = np.load(os.path.join(HERE, "golden", "c1_receivers.npz"))
    conf = W.c1(500)
    k, dev = device(conf)
    got = dev.advance(500)
    assert rel(got, g["fp64"]) <= 1e-4


@pytest.mark.parametrize("idx", range(4))
def test_c4_scenes(idx):
    conf = W.c4(7, 40, n_scenes=4, hi=300)[idx]
    k, dev = device(conf)
    got = dev.advance(40)
    o = ob.OracleScene(conf, 64, THREADS)
    want = o.step(40)
    compare_fields(k, dev, o, 2e-5)
    assert rel(got, want) <= 1e-4


def test_default_conf_partial_pmls():
    """openPSTD's own default scene: partial-edge and overlapping PML domains."""
    conf = P.PSTDConfiguration.CreateDefaultConf()
    s = conf.Settings
    s.SetRenderTime(float(np.float32(20.5 * s.GetTimeStep())))
    k, dev = device(conf)
    dev.advance(20)
    o = ob.OracleScene(conf, 64)
    o.step(20)
    compare_fields(k, dev, o, 2e-5)


@pytest.mark.parametrize("mode", [(4, 0), (3, 2)])
def test_rigid_box_standing_wave(mode):
    """Closed form (see tests/test_oracle_kat.py): a rigid-wall cavity mode."""
    conf, p, expect = rigid_box(steps=60, mode=mode)
    k, dev = device(conf)
    dev.set_field(0, "p0", p)
    dev.set_field(0, "px0", p)
    got = dev.advance(expect.shape[0])
    assert np.abs(got - expect).max() < 5e-6


@pytest.mark.parametrize("n", [2000, 4100])
def test_long_lines(n):
    """F = 4096 and F = 8192 lines (the long-line kernels)."""
    conf = P.PSTDConfiguration()
    W.base_settings(conf, 4)
    conf.Domains = [W.domain(0, 0, n, 60, 1.0)]
    conf.Speakers = [(W._world_pos(n / 2), W._world_pos(30.0), 0.0)]
    conf.Receivers = [(W._world_pos(n / 2 + 7.25), W._world_pos(33.25), 0.0)]
    k, dev = device(conf)
    dev.advance(4)
    o = ob.OracleScene(conf, 64, THREADS)
    o.step(4)
    compare_fields(k, dev, o, 5e-6)


def test_run_callbacks_contract():
    """PSTDKernel.run: the callback sequence of SingleThreadSolver (kernel/Solver.cpp:62-93)."""
    steps = 12
    conf = W.c2(1, steps)
    conf.Settings.SetSaveNth(5)
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(conf)
    cb = P.RecordingCallback()
    k.run(cb)
    assert [f for f, _, _ in cb.frames] == [0, 0, 0, 5, 5, 5, 10, 10, 10]
    assert [d for _, d, _ in cb.frames] == [0, 1, 2] * 3
    assert sorted(cb.samples) == [0, 1, 2, 3]
    for r in cb.samples:
        assert [s for s, _ in cb.samples[r]] == [0, 5, 10]
    infos = [m for lvl, m in cb.messages if lvl == "Info"]
    assert infos[0] == "Starting simulation" and infos[-1] == "Succesfully finished simulation"
    assert infos[1:-1] == [f"Finished frame: {i}" for i in range(steps)]
    o = ob.OracleScene(conf, 64)
    frames = []
    series = o.step(steps, frames)
    for (f1, d1, a), (f2, d2, b) in zip(cb.frames, frames):
        assert (f1, d1) == (f2, d2)
        assert rel(a, b) < 1e-5
    for r in cb.samples:
        for s, v in cb.samples[r]:
            assert v == pytest.approx(series[s, r], rel=1e-4, abs=1e-9)
    # device state stays readable after run (last frame)
    assert rel(k.field(0, "p0").ravel(), o.field(0, "p0").ravel()) < 1e-5


def test_reference_receiver_mode():
    conf = W.c2(0, 30)
    conf.receiver_mode = T.PSTD_RECEIVER_REFERENCE
    conf.Receivers = [(W._world_pos(40.25), W._world_pos(30.25), 0.0), (W._world_pos(100.25), W._world_pos(20.25), 0.0)]
    k, dev = device(conf)
    got = dev.advance(30)
    want = ob.OracleScene(conf, 64).step(30)
    assert rel(got, want) <= 1e-4


def test_set_field_midrun_p0():
    """Host writes between steps keep openPSTD semantics (p0 is its own array)."""
    conf = W.c1(10)
    k, dev = device(conf)
    o = ob.OracleScene(conf, 64)
    dev.advance(3)
    o.step(3)
    p = dev.field(1, "p0") * 0.5
    dev.set_field(1, "p0", p)
    o.set_field(1, "p0", p)
    dev.advance(4)
    o.step(4)
    compare_fields(k, dev, o, 5e-6)
