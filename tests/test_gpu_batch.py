"""Batched scenes, hipGraph replay and tile order on the GPU.

A batch advances every scene with the same per-line arithmetic as a single-
scene context, so results must be bit-identical to running the scenes one by
one; graph replay and the XCD tile order change only how launches are issued
and which workgroup takes which tile, so they must be bit-identical too.  One
scene of the batch is also checked against the FP64 oracle (same tolerances as
tests/test_gpu_parity.py).
"""
import os

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import workloads as W
from oracle import binding as ob

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def single(conf):
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(conf)
    return k, k.prepare()


def test_batch_equals_scenes_run_alone():
    steps = 30
    confs = W.scene_batch(7, steps, n_scenes=5, hi=300)
    print("batch of", len(confs), "scenes,", steps, "steps")
    b = P.BatchSolver(confs)
    got = b.advance(steps)
    assert got.shape == (steps, b.n_receivers)
    for i, conf in enumerate(confs):
        k, dev = single(conf)
        want = dev.advance(steps)
        assert np.array_equal(b.scene_samples(got, i), want), i
        for d in range(k.domain_count()):
            for f in ("p0", "px0", "py0", "vx0", "vy0"):
                assert np.array_equal(b.field(i, d, f), dev.field(d, f)), (i, d, f)


def test_batch_scene_against_oracle():
    steps = 40
    confs = W.scene_batch(3, steps, n_scenes=3, hi=260)
    b = P.BatchSolver(confs)
    got = b.advance(steps)
    for i in (0, 2):
        want = ob.OracleScene(confs[i], 64, THREADS).step(steps)
        g = b.scene_samples(got, i)
        assert np.abs(g - want).max() <= 2e-6
        if np.abs(want).max() > 1e-2:
            assert rel(g, want) <= 1e-4


def test_batch_rejects_mixed_time_steps():
    a = W.room(10)
    c = W.room(10)
    c.Settings.SetFactRK(0.4)
    with pytest.raises(P.PSTDError):
        P.BatchSolver([a, c])


@pytest.mark.parametrize("graph_steps", [1, 7, 20])
def test_graph_replay_is_bitwise_eager(graph_steps):
    """Replays with eager work, host reads and synchronisations in between
    (20 is the default)."""
    conf = W.lshape(1, 80)
    _, eager = single(conf)
    eager.set_graph_steps(0)
    k, g = single(conf)
    g.set_graph_steps(graph_steps)
    with pytest.raises(P.PSTDError):
        g.capture_graph()  # p0 still explicit: no graph before the first step
    for n in (1, 20, 3, 30, 26):
        print("advance", n, "graph", graph_steps)
        a = eager.advance(n)
        b = g.advance(n)
        if n == 1:
            g.capture_graph()  # captured ahead of its first replay
        assert np.array_equal(a, b)
        for d in range(k.domain_count()):
            assert np.array_equal(eager.field(d, "p0"), g.field(d, "p0")), d
    p = g.field(0, "p0") * 0.5
    eager.set_field(0, "p0", p)
    g.set_field(0, "p0", p)
    assert np.array_equal(eager.advance(9), g.advance(9))
    assert np.array_equal(eager.field(1, "vx0"), g.field(1, "vx0"))


def test_graph_samples_across_sample_buffer_chunks():
    """4100 steps in one call: more than the 4096-step sample buffer, so two
    chunks, each of an eager first piece, graph replays whose samples are
    copied into place on the device, and an eager remainder."""
    steps = 4100
    conf = W.lshape(3, steps)
    _, eager = single(conf)
    eager.set_graph_steps(0)
    _, g = single(conf)
    a = eager.advance(steps)
    b = g.advance(steps)
    print("graph vs eager,", steps, "steps: max |diff|", np.abs(a - b).max(), "max |p|", np.abs(a).max())
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)


def test_autotune_cache_per_plan():
    """A second context of the same plan takes the first one's launch shape
    without trials; a different plan is tuned afresh."""
    confs = W.scene_batch(5, 10, n_scenes=4, hi=260)
    a = P.BatchSolver(confs)
    ta = a.tuning()
    b = P.BatchSolver(confs)
    tb = b.tuning()
    print("first:", ta, "second:", tb)
    assert tb["cached"] and tb["trial_steps"] == 0
    assert (tb["lean"], tb["xcd_order"], tb["tile_threads"]) == (ta["lean"], ta["xcd_order"], ta["tile_threads"])
    assert ta["cached"] or ta["trial_steps"] > 0
    c = P.BatchSolver(confs[:3])
    tc = c.tuning()
    assert tc["cached"] or tc["trial_steps"] > 0
    assert np.array_equal(a.advance(10), b.advance(10))


@pytest.mark.parametrize("shape", [(0, 1, 256), (1, 0, 256), (1, 1, 256), (2, 0, 256), (2, 1, 256), (0, 0, 512),
                                   (1, 1, 512), (2, 1, 512), (1, 1, 128), (0, 0, 128), (2, 1, 128)])
def test_launch_shape_is_bitwise_neutral(shape):
    """Every (kernel flavour, tile order, tile threads) the autotuner may pick gives the same bits."""
    conf = W.lshape(2, 25)
    k, a = single(conf)
    _, b = single(conf)
    print("autotuned:", a.tuning())
    a.set_kernel_flavour(0)
    a.set_tile_order(False)
    a.set_tile_threads(256)
    b.set_kernel_flavour(shape[0])
    b.set_tile_order(shape[1])
    b.set_tile_threads(shape[2])
    assert np.array_equal(a.advance(25), b.advance(25))
    for d in range(k.domain_count()):
        for f in ("p0", "vx0", "py0"):
            assert np.array_equal(a.field(d, f), b.field(d, f)), (d, f)


def test_long_lines_lean_flavour_bitwise():
    """F = 8192 lines in two flavours of the long-line kernel."""
    conf = P.PSTDConfiguration()
    W.base_settings(conf, 3)
    conf.Domains = [W.domain(0, 0, 4100, 40, 1.0)]
    conf.Speakers = [(W._world_pos(2050), W._world_pos(20.0), 0.0)]
    k, a = single(conf)
    _, b = single(conf)
    a.set_kernel_flavour(0)
    b.set_kernel_flavour(2)
    a.advance(3)
    b.advance(3)
    for d in range(k.domain_count()):
        assert np.array_equal(a.field(d, "p0"), b.field(d, "p0")), d


@pytest.mark.parametrize("nx, ny", [(301, 1000), (1000, 301), (77, 129)])
def test_tile_threads_bitwise(nx, ny):
    """Block sizes 128 (where every line has F <= 1024), 256 and 512 of the
    F <= 2048 class give the same bits: the line pairs are the same, only the
    number of pairs per tile differs (odd line counts leave a lone last line
    in all)."""
    conf = P.PSTDConfiguration()
    W.base_settings(conf, 4)
    conf.Domains = [W.domain(0, 0, nx, ny, 1.0)]
    conf.Speakers = [(W._world_pos(nx / 2), W._world_pos(ny / 2), 0.0)]
    k, a = single(conf)
    _, b = single(conf)
    a.set_tile_threads(256)
    b.set_tile_threads(512)
    assert b.tuning()["tile_threads"] == 512
    with pytest.raises(P.PSTDError):
        b.set_tile_threads(384)  # only 128, 256 and 512 exist
    assert b.tuning()["tile_threads"] == 512
    small = max(nx, ny) + 1 + 64 <= 1024  # every line of FFT length <= 1024
    _, c = single(conf)
    if small:
        c.set_tile_threads(128)
    else:
        with pytest.raises(P.PSTDError):
            c.set_tile_threads(128)
        c.set_tile_threads(256)
    a.advance(4)
    b.advance(4)
    c.advance(4)
    for d in range(k.domain_count()):
        for f in ("p0", "vx0", "vy0"):
            assert np.array_equal(a.field(d, f), b.field(d, f)), (d, f)
            assert np.array_equal(a.field(d, f), c.field(d, f)), (d, f)


def test_batch64_scenes_vs_golden():
    """The 64-scene batch at full length (2000 steps, FP32 FFT form, one launch
    per RK stage, replayed as 20-step hipGraphs): four of its scenes (one to
    four domains) against committed FP64 oracle series; the same four as a
    GEMM-form batch."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "batch64_receivers.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (tests/golden/make_fixtures.py --batch64)")
    g = np.load(path)
    steps = 2000
    confs = W.scene_batch(0, steps)
    picked = sorted(int(k[len("fp64_scene"):]) for k in g.files)
    assert len(picked) == 4
    full = P.BatchSolver(confs)
    full.set_graph_steps(20)  # the default, stated: 2000 steps as one eager step, replays and a remainder
    got_full = full.advance(steps)
    sub = P.BatchSolver([confs[i] for i in picked], form="gemm")
    sub.set_graph_steps(20)
    got_sub = sub.advance(steps)
    for j, i in enumerate(picked):
        want = g[f"fp64_scene{i}"]
        assert np.abs(want).max() > 1e-3
        for form, got in (("fft", full.scene_samples(got_full, i)), ("gemm", sub.scene_samples(got_sub, j))):
            e, m = rel(got, want), np.abs(got - want).max()
            print(f"batch scene {i} ({len(confs[i].Domains)} domains), {form} form, {steps} steps: "
                  f"rel-L2 vs FP64 {e:.2e}, max-abs {m:.2e}")
            assert e <= 1e-4 and m <= 5e-6, (i, form, e, m)


@pytest.mark.parametrize("n_scenes", [4, 16])
def test_auto_form_picks_a_form_and_matches_it_bitwise(n_scenes):
    """PSTD_FORM_AUTO uploads the plan in the FFT and the FP16 GEMM form and
    keeps the faster; its results are the bits of that form chosen
    explicitly, and a second upload of the same plan takes the choice from the
    cache (fewer trial steps, same form)."""
    steps = 40
    confs = W.scene_batch(0, steps, n_scenes=n_scenes)
    auto = P.BatchSolver(confs, form="auto")
    chosen = auto.form()
    assert chosen in ("fft", "gemm_f16")
    t1 = auto.tuning()["trial_steps"]
    a = auto.advance(steps)
    ref = P.BatchSolver(confs, form=chosen)
    b = ref.advance(steps)
    print(n_scenes, "scenes: auto chose", chosen, "trial steps", t1, flush=True)
    assert np.array_equal(a, b)
    again = P.BatchSolver(confs, form="auto")
    assert again.form() == chosen
    assert again.tuning()["trial_steps"] < t1
