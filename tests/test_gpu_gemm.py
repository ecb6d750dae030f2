"""GEMM form (dense derivative matrices on MFMA, 16-bit hi+lo split operands,
FP32 accumulation) against the FP64 oracle and against the FFT form.

Two operand types: FP16 (22-bit significands after the split; inputs scaled
per scene by a power of two) and BF16 (16-bit significands).
Tolerances (the source's peak pressure is 1): FP16 fields rel-L2 <= 2e-5 after
a few steps; receiver series rel-L2 <= 1e-4 vs the FP64 oracle (the accuracy
bar of the split-precision form) and max-abs <= 5e-6.  BF16 keeps 6 bits
less: its bounds (TOL) are what it reaches, stated per test.
"""
import os

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import workloads as W
from oracle import binding as ob

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
FORMS = ["gemm_f16", "gemm_bf16"]
# (fields after a few steps, receiver rel-L2, receiver max-abs) per operand type
TOL = {"gemm_f16": (2e-5, 1e-4, 5e-6), "gemm_bf16": (2e-3, 5e-3, 2e-4)}


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def gemm_pair(steps, seed=0):
    """Two 512 x 384-cell domains side by side (edge absorption 0.3 / 1.0,
    PML around): both line directions of both domains fill the 128-line x
    256-output GEMM tiles at least to 3/4 (X: 384 lines, 512 outputs; Y: 512
    lines, 384 outputs), and the shared edge couples them through the side
    halos."""
    rng = np.random.RandomState(seed)
    conf = P.PSTDConfiguration()
    W.base_settings(conf, steps)
    conf.Domains = [W.domain(0, 0, 512, 384, {"T": 1.0, "L": 0.3, "B": 1.0, "R": 0.0}),
                    W.domain(512, 0, 512, 384, {"T": 0.3, "L": 0.0, "B": 1.0, "R": 1.0})]
    conf.Speakers = [(W._world_pos(440.5), W._world_pos(190.5), 0.0)]
    for _ in range(4):  # within reach of the wave in a few hundred steps, on both sides of the shared edge
        conf.Receivers.append((W._world_pos(rng.randint(380, 600) + 0.25), W._world_pos(rng.randint(140, 240) + 0.25),
                               0.0))
    return conf


def device(conf, form):
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(conf)
    k.set_form(form)
    return k, k.prepare()


def field_errors(k, dev, o, names=("p0", "vx0", "vy0")):
    out = {}
    for f in names:
        a = np.concatenate([dev.field(i, f).ravel() for i in range(k.domain_count())])
        b = np.concatenate([o.field(i, f).ravel() for i in range(k.domain_count())])
        out[f] = rel(a, b)
    return out


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("steps", [1, 4])
def test_pair_fields_gemm(steps, form):
    conf = gemm_pair(steps)
    k, dev = device(conf, form)
    flops, outs = dev.gemm_work()
    assert flops > 0 and outs >= 2 * 2 * 2 * 512 * 384  # both directions and fields of both domains
    dev.advance(steps)
    o = ob.OracleScene(conf, 64, THREADS)
    o.step(steps)
    errs = field_errors(k, dev, o)
    print(form, "two-domain pair", steps, "steps, field rel-L2 vs FP64:", errs)
    assert max(errs.values()) <= TOL[form][0]


@pytest.mark.parametrize("form", FORMS)
def test_pair_series_gemm(form):
    steps = 300
    conf = gemm_pair(steps, seed=1)
    k, dev = device(conf, form)
    got = dev.advance(steps)
    want = ob.OracleScene(conf, 64, THREADS).step(steps)
    _, fft = device(conf, "fft")
    got_fft = fft.advance(steps)
    e, e_fft = rel(got, want), rel(got_fft, want)
    print(f"two-domain pair {steps} steps: receiver rel-L2 vs FP64: {form} {e:.2e}, FFT form {e_fft:.2e}; "
          f"max-abs {np.abs(got - want).max():.2e}")
    assert np.abs(want).max() > 1e-2
    assert e <= TOL[form][1]
    assert np.abs(got - want).max() <= TOL[form][2]


def test_lshape_mixed_forms_vs_golden():
    """On the L-shape only the velocity X segments fill GEMM blocks (256
    outputs); the other segments keep the FFT form in the same stage launches."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lshape_receivers.npz"))
    steps = 600
    conf = W.lshape(0, 2000)
    _, dev = device(conf, "gemm")
    assert dev.gemm_work()[0] > 0
    got = dev.advance(steps)
    want = g["fp64"][:steps]
    e = rel(got, want)
    print(f"mixed forms, L-shape {steps} steps: receiver rel-L2 vs FP64 {e:.2e}, max-abs {np.abs(got - want).max():.2e}")
    assert e <= 1e-4
    assert np.abs(got - want).max() <= 5e-6


def test_gemm_batch_is_bitwise_single():
    confs = [gemm_pair(10, seed=s) for s in range(2)] + W.scene_batch(11, 10, n_scenes=2, hi=300)
    b = P.BatchSolver(confs, form="gemm")
    assert b.gemm_work()[0] > 0
    got = b.advance(10)
    for i, conf in enumerate(confs):
        k, dev = device(conf, "gemm")
        assert np.array_equal(b.scene_samples(got, i), dev.advance(10)), i
        for d in range(k.domain_count()):
            assert np.array_equal(b.field(i, d, "p0"), dev.field(d, "p0")), (i, d)


@pytest.mark.parametrize("graph_steps", [0, 20])
def test_stage_overlap_is_bitwise(graph_steps):
    """The FFT-form tiles of a GEMM-form stage on the side stream, beside the
    GEMM kernel, give the same bits as the two issued in order (eagerly and
    replayed as graphs); a run longer than one graph."""
    confs = [gemm_pair(45, seed=s) for s in range(2)] + W.scene_batch(11, 45, n_scenes=2, hi=300)
    runs = []
    for overlap in (True, False):
        b = P.BatchSolver(confs, form="gemm")
        b.set_graph_steps(graph_steps)
        b.set_stage_overlap(overlap)
        got = b.advance(45)
        runs.append((got, [b.field(i, 0, "p0") for i in range(len(confs))]))
        print("overlap", overlap, "graph", graph_steps, "max |p|", np.abs(got).max())
    assert np.abs(runs[0][0]).max() > 0
    assert np.array_equal(runs[0][0], runs[1][0])
    for x, y in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(x, y)


def test_square1024_forms_vs_golden():
    """A 1024 x 1024-cell domain (GEMM tiles filled in both directions) against
    the committed FP64 oracle series (300 steps, 16 receivers near the source),
    in every form; the FP16 GEMM form within the 1e-4 bar."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "square1024_receivers.npz"))
    conf = W.square(0, 300, n=1024, n_receivers=16, receiver_radius=120)
    errs = {}
    for form in FORMS + ["fft"]:
        _, dev = device(conf, form)
        if form != "fft":
            assert dev.gemm_work()[0] > 0
        got = dev.advance(300)
        errs[form] = (rel(got, g["fp64"]), np.abs(got - g["fp64"]).max())
    e32 = rel(g["fp32"], g["fp64"])
    print("1024^2, 300 steps, receiver rel-L2 / max-abs vs FP64: " +
          ", ".join(f"{f} {e:.2e} / {m:.2e}" for f, (e, m) in errs.items()) + f", CPU FP32 oracle {e32:.2e}")
    for form in FORMS:
        assert errs[form][0] <= TOL[form][1] and errs[form][1] <= TOL[form][2], form
    assert errs["fft"][0] <= 1e-4


def test_sources1024_per_source_vs_golden():
    """The 256-source geometry (1024 x 1024 cells as two 512 x 1024 domains, PML
    around): its first seeded source positions (as many as the fixture holds,
    at least two) advanced as one batch for the full 3000 steps, 8 seeded
    receivers each, against the committed FP64 oracle series, error stated per
    source; FP16 GEMM and FFT within the 1e-4 bar, BF16 within its own."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sources1024_receivers.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (tests/golden/make_fixtures.py --sources)")
    g = np.load(path)
    steps = 3000
    n_src = len([k for k in g.files if k.startswith("fp64_source")])
    assert n_src >= 2
    confs = W.source_batch(0, steps, n_sources=n_src)
    errs = {}
    for form in FORMS + ["fft"]:
        b = P.BatchSolver(confs, form=form)
        if form != "fft":
            assert b.gemm_work()[0] > 0
        got = b.advance(steps)
        for i in range(len(confs)):
            s, want = b.scene_samples(got, i), g[f"fp64_source{i}"]
            assert np.abs(want).max() > 1e-3
            errs[(form, i)] = (rel(s, want), np.abs(s - want).max())
    for (form, i), (e, m) in sorted(errs.items()):
        print(f"source {i}, {form} form, 3000 steps: receiver rel-L2 vs FP64 {e:.2e}, max-abs {m:.2e}")
    for (form, i), (e, m) in errs.items():
        t = TOL.get(form, (0, 1e-4, 5e-6))
        assert e <= t[1], (form, i, e)
        assert m <= t[2], (form, i, m)


def test_square4096_forms_vs_golden():
    """The 4096 x 4096-cell bench domain (F = 8192 lines), 500 steps, 16
    receivers within 200 cells of the centre source, against the committed FP64
    oracle series; both forms within the 1e-4 bar."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "square4096_receivers.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (tests/golden/make_fixtures.py --square4096)")
    want = np.load(path)["fp64"]
    steps = 500
    conf = W.square(0, steps, receiver_radius=200)
    assert np.abs(want).max() > 1e-3
    for form in FORMS + ["fft"]:
        _, dev = device(conf, form)
        if form != "fft":
            assert dev.gemm_work()[0] > 0
        got = dev.advance(steps)
        e, m = rel(got, want), np.abs(got - want).max()
        print(f"4096^2, {steps} steps, {form} form: receiver rel-L2 vs FP64 {e:.2e}, max-abs {m:.2e}")
        t = TOL.get(form, (0, 1e-4, 5e-6))
        assert e <= t[1] and m <= t[2], (form, e, m)


def test_square4096_long_vs_golden():
    """The 4096 x 4096-cell domain at its configured 1000 steps (16 receivers
    within 400 cells of the centre source) against the FP64 oracle series, as
    far as the committed fixture reaches (the CPU oracle needs ~18 s per step
    of this scene on 8 cores); FFT form and both GEMM forms."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "square4096_long_receivers.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (tests/golden/make_fixtures.py --square4096-long)")
    want = np.load(path)["fp64"]
    steps = len(want)
    conf = W.square(0, 1000, receiver_radius=400)
    assert np.abs(want).max() > 1e-3
    for form in FORMS + ["fft"]:
        _, dev = device(conf, form)
        got = dev.advance(steps)
        e, m = rel(got, want), np.abs(got - want).max()
        print(f"4096^2, {steps} of 1000 steps, {form} form: receiver rel-L2 vs FP64 {e:.2e}, max-abs {m:.2e}")
        t = TOL.get(form, (0, 1e-4, 5e-6))
        assert e <= t[1] and m <= t[2], (form, e, m)


def test_sources256_one_gemm_batch():
    """All 256 seeded source positions of the 1024 x 1024 geometry in one
    GEMM-form batch (one launch per kernel class and RK stage), 120 steps:
    every series finite; sources 0 and 1 bitwise equal to a batch of only
    them (the per-scene scaling makes scenes independent); every source's p0
    fields within 1e-4 rel-L2 of its FFT-form run."""
    steps = 120
    confs = W.source_batch(0, steps)
    assert len(confs) == 256
    ref = P.BatchSolver(confs, form="fft")
    want = ref.advance(steps)
    p_fft = [np.concatenate([ref.field(i, d, "p0").ravel() for d in range(2)]) for i in range(256)]
    del ref
    b = P.BatchSolver(confs, form="gemm_f16")
    assert b.gemm_work()[0] > 0
    got = b.advance(steps)
    assert np.isfinite(got).all()
    errs = []
    for i in range(256):
        p = np.concatenate([b.field(i, d, "p0").ravel() for d in range(2)])
        errs.append(rel(p, p_fft[i]))
    small = P.BatchSolver(confs[:2], form="gemm_f16")
    got2 = small.advance(steps)
    for i in range(2):
        assert np.array_equal(b.scene_samples(got, i), small.scene_samples(got2, i)), i
        for d in range(2):
            assert np.array_equal(b.field(i, d, "p0"), small.field(i, d, "p0")), (i, d)
    errs = np.array(errs)
    print(f"256 sources, {steps} steps, FP16 GEMM vs FFT form p0 rel-L2: max {errs.max():.2e}, median "
          f"{np.median(errs):.2e}; receiver max-abs difference {np.abs(got - want).max():.2e}")
    assert errs.max() <= 1e-4


@pytest.mark.skipif(not os.environ.get("PSTD_FULL_C5"),
                    reason="all 256 sources for 3000 steps in two forms: ~6 minutes of GPU time; set PSTD_FULL_C5=1")
def test_sources256_full_length_gemm_vs_fft():
    """The headline configuration as configured: all 256 seeded source
    positions of the 1024 x 1024 geometry for the full 3000 steps, as one
    FP16 GEMM batch and as one FFT batch; every source's 8 receiver series
    in the two forms within 1e-4 rel-L2 of each other (the FFT form itself is
    within ~7.5e-6 of the FP64 oracle on the sources the fixture holds,
    which are checked here as well)."""
    steps, chunk = 3000, 500
    confs = W.source_batch(0, steps)
    assert len(confs) == 256
    series = {}
    for form in ("fft", "gemm_f16"):
        b = P.BatchSolver(confs, form=form)
        parts = []
        for s0 in range(0, steps, chunk):
            parts.append(b.advance(chunk))
            print(f"{form}: {s0 + chunk} of {steps} steps", flush=True)
        got = np.concatenate(parts)
        assert np.isfinite(got).all()
        series[form] = [b.scene_samples(got, i).copy() for i in range(256)]
        del b
    errs = np.array([rel(series["gemm_f16"][i], series["fft"][i]) for i in range(256)])
    print(f"256 sources, {steps} steps: FP16 GEMM vs FFT receiver rel-L2 max {errs.max():.2e} "
          f"(source {errs.argmax()}), median {np.median(errs):.2e}")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sources1024_receivers.npz")
    if os.path.exists(path):
        g = np.load(path)
        for k in sorted(g.files):
            i = int(k[len("fp64_source"):])
            for form in ("fft", "gemm_f16"):
                e = rel(series[form][i], g[k])
                print(f"source {i}, {form} form in the 256-source batch: rel-L2 vs FP64 {e:.2e}")
                assert e <= 1e-4, (form, i, e)
    assert errs.max() <= 1e-4
