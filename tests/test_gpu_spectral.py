"""Spectral-interpolation receivers (PSTD_RECEIVER_SPECTRAL) on the GPU.

openPSTD's Receiver::compute_with_si (kernel/core/Receiver.cpp:85-121) is
switched off there (:48).  The device evaluates it after every step as two
dot-product passes over the raw stage-5 output (the X shift of every line the
receiver's Y line reads, then the Y shift of the receiver's column), with
weights built in double from each line's segment.  Checked against:

* the FP64 oracle's series of the L-shaped scene with ten spectral receivers
  (tests/golden/lshape_spectral_receivers.npz, 500 steps; the oracle itself
  is pinned by tests/test_oracle_spectral.py): receiver series with signal
  rel-L2 <= 1e-4, every series max-abs <= 2e-6, as the nearest-cell ones;
* the nearest-cell receivers: at integer locations the two modes agree to
  FP32 rounding of the dot products (<= 1e-6 of the peak);
* graph replay and eager launches: bitwise equal;
* the GEMM form's state: the same spectral receivers of the 1024 x 1024
  domain in the FFT and the FP16 GEMM form agree to the forms' own
  difference (rel-L2 <= 1e-4);
* batching: two scenes with spectral receivers advanced together give the
  bits each gives alone.
"""
import os

import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import abi_types as T
from openpstd_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def device(conf, form="fft"):
    k = P.PSTDKernel(GPU=True)
    k.initialize_kernel(conf)
    if form != "fft":
        k.set_form(form)
    return k, k.prepare()


def test_lshape_spectral_vs_golden():
    path = os.path.join(HERE, "golden", "lshape_spectral_receivers.npz")
    if not os.path.exists(path):
        pytest.skip("FP64 fixture not generated (make_fixtures.py --lshape-spectral)")
    g = np.load(path)
    steps = g["fp64"].shape[0]
    conf = W.lshape_spectral(steps)
    k, dev = device(conf)
    print("L-shape, spectral receivers,", steps, "steps", flush=True)
    got = dev.advance(steps)
    want = g["fp64"]
    assert got.shape == want.shape
    assert np.isfinite(got).all()
    for r in range(want.shape[1]):
        e = rel(got[:, r], want[:, r])
        e32 = rel(g["fp32"][:, r], want[:, r])
        m = np.abs(got[:, r] - want[:, r]).max()
        print(f"  receiver {r}: peak {np.abs(want[:, r]).max():.2e} rel-L2 {e:.2e} (CPU FP32 {e32:.2e}) max-abs {m:.2e}")
        assert m <= 2e-6
        if np.abs(want[:, r]).max() > 1e-2:
            assert e <= 1e-4


def test_zero_offsets_equal_nearest_cell():
    from test_oracle_spectral import conf_of
    doms = [W.domain(0, 0, 66, 48, 1.0), W.domain(66, 0, 66, 48, 1.0)]
    rec = [(40.0, 20.0), (65.0, 30.0), (66.0, 12.0), (100.0, 46.0), (3.0, 0.0)]
    spk = (W._world_pos(46.5), W._world_pos(24.5), 0.0)
    steps = 200
    out = {}
    for mode in (T.PSTD_RECEIVER_SPECTRAL, T.PSTD_RECEIVER_PHYSICAL):
        k, dev = device(conf_of(doms, rec, mode, steps, spk, exact=True))
        out[mode] = dev.advance(steps)
    a, b = out[T.PSTD_RECEIVER_SPECTRAL], out[T.PSTD_RECEIVER_PHYSICAL]
    print("spectral vs nearest cell at integer locations: max |diff|", np.abs(a - b).max(), "peak", np.abs(b).max())
    assert np.abs(b).max() > 1e-2
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_spectral_graph_replay_is_bitwise_eager():
    steps = 130
    conf = W.lshape_spectral(steps)
    _, eager = device(conf)
    eager.set_graph_steps(0)
    _, g = device(conf)
    g.set_graph_steps(20)
    a = eager.advance(steps)
    b = g.advance(steps)
    assert np.array_equal(a, b)


def test_spectral_receivers_in_gemm_form():
    steps = 200
    conf = W.square(0, steps, n=1024, n_receivers=8, receiver_radius=60)
    rng = np.random.RandomState(3)
    conf.Receivers = [(x + float(rng.uniform(0, 1)) * W.DX, y + float(rng.uniform(0, 1)) * W.DX, z)
                      for x, y, z in conf.Receivers]
    conf.receiver_mode = T.PSTD_RECEIVER_SPECTRAL
    series = {}
    for form in ("fft", "gemm"):
        k, dev = device(conf, form)
        print("1024^2 spectral receivers, form", form, flush=True)
        series[form] = dev.advance(steps)
    a, b = series["gemm"], series["fft"]
    assert np.isfinite(a).all()
    e = rel(a, b)
    print("GEMM vs FFT form, spectral receivers: rel-L2", e)
    assert np.abs(b).max() > 1e-2
    assert e <= 1e-4


def test_spectral_receivers_in_a_batch_are_bitwise_alone():
    """Two scenes with spectral receivers advanced as one batch (one launch per
    RK stage, the interpolation passes of both) give the bits each gives alone."""
    steps = 90
    a = W.lshape_spectral(steps)
    b = W.lshape_spectral(steps)
    b.Speakers = [(W._world_pos(300.5), W._world_pos(60.5), 0.0)]
    batch = P.BatchSolver([a, b])
    both = batch.advance(steps)
    for i, conf in enumerate((a, b)):
        _, dev = device(conf)
        alone = dev.advance(steps)
        assert np.array_equal(batch.scene_samples(both, i), alone), i
