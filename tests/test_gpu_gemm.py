"""GEMM form (dense derivative matrices on MFMA, FP16 hi+lo split operands,
FP32 accumulation) against the FP64 oracle and against the FFT form.

Tolerances (the source's peak pressure is 1): fields rel-L2 <= 2e-5 after a few
steps; receiver series rel-L2 <= 1e-4 vs the FP64 oracle (the accuracy bar of
the split-precision form) and max-abs <= 5e-6.
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


@pytest.mark.parametrize("steps", [1, 4])
def test_lshape_fields_gemm(steps):
    conf = W.lshape(0, steps)
    k, dev = device(conf, "gemm")
    dev.advance(steps)
    o = ob.OracleScene(conf, 64, THREADS)
    o.step(steps)
    errs = field_errors(k, dev, o)
    print("GEMM form, L-shape", steps, "steps, field rel-L2 vs FP64:", errs)
    assert max(errs.values()) <= 2e-5


def test_room_series_gemm():
    steps = 300
    conf = W.room(steps)
    k, dev = device(conf, "gemm")
    got = dev.advance(steps)
    want = ob.OracleScene(conf, 64, THREADS).step(steps)
    e = rel(got, want)
    print(f"GEMM form, room {steps} steps: receiver rel-L2 vs FP64 {e:.2e}, max-abs {np.abs(got - want).max():.2e}")
    assert e <= 1e-4
    assert np.abs(got - want).max() <= 5e-6


def test_lshape_series_gemm_vs_golden():
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lshape_receivers.npz"))
    steps = 600
    conf = W.lshape(0, 2000)
    k, dev = device(conf, "gemm")
    got = dev.advance(steps)
    want = g["fp64"][:steps]
    e = rel(got, want)
    print(f"GEMM form, L-shape {steps} steps: receiver rel-L2 vs FP64 {e:.2e}, max-abs {np.abs(got - want).max():.2e}")
    assert e <= 1e-4
    assert np.abs(got - want).max() <= 5e-6


def test_gemm_batch_is_bitwise_single():
    confs = W.scene_batch(11, 10, n_scenes=3, hi=300)
    b = P.BatchSolver(confs, form="gemm")
    got = b.advance(10)
    for i, conf in enumerate(confs):
        k, dev = device(conf, "gemm")
        assert np.array_equal(b.scene_samples(got, i), dev.advance(10)), i
        for d in range(k.domain_count()):
            assert np.array_equal(b.field(i, d, "p0"), dev.field(d, "p0")), (i, d)
