"""The allocation-free advance (pstd_ctx_advance_static) against
pstd_ctx_advance_async: two calls in a row on two contexts, the second scene
smaller, each from host-written state (bitwise equal), then a continuation
(within the split form's tolerance); and the scene limit of its scratch."""
import numpy as np
import pytest

import openpstd_amd as P
from openpstd_amd import workloads as W

pytestmark = pytest.mark.gpu
FIELDS = ("p0", "px0", "py0", "vx0", "vy0")


def state(k, dev):
    return {(d, f): dev.field(d, f) for d in range(k.domain_count()) for f in FIELDS}


def pair(conf, form="gemm_f16"):
    out = []
    for _ in range(2):
        k = P.PSTDKernel(GPU=True)
        k.initialize_kernel(conf)
        k.set_form(form)
        dev = k.prepare()
        dev.set_graph_steps(0)
        out.append((k, dev))
    return out


def square(n, steps):
    return W.square(0, steps, n=n, n_receivers=4, receiver_radius=n // 4)


@pytest.mark.parametrize("n,steps", [(1024, 12), (512, 7)])  # the second call: a smaller scene
def test_static_advance_matches_async(n, steps):
    (k, ref), (_, dev) = pair(square(n, 3 * steps))
    assert dev.gemm_work()[0] > 0
    ref.advance_async(steps)
    dev.advance_static(steps)
    ref.synchronize()
    dev.synchronize()
    a, b = state(k, ref), state(k, dev)
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    # continuing: the static call recomputes the maxima from the state, the
    # regular one keeps the previous stage's (a scale that may differ by a
    # power of two): equal to the split form's tolerance
    ref.advance_async(steps)
    dev.advance_static(steps)
    a, b = state(k, ref), state(k, dev)
    num = sum(np.linalg.norm(a[key] - b[key]) ** 2 for key in a) ** 0.5
    den = sum(np.linalg.norm(a[key]) ** 2 for key in a) ** 0.5
    assert num <= 1e-5 * den, num / den
    # and the regular path after a static one recomputes its own maxima
    ref.advance_async(2)
    dev.advance_async(2)
    a, b = state(k, ref), state(k, dev)
    num = sum(np.linalg.norm(a[key] - b[key]) ** 2 for key in a) ** 0.5
    assert num <= 1e-5 * den, num / den


def test_static_advance_fft_form_matches():
    (k, ref), (_, dev) = pair(W.lshape(0, 20), form="fft")
    ref.advance_async(9)
    dev.advance_static(9)
    a, b = state(k, ref), state(k, dev)
    for key in a:
        assert np.array_equal(a[key], b[key]), key


def test_static_advance_rejects_more_scenes_than_its_scratch():
    confs = []
    for _ in range(1025):
        c = P.PSTDConfiguration()
        W.base_settings(c, 2)
        c.Domains = [W.domain(0, 0, 256, 128, 1.0)]  # 128 rows of 256 cells: GEMM tiles
        confs.append(c)
    b = P.BatchSolver(confs, form="gemm_f16")
    assert b.gemm_work()[0] > 0
    with pytest.raises(P.PSTDError, match="at most 1024"):
        b.advance_static(1)
    b.advance_async(1)  # the regular path is unaffected
    b.synchronize()
