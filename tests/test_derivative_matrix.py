"""The dense derivative matrix of the GEMM form against the oracle's spatderp3.

D (built in FP64 by the product library, pstd_derivative_matrix) applied to
[side1 tail | main | side2 head] must give spatderp3's FP64 output for the same
line (kernel/core/kernel_functions.cpp:123-318) to round-off, for pressure and
velocity lines, asymmetric stencils, and the half window of short domains.
"""
import ctypes as C
import math

import numpy as np
import pytest

from openpstd_amd._lib import lib
from oracle import binding as ob


def product_matrix(F, L, field, dx, w, n_main, coef, pe=70.0):
    c = (C.c_double * 4)(*coef)
    n = lib().pstd_derivative_matrix(F, L, field, dx, w, n_main, pe, c, None, 0)
    assert n > 0
    out = np.zeros(n)
    assert lib().pstd_derivative_matrix(F, L, field, dx, w, n_main, pe, c,
                                        out.ctypes.data_as(C.POINTER(C.c_double)), n) == n
    n_out = n_main + 1 if field == 0 else n_main - 1
    return out.reshape(n_out, n_main + 2 * w)


CASES = [
    # n, field, w, rho1, rho2
    (100, 0, 32, 1.2, 1.2),
    (100, 1, 32, 1.2, 1.2),
    (77, 0, 32, 2.5, 0.6),
    (77, 1, 32, 2.5, 0.6),
    (50, 0, 16, 1e30, 3.0),   # rigid left neighbour, half window
    (50, 1, 16, 1e30, 3.0),
    (300, 1, 32, 0.9, 1e30),
]


@pytest.mark.parametrize("n, field, w, rho1, rho2", CASES)
def test_matrix_matches_spatderp3(n, field, w, rho1, rho2):
    rng = np.random.RandomState(n + 7 * field)
    dx = 0.2
    nm = n + field
    F = 2 ** math.ceil(math.log2(nm + 2 * w))
    p1 = rng.randn(3, 40 + field)
    p2 = rng.randn(3, nm)
    p3 = rng.randn(3, 45 + field)
    disc = ob.discretization(dx, F, 64)
    win = ob.window(w, 70, 64)
    rp, rv = ob.rho_array(rho1, 1.2, rho2, 64)
    R = rp if field == 0 else rv
    want = ob.spatderp3(p1, p2, p3, disc["pressure" if field == 0 else "velocity"], rp, rv, win, w, field, 0, 64)
    D = product_matrix(F, len(disc["pressure"]), field, dx, w, nm, [R[2][1], R[0][0], R[3][1], R[1][0]])
    if field == 0:
        u = np.concatenate([p1[:, -w:], p2, p3[:, :w]], axis=1)
    else:
        u = np.concatenate([p1[:, -(w + 1):-1], p2, p3[:, 1:w + 1]], axis=1)
    got = u @ D.T
    assert got.shape == want.shape
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-11 * scale


def test_interior_rows_are_toeplitz():
    """Away from the window the operator is a convolution: rows shift by one column."""
    F, n, w = 512, 300, 32
    D = product_matrix(F, F, 0, 0.2, w, n, [0.0, 0.0, 0.0, 0.0])
    i = 150
    assert np.allclose(D[i, w + 50:w + 250], D[i + 1, w + 51:w + 251], rtol=0, atol=1e-14)


def test_rejects_inconsistent_shape():
    c = (C.c_double * 4)(0, 0, 0, 0)
    assert lib().pstd_derivative_matrix(100, 128, 0, 0.2, 32, 100, 70.0, c, None, 0) < 0


def product_shift_row(F, L, dx, w, n, coef, offset, row, pe=70.0):
    c = (C.c_double * 4)(*coef)
    K = lib().pstd_shift_operator_row(F, L, dx, w, n, pe, c, offset, row, None, 0)
    assert K == n + 2 * w
    out = np.zeros(K)
    assert lib().pstd_shift_operator_row(F, L, dx, w, n, pe, c, offset, row,
                                         out.ctypes.data_as(C.POINTER(C.c_double)), K) == K
    return out


SHIFT_CASES = [
    # n, w, rho1, rho2, offset, row
    (100, 32, 1.2, 1.2, 0.37, 40),
    (77, 32, 2.5, 0.6, 0.9, 0),
    (77, 32, 2.5, 0.6, 0.05, 76),
    (50, 16, 1e30, 3.0, 0.5, 3),   # rigid left neighbour, half window
    (300, 32, 0.9, 1e30, 0.71, 299),
]


@pytest.mark.parametrize("n, w, rho1, rho2, offset, row", SHIFT_CASES)
def test_shift_row_matches_spatderp3(n, w, rho1, rho2, offset, row):
    """The spectral receivers' weights: spatderp3 of a pressure line with the
    shift factors of Receiver::get_fft_factors in place of derfact, one output."""
    rng = np.random.RandomState(n + row)
    dx = 0.2
    F = 2 ** math.ceil(math.log2(n + 2 * w))
    p1, p2, p3 = rng.randn(3, 40), rng.randn(3, n), rng.randn(3, 45)
    fac = ob.shift_factors(dx, F, offset)
    assert len(fac) == F
    win = ob.window(w, 70, 64)
    rp, rv = ob.rho_array(rho1, 1.2, rho2, 64)
    want = ob.spatderp3(p1, p2, p3, fac, rp, rv, win, w, 0, 0, 64)[:, row]
    r = product_shift_row(F, F, dx, w, n, [rp[2][1], rp[0][0], rp[3][1], rp[1][0]], offset, row)
    got = np.concatenate([p1[:, -w:], p2, p3[:, :w]], axis=1) @ r
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_shift_row_zero_offset_is_the_cell():
    r = product_shift_row(256, 256, 0.2, 32, 150, [0.0, 0.0, 0.0, 0.0], 0.0, 70)
    e = np.zeros_like(r)
    e[32 + 70] = 1.0
    assert np.abs(r - e).max() < 1e-14
