"""The pair GEMM on the fp16 matrix cores (precisions FP16X2 and FP16)
against the oracle, and the mixed-precision solve built on it.

Bars: FP16X2 (pair factors and x each as two fp16 terms) within 2e-5 of the
oracle's FP64 particle_interaction_scalar, the complex64 GEMM's bar; FP16
(one term) within 2e-3; solves with either inner operator meet the FP64
criterion ||b - A x|| <= 1e-6 ||b|| against the oracle's operator."""

import numpy as np
import pytest

from oracle import restate as R
from yasf_amd import _native
from yasf_amd.clusters import cluster

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("lmax", [1, 2, 3, 4, 5, 6, 7, 8])
def test_half_matvec_all_orders(lmax):
    rng = np.random.default_rng(300 + lmax)
    N = 37  # not a multiple of the 16-particle tile or the 32-source K-step
    pos, _ = cluster(40 + lmax, N, 0.1)
    k = 2 * np.pi / 520.0
    n = N * 2 * lmax * (lmax + 2)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    s = _native.Session(pos, lmax, [k])
    s.set_matvec_mode(_native.MATVEC_TABLE)
    ref = R.coupling_matvec(lmax, pos, k, x)
    y16x2 = s.coupling_matvec(x, precision=_native.FP16X2)
    y16 = s.coupling_matvec(x, precision=_native.FP16)
    assert rel(y16x2, ref) < 2e-5
    assert rel(y16, ref) < 2e-3
    # again (the one-term request reuses the two-term table), and repeatable
    assert np.array_equal(s.coupling_matvec(x, precision=_native.FP16X2), y16x2)
    s.close()


def test_half_matvec_scaled_columns():
    # columns of very different magnitude (like the high-order modes of a
    # Krylov vector, ~|t_l|) keep their relative accuracy: each mode column
    # has its own power-of-two scale
    lmax, N = 6, 50
    pos, radii = cluster(7, N, 0.1)
    k = 2 * np.pi / 600.0
    t = R.mie_coefficients(lmax, radii, np.full(N, 1.5 + 0.01j), np.array([600.0]), 1.0)[:, :, 0].ravel()
    rng = np.random.default_rng(8)
    x = t * (rng.normal(size=t.size) + 1j * rng.normal(size=t.size))
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    s.set_matvec_mode(_native.MATVEC_TABLE)
    ref = R.master_matvec(lmax, pos, k, t, x)
    got = s.master_matvec(x, precision=_native.FP16X2)
    assert rel(got, ref) < 2e-5
    # column-wise: the modes of the largest order l = lmax, tiny in x
    nmax = 2 * lmax * (lmax + 2)
    cols = [n for n in range(nmax) if n % (nmax // 2) >= lmax * lmax - 1]
    idx = (np.arange(N)[:, None] * nmax + np.array(cols)[None, :]).ravel()
    assert rel(got[idx], ref[idx]) < 2e-5
    s.close()


@pytest.mark.parametrize("precision", [_native.FP16X2, _native.FP16], ids=["fp16x2", "fp16"])
def test_refined_solve_with_half_inner_operator(precision):
    lmax, N, lam = 4, 50, 550.0
    pos, radii = cluster(1, N, 0.1)
    k = 2 * np.pi / lam
    t = R.mie_coefficients(lmax, radii, np.full(N, 1.5 + 0.01j), np.array([lam]), 1.0)[:, :, 0].ravel()
    a = R.planewave_coefficients(lmax, pos, np.array([k]), 0.0, 0.0, 1)[:, :, 0].ravel()
    b = t * a
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    s.set_inner_precision(precision)
    x, info, steps, iters, res, hist = s.gmres_refined(b, b, rtol=1e-6, inner_rtol=1e-4, restart=50)
    assert info == 0 and res <= 1e-6
    true_res = np.linalg.norm(b - R.master_matvec(lmax, pos, k, t, x)) / np.linalg.norm(b)
    assert true_res <= 1e-6
    x64, info64, _, _, _ = s.gmres(b, b, rtol=1e-10, restart=100, maxiter=100)
    assert info64 == 0
    assert rel(x, x64) < 1e-5
    c16 = s.cross_sections(a, x)
    c64 = s.cross_sections(a, x64)
    assert abs(c16[0] - c64[0]) <= 1e-6 * abs(c64[0]) and abs(c16[1] - c64[1]) <= 1e-6 * abs(c64[1])
    s.close()


def test_half_precision_needs_a_resident_table():
    pos, _ = cluster(3, 20, 0.1)
    s = _native.Session(pos, 3, [0.01])
    s.set_matvec_mode(_native.MATVEC_ON_THE_FLY)
    with pytest.raises(ValueError):
        s.coupling_matvec(np.ones(20 * 30, dtype=complex), precision=_native.FP16X2)
    with pytest.raises(ValueError):
        s.set_inner_precision(_native.FP64)
    s.close()
