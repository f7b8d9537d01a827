"""Oracle (test infrastructure only): CPU restatement of the reference's
multiple-sphere T-matrix solve path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the reference that results are
compared against or as the timed CPU baseline.  The product package (``yasf_amd``) never imports it.

Every function cites the reference file:line it restates.  The pairwise
special-function tables use the same library routines the reference calls
(``scipy.special.spherical_jn/yn``); the Wigner-3j symbols are computed in exact
rational arithmetic (Racah's formula) instead of the reference's pywigxjpf,
which is absent here.  The O(N^2) coupling loop itself runs in ``matvec.c``
(loaded through ctypes, see ``oracle/__init__.py``).

How each function is pinned (tests/test_oracle_*.py):
  idx_lookup / multi2single_index   multi2single_index_data.mat
  legendre_normalized_trigon        legendre_normalized_trigon_data.mat + scipy lpmv closed form
  mutual_lookup (h_p)               sph_bessel_data.mat + SciPy spherical_jn/yn (library)
  wigner3j                          SymPy wigner_3j (exact) + orthogonality relations
  translation_table                 translation_table_ab_data.mat (L=3)
  coupling_matvec                   coupling_matrix_multiply_data.mat (fp32 MATLAB, rtol 2e-4),
                                    N=1 -> 0, z-axis pair selection rule, reciprocity symmetry
  mie_coefficients                  t_entry_data.mat, full_simulation_data.mat
  planewave_coefficients            full_simulation_data.mat, CELES FITS initial coefficients
  solve (GMRES, SciPy)              full_simulation_data.mat (1e-3), CELES 500-sphere FITS (2e-3)
  block preconditioner              one box == exact inverse (1 iteration), M_box M_box^-1 = I
  cross sections                    lossless energy conservation C_ext == C_sca, Mie series (N=1)
"""

from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

import numpy as np
from scipy.special import spherical_jn, spherical_yn

# ---------------------------------------------------------------- indexing


def nmax_of(lmax: int) -> int:
    """Coefficients per particle, ``2 L (L + 2)`` (misc.py:9-21)."""
    return 2 * lmax * (lmax + 2)


def multi2single_index(s: int, tau: int, l: int, m: int, lmax: int) -> int:
    """Flat index of (s, tau, l, m) (yasfpy/functions/misc.py:23-43)."""
    return s * 2 * lmax * (lmax + 2) + (tau - 1) * lmax * (lmax + 2) + (l - 1) * (l + 1) + m + l


def idx_lookup(lmax: int, particle_number: int) -> np.ndarray:
    """(s, n, tau, l, m) per flat index (yasfpy/functions/cpu_numba.py:370-407)."""
    nmax = nmax_of(lmax)
    idx = np.zeros((nmax * particle_number, 5), dtype=np.int64)
    for s in range(particle_number):
        for tau in (1, 2):
            for l in range(1, lmax + 1):
                for m in range(-l, l + 1):
                    n = (tau - 1) * lmax * (lmax + 2) + (l - 1) * (l + 1) + l + m
                    idx[n + s * nmax] = (s, n, tau, l, m)
    return idx


# ---------------------------------------------------------------- Legendre


def legendre_normalized_trigon(lmax: int, ct: np.ndarray, st: np.ndarray) -> np.ndarray:
    """Normalized associated Legendre functions without Condon-Shortley phase,
    ``plm[l, m, ...]`` (yasfpy/functions/legendre_normalized_trigon.py:59-91)."""
    size = np.shape(ct)
    ct = np.ravel(ct).astype(float)
    st = np.ravel(st).astype(float)
    plm = np.zeros((lmax + 1, lmax + 1, ct.size))
    plm[0, 0] = np.sqrt(1 / 2)
    if lmax >= 1:
        plm[1, 0] = np.sqrt(3 / 2) * ct
    for l in range(1, lmax):
        plm[l + 1, 0] = (
            1 / (l + 1) * np.sqrt((2 * l + 1) * (2 * l + 3)) * plm[l, 0] * ct
            - l / (l + 1) * np.sqrt((2 * l + 3) / (2 * l - 1)) * plm[l - 1, 0]
        )
    for m in range(1, lmax + 1):
        plm[m - 1, m] = 0.0
        plm[m, m] = (
            np.sqrt((2 * m + 1) / 2 / math.factorial(2 * m))
            * np.prod(np.arange(1, 2 * m, 2))
            * np.power(st, m)
        )
        for l in range(m, lmax):
            plm[l + 1, m] = (
                np.sqrt((2 * l + 1) * (2 * l + 3) / (l + 1 - m) / (l + 1 + m)) * ct * plm[l, m]
                - np.sqrt((2 * l + 3) * (l - m) * (l + m) / (2 * l - 1) / (l + 1 - m) / (l + 1 + m))
                * plm[l - 1, m]
            )
    return plm.reshape((lmax + 1, lmax + 1) + size)


# ---------------------------------------------------------------- pair tables


def mutual_lookup(lmax: int, pos1: np.ndarray, pos2: np.ndarray, k_medium: np.ndarray):
    """Pairwise tables (yasfpy/functions/misc.py:117-243, non-derivative branch).

    Returns ``sph_j, sph_h`` with shape ``[2L+1, n1, n2, W]``, ``e_j_dm_phi``
    ``[4L+1, n1, n2]`` and the flattened ``plm`` ``[(2L+1)(2L+2)/2, n1, n2]``.
    Geometry follows the reference: ``d = r_s1 - r_s2`` (:149), cos(theta)
    clipped to [-1, 1] (:163-165), ``phi = atan2(dy, dx)`` (:166), ``y_p``
    masked to 0 at zero distance (:187-193).
    """
    k_medium = np.atleast_1d(np.asarray(k_medium))
    diff = pos1[:, None, :] - pos2[None, :, :]
    dist = np.sqrt(np.sum(diff ** 2, axis=2))[:, :, None]
    e_r = np.divide(diff, dist, out=np.zeros_like(diff), where=dist != 0)
    ct = e_r[:, :, 2].copy()
    ct[ct < -1] = -1
    ct[ct > 1] = 1
    st = np.sqrt(1 - ct ** 2)
    phi = np.arctan2(diff[:, :, 1], diff[:, :, 0])
    size_parameter = dist * k_medium[None, None, :]

    sph_h = np.zeros((2 * lmax + 1,) + size_parameter.shape, dtype=complex)
    sph_j = np.zeros_like(sph_h)
    nz = size_parameter != 0
    for p in range(2 * lmax + 1):
        sph_j[p] = spherical_jn(p, size_parameter)
        y = np.zeros_like(size_parameter, dtype=complex)
        y[nz] = spherical_yn(p, size_parameter[nz])
        sph_h[p] = sph_j[p] + 1j * y

    plm_full = legendre_normalized_trigon(2 * lmax, ct, st)
    plm = np.zeros(((2 * lmax + 1) * (2 * lmax + 2) // 2,) + ct.shape)
    for p in range(2 * lmax + 1):
        for absdm in range(p + 1):
            plm[p * (p + 1) // 2 + absdm] = plm_full[p, absdm]

    p_range = np.arange(-2 * lmax, 2 * lmax + 1)[:, None, None]
    e_j_dm_phi = np.exp(1j * p_range * phi[None, :, :])
    return sph_j, sph_h, e_j_dm_phi, plm


# ---------------------------------------------------------------- Wigner 3j


@lru_cache(maxsize=None)
def _fact(n: int) -> int:
    return math.factorial(n)


@lru_cache(maxsize=None)
def wigner3j_squared_signed(j1: int, j2: int, j3: int, m1: int, m2: int, m3: int):
    """(sign, value^2) of the Wigner 3j symbol, exact (Racah's formula)."""
    if m1 + m2 + m3 != 0:
        return 0, Fraction(0)
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0, Fraction(0)
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0, Fraction(0)
    delta = Fraction(_fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3),
                     _fact(j1 + j2 + j3 + 1))
    pref = (_fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2)
            * _fact(j3 + m3) * _fact(j3 - m3))
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    total = Fraction(0)
    for k in range(kmin, kmax + 1):
        den = (_fact(k) * _fact(j3 - j2 + k + m1) * _fact(j3 - j1 + k - m2)
               * _fact(j1 + j2 - j3 - k) * _fact(j1 - k - m1) * _fact(j2 - k + m2))
        total += Fraction((-1) ** k, den)
    if total == 0:
        return 0, Fraction(0)
    sign = (-1) ** ((j1 - j2 - m3) % 2) * (1 if total > 0 else -1)
    return sign, delta * pref * total * total


def wigner3j(j1, j2, j3, m1, m2, m3) -> float:
    sign, sq = wigner3j_squared_signed(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3))
    if sign == 0:
        return 0.0
    return sign * math.sqrt(sq)


# ---------------------------------------------------------------- ab5 table

_IPOW = (1.0 + 0.0j, 1.0j, -1.0 + 0.0j, -1.0j)


@lru_cache(maxsize=None)
def translation_table(lmax: int) -> np.ndarray:
    """``ab5[j1, j2, p]`` (yasfpy/numerics.py:224-272), complex128."""
    nmax = nmax_of(lmax)
    table = np.zeros((nmax, nmax, 2 * lmax + 1), dtype=complex)
    modes = []
    for tau in (1, 2):
        for l in range(1, lmax + 1):
            for m in range(-l, l + 1):
                modes.append((tau, l, m))
    for j1, (tau1, l1, m1) in enumerate(modes):
        for j2, (tau2, l2, m2) in enumerate(modes):
            for p in range(0, 2 * lmax + 1):
                phase = _IPOW[(abs(m1 - m2) - abs(m1) - abs(m2) + l2 - l1 + p) % 4]
                common = (phase * (-1.0) ** (m1 - m2)
                          * math.sqrt((2 * l1 + 1) * (2 * l2 + 1) / (2 * l1 * (l1 + 1) * l2 * (l2 + 1))))
                if tau1 == tau2:
                    w_a = wigner3j(l1, l2, p, m1, -m2, -m1 + m2)
                    w_b = wigner3j(l1, l2, p, 0, 0, 0)
                    table[j1, j2, p] = (common * (l1 * (l1 + 1) + l2 * (l2 + 1) - p * (p + 1))
                                        * math.sqrt(2 * p + 1) * w_a * w_b)
                elif p > 0:
                    rad = ((l1 + l2 + 1 + p) * (l1 + l2 + 1 - p) * (p + l1 - l2)
                           * (p - l1 + l2) * (2 * p + 1))
                    root = math.sqrt(rad) if rad >= 0 else 1j * math.sqrt(-rad)
                    w_a = wigner3j(l1, l2, p, m1, -m2, -m1 + m2)
                    w_b = wigner3j(l1, l2, p - 1, 0, 0, 0)
                    table[j1, j2, p] = common * root * w_a * w_b
    table.setflags(write=False)
    return table


# ---------------------------------------------------------------- Mie


def mie_coefficients(lmax: int, radius: np.ndarray, n_particle: np.ndarray,
                     wavelength: np.ndarray, n_medium: np.ndarray) -> np.ndarray:
    """Per-particle diagonal T entries ``[N, nmax, W]``
    (yasfpy/mie_coefficients.py:88-161; tau=1 -> -b_l, tau=2 -> -a_l)."""
    radius = np.atleast_1d(radius).astype(float)
    n_particle = np.atleast_1d(n_particle).astype(complex)
    wavelength = np.atleast_1d(wavelength).astype(float)
    omega = 2 * np.pi / wavelength
    k_medium = omega * np.atleast_1d(n_medium)
    nmax = nmax_of(lmax)
    out = np.zeros((radius.size, nmax, wavelength.size), dtype=complex)
    l_grid = np.arange(1, lmax + 1)[:, None]
    for s in range(radius.size):
        k_sphere = omega * n_particle[s]
        m = k_sphere / k_medium
        x = (k_medium * radius[s])[None, :]
        mx = (k_sphere * radius[s])[None, :]
        jx = spherical_jn(l_grid, x)
        jxp = spherical_jn(l_grid, x, derivative=True)
        yx = spherical_yn(l_grid, x)
        yxp = spherical_yn(l_grid, x, derivative=True)
        jmx = spherical_jn(l_grid, mx)
        jmxp = spherical_jn(l_grid, mx, derivative=True)
        hx = jx + 1j * yx
        hxp = jxp + 1j * yxp
        djx = jx + x * jxp
        djmx = jmx + mx * jmxp
        dhx = hx + x * hxp
        t1 = -(jmx * djx - jx * djmx) / (jmx * dhx - hx * djmx)
        t2 = -(m ** 2 * jmx * djx - jx * djmx) / (m ** 2 * jmx * dhx - hx * djmx)
        for li, l in enumerate(range(1, lmax + 1)):
            a = l * l - 1
            out[s, a:a + 2 * l + 1, :] = t1[li, :]
            b = lmax * (lmax + 2) + l * l - 1
            out[s, b:b + 2 * l + 1, :] = t2[li, :]
    return out


# ---------------------------------------------------------------- plane wave


def spherical_functions_trigon(lmax: int, ct, st, all_angles: bool = False):
    """pi_lm, tau_lm angular functions (yasfpy/functions/spherical_functions_trigon.py:39-93);
    ``[lmax+1, lmax+1]`` for the first angle, or ``[lmax+1, lmax+1, A]`` with all_angles."""
    ct = np.atleast_1d(ct).astype(float)
    st = np.atleast_1d(st).astype(float)
    shape = (lmax + 1, lmax + 1, ct.size)
    plm = np.zeros(shape)
    pilm = np.zeros(shape)
    taulm = np.zeros(shape)
    pp0 = np.zeros((lmax + 1, ct.size))
    plm[0, 0] = np.sqrt(1 / 2)
    plm[1, 0] = np.sqrt(3 / 2) * ct
    pp0[1] = np.sqrt(3) * plm[0, 0]
    taulm[0, 0] = -st * pp0[0]
    taulm[1, 0] = -st * pp0[1]
    for l in range(1, lmax):
        plm[l + 1, 0] = (1 / (l + 1) * np.sqrt((2 * l + 1) * (2 * l + 3)) * ct * plm[l, 0]
                         - l / (l + 1) * np.sqrt((2 * l + 3) / (2 * l - 1)) * plm[l - 1, 0])
        pp0[l + 1] = np.sqrt((2 * l + 3) / (2 * l + 1)) * ((l + 1) * plm[l, 0] + ct * pp0[l])
        taulm[l + 1, 0] = -st * pp0[l + 1]
    for m in range(1, lmax + 1):
        coeff = np.sqrt((2 * m + 1) / 2 / math.factorial(2 * m)) * np.prod(np.arange(1, 2 * m, 2))
        plm[m, m] = coeff * np.power(st, m)
        pilm[m, m] = coeff * np.power(st, m - 1)
        taulm[m, m] = m * ct * pilm[m, m]
        for l in range(m, lmax):
            c1 = np.sqrt((2 * l + 1) * (2 * l + 3) / (l + 1 - m) / (l + 1 + m)) * ct
            c2 = np.sqrt((2 * l + 3) * (l - m) * (l + m) / (2 * l - 1) / (l + 1 - m) / (l + 1 + m))
            plm[l + 1, m] = c1 * plm[l, m] - c2 * plm[l - 1, m]
            pilm[l + 1, m] = c1 * pilm[l, m] - c2 * pilm[l - 1, m]
            taulm[l + 1, m] = ((l + 1) * ct * pilm[l + 1, m]
                               - (l + 1 + m) * np.sqrt((2 * l + 3) * (l + 1 - m) / (2 * l + 1) / (l + 1 + m))
                               * pilm[l, m])
    if all_angles:
        return pilm, taulm
    return pilm[:, :, 0], taulm[:, :, 0]


def transformation_coefficients(pilm, taulm, tau, l, m, pol, dagger=False):
    """B_{tau l m}(beta) (yasfpy/functions/misc.py:70-114)."""
    ifac = -1j if dagger else 1j
    if pol in (1, 2):
        spher = taulm[l, abs(m)] if tau == pol else m * pilm[l, abs(m)]
        return -1 / ifac ** (l + 1) / np.sqrt(2 * l * (l + 1)) * (ifac * (pol == 1) + (pol == 2)) * spher
    return (-1 / ifac ** (l + 1) / np.sqrt(2 * l * (l + 1)) * (ifac + 1)
            * (taulm[l, abs(m)] + m * pilm[l, abs(m)]) / 2)


def planewave_coefficients(lmax: int, positions: np.ndarray, k_medium: np.ndarray,
                           polar: float, azimuthal: float, pol: int,
                           amplitude: float = 1.0, focal_point=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Incident plane-wave coefficients ``[N, nmax, W]``
    (yasfpy/initial_field_coefficients.py:99-142)."""
    k = np.atleast_1d(k_medium)
    cb, sb = np.cos(polar), np.sin(polar)
    pilm, taulm = spherical_functions_trigon(lmax, cb, sb)
    rel = positions - np.asarray(focal_point, dtype=float)
    kvec = np.outer(np.array((sb * np.cos(azimuthal), sb * np.sin(azimuthal), cb)), k)
    eikr = np.exp(1j * rel @ kvec)
    out = np.zeros((positions.shape[0], nmax_of(lmax), k.size), dtype=complex)
    for m in range(-lmax, lmax + 1):
        for tau in (1, 2):
            for l in range(max(1, abs(m)), lmax + 1):
                n = multi2single_index(0, tau, l, m, lmax)
                out[:, n, :] = (4 * amplitude * np.exp(-1j * m * azimuthal) * eikr
                                * transformation_coefficients(pilm, taulm, tau, l, m, pol, dagger=True))
    return out


# ---------------------------------------------------------------- operators


def coupling_matvec(lmax: int, positions: np.ndarray, k_medium: float, x: np.ndarray,
                    s1_lo: int = 0, n_s1: int | None = None,
                    near_field_radius: float | None = None) -> np.ndarray:
    """``(W x)`` rows of particles ``[s1_lo, s1_lo+n_s1)`` for one channel
    (yasfpy/coupling/cpu_dense.py:92-104 -> cpu_numba.py:21-95).

    near_field_radius: the near/far backend's approximation
    (yasfpy/coupling/backends.py:300-345, 402-457): only pairs with
    ``|r_s1 - r_s2|^2 <= radius^2`` couple; the far pairs' Hankel factors are
    zeroed, which drops exactly their terms from the sum."""
    from oracle import load_native

    lib = load_native()
    N = positions.shape[0]
    n_s1 = N - s1_lo if n_s1 is None else n_s1
    nmax = nmax_of(lmax)
    _, sph_h, ephi, plm = mutual_lookup(lmax, positions[s1_lo:s1_lo + n_s1], positions,
                                        np.array([k_medium]))
    if near_field_radius is not None:
        d = positions[s1_lo:s1_lo + n_s1, None, :] - positions[None, :, :]
        far = np.sum(d * d, axis=2) > float(near_field_radius) ** 2
        sph_h = np.where(far[None, :, :, None], 0.0, sph_h)
    idx = idx_lookup(lmax, N)
    ab5 = translation_table(lmax)
    x = np.ascontiguousarray(x, dtype=complex)
    out = np.zeros(n_s1 * nmax, dtype=complex)
    lib.matvec(lmax, N, s1_lo, n_s1, idx, x, np.ascontiguousarray(ab5),
               np.ascontiguousarray(plm), np.ascontiguousarray(sph_h[..., 0]),
               np.ascontiguousarray(ephi), out)
    return out


def master_matvec(lmax, positions, k_medium, t_flat, x):
    """``x - t * (W x)`` (yasfpy/matrix_apply.py:104-110)."""
    return x - t_flat * coupling_matvec(lmax, positions, k_medium, x)


# ---------------------------------------------------------------- solver


def gmres_solve(matvec, b, x0=None, rtol=1e-6, restart=100, maxiter=10000, psolve=None):
    """SciPy GMRES as called by the reference (yasfpy/solver.py:108-121):
    ``rtol=tol, atol=0``; x0 defaults to b (scattered_field_coefficients.py:131-136)."""
    from scipy.sparse.linalg import LinearOperator, gmres

    n = b.size
    A = LinearOperator((n, n), matvec=matvec, dtype=complex)
    M = None if psolve is None else LinearOperator((n, n), matvec=psolve, dtype=complex)
    counter = {"it": 0}

    def cb(_):
        counter["it"] += 1

    x, info = gmres(A, b, b.copy() if x0 is None else x0, M=M, restart=restart, rtol=rtol,
                    atol=0, maxiter=maxiter, callback=cb, callback_type="pr_norm")
    return x, info, counter["it"]


def make_particle_partition(positions: np.ndarray, edge_sizes) -> list[np.ndarray]:
    """Cuboid binning (yasfpy/preconditioner.py:222-265)."""
    edge_sizes = np.asarray(edge_sizes, dtype=float)
    lo = positions.min(axis=0) - 1.0
    hi = positions.max(axis=0) + 1.0
    bins = [np.digitize(positions[:, d], np.arange(lo[d], hi[d] + edge_sizes[d], edge_sizes[d]))
            for d in range(3)]
    groups: dict = {}
    for i in range(positions.shape[0]):
        groups.setdefault((bins[0][i], bins[1][i], bins[2][i]), []).append(i)
    return [np.array(v, dtype=int) for v in groups.values()]


def block_preconditioner(lmax, positions, k_medium, t_flat, partition):
    """Intended block-diagonal preconditioner (preconditioner.py:1-14, 92-219
    docstrings): per box, the inverse of ``I - T W`` restricted to intra-box
    pairs, particle-major indices, receiver T, float64, LU with P^T applied.
    Dense W_box is assembled column by column from the oracle matvec."""
    import scipy.linalg as sla

    nmax = nmax_of(lmax)
    factors = []
    for box in partition:
        m = box.size * nmax
        cols = np.arange(m)
        idx = (box[:, None] * nmax + np.arange(nmax)[None, :]).ravel()
        W = np.zeros((m, m), dtype=complex)
        e = np.zeros(m, dtype=complex)
        for c in cols:
            e[:] = 0
            e[c] = 1
            W[:, c] = coupling_matvec(lmax, positions[box], k_medium, e)
        M = np.eye(m) - t_flat[idx][:, None] * W
        factors.append((idx, sla.lu_factor(M)))

    def apply(r):
        out = np.zeros_like(r)
        for idx, lu in factors:
            out[idx] = sla.lu_solve(lu, r[idx])
        return out

    return apply


# ---------------------------------------------------------------- observables


def cross_sections(lmax, positions, k_medium, a_inc, f):
    """C_ext and C_sca for one channel (yasfpy/optics.py:62-66, 136-156 and
    cpu_numba.py:410-457): C_ext = -Re sum conj(a) f * pi/|k|^2,
    C_sca = |sum ...| * pi/|k|^2 with the J-translation (sph_j) double sum."""
    from oracle import load_native

    lib = load_native()
    N = positions.shape[0]
    sph_j, _, ephi, plm = mutual_lookup(lmax, positions, positions, np.array([k_medium]))
    idx = idx_lookup(lmax, N)
    out = np.zeros(1, dtype=complex)
    lib.cross_section(lmax, N, idx, np.ascontiguousarray(f.ravel(), dtype=complex),
                      np.ascontiguousarray(translation_table(lmax)), np.ascontiguousarray(plm),
                      np.ascontiguousarray(sph_j[..., 0]), np.ascontiguousarray(ephi), out)
    k2 = abs(k_medium) ** 2
    c_ext = float(np.real(-np.sum(np.conj(a_inc) * f)) / k2 * np.pi)
    c_sca = float(abs(out[0]) / k2 * np.pi)
    return c_ext, c_sca


def far_field(lmax, positions, k_medium, f, polar, azimuthal):
    """(E_theta, E_phi) at directions (polar[a], azimuthal[a]) for one channel:
    compute_electric_field_angle_components (yasfpy/functions/cpu_numba.py:535-606)
    with e_r of numerics.py:314-330; the sum over particles is vectorised."""
    polar = np.atleast_1d(np.asarray(polar, dtype=float))
    azimuthal = np.atleast_1d(np.asarray(azimuthal, dtype=float))
    N = positions.shape[0]
    nmax = nmax_of(lmax)
    f = np.asarray(f).reshape(N, nmax)
    pilm, taulm = spherical_functions_trigon(lmax, np.cos(polar), np.sin(polar), all_angles=True)
    e_r = np.stack([np.sin(polar) * np.cos(azimuthal), np.sin(polar) * np.sin(azimuthal),
                    np.cos(polar)], axis=1)
    phase = np.exp(-1j * k_medium * (positions @ e_r.T))  # [N, A]
    e_theta = np.zeros(polar.size, dtype=complex)
    e_phi = np.zeros(polar.size, dtype=complex)
    for tau in (1, 2):
        for l in range(1, lmax + 1):
            for m in range(-l, l + 1):
                n = multi2single_index(0, tau, l, m, lmax)
                t = (1j ** (tau - l - 2) * (f[:, n] @ phase) / np.sqrt(2 * l * (l + 1))
                     * np.exp(1j * m * azimuthal))
                if tau == 1:
                    e_theta += t * pilm[l, abs(m)] * 1j * m
                    e_phi -= t * taulm[l, abs(m)]
                else:
                    e_theta += t * taulm[l, abs(m)]
                    e_phi += t * pilm[l, abs(m)] * 1j * m
    return e_theta, e_phi


def polarization_components(e_theta, e_phi):
    """I, DOP, DOLP, DOLQ, DOLU, DOCP (yasfpy/functions/cpu_numba.py:609-671)."""
    et2 = np.abs(e_theta) ** 2
    ep2 = np.abs(e_phi) ** 2
    cross = e_theta * np.conj(e_phi)
    I = et2 + ep2
    Q = et2 - ep2
    U = -2 * cross.real
    V = 2 * cross.imag
    return (I, np.sqrt(Q ** 2 + U ** 2 + V ** 2) / I, np.sqrt(Q ** 2 + U ** 2) / I, -Q / I, U / I, V / I)


def phase_function(lmax, positions, k_medium, f, c_sca, sampling_points_number):
    """Azimuthally averaged phase function on the [n_phi, n_theta] grid of
    numerics.py:91-108 (optics.py:437-482): I 4 pi / |k|^2 / C_sca."""
    n_phi, n_theta = (int(v) for v in sampling_points_number)
    th = np.pi * np.linspace(0, 1, n_theta)
    ph = 2 * np.pi * np.linspace(0, 1, n_phi + 1)[:-1]
    pa, aa = np.meshgrid(th, ph, indexing="xy")
    et, ep = far_field(lmax, positions, k_medium, f, pa.ravel(), aa.ravel())
    I = polarization_components(et, ep)[0]
    p3 = I * 4 * np.pi / abs(k_medium) ** 2 / c_sca
    return th, p3.reshape(n_phi, n_theta).mean(axis=0)
