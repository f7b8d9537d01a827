"""FP64 host set-up of the linear system: Mie T-matrix entries and plane-wave
incident coefficients.

* ``mie_coefficients`` follows yasfpy/mie_coefficients.py:88-161 (tau=1 -> -b_l,
  tau=2 -> -a_l, replicated over m), evaluated per unique (radius, index) pair.
* ``planewave_coefficients`` follows yasfpy/initial_field_coefficients.py:99-142
  with the angular functions of spherical_functions_trigon.py:39-93 and the
  transformation coefficients of misc.py:70-114 (dagger form).
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import spherical_jn, spherical_yn


def nmax_of(lmax: int) -> int:
    return 2 * lmax * (lmax + 2)


def mode_index(tau: int, l: int, m: int, lmax: int) -> int:
    """n = (tau-1) L(L+2) + (l-1)(l+1) + l + m (misc.py:23-43)."""
    return (tau - 1) * lmax * (lmax + 2) + (l - 1) * (l + 1) + l + m


def mie_coefficients(lmax: int, unique_pairs: np.ndarray, omega: np.ndarray,
                     k_medium: np.ndarray) -> np.ndarray:
    """``[U, nmax, W]`` T-matrix diagonal for unique (radius, index) pairs
    (vectorised over pairs, orders and wavelengths)."""
    U = unique_pairs.shape[0]
    W = np.size(omega)
    out = np.zeros((U, nmax_of(lmax), W), dtype=complex)
    if U == 0:
        return out
    ls = np.arange(1, lmax + 1)[:, None, None]                          # [L, 1, 1]
    radius = np.real(unique_pairs[:, 0]).astype(float)[None, :, None]   # [1, U, 1]
    k_sphere = np.asarray(omega)[None, None, :] * unique_pairs[:, 1][None, :, None]
    km = np.asarray(k_medium)[None, None, :]
    rel = k_sphere / km
    x = km * radius
    mx = k_sphere * radius
    jx, djx_ = spherical_jn(ls, x), spherical_jn(ls, x, derivative=True)
    yx, dyx_ = spherical_yn(ls, x), spherical_yn(ls, x, derivative=True)
    jmx, djmx_ = spherical_jn(ls, mx), spherical_jn(ls, mx, derivative=True)
    hx = jx + 1j * yx
    dhx = hx + x * (djx_ + 1j * dyx_)          # d/dx [x h_l(x)]
    djx = jx + x * djx_                        # d/dx [x j_l(x)]
    djmx = jmx + mx * djmx_
    b_l = -(jmx * djx - jx * djmx) / (jmx * dhx - hx * djmx)                                  # [L, U, W]
    a_l = -(rel ** 2 * jmx * djx - jx * djmx) / (rel ** 2 * jmx * dhx - hx * djmx)
    for li, l in enumerate(range(1, lmax + 1)):
        s1 = l * l - 1
        out[:, s1:s1 + 2 * l + 1, :] = b_l[li][:, None, :]
        s2 = lmax * (lmax + 2) + l * l - 1
        out[:, s2:s2 + 2 * l + 1, :] = a_l[li][:, None, :]
    return out


def angular_functions(lmax: int, ct: float, st: float):
    """pi_lm(theta), tau_lm(theta), normalised, ``[lmax+1, lmax+1]``."""
    pilm = np.zeros((lmax + 1, lmax + 1))
    taulm = np.zeros((lmax + 1, lmax + 1))
    plm = np.zeros((lmax + 1, lmax + 1))
    dp0 = np.zeros(lmax + 1)
    plm[0, 0] = math.sqrt(0.5)
    plm[1, 0] = math.sqrt(1.5) * ct
    dp0[1] = math.sqrt(3.0) * plm[0, 0]
    taulm[1, 0] = -st * dp0[1]
    for l in range(1, lmax):
        plm[l + 1, 0] = (math.sqrt((2 * l + 1) * (2 * l + 3)) / (l + 1) * ct * plm[l, 0]
                         - l / (l + 1) * math.sqrt((2 * l + 3) / (2 * l - 1)) * plm[l - 1, 0])
        dp0[l + 1] = math.sqrt((2 * l + 3) / (2 * l + 1)) * ((l + 1) * plm[l, 0] + ct * dp0[l])
        taulm[l + 1, 0] = -st * dp0[l + 1]
    for m in range(1, lmax + 1):
        c = math.sqrt((2 * m + 1) / 2 / math.factorial(2 * m)) * math.prod(range(1, 2 * m, 2))
        plm[m, m] = c * st ** m
        pilm[m, m] = c * st ** (m - 1)
        taulm[m, m] = m * ct * pilm[m, m]
        for l in range(m, lmax):
            a = math.sqrt((2 * l + 1) * (2 * l + 3) / ((l + 1 - m) * (l + 1 + m))) * ct
            b = math.sqrt((2 * l + 3) * (l - m) * (l + m) / ((2 * l - 1) * (l + 1 - m) * (l + 1 + m)))
            plm[l + 1, m] = a * plm[l, m] - b * plm[l - 1, m]
            pilm[l + 1, m] = a * pilm[l, m] - b * pilm[l - 1, m]
            taulm[l + 1, m] = ((l + 1) * ct * pilm[l + 1, m] - (l + 1 + m)
                               * math.sqrt((2 * l + 3) * (l + 1 - m) / ((2 * l + 1) * (l + 1 + m)))
                               * pilm[l, m])
    return pilm, taulm


def _b_dagger(pilm, taulm, tau, l, m, pol):
    ifac = -1j
    if pol in (1, 2):
        f = taulm[l, abs(m)] if tau == pol else m * pilm[l, abs(m)]
        return -1 / ifac ** (l + 1) / math.sqrt(2 * l * (l + 1)) * (ifac * (pol == 1) + (pol == 2)) * f
    return (-1 / ifac ** (l + 1) / math.sqrt(2 * l * (l + 1)) * (ifac + 1)
            * (taulm[l, abs(m)] + m * pilm[l, abs(m)]) / 2)


def planewave_coefficients(lmax: int, positions: np.ndarray, k_medium: np.ndarray, polar: float,
                           azimuthal: float, pol: int, amplitude: float = 1.0,
                           focal_point=(0.0, 0.0, 0.0)) -> np.ndarray:
    """``[N, nmax, W]`` incident coefficients of a plane wave."""
    k = np.atleast_1d(k_medium)
    cb, sb = math.cos(polar), math.sin(polar)
    pilm, taulm = angular_functions(lmax, cb, sb)
    rel = positions - np.asarray(focal_point, dtype=float)
    direction = np.array((sb * math.cos(azimuthal), sb * math.sin(azimuthal), cb))
    eikr = np.exp(1j * np.outer(rel @ direction, k))
    out = np.zeros((positions.shape[0], nmax_of(lmax), k.size), dtype=complex)
    for m in range(-lmax, lmax + 1):
        ph = 4 * amplitude * np.exp(-1j * m * azimuthal)
        for tau in (1, 2):
            for l in range(max(1, abs(m)), lmax + 1):
                out[:, mode_index(tau, l, m, lmax), :] = ph * eikr * _b_dagger(pilm, taulm, tau, l, m, pol)
    return out
