"""Pins of the CPU oracle against the reference's own fixtures, library routines
and closed forms (no GPU).  Each test names what it pins."""

import math

import numpy as np
import numpy.testing as npt
import pytest
from scipy.special import lpmv, spherical_jn, spherical_yn

from oracle import restate as R


def test_multi2single_index(golden):
    d = golden("index")  # MATLAB 1-based (tests/test_comparison.py:59-67)
    for (s, tau, l, m), expect in zip(d["inputs"], d["outputs"]):
        assert R.multi2single_index(s - 1, tau, l, m, int(d["lmax"])) == expect - 1
    idx = R.idx_lookup(3, 2)
    for j, (s, n, tau, l, m) in enumerate(idx):
        assert R.multi2single_index(s, tau, l, m, 3) == j


def test_legendre_fixture(golden):
    d = golden("legendre")
    theta = d["theta"]
    plm = R.legendre_normalized_trigon(int(d["lmax"]), np.cos(theta), np.sin(theta))
    npt.assert_allclose(plm, d["plm"], rtol=1e-5, atol=1e-12)


def test_legendre_closed_form():
    """Normalised P_l^m without Condon-Shortley phase = (-1)^m norm * scipy lpmv."""
    th = np.linspace(0.05, 3.1, 17)
    ct, st = np.cos(th), np.sin(th)
    plm = R.legendre_normalized_trigon(16, ct, st)
    for l in range(17):
        for m in range(l + 1):
            norm = math.sqrt((2 * l + 1) / 2 * math.factorial(l - m) / math.factorial(l + m))
            npt.assert_allclose(plm[l, m], (-1) ** m * norm * lpmv(m, l, ct), rtol=1e-10, atol=1e-13)


def test_hankel_tables(golden):
    d = golden("sph_bessel")
    npt.assert_allclose(spherical_jn(int(d["l"]), d["z"]), d["j"], rtol=1e-5)
    npt.assert_allclose(spherical_jn(int(d["l"]), d["z"]) + 1j * spherical_yn(int(d["l"]), d["z"]),
                        d["h"], rtol=1e-5)
    # closed forms h_0 = -i e^{ix}/x, h_1 = -(x + i) e^{ix}/x^2
    pos = np.array([[0.0, 0.0, 0.0], [30.0, -40.0, 120.0]])
    k = 0.011
    _, sph_h, _, _ = R.mutual_lookup(2, pos, pos, np.array([k]))
    x = k * np.linalg.norm(pos[0] - pos[1])
    npt.assert_allclose(sph_h[0, 0, 1, 0], -1j * np.exp(1j * x) / x, rtol=1e-14)
    npt.assert_allclose(sph_h[1, 0, 1, 0], -(x + 1j) * np.exp(1j * x) / x ** 2, rtol=1e-14)


def test_wigner3j_against_sympy():
    from sympy import Rational
    from sympy.physics.wigner import wigner_3j

    rng = np.random.default_rng(3)
    for _ in range(200):
        j1, j2 = rng.integers(0, 8, 2)
        j3 = rng.integers(abs(j1 - j2), j1 + j2 + 1)
        m1 = rng.integers(-j1, j1 + 1)
        m2 = rng.integers(-j2, j2 + 1)
        m3 = -m1 - m2
        if abs(m3) > j3:
            continue
        exact = float(wigner_3j(int(j1), int(j2), int(j3), int(m1), int(m2), int(m3)))
        assert R.wigner3j(j1, j2, j3, m1, m2, m3) == pytest.approx(exact, rel=1e-14, abs=1e-16)
    assert R.wigner3j(1, 1, 1, 0, 0, 0) == 0.0  # odd sum with zero m vanishes
    del Rational


def test_wigner3j_orthogonality():
    """sum_{m1} (2 j3 + 1) (j1 j2 j3; m1 m2 m3)(j1 j2 j3'; m1 m2 m3) = delta_{j3 j3'} at fixed m3."""
    for j1, j2 in [(2, 3), (4, 4), (5, 2)]:
        for m3 in (-1, 0, 2):
            js = [j for j in range(abs(j1 - j2), j1 + j2 + 1) if abs(m3) <= j]
            for j3 in js:
                for j3b in js:
                    s = sum((2 * j3 + 1) * R.wigner3j(j1, j2, j3, m1, -m1 - m3, m3)
                            * R.wigner3j(j1, j2, j3b, m1, -m1 - m3, m3)
                            for m1 in range(-j1, j1 + 1) if abs(m1 + m3) <= j2)
                    assert s == pytest.approx(1.0 if j3 == j3b else 0.0, abs=1e-13)


def test_translation_table_fixture(golden):
    d = golden("ab5_lmax3")
    npt.assert_allclose(R.translation_table(3), d["ab5"], rtol=1e-5, atol=1e-7)


def test_mie_t_entry(golden):
    d = golden("t_entry")  # t_entry(tau, l, kM, kS, R) (tests/test_comparison.py:92-110)
    lam = 2 * np.pi / d["k_medium"]
    mie = R.mie_coefficients(1, np.array([d["radius"]]), np.array([d["k_sphere"] / d["k_medium"]]),
                             np.array([lam]), 1.0)
    l = int(d["l"])
    npt.assert_allclose(mie[0, l * l - 1, 0], d["scattered_1"], rtol=1e-5)
    npt.assert_allclose(mie[0, 3 + l * l - 1, 0], d["scattered_2"], rtol=1e-5)


def test_mie_series_single_sphere():
    """Q_ext, Q_sca of one sphere from the textbook Mie series vs -Re sum(conj(a) f)."""
    lmax, r, n, lam = 10, 150.0, 1.5 + 0.02j, 600.0
    k = 2 * np.pi / lam
    mie = R.mie_coefficients(lmax, np.array([r]), np.array([n]), np.array([lam]), 1.0)[0, :, 0]
    a = R.planewave_coefficients(lmax, np.zeros((1, 3)), np.array([k]), 0.0, 0.0, 1)[0, :, 0]
    f = mie * a
    c_ext = float(np.real(-np.sum(np.conj(a) * f))) / k ** 2 * np.pi
    x = k * r
    qs = 0.0
    for l in range(1, lmax + 1):
        b_l = -mie[l * l - 1]
        a_l = -mie[lmax * (lmax + 2) + l * l - 1]
        qs += (2 * l + 1) * np.real(a_l + b_l)
    q_ext_series = 2 / x ** 2 * qs
    assert c_ext / (np.pi * r ** 2) == pytest.approx(q_ext_series, rel=1e-10)


def test_planewave_fixtures(golden):
    d = golden("full_simulation")
    k = 2 * np.pi / float(d["wavelength"]) * float(d["medium_refractive_index"])
    a = R.planewave_coefficients(int(d["lmax"]), d["positions"], np.array([k]), 0.0, 0.0, 1)[:, :, 0]
    npt.assert_allclose(a, d["initial_field_coefficients"], rtol=1e-4, atol=1e-6)
    c = golden("celes500")
    k = (2 * np.pi / c["wavelength"] * c["medium_refractive_index"]).real
    a = R.planewave_coefficients(int(c["lmax"]), c["positions"], k, 0.0, 0.0, 1)[:, :, 0]
    npt.assert_allclose(a, c["initial_field_coefficients"], rtol=2e-3)


def test_mie_fixture(golden):
    d = golden("full_simulation")
    mie = R.mie_coefficients(int(d["lmax"]), d["radii"], d["refractive_index"],
                             np.array([float(d["wavelength"])]), float(d["medium_refractive_index"]))
    npt.assert_allclose(mie[:, :, 0], d["mie_coefficients"], rtol=1e-4, atol=1e-6)


def test_coupling_matvec_fixture(golden):
    d = golden("coupling_matvec")
    k = 2 * np.pi / float(d["wavelength"]) * float(d["medium_refractive_index"])
    wx = R.coupling_matvec(int(d["lmax"]), d["positions"], k, d["x"])
    npt.assert_allclose(wx, d["wx"], rtol=2e-4, atol=1e-5)


def _dense_w(lmax, pos, k):
    n = pos.shape[0] * R.nmax_of(lmax)
    W = np.zeros((n, n), dtype=complex)
    for c in range(n):
        e = np.zeros(n, dtype=complex)
        e[c] = 1.0
        W[:, c] = R.coupling_matvec(lmax, pos, k, e)
    return W


def test_coupling_reciprocity():
    """W[j1, j2] = W[P j2, P j1] with P: m -> -m (reciprocity of the translation)."""
    lmax = 3
    pos = np.random.default_rng(1).normal(size=(3, 3)) * 250.0
    W = _dense_w(lmax, pos, 2 * np.pi / 480.0)
    idx = R.idx_lookup(lmax, 3)
    perm = np.array([R.multi2single_index(s, t, l, -m, lmax) for s, _, t, l, m in idx])
    npt.assert_allclose(W.T[perm][:, perm], W, rtol=0, atol=1e-14 * np.abs(W).max())


def test_coupling_axis_selection_rule():
    """Spheres on the z axis couple only equal m (P_p^{|dm|}(+-1) = 0 for dm != 0)."""
    lmax = 3
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 350.0]])
    W = _dense_w(lmax, pos, 2 * np.pi / 550.0)
    m = R.idx_lookup(lmax, 2)[:, 4]
    assert np.abs(W[m[:, None] != m[None, :]]).max() == 0.0
    assert np.abs(W[m[:, None] == m[None, :]]).max() > 1e-3


def test_near_field_cutoff():
    """The near/far restriction (backends.py:300-345): the reference's own
    cases -- a radius beyond every pair equals the exact W (1e-12), radius 0
    gives W = 0 -- and, for a radius between the pair distances, the sum of
    the exact two-sphere operators of the near pairs alone."""
    lmax = 2
    pos = np.array([[0.0, 0.0, 0.0], [400.0, 0.0, 0.0], [100.0, 200.0, 0.0], [900.0, 300.0, 50.0]])
    k = 2 * np.pi / 520.0
    N, nmax = pos.shape[0], R.nmax_of(lmax)
    x = np.random.default_rng(5).normal(size=N * nmax) + 1j * np.random.default_rng(6).normal(size=N * nmax)
    exact = R.coupling_matvec(lmax, pos, k, x)
    npt.assert_allclose(R.coupling_matvec(lmax, pos, k, x, near_field_radius=1e4), exact, rtol=1e-12, atol=0)
    assert np.all(R.coupling_matvec(lmax, pos, k, x, near_field_radius=0.0) == 0)
    radius = 450.0
    ref = np.zeros_like(exact)
    xs = x.reshape(N, nmax)
    for a in range(N):
        for b in range(a + 1, N):
            if np.linalg.norm(pos[a] - pos[b]) > radius:
                continue
            y = R.coupling_matvec(lmax, pos[[a, b]], k, np.concatenate([xs[a], xs[b]])).reshape(2, nmax)
            ref.reshape(N, nmax)[a] += y[0]
            ref.reshape(N, nmax)[b] += y[1]
    got = R.coupling_matvec(lmax, pos, k, x, near_field_radius=radius)
    npt.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    assert np.linalg.norm(got - exact) > 1e-3 * np.linalg.norm(exact)  # some pairs were dropped


def test_full_simulation_solve(golden):
    d = golden("full_simulation")
    L = int(d["lmax"])
    lam = float(d["wavelength"])
    k = 2 * np.pi / lam * float(d["medium_refractive_index"])
    mie = R.mie_coefficients(L, d["radii"], d["refractive_index"], np.array([lam]),
                             float(d["medium_refractive_index"]))[:, :, 0]
    a = R.planewave_coefficients(L, d["positions"], np.array([k]), 0.0, 0.0, 1)[:, :, 0]
    t = mie.ravel()
    x, info, _ = R.gmres_solve(lambda v: R.master_matvec(L, d["positions"], k, t, v), t * a.ravel(),
                               rtol=1e-4, restart=1000, maxiter=1000)
    assert info == 0
    npt.assert_allclose(x.reshape(a.shape), d["scattered_field_coefficients"], rtol=1e-3, atol=1e-5)


def test_block_preconditioner_single_box_is_exact():
    """One box holding every particle: M^-1 is the exact inverse (1 iteration)."""
    L = 2
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(4, 3)) * 300.0
    lam = 600.0
    k = 2 * np.pi / lam
    mie = R.mie_coefficients(L, np.full(4, 90.0), np.full(4, 1.5 + 0.01j), np.array([lam]), 1.0)[:, :, 0]
    t = mie.ravel()
    apply = R.block_preconditioner(L, pos, k, t, [np.arange(4)])
    b = rng.normal(size=t.size) + 1j * rng.normal(size=t.size)
    z = apply(b)
    npt.assert_allclose(R.master_matvec(L, pos, k, t, z), b, rtol=1e-12, atol=1e-12)


def _bh_mie(lmax, x, m):
    """Textbook Mie a_n, b_n (Bohren & Huffman eq. 4.53), independent of the oracle."""
    n = np.arange(1, lmax + 1)
    def psi(z):
        return z * spherical_jn(n, z)
    def dpsi(z):
        return spherical_jn(n, z) + z * spherical_jn(n, z, derivative=True)
    h = spherical_jn(n, x) + 1j * spherical_yn(n, x)
    dh = (spherical_jn(n, x, derivative=True) + 1j * spherical_yn(n, x, derivative=True))
    xi, dxi = x * h, h + x * dh
    mx = m * x
    a = (m * psi(mx) * dpsi(x) - psi(x) * dpsi(mx)) / (m * psi(mx) * dxi - xi * dpsi(mx))
    b = (psi(mx) * dpsi(x) - m * psi(x) * dpsi(mx)) / (psi(mx) * dxi - m * xi * dpsi(mx))
    return a, b


def test_phase_function_single_sphere_textbook():
    """Far field (cpu_numba.py:535-606) + C_sca (cpu_numba.py:410-457) of one sphere
    reproduce the textbook Mie phase function (|S1|^2 + |S2|^2) / sum (2n+1)(|a|^2+|b|^2),
    S1, S2 from the pi_n, tau_n recurrences (Bohren & Huffman 4.47)."""
    lmax, r, m, lam = 8, 180.0, 1.6 + 0.05j, 550.0
    k = 2 * np.pi / lam
    pos = np.zeros((1, 3))
    mie = R.mie_coefficients(lmax, np.array([r]), np.array([m]), np.array([lam]), 1.0)[0, :, 0]
    a_inc = R.planewave_coefficients(lmax, pos, np.array([k]), 0.0, 0.0, 1)[0, :, 0]
    f = mie * a_inc
    _, c_sca = R.cross_sections(lmax, pos, k, a_inc, f)
    th, p = R.phase_function(lmax, pos, k, f, c_sca, (8, 37))
    an, bn = _bh_mie(lmax, k * r, m)
    mu = np.cos(th)
    pi_prev, pi_cur = np.zeros_like(mu), np.ones_like(mu)
    s1 = np.zeros_like(mu, dtype=complex)
    s2 = np.zeros_like(mu, dtype=complex)
    for nn in range(1, lmax + 1):
        if nn > 1:
            pi_prev, pi_cur = pi_cur, ((2 * nn - 1) * mu * pi_cur - nn * pi_prev) / (nn - 1)
        tau = nn * mu * pi_cur - (nn + 1) * pi_prev
        c = (2 * nn + 1) / (nn * (nn + 1))
        s1 += c * (an[nn - 1] * pi_cur + bn[nn - 1] * tau)
        s2 += c * (an[nn - 1] * tau + bn[nn - 1] * pi_cur)
    norm = np.sum((2 * np.arange(1, lmax + 1) + 1) * (np.abs(an) ** 2 + np.abs(bn) ** 2))
    npt.assert_allclose(p, (np.abs(s1) ** 2 + np.abs(s2) ** 2) / norm, rtol=1e-9)
    # C_sca of one sphere equals the textbook 2 pi / k^2 sum (2n+1)(|a|^2 + |b|^2)
    assert c_sca == pytest.approx(2 * np.pi / k ** 2 * norm, rel=1e-10)


def _solved_cluster(n_index, seed=21, N=7, lmax=3, lam=500.0):
    from yasf_amd.clusters import cluster
    pos, radii = cluster(seed, N, 0.1)
    k = 2 * np.pi / lam
    mie = R.mie_coefficients(lmax, radii, np.full(N, n_index), np.array([lam]), 1.0)[:, :, 0]
    t = mie.ravel()
    a_inc = R.planewave_coefficients(lmax, pos, np.array([k]), 0.3, 1.1, 1)[:, :, 0].ravel()
    f, info, _ = R.gmres_solve(lambda v: R.master_matvec(lmax, pos, k, t, v), t * a_inc, rtol=1e-12,
                               restart=200)
    assert info == 0
    return pos, k, lmax, a_inc, f


def test_cross_section_equals_far_field_integral():
    """C_sca of a 7-sphere cluster (J-translation pair sum, self terms included) equals
    the far-field integral 1/k^2 int |E|^2 dOmega (Gauss-Legendre in cos theta,
    uniform in phi; exact for this band limit): int p dOmega = 4 pi (test_compare_mstm4.py:37-47)."""
    pos, k, lmax, a_inc, f = _solved_cluster(1.5 + 0.02j)
    _, c_sca = R.cross_sections(lmax, pos, k, a_inc, f)
    x, wx = np.polynomial.legendre.leggauss(48)
    nphi = 96
    ph = 2 * np.pi * np.arange(nphi) / nphi
    TH, PH = np.meshgrid(np.arccos(x), ph, indexing="ij")
    et, ep = R.far_field(lmax, pos, k, f, TH.ravel(), PH.ravel())
    I = (np.abs(et) ** 2 + np.abs(ep) ** 2).reshape(TH.shape)
    integral = np.sum(wx[:, None] * I) * (2 * np.pi / nphi)
    assert integral / k ** 2 == pytest.approx(c_sca, rel=1e-9)


def test_lossless_cluster_conserves_energy():
    """Real refractive index: C_ext = C_sca (optical theorem) to solver tolerance."""
    pos, k, lmax, a_inc, f = _solved_cluster(1.45 + 0.0j, seed=22)
    c_ext, c_sca = R.cross_sections(lmax, pos, k, a_inc, f)
    assert c_ext == pytest.approx(c_sca, rel=1e-9)


def test_spherical_functions_trigon_fixture(golden):
    """pi_lm, tau_lm (spherical_functions_trigon.py:39-93) against the
    reference's MATLAB fixture (tests/test_comparison.py:79-89, rtol 1e-5)."""
    d = golden("spherical_functions_trigon")
    lmax = int(d["lmax_sft"][0, 0])
    theta = d["theta_sft"].ravel()
    pilm, taulm = R.spherical_functions_trigon(lmax, np.cos(theta), np.sin(theta), all_angles=True)
    npt.assert_allclose(pilm, d["pilm_sft"], rtol=1e-5, atol=1e-12)
    npt.assert_allclose(taulm, d["taulm_sft"], rtol=1e-5, atol=1e-12)
    assert np.count_nonzero(d["pilm_sft"]) > 0 and np.count_nonzero(d["taulm_sft"]) > 0


@pytest.mark.parametrize("tau_key, pol_key, dagger, expected_key", [
    ("tau_tc_1", "pol_tc_1", False, "B_11"), ("tau_tc_1", "pol_tc_2", False, "B_12"),
    ("tau_tc_2", "pol_tc_1", False, "B_21"), ("tau_tc_2", "pol_tc_2", False, "B_22"),
    ("tau_tc_1", "pol_tc_1", True, "Bd_11"), ("tau_tc_1", "pol_tc_2", True, "Bd_12"),
    ("tau_tc_2", "pol_tc_1", True, "Bd_21"), ("tau_tc_2", "pol_tc_2", True, "Bd_22")])
def test_transformation_coefficients_fixture(golden, tau_key, pol_key, dagger, expected_key):
    """B_{tau l m}(beta) and its dagger (misc.py:70-114) against the MATLAB
    fixture (tests/test_comparison.py:157-185, rtol 1e-5)."""
    d = golden("transformation_coefficients")
    lmax = int(d["lmax_tc"][0, 0])
    theta = float(d["theta_tc"][0, 0])
    pilm, taulm = R.spherical_functions_trigon(lmax, np.cos(theta), np.sin(theta))
    out = R.transformation_coefficients(pilm, taulm, int(d[tau_key][0, 0]), int(d["l_tc"][0, 0]),
                                        int(d["m_tc"][0, 0]), int(d[pol_key][0, 0]), dagger=dagger)
    npt.assert_allclose(out, np.squeeze(d[expected_key]), rtol=1e-5)


def test_polarization_degrees_single_sphere_textbook():
    """Degrees of polarisation (cpu_numba.py:609-671) of one sphere's far field
    against the textbook Stokes parameters of the scattered light built from
    S1, S2 (Bohren & Huffman 3.16, 4.74-4.77): for incident light linearly
    polarised at angle psi to the scattering plane,
      I ~ S11 + S12 cos 2psi,  Q ~ S12 + S11 cos 2psi,  U ~ S33 sin 2psi,
      V ~ S34 sin 2psi (up to sign),
    with S11 = (|S2|^2 + |S1|^2)/2, S12 = (|S2|^2 - |S1|^2)/2, S33 = Re(S2 S1*),
    S34 = Im(S2 S1*).  The TE plane wave along +z is polarised along y, so
    psi = phi - pi/2.  A single coherent field is fully polarised (DOP = 1).
    The product's host function is held to the same values."""
    from yasf_amd.optics import polarization_components as product_components

    lmax, r, m, lam = 8, 180.0, 1.6 + 0.05j, 550.0
    k = 2 * np.pi / lam
    pos = np.zeros((1, 3))
    mie = R.mie_coefficients(lmax, np.array([r]), np.array([m]), np.array([lam]), 1.0)[0, :, 0]
    a_inc = R.planewave_coefficients(lmax, pos, np.array([k]), 0.0, 0.0, 1)[0, :, 0]
    th = np.linspace(0.1, 3.0, 13)
    ph = np.linspace(0.05, 2 * np.pi - 0.05, 11)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    et, ep = R.far_field(lmax, pos, k, mie * a_inc, TH.ravel(), PH.ravel())
    an, bn = _bh_mie(lmax, k * r, m)
    mu = np.cos(TH.ravel())
    pi_prev, pi_cur = np.zeros_like(mu), np.ones_like(mu)
    s1 = np.zeros_like(mu, dtype=complex)
    s2 = np.zeros_like(mu, dtype=complex)
    for nn in range(1, lmax + 1):
        if nn > 1:
            pi_prev, pi_cur = pi_cur, ((2 * nn - 1) * mu * pi_cur - nn * pi_prev) / (nn - 1)
        tau = nn * mu * pi_cur - (nn + 1) * pi_prev
        c = (2 * nn + 1) / (nn * (nn + 1))
        s1 += c * (an[nn - 1] * pi_cur + bn[nn - 1] * tau)
        s2 += c * (an[nn - 1] * tau + bn[nn - 1] * pi_cur)
    s11 = (np.abs(s2) ** 2 + np.abs(s1) ** 2) / 2
    s12 = (np.abs(s2) ** 2 - np.abs(s1) ** 2) / 2
    s33 = np.real(s2 * np.conj(s1))
    s34 = np.imag(s2 * np.conj(s1))
    psi = PH.ravel() - np.pi / 2
    c2, sn2 = np.cos(2 * psi), np.sin(2 * psi)
    i_t = s11 + s12 * c2
    q_t, u_t, v_t = s12 + s11 * c2, s33 * sn2, s34 * sn2
    for comp in (R.polarization_components, product_components):
        I, dop, dolp, dolq, dolu, docp = comp(et, ep)
        npt.assert_allclose(I / I.max(), i_t / i_t.max(), rtol=1e-9, atol=1e-12)
        npt.assert_allclose(dop, 1.0, rtol=1e-12)
        npt.assert_allclose(dolp, np.sqrt(q_t ** 2 + u_t ** 2) / i_t, rtol=1e-9, atol=1e-12)
        npt.assert_allclose(dolq, -q_t / i_t, rtol=1e-9, atol=1e-12)
        npt.assert_allclose(np.abs(dolu), np.abs(u_t) / i_t, rtol=1e-9, atol=1e-12)
        npt.assert_allclose(np.abs(docp), np.abs(v_t) / i_t, rtol=1e-9, atol=1e-12)
