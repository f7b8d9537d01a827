"""GPU parity of the observables (C_ext, C_sca, far field, phase function)
through the C ABI and the Optics mirror, against the oracle's restatements of
optics.py:55-156, 269-551 and cpu_numba.py:410-457, 535-671.

Bar: FP64 device results within 1e-10 relative of the oracle on the same
coefficients; the mixed-precision solve's cross sections within 1e-6 of the
all-FP64 solve's."""

import numpy as np
import numpy.testing as npt
import pytest

from oracle import restate as R
from yasf_amd import InitialField, Numerics, Parameters, Particles, Simulation, Solver
from yasf_amd import _native
from yasf_amd.clusters import cluster
from yasf_amd.optics import Optics

pytestmark = pytest.mark.gpu


def _coefficients(seed, N, lmax, lam, n_index):
    pos, radii = cluster(seed, N, 0.1)
    k = 2 * np.pi / lam
    t = R.mie_coefficients(lmax, radii, np.full(N, n_index), np.array([lam]), 1.0)[:, :, 0].ravel()
    a = R.planewave_coefficients(lmax, pos, np.array([k]), 0.4, 2.0, 2)[:, :, 0].ravel()
    rng = np.random.default_rng(seed)
    # arbitrary (unsolved) scattered coefficients exercise every term
    f = t * a + 0.1 * (rng.normal(size=t.size) + 1j * rng.normal(size=t.size)) * np.abs(t * a).max()
    return pos, k, t, a, f


METHODS = [(_native.CSCA_FAR_FIELD, "far_field"), (_native.CSCA_PAIR_SUM, "pair_sum")]


@pytest.mark.parametrize("method", [m for m, _ in METHODS], ids=[i for _, i in METHODS])
@pytest.mark.parametrize("lmax", [1, 3, 4, 8])
def test_cross_sections_match_oracle(lmax, method):
    # C_sca by far-field quadrature (default) and by the reference's regular
    # pair sum: both within 1e-10 of the oracle's pair sum
    N = 21 if lmax < 8 else 9
    pos, k, t, a, f = _coefficients(30 + lmax, N, lmax, 520.0, 1.5 + 0.02j)
    s = _native.Session(pos, lmax, [k])
    s.set_scattering_method(method)
    c_ext, c_sca = s.cross_sections(a, f)
    r_ext, r_sca = R.cross_sections(lmax, pos, k, a, f)
    assert c_ext == pytest.approx(r_ext, rel=1e-10)
    assert c_sca == pytest.approx(r_sca, rel=1e-10)


@pytest.mark.parametrize("mode,slab_tiles", [(_native.MATVEC_ON_THE_FLY, 0), (_native.MATVEC_TABLE, 0),
                                              (_native.MATVEC_SLAB, 1), (_native.MATVEC_SLAB, 2),
                                              (_native.MATVEC_AUTO, 0)],
                         ids=["on_the_fly", "table", "slab1", "slab2", "auto"])
def test_cross_sections_every_matvec_mode(mode, slab_tiles):
    # C_sca's regular pair GEMM runs over slabs of the regular table (SLAB)
    # or on the fly (TABLE, ON_THE_FLY, and AUTO with its table resident);
    # 37 spheres = 3 target tiles, the last one ragged, so one-tile slabs
    # make three builds per product
    lmax = 6
    pos, k, t, a, f = _coefficients(77, 37, lmax, 560.0, 1.6 + 0.05j)
    s = _native.Session(pos, lmax, [k])
    s.set_scattering_method(_native.CSCA_PAIR_SUM)
    s.set_matvec_mode(mode)
    if slab_tiles:
        s.set_slab_tiles(slab_tiles)
    got = s.cross_sections(a, f)
    npt.assert_allclose(got, R.cross_sections(lmax, pos, k, a, f), rtol=1e-10)
    # a matvec in between (the FP64 slab then holds irregular factors) and again
    s.coupling_matvec(f)
    npt.assert_allclose(s.cross_sections(a, f), got, rtol=1e-13)


@pytest.mark.parametrize("method", [m for m, _ in METHODS], ids=[i for _, i in METHODS])
def test_cross_sections_single_sphere_and_far_pairs(method):
    # one sphere: only the coincident (self) term of the J-translation sum
    pos = np.zeros((1, 3))
    lmax, lam, k = 6, 600.0, 2 * np.pi / 600.0
    t = R.mie_coefficients(lmax, np.array([200.0]), np.array([1.7 + 0.1j]), np.array([lam]), 1.0)[0, :, 0]
    a = R.planewave_coefficients(lmax, pos, np.array([k]), 0.0, 0.0, 1)[0, :, 0]
    s = _native.Session(pos, lmax, [k])
    s.set_scattering_method(method)
    got = s.cross_sections(a, t * a)
    ref = R.cross_sections(lmax, pos, k, a, t * a)
    npt.assert_allclose(got, ref, rtol=1e-12)
    # two spheres far apart (k d ~ 60): the upward j_p recurrence regime of
    # the pair sum; a wide band (k R ~ 55) for the quadrature
    pos2 = np.array([[0.0, 0.0, 0.0], [1500.0, 3000.0, -4000.0]])
    f2 = np.tile(t * a, 2)
    s2 = _native.Session(pos2, lmax, [k])
    s2.set_scattering_method(method)
    npt.assert_allclose(s2.cross_sections(np.tile(a, 2), f2),
                        R.cross_sections(lmax, pos2, k, np.tile(a, 2), f2), rtol=1e-11)


@pytest.mark.parametrize("lmax", [2, 5, 8])
def test_far_field_matches_oracle(lmax):
    N = 17 if lmax < 8 else 6
    pos, k, t, a, f = _coefficients(40 + lmax, N, lmax, 480.0, 1.4 + 0.0j)
    rng = np.random.default_rng(lmax)
    th = np.concatenate([[0.0, np.pi, np.pi / 2], rng.uniform(0, np.pi, 50)])
    ph = np.concatenate([[0.0, 1.0, 2 * np.pi - 1e-3], rng.uniform(0, 2 * np.pi, 50)])
    s = _native.Session(pos, lmax, [k])
    et, ep = s.far_field(f, th, ph)
    rt, rp = R.far_field(lmax, pos, k, f, th, ph)
    scale = max(np.abs(rt).max(), np.abs(rp).max())
    npt.assert_allclose(et, rt, rtol=1e-10, atol=1e-12 * scale)
    npt.assert_allclose(ep, rp, rtol=1e-10, atol=1e-12 * scale)


def _sim(pos, radii, n, lam, lmax, mixed=False, spn=(72, 181)):
    particles = Particles(pos, radii, n)
    field = InitialField(beam_width=0, focal_point=np.zeros(3), polar_angle=0.0, azimuthal_angle=0.0,
                         polarization="TE")
    params = Parameters(np.atleast_1d(lam), np.atleast_1d(1.0), particles, field)
    solver = Solver("gmres", tolerance=1e-6, max_iter=1000, restart=100, mixed_precision=mixed)
    sim = Simulation(params, Numerics(lmax=lmax, sampling_points_number=np.array(spn), solver=solver,
                                      coupling_backend="hip"))
    sim.compute_mie_coefficients()
    sim.compute_initial_field_coefficients()
    sim.compute_right_hand_side()
    sim.compute_scattered_field_coefficients()
    return sim


def test_cfg1_observables_match_oracle():
    """The CPU-runnable workload (50 spheres, lmax 4, lambda 550 nm): C_ext, C_sca and
    the 181-angle phase function of the device solution vs the oracle on the same
    coefficients, and against the oracle's own GMRES solution to the solve tolerance."""
    pos, radii = cluster(1, 50, 0.1)
    n = np.full(50, 1.5 + 0.01j)
    sim = _sim(pos, radii, n, 550.0, 4)
    opt = Optics(sim)
    opt.compute_cross_sections()
    opt.compute_phase_funcition()
    k = 2 * np.pi / 550.0
    a = sim.initial_field_coefficients[:, :, 0].ravel()
    f = sim.scattered_field_coefficients[:, :, 0].ravel()
    r_ext, r_sca = R.cross_sections(4, pos, k, a, f)
    assert opt.c_ext[0] == pytest.approx(r_ext, rel=1e-10)
    assert opt.c_sca[0] == pytest.approx(r_sca, rel=1e-10)
    th, p = R.phase_function(4, pos, k, f, r_sca, (72, 181))
    npt.assert_allclose(opt.scattering_angles, th)
    npt.assert_allclose(opt.phase_function[:, 0], p, rtol=1e-9)
    assert opt._check_phase_function(precision=1e-3)
    # oracle solve (SciPy GMRES over the oracle operator) -> observables at 1e-6
    t = sim.tmatrix_flat()[0]
    fr, info, _ = R.gmres_solve(lambda v: R.master_matvec(4, pos, k, t, v), t * a, rtol=1e-8,
                                restart=100)
    o_ext, o_sca = R.cross_sections(4, pos, k, a, fr)
    assert abs(opt.c_ext[0] - o_ext) / o_ext < 1e-6
    assert abs(opt.c_sca[0] - o_sca) / o_sca < 1e-6


def test_mixed_precision_observables():
    """Mixed-precision solve: cross sections and phase function within 1e-6 of
    the all-FP64 solve."""
    pos, radii = cluster(5, 120, 0.15)
    n = np.full(120, 1.33 + 0.001j)
    out = {}
    for mixed in (False, True):
        sim = _sim(pos, radii, n, 600.0, 5, mixed=mixed, spn=(24, 91))
        assert sim.scattered_field_err_codes[0] == 0
        opt = Optics(sim)
        opt.compute_cross_sections()
        opt.compute_phase_funcition()
        out[mixed] = (opt.c_ext[0], opt.c_sca[0], opt.phase_function[:, 0])
    for i in range(2):
        assert abs(out[True][i] - out[False][i]) / abs(out[False][i]) < 1e-6
    assert np.max(np.abs(out[True][2] - out[False][2]) / out[False][2]) < 1e-5
