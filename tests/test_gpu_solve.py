"""GPU solve parity: device GMRES (and block preconditioner) through the
Simulation API and the C ABI against the reference's fixtures and SciPy GMRES
driving the oracle."""

import numpy as np
import numpy.testing as npt
import pytest

from oracle import restate as R
from yasf_amd import (BlockDiagonalPreconditioner, InitialField, Numerics, Parameters, Particles,
                      Simulation, Solver, edge_size_for)
from yasf_amd import _native
from yasf_amd.clusters import cluster

pytestmark = pytest.mark.gpu


def build_sim(positions, radii, n, lam, nmed, lmax, tol, restart=100, maxiter=1000, pc=None,
              mixed=False):
    particles = Particles(positions, radii, n)
    field = InitialField(beam_width=0, focal_point=np.zeros(3), polar_angle=0.0,
                         azimuthal_angle=0.0, polarization="TE")
    params = Parameters(np.atleast_1d(lam), np.atleast_1d(nmed), particles, field)
    solver = Solver("gmres", tolerance=tol, max_iter=maxiter, restart=restart, preconditioner=pc,
                    mixed_precision=mixed)
    sim = Simulation(params, Numerics(lmax=lmax, solver=solver, coupling_backend="hip"))
    sim.compute_mie_coefficients()
    sim.compute_initial_field_coefficients()
    sim.compute_right_hand_side()
    return sim


def test_full_simulation_fixture(golden):
    d = golden("full_simulation")
    sim = build_sim(d["positions"], d["radii"], d["refractive_index"], float(d["wavelength"]),
                    float(d["medium_refractive_index"]), int(d["lmax"]), 1e-4, restart=1000)
    sim.compute_scattered_field_coefficients()
    assert sim.scattered_field_err_codes[0] == 0
    npt.assert_allclose(sim.scattered_field_coefficients[:, :, 0], d["scattered_field_coefficients"],
                        rtol=1e-3, atol=1e-5)


def test_celes_500_fixture(golden):
    c = golden("celes500")
    sim = build_sim(c["positions"], c["radii"], c["refractive_index"], c["wavelength"],
                    c["medium_refractive_index"].real, int(c["lmax"]), float(c["tolerance"]),
                    restart=int(c["restart"]), maxiter=int(c["max_iter"]))
    sim.compute_scattered_field_coefficients()
    npt.assert_allclose(sim.initial_field_coefficients[:, :, 0], c["initial_field_coefficients"],
                        rtol=2e-3)
    npt.assert_allclose(sim.scattered_field_coefficients[:, :, 0], c["scattered_field_coefficients"],
                        rtol=2e-3)


def _cfg1():
    pos, radii = cluster(1, 50, 0.1)
    return pos, radii, np.full(50, 1.5 + 0.01j), 550.0, 1.0, 4


def test_gmres_matches_scipy_over_oracle():
    pos, radii, n, lam, nmed, L = _cfg1()
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-6, restart=100)
    sim.compute_scattered_field_coefficients()
    x = sim.scattered_field_coefficients[:, :, 0].ravel()
    b = sim.right_hand_side[:, :, 0].ravel()
    k = 2 * np.pi / lam
    t = sim.tmatrix_flat()[0]
    res = np.linalg.norm(b - R.master_matvec(L, pos, k, t, x)) / np.linalg.norm(b)
    assert res <= 1e-6
    xr, info, it_ref = R.gmres_solve(lambda v: R.master_matvec(L, pos, k, t, v), b, rtol=1e-6,
                                     restart=100)
    assert info == 0
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5
    its = sim.numerics.solver.last_iterations[0]
    assert abs(its - it_ref) <= 1


def test_gmres_history_matches_scipy_over_oracle():
    """Several restart cycles (restart 10): the device cycles' residual
    estimates follow SciPy's pr_norm history over the oracle operator (CGS2 on
    the device, MGS in SciPy: equal to rounding), the cycle count and the
    iterations agree within one per cycle, and the solutions agree."""
    pos, radii, n, lam, nmed, L = _cfg1()
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-8, restart=10)
    b = sim.right_hand_side[:, :, 0].ravel()
    k = 2 * np.pi / lam
    t = sim.tmatrix_flat()[0]
    x, info, iters, res, hist = sim.session.gmres(b, b, rtol=1e-8, restart=10, maxiter=100, history_cap=4096)
    assert info == 0 and res <= 1e-8
    from scipy.sparse.linalg import LinearOperator, gmres

    ref_hist = []
    A = LinearOperator((b.size, b.size), matvec=lambda v: R.master_matvec(L, pos, k, t, v), dtype=complex)
    xr, info_r = gmres(A, b, b.copy(), restart=10, rtol=1e-8, atol=0, maxiter=100,
                       callback=ref_hist.append, callback_type="pr_norm")
    assert info_r == 0
    ref_hist = np.asarray(ref_hist)
    hist = np.asarray(hist)
    assert len(hist) == iters
    cycles = -(-iters // 10)
    assert abs(iters - len(ref_hist)) <= max(1, cycles)
    m = min(len(hist), len(ref_hist), 30)
    npt.assert_allclose(hist[:m], ref_hist[:m], rtol=1e-6)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-6
    with pytest.raises(ValueError):
        sim.session.gmres(b, b, restart=801)


def test_block_preconditioner_apply_and_solve():
    pos, radii, n, lam, nmed, L = _cfg1()
    pc = BlockDiagonalPreconditioner(edge_size_for(pos, 8))
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-6, restart=100, pc=pc)
    sim.session.set_precond_factor_precision(_native.FP64)  # M_box^-1 to FP64 rounding
    pc.prepare(sim)
    k = 2 * np.pi / lam
    t = sim.tmatrix_flat()[0]
    ref_apply = R.block_preconditioner(L, pos, k, t, pc.partitioning)
    r = np.random.default_rng(9).normal(size=t.size) + 1j
    z_ref = ref_apply(r)
    npt.assert_allclose(pc.apply(r), z_ref, rtol=1e-10, atol=1e-10 * np.abs(r).max())
    # default: factors stored in complex64, substitutions in FP64 -- the apply is
    # M_box^-1 to about the factors' rounding, and the same operator every time
    sim.session.set_precond_factor_precision(_native.FP32)
    pc.prepare(sim)
    z32 = pc.apply(r)
    assert np.linalg.norm(z32 - z_ref) / np.linalg.norm(z_ref) < 1e-5
    assert np.array_equal(pc.apply(r), z32)
    sim.compute_scattered_field_coefficients()
    x_pc = sim.scattered_field_coefficients[:, :, 0].ravel().copy()
    b = sim.right_hand_side[:, :, 0].ravel()
    assert np.linalg.norm(b - R.master_matvec(L, pos, k, t, x_pc)) / np.linalg.norm(b) <= 1e-6
    its_pc = sim.numerics.solver.last_iterations[0]
    sim2 = build_sim(pos, radii, n, lam, nmed, L, 1e-6, restart=100)
    sim2.compute_scattered_field_coefficients()
    x = sim2.scattered_field_coefficients[:, :, 0].ravel()
    assert np.linalg.norm(x - x_pc) / np.linalg.norm(x) < 1e-5
    assert its_pc <= sim2.numerics.solver.last_iterations[0]


def test_single_box_preconditioner_is_exact():
    pos, radii = cluster(11, 6, 0.1)
    sim = build_sim(pos, radii, np.full(6, 1.6 + 0.02j), 600.0, 1.0, 3, 1e-8, restart=50)
    s = sim.session
    s.set_precond_factor_precision(_native.FP64)
    s.precond_build(np.zeros(6, dtype=np.int32))
    b = sim.right_hand_side[:, :, 0].ravel()
    x, info, iters, res, _ = s.gmres(b, b, rtol=1e-8, restart=50, maxiter=10, use_precond=True)
    assert info == 0 and iters == 1 and res < 1e-8


def test_gmres_device_vectors_match_host_vectors():
    # The device-pointer entry point runs the same solve on vectors in HBM:
    # identical solution, iterations, residual and history.  Its vectors are
    # torch tensors in this process (the library shares torch's HIP runtime).
    import ctypes

    import torch

    assert len(_native.hip_runtimes()) == 1
    pos, radii = cluster(12, 30, 0.1)
    lmax, lam = 4, 550.0
    k = 2 * np.pi / lam
    t = R.mie_coefficients(lmax, radii, np.full(30, 1.5 + 0.01j), np.array([lam]), 1.0)[:, :, 0].ravel()
    a = R.planewave_coefficients(lmax, pos, np.array([k]), 0.0, 0.0, 1)[:, :, 0].ravel()
    b = t * a
    s = _native.Session(pos, lmax, [k])
    s.set_tmatrix(t)
    s.precond_build(np.arange(30, dtype=np.int32) // 8)
    x, info, iters, res, hist = s.gmres(b, b, rtol=1e-7, restart=30, maxiter=20, use_precond=True)
    b_dev = torch.from_numpy(b.copy()).to("cuda")
    x_dev = b_dev.clone()
    info_d, iters_d, res_d, hist_d = s.gmres_device(b_dev, x_dev, rtol=1e-7, restart=30, maxiter=20,
                                                   use_precond=True)
    assert info == info_d == 0 and iters == iters_d and res == res_d
    assert np.array_equal(hist, hist_d)
    assert np.array_equal(x_dev.cpu().numpy(), x)
    with pytest.raises(ValueError):
        s.gmres_device(b_dev.cpu(), x_dev)  # host tensors are refused before any device work
    ic, itc, rc = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()

    def raw(bp, xp):
        return s._lib.yasfhip_gmres_solve_device(s._h, 0, bp, xp, 1e-7, 30, 20, 1, ctypes.byref(ic),
                                                 ctypes.byref(itc), ctypes.byref(rc), None, 0)

    # a host pointer at the C ABI is refused (YASFHIP_INVALID_ARGUMENT)
    assert raw(b.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(x_dev.data_ptr())) == 1
    # so is a device allocation shorter than n complex128 values (a direct
    # hipMalloc of 256 bytes from the runtime torch and the library share)
    hip = ctypes.CDLL("libamdhip64.so.7")
    small = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(small), ctypes.c_size_t(256)) == 0
    try:
        assert raw(ctypes.c_void_p(b_dev.data_ptr()), small) == 1
        assert "inside one device allocation" in _native.library().yasfhip_last_error().decode()
    finally:
        assert hip.hipFree(small) == 0
    x_dev.copy_(b_dev)  # the session still works after the refusals
    assert s.gmres_device(b_dev, x_dev, rtol=1e-7, restart=30, maxiter=20, use_precond=True)[1] == iters
    s.close()


def test_gmres_errors_and_nonconvergence():
    pos, radii, n, lam, nmed, L = _cfg1()
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-12, restart=2, maxiter=1)
    sim.compute_scattered_field_coefficients()
    assert sim.scattered_field_err_codes[0] == 1  # SciPy: info = maxiter when not converged
    with pytest.raises(ValueError):
        sim.session.gmres(sim.right_hand_side[:, :, 0].ravel(), channel=3)
    with pytest.raises(ValueError):
        sim.session.gmres(sim.right_hand_side[:, :, 0].ravel(), use_precond=True)
    zero = np.zeros(sim.session.n, dtype=complex)
    x, info, iters, res, _ = sim.session.gmres(zero)
    assert info == 0 and np.all(x == 0)


@pytest.mark.parametrize("use_pc", [False, True], ids=["plain", "block_precond"])
def test_mixed_precision_refinement(use_pc):
    """FP64 refinement around complex64-operator GMRES reaches the FP64
    criterion ||b - A x|| <= 1e-6 ||b|| against the oracle's operator and
    agrees with the all-FP64 solve to the solve tolerance."""
    pos, radii, n, lam, nmed, L = _cfg1()
    pc = BlockDiagonalPreconditioner(edge_size_for(pos, 8)) if use_pc else None
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-6, restart=50, pc=pc, mixed=True)
    if pc is not None:
        pc.prepare(sim)
    sim.compute_scattered_field_coefficients()
    assert sim.scattered_field_err_codes[0] == 0
    x = sim.scattered_field_coefficients[:, :, 0].ravel()
    b = sim.right_hand_side[:, :, 0].ravel()
    k = 2 * np.pi / lam
    t = sim.tmatrix_flat()[0]
    res = np.linalg.norm(b - R.master_matvec(L, pos, k, t, x)) / np.linalg.norm(b)
    assert res <= 1e-6
    xr, info, _ = R.gmres_solve(lambda v: R.master_matvec(L, pos, k, t, v), b, rtol=1e-8, restart=100)
    assert info == 0
    # forward error of a 1e-6-residual solution, bounded through the
    # conditioning of this small cluster (well below 1e-4)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5


def test_mixed_precision_history_and_errors():
    pos, radii, n, lam, nmed, L = _cfg1()
    sim = build_sim(pos, radii, n, lam, nmed, L, 1e-6, restart=50)
    s = sim.session
    b = sim.right_hand_side[:, :, 0].ravel()
    x, info, steps, iters, res, hist = s.gmres_refined(b, rtol=1e-10, inner_rtol=1e-5, restart=50)
    # each step reduces the FP64 residual; 1e-10 needs more than one step
    assert steps >= 2 and np.all(np.diff(hist) < 0)
    assert hist[-1] == pytest.approx(res)
    assert info == 0 and res <= 1e-10
    # one refinement step with inner_rtol 1e-3 cannot reach 1e-6 from x0 = b
    x, info, steps, iters, res, hist = s.gmres_refined(b, rtol=1e-6, inner_rtol=1e-3,
                                                       max_refinements=1)
    assert steps == 1 and info == 1 and res > 1e-6
    with pytest.raises(ValueError):
        s.gmres_refined(b, inner_rtol=0.0)
    with pytest.raises(ValueError):
        s.gmres_refined(b, max_refinements=0)
    x, info, steps, iters, res, hist = s.gmres_refined(np.zeros_like(b))
    assert info == 0 and steps == 0 and np.all(x == 0)


@pytest.mark.parametrize("restart", [1, 2, 7])
def test_device_cycles_short_restarts_match_scipy(restart):
    """Restart cycles of one, two and seven Arnoldi steps (every step ends or
    nearly ends a cycle, so the device stop decision, the gated steps queued
    past it and the per-cycle true residual are all exercised): the residual
    history and the solution follow SciPy's GMRES over the oracle operator."""
    from scipy.sparse.linalg import LinearOperator, gmres

    pos, radii = cluster(3, 12, 0.1)
    L, lam = 3, 600.0
    sim = build_sim(pos, radii, np.full(12, 1.45 + 0.01j), lam, 1.0, L, 1e-7, restart=restart)
    b = sim.right_hand_side[:, :, 0].ravel()
    k = 2 * np.pi / lam
    t = sim.tmatrix_flat()[0]
    x, info, iters, res, hist = sim.session.gmres(b, b, rtol=1e-7, restart=restart, maxiter=400,
                                                   history_cap=100000)
    ref_hist = []
    A = LinearOperator((b.size, b.size), matvec=lambda v: R.master_matvec(L, pos, k, t, v), dtype=complex)
    xr, info_r = gmres(A, b, b.copy(), restart=restart, rtol=1e-7, atol=0, maxiter=400,
                       callback=ref_hist.append, callback_type="pr_norm")
    assert info == 0 and info_r == 0 and res <= 1e-7
    assert len(hist) == iters
    assert abs(iters - len(ref_hist)) <= max(2, iters // 10)
    m = min(len(hist), len(ref_hist), 12)
    npt.assert_allclose(np.asarray(hist[:m]), np.asarray(ref_hist[:m]), rtol=1e-6)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-6


def test_device_cycles_single_sphere_and_breakdown():
    """One sphere: W = 0, so A = I and the first Arnoldi step breaks down
    (h[1,0] = 0): one step, x = b, as SciPy; and the smallest vector the
    Arnoldi blocks handle (one partial block)."""
    pos = np.zeros((1, 3))
    sim = build_sim(pos, np.array([120.0]), np.array([1.5 + 0.02j]), 550.0, 1.0, 2, 1e-8, restart=20)
    b = sim.right_hand_side[:, :, 0].ravel()
    x, info, iters, res, hist = sim.session.gmres(b, np.zeros_like(b), rtol=1e-8, restart=20, maxiter=5)
    assert info == 0 and iters == 1 and res < 1e-12
    npt.assert_allclose(x, b, rtol=1e-13, atol=1e-15 * np.abs(b).max())
