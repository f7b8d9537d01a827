"""GPU parity of the coupling operator (through the C ABI) against the oracle.

Bar: FP64 matvec within 1e-12 relative (2-norm) of the oracle's restatement of
particle_interaction_scalar; FP32 pair GEMM within 2e-5; MATLAB fixture at
the reference's own rtol 2e-4 / atol 1e-5."""

import numpy as np
import numpy.testing as npt
import pytest

from oracle import restate as R
from yasf_amd import _native
from yasf_amd.clusters import cluster

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def rand_vec(rng, n):
    return rng.normal(size=n) + 1j * rng.normal(size=n)


def test_fixture_matvec(golden):
    d = golden("coupling_matvec")
    L = int(d["lmax"])
    k = 2 * np.pi / float(d["wavelength"]) * float(d["medium_refractive_index"])
    s = _native.Session(d["positions"], L, [k])
    y = s.coupling_matvec(d["x"])
    npt.assert_allclose(y, d["wx"], rtol=2e-4, atol=1e-5)
    assert rel(y, R.coupling_matvec(L, d["positions"], k, d["x"])) < 1e-12


MODES = [_native.MATVEC_ON_THE_FLY, _native.MATVEC_TABLE, _native.MATVEC_SLAB]
MODE_IDS = ["on_the_fly", "table", "slab"]


def set_mode(s, mode, slab_tiles=1):
    s.set_matvec_mode(mode)
    if mode == _native.MATVEC_SLAB:
        s.set_slab_tiles(slab_tiles)  # several slabs per matvec


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("lmax", [1, 2, 3, 4, 5, 6, 7, 8])
def test_matvec_all_orders(lmax, mode):
    rng = np.random.default_rng(100 + lmax)
    N = 21  # not a multiple of the 16-particle tile or the 4-source step
    pos, _ = cluster(lmax, N, 0.1)
    k = 2 * np.pi / 550.0
    x = rand_vec(rng, N * 2 * lmax * (lmax + 2))
    s = _native.Session(pos, lmax, [k])
    set_mode(s, mode)
    y = s.coupling_matvec(x)
    ref = R.coupling_matvec(lmax, pos, k, x)
    assert rel(y, ref) < 1e-12
    y32 = s.coupling_matvec(x, precision=_native.FP32)
    assert rel(y32, ref) < 2e-5


def test_single_particle_is_zero():
    s = _native.Session(np.zeros((1, 3)), 3, [0.01])
    y = s.coupling_matvec(np.ones(30, dtype=complex))
    assert np.all(y == 0)


def test_small_cluster_config():
    """50 spheres, lmax 4, lambda 550 nm (the CPU-runnable workload)."""
    pos, _ = cluster(1, 50, 0.1)
    k = 2 * np.pi / 550.0
    x = rand_vec(np.random.default_rng(1), 50 * 48)
    s = _native.Session(pos, 4, [k])
    assert rel(s.coupling_matvec(x), R.coupling_matvec(4, pos, k, x)) < 1e-12


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_full_size_sampled_rows(mode):
    """2000 spheres, lmax 6: full matvec on the GPU, oracle on sampled target rows."""
    pos, _ = cluster(2, 2000, 0.2)
    k = 2 * np.pi / 550.0
    nmax = 96
    x = rand_vec(np.random.default_rng(2), 2000 * nmax)
    s = _native.Session(pos, 6, [k])
    set_mode(s, mode, slab_tiles=7)  # 125 target tiles: 18 slabs, the last one short
    y = s.coupling_matvec(x)
    for s1 in (0, 777, 1999):
        ref = R.coupling_matvec(6, pos, k, x, s1_lo=s1, n_s1=1)
        assert rel(y[s1 * nmax:(s1 + 1) * nmax], ref) < 1e-12
    # linearity at full size
    x2 = rand_vec(np.random.default_rng(3), x.size)
    a, b = 0.3 - 1.1j, -2.0 + 0.5j
    y2 = s.coupling_matvec(x2)
    assert rel(s.coupling_matvec(a * x + b * x2), a * y + b * y2) < 1e-13


def test_modes_agree_and_table_follows_channel():
    pos, _ = cluster(8, 40, 0.15)
    k = 2 * np.pi / np.array([400.0, 650.0])
    s = _native.Session(pos, 5, k)
    x = rand_vec(np.random.default_rng(8), 40 * 70)
    out = {}
    for mode in MODES:
        set_mode(s, mode)
        out[mode] = [s.coupling_matvec(x, channel=w) for w in (0, 1, 0)]
    for w in range(3):
        assert rel(out[MODES[1]][w], out[MODES[0]][w]) < 1e-13
        assert np.array_equal(out[MODES[2]][w], out[MODES[1]][w])  # slab = table, bit for bit
    assert rel(out[MODES[1]][0], out[MODES[1]][1]) > 1e-3  # the channels differ
    with pytest.raises(ValueError):
        s.set_matvec_mode(7)


def test_master_and_multichannel():
    pos, radii = cluster(7, 30, 0.1)
    lam = np.array([450.0, 700.0])
    k = 2 * np.pi / lam
    L = 3
    n = 30 * 30
    mie = R.mie_coefficients(L, radii, np.full(30, 1.5 + 0.01j), lam, 1.0)  # [N, nmax, W]
    t = np.moveaxis(mie, 2, 0).reshape(2, n)
    s = _native.Session(pos, L, k)
    s.set_tmatrix(t)
    rng = np.random.default_rng(4)
    for w in range(2):
        x = rand_vec(rng, n)
        assert rel(s.master_matvec(x, channel=w), R.master_matvec(L, pos, k[w], t[w], x)) < 1e-12


@pytest.mark.parametrize("lmax", [3, 6, 8])
def test_table_groupings_agree(lmax):
    """The table-GEMM tuner times every offered column-tile grouping (2/4/8
    tiles in both precisions; FP64's 8-tile jobs run one order per wave) and
    the matvec with its choice keeps oracle parity."""
    pos, _ = cluster(20 + lmax, 45, 0.1)
    k = 2 * np.pi / 500.0
    x = rand_vec(np.random.default_rng(lmax), 45 * 2 * lmax * (lmax + 2))
    s = _native.Session(pos, lmax, [k])
    s.set_matvec_mode(_native.MATVEC_TABLE)
    for prec, offered in ((_native.FP64, (0, 1, 2, 3)), (_native.FP32, (0, 1, 2, 3, 4))):
        g, ms = s.tune_matvec(precision=prec, reps=1)
        assert g in offered and all((ms[i] > 0) == (i in offered) for i in range(6))
        # every offered shape computes the same C tiles; the orders a block
        # cuts along K depend on its job's order count, which the column-tile
        # grouping sets, so shapes agree to rounding
        outs = []
        for shape in offered:
            s.set_table_shape(shape, precision=prec)
            outs.append(s.coupling_matvec(x, precision=prec))
        agree = 1e-14 if prec == _native.FP64 else 1e-6
        assert all(rel(o, outs[0]) < agree for o in outs[1:])
        # a repeat gives the same bits: the pieces are summed in a fixed order
        assert np.array_equal(s.coupling_matvec(x, precision=prec), outs[-1])
        with pytest.raises(ValueError):
            s.set_table_shape(5, precision=prec)
        s.set_table_shape(g, precision=prec)
        y = s.coupling_matvec(x, precision=prec)
        oracle_y = R.coupling_matvec(lmax, pos, k, x)
        assert rel(y, oracle_y) < (1e-12 if prec == _native.FP64 else 2e-5)


@pytest.mark.parametrize("N", [2, 5, 17])
def test_table_shapes_few_particles(N):
    """Blocks cut their leftover orders into K pieces of a quarter of the
    sources each; with fewer than 16 sources some pieces are empty.  Every
    shape keeps oracle parity there in both precisions."""
    pos, _ = cluster(50 + N, N, 0.1)
    L, k = 5, 2 * np.pi / 520.0
    x = rand_vec(np.random.default_rng(N), N * 2 * L * (L + 2))
    ref = R.coupling_matvec(L, pos, k, x)
    s = _native.Session(pos, L, [k])
    s.set_matvec_mode(_native.MATVEC_TABLE)
    for prec, offered, tol in ((_native.FP64, (0, 1, 2, 3), 1e-12), (_native.FP32, (0, 1, 2, 3, 4), 2e-5)):
        for shape in offered:
            s.set_table_shape(shape, precision=prec)
            assert rel(s.coupling_matvec(x, precision=prec), ref) < tol, (prec, shape)
    s.close()


def test_slab_matches_table_and_tunes():
    """Slabs of any size give the resident table's result bit for bit in both
    precisions (same build and GEMM shape per target tile); the tuner times
    the first slab when no table is kept; AUTO sizes its own slab."""
    pos, _ = cluster(31, 70, 0.15)   # 5 target tiles
    k = 2 * np.pi / 480.0
    L = 4
    x = rand_vec(np.random.default_rng(31), 70 * 48)
    s = _native.Session(pos, L, [k])
    s.set_matvec_mode(_native.MATVEC_TABLE)
    ref = {p: s.coupling_matvec(x, precision=p) for p in (_native.FP64, _native.FP32)}
    assert rel(ref[_native.FP64], R.coupling_matvec(L, pos, k, x)) < 1e-12
    s.set_matvec_mode(_native.MATVEC_SLAB)
    for tiles in (1, 2, 4, 5, 0):
        s.set_slab_tiles(tiles)
        for p in (_native.FP64, _native.FP32):
            assert np.array_equal(s.coupling_matvec(x, precision=p), ref[p]), (tiles, p)
    s.set_slab_tiles(2)
    for p, offered in ((_native.FP64, (0, 1, 2, 3)), (_native.FP32, (0, 1, 2, 3, 4))):
        g, ms = s.tune_matvec(precision=p, reps=1)
        assert g in offered and all((ms[i] > 0) == (i in offered) for i in range(6))
        # the tuned shape may cut leftover orders differently: equal to rounding
        assert rel(s.coupling_matvec(x, precision=p), ref[p]) < (1e-14 if p == _native.FP64 else 1e-6)
    with pytest.raises(ValueError):
        s.set_slab_tiles(-1)
    s.close()


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_near_field_cutoff(mode):
    """Near/far approximation (backends.py:259-459) against the oracle's
    restatement at 1e-12 in every pair-factor mode; radius 0 removes all
    coupling; a radius beyond every pair, or none, gives the exact W."""
    pos, _ = cluster(40, 40, 0.15)
    L, k = 3, 2 * np.pi / 530.0
    x = rand_vec(np.random.default_rng(40), 40 * 30)
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    radius = float(np.median(d[np.triu_indices(40, 1)]))
    s = _native.Session(pos, L, [k])
    set_mode(s, mode)
    exact = s.coupling_matvec(x)
    s.set_near_field_radius(radius)
    for prec, tol in ((_native.FP64, 1e-12), (_native.FP32, 2e-5)):
        y = s.coupling_matvec(x, precision=prec)
        assert rel(y, R.coupling_matvec(L, pos, k, x, near_field_radius=radius)) < tol
    assert rel(y, exact) > 1e-2
    s.set_near_field_radius(0.0)
    assert np.all(s.coupling_matvec(x) == 0)
    s.set_near_field_radius(10 * d.max())
    assert np.array_equal(s.coupling_matvec(x), exact)
    s.set_near_field_radius(None)
    assert np.array_equal(s.coupling_matvec(x), exact)
    s.close()


def test_reference_backend_names_on_device():
    """The reference's backend tests through the Simulation mirror: nearfar
    with a radius beyond every pair equals dense (idx=None with one x gives
    every channel's W x, and idx=1 the slice), radius 0 gives W x = 0
    (test_nearfar_backend.py, test_tiled_dense_backend.py)."""
    from yasf_amd import InitialField, Numerics, Parameters, Simulation
    from yasf_amd.model import Particles

    def make(backend, radius):
        p = Particles(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.5, 0.0]]), np.full(3, 0.1),
                      np.full(3, 1.5 + 0j))
        params = Parameters(np.array([1.0, 1.3]), np.array([1.0 + 0j, 1.0 + 0j]), p, InitialField(beam_width=0))
        return Simulation(params, Numerics(lmax=1, coupling_backend=backend, coupling_near_field_radius=radius))

    dense, tiled, nf, nf0 = make("dense", None), make("tiled_dense", None), make("nearfar", 10.0), make("nearfar", 0.0)
    x = rand_vec(np.random.default_rng(0), 3 * 6)
    wd = dense.coupling_matrix_multiply(x)
    assert wd.shape == (18, 2)
    npt.assert_allclose(tiled.coupling_matrix_multiply(x), wd, rtol=1e-12, atol=1e-12)
    npt.assert_allclose(nf.coupling_matrix_multiply(x), wd, rtol=1e-12, atol=1e-12)
    npt.assert_allclose(nf.coupling_matrix_multiply(x, idx=1), wd[:, 1], rtol=1e-12, atol=1e-12)
    assert np.all(nf0.coupling_matrix_multiply(x) == 0)
    for w in range(2):
        k = 2 * np.pi / [1.0, 1.3][w]
        assert rel(wd[:, w], R.coupling_matvec(1, dense.parameters.particles.position, k, x)) < 1e-12
