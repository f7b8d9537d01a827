"""Per-request set-up on the device (yasfhip_mie_coefficients,
yasfhip_planewave_coefficients, yasfhip_server_submit_spheres) against the
host restatements, which are pinned to the reference's fixtures and the
textbook Mie series (test_oracle.py, test_native_host.py)."""

import math

import numpy as np
import numpy.testing as npt
import pytest

from yasf_amd import _native
from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
from yasf_amd.serving import make_request, serve

pytestmark = pytest.mark.gpu


def _host_t(lmax, radii, rel, k):
    pairs = np.stack([radii, rel], axis=1).astype(complex)
    return mie_coefficients(lmax, pairs, np.array([k]), np.array([k]))[:, :, 0].ravel()


@pytest.mark.parametrize("lmax", [1, 3, 5, 8])
def test_device_mie_matches_host(lmax):
    """Size parameters x >= 0.2: the host's SciPy form agrees to 2e-10 (its
    own error grows as 1/x^2 below that, from the cancellation in the Mie
    quotients)."""
    rng = np.random.default_rng(lmax)
    N = 300
    radii = np.exp(rng.normal(math.log(100.0), 0.4, N))
    radii[:2] = [32.0, 450.0]           # x from 0.2 to about 9.4
    rel = rng.uniform(1.3, 1.8, N) + 1j * rng.uniform(0.0, 0.1, N)
    rel[2] = 1.7 + 0.1j
    for lam in (300.0, 550.0, 1000.0):
        k = 2 * np.pi / lam
        keep = k * radii >= 0.2
        dev = _native.mie_coefficients(lmax, radii[keep], rel[keep], k)
        host = _host_t(lmax, radii[keep], rel[keep], k)
        assert np.max(np.abs(dev - host) / np.abs(host)) < 2e-10


def test_device_mie_small_spheres_high_precision():
    """x from 1e-3 to 0.2 against the same quotients in 40-digit arithmetic
    (mpmath Bessel functions).  The quotients cancel to O(x^2), so double
    precision bounds their relative accuracy by about eps / x^2."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    lmax = 6
    k = 2 * np.pi / 700.0
    radii = np.array([0.12, 1.0, 5.0, 22.0])
    rel = np.array([1.33 + 0.0j, 1.5 + 0.01j, 1.7 + 0.1j, 1.8 + 0.05j])
    dev = _native.mie_coefficients(lmax, radii, rel, k).reshape(4, -1)

    def sj(l, z):
        return mp.sqrt(mp.pi / (2 * z)) * mp.besselj(l + mp.mpf(1) / 2, z)

    def sy(l, z):
        return mp.sqrt(mp.pi / (2 * z)) * mp.bessely(l + mp.mpf(1) / 2, z)

    for s in range(4):
        tol = max(1e-10, 5e-14 / (k * radii[s]) ** 2)
        X = mp.mpf(k * radii[s])
        M = mp.mpc(rel[s].real, rel[s].imag)
        Z = M * X
        for l in range(1, lmax + 1):
            jx, jz = sj(l, X), sj(l, Z)
            hx, hxm = jx + 1j * sy(l, X), sj(l - 1, X) + 1j * sy(l - 1, X)
            djx, dhx = X * sj(l - 1, X) - l * jx, X * hxm - l * hx
            djz = Z * sj(l - 1, Z) - l * jz
            b = complex((jx * djz - jz * djx) / (jz * dhx - hx * djz))
            a = complex((jx * djz - M ** 2 * jz * djx) / (M ** 2 * jz * dhx - hx * djz))
            n = l * l - 1
            assert abs(dev[s, n] - b) <= tol * abs(b), (s, l, "b")
            assert abs(dev[s, lmax * (lmax + 2) + n] - a) <= tol * abs(a), (s, l, "a")


def test_device_planewave_matches_host():
    rng = np.random.default_rng(3)
    pos = rng.normal(size=(257, 3)) * 800.0
    for lmax in (1, 4, 8):
        for pol in (1, 2):
            polar, azimuthal = float(rng.uniform(0, np.pi)), float(rng.uniform(0, 2 * np.pi))
            k = 2 * np.pi / 633.0
            focal = rng.normal(size=3) * 100.0
            dev = _native.planewave_coefficients(lmax, pos, k, polar, azimuthal, pol, 1.7, focal)
            host = planewave_coefficients(lmax, pos, np.array([k]), polar, azimuthal, pol, 1.7, focal)[:, :, 0]
            npt.assert_allclose(dev, host.ravel(), rtol=1e-12, atol=1e-12 * np.abs(host).max())
    with pytest.raises(ValueError):
        _native.planewave_coefficients(2, pos, k, 0.1, 0.2, 3)


def test_served_device_setup_matches_host_setup():
    """Requests submitted as spheres + incidence (t and a_inc set up on the
    device at admission) give the host-set-up solutions to rounding, also
    when they share an operator by key."""
    from yasf_amd.serving import make_aggregate_workload

    rng = np.random.default_rng(21)
    reqs = [make_request(rng, i, n_range=(20, 150)) for i in range(6)]
    for r in make_aggregate_workload([60], [500.0], 3, seed=8):
        r.index += 100
        reqs.append(r)
    host = serve(_native.Server(memory_bytes=1 << 30), reqs, keep_solutions=True)
    dev = serve(_native.Server(memory_bytes=1 << 30), reqs, keep_solutions=True, device_setup=True)
    for r in reqs:
        a, b = host.results[r.index], dev.results[r.index]
        assert b["info"] == 0
        assert np.linalg.norm(b["x"] - a["x"]) <= 1e-9 * np.linalg.norm(a["x"])
        assert b["c_ext"] == pytest.approx(a["c_ext"], rel=1e-9)
        assert b["c_sca"] == pytest.approx(a["c_sca"], rel=1e-9)
    srv = _native.Server(memory_bytes=1 << 30)
    r = reqs[0]
    with pytest.raises(ValueError):
        srv.submit_spheres(r.positions, -r.radii, r.refractive_index, r.lmax, r.k, 0.1, 0.2)
    with pytest.raises(ValueError):
        srv.submit_spheres(r.positions, r.radii, r.refractive_index, r.lmax, r.k, 0.1, 0.2, polarization=5)
    srv.close()
