"""Host-side checks of the product (no GPU): the C ABI library loads and exports
every declared symbol, the host tables match the oracle, argument errors are
raised before any device call, and the host set-up matches the fixtures."""

import re
from pathlib import Path

import numpy as np
import numpy.testing as npt
import pytest

from oracle import restate as R
from yasf_amd import _native
from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
from yasf_amd.model import Particles

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "yasfhip.h").read_text()
    declared = set(re.findall(r"\b(yasfhip_[a-z0-9_]+)\s*\(", header))
    lib = _native.library()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_native.exported_symbols())


@pytest.mark.parametrize("lmax", range(1, 9))
def test_translation_table_matches_exact_oracle(lmax):
    ours = _native.translation_table(lmax)
    exact = R.translation_table(lmax)
    npt.assert_allclose(ours, exact, rtol=0, atol=2e-15)
    assert np.count_nonzero(ours) == np.count_nonzero(exact)


def test_translation_table_fixture(golden):
    d = golden("ab5_lmax3")
    npt.assert_allclose(_native.translation_table(3), d["ab5"], rtol=1e-5, atol=1e-7)


def test_plan_covers_every_ab5_entry():
    # nnz reported by the plan equals the exact count; pair work per pair is
    # below the reference loop's direct count 8*T_nz + 8*nmax^2.  Building the
    # plan also runs its own checks (an error if one fails): every fp16 bin
    # list computes each of its jobs' (order, column tile) products once and
    # every Q slot the contraction reads, and the contraction rows pack each
    # m1's modes into one block.
    for lmax in range(1, 9):
        macs, nnz = _native.plan_stats(lmax)
        assert nnz == np.count_nonzero(R.translation_table(lmax))
        nmax = 2 * lmax * (lmax + 2)
        if lmax >= 3:  # below that the 16-column tile padding dominates
            assert macs < nnz + nmax * nmax


def test_argument_errors_before_device():
    with pytest.raises(ValueError):
        _native.translation_table(0)
    with pytest.raises(ValueError):
        _native.translation_table(9)
    with pytest.raises(ValueError):
        _native.Session(np.zeros((2, 3)), 3, [-1.0])
    with pytest.raises(NotImplementedError):
        _native.Session(np.zeros((2, 3)), 3, [1.0 + 0.1j])


def test_host_mie_and_planewave_fixtures(golden):
    d = golden("full_simulation")
    lam = float(d["wavelength"])
    nmed = float(d["medium_refractive_index"])
    p = Particles(d["positions"], d["radii"], d["refractive_index"])
    omega = np.array([2 * np.pi / lam])
    mie = mie_coefficients(int(d["lmax"]), p.unique_radius_index_pairs, omega, omega * nmed)
    npt.assert_allclose(mie[p.single_unique_array_idx, :, 0], d["mie_coefficients"], rtol=1e-4, atol=1e-6)
    a = planewave_coefficients(int(d["lmax"]), d["positions"], omega * nmed, 0.0, 0.0, 1)
    npt.assert_allclose(a[:, :, 0], d["initial_field_coefficients"], rtol=1e-4, atol=1e-6)


def test_host_mie_t_entry(golden):
    d = golden("t_entry")
    pairs = np.array([[d["radius"], d["k_sphere"] / d["k_medium"]]])
    mie = mie_coefficients(1, pairs, np.array([float(d["k_medium"])]), np.array([float(d["k_medium"])]))
    npt.assert_allclose(mie[0, 0, 0], d["scattered_1"], rtol=1e-5)
    npt.assert_allclose(mie[0, 3, 0], d["scattered_2"], rtol=1e-5)


def test_backend_selection():
    from yasf_amd import InitialField, Numerics, Parameters, Simulation

    p = Particles(np.zeros((1, 3)), np.array([50.0]), np.array([1.5 + 0j]))
    params = Parameters(np.array([500.0]), np.array([1.0]), p, InitialField(beam_width=0))
    with pytest.raises(ValueError):
        Simulation(params, Numerics(lmax=2, coupling_backend="nope"))
    # every reference name selects the device operator, without pair tables
    # (test_coupling_backend_selection.py); the near/far one needs a radius
    # when used (test_coupling_backend_selection.py:66-74)
    for name, near in (("HIP", None), ("dense", None), ("tiled_dense", None), ("Tiled-Dense", None),
                       ("nearfar", 250.0), ("near_far", 250.0)):
        sim = Simulation(params, Numerics(lmax=2, coupling_backend=name, coupling_near_field_radius=near))
        assert sim.coupling_backend.requires_pairwise_lookups is False
        assert sim.plm is None and sim.sph_h is None and sim.e_j_dm_phi is None
        assert sim.coupling_backend.near_field_radius == near
    sim = Simulation(params, Numerics(lmax=2, coupling_backend="nearfar"))
    with pytest.raises(ValueError):
        sim.coupling_matrix_multiply(np.zeros(16, dtype=complex))


def test_hip_runtime_list_from_maps_text():
    maps = "\n".join([
        "7f00-7f10 r-xp 00000000 08:01 1 /usr/lib/x/torch/lib/libamdhip64.so",
        "7f10-7f20 r--p 00010000 08:01 1 /usr/lib/x/torch/lib/libamdhip64.so",
        "7f20-7f30 r-xp 00000000 08:01 2 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.0",
        "7f30-7f40 rw-p 00000000 00:00 0 ",
        "7f40-7f50 r-xp 00000000 08:01 3 /usr/lib/libc.so.6",
    ])
    assert _native.hip_runtimes_in_maps(maps) == ["/usr/lib/x/torch/lib/libamdhip64.so",
                                                  "/opt/rocm-7.2.0/lib/libamdhip64.so.7.2.0"]
    assert _native.hip_runtimes_in_maps("") == []


def test_library_loads_into_torch_hip_runtime():
    # Loading the library first must not leave two HIP runtimes in a process
    # that uses torch: _load imports torch before the library, so the
    # library's libamdhip64.so.7 resolves to the runtime torch already mapped.
    import subprocess
    import sys

    code = ("from yasf_amd import _native\n"
            "_native.library()\n"
            "import torch\n"
            "print('\\n'.join(_native.hip_runtimes()))\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-3000:]
    runtimes = [line for line in out.stdout.splitlines() if line.strip()]
    assert len(runtimes) == 1, runtimes
    assert "torch" in runtimes[0]


def test_serve_refuses_duplicate_request_indices():
    # results are keyed by request index: duplicates would never all arrive
    from yasf_amd.serving import ServeRequest, serve

    r = ServeRequest(0, np.zeros((1, 3)), 3, 500.0, 0.01, np.zeros(30, complex), np.zeros(30, complex), 1.5, 0.0, 0.0)
    with pytest.raises(ValueError):
        serve(None, [r, r])
