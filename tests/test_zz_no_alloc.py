"""Allocation-free coupling matvec (yasfhip_coupling_matvec_noalloc), GPU.

Called through the C ABI on device tensors, twice in a row with the second
call smaller than the first, and on a size that takes several target-tile
chunks of the scratch array; each result must equal, bit for bit, the
session's FP64 matvec with on-the-fly pair factors (the same kernels), which
the oracle pins at 1e-12."""

import ctypes

import numpy as np
import pytest

from oracle import restate as R
from yasf_amd import _native
from yasf_amd.clusters import cluster

pytestmark = pytest.mark.gpu


def _noalloc(pos_d, N, lmax, k, x_d, y_d, stream=None):
    import torch

    lib = _native.library()
    handle = torch.cuda.current_stream().cuda_stream if stream is None else stream
    status = lib.yasfhip_coupling_matvec_noalloc(
        ctypes.c_void_p(pos_d.data_ptr()), N, lmax, k, ctypes.c_void_p(x_d.data_ptr()),
        ctypes.c_void_p(y_d.data_ptr()), ctypes.c_void_p(handle))
    return status, lib.yasfhip_last_error().decode()


def _session_matvec(pos, lmax, k, x):
    s = _native.Session(pos, lmax, [k])
    s.set_matvec_mode(_native.MATVEC_ON_THE_FLY)
    y = s.coupling_matvec(x)
    s.close()
    return y


@pytest.mark.parametrize("cases", [[(1000, 6, 21), (40, 4, 22)], [(200, 8, 23), (90, 3, 24)]],
                         ids=["lmax6_then_lmax4", "lmax8_then_lmax3"])
def test_noalloc_matvec_matches_session(cases):
    import torch

    for N, lmax, seed in cases:  # the second call is smaller than the first
        pos, _ = cluster(seed, N, 0.1)
        k = 2 * np.pi / 550.0
        n = N * 2 * lmax * (lmax + 2)
        rng = np.random.default_rng(seed)
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        pos_d = torch.from_numpy(np.ascontiguousarray(pos)).to("cuda")
        x_d = torch.from_numpy(x).to("cuda")
        y_d = torch.full_like(x_d, complex(np.nan, np.nan))
        status, msg = _noalloc(pos_d, N, lmax, k, x_d, y_d)
        assert status == 0, msg
        torch.cuda.synchronize()
        y = y_d.cpu().numpy()
        ref = _session_matvec(pos, lmax, k, x)
        assert np.array_equal(y, ref), (N, lmax, np.abs(y - ref).max())
        if N <= 200:
            orc = R.coupling_matvec(lmax, pos, k, x, s1_lo=0, n_s1=3)
            assert np.linalg.norm(y[:orc.size] - orc) / np.linalg.norm(orc) <= 1e-12


def test_noalloc_rejects_bad_arguments():
    import torch

    pos, _ = cluster(25, 30, 0.1)
    k = 2 * np.pi / 550.0
    n = 30 * 48
    pos_d = torch.from_numpy(pos).to("cuda")
    x_d = torch.zeros(n, dtype=torch.complex128, device="cuda")
    y_d = torch.zeros_like(x_d)
    status, msg = _noalloc(pos_d, 32769, 4, k, x_d, y_d)  # beyond the scratch array's limit
    assert status == 1 and "particle_number" in msg
    status, msg = _noalloc(pos_d, 30, 9, k, x_d, y_d)
    assert status == 1 and "lmax" in msg
    lib = _native.library()
    xh = np.zeros(n, dtype=complex)
    status = lib.yasfhip_coupling_matvec_noalloc(ctypes.c_void_p(pos_d.data_ptr()), 30, 4, k,
                                                 xh.ctypes.data_as(ctypes.c_void_p),
                                                 ctypes.c_void_p(y_d.data_ptr()), None)
    assert status == 1 and "device memory" in lib.yasfhip_last_error().decode()
    # a valid call still works after the refusals
    x_d.copy_(torch.from_numpy(np.random.default_rng(3).normal(size=n) + 0j))
    status, msg = _noalloc(pos_d, 30, 4, k, x_d, y_d)
    assert status == 0, msg
    torch.cuda.synchronize()
    assert np.array_equal(y_d.cpu().numpy(), _session_matvec(pos, 4, k, x_d.cpu().numpy()))
