"""Oracle: CPU restatement of the reference's T-matrix solve path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg may import, call or load anything under
``oracle/``, and only as the reference that results are compared against (or the
timed CPU baseline), never as the thing measured or shipped.  The product package ``yasf_amd`` never imports it.

``restate.py`` holds the numpy restatement, ``matvec.c`` the O(N^2) loop nests
(built by ``Makefile`` into ``liboracle.so``).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None


def build(force: bool = False) -> Path:
    """Compile ``liboracle.so`` with gcc (the ``Makefile`` recipe)."""
    so = _HERE / "liboracle.so"
    if force and so.exists():
        so.unlink()
    subprocess.run(["make", "-s", "-C", str(_HERE), "liboracle.so"], check=True)
    return so


class _Native:
    def __init__(self, lib):
        self._lib = lib
        P = ctypes.c_void_p
        i = ctypes.c_int
        lib.oracle_particle_interaction_scalar.argtypes = [i, i, i, i, P, P, P, P, P, P, P]
        lib.oracle_particle_interaction_scalar.restype = None
        lib.oracle_scattering_cross_section.argtypes = [i, i, P, P, P, P, P, P, P]
        lib.oracle_scattering_cross_section.restype = None
        lib.oracle_set_threads.argtypes = [i]
        lib.oracle_set_threads.restype = i

    def set_threads(self, n: int) -> int:
        """Set the OpenMP thread count; returns the count now in effect."""
        return int(self._lib.oracle_set_threads(int(n)))

    @staticmethod
    def _p(a: np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data_as(ctypes.c_void_p)

    def matvec(self, lmax, N, s1_lo, n_s1, idx, x, ab5, plm, sph_h, ephi, out):
        p = self._p
        self._lib.oracle_particle_interaction_scalar(
            lmax, N, s1_lo, n_s1, p(idx), p(x), p(ab5), p(plm), p(sph_h), p(ephi), p(out))

    def cross_section(self, lmax, N, idx, f, ab5, plm, sph_j, ephi, out):
        p = self._p
        self._lib.oracle_scattering_cross_section(
            lmax, N, p(idx), p(f), p(ab5), p(plm), p(sph_j), p(ephi), p(out))


def load_native() -> _Native:
    global _LIB
    if _LIB is None:
        so = _HERE / "liboracle.so"
        if not so.exists():
            build()
        _LIB = _Native(ctypes.CDLL(str(so)))
    return _LIB
