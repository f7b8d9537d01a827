"""Coupling backends: the reference's plugin seam with the new ``hip`` backend.

``CouplingBackend.multiply(x, idx)`` keeps the contract of
yasfpy/coupling/backends.py:36-66: ``x`` complex128 particle-major ``(n,)``
for one channel ``idx`` (returns ``(n,)``) or ``(n, W)`` with ``idx=None``
(returns ``(n, W)``).  ``HipCouplingBackend`` is matrix-free
(``requires_pairwise_lookups = False``) and returns fresh arrays.
"""

from __future__ import annotations

from typing import TYPE_CHECKING

import numpy as np

from yasf_amd import _native

if TYPE_CHECKING:
    from yasf_amd.simulation import Simulation


class CouplingBackend:
    requires_pairwise_lookups: bool = False

    def multiply(self, x: np.ndarray, idx: int | None = None) -> np.ndarray:  # pragma: no cover
        raise NotImplementedError


class HipCouplingBackend(CouplingBackend):
    """W x on MI355X through libyasfhip (pair-GEMM on MFMA + ab5 contraction)."""

    requires_pairwise_lookups = False

    def __init__(self, simulation: "Simulation", precision: int = _native.FP64):
        self._sim = simulation
        self.precision = precision

    @property
    def session(self) -> _native.Session:
        return self._sim.session

    def multiply(self, x: np.ndarray, idx: int | None = None) -> np.ndarray:
        s = self.session
        x = np.asarray(x)
        if idx is not None:
            return s.coupling_matvec(x.reshape(-1), channel=int(idx), precision=self.precision)
        x2 = x.reshape(s.n, -1)
        if x2.shape[1] != s.channels:
            raise ValueError(f"x has {x2.shape[1]} channels, expected {s.channels}")
        return np.stack([s.coupling_matvec(np.ascontiguousarray(x2[:, w]), channel=w,
                                           precision=self.precision)
                         for w in range(s.channels)], axis=1)


_BACKENDS = {"hip": HipCouplingBackend}
_REFERENCE_ONLY = {"dense", "tiled_dense", "tiled-dense", "tileddense", "nearfar", "near_far",
                   "near-far"}


def make_backend(name: str, simulation: "Simulation") -> CouplingBackend:
    """Backend selection as in yasfpy/simulation.py:95-106, plus ``hip``."""
    key = str(name).lower()
    if key in _BACKENDS:
        return _BACKENDS[key](simulation)
    if key in _REFERENCE_ONLY:
        raise ValueError(f"Coupling backend {key!r} is the reference's CPU/CUDA table backend; "
                         "this package provides {'hip'}.")
    raise ValueError(f"Unsupported coupling backend: {key!r}. Expected one of {{'hip'}}.")
