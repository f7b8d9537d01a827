"""Coupling backends: the reference's plugin seam with the new ``hip`` backend.

``CouplingBackend.multiply(x, idx)`` keeps the contract of
yasfpy/coupling/backends.py:36-66: ``x`` complex128 particle-major ``(n,)``
for one channel ``idx`` (returns ``(n,)``) or ``(n, W)`` with ``idx=None``
(returns ``(n, W)``).  ``HipCouplingBackend`` is matrix-free
(``requires_pairwise_lookups = False``) and returns fresh arrays.

The reference's backend names select the same device operator: ``dense`` and
``tiled_dense`` compute the exact W (backends.py:99-236, which differ only in
how their CPU tables are held), ``nearfar`` the near-field approximation of
backends.py:259-459 -- pairs farther apart than
``numerics.coupling_near_field_radius`` do not couple.
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

    def __init__(self, simulation: "Simulation", precision: int = _native.FP64, near_far: bool = False):
        self._sim = simulation
        self.precision = precision
        self.near_far = near_far

    @property
    def near_field_radius(self) -> float | None:
        """Cut-off of the near/far approximation (None: exact W).  As in
        backends.py:315-319 and 394-398, a near/far backend without
        ``numerics.coupling_near_field_radius`` raises ValueError when used."""
        if not self.near_far:
            return None
        radius = getattr(self._sim.numerics, "coupling_near_field_radius", None)
        if radius is None:
            raise ValueError("nearfar backend requires numerics.coupling_near_field_radius to be set.")
        return float(radius)

    @property
    def session(self) -> _native.Session:
        return self._sim.session

    def multiply(self, x: np.ndarray, idx: int | None = None) -> np.ndarray:
        self.near_field_radius  # noqa: B018 -- raises for a near/far backend without a radius
        s = self.session
        x = np.asarray(x)
        if idx is not None:
            return s.coupling_matvec(x.reshape(-1), channel=int(idx), precision=self.precision)
        x2 = x.reshape(s.n, -1)
        if x2.shape[1] == 1:
            # the reference's all-channel product: one x, W x at every channel's
            # wave number (cpu_numba.py:98-181) -> (n, W)
            xv = np.ascontiguousarray(x2[:, 0])
            return np.stack([s.coupling_matvec(xv, channel=w, precision=self.precision)
                             for w in range(s.channels)], axis=1)
        if x2.shape[1] != s.channels:
            raise ValueError(f"x has {x2.shape[1]} channels, expected 1 or {s.channels}")
        return np.stack([s.coupling_matvec(np.ascontiguousarray(x2[:, w]), channel=w,
                                           precision=self.precision)
                         for w in range(s.channels)], axis=1)


_EXACT = {"hip", "dense", "tiled_dense", "tiled-dense", "tileddense"}
_NEAR_FAR = {"nearfar", "near_far", "near-far"}


def make_backend(name: str, simulation: "Simulation") -> CouplingBackend:
    """Backend selection as in yasfpy/simulation.py:95-106, plus ``hip``; every
    name runs on the device."""
    key = str(name).lower()
    if key in _EXACT:
        return HipCouplingBackend(simulation)
    if key in _NEAR_FAR:
        return HipCouplingBackend(simulation, near_far=True)
    raise ValueError(f"Unsupported coupling backend: {key!r}. Expected one of "
                     "{'hip', 'dense', 'tiled_dense', 'nearfar'}.")
