"""Simulation orchestration (mirror of yasfpy/simulation.py:33-249 and the
solve driver yasfpy/scattered_field_coefficients.py:43-156).

The ``hip`` backend needs no O(N^2) lookup tables: ``sph_j``, ``sph_h``,
``e_j_dm_phi`` and ``plm`` stay ``None`` (simulation.py:160-167).  One device
session per Simulation holds geometry, T-matrix and solver state.
"""

from __future__ import annotations

import logging
from time import perf_counter

import numpy as np

from yasf_amd import _native
from yasf_amd.coefficients import mie_coefficients, planewave_coefficients
from yasf_amd.coupling import make_backend
from yasf_amd.model import Parameters
from yasf_amd.numerics import Numerics


class Simulation:
    def __init__(self, parameters: Parameters, numerics: Numerics):
        self.parameters = parameters
        self.numerics = numerics
        self.log = logging.getLogger(self.__class__.__module__)
        self.coupling_backend = make_backend(getattr(numerics, "coupling_backend", "hip"), self)
        self.sph_j = self.sph_h = self.e_j_dm_phi = self.plm = None
        self._session: _native.Session | None = None
        self._near_field_radius: float | None = None
        self._tmatrix_uploaded = False
        self.mie_coefficients = None
        self.initial_field_coefficients = None
        self.right_hand_side = None

    # ------------------------------------------------------------ device state
    @property
    def session(self) -> _native.Session:
        if self._session is None:
            self._session = _native.Session(self.parameters.particles.position, self.numerics.lmax,
                                            self.parameters.k_medium)
            self._near_field_radius = None
        # the near/far cut-off as numerics holds it now (a change rebuilds the pair tables)
        radius = getattr(self.coupling_backend, "near_field_radius", None)
        if radius != self._near_field_radius:
            self._session.set_near_field_radius(radius)
            self._near_field_radius = radius
        if not self._tmatrix_uploaded and self.mie_coefficients is not None:
            self._session.set_tmatrix(self.tmatrix_flat())
            self._tmatrix_uploaded = True
        return self._session

    def tmatrix_flat(self) -> np.ndarray:
        """``t[w, s*nmax + n]`` = mie[single_unique_array_idx[s], n, w] (matrix_apply.py:104-107)."""
        idx = self.parameters.particles.single_unique_array_idx
        t = self.mie_coefficients[idx, :, :]  # [N, nmax, W]
        return np.ascontiguousarray(np.moveaxis(t, 2, 0).reshape(t.shape[2], -1))

    # ------------------------------------------------------------ set-up
    def compute_mie_coefficients(self):
        """mie_coefficients.py:29-161."""
        p = self.parameters
        self.mie_coefficients = mie_coefficients(self.numerics.lmax,
                                                 p.particles.unique_radius_index_pairs, p.omega,
                                                 p.k_medium)
        self._tmatrix_uploaded = False

    def compute_initial_field_coefficients(self):
        """initial_field_coefficients.py:34-142 (plane wave)."""
        f = self.parameters.initial_field
        if not f.is_plane_wave:
            raise NotImplementedError("only plane-wave incidence is provided")
        self.initial_field_coefficients = planewave_coefficients(
            self.numerics.lmax, self.parameters.particles.position, self.parameters.k_medium,
            f.polar_angle, f.azimuthal_angle, f.pol, f.amplitude, f.focal_point)

    def compute_right_hand_side(self):
        """b = T a_inc (simulation.py:189-205)."""
        idx = self.parameters.particles.single_unique_array_idx
        self.right_hand_side = self.mie_coefficients[idx, :] * self.initial_field_coefficients

    # ------------------------------------------------------------ operators
    def coupling_matrix_multiply(self, x: np.ndarray, idx: int | None = None):
        """W x (matrix_apply.py:29-63)."""
        return self.coupling_backend.multiply(x, idx)

    def master_matrix_multiply(self, value: np.ndarray, idx: int):
        """(I - T W) x (matrix_apply.py:66-120)."""
        return self.session.master_matvec(np.asarray(value).reshape(-1), channel=int(idx),
                                          precision=self.coupling_backend.precision)

    # ------------------------------------------------------------ solve
    def compute_scattered_field_coefficients(self, guess: np.ndarray | None = None):
        """Per-wavelength solve with warm start (scattered_field_coefficients.py:95-156)."""
        solver = self.numerics.solver
        if solver is None:
            raise RuntimeError("numerics.solver is not set")
        shape = self.right_hand_side.shape
        self.scattered_field_coefficients = np.zeros_like(self.initial_field_coefficients)
        self.scattered_field_err_codes = np.zeros(self.parameters.wavelengths_number)
        self.solve_seconds = []
        prev = None
        session = self.session
        for w in range(self.parameters.wavelengths_number):
            b = np.ascontiguousarray(self.right_hand_side[:, :, w]).ravel()
            if guess is not None:
                x0 = np.ascontiguousarray(guess[:, :, w]).ravel()
            elif solver.warm_start and prev is not None:
                x0 = prev
            else:
                x0 = b
            t0 = perf_counter()
            x, err = solver.run_device(session, w, b, x0)
            self.solve_seconds.append(perf_counter() - t0)
            self.scattered_field_coefficients[:, :, w] = x.reshape(shape[:2])
            self.scattered_field_err_codes[w] = err
            prev = x
