"""Numerical settings (mirror of yasfpy/numerics.py:23-345, reduced to the solve path).

``Numerics(lmax, ..., solver, coupling_backend="hip")`` keeps the reference's
constructor names.  The translation table comes from the native library
(Wigner 3j by Racah's formula), replacing pywigxjpf and the on-disk pickle
cache of numerics.py:187-280.
"""

from __future__ import annotations

import logging

import numpy as np

from yasf_amd import _native


class Numerics:
    def __init__(self, lmax: int, sampling_points_number=None, polar_angles=None,
                 polar_weight_func=lambda x: x, azimuthal_angles=None, gpu: bool = True,
                 particle_distance_resolution: float = 10.0, solver=None,
                 coupling_backend: str = "hip", coupling_tile_size: int = 64,
                 coupling_near_field_radius=None):
        self.log = logging.getLogger(self.__class__.__module__)
        if not 1 <= int(lmax) <= 8:
            raise ValueError("lmax must be in [1, 8]")
        self.lmax = int(lmax)
        self.nmax = 2 * self.lmax * (self.lmax + 2)
        self.gpu = gpu
        self.particle_distance_resolution = particle_distance_resolution
        self.solver = solver
        self.coupling_backend = coupling_backend
        self.coupling_tile_size = int(coupling_tile_size)
        self.coupling_near_field_radius = coupling_near_field_radius
        # Sampling grid [n_phi, n_theta] as numerics.py:91-108 (used by the phase function).
        self.sampling_points_number = sampling_points_number
        if polar_angles is None and sampling_points_number is not None:
            spn = np.atleast_1d(sampling_points_number)
            if spn.size == 2:
                self.polar_angles_linspace = np.pi * polar_weight_func(np.linspace(0, 1, int(spn[1])))
                self.azimuthal_angles_linspace = 2 * np.pi * np.linspace(0, 1, int(spn[0]) + 1)[:-1]
                pa, aa = np.meshgrid(self.polar_angles_linspace, self.azimuthal_angles_linspace,
                                     indexing="xy")
                polar_angles, azimuthal_angles = pa.ravel(), aa.ravel()
        self.polar_angles = polar_angles
        self.azimuthal_angles = azimuthal_angles
        self.translation_ab5 = None

    def compute_translation_table(self, force_compute: bool = False):
        """ab5 for this lmax (numerics.py:187-280)."""
        if self.translation_ab5 is None or force_compute:
            self.translation_ab5 = _native.translation_table(self.lmax)
