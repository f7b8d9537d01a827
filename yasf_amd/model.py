"""Physical-model containers mirroring the reference's public objects.

``Particles`` (yasfpy/particles.py:24-247), ``InitialField``
(yasfpy/initial_field.py:14-128) and ``Parameters`` (yasfpy/parameters.py:15-132),
reduced to what the coupling/solve path consumes: positions, radii,
refractive indices, unique (radius, index) pairs, the incident-field
description and the medium wave numbers.
"""

from __future__ import annotations

import logging

import numpy as np


class Particles:
    """Spheres: positions ``(N, 3)``, radii ``(N,)``, complex refractive indices ``(N,)``
    (or a two-column ``[n, k]`` array, yasfpy/particles.py:56-79)."""

    def __init__(self, position, r, refractive_index, refractive_index_table=None,
                 shape_type: str = "sphere"):
        self.log = logging.getLogger(self.__class__.__module__)
        self.position = np.ascontiguousarray(position, dtype=float)
        self.r = np.asarray(r, dtype=float).ravel()
        n = np.asarray(refractive_index)
        if refractive_index_table is not None:
            raise NotImplementedError("refractive-index tables are not supported; pass n + ik")
        if n.ndim > 2 or (n.ndim == 2 and n.shape[1] > 2):
            raise ValueError("Refractive index should be either an integer array, complex array, "
                             "or a two column float matrix!")
        if n.ndim == 2 and n.shape[1] == 2:
            n = n[:, 0] + 1j * n[:, 1]
        self.refractive_index = np.asarray(n, dtype=complex).ravel()
        self.type = shape_type
        self.number = self.r.shape[0]
        if self.position.shape != (self.number, 3) or self.refractive_index.size != self.number:
            raise ValueError("position, r and refractive_index must describe the same particles")
        self._setup()

    def _setup(self):
        # unique (radius, index) pairs -> per-particle index (particles.py:114-128)
        stacked = np.column_stack((self.r, self.refractive_index.real, self.refractive_index.imag))
        uniq, inverse = np.unique(stacked, axis=0, return_inverse=True)
        self.unique_radius_index_pairs = np.column_stack((uniq[:, 0], uniq[:, 1] + 1j * uniq[:, 2]))
        self.single_unique_array_idx = inverse.ravel()
        self.num_unique_pairs = uniq.shape[0]
        # volume-equivalent projection (particles.py:161-165)
        self.geometric_projection = np.pi * np.power(np.sum(self.r ** 3), 2 / 3)


class InitialField:
    """Incident field (yasfpy/initial_field.py:17-128).  Plane waves are used
    when ``beam_width`` is 0 or infinite (initial_field_coefficients.py:63-75)."""

    def __init__(self, beam_width=0, focal_point=(0.0, 0.0, 0.0), field_type: str = "gaussian",
                 amplitude: float = 1.0, polar_angle: float = 0.0, azimuthal_angle: float = 0.0,
                 polarization="TE"):
        self.log = logging.getLogger(self.__class__.__module__)
        self.field_type = field_type
        self.amplitude = amplitude
        self.polar_angle = polar_angle
        self.azimuthal_angle = azimuthal_angle
        self.beam_width = beam_width
        self.focal_point = np.asarray(focal_point, dtype=float)
        self.set_polarization(polarization)

    def set_polarization(self, polarization):
        self.polarization = polarization
        p = polarization.lower() if isinstance(polarization, str) else polarization
        if p in ("unp", 0):
            self.pol = 0
        elif p in ("te", 1):
            self.pol = 1
        elif p in ("tm", 2):
            self.pol = 2
        else:
            self.pol = 0
            self.log.warning("%s is not a valid polarization type; using unpolarized", polarization)
        self.normal_incidence = bool(np.abs(np.sin(self.polar_angle)) < 1e-5)

    @property
    def is_plane_wave(self) -> bool:
        return not (np.isfinite(self.beam_width) and self.beam_width > 0)


class Parameters:
    """Wavelengths, medium index and derived wave numbers (yasfpy/parameters.py:22-100):
    ``omega = 2 pi / lambda``, ``k_medium = omega * n_medium``,
    ``k_particle = outer(n_particle, omega)``."""

    def __init__(self, wavelength, medium_refractive_index, particles: Particles,
                 initial_field: InitialField):
        self.wavelength = np.atleast_1d(np.asarray(wavelength, dtype=float))
        self.medium_refractive_index = np.atleast_1d(np.asarray(medium_refractive_index))
        self.wavelengths_number = self.wavelength.size
        self.particles = particles
        self.initial_field = initial_field
        self.omega = 2 * np.pi / self.wavelength
        self.k_medium = self.omega * self.medium_refractive_index
        self.k_particle = np.outer(particles.refractive_index, self.omega)
