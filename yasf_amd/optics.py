"""Optical observables (mirror of yasfpy/optics.py:20-551, device-evaluated).

``Optics(simulation)`` keeps the reference's attribute names: ``c_ext``,
``c_sca``, ``albedo``, ``q_ext``, ``q_sca``, ``scattering_angles``,
``phase_function``, ``phase_function_3d``, the degrees of polarisation and
``phase_function_legendre_coefficients``.  The O(N^2 nmax^2) scattering cross
section (regular J-translation pair sum) and the O(directions x N x nmax) far
field run in ``libyasfhip.so`` (``yasfhip_cross_sections``,
``yasfhip_far_field``); only elementwise work on the direction grid stays on
the host.  The ``hip`` backend keeps no pairwise tables, so these replace the
table-based paths that the reference reserves for its dense backend
(optics.py:81-89).
"""

from __future__ import annotations

import logging

import numpy as np


def polarization_components(e_theta: np.ndarray, e_phi: np.ndarray):
    """I, DOP, DOLP, DOLQ, DOLU, DOCP from the far-field angle components
    (compute_polarization_components, cpu_numba.py:609-671)."""
    et2 = e_theta.real ** 2 + e_theta.imag ** 2
    ep2 = e_phi.real ** 2 + e_phi.imag ** 2
    cross = e_theta * np.conj(e_phi)
    intensity = et2 + ep2
    q = et2 - ep2
    u = -2 * cross.real
    v = 2 * cross.imag
    return (intensity, np.sqrt(q ** 2 + u ** 2 + v ** 2) / intensity,
            np.sqrt(q ** 2 + u ** 2) / intensity, -q / intensity, u / intensity, v / intensity)


class Optics:
    def __init__(self, simulation):
        self.simulation = simulation
        w = simulation.parameters.wavelengths_number
        self.c_ext = np.zeros(w, dtype=complex)
        self.c_sca = np.zeros(w, dtype=complex)
        self.log = logging.getLogger(self.__class__.__module__)

    def _coefficients(self, name: str, w: int) -> np.ndarray:
        arr = getattr(self.simulation, name)
        if arr is None:
            raise RuntimeError(f"simulation.{name} is not computed")
        return np.ascontiguousarray(arr[:, :, w]).ravel()

    def compute_cross_sections(self):
        """C_ext = -Re sum conj(a) f pi/|k|^2 and C_sca = |sum conj(f) W_J f| pi/|k|^2
        per wavelength (optics.py:55-156)."""
        sim = self.simulation
        session = sim.session
        c_ext = np.zeros(sim.parameters.wavelengths_number)
        c_sca = np.zeros(sim.parameters.wavelengths_number)
        for w in range(sim.parameters.wavelengths_number):
            c_ext[w], c_sca[w] = session.cross_sections(
                self._coefficients("initial_field_coefficients", w),
                self._coefficients("scattered_field_coefficients", w), channel=w)
        self.c_ext = c_ext
        self.c_sca = c_sca
        self.albedo = self.c_sca / self.c_ext

    def compute_efficiencies(self):
        """Cross sections over the cluster's geometric projection (optics.py:158-175)."""
        proj = self.simulation.parameters.particles.geometric_projection
        nmed = np.abs(self.simulation.parameters.medium_refractive_index)
        self.q_sca = self.c_sca / proj * nmed
        self.q_ext = self.c_ext / proj * nmed

    def compute_scattered_e_field_angle_components(self):
        """(E_theta, E_phi), ``[directions, wavelengths]`` (optics.py:177-266)."""
        sim = self.simulation
        polar = np.asarray(sim.numerics.polar_angles, dtype=float)
        azimuthal = np.asarray(sim.numerics.azimuthal_angles, dtype=float)
        if polar is None or azimuthal is None or polar.size == 0:
            raise RuntimeError("numerics has no sampling directions (sampling_points_number)")
        W = sim.parameters.wavelengths_number
        e_theta = np.zeros((polar.size, W), dtype=complex)
        e_phi = np.zeros_like(e_theta)
        for w in range(W):
            e_theta[:, w], e_phi[:, w] = sim.session.far_field(
                self._coefficients("scattered_field_coefficients", w), polar, azimuthal, channel=w)
        return e_theta, e_phi

    def compute_phase_funcition(self, legendre_coefficients_number: int = 15, c_and_b=False,
                                check_phase_function: bool = False):
        """Phase function I 4 pi / |k|^2 / C_sca, its azimuthal mean on an
        [n_phi, n_theta] grid and the degrees of polarisation (optics.py:269-551).
        The name keeps the reference's spelling; ``compute_phase_function`` is an alias."""
        if c_and_b:
            raise NotImplementedError("c_and_b bounds fitting is not provided")
        sim = self.simulation
        e_theta, e_phi = self.compute_scattered_e_field_angle_components()
        intensity, dop, dolp, dolq, dolu, docp = polarization_components(e_theta, e_phi)
        self.scattering_angles = np.asarray(sim.numerics.polar_angles, dtype=float)
        k = np.abs(np.asarray(sim.parameters.k_medium))
        self.phase_function_3d = intensity * 4 * np.pi / k[np.newaxis, :] ** 2 / self.c_sca[np.newaxis, :]
        if check_phase_function and not self._check_phase_function():
            raise AssertionError("the phase function integral deviates from 4 pi; use more angles")
        self.phase_function_legendre_coefficients = np.polynomial.legendre.legfit(
            np.cos(self.scattering_angles), self.phase_function_3d, legendre_coefficients_number)
        fields = {"degree_of_polarization": dop, "degree_of_linear_polarization": dolp,
                  "degree_of_linear_polarization_q": dolq, "degree_of_linear_polarization_u": dolu,
                  "degree_of_circular_polarization": docp}
        for name, val in fields.items():
            setattr(self, name + "_3d", val)
        spn = sim.numerics.sampling_points_number
        if spn is not None and np.size(spn) == 2:
            shape = (int(np.ravel(spn)[0]), int(np.ravel(spn)[1]), sim.parameters.wavelengths_number)
            self.phase_function = self.phase_function_3d.reshape(shape).mean(axis=0)
            for name, val in fields.items():
                setattr(self, name, val.reshape(shape).mean(axis=0))
            self.scattering_angles = self.scattering_angles.reshape(shape[:2])[0, :]
        else:
            self.phase_function = self.phase_function_3d
            for name, val in fields.items():
                setattr(self, name, val)

    compute_phase_function = compute_phase_funcition

    def _check_phase_function(self, precision: float = 1e-4) -> bool:
        """Trapezoidal integral of the phase function over the [n_phi, n_theta]
        grid equals 4 pi (optics.py:553-584)."""
        spn = self.simulation.numerics.sampling_points_number
        if spn is None or np.size(spn) != 2:
            return True
        n_phi, n_theta = (int(v) for v in np.ravel(spn))
        th = np.asarray(self.simulation.numerics.polar_angles).reshape(n_phi, n_theta)[0, :]
        p = self.phase_function_3d.reshape(n_phi, n_theta, -1)
        integrand = np.trapezoid(p * np.sin(th)[None, :, None], th, axis=1).mean(axis=0) * 2 * np.pi
        return bool(np.all(np.abs(integrand - 4 * np.pi) / (4 * np.pi) < precision))
