"""ctypes binding of ``libyasfhip.so`` (the C ABI declared in include/yasfhip.h).

The library is required: importing a function that needs it raises
``NativeLibraryError`` when it is missing, and every device call raises when
no GPU is usable.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libyasfhip.so"

FP64 = 0
FP32 = 1
FP16X2 = 2  # pair GEMM on the fp16 matrix cores, operands as two fp16 terms (complex64 accuracy)
FP16 = 3    # one fp16 term (relative error ~1e-4)

MATVEC_AUTO = 0
MATVEC_ON_THE_FLY = 1
MATVEC_TABLE = 2
MATVEC_SLAB = 3

CSCA_FAR_FIELD = 0  # C_sca by quadrature of the far field (default)
CSCA_PAIR_SUM = 1   # C_sca by the regular (J) pair sum of cpu_numba.py:410-457

_lib = None
_lock = threading.Lock()


class NativeLibraryError(RuntimeError):
    """libyasfhip.so is missing or a call into it failed."""


class DeviceError(NativeLibraryError):
    """A HIP call inside the library failed (no GPU, out of memory, ...)."""


def hip_runtimes_in_maps(maps_text: str) -> list[str]:
    """Distinct paths of HIP runtime libraries (libamdhip64) in the text of
    /proc/<pid>/maps."""
    paths = []
    for line in maps_text.splitlines():
        fields = line.split()
        if len(fields) >= 6 and "libamdhip64" in Path(fields[-1]).name and fields[-1] not in paths:
            paths.append(fields[-1])
    return paths


def hip_runtimes() -> list[str]:
    """HIP runtime libraries mapped into this process."""
    try:
        return hip_runtimes_in_maps(Path("/proc/self/maps").read_text())
    except OSError:
        return []


def _load():
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise NativeLibraryError(
                f"{LIB_PATH} is not built; run `python -m yasf_amd.build` (hipcc, gfx950)")
        # One HIP runtime per process.  torch ships its own libamdhip64 (same
        # SONAME as ROCm's) and its libraries ask for it by file name: loaded
        # first, it also serves this library; loaded after this library, it
        # would come in as a second runtime that cannot see the first one's
        # device memory.  So torch, where installed, is imported first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        i = ctypes.c_int
        d = ctypes.c_double
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        llp = ctypes.POINTER(ctypes.c_longlong)
        sig = {
            "yasfhip_last_error": ([], ctypes.c_char_p),
            "yasfhip_translation_table": ([i, P], i),
            "yasfhip_plan_stats": ([i, llp, llp], i),
            "yasfhip_create": ([P, i, i, P, i, ctypes.POINTER(P)], i),
            "yasfhip_destroy": ([P], i),
            "yasfhip_set_tmatrix": ([P, P], i),
            "yasfhip_coupling_matvec": ([P, i, P, P, i], i),
            "yasfhip_master_matvec": ([P, i, P, P, i], i),
            "yasfhip_gmres_solve": ([P, i, P, P, d, i, i, i, ip, ip, dp, P, i], i),
            "yasfhip_gmres_solve_device": ([P, i, P, P, d, i, i, i, ip, ip, dp, P, i], i),
            "yasfhip_gmres_solve_refined": ([P, i, P, P, d, d, i, i, i, i, ip, ip, ip, dp, P, i], i),
            "yasfhip_cross_sections": ([P, i, P, P, dp, dp], i),
            "yasfhip_far_field": ([P, i, P, i, P, P, P, P], i),
            "yasfhip_precond_build": ([P, i, P, i], i),
            "yasfhip_precond_apply": ([P, P, P], i),
            "yasfhip_set_precond_factor_precision": ([P, i], i),
            "yasfhip_time_matvec": ([P, i, i, i, dp, dp], i),
            "yasfhip_server_create": ([ctypes.c_size_t, i, d, i, ctypes.POINTER(P)], i),
            "yasfhip_server_destroy": ([P], i),
            "yasfhip_server_submit": ([P, P, i, i, d, P, P, d, i, i, ctypes.c_longlong, llp], i),
            "yasfhip_server_step": ([P, ip, ip, llp], i),
            "yasfhip_server_poll": ([P, llp, P, ctypes.c_longlong, ip, ip, ip, dp, dp, dp, dp], i),
            "yasfhip_server_autotune": ([P, i, i, i, ip, ip, P], i),
            "yasfhip_server_set_tile_shape": ([P, i, i, i], i),
            "yasfhip_server_set_inner_rtol": ([P, d], i),
            "yasfhip_server_set_table_policy": ([P, i, i], i),
            "yasfhip_set_slab_tiles": ([P, i], i),
            "yasfhip_set_near_field_radius": ([P, d], i),
            "yasfhip_set_table_shape": ([P, i, i], i),
            "yasfhip_mie_coefficients": ([i, i, P, P, d, P], i),
            "yasfhip_server_submit_spheres": ([P, P, P, P, i, i, d, d, d, i, d, P, d, i, i, ctypes.c_longlong,
                                               llp], i),
            "yasfhip_planewave_coefficients": ([i, i, P, d, d, d, i, d, P, P], i),
            "yasfhip_set_matvec_mode": ([P, i], i),
            "yasfhip_tune_matvec": ([P, i, i, i, ip, P], i),
            "yasfhip_coupling_matvec_noalloc": ([P, i, i, d, P, P, P], i),
            "yasfhip_set_scattering_method": ([P, i], i),
            "yasfhip_set_inner_precision": ([P, i], i),
            "yasfhip_server_set_inner_precision": ([P, i], i),
            "yasfhip_server_set_scattering_method": ([P, i], i),
            "yasfhip_server_set_residual_pool": ([P, ctypes.c_size_t], i),
            "yasfhip_server_set_iterations_per_step": ([P, i], i),
            "yasfhip_server_residual_forms": ([P, llp], i),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
        return lib


def library():
    """The loaded ctypes library (loads it on first use)."""
    return _load()


def exported_symbols() -> list[str]:
    return [
        "yasfhip_last_error", "yasfhip_translation_table", "yasfhip_plan_stats", "yasfhip_create",
        "yasfhip_destroy", "yasfhip_set_tmatrix", "yasfhip_coupling_matvec", "yasfhip_master_matvec",
        "yasfhip_gmres_solve", "yasfhip_gmres_solve_device", "yasfhip_gmres_solve_refined", "yasfhip_cross_sections",
        "yasfhip_far_field", "yasfhip_precond_build", "yasfhip_precond_apply",
        "yasfhip_time_matvec", "yasfhip_set_matvec_mode", "yasfhip_tune_matvec", "yasfhip_server_create",
        "yasfhip_server_destroy", "yasfhip_server_submit", "yasfhip_server_step", "yasfhip_server_poll",
        "yasfhip_server_autotune", "yasfhip_server_set_tile_shape", "yasfhip_server_set_table_policy",
        "yasfhip_server_set_inner_rtol",
        "yasfhip_set_slab_tiles", "yasfhip_set_near_field_radius", "yasfhip_mie_coefficients",
        "yasfhip_planewave_coefficients", "yasfhip_server_submit_spheres", "yasfhip_set_table_shape",
        "yasfhip_set_precond_factor_precision", "yasfhip_coupling_matvec_noalloc",
        "yasfhip_set_scattering_method", "yasfhip_server_set_scattering_method", "yasfhip_set_inner_precision",
        "yasfhip_server_set_inner_precision", "yasfhip_server_set_residual_pool", "yasfhip_server_residual_forms",
        "yasfhip_server_set_iterations_per_step",
    ]


def _check(status: int):
    if status == 0:
        return
    msg = _load().yasfhip_last_error().decode(errors="replace")
    if status == 1:
        raise ValueError(msg)
    if status == 3:
        raise RuntimeError(msg)
    if status == 2:
        raise DeviceError(msg)
    raise NativeLibraryError(msg)


def _ptr(a: np.ndarray):
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def translation_table(lmax: int) -> np.ndarray:
    """ab5[j1, j2, p] as complex128 ``(nmax, nmax, 2 lmax + 1)``."""
    nmax = 2 * lmax * (lmax + 2)
    out = np.zeros((nmax, nmax, 2 * lmax + 1), dtype=np.complex128)
    _check(_load().yasfhip_translation_table(int(lmax), _ptr(out)))
    return out


def plan_stats(lmax: int) -> tuple[int, int]:
    """(complex MACs per particle pair issued by the pair kernel, nnz(ab5))."""
    a = ctypes.c_longlong()
    b = ctypes.c_longlong()
    _check(_load().yasfhip_plan_stats(int(lmax), ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def mie_coefficients(lmax: int, radius, relative_index, k_medium: float) -> np.ndarray:
    """Mie T-matrix diagonal ``[N * nmax]`` on the device (one wavelength)."""
    r = np.ascontiguousarray(radius, dtype=np.float64).ravel()
    m = np.ascontiguousarray(np.broadcast_to(relative_index, r.shape), dtype=np.complex128)
    out = np.empty(r.size * 2 * lmax * (lmax + 2), dtype=np.complex128)
    _check(_load().yasfhip_mie_coefficients(int(lmax), r.size, _ptr(r), _ptr(m), float(k_medium), _ptr(out)))
    return out


def planewave_coefficients(lmax: int, positions, k_medium: float, polar: float, azimuthal: float,
                           polarization: int = 1, amplitude: float = 1.0, focal_point=None) -> np.ndarray:
    """Plane-wave incident coefficients ``[N * nmax]`` on the device (one wavelength)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    f = None if focal_point is None else np.ascontiguousarray(focal_point, dtype=np.float64)
    out = np.empty(pos.shape[0] * 2 * lmax * (lmax + 2), dtype=np.complex128)
    _check(_load().yasfhip_planewave_coefficients(int(lmax), pos.shape[0], _ptr(pos), float(k_medium),
                                                  float(polar), float(azimuthal), int(polarization),
                                                  float(amplitude), _ptr(f) if f is not None else None,
                                                  _ptr(out)))
    return out


class Session:
    """One device session: geometry, expansion order and channels.

    Owns the device buffers; ``close()`` (or garbage collection) frees them.
    """

    def __init__(self, positions: np.ndarray, lmax: int, k_medium):
        lib = _load()
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("positions must have shape (N, 3)")
        k = np.atleast_1d(np.asarray(k_medium))
        if np.iscomplexobj(k):
            if np.any(k.imag != 0):
                raise NotImplementedError("the hip backend supports real medium wave numbers only")
            k = k.real
        k = np.ascontiguousarray(k, dtype=np.float64)
        self.N = pos.shape[0]
        self.lmax = int(lmax)
        self.nmax = 2 * self.lmax * (self.lmax + 2)
        self.n = self.N * self.nmax
        self.channels = k.size
        h = ctypes.c_void_p()
        _check(lib.yasfhip_create(_ptr(pos), self.N, self.lmax, _ptr(k), self.channels, ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.yasfhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _vec(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128).ravel()
        if x.size != self.n:
            raise ValueError(f"vector has {x.size} entries, expected {self.n}")
        return x

    def set_tmatrix(self, t: np.ndarray):
        t = np.ascontiguousarray(t, dtype=np.complex128).reshape(self.channels, self.n)
        _check(self._lib.yasfhip_set_tmatrix(self._h, _ptr(t)))

    def coupling_matvec(self, x, channel: int = 0, precision: int = FP64) -> np.ndarray:
        x = self._vec(x)
        y = np.empty_like(x)
        _check(self._lib.yasfhip_coupling_matvec(self._h, int(channel), _ptr(x), _ptr(y), int(precision)))
        return y

    def master_matvec(self, x, channel: int = 0, precision: int = FP64) -> np.ndarray:
        x = self._vec(x)
        y = np.empty_like(x)
        _check(self._lib.yasfhip_master_matvec(self._h, int(channel), _ptr(x), _ptr(y), int(precision)))
        return y

    def gmres(self, b, x0=None, channel: int = 0, rtol: float = 1e-6, restart: int = 20,
              maxiter: int = 1000, use_precond: bool = False, history_cap: int = 100000):
        """Returns (x, info, inner_iterations, relative_residual, pr_norm history)."""
        b = self._vec(b)
        x = self._vec(b if x0 is None else x0).copy()
        info = ctypes.c_int()
        iters = ctypes.c_int()
        res = ctypes.c_double()
        hist = np.zeros(max(1, int(history_cap)), dtype=np.float64)
        _check(self._lib.yasfhip_gmres_solve(
            self._h, int(channel), _ptr(b), _ptr(x), float(rtol), int(restart), int(maxiter),
            int(bool(use_precond)), ctypes.byref(info), ctypes.byref(iters), ctypes.byref(res),
            _ptr(hist), hist.size))
        return x, info.value, iters.value, res.value, hist[: min(iters.value, hist.size)].copy()

    def gmres_device(self, b, x, channel: int = 0, rtol: float = 1e-6, restart: int = 20,
                     maxiter: int = 1000, use_precond: bool = False, history_cap: int = 100000):
        """gmres on vectors already in device memory: b and x are contiguous
        complex128 torch tensors of n values on the session's device; x holds
        x0 on entry and the solution on exit (torch's current stream is
        synchronised first; the solution is complete on return).  Returns
        (info, inner_iterations, relative_residual, pr_norm history)."""
        import torch  # device vectors only; the rest of the module needs no torch

        runtimes = hip_runtimes()
        if len(runtimes) > 1:
            raise NativeLibraryError(
                "two HIP runtimes are loaded in this process (" + ", ".join(runtimes) + "): torch's device "
                "memory is not visible to libyasfhip.so; import torch before anything loads ROCm's libamdhip64")
        for name, v in (("b", b), ("x", x)):
            if v.dtype != torch.complex128 or not v.is_cuda or not v.is_contiguous() or v.numel() != self.n:
                raise ValueError(f"{name} must be a contiguous complex128 device tensor of {self.n} values")
        torch.cuda.current_stream(b.device).synchronize()  # torch's writes to b and x are done
        info = ctypes.c_int()
        iters = ctypes.c_int()
        res = ctypes.c_double()
        hist = np.zeros(max(1, int(history_cap)), dtype=np.float64)
        _check(self._lib.yasfhip_gmres_solve_device(
            self._h, int(channel), ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()), float(rtol),
            int(restart), int(maxiter), int(bool(use_precond)), ctypes.byref(info), ctypes.byref(iters),
            ctypes.byref(res), _ptr(hist), hist.size))
        return info.value, iters.value, res.value, hist[: min(iters.value, hist.size)].copy()

    def gmres_refined(self, b, x0=None, channel: int = 0, rtol: float = 1e-6, inner_rtol: float = 1e-5,
                      restart: int = 50, inner_maxiter: int = 1000, max_refinements: int = 10,
                      use_precond: bool = False):
        """FP64 iterative refinement around complex64-operator GMRES.

        Returns (x, info, refinements, inner_iterations, relative_residual,
        relative residual before the first and after each refinement)."""
        b = self._vec(b)
        x = self._vec(b if x0 is None else x0).copy()
        info = ctypes.c_int()
        steps = ctypes.c_int()
        iters = ctypes.c_int()
        res = ctypes.c_double()
        hist = np.zeros(int(max_refinements) + 1, dtype=np.float64)
        _check(self._lib.yasfhip_gmres_solve_refined(
            self._h, int(channel), _ptr(b), _ptr(x), float(rtol), float(inner_rtol), int(restart),
            int(inner_maxiter), int(max_refinements), int(bool(use_precond)), ctypes.byref(info),
            ctypes.byref(steps), ctypes.byref(iters), ctypes.byref(res), _ptr(hist), hist.size))
        return x, info.value, steps.value, iters.value, res.value, hist[: steps.value + 1].copy()

    def cross_sections(self, a_inc, f, channel: int = 0) -> tuple[float, float]:
        """(C_ext, C_sca) of one channel from incident and scattered coefficients."""
        a = self._vec(a_inc)
        fv = self._vec(f)
        ce = ctypes.c_double()
        cs = ctypes.c_double()
        _check(self._lib.yasfhip_cross_sections(self._h, int(channel), _ptr(a), _ptr(fv),
                                                ctypes.byref(ce), ctypes.byref(cs)))
        return ce.value, cs.value

    def far_field(self, f, polar, azimuthal, channel: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """(E_theta, E_phi) of the scattered far field at directions (polar[a], azimuthal[a])."""
        fv = self._vec(f)
        th = np.ascontiguousarray(np.ravel(polar), dtype=np.float64)
        ph = np.ascontiguousarray(np.ravel(azimuthal), dtype=np.float64)
        if th.size != ph.size or th.size == 0:
            raise ValueError("polar and azimuthal must be non-empty and of equal size")
        et = np.empty(th.size, dtype=np.complex128)
        ep = np.empty(th.size, dtype=np.complex128)
        _check(self._lib.yasfhip_far_field(self._h, int(channel), _ptr(fv), th.size, _ptr(th), _ptr(ph),
                                           _ptr(et), _ptr(ep)))
        return et, ep

    def precond_build(self, box_of_particle, channel: int = 0):
        box = np.ascontiguousarray(box_of_particle, dtype=np.int32)
        if box.size != self.N:
            raise ValueError("box_of_particle must have one entry per particle")
        nbox = int(box.max()) + 1 if box.size else 0
        _check(self._lib.yasfhip_precond_build(self._h, int(channel), _ptr(box), nbox))

    def set_precond_factor_precision(self, precision: int):
        """FP32 (default): the preconditioner's LU factors are stored in
        complex64 for the apply (half the bytes; FP64 arithmetic); FP64 keeps
        them in complex128.  Takes effect at the next precond_build."""
        _check(self._lib.yasfhip_set_precond_factor_precision(self._h, int(precision)))

    def precond_apply(self, r) -> np.ndarray:
        r = self._vec(r)
        z = np.empty_like(r)
        _check(self._lib.yasfhip_precond_apply(self._h, _ptr(r), _ptr(z)))
        return z

    def set_inner_precision(self, precision: int):
        """Operator of gmres_refined's inner GMRES: FP32 (default), FP16X2 or FP16."""
        _check(self._lib.yasfhip_set_inner_precision(self._h, int(precision)))

    def set_scattering_method(self, method: int):
        """CSCA_FAR_FIELD (default) or CSCA_PAIR_SUM for cross_sections."""
        _check(self._lib.yasfhip_set_scattering_method(self._h, int(method)))

    def set_matvec_mode(self, mode: int):
        """MATVEC_AUTO, MATVEC_ON_THE_FLY, MATVEC_TABLE (resident pair-factor
        table) or MATVEC_SLAB (the table built and streamed a slab of target
        tiles at a time inside every matvec)."""
        _check(self._lib.yasfhip_set_matvec_mode(self._h, int(mode)))

    def set_near_field_radius(self, radius):
        """Pairs farther apart than radius do not couple (0: none; None / inf: exact W)."""
        r = -1.0 if radius is None else float(radius)
        _check(self._lib.yasfhip_set_near_field_radius(self._h, r))

    def set_slab_tiles(self, tiles: int):
        """Target tiles per slab (0: as many as fit in a quarter of the free memory)."""
        _check(self._lib.yasfhip_set_slab_tiles(self._h, int(tiles)))

    def tune_matvec(self, channel: int = 0, precision: int = FP64, reps: int = 3):
        """Picks the table GEMM's shape; returns (index, ms per shape: 8 waves with
        2/4/8 column tiles, then 4 waves with 2/4/8; -1 where not offered).
        FP64 shape 2 runs 8 waves that compute their orders one at a time over
        jobs of up to 6 column tiles."""
        g = ctypes.c_int()
        ms = np.zeros(6)
        _check(self._lib.yasfhip_tune_matvec(self._h, int(channel), int(precision), int(reps),
                                             ctypes.byref(g), _ptr(ms)))
        return g.value, ms

    TABLE_SHAPES = ((8, 2), (8, 4), (8, 8), (4, 2), (4, 4), (4, 8))  # (waves, column tiles) per index

    def set_table_shape(self, shape: int, precision: int = FP64):
        """Fixes the table GEMM shape (index into TABLE_SHAPES) without timing."""
        _check(self._lib.yasfhip_set_table_shape(self._h, int(precision), int(shape)))

    def time_matvec(self, channel: int = 0, precision: int = FP64, reps: int = 5) -> tuple[float, float]:
        """Mean (pair-kernel ms, whole-matvec ms) over `reps` device-resident matvecs."""
        a = ctypes.c_double()
        b = ctypes.c_double()
        _check(self._lib.yasfhip_time_matvec(self._h, int(channel), int(precision), int(reps),
                                             ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Server:
    """Batched serving engine (yasfhip_server_*): many independent solves
    advanced together, memory-admitted into a device arena."""

    def __init__(self, memory_bytes: int = 0, restart: int = 50, inner_rtol: float = 1e-4,
                 use_graph: bool = True):
        lib = _load()
        h = ctypes.c_void_p()
        _check(lib.yasfhip_server_create(int(memory_bytes), int(restart), float(inner_rtol),
                                         int(bool(use_graph)), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self._sizes: dict[int, int] = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.yasfhip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, positions, lmax: int, k_medium: float, t, a_inc, rtol: float = 1e-6,
               max_cycles: int = 100, observables: bool = True, operator_key: int = 0) -> int:
        """Queues one solve; requests with the same nonzero operator_key share
        the operator (geometry, lmax, wave number, T-matrix) on the device."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        n = pos.shape[0] * 2 * lmax * (lmax + 2)
        t = np.ascontiguousarray(t, dtype=np.complex128).ravel()
        a = np.ascontiguousarray(a_inc, dtype=np.complex128).ravel()
        if t.size != n or a.size != n:
            raise ValueError(f"t and a_inc need {n} entries")
        rid = ctypes.c_longlong()
        _check(self._lib.yasfhip_server_submit(self._h, _ptr(pos), pos.shape[0], int(lmax), float(k_medium),
                                               _ptr(t), _ptr(a), float(rtol), int(max_cycles),
                                               int(bool(observables)), int(operator_key), ctypes.byref(rid)))
        self._sizes[rid.value] = n
        return rid.value

    def submit_spheres(self, positions, radius, relative_index, lmax: int, k_medium: float, polar: float,
                       azimuthal: float, polarization: int = 1, amplitude: float = 1.0, focal_point=None,
                       rtol: float = 1e-6, max_cycles: int = 100, observables: bool = True,
                       operator_key: int = 0) -> int:
        """Queues one solve whose T-matrix (Mie, from radius and relative
        refractive index per sphere) and plane-wave coefficients are set up on
        the device at admission."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("positions must have shape (N, 3)")
        N = pos.shape[0]
        r = np.ascontiguousarray(np.broadcast_to(radius, (N,)), dtype=np.float64)
        m = np.ascontiguousarray(np.broadcast_to(relative_index, (N,)), dtype=np.complex128)
        f = None if focal_point is None else np.ascontiguousarray(focal_point, dtype=np.float64)
        rid = ctypes.c_longlong()
        _check(self._lib.yasfhip_server_submit_spheres(
            self._h, _ptr(pos), _ptr(r), _ptr(m), N, int(lmax), float(k_medium), float(polar), float(azimuthal),
            int(polarization), float(amplitude), _ptr(f) if f is not None else None, float(rtol),
            int(max_cycles), int(bool(observables)), int(operator_key), ctypes.byref(rid)))
        self._sizes[rid.value] = N * 2 * lmax * (lmax + 2)
        return rid.value

    def step(self) -> tuple[int, int, int]:
        """(active, queued, completed so far) after one engine step."""
        a = ctypes.c_int()
        q = ctypes.c_int()
        c = ctypes.c_longlong()
        _check(self._lib.yasfhip_server_step(self._h, ctypes.byref(a), ctypes.byref(q), ctypes.byref(c)))
        return a.value, q.value, c.value

    def poll(self, with_solution: bool = True):
        """Next completed request as a dict, or None."""
        rid = ctypes.c_longlong()
        info = ctypes.c_int()
        cycles = ctypes.c_int()
        iters = ctypes.c_int()
        res = ctypes.c_double()
        ce = ctypes.c_double()
        cs = ctypes.c_double()
        lat = ctypes.c_double()
        # the next completion's size is unknown here: pass the largest pending buffer
        cap = max(self._sizes.values()) if (with_solution and self._sizes) else 0
        x = np.empty(max(cap, 1), dtype=np.complex128)
        _check(self._lib.yasfhip_server_poll(self._h, ctypes.byref(rid), _ptr(x) if cap else None, cap,
                                             ctypes.byref(info), ctypes.byref(cycles), ctypes.byref(iters),
                                             ctypes.byref(res), ctypes.byref(ce), ctypes.byref(cs),
                                             ctypes.byref(lat)))
        if rid.value < 0:
            return None
        n = self._sizes.pop(rid.value, 0)
        return {"id": rid.value, "x": x[:n].copy() if cap else None, "info": info.value,
                "cycles": cycles.value, "iterations": iters.value, "residual": res.value,
                "c_ext": ce.value, "c_sca": cs.value, "latency_s": lat.value}

    TILE_SHAPES = ((2, 2), (4, 2), (8, 2), (4, 4), (8, 4), (8, 8))

    def autotune(self, lmax: int, particles: int = 600, reps: int = 5):
        """Times every tile shape on a synthetic request of `particles` spheres
        and keeps the fastest for lmax and that request size class (fewer than
        1000 particles, or at least 1000).  Returns ((waves, column tiles)
        chosen, ms per candidate of TILE_SHAPES; -1 where not usable)."""
        w = ctypes.c_int()
        ct = ctypes.c_int()
        ms = np.zeros(len(self.TILE_SHAPES))
        _check(self._lib.yasfhip_server_autotune(self._h, int(lmax), int(particles), int(reps),
                                                 ctypes.byref(w), ctypes.byref(ct), _ptr(ms)))
        return (w.value, ct.value), ms

    def set_tile_shape(self, lmax: int, waves: int, column_tiles: int = 2):
        _check(self._lib.yasfhip_server_set_tile_shape(self._h, int(lmax), int(waves), int(column_tiles)))

    TABLES_ON_THE_FLY, TABLES_AUTO, TABLES_SLABS = 0, 1, 2

    def set_inner_rtol(self, inner_rtol: float):
        """Reduction asked of each restart cycle started from now on (at least)."""
        _check(self._lib.yasfhip_server_set_inner_rtol(self._h, float(inner_rtol)))

    def set_inner_precision(self, precision: int):
        """Pair GEMM of the inner iterations of operators admitted from now on:
        FP32 (default), FP16X2 or FP16.  Needs an idle server."""
        _check(self._lib.yasfhip_server_set_inner_precision(self._h, int(precision)))

    def set_scattering_method(self, method: int):
        """CSCA_FAR_FIELD (default) or CSCA_PAIR_SUM for the completions' C_sca."""
        _check(self._lib.yasfhip_server_set_scattering_method(self._h, int(method)))

    def set_table_policy(self, policy: int, slab_tiles: int = 0):
        """TABLES_AUTO (default): resident pair tables where they fit, else
        slabs; TABLES_SLABS: slabs only; TABLES_ON_THE_FLY: neither.
        slab_tiles > 0 fixes the FP64 target tiles per slab."""
        _check(self._lib.yasfhip_server_set_table_policy(self._h, int(policy), int(slab_tiles)))

    def set_residual_pool(self, nbytes: int):
        """Reserves an FP64 slab of nbytes from the arena for the FP64
        residuals of operators without an FP64 table or slab of their own (0
        releases it).  Arenas of 32 GiB or more start with memory_bytes / 32."""
        _check(self._lib.yasfhip_server_set_residual_pool(self._h, int(nbytes)))

    def set_iterations_per_step(self, iterations: int):
        """Batched Arnoldi iterations per step (1..16) between host synchronisations."""
        _check(self._lib.yasfhip_server_set_iterations_per_step(self._h, int(iterations)))

    def residual_forms(self) -> dict:
        """FP64 residual pair GEMMs run so far, by form."""
        forms = (ctypes.c_longlong * 4)()
        _check(self._lib.yasfhip_server_residual_forms(self._h, forms))
        return dict(zip(("table", "slab", "pool", "on_the_fly"), (int(v) for v in forms)))
