"""Krylov solver settings (mirror of yasfpy/solver.py:27-181).

With the ``hip`` coupling backend the solve runs entirely on the device
(``yasfhip_gmres_solve``): SciPy's GMRES algorithm with ``rtol=tolerance``,
``atol=0``, ``restart`` and ``maxiter`` counted in restart cycles.  Only
``solver_type="gmres"`` is provided.

``mixed_precision=True`` selects FP64 iterative refinement around GMRES whose
operator runs the pair GEMM in complex64 (``yasfhip_gmres_solve_refined``):
the convergence criterion ``||b - A x|| <= tolerance ||b||`` is still checked
with the FP64 operator; ``max_iter`` then bounds the restart cycles of each
inner solve and ``inner_tolerance`` the reduction one inner solve is asked for.
"""

from __future__ import annotations

import logging


class Solver:
    def __init__(self, solver_type: str = "gmres", tolerance: float = 1e-4, max_iter: float = 1e4,
                 restart: float = 1e2, preconditioner=None, *, warm_start: bool = True,
                 mixed_precision: bool = False, inner_tolerance: float = 1e-5,
                 max_refinements: int = 10):
        self.type = solver_type.lower()
        if self.type != "gmres":
            raise NotImplementedError(f"solver_type {solver_type!r}: only 'gmres' is provided")
        self.tolerance = float(tolerance)
        self.max_iter = int(max_iter)
        self.restart = int(restart)
        self.preconditioner = preconditioner
        self.warm_start = bool(warm_start)
        self.mixed_precision = bool(mixed_precision)
        self.inner_tolerance = float(inner_tolerance)
        self.max_refinements = int(max_refinements)
        self.log = logging.getLogger(self.__class__.__module__)
        self.last_iterations: list[int] = []
        self.last_residuals: list[float] = []
        self.last_histories: list = []

    def run_device(self, session, channel: int, b, x0):
        """GMRES on the device for one channel; returns (x, err_code)."""
        use_pc = False
        if self.preconditioner is not None:
            self.preconditioner.ensure_prepared(channel)
            use_pc = True
        if self.mixed_precision:
            x, info, _, iters, res, hist = session.gmres_refined(
                b, x0, channel=channel, rtol=self.tolerance, inner_rtol=self.inner_tolerance,
                restart=self.restart, inner_maxiter=self.max_iter,
                max_refinements=self.max_refinements, use_precond=use_pc)
        else:
            x, info, iters, res, hist = session.gmres(
                b, x0, channel=channel, rtol=self.tolerance, restart=self.restart,
                maxiter=self.max_iter, use_precond=use_pc)
        self.last_iterations.append(iters)
        self.last_residuals.append(res)
        self.last_histories.append(hist)
        self.log.debug("gmres channel %d: info=%d iterations=%d residual=%.3e", channel, info, iters, res)
        return x, info
