"""Preconditioned conjugate gradient (reference ``solvers/pcg.py``).

Same constructor and semantics as ``PCGSolver`` (:64-187): DIC (default),
DILU or no preconditioner with the reference's diagonal apply, x0 cloned,
breakdown when |p.Ap| < 1e-30 (returns unconverged), the residual monitor's
stopping rule after every update.  The whole iteration runs on the device
(``pfm_pcg_solve``): one fused SpMV + p.Ap kernel, one update kernel
(x, r, ||r||, r.z, beta) and one direction kernel per iteration, with no host
synchronisation inside the loop.

``precision="mixed"`` (additive keyword) solves with FP32 coefficients, FP32
SpMV and FP32 diagonal DIC apply in an inner PCG loop (dot products and norms
accumulated in FP64) inside FP64 iterative refinement: each refinement step
solves A e = r to ``inner_rel_tol`` and updates x += e, and the reference's
stopping rule is applied to the FP64 true residual ``b - A x``.  The residual
history then holds the true residual after each refinement step;
``last_info.inner_iterations`` counts the inner iterations.
"""

from __future__ import annotations

import torch

from ..core.ldu_matrix import LduMatrix
from ._krylov import KrylovWorkspace, device_solve, device_solve_mixed
from .linear_solver import LinearSolverBase
from .preconditioners import make_preconditioner
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["PCGSolver"]


class PCGSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 1000,
                 min_iter: int = 0, verbose: bool = False, preconditioner: str = "DIC",
                 precision: str = "double", inner_rel_tol: float = 1e-3,
                 max_outer: int = 50) -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._precond_type = preconditioner.upper()
        if precision not in ("double", "mixed"):
            raise ValueError(f"Unknown precision '{precision}'. Use 'double' or 'mixed'.")
        self._precision = precision
        self._inner_rel_tol = float(inner_rel_tol)
        self._max_outer = int(max_outer)
        self._ws = KrylovWorkspace()

    def _make_preconditioner(self, matrix: LduMatrix):
        return make_preconditioner(self._precond_type, matrix, default_names=("DIC", "DILU", "none"))

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        precond = self._make_preconditioner(matrix)
        rD = precond.diagonal if precond is not None else None
        if self._precision == "mixed":
            return device_solve_mixed(matrix, source, x0, monitor, max_iter, rD, self._ws,
                                      self._inner_rel_tol, self._max_outer)
        return device_solve("pcg", matrix, source, x0, monitor, max_iter, rD, self._ws)
