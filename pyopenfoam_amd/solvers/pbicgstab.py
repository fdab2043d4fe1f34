"""Preconditioned BiCGStab (reference ``solvers/pbicgstab.py``).

Same constructor and semantics as ``PBiCGSTABSolver`` (:69-239), including
the reference's details: shadow residual r~ = r0, breakdown tests on rho,
r~.v, omega (1e-30), the half-step exit on ||s|| (x += alpha*y), the t.t ~ 0
branch, and omega = (t.z)/(t.t) with z = M^-1 s as the reference computes it.
The iteration runs on the device (``pfm_pbicgstab_solve``), five kernels per
iteration, no host synchronisation inside the loop.
"""

from __future__ import annotations

import torch

from ..core.ldu_matrix import LduMatrix
from ._krylov import KrylovWorkspace, device_solve
from .linear_solver import LinearSolverBase
from .preconditioners import make_preconditioner
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["PBiCGSTABSolver"]


class PBiCGSTABSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 1000,
                 min_iter: int = 0, verbose: bool = False, preconditioner: str = "DILU") -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._precond_type = preconditioner.upper()
        self._ws = KrylovWorkspace()

    def _make_preconditioner(self, matrix: LduMatrix):
        return make_preconditioner(self._precond_type, matrix, default_names=("DILU", "DIC", "none"))

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        precond = self._make_preconditioner(matrix)
        rD = precond.diagonal if precond is not None else None
        return device_solve("pbicgstab", matrix, source, x0, monitor, max_iter, rD, self._ws)
