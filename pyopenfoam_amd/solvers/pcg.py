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

``preconditioner="GAMG"`` (additive; OpenFOAM's PCG with a GAMG
preconditioner) applies z = M^-1 r as one V-cycle of a GAMG hierarchy from
zero, in the same PCG loop and stopping rule.  ``gamg_options`` are
``GAMGSolver`` keywords for the hierarchy (default: pair agglomeration,
Galerkin coarse diagonal, 20 levels); with ``precision="mixed"`` the cycle's
finest ``fp32_levels`` (default 2) levels run in FP32 while the Krylov
vectors, dot products and the outer SpMV stay FP64.
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
                 max_outer: int = 50, gamg_options: dict | None = None,
                 fp32_levels: int = 2) -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._precond_type = preconditioner.upper()
        if precision not in ("double", "mixed"):
            raise ValueError(f"Unknown precision '{precision}'. Use 'double' or 'mixed'.")
        self._precision = precision
        self._inner_rel_tol = float(inner_rel_tol)
        self._max_outer = int(max_outer)
        self._fp32_levels = int(fp32_levels)
        self._gamg = None
        if self._precond_type == "GAMG":
            from .gamg import GAMGSolver
            opts = dict(agglomeration="pair", coarse="galerkin", max_levels=20)
            opts.update(gamg_options or {})
            self._gamg = GAMGSolver(**opts)
        self._ws = KrylovWorkspace()

    def _make_preconditioner(self, matrix: LduMatrix):
        return make_preconditioner(self._precond_type, matrix, default_names=("DIC", "DILU", "none"))

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        if self._gamg is not None:
            dev, dt = matrix.device, matrix.dtype
            b = source.to(device=dev, dtype=dt).contiguous()
            if b.shape[0] != matrix.n_cells:
                raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
            x = x0.to(device=dev, dtype=dt).clone().contiguous()
            if int(max_iter) < 0:
                raise ValueError("max_iter must be >= 0")
            nfp32 = self._fp32_levels if self._precision == "mixed" else 0
            return self._gamg.pcg_solve(matrix, b, x, monitor, int(max_iter), nfp32)
        precond = self._make_preconditioner(matrix)
        rD = precond.diagonal if precond is not None else None
        if self._precision == "mixed":
            return device_solve_mixed(matrix, source, x0, monitor, max_iter, rD, self._ws,
                                      self._inner_rel_tol, self._max_outer)
        return device_solve("pcg", matrix, source, x0, monitor, max_iter, rD, self._ws)
