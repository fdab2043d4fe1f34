"""GAMG solver (reference ``solvers/gamg.py``); device V-cycle in progress."""

from __future__ import annotations

from .linear_solver import LinearSolverBase

__all__ = ["GAMGSolver"]


class GAMGSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 100,
                 min_iter: int = 0, verbose: bool = False, n_pre_smooth: int = 2,
                 n_post_smooth: int = 2, max_levels: int = 10, min_cells_coarse: int = 10,
                 smoother: str = "PCG", agglomeration: str = "greedy") -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._n_pre_smooth = n_pre_smooth
        self._n_post_smooth = n_post_smooth
        self._max_levels = max_levels
        self._min_cells_coarse = min_cells_coarse
        self._smoother_type = smoother.upper()
        self._agglomeration = agglomeration

    def _solve(self, matrix, source, x0, monitor, max_iter):
        raise NotImplementedError("GAMG device V-cycle is not built yet")
