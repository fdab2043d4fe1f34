"""Diagonal solver (reference ``solvers/diagonal.py``).

``DiagonalSolver`` (:20-85) solves explicit systems: x = b / max(|diag|,
1e-30) in one pass, the off-diagonal coefficients ignored, and reports the
residual ||b - diag x|| as its single history entry (converged).  One device
kernel (``pfm_diagonal_solve``).
"""

from __future__ import annotations

import ctypes as C

import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.ldu_matrix import LduMatrix
from .linear_solver import LinearSolverBase
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["DiagonalSolver"]


class DiagonalSolver(LinearSolverBase):
    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        dev, dt = matrix.device, matrix.dtype
        require_gpu(matrix.diag)
        b = source.to(device=dev, dtype=dt).contiguous()
        if b.shape[0] != matrix.n_cells:
            raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
        x = torch.empty_like(b)
        work = torch.empty(matrix.n_cells // 256 + 3, device=dev, dtype=torch.float64)
        res = C.c_double()
        if matrix.n_cells:
            d = matrix.diag.detach().contiguous()
            _lib.check(_lib.lib().pfm_diagonal_solve(matrix.addressing.handle, _lib.ptr(d),
                                                     _lib.ptr(b), _lib.ptr(x), _lib.ptr(work),
                                                     C.byref(res), current_stream(dev)),
                       "pfm_diagonal_solve")
        monitor.load_history([res.value])
        return x, monitor.build_info(converged=True)
