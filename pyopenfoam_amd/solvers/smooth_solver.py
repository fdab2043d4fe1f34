"""SmoothSolver (reference ``solvers/smooth_solver.py``).

``SmoothSolver(tolerance, rel_tol, max_iter, min_iter, verbose,
smoother="GaussSeidel", n_sweeps=2, omega=1.0)`` (:40-95): the residual
monitor on b - A x0, then per iteration ``n_sweeps`` smoothing passes and the
monitor on b - A x.  ``GaussSeidel`` sweeps cells 0..n-1 with the latest
values and assigns x[c] = (b[c] - off) / diag[c] (:130-180); ``DIC`` / ``DILU``
run preconditioned Richardson x = x + omega apply_full(b - A x) (:198-232).
The whole loop runs on the device (``pfm_smooth_solve``): level-scheduled
sweeps, the reference's exact arithmetic, the device residual monitor.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.ldu_matrix import LduMatrix
from ._krylov import KrylovWorkspace
from .linear_solver import LinearSolverBase
from .preconditioners import DICPreconditioner, DILUPreconditioner
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["SmoothSolver"]

_SMOOTHERS = {"GaussSeidel": _lib.PFM_SMOOTH_GAUSS_SEIDEL, "DIC": _lib.PFM_SMOOTH_DIC,
              "DILU": _lib.PFM_SMOOTH_DILU}


class SmoothSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 1000,
                 min_iter: int = 0, verbose: bool = False, smoother: str = "GaussSeidel",
                 n_sweeps: int = 2, omega: float = 1.0) -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._smoother = smoother
        self._n_sweeps = int(n_sweeps)
        self._omega = float(omega)
        self._ws = KrylovWorkspace()

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        if self._smoother not in _SMOOTHERS:
            raise ValueError(f"Unknown smoother '{self._smoother}'. "
                             "Use 'GaussSeidel', 'DIC', or 'DILU'.")
        dev, dt = matrix.device, matrix.dtype
        require_gpu(matrix.diag)
        b = source.to(device=dev, dtype=dt).contiguous()
        if b.shape[0] != matrix.n_cells:
            raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
        x = x0.to(device=dev, dtype=dt).clone().contiguous()
        max_iter = int(max_iter)
        if max_iter < 0:
            raise ValueError("max_iter must be >= 0")
        kind = _SMOOTHERS[self._smoother]
        rD = None
        if kind == _lib.PFM_SMOOTH_DIC:
            rD = DICPreconditioner(matrix).diagonal
        elif kind == _lib.PFM_SMOOTH_DILU:
            rD = DILUPreconditioner(matrix).diagonal
        L = _lib.lib()
        addr = matrix.addressing
        work = self._ws.get(int(L.pfm_smooth_workspace_bytes(addr.handle, max_iter)), dev)
        prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), max_iter,
                               int(monitor.min_iter), 0, 0)
        res = _lib.SolveResult()
        hist = np.zeros(max_iter + 1)
        d, lo, up = matrix.coefficients()
        pk = matrix.packed(False)
        _lib.check(L.pfm_smooth_solve(addr.handle, C.byref(prm), kind, self._n_sweeps, self._omega,
                                      _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up),
                                      _lib.ptr(pk), _lib.ptr(rD), _lib.ptr(b),
                                      _lib.ptr(x), _lib.ptr(work), hist.ctypes.data, C.byref(res),
                                      current_stream(dev)), "pfm_smooth_solve")
        monitor.load_history(hist[: min(res.iterations, max_iter + 1)].tolist())
        info = monitor.build_info(converged=bool(res.converged))
        info.status = int(res.status)
        return x, info
