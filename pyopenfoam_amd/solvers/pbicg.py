"""Preconditioned bi-conjugate gradient (reference ``solvers/pbicg.py``).

``PBiCGSolver(tolerance, rel_tol, max_iter, min_iter, verbose,
preconditioner="DILU")`` (:55-105): the Lanczos form of BiCG (:107-220) with
the shadow system on A^T (``Ax_T``) and the same diagonal preconditioner for
both systems, breakdown (not converged) when |rho| or |z~.v| < 1e-30, the
residual monitor after every update.  As in the reference, one search
direction p serves both recurrences (r~ -= alpha A^T p, no separate shadow
direction), so the method is BiCG when A is symmetric and stalls on
non-symmetric systems such as the per-volume Laplacian of a distorted mesh;
the device loop keeps that behaviour.  The loop runs on the device
(``pfm_pbicg_solve``): one kernel for A p and A^T p together, one for the
updates, one for the monitor and the new direction.
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
from .preconditioners import make_preconditioner
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["PBiCGSolver"]


class PBiCGSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 1000,
                 min_iter: int = 0, verbose: bool = False, preconditioner: str = "DILU") -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._precond_type = preconditioner.upper()
        self._ws = KrylovWorkspace()

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        precond = make_preconditioner(self._precond_type, matrix,
                                      default_names=("DILU", "DIC", "none"))
        rD = precond.diagonal if precond is not None else None
        dev, dt = matrix.device, matrix.dtype
        require_gpu(matrix.diag)
        b = source.to(device=dev, dtype=dt).contiguous()
        if b.shape[0] != matrix.n_cells:
            raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
        x = x0.to(device=dev, dtype=dt).clone().contiguous()
        max_iter = int(max_iter)
        if max_iter < 0:
            raise ValueError("max_iter must be >= 0")
        L = _lib.lib()
        addr = matrix.addressing
        work = self._ws.get(int(L.pfm_pbicg_workspace_bytes(addr.handle, max_iter)), dev)
        prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), max_iter,
                               int(monitor.min_iter),
                               _lib.PFM_PRECOND_DIAG if rD is not None else _lib.PFM_PRECOND_NONE, 0)
        res = _lib.SolveResult()
        cap = 2 * max_iter + 2
        hist = np.zeros(cap)
        d, pk, pkt = matrix.diag.detach().contiguous(), matrix.packed(False), matrix.packed(True)
        _lib.check(L.pfm_pbicg_solve(addr.handle, C.byref(prm), _lib.ptr(d),
                                     _lib.ptr(pk), _lib.ptr(pkt),
                                     _lib.ptr(rD), _lib.ptr(b), _lib.ptr(x), _lib.ptr(work),
                                     hist.ctypes.data, C.byref(res), current_stream(dev)),
                   "pfm_pbicg_solve")
        monitor.load_history(hist[: min(res.iterations, cap)].tolist())
        info = monitor.build_info(converged=bool(res.converged))
        info.status = int(res.status)
        return x, info
