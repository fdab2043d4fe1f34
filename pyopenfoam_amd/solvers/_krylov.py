"""Shared driver of the device Krylov solves (pfm_pcg_solve / pfm_pbicgstab_solve)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.ldu_matrix import LduMatrix
from .residual import ConvergenceInfo, ResidualMonitor

_SOLVE = {"pcg": ("pfm_pcg_solve", "pfm_pcg_workspace_bytes"),
          "pbicgstab": ("pfm_pbicgstab_solve", "pfm_pbicgstab_workspace_bytes")}


class KrylovWorkspace:
    """Device scratch for one solver kind; grown on demand, reused across solves."""

    def __init__(self) -> None:
        self._buf: torch.Tensor | None = None

    def get(self, nbytes: int, device: torch.device) -> torch.Tensor:
        if self._buf is None or self._buf.numel() < nbytes or self._buf.device != device:
            self._buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        return self._buf


def device_solve(kind: str, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
                 monitor: ResidualMonitor, max_iter: int, rD: torch.Tensor | None,
                 workspace: KrylovWorkspace, *, transpose: bool = False,
                 x_out: torch.Tensor | None = None) -> tuple[torch.Tensor, ConvergenceInfo]:
    """Run one solve on the device; returns (x, ConvergenceInfo)."""
    dev, dt = matrix.device, matrix.dtype
    require_gpu(matrix.diag)
    b = source.to(device=dev, dtype=dt).contiguous()
    if b.shape[0] != matrix.n_cells:
        raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
    x = x_out if x_out is not None else torch.empty(matrix.n_cells, device=dev, dtype=dt)
    x.copy_(x0.to(device=dev, dtype=dt).reshape(-1))
    max_iter = int(max_iter)
    if max_iter < 0:
        raise ValueError("max_iter must be >= 0")
    fn_name, ws_name = _SOLVE[kind]
    L = _lib.lib()
    addr = matrix.addressing
    pk = matrix.packed(transpose)
    nbytes = int(getattr(L, ws_name)(addr.handle, max_iter))
    if nbytes < 0:
        raise _lib.PfmError("workspace query failed")
    work = workspace.get(nbytes, dev)
    prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), max_iter,
                           int(monitor.min_iter),
                           _lib.PFM_PRECOND_DIAG if rD is not None else _lib.PFM_PRECOND_NONE, 0)
    res = _lib.SolveResult()
    cap = 2 * max_iter + 2
    hist = np.zeros(cap, dtype=np.float64)
    d = matrix.diag.detach().contiguous()
    _lib.check(getattr(L, fn_name)(addr.handle, C.byref(prm), _lib.ptr(d),
                                   _lib.ptr(pk), _lib.ptr(rD), _lib.ptr(b), _lib.ptr(x),
                                   _lib.ptr(work), hist.ctypes.data, C.byref(res),
                                   current_stream(dev)), fn_name)
    monitor.load_history(hist[: min(res.iterations, cap)].tolist())
    info = monitor.build_info(converged=bool(res.converged))
    info.status = int(res.status)
    return x, info


def device_solve_mixed(matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
                       monitor: ResidualMonitor, max_iter: int, rD: torch.Tensor | None,
                       workspace: KrylovWorkspace, inner_rel_tol: float, max_outer: int,
                       *, transpose: bool = False) -> tuple[torch.Tensor, ConvergenceInfo]:
    """FP32-inner / FP64-outer PCG (pfm_pcg_mixed_solve)."""
    dev, dt = matrix.device, matrix.dtype
    require_gpu(matrix.diag)
    b = source.to(device=dev, dtype=dt).contiguous()
    if b.shape[0] != matrix.n_cells:
        raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
    x = x0.to(device=dev, dtype=dt).clone().contiguous()
    L = _lib.lib()
    addr = matrix.addressing
    pk = matrix.packed(transpose)
    pk32 = matrix.packed32(transpose)
    work = workspace.get(int(L.pfm_pcg_mixed_workspace_bytes(addr.handle, int(max_outer))), dev)
    prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), int(max_iter),
                           int(monitor.min_iter),
                           _lib.PFM_PRECOND_DIAG if rD is not None else _lib.PFM_PRECOND_NONE, 0)
    res = _lib.SolveResult()
    hist = np.zeros(int(max_outer) + 2)
    inner = C.c_int32()
    d = matrix.diag.detach().contiguous()
    _lib.check(L.pfm_pcg_mixed_solve(addr.handle, C.byref(prm), float(inner_rel_tol),
                                     int(max_outer), _lib.ptr(d),
                                     _lib.ptr(pk), _lib.ptr(pk32), _lib.ptr(rD), _lib.ptr(b),
                                     _lib.ptr(x), _lib.ptr(work), hist.ctypes.data, C.byref(res),
                                     C.byref(inner), current_stream(dev)), "pfm_pcg_mixed_solve")
    monitor.load_history(hist[: min(res.iterations, len(hist))].tolist())
    info = monitor.build_info(converged=bool(res.converged))
    info.status = int(res.status)
    info.inner_iterations = int(inner.value)
    return x, info
