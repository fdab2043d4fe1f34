"""Batched PCG-DIC over many independent systems (configs of uneven sizes).

``BatchedLdu`` describes a batch (per-system local owner/neighbour arrays,
concatenated) and holds the native ``pfm_batch`` handle: the systems padded to
whole 64-row slices in one block-diagonal row-slice layout, and the work tiles
(chosen by the autotuner unless given).  ``BatchedPCGSolver.solve`` runs the
reference PCG (``solvers/pcg.py:97-187``, diagonal DIC) on every system at
once with per-system stopping (``solvers/residual.py:112-145``); iterations are
captured in a hipGraph and replayed until every system has stopped.
``batch_poisson_solve`` is the differentiable form for assembled Laplacian
systems: its backward runs the batched adjoint (A^T solve of every system) and
one fused face-gradient kernel over the whole batch.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..core.addressing import LduAddressing
from ..core.device import current_stream

__all__ = ["BatchedLdu", "BatchedPCGSolver", "BatchResult", "batch_poisson_solve"]


class BatchedLdu:
    def __init__(self, cell_offsets, face_offsets, owner_local, neighbour_local,
                 device: torch.device, tile_slices: int = 0, global_owner=None,
                 global_neighbour=None):
        co = _lib.host_i32(cell_offsets)
        fo = _lib.host_i32(face_offsets)
        ol = _lib.host_i32(owner_local)
        nl = _lib.host_i32(neighbour_local)
        self.n_systems = int(co.shape[0] - 1)
        self.cell_offsets = co
        self.face_offsets = fo
        self.device = torch.device(device)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().pfm_batch_create(self.n_systems, co.ctypes.data, fo.ctypes.data,
                                                   ol.ctypes.data, nl.ctypes.data, int(tile_slices),
                                                   C.byref(h)), "pfm_batch_create")
        self._h = h
        npad, nt, ts = C.c_int32(), C.c_int32(), C.c_int32()
        tuned = np.zeros(6, np.float32)
        _lib.check(_lib.lib().pfm_batch_info(h, C.byref(npad), C.byref(nt), C.byref(ts),
                                             tuned.ctypes.data))
        self.n_padded, self.n_tiles, self.tile_slices = npad.value, nt.value, ts.value
        self.autotune_us = {s: float(t) for s, t in zip((1, 2, 4, 8, 16, 32), tuned) if t > 0}
        self.n_cells = int(co[-1])
        self.n_faces = int(fo[-1])
        self._global = None
        if global_owner is not None:
            self._global = LduAddressing(self.n_cells, global_owner, global_neighbour, self.device)

    @property
    def handle(self):
        return self._h

    @property
    def global_addressing(self) -> LduAddressing:
        """Unpadded block-diagonal topology (assembly and face-gradient kernels)."""
        if self._global is None:
            raise RuntimeError("BatchedLdu was built without global owner/neighbour arrays")
        return self._global

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pfm_batch_destroy(h)
            except Exception:
                pass
            self._h = None


class BatchResult:
    def __init__(self, iterations, status, initial, final):
        self.iterations = iterations
        self.status = status
        self.initial_residual = initial
        self.final_residual = final

    @property
    def converged(self) -> np.ndarray:
        return self.status == _lib.PFM_SOLVE_CONVERGED


class BatchedPCGSolver:
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 1000,
                 min_iter: int = 0, preconditioner: str = "DIC", graph_chunk: int = 16):
        k = preconditioner.upper()
        if k not in ("DIC", "NONE"):
            raise ValueError(f"Unknown preconditioner '{preconditioner}'. Use 'DIC' or 'none'.")
        self.tolerance, self.rel_tol = tolerance, rel_tol
        self.max_iter, self.min_iter = int(max_iter), int(min_iter)
        self.precond = _lib.PFM_PRECOND_DIAG if k == "DIC" else _lib.PFM_PRECOND_NONE
        self.graph_chunk = int(graph_chunk)
        self._ws: torch.Tensor | None = None
        self.last: BatchResult | None = None

    def _work(self, batch: BatchedLdu) -> torch.Tensor:
        nb = int(_lib.lib().pfm_batch_workspace_bytes(batch.handle))
        if self._ws is None or self._ws.numel() < nb or self._ws.device != batch.device:
            self._ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=batch.device)
        return self._ws

    def solve(self, batch: BatchedLdu, diag, lower, upper, b, x0=None, *, transpose=False,
              graph_chunk: int | None = None):
        dev = batch.device
        f64 = dict(device=dev, dtype=torch.float64)
        d = diag.to(**f64).contiguous()
        lo = lower.to(**f64).contiguous()
        up = upper.to(**f64).contiguous()
        bb = b.to(**f64).contiguous()
        if d.shape[0] != batch.n_cells or bb.shape[0] != batch.n_cells:
            raise ValueError("diag/b must have the batch's total cell count")
        if lo.shape[0] != batch.n_faces or up.shape[0] != batch.n_faces:
            raise ValueError("lower/upper must have the batch's total face count")
        x = (torch.zeros(batch.n_cells, **f64) if x0 is None else x0.to(**f64).clone()).contiguous()
        ns = batch.n_systems
        it = np.zeros(ns, np.int32)
        st = np.zeros(ns, np.int32)
        r0 = np.zeros(ns)
        rf = np.zeros(ns)
        prm = _lib.SolveParams(float(self.tolerance), float(self.rel_tol), self.max_iter,
                               self.min_iter, self.precond, 0)
        chunk = self.graph_chunk if graph_chunk is None else int(graph_chunk)
        work = self._work(batch)
        _lib.check(_lib.lib().pfm_batch_pcg_solve(batch.handle, C.byref(prm), int(transpose),
                                                  _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up),
                                                  _lib.ptr(bb), _lib.ptr(x), _lib.ptr(work), chunk,
                                                  it.ctypes.data, st.ctypes.data, r0.ctypes.data,
                                                  rf.ctypes.data, current_stream(dev)),
                   "pfm_batch_pcg_solve")
        self.last = BatchResult(it, st, r0, rf)
        return x, self.last


class _BatchPoisson(torch.autograd.Function):
    @staticmethod
    def forward(ctx, gamma, dir_gamma, b, solver, batch, geo):
        from ..discretisation.fvm import LaplacianAssemble
        from ..differentiable.linear_solver import _Ctx

        addr = batch.global_addressing
        mag, delta, vol, cells, dmag, ddel = geo
        diag, lower, upper = LaplacianAssemble.forward(_Ctx(), gamma.detach(), dir_gamma.detach(),
                                                       addr, mag, delta, vol, cells, dmag, ddel)
        x, res = solver.solve(batch, diag, lower, upper, b.detach())
        ctx.save_for_backward(x)
        ctx.coeffs = (diag, lower, upper)
        ctx.solver, ctx.batch, ctx.geo = solver, batch, geo
        ctx.result = res
        return x

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        diag, lower, upper = ctx.coeffs
        lam, _ = ctx.solver.solve(ctx.batch, diag, lower, upper, g.contiguous(), transpose=True)
        mag, delta, vol, cells, dmag, ddel = ctx.geo
        addr = ctx.batch.global_addressing
        nd = int(cells.numel())
        gg = torch.empty(addr.n_faces, dtype=torch.float64, device=x.device)
        gb = torch.empty(nd, dtype=torch.float64, device=x.device)
        _lib.check(_lib.lib().pfm_adjoint_gamma_grad(
            addr.handle, _lib.ptr(lam), _lib.ptr(x), _lib.ptr(mag), _lib.ptr(delta), _lib.ptr(vol),
            nd, _lib.ptr(cells), _lib.ptr(dmag), _lib.ptr(ddel), _lib.ptr(gg), _lib.ptr(gb),
            current_stream(x.device)), "pfm_adjoint_gamma_grad")
        return gg, gb, lam, None, None, None


def batch_poisson_solve(gamma, dir_gamma, b, solver: BatchedPCGSolver, batch: BatchedLdu, geo):
    """x(gamma) for every system of an assembled batch; differentiable in gamma,
    dir_gamma and b.  ``geo`` = (|S_f|, delta_f, V, dirichlet cells int32,
    dirichlet |S_b|, dirichlet delta_b) as device tensors of the concatenation."""
    return _BatchPoisson.apply(gamma, dir_gamma, b, solver, batch, geo)
