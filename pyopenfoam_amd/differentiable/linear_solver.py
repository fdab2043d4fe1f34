"""Implicit-differentiation linear solve (reference ``differentiable/linear_solver.py``).

``DifferentiableLinearSolve.apply(A_diag, A_lower, A_upper, b, solver, owner,
neighbour, x0, tolerance=1e-6, max_iter=1000)`` has the reference's forward
(:47-96) and backward (:99-151): the adjoint system is the LDU with lower and
upper swapped, solved by the same solver object from a zero initial guess, and
  grad_diag = -lam*x, grad_lower = -lam[own]*x[nei], grad_upper = -lam[nei]*x[own],
  grad_b = lam,
computed by the ``pfm_adjoint_ldu_grads`` kernel.  No Krylov history is
stored; only x is saved.

``DifferentiablePoissonSolve`` is the fused form for the assembled pressure
equation: gamma_f -> (assembly kernel) -> solve -> x, whose backward solves
A^T lam = dJ/dx and scatters dJ/dgamma_f = -|S_f| delta_f (x_P - x_N)
(lam_P/V_P - lam_N/V_N) (and the Dirichlet-face term) in one kernel
(``pfm_adjoint_gamma_grad``) instead of materialising the three LDU gradients.
"""

from __future__ import annotations

from typing import Any

import torch

from .. import _lib
from ..core.device import current_stream
from ..core.ldu_matrix import LduMatrix

__all__ = ["DifferentiableLinearSolve", "DifferentiablePoissonSolve", "poisson_solve"]


class DifferentiableLinearSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx: Any, A_diag, A_lower, A_upper, b, solver, owner, neighbour, x0,
                tolerance: float = 1e-6, max_iter: int = 1000):
        n = A_diag.shape[0]
        m = LduMatrix(n, owner, neighbour, device=A_diag.device, dtype=A_diag.dtype)
        m.diag = A_diag.detach()
        m.lower = A_lower.detach()
        m.upper = A_upper.detach()
        x, iters, res = solver(m, b.detach(), x0.detach(), tolerance, max_iter)
        ctx.save_for_backward(A_diag, A_lower, A_upper, x)
        ctx.solver, ctx.tolerance, ctx.max_iter, ctx.n = solver, tolerance, max_iter, n
        ctx.addr = m.addressing
        ctx.topo = (owner, neighbour)
        return x

    @staticmethod
    def backward(ctx: Any, grad_output):
        A_diag, A_lower, A_upper, x = ctx.saved_tensors
        owner, neighbour = ctx.topo
        dev, dt = grad_output.device, grad_output.dtype
        mt = LduMatrix(ctx.n, owner, neighbour, device=dev, dtype=dt)
        mt._addr = ctx.addr
        mt.diag = A_diag.detach().clone()
        mt.lower = A_upper.detach().clone()
        mt.upper = A_lower.detach().clone()
        lam, _, _ = ctx.solver(mt, grad_output.contiguous(), torch.zeros(ctx.n, device=dev, dtype=dt),
                               ctx.tolerance, ctx.max_iter)
        gd = torch.empty(ctx.n, device=dev, dtype=dt)
        gl = torch.empty(A_lower.shape[0], device=dev, dtype=dt)
        gu = torch.empty(A_upper.shape[0], device=dev, dtype=dt)
        _lib.check(_lib.lib().pfm_adjoint_ldu_grads(ctx.addr.handle, _lib.ptr(lam), _lib.ptr(x),
                                                    _lib.ptr(gd), _lib.ptr(gl), _lib.ptr(gu),
                                                    current_stream(dev)), "pfm_adjoint_ldu_grads")
        return gd, gl, gu, lam, None, None, None, None, None, None


class DifferentiablePoissonSolve(torch.autograd.Function):
    """(gamma_internal, gamma_dirichlet, b) -> x for an assembled Laplacian system."""

    @staticmethod
    def forward(ctx: Any, gamma, dir_gamma, b, solver, mesh, dir_data, x0, tolerance, max_iter):
        from ..discretisation.fvm import LaplacianAssemble
        from ..core.fv_matrix import FvMatrix

        dev = gamma.device
        nfi = mesh.n_internal_faces
        m = FvMatrix(mesh.n_cells, mesh.owner[:nfi], mesh.neighbour, device=dev)
        if not hasattr(mesh, "_ldu_addr"):
            mesh._ldu_addr = m.addressing
        m._addr = mesh._ldu_addr
        cells, dmag, ddel = dir_data
        mag = mesh.face_areas_magnitude[:nfi].contiguous()
        delta = mesh.delta_coefficients[:nfi].contiguous()
        vol = mesh.cell_volumes
        diag, lower, upper = LaplacianAssemble.forward(_Ctx(), gamma.detach(),
                                                       dir_gamma.detach(), m._addr, mag, delta,
                                                       vol, cells, dmag, ddel)
        m._diag, m._lower, m._upper = diag, lower, upper
        x, iters, res = solver(m, b.detach(), x0.detach(), tolerance, max_iter)
        ctx.save_for_backward(x)
        ctx.matrix = m
        ctx.solver, ctx.tolerance, ctx.max_iter = solver, tolerance, max_iter
        ctx.geo = (mag, delta, vol, cells, dmag, ddel)
        ctx.iterations = iters
        return x

    @staticmethod
    def backward(ctx: Any, grad_output):
        (x,) = ctx.saved_tensors
        m = ctx.matrix
        dev, dt = grad_output.device, grad_output.dtype
        mt = m.transposed()
        lam, _, _ = ctx.solver(mt, grad_output.contiguous(), torch.zeros(m.n_cells, device=dev, dtype=dt),
                               ctx.tolerance, ctx.max_iter)
        mag, delta, vol, cells, dmag, ddel = ctx.geo
        nd = int(cells.numel()) if cells is not None else 0
        gg = torch.empty(m.n_internal_faces, device=dev, dtype=dt)
        gb = torch.empty(nd, device=dev, dtype=dt)
        _lib.check(_lib.lib().pfm_adjoint_gamma_grad(
            m.addressing.handle, _lib.ptr(lam), _lib.ptr(x), _lib.ptr(mag), _lib.ptr(delta),
            _lib.ptr(vol), nd, _lib.ptr(cells) if nd else None, _lib.ptr(dmag) if nd else None,
            _lib.ptr(ddel) if nd else None, _lib.ptr(gg), _lib.ptr(gb) if nd else None,
            current_stream(dev)), "pfm_adjoint_gamma_grad")
        return gg, gb, lam, None, None, None, None, None, None


class _Ctx:
    """Stand-in autograd context for calling LaplacianAssemble.forward directly."""

    def save_for_backward(self, *a):
        pass


def poisson_solve(gamma, dir_gamma, b, solver, mesh, dir_data, x0=None, tolerance=None,
                  max_iter=None):
    """x(gamma) for the assembled Poisson system, differentiable in gamma, dir_gamma, b."""
    if x0 is None:
        x0 = torch.zeros(mesh.n_cells, device=gamma.device, dtype=torch.float64)
    tol = solver.tolerance if tolerance is None else tolerance
    mi = solver.max_iter if max_iter is None else max_iter
    return DifferentiablePoissonSolve.apply(gamma, dir_gamma, b, solver, mesh, dir_data, x0, tol, mi)
