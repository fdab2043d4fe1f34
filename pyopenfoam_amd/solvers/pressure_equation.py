"""The pressure-equation step around the solve (reference ``solvers/pressure_equation.py``).

Same functions, arguments and results as the reference:

* ``assemble_pressure_equation(phiHbyA, A_p, mesh, face_weights=None)``
  (:49-152) -> FvMatrix of laplacian(1/A_p, p) = div(phiHbyA), per unit
  volume: ``pfm_pressure_assemble`` (face coefficients, then one pass per
  row for the diagonal and the source).
* ``solve_pressure_equation(p_eqn, p, solver, tolerance, max_iter,
  reference_cell)`` (:155-184): pins the reference cell, then solves.
* ``correct_velocity(U, HbyA, p, A_p, mesh, alpha=1.0)`` (:187-267):
  ``pfm_correct_velocity`` (Gauss gradient, correction, |U| <= 1e4 limiter).
* ``correct_face_flux(phi, p, A_p, mesh, face_weights=None)`` (:270-332):
  ``pfm_correct_face_flux``.
* ``adjust_phi(phi, mesh, closed=True)`` (:335-377): ``pfm_adjust_phi``,
  in place, returns whether phi changed.

Every face sum runs in the reference's scatter order, so the results equal
the reference's bit for bit given the same inputs, except adjust_phi's net
boundary flux, whose summation order torch leaves open.  ``mesh`` is an
``FvMesh`` (face weights, areas, deltas and volumes on the device).
"""

from __future__ import annotations

from typing import Any

import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.fv_matrix import FvMatrix
from .linear_solver import LinearSolverBase

__all__ = ["assemble_pressure_equation", "solve_pressure_equation", "correct_velocity",
           "correct_face_flux", "adjust_phi"]


def _dev(t: torch.Tensor, mesh, shape=None) -> torch.Tensor:
    t = torch.as_tensor(t).to(device=mesh.device, dtype=torch.float64).contiguous()
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t


def _weights(mesh, face_weights) -> torch.Tensor:
    w = mesh.face_weights if face_weights is None else face_weights
    return _dev(w, mesh)[: mesh.n_internal_faces].contiguous()


def assemble_pressure_equation(phiHbyA: torch.Tensor, A_p: torch.Tensor, mesh: Any,
                               face_weights: torch.Tensor | None = None) -> FvMatrix:
    require_gpu(torch.empty(0, device=mesh.device))
    n, nfi, nfa = mesh.n_cells, mesh.n_internal_faces, mesh.n_faces
    phi = _dev(phiHbyA, mesh, (nfa,))
    ap = _dev(A_p, mesh, (n,))
    w = _weights(mesh, face_weights)
    mat = FvMatrix(n, mesh.owner[:nfi], mesh.neighbour, device=mesh.device)
    addr = mesh.ldu_addressing()
    mat._addr = addr
    f64 = dict(dtype=torch.float64, device=mesh.device)
    diag, lower, upper, source = (torch.empty(n, **f64), torch.empty(nfi, **f64),
                                  torch.empty(nfi, **f64), torch.empty(n, **f64))
    bptr, bface = mesh.boundary_csr
    p = _lib.ptr
    _lib.check(_lib.lib().pfm_pressure_assemble(
        addr.handle, nfa, p(bptr), p(bface), p(phi), p(ap), p(w),
        p(mesh.face_areas_magnitude[:nfi].contiguous()),
        p(mesh.delta_coefficients[:nfi].contiguous()), p(mesh.cell_volumes), p(diag), p(lower),
        p(upper), p(source), current_stream(mesh.device)), "pfm_pressure_assemble")
    mat._diag, mat._lower, mat._upper = diag, lower, upper
    mat._source = source
    return mat


def solve_pressure_equation(p_eqn: FvMatrix, p: torch.Tensor, solver: LinearSolverBase,
                            tolerance: float = 1e-6, max_iter: int = 1000,
                            reference_cell: int = 0) -> tuple[torch.Tensor, int, float]:
    p_eqn.set_reference(reference_cell, value=0.0)
    return p_eqn.solve(solver, p, tolerance=tolerance, max_iter=max_iter)


def correct_velocity(U: torch.Tensor, HbyA: torch.Tensor, p: torch.Tensor, A_p: torch.Tensor,
                     mesh: Any, alpha: float = 1.0) -> torch.Tensor:
    n, nfa = mesh.n_cells, mesh.n_faces
    hb = _dev(HbyA, mesh, (n, 3))
    pp = _dev(p, mesh, (n,))
    ap = _dev(A_p, mesh, (n,))
    w = _weights(mesh, None)
    out = torch.empty(n, 3, dtype=torch.float64, device=mesh.device)
    bptr, bface = mesh.boundary_csr
    q = _lib.ptr
    _lib.check(_lib.lib().pfm_correct_velocity(
        mesh.ldu_addressing().handle, nfa, q(bptr), q(bface), q(hb), q(pp), q(ap), q(w),
        q(mesh.face_areas.contiguous()), q(mesh.cell_volumes), float(alpha), q(out),
        current_stream(mesh.device)), "pfm_correct_velocity")
    return out


def correct_face_flux(phi: torch.Tensor, p: torch.Tensor, A_p: torch.Tensor, mesh: Any,
                      face_weights: torch.Tensor | None = None) -> torch.Tensor:
    n, nfi, nfa = mesh.n_cells, mesh.n_internal_faces, mesh.n_faces
    ph = _dev(phi, mesh, (nfa,))
    pp = _dev(p, mesh, (n,))
    ap = _dev(A_p, mesh, (n,))
    w = _weights(mesh, face_weights)
    out = torch.empty_like(ph)
    q = _lib.ptr
    _lib.check(_lib.lib().pfm_correct_face_flux(
        mesh.ldu_addressing().handle, nfa, q(ph), q(pp), q(ap), q(w),
        q(mesh.face_areas_magnitude[:nfi].contiguous()),
        q(mesh.delta_coefficients[:nfi].contiguous()), q(out), current_stream(mesh.device)),
        "pfm_correct_face_flux")
    return out


def adjust_phi(phi: torch.Tensor, mesh: Any, closed: bool = True) -> bool:
    """In place on ``phi`` (a contiguous float64 device tensor of n_faces)."""
    import ctypes as C
    nfi, nfa = mesh.n_internal_faces, mesh.n_faces
    if not (isinstance(phi, torch.Tensor) and phi.dtype == torch.float64 and phi.is_contiguous()
            and phi.device.type == "cuda" and phi.shape == (nfa,)):
        raise ValueError("adjust_phi works in place on a contiguous float64 device tensor "
                         f"of {nfa} face fluxes")
    work = torch.empty(1, dtype=torch.float64, device=phi.device)
    adjusted = C.c_int32(0)
    net = C.c_double(0.0)
    _lib.check(_lib.lib().pfm_adjust_phi(nfa, nfi, _lib.ptr(phi), int(bool(closed)),
                                         _lib.ptr(work), C.byref(net), C.byref(adjusted),
                                         current_stream(phi.device)), "pfm_adjust_phi")
    return bool(adjusted.value)
