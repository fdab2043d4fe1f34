"""Implicit Laplacian assembly on the GPU (reference ``discretisation/operators.py``).

``laplacian(D, phi, scheme, *, mesh, dirichlet=None)`` follows ``fvm.laplacian``
(:432-518): D may be a scalar, an n_faces (or n_internal_faces) face tensor
used as is, or a cell field (interpolated linearly, ``interpolation.py:101-144``:
w = |d_N|/(|d_P|+|d_N|), boundary faces take the owner value).  The scheme
string is accepted and ignored, as in the reference (no non-orthogonal
correction).  Coefficients:
  c_f = D_f |S_f| delta_f,  lower_f = -c_f/V_P,  upper_f = -c_f/V_N,
  diag_i = sum_owner-faces c/V_P + sum_neighbour-faces c/V_N,
computed by the ``pfm_laplacian_assemble`` kernel, bit-identical to the
reference's torch ops.  The reference's generic boundary term is zero
(boundary delta = 0).  ``dirichlet`` (additive keyword) names patches held at
a fixed value, contributing laplacianFoam's term
(``applications/laplacian_foam.py:342-366``): diag_P += D_b |S_b| delta_b / V_P
and source_P += the same times the value.

When D is a tensor that requires grad, the assembly is an autograd node whose
backward (``pfm_laplacian_backward``) returns dJ/dD_f from the gradients of
(diag, lower, upper).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from ..core.device import current_stream
from ..core.fv_matrix import FvMatrix

__all__ = ["laplacian", "assemble", "LaplacianAssemble"]


def _ptr(t):
    return _lib.ptr(t)


class LaplacianAssemble(torch.autograd.Function):
    """(gamma_internal, gamma_dirichlet) -> (diag, lower, upper).

    The Dirichlet arrays (``dir_cells`` int32 on the device, ``dir_gamma``,
    ``dir_mag``, ``dir_delta``) hold one entry per face in the reference's
    order: patch after patch (``patch_sizes`` faces each; default one patch),
    faces in face order.
    """

    @staticmethod
    def forward(ctx, gamma, dir_gamma, addr, mag, delta, vol, dir_cells, dir_mag, dir_delta,
                patch_sizes=None):
        dev = gamma.device
        n, nf = addr.n_cells, addr.n_faces
        diag = torch.empty(n, device=dev, dtype=torch.float64)
        lower = torch.empty(nf, device=dev, dtype=torch.float64)
        upper = torch.empty(nf, device=dev, dtype=torch.float64)
        nd = 0 if dir_cells is None else int(dir_cells.numel())
        g = gamma.detach().contiguous()
        dg = dir_gamma.detach().contiguous() if nd else None
        plan = addr.dirichlet_plan(dir_cells, patch_sizes) if nd else None
        _lib.check(_lib.lib().pfm_laplacian_assemble(
            addr.handle, _ptr(g), _ptr(mag), _ptr(delta), _ptr(vol),
            plan.handle if nd else None, _ptr(dg) if nd else None,
            _ptr(dir_mag) if nd else None, _ptr(dir_delta) if nd else None,
            _ptr(diag), _ptr(lower), _ptr(upper), current_stream(dev)), "pfm_laplacian_assemble")
        ctx.addr = addr
        ctx.nd = nd
        ctx.save_for_backward(mag, delta, vol, *( (dir_cells, dir_mag, dir_delta) if nd else ()))
        return diag, lower, upper

    @staticmethod
    def backward(ctx, g_diag, g_lower, g_upper):
        addr = ctx.addr
        saved = ctx.saved_tensors
        mag, delta, vol = saved[:3]
        dev = mag.device
        n, nf = addr.n_cells, addr.n_faces
        z = lambda k: torch.zeros(k, device=dev, dtype=torch.float64)
        gd = (g_diag if g_diag is not None else z(n)).contiguous()
        gl = (g_lower if g_lower is not None else z(nf)).contiguous()
        gu = (g_upper if g_upper is not None else z(nf)).contiguous()
        gg = torch.empty(nf, device=dev, dtype=torch.float64)
        gdir = None
        nd = ctx.nd
        if nd:
            dir_cells, dir_mag, dir_delta = saved[3:]
            gdir = torch.empty(nd, device=dev, dtype=torch.float64)
        _lib.check(_lib.lib().pfm_laplacian_backward(
            addr.handle, _ptr(gd), _ptr(gl), _ptr(gu), _ptr(mag), _ptr(delta), _ptr(vol), nd,
            _ptr(saved[3]) if nd else None, _ptr(saved[4]) if nd else None,
            _ptr(saved[5]) if nd else None, _ptr(gg), _ptr(gdir), current_stream(dev)),
            "pfm_laplacian_backward")
        return gg, gdir, None, None, None, None, None, None, None, None


def _face_D(D, mesh) -> torch.Tensor:
    dev, nfa, n = mesh.device, mesh.n_faces, mesh.n_cells
    if isinstance(D, (int, float)):
        return torch.full((nfa,), float(D), dtype=torch.float64, device=dev)
    if hasattr(D, "internal_field"):
        D = D.internal_field
    D = D if isinstance(D, torch.Tensor) else torch.tensor(D, dtype=torch.float64)
    D = D.to(device=dev, dtype=torch.float64)
    if D.dim() == 0:
        return D.expand(nfa)
    if D.shape[0] == n and n != nfa:
        # linear interpolation of a cell field, boundary faces take the owner value
        nfi = mesh.n_internal_faces
        own = mesh.owner
        nei = mesh.neighbour
        cc, fc = mesh.cell_centres, mesh.face_centres
        dP = (fc[:nfi] - cc[own[:nfi]]).norm(dim=1)
        dN = (fc[:nfi] - cc[nei]).norm(dim=1)
        den = dP + dN
        w = dN / torch.where(den > 1e-30, den, torch.ones_like(den))
        Df = torch.cat([w * D[own[:nfi]] + (1 - w) * D[nei], D[own[nfi:]]])
        return Df
    return D


def laplacian(D, phi=None, scheme: str = "Gauss linear corrected", *, mesh=None,
              dirichlet: dict | None = None) -> FvMatrix:
    if mesh is None:
        if hasattr(phi, "mesh"):
            mesh = phi.mesh
        else:
            raise ValueError("mesh is required when phi is not a GeometricField")
    dev = mesh.device
    nfi = mesh.n_internal_faces
    Df = _face_D(D, mesh)
    mat = FvMatrix(mesh.n_cells, mesh.owner[:nfi], mesh.neighbour, device=dev)
    # share the cached native topology across assemblies of one mesh
    if not hasattr(mesh, "_ldu_addr"):
        mesh._ldu_addr = mat.addressing
    mat._addr = mesh._ldu_addr
    addr = mat._addr
    mag = mesh.face_areas_magnitude
    delta = mesh.delta_coefficients
    vol = mesh.cell_volumes
    g_int = Df[:nfi] if Df.shape[0] != nfi else Df
    dir_cells = dir_mag = dir_delta = sizes = None
    g_dir = torch.zeros(0, dtype=torch.float64, device=dev)
    src_terms = []
    if dirichlet:
        cells_l, mag_l, dlt_l, g_l = [], [], [], []
        for patch, value in dirichlet.items():
            faces, cells, m_b, d_b = mesh.dirichlet_data(patch)
            cells_l.append(cells)
            mag_l.append(m_b)
            dlt_l.append(d_b)
            gf = Df[torch.as_tensor(faces, device=dev)] if Df.shape[0] == mesh.n_faces else \
                torch.ones(len(faces), dtype=torch.float64, device=dev) * (Df[0] if Df.numel() else 1.0)
            g_l.append(gf)
            src_terms.append((cells, m_b, d_b, gf, float(value)))
        # the reference's order: patch after patch (laplacian_foam.py:318-366)
        sizes = tuple(int(c.numel()) for c in cells_l)
        dir_cells = torch.cat(cells_l).to(torch.int32).contiguous()
        dir_mag, dir_delta = torch.cat(mag_l).contiguous(), torch.cat(dlt_l).contiguous()
        g_dir = torch.cat(g_l)
    diag, lower, upper = LaplacianAssemble.apply(g_int.contiguous(), g_dir.contiguous(), addr,
                                                 mag[:nfi].contiguous(), delta[:nfi].contiguous(),
                                                 vol, dir_cells, dir_mag, dir_delta, sizes)
    mat._diag, mat._lower, mat._upper = diag, lower, upper
    for cells, m_b, d_b, gf, value in src_terms:
        if value != 0.0:
            coeff = gf * m_b * d_b * value / vol[cells.long()]
            mat._source = mat._source.index_add(0, cells.long(), coeff)
    return mat


def assemble(case, mesh, *, gamma=None, dir_gamma=None):
    """Assemble a PoissonCase on the GPU: (FvMatrix, gamma_internal, gamma_dirichlet).

    gamma / dir_gamma default to the case's diffusivities (as device tensors);
    pass leaf tensors with requires_grad for the differentiable path.
    """
    dev = mesh.device
    if gamma is None:
        gamma = torch.from_numpy(case.gamma).to(dev)
    if dir_gamma is None:
        dir_gamma = torch.from_numpy(case.dir_gamma).to(dev)
    nfi = mesh.n_internal_faces
    mat = FvMatrix(mesh.n_cells, mesh.owner[:nfi], mesh.neighbour, device=dev)
    if not hasattr(mesh, "_ldu_addr"):
        mesh._ldu_addr = mat.addressing
    mat._addr = mesh._ldu_addr
    cells = torch.from_numpy(case.dir_cells.astype(np.int32)).to(dev)
    dmag = torch.from_numpy(case.dir_mag).to(dev)
    ddel = torch.from_numpy(case.dir_delta).to(dev)
    diag, lower, upper = LaplacianAssemble.apply(
        gamma, dir_gamma, mat._addr, mesh.face_areas_magnitude[:nfi].contiguous(),
        mesh.delta_coefficients[:nfi].contiguous(), mesh.cell_volumes, cells, dmag, ddel)
    mat._diag, mat._lower, mat._upper = diag, lower, upper
    mat.source = torch.from_numpy(case.b).to(dev)
    return mat
