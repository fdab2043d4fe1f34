"""Finite-volume geometry of polyMeshes, computed on the device.

Formulas and operation order are those of the reference's
``src/pyfoam/mesh/mesh_geometry.py`` (``pfm_mesh_geometry``):

* face centre = mean of the face vertices                         (:35-63)
* face area vector S_f = 1/2 sum_j v_j x v_{j+1}                    (:66-108)
* cell volume = sum over the cell's faces and face edges of
  |tet(ref, fc, v_j, v_{j+1})|, ref = mean of the cell's distinct
  vertices, faces seen from the neighbour side reversed; cell centre =
  volume-weighted tet centroids; degenerate cells get V = 0 and the
  mean of their face centres                                        (:139-254)
* delta_f = 1/|(C_N - C_P) . n_hat| on internal faces, 0 on
  boundary faces                                                    (:330-395)
* boundary delta for a Dirichlet patch (laplacianFoam convention)
  1/max(|(x_f - C_P) . n_hat|, 1e-30)   (``src/pyfoam/applications/laplacian_foam.py:342-358``)

The reference evaluates these with per-face and per-cell Python loops (hours
at 2M cells); here the cell -> face lists are one host counting sort
(``pfm_cell_faces``) and the geometry is three kernels, one thread per face
or cell.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.device import current_stream, get_device
from .hexmesh import HexMesh

__all__ = ["MeshGeometry", "compute_geometry", "dirichlet_delta", "face_csr", "cell_faces"]


@dataclass
class MeshGeometry:
    face_centres: np.ndarray       # (n_faces, 3)
    face_areas: np.ndarray         # (n_faces, 3) area vectors
    mag_sf: np.ndarray             # (n_faces,)
    cell_volumes: np.ndarray       # (n_cells,)
    cell_centres: np.ndarray       # (n_cells, 3)
    delta: np.ndarray              # (n_faces,) internal 1/|d.n|, boundary 0


def face_csr(faces) -> tuple[np.ndarray, np.ndarray]:
    """(face_ptr, face_verts) int32 from an (n_faces, n_v) array, a list of
    vertex loops or a (face_ptr, face_verts) pair."""
    if isinstance(faces, tuple) and len(faces) == 2:
        return np.asarray(faces[0], np.int32), np.asarray(faces[1], np.int32)
    if isinstance(faces, np.ndarray) and faces.ndim == 2:
        nf, nv = faces.shape
        return (np.arange(nf + 1, dtype=np.int64) * nv).astype(np.int32), \
            np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
    lens = np.array([len(f) for f in faces], dtype=np.int64)
    ptr = np.zeros(len(faces) + 1, dtype=np.int64)
    np.cumsum(lens, out=ptr[1:])
    verts = np.concatenate([np.asarray(f, dtype=np.int64) for f in faces]) if len(faces) else \
        np.zeros(0, np.int64)
    return ptr.astype(np.int32), verts.astype(np.int32)


def cell_faces(n_cells: int, owner, neighbour, face_ptr) -> tuple[np.ndarray, np.ndarray, int]:
    """(cell_ptr, cell_faces, max_cell_verts) from ``pfm_cell_faces`` (host):
    owned faces ascending, then neighboured faces ascending as -(f + 1)."""
    o = _lib.host_i32(owner)
    q = _lib.host_i32(neighbour)
    fp = _lib.host_i32(face_ptr)
    nf, nfi = o.shape[0], q.shape[0]
    cptr = np.empty(n_cells + 1, np.int32)
    cf = np.empty(nf + nfi, np.int32)
    mx = C.c_int32(0)
    _lib.check(_lib.lib().pfm_cell_faces(n_cells, nf, nfi, _lib.ptr(o), _lib.ptr(q), _lib.ptr(fp),
                                         _lib.ptr(cptr), _lib.ptr(cf), C.byref(mx)),
               "pfm_cell_faces")
    return cptr, cf, int(mx.value)


def compute_geometry(mesh: HexMesh, progress=None, device=None,
                     keep_device: bool = False) -> MeshGeometry:
    """Geometry of ``mesh`` on the HIP device ``device`` (default: current).

    ``progress(msg)``, if given, is called after the host and device stages.
    With ``keep_device`` the result also carries the device tensors
    (attribute ``device_tensors``: dict of the six outputs)."""
    note = progress or (lambda msg: None)
    dev = torch.device(device) if device is not None else get_device()
    if dev.type != "cuda":
        raise RuntimeError("compute_geometry runs on a HIP device (pfm_mesh_geometry); "
                           f"got {dev}")
    fptr, fverts = face_csr(mesh.faces)
    n_points = int(mesh.points.shape[0])
    if fverts.size and (fverts.min() < 0 or fverts.max() >= n_points):
        raise ValueError(f"face vertex ids must lie in [0, {n_points})")
    nf, nfi, nc = fptr.shape[0] - 1, mesh.n_internal_faces, mesh.n_cells
    cptr, cf, max_cv = cell_faces(nc, mesh.owner, mesh.neighbour, fptr)
    if max_cv > 64:
        raise ValueError(f"a cell references {max_cv} vertices; at most 64 are supported")
    note("geometry: cell -> face lists (host)")

    def up(a, dt):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dt)

    pts = up(mesh.points, torch.float64)
    d_fptr, d_fv = up(fptr, torch.int32), up(fverts, torch.int32)
    d_own = up(mesh.owner, torch.int32)
    d_nei = up(mesh.neighbour, torch.int32)
    d_cptr, d_cf = up(cptr, torch.int32), up(cf, torch.int32)
    f64 = dict(dtype=torch.float64, device=dev)
    fc = torch.empty(nf, 3, **f64)
    sf = torch.empty(nf, 3, **f64)
    mag = torch.empty(nf, **f64)
    vol = torch.empty(nc, **f64)
    cc = torch.empty(nc, 3, **f64)
    delta = torch.empty(nf, **f64)
    p = _lib.ptr
    _lib.check(_lib.lib().pfm_mesh_geometry(
        n_points, p(pts), nf, p(d_fptr), p(d_fv), nfi, p(d_own), p(d_nei), nc, p(d_cptr),
        p(d_cf), max_cv, p(fc), p(sf), p(mag), p(vol), p(cc), p(delta), current_stream(dev)),
        "pfm_mesh_geometry")
    geo = MeshGeometry(face_centres=fc.cpu().numpy(), face_areas=sf.cpu().numpy(),
                       mag_sf=mag.cpu().numpy(), cell_volumes=vol.cpu().numpy(),
                       cell_centres=cc.cpu().numpy(), delta=delta.cpu().numpy())
    if keep_device:
        geo.device_tensors = dict(face_centres=fc, face_areas=sf, mag_sf=mag, cell_volumes=vol,
                                  cell_centres=cc, delta=delta)
    note("geometry: device kernels")
    return geo


def dirichlet_delta(mesh: HexMesh, geo: MeshGeometry, face_ids: np.ndarray) -> np.ndarray:
    """Boundary delta 1/max(|(x_f - C_P).n_hat|, 1e-30) on the given boundary
    faces (a few per cent of the faces; host), as
    src/applications/laplacian_foam.py:341-353 forms it: n_hat = S_f/|S_f|, and
    (1, 1, 1)/sqrt(3) for a face with |S_f| <= 1e-30."""
    face_ids = np.asarray(face_ids, dtype=np.int64)
    d = geo.face_centres[face_ids] - geo.cell_centres[mesh.owner[face_ids]]
    s = geo.face_areas[face_ids]
    m = geo.mag_sf[face_ids]
    ok = (m > 1e-30)[:, None]
    nh = np.where(ok, s / np.where(ok, m[:, None], 1.0), 1.0 / np.sqrt(3.0))
    dn = np.abs((d[:, 0] * nh[:, 0] + d[:, 1] * nh[:, 1]) + d[:, 2] * nh[:, 2])
    return 1.0 / np.maximum(dn, 1e-30)
