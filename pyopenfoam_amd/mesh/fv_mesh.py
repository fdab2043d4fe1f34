"""Finite-volume mesh with device-resident geometry (reference ``mesh/fv_mesh.py``).

Holds the polyMesh arrays and the geometry the Laplacian assembly consumes
(cell volumes, face area vectors and magnitudes, delta coefficients), with
the reference's property names (``fv_mesh.py:127-190``).  Geometry is
computed once on the host (``geometry.compute_geometry``) and kept as float64
tensors on the mesh's device.
"""

from __future__ import annotations

import numpy as np
import torch

from ..core.device import CFD_DTYPE, INDEX_DTYPE, get_device
from .geometry import MeshGeometry, compute_geometry, dirichlet_delta
from .hexmesh import HexMesh

__all__ = ["FvMesh"]


class FvMesh:
    def __init__(self, mesh: HexMesh, *, device: torch.device | str | None = None,
                 geometry: MeshGeometry | None = None) -> None:
        self._mesh = mesh
        self._device = torch.device(device) if device is not None else get_device()
        self._geo = geometry
        self._t: dict = {}

    @classmethod
    def from_case(cls, case, device=None) -> "FvMesh":
        return cls(case.mesh, device=device, geometry=case.geo)

    # topology
    @property
    def device(self) -> torch.device:
        return self._device

    @property
    def dtype(self) -> torch.dtype:
        return CFD_DTYPE

    @property
    def n_cells(self) -> int:
        return self._mesh.n_cells

    @property
    def n_faces(self) -> int:
        return self._mesh.n_faces

    @property
    def n_internal_faces(self) -> int:
        return self._mesh.n_internal_faces

    @property
    def boundary(self) -> list:
        return self._mesh.patches

    @property
    def host(self) -> HexMesh:
        return self._mesh

    def _tensor(self, key, fn, dtype):
        if key not in self._t:
            self._t[key] = torch.as_tensor(np.ascontiguousarray(fn()), dtype=dtype).to(self._device)
        return self._t[key]

    @property
    def geometry(self) -> MeshGeometry:
        if self._geo is None:
            self._geo = compute_geometry(self._mesh)
        return self._geo

    @property
    def owner(self) -> torch.Tensor:
        return self._tensor("owner", lambda: self._mesh.owner, INDEX_DTYPE)

    @property
    def neighbour(self) -> torch.Tensor:
        return self._tensor("neighbour", lambda: self._mesh.neighbour, INDEX_DTYPE)

    @property
    def cell_volumes(self) -> torch.Tensor:
        return self._tensor("vol", lambda: self.geometry.cell_volumes, CFD_DTYPE)

    @property
    def cell_centres(self) -> torch.Tensor:
        return self._tensor("cc", lambda: self.geometry.cell_centres, CFD_DTYPE)

    @property
    def face_centres(self) -> torch.Tensor:
        return self._tensor("fc", lambda: self.geometry.face_centres, CFD_DTYPE)

    @property
    def face_areas(self) -> torch.Tensor:
        return self._tensor("sf", lambda: self.geometry.face_areas, CFD_DTYPE)

    @property
    def face_areas_magnitude(self) -> torch.Tensor:
        return self._tensor("magsf", lambda: self.geometry.mag_sf, CFD_DTYPE)

    @property
    def delta_coefficients(self) -> torch.Tensor:
        return self._tensor("delta", lambda: self.geometry.delta, CFD_DTYPE)

    @property
    def face_weights(self) -> torch.Tensor:
        """Linear-interpolation weights (n_faces,), ``pfm_face_weights``
        (mesh_geometry.py:262-327): |x_f - C_N| / (|x_f - C_P| + |x_f - C_N|),
        1 on boundary faces."""
        if "w" not in self._t:
            from .. import _lib
            from ..core.device import current_stream
            nf, nfi = self.n_faces, self.n_internal_faces
            w = torch.empty(nf, dtype=CFD_DTYPE, device=self._device)
            own, nei = self.owner_i32, self.neighbour_i32
            fc, cc = self.face_centres.contiguous(), self.cell_centres.contiguous()
            _lib.check(_lib.lib().pfm_face_weights(
                nf, nfi, _lib.ptr(own), _lib.ptr(nei), _lib.ptr(fc), _lib.ptr(cc), _lib.ptr(w),
                current_stream(self._device)), "pfm_face_weights")
            self._t["w"] = w
        return self._t["w"]

    @property
    def owner_i32(self) -> torch.Tensor:
        return self._tensor("own32", lambda: self._mesh.owner, torch.int32)

    @property
    def neighbour_i32(self) -> torch.Tensor:
        return self._tensor("nei32", lambda: self._mesh.neighbour, torch.int32)

    @property
    def boundary_csr(self) -> tuple[torch.Tensor, torch.Tensor]:
        """(bnd_ptr (n_cells + 1), bnd_face): per cell, in face order, the
        boundary faces it owns as offsets past the internal faces (int32)."""
        if "bcsr" not in self._t:
            nfi = self.n_internal_faces
            bo = np.asarray(self._mesh.owner[nfi:], dtype=np.int64)
            order = np.argsort(bo, kind="stable").astype(np.int32)
            ptr = np.zeros(self.n_cells + 1, np.int64)
            np.cumsum(np.bincount(bo, minlength=self.n_cells), out=ptr[1:])
            self._t["bcsr"] = (torch.from_numpy(ptr.astype(np.int32)).to(self._device),
                               torch.from_numpy(order).to(self._device))
        return self._t["bcsr"]

    def ldu_addressing(self):
        """The native LDU topology of the internal faces, shared by every
        matrix assembled on this mesh."""
        if not hasattr(self, "_ldu_addr"):
            from ..core.addressing import addressing_for
            nfi = self.n_internal_faces
            self._ldu_addr = addressing_for(self.n_cells, self.owner[:nfi], self.neighbour,
                                            self._device)
        return self._ldu_addr

    def patch_faces(self, name: str) -> np.ndarray:
        p = self._mesh.patch(name)
        return np.arange(p["start"], p["start"] + p["size"])

    def dirichlet_data(self, patch: str):
        """(faces, cells, |S_b|, delta_b) of a patch in face order; cells, |S_b|
        and delta_b on the device."""
        key = ("dir", patch)
        if key not in self._t:
            faces = self.patch_faces(patch)
            cells = self._mesh.owner[faces]
            geo = self.geometry
            self._t[key] = (
                faces,
                torch.from_numpy(cells.astype(np.int32)).to(self._device),
                torch.from_numpy(geo.mag_sf[faces].copy()).to(self._device),
                torch.from_numpy(dirichlet_delta(self._mesh, geo, faces)).to(self._device),
            )
        return self._t[key]
