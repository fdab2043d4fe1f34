"""General polyhedral meshes read from OpenFOAM polyMesh files.

``PolyMesh`` carries the arrays ``FvMesh`` and the device geometry take --
points, faces as CSR (face_ptr / face_verts), owner, neighbour, patches --
with the attribute names of ``HexMesh``, so a case read from disk goes
straight into the geometry kernels and the int32 LDU handle.
``read_polymesh(dir)`` = ``io.read_mesh`` + ``PolyMesh.from_mesh_data``
(the reference's ``MeshData`` -> ``PolyMesh`` of ``mesh/poly_mesh.py:62-101``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Union

import numpy as np

__all__ = ["PolyMesh", "read_polymesh"]


@dataclass
class PolyMesh:
    points: np.ndarray                 # (n_points, 3) float64
    faces: tuple                       # (face_ptr int32 (n_faces + 1), face_verts int32)
    owner: np.ndarray                  # (n_faces,) int64
    neighbour: np.ndarray              # (n_internal_faces,) int64
    n_cells: int
    patches: list = field(default_factory=list)

    @property
    def n_faces(self) -> int:
        return int(self.faces[0].shape[0]) - 1

    @property
    def n_internal_faces(self) -> int:
        return int(self.neighbour.shape[0])

    def patch(self, name: str) -> dict:
        for p in self.patches:
            if p["name"] == name:
                return p
        raise KeyError(name)

    @classmethod
    def from_mesh_data(cls, md) -> "PolyMesh":
        owner = md.owner.cpu().numpy().astype(np.int64)
        neighbour = md.neighbour.cpu().numpy().astype(np.int64)
        nc = md.n_cells
        if neighbour.size:
            nc = max(nc, int(neighbour.max()) + 1)
        patches = [dict(name=p.name, type=p.patch_type, start=p.start_face, size=p.n_faces)
                   for p in md.boundary]
        return cls(points=md.points.cpu().numpy().astype(np.float64),
                   faces=(np.asarray(md.face_ptr, np.int32), np.asarray(md.face_verts, np.int32)),
                   owner=owner, neighbour=neighbour, n_cells=nc, patches=patches)


def read_polymesh(mesh_dir: Union[str, Path]) -> PolyMesh:
    from ..io.mesh_io import read_mesh
    return PolyMesh.from_mesh_data(read_mesh(mesh_dir))
