"""Hex-mesh generation and finite-volume geometry (inputs of the assembly)."""

from .geometry import MeshGeometry, compute_geometry, dirichlet_delta
from .hexmesh import HexMesh, jitter_points, relabel_cells, structured_hex

__all__ = ["MeshGeometry", "compute_geometry", "dirichlet_delta", "HexMesh", "jitter_points",
           "relabel_cells", "structured_hex"]
