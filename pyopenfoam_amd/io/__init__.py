"""OpenFOAM files: FoamFile headers, polyMesh lists (ASCII and binary) and
dictionaries (reference ``io/``)."""

from .dictionary import FoamDict, FoamList, expand_macros, parse_dict, parse_dict_file, solver_dicts
from .foam_file import FileFormat, FoamFileHeader, parse_header, read_foam_file, write_foam_file
from .mesh_io import (BoundaryPatch, MeshData, read_boundary, read_faces, read_faces_csr, read_mesh,
                      read_neighbour, read_owner, read_points, write_boundary, write_faces,
                      write_neighbour, write_owner, write_points)

__all__ = ["FoamDict", "FoamList", "expand_macros", "parse_dict", "parse_dict_file",
           "solver_dicts", "FileFormat", "FoamFileHeader", "parse_header", "read_foam_file",
           "write_foam_file", "BoundaryPatch", "MeshData", "read_boundary", "read_faces",
           "read_faces_csr", "read_mesh", "read_neighbour", "read_owner", "read_points",
           "write_boundary", "write_faces", "write_neighbour", "write_owner", "write_points"]
