"""Shared builders for the tests: seeded Poisson systems assembled by the oracle."""

from __future__ import annotations

import functools

import numpy as np

from oracle import ldu as oldu
from oracle import mesh as omesh
from pyopenfoam_amd.cases import poisson_case
from pyopenfoam_amd.mesh.geometry import MeshGeometry


def oracle_geometry(mesh, progress=None) -> MeshGeometry:
    """Mesh geometry from the oracle's vectorised restatement (host)."""
    g = omesh.geometry_vec(mesh.points, mesh.faces, mesh.owner, mesh.neighbour, mesh.n_cells)
    return MeshGeometry(face_centres=g["face_centres"], face_areas=g["face_areas"],
                        mag_sf=g["mag_sf"], cell_volumes=g["cell_volumes"],
                        cell_centres=g["cell_centres"], delta=g["delta"])


def make_case(n: int, **kw):
    """``poisson_case`` with the oracle's geometry (no device needed)."""
    return poisson_case(n, geometry=oracle_geometry, **kw)


@functools.lru_cache(maxsize=16)
def case(n: int, jitter: float = 0.0, sigma: float = 0.0, seed: int = 0):
    return make_case(n, jitter=jitter, sigma=sigma, seed=seed)


@functools.lru_cache(maxsize=16)
def system(n: int, jitter: float = 0.0, sigma: float = 0.0, seed: int = 0):
    """(case, diag, lower, upper) assembled by the oracle restatement."""
    c = case(n, jitter, sigma, seed)
    nfi = c.mesh.n_internal_faces
    d, lo, up = oldu.laplacian(c.gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                               c.owner, c.neighbour)
    d = oldu.dirichlet_diag(d, c.dir_gamma, c.dir_mag, c.dir_delta, c.geo.cell_volumes,
                            c.dir_cells)
    return c, d, lo, up


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))
