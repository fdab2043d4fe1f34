"""Structured hexahedral meshes in OpenFOAM polyMesh form, plus cell relabelling.

The solver path consumes an unstructured LDU system, so a structured block is
only a convenient way to make one: every array produced here (points, quad
faces, owner, neighbour, patches) has exactly the layout of an OpenFOAM
``polyMesh`` as read by the reference's ``PolyMesh`` (``src/pyfoam/mesh/poly_mesh.py:62-101``):
internal faces first, then boundary faces grouped by patch, face vertices
ordered so that the right-hand-rule normal points from owner to neighbour
(outwards on boundary faces).

Relabelling follows the convention the hot path relies on: after a cell
permutation every internal face is re-oriented so that owner < neighbour (its
vertex loop reversed, flipping the normal) and internal faces are sorted by
(owner, neighbour), the upper-triangular order used by OpenFOAM's
``renumberMesh``.  The reference only validates owner < neighbour
(``src/pyfoam/mesh/topology.py:81-87``); the sort is fixed here so that every
face-ordered algorithm (DILU, GAMG, pair agglomeration, Gauss-Seidel) sees one
well-defined order.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["HexMesh", "structured_hex", "relabel_cells", "jitter_points"]


@dataclass
class HexMesh:
    """All-quad polyMesh arrays.

    points     (n_points, 3) float64
    faces      (n_faces, 4) int64 vertex loops
    owner      (n_faces,) int64
    neighbour  (n_internal_faces,) int64
    patches    list of dicts {name, type, start, size}
    """

    points: np.ndarray
    faces: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray
    n_cells: int
    patches: list = field(default_factory=list)

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])

    @property
    def n_internal_faces(self) -> int:
        return int(self.neighbour.shape[0])

    def patch(self, name: str) -> dict:
        for p in self.patches:
            if p["name"] == name:
                return p
        raise KeyError(name)


def _pid(n: int, i, j, k):
    return i + (n + 1) * (j + (n + 1) * k)


def _cid(n: int, i, j, k):
    return i + n * (j + n * k)


def structured_hex(n: int, *, length: float = 1.0) -> HexMesh:
    """Unit-cube block of n^3 hex cells in natural (i fastest) order.

    Patches: ``x0`` (x = 0), ``x1``, ``y0``, ``y1``, ``z0``, ``z1``; all are
    created with type ``zeroGradient``; callers set the Dirichlet one.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    h = length / n
    g = np.arange(n + 1, dtype=np.float64) * h
    kk, jj, ii = np.meshgrid(np.arange(n + 1), np.arange(n + 1), np.arange(n + 1), indexing="ij")
    points = np.stack([g[ii.ravel()], g[jj.ravel()], g[kk.ravel()]], axis=1)

    faces, owner, neighbour = [], [], []
    c = np.arange(n)
    # x-normal internal faces between (i-1,j,k) and (i,j,k)
    if n > 1:
        k3, j3, i3 = np.meshgrid(c, c, np.arange(1, n), indexing="ij")
        i3, j3, k3 = i3.ravel(), j3.ravel(), k3.ravel()
        faces.append(np.stack([_pid(n, i3, j3, k3), _pid(n, i3, j3 + 1, k3),
                               _pid(n, i3, j3 + 1, k3 + 1), _pid(n, i3, j3, k3 + 1)], 1))
        owner.append(_cid(n, i3 - 1, j3, k3)); neighbour.append(_cid(n, i3, j3, k3))
        # y-normal
        k3, j3, i3 = np.meshgrid(c, np.arange(1, n), c, indexing="ij")
        i3, j3, k3 = i3.ravel(), j3.ravel(), k3.ravel()
        faces.append(np.stack([_pid(n, i3, j3, k3), _pid(n, i3, j3, k3 + 1),
                               _pid(n, i3 + 1, j3, k3 + 1), _pid(n, i3 + 1, j3, k3)], 1))
        owner.append(_cid(n, i3, j3 - 1, k3)); neighbour.append(_cid(n, i3, j3, k3))
        # z-normal
        k3, j3, i3 = np.meshgrid(np.arange(1, n), c, c, indexing="ij")
        i3, j3, k3 = i3.ravel(), j3.ravel(), k3.ravel()
        faces.append(np.stack([_pid(n, i3, j3, k3), _pid(n, i3 + 1, j3, k3),
                               _pid(n, i3 + 1, j3 + 1, k3), _pid(n, i3, j3 + 1, k3)], 1))
        owner.append(_cid(n, i3, j3, k3 - 1)); neighbour.append(_cid(n, i3, j3, k3))
    n_int = int(sum(len(o) for o in owner))

    patches = []
    start = n_int
    b, a = np.meshgrid(c, c, indexing="ij")  # a fastest
    a, b = a.ravel(), b.ravel()
    zero = np.zeros_like(a)
    last = np.full_like(a, n)
    # (name, vertex loop with outward normal, owner cell)
    spec = [
        ("x0", [_pid(n, zero, a, b), _pid(n, zero, a, b + 1), _pid(n, zero, a + 1, b + 1), _pid(n, zero, a + 1, b)],
         _cid(n, zero, a, b)),
        ("x1", [_pid(n, last, a, b), _pid(n, last, a + 1, b), _pid(n, last, a + 1, b + 1), _pid(n, last, a, b + 1)],
         _cid(n, last - 1, a, b)),
        ("y0", [_pid(n, a, zero, b), _pid(n, a + 1, zero, b), _pid(n, a + 1, zero, b + 1), _pid(n, a, zero, b + 1)],
         _cid(n, a, zero, b)),
        ("y1", [_pid(n, a, last, b), _pid(n, a, last, b + 1), _pid(n, a + 1, last, b + 1), _pid(n, a + 1, last, b)],
         _cid(n, a, last - 1, b)),
        ("z0", [_pid(n, a, b, zero), _pid(n, a, b + 1, zero), _pid(n, a + 1, b + 1, zero), _pid(n, a + 1, b, zero)],
         _cid(n, a, b, zero)),
        ("z1", [_pid(n, a, b, last), _pid(n, a + 1, b, last), _pid(n, a + 1, b + 1, last), _pid(n, a, b + 1, last)],
         _cid(n, a, b, last - 1)),
    ]
    for name, loop, own in spec:
        faces.append(np.stack(loop, 1))
        owner.append(own)
        patches.append({"name": name, "type": "zeroGradient", "start": start, "size": int(own.shape[0])})
        start += int(own.shape[0])

    return HexMesh(
        points=points,
        faces=np.concatenate(faces).astype(np.int64),
        owner=np.concatenate(owner).astype(np.int64),
        neighbour=(np.concatenate(neighbour) if neighbour else np.zeros(0)).astype(np.int64),
        n_cells=n ** 3,
        patches=patches,
    )


def jitter_points(mesh: HexMesh, n: int, fraction: float, rng: np.random.Generator,
                  *, length: float = 1.0) -> HexMesh:
    """Displace every vertex by U[-fraction*h, fraction*h] per coordinate.

    Vertices on the cube's boundary move only tangentially (the coordinate that
    lies on a boundary plane is kept), so the domain stays the unit cube.
    The random draw is one (n_points, 3) uniform block, consumed before any
    other use of ``rng``.
    """
    h = length / n
    d = rng.uniform(-fraction * h, fraction * h, size=mesh.points.shape)
    p = mesh.points
    tol = 1e-12 * length
    on_bnd = (np.abs(p) < tol) | (np.abs(p - length) < tol)
    d[on_bnd] = 0.0
    return HexMesh(points=p + d, faces=mesh.faces, owner=mesh.owner, neighbour=mesh.neighbour,
                   n_cells=mesh.n_cells, patches=[dict(q) for q in mesh.patches])


def relabel_cells(mesh: HexMesh, new_of_old: np.ndarray) -> HexMesh:
    """Apply a cell permutation (``new_of_old[old] = new``).

    Internal faces are re-oriented to owner < neighbour (vertex loop reversed
    when swapped) and sorted by (owner, neighbour) with ties in face order
    (``pfm_relabel_faces``, two counting sorts on the host); boundary faces
    keep their patch grouping and order and only get the new owner label.
    """
    from .. import _lib

    new_of_old = np.asarray(new_of_old, dtype=np.int64)
    if new_of_old.shape != (mesh.n_cells,):
        raise ValueError("permutation has the wrong length")
    nfi = mesh.n_internal_faces
    o = _lib.host_i32(mesh.owner[:nfi])
    q = _lib.host_i32(mesh.neighbour)
    perm = _lib.host_i32(new_of_old)
    order = np.empty(nfi, np.int32)
    flip = np.empty(nfi, np.uint8)
    own_i = np.empty(nfi, np.int32)
    nei = np.empty(nfi, np.int32)
    p = _lib.ptr
    _lib.check(_lib.lib().pfm_relabel_faces(mesh.n_cells, nfi, p(o), p(q), p(perm), p(order),
                                            p(flip), p(own_i), p(nei)), "pfm_relabel_faces")
    faces = mesh.faces.copy()
    fi = mesh.faces[:nfi][order]
    sw = flip.astype(bool)
    fi[sw] = fi[sw][:, ::-1]
    faces[:nfi] = fi
    own = new_of_old[mesh.owner]
    own[:nfi] = own_i
    return HexMesh(points=mesh.points, faces=faces, owner=own, neighbour=nei.astype(np.int64),
                   n_cells=mesh.n_cells, patches=[dict(q) for q in mesh.patches])
