"""Seeded Poisson test systems on unit-cube hex meshes.

``poisson_case(n, ...)`` builds the benchmark systems: an n^3 hex block,
optionally jittered (every vertex moved by U[-j*h, j*h] per coordinate,
boundary vertices only tangentially), cells randomly permuted (seeded), then
RCM-renumbered (``pfm_rcm``, the reference's tie-breaks), faces re-oriented
owner < neighbour and sorted (owner, neighbour).  Face diffusivity is 1 or
lognormal(0, sigma) on every face (internal and boundary, drawn in final face
order), p = 0 (Dirichlet) on the x = 0 patch, zeroGradient elsewhere, and the
right-hand side is U[-1, 1] per cell made zero-mean.

Random draws come from one ``np.random.default_rng(seed)`` in a fixed order:
jitter (n_points x 3), permutation (n_cells), diffusivity (n_faces),
right-hand side (n_cells).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .mesh.geometry import MeshGeometry, compute_geometry, dirichlet_delta
from .mesh.hexmesh import HexMesh, jitter_points, relabel_cells, structured_hex

__all__ = ["PoissonCase", "poisson_case", "rcm_permutation"]


@dataclass
class PoissonCase:
    n: int
    mesh: HexMesh
    geo: MeshGeometry
    gamma: np.ndarray          # (n_internal_faces,)
    dir_faces: np.ndarray      # boundary face ids of the Dirichlet patch, in face order
    dir_cells: np.ndarray      # owner of each Dirichlet face
    dir_gamma: np.ndarray
    dir_mag: np.ndarray
    dir_delta: np.ndarray
    b: np.ndarray
    seed: int

    @property
    def n_cells(self) -> int:
        return self.mesh.n_cells

    @property
    def owner(self) -> np.ndarray:
        return self.mesh.owner[: self.mesh.n_internal_faces]

    @property
    def neighbour(self) -> np.ndarray:
        return self.mesh.neighbour


def rcm_permutation(n_cells: int, owner, neighbour) -> np.ndarray:
    """RCM order perm[new] = old via the native pfm_rcm (host, no GPU)."""
    o = _lib.host_i32(owner)
    q = _lib.host_i32(neighbour)
    perm = np.empty(n_cells, np.int32)
    _lib.check(_lib.lib().pfm_rcm(n_cells, o.shape[0], _lib.ptr(o), _lib.ptr(q), _lib.ptr(perm)),
               "pfm_rcm")
    return perm.astype(np.int64)


def _case_mesh(n: int, jitter: float, seed: int, permute: bool, renumber: bool, note):
    """Mesh stages: structured block, jitter, permutation, RCM.  Returns the
    mesh and the generator after its mesh draws."""
    rng = np.random.default_rng(seed)
    mesh = structured_hex(n)
    note(f"case {n}^3: structured mesh")
    if jitter > 0.0:
        mesh = jitter_points(mesh, n, jitter, rng)
    if permute:
        mesh = relabel_cells(mesh, rng.permutation(mesh.n_cells))
        note("case: cells permuted")
    if renumber:
        nfi = mesh.n_internal_faces
        perm = rcm_permutation(mesh.n_cells, mesh.owner[:nfi], mesh.neighbour)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(mesh.n_cells)
        mesh = relabel_cells(mesh, inv)
        note("case: RCM renumbering")
    return mesh, rng


def _finish_case(n, mesh, rng, sigma, seed, dirichlet_patch, geo) -> PoissonCase:
    if sigma > 0.0:
        gamma_all = rng.lognormal(0.0, sigma, size=mesh.n_faces)
    else:
        gamma_all = np.ones(mesh.n_faces)
    b = rng.uniform(-1.0, 1.0, size=mesh.n_cells)
    b = b - b.mean()
    nfi = mesh.n_internal_faces
    patch = mesh.patch(dirichlet_patch)
    faces = np.arange(patch["start"], patch["start"] + patch["size"])
    cells = mesh.owner[faces]
    return PoissonCase(
        n=n, mesh=mesh, geo=geo, gamma=gamma_all[:nfi].copy(), dir_faces=faces, dir_cells=cells,
        dir_gamma=gamma_all[faces].copy(), dir_mag=geo.mag_sf[faces].copy(),
        dir_delta=dirichlet_delta(mesh, geo, faces), b=b, seed=seed)


def poisson_case(n: int, *, jitter: float = 0.0, sigma: float = 0.0, seed: int = 0,
                 permute: bool = True, renumber: bool = True,
                 dirichlet_patch: str = "x0", progress=None, geometry=None) -> PoissonCase:
    """``progress(msg)``, if given, is called after each construction stage.
    ``geometry(mesh, progress)`` computes the mesh geometry (default: the
    device kernels, ``compute_geometry``)."""
    note = progress or (lambda msg: None)
    mesh, rng = _case_mesh(n, jitter, seed, permute, renumber, note)
    geo = (geometry or compute_geometry)(mesh, progress)
    return _finish_case(n, mesh, rng, sigma, seed, dirichlet_patch, geo)


@dataclass
class BatchCase:
    """Many Poisson systems concatenated block-diagonally (global cell indices)."""
    sizes: np.ndarray          # n per system (cube side)
    cell_offsets: np.ndarray   # (ns+1,)
    face_offsets: np.ndarray   # (ns+1,)
    owner: np.ndarray          # global, (F,)
    neighbour: np.ndarray
    owner_local: np.ndarray    # per-system local indices
    neighbour_local: np.ndarray
    vol: np.ndarray
    mag: np.ndarray            # (F,)
    delta: np.ndarray
    gamma: np.ndarray
    dir_cells: np.ndarray      # global; each system's patch in face order
    dir_gamma: np.ndarray
    dir_mag: np.ndarray
    dir_delta: np.ndarray
    b: np.ndarray
    seed: int
    systems: list | None = None  # the per-system PoissonCases (keep_systems=True)

    @property
    def n_systems(self) -> int:
        return int(self.sizes.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cell_offsets[-1])


def _batch_mesh(args):
    """Mesh stages of one batch member (host; runs in worker processes)."""
    n, seed, jitter, sigma = args
    return _case_mesh(n, jitter, seed, True, True, lambda msg: None)


def poisson_batch(n_systems: int = 512, n_min: int = 8, n_max: int = 48, *, seed: int = 0,
                  jitter: float = 0.2, sigma: float = 1.0, workers: int = 8,
                  geometry=None, keep_systems: bool = False) -> BatchCase:
    """Batch of independent cube systems: side n_k ~ U{n_min..n_max} from
    ``default_rng(seed)``; system k is ``poisson_case(n_k, seed=seed*100003 + k)``
    (jittered, permuted, RCM-renumbered, lognormal diffusivity, zero-mean RHS).
    The meshes are built in ``workers`` host processes; the geometry
    (``geometry``, default the device kernels) runs in this process.
    ``keep_systems`` keeps the per-system cases in ``BatchCase.systems``."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=n_systems)
    jobs = [(int(n), seed * 100003 + k, jitter, sigma) for k, n in enumerate(sizes)]
    if workers > 1 and n_systems > 1:
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=workers) as ex:
            meshes = list(ex.map(_batch_mesh, jobs, chunksize=4))
    else:
        meshes = [_batch_mesh(j) for j in jobs]
    geometry = geometry or compute_geometry
    cases = [_finish_case(n, mesh, rng, sigma, sd, "x0", geometry(mesh, None))
             for (n, sd, _, _), (mesh, rng) in zip(jobs, meshes)]
    cell_off = np.zeros(n_systems + 1, np.int64)
    face_off = np.zeros(n_systems + 1, np.int64)
    for k, c in enumerate(cases):
        cell_off[k + 1] = cell_off[k] + c.n_cells
        face_off[k + 1] = face_off[k] + c.mesh.n_internal_faces
    cat = np.concatenate
    nfi = [c.mesh.n_internal_faces for c in cases]
    return BatchCase(
        sizes=sizes, cell_offsets=cell_off, face_offsets=face_off,
        owner=cat([c.owner + cell_off[k] for k, c in enumerate(cases)]),
        neighbour=cat([c.neighbour + cell_off[k] for k, c in enumerate(cases)]),
        owner_local=cat([c.owner for c in cases]), neighbour_local=cat([c.neighbour for c in cases]),
        vol=cat([c.geo.cell_volumes for c in cases]),
        mag=cat([c.geo.mag_sf[:f] for c, f in zip(cases, nfi)]),
        delta=cat([c.geo.delta[:f] for c, f in zip(cases, nfi)]),
        gamma=cat([c.gamma for c in cases]),
        dir_cells=cat([c.dir_cells + cell_off[k] for k, c in enumerate(cases)]),
        dir_gamma=cat([c.dir_gamma for c in cases]), dir_mag=cat([c.dir_mag for c in cases]),
        dir_delta=cat([c.dir_delta for c in cases]), b=cat([c.b for c in cases]), seed=seed,
        systems=cases if keep_systems else None)
