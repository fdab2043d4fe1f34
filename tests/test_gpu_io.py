"""A case read from polyMesh files goes straight into the device path.

Bar: a jittered 10^3 case written as a binary polyMesh and read back
(``read_polymesh``) gives the same device geometry, the same assembled
Laplacian and the same PCG solution, bit for bit, as the in-memory mesh."""

import numpy as np
import pytest
import torch

from pyopenfoam_amd import PCGSolver
from pyopenfoam_amd.discretisation.fvm import laplacian
from pyopenfoam_amd.io import FileFormat, FoamFileHeader, write_boundary, write_faces
from pyopenfoam_amd.io import write_neighbour, write_owner, write_points
from pyopenfoam_amd.io.mesh_io import BoundaryPatch
from pyopenfoam_amd.mesh.fv_mesh import FvMesh
from pyopenfoam_amd.mesh.geometry import compute_geometry
from pyopenfoam_amd.mesh.polymesh import read_polymesh
from tests.helpers import bits_equal, make_case

pytestmark = pytest.mark.gpu


def test_polymesh_case_on_device(gpu_device, tmp_path):
    c = make_case(10, jitter=0.2, sigma=1.0, seed=5)
    m = c.mesh
    h = lambda cls, obj: FoamFileHeader(format=FileFormat.BINARY, class_name=cls, object=obj)
    write_points(tmp_path / "points", h("vectorField", "points"), torch.from_numpy(m.points))
    write_faces(tmp_path / "faces", h("faceList", "faces"), list(m.faces))
    write_owner(tmp_path / "owner", h("labelList", "owner"), torch.from_numpy(m.owner))
    write_neighbour(tmp_path / "neighbour", h("labelList", "neighbour"),
                    torch.from_numpy(m.neighbour))
    write_boundary(tmp_path / "boundary", h("polyBoundaryMesh", "boundary"),
                   [BoundaryPatch(p["name"], p["type"], p["size"], p["start"]) for p in m.patches])
    pm = read_polymesh(tmp_path)
    assert pm.n_cells == m.n_cells and np.array_equal(pm.owner, m.owner)

    g1 = compute_geometry(pm, device=gpu_device)
    g0 = compute_geometry(m, device=gpu_device)
    for k in ("face_centres", "face_areas", "mag_sf", "cell_volumes", "cell_centres", "delta"):
        assert bits_equal(getattr(g1, k), getattr(g0, k)), k

    mats = []
    for mesh in (FvMesh(pm, device=gpu_device, geometry=g1), FvMesh(m, device=gpu_device, geometry=g0)):
        gam = torch.from_numpy(np.concatenate([c.gamma, np.ones(m.n_faces - m.n_internal_faces)]))
        mat = laplacian(gam.to(gpu_device), mesh=mesh)
        mat.source = torch.from_numpy(c.b).to(gpu_device)
        mats.append(mat)
    for k in ("diag", "lower", "upper"):
        assert bits_equal(getattr(mats[0], k).cpu().numpy(), getattr(mats[1], k).cpu().numpy()), k
    xs = []
    for mat in mats:
        mat.diag = mat.diag + 1e-3  # a non-singular system (no Dirichlet patch here)
        s = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=3000)
        x, it, _ = s(mat, mat.source, torch.zeros(m.n_cells, dtype=torch.float64, device=gpu_device))
        xs.append((x.cpu().numpy(), it, s.last_info.converged))
    assert xs[0][1] == xs[1][1] and xs[0][2] == xs[1][2] and bits_equal(xs[0][0], xs[1][0])
