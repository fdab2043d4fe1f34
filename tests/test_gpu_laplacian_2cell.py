"""fvm.laplacian on the reference tests' two-cell hex mesh (two unit cubes
stacked in z, tests/unit/discretisation/test_laplacian.py:18-65), through the
device path: the polyMesh arrays -> device geometry -> the assembly kernels.

The reference's assertions (test_laplacian.py:82-170: an FvMatrix of the
right size, |lower| = |upper|, negative off-diagonals, positive diagonal,
A 1 = 0, coefficients linear in D, D given per cell), with the known values of
this mesh (|S| = delta = V = 1: lower = upper = -1, diag = [1, 1] for D = 1,
to the rounding of the geometry) and, exactly, what the arithmetic fixes:
A 1 = 0 (each diagonal entry and its coupling are the same quotient c/V with
opposite signs) and coefficients for 2D = 2 x those for D bit for bit (scaling
by 2 commutes with rounding)."""

import numpy as np
import pytest
import torch

from pyopenfoam_amd.core.fv_matrix import FvMatrix
from pyopenfoam_amd.discretisation import fvm
from pyopenfoam_amd.mesh.fv_mesh import FvMesh
from pyopenfoam_amd.mesh.geometry import compute_geometry
from pyopenfoam_amd.mesh.polymesh import PolyMesh
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

_POINTS = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
           [0, 1, 1], [0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2]]
_FACES = [[4, 5, 6, 7], [0, 3, 2, 1], [0, 1, 5, 4], [3, 7, 6, 2], [0, 4, 7, 3], [1, 2, 6, 5],
          [8, 9, 10, 11], [4, 5, 9, 8], [7, 11, 10, 6], [4, 8, 11, 7], [5, 6, 10, 9]]


def _mesh(dev):
    ptr = np.cumsum([0] + [len(f) for f in _FACES]).astype(np.int32)
    pm = PolyMesh(points=np.array(_POINTS, dtype=np.float64),
                  faces=(ptr, np.concatenate(_FACES).astype(np.int32)),
                  owner=np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1], np.int64),
                  neighbour=np.array([1], np.int64), n_cells=2,
                  patches=[dict(name="bottom", type="wall", start=1, size=5),
                           dict(name="top", type="wall", start=6, size=5)])
    return FvMesh(pm, device=dev, geometry=compute_geometry(pm, device=dev))


def _np(t):
    return t.detach().cpu().numpy()


def test_two_cell_laplacian_matrix(gpu_device):
    mesh = _mesh(gpu_device)
    phi = torch.tensor([1.0, 2.0], dtype=torch.float64)
    mat = fvm.laplacian(1.0, phi, mesh=mesh)
    assert isinstance(mat, FvMatrix)
    assert mat.n_cells == 2 and mat.n_internal_faces == 1
    lo, up, d = _np(mat.lower), _np(mat.upper), _np(mat.diag)
    np.testing.assert_allclose(lo, [-1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(up, [-1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(d, [1.0, 1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.abs(lo), np.abs(up), rtol=0, atol=1e-12)
    assert (lo < 0).all() and (up < 0).all() and (d > 0).all()
    ones = torch.ones(2, dtype=torch.float64, device=gpu_device)
    assert _np(mat.Ax(ones)).tolist() == [0.0, 0.0]
    y = _np(mat.Ax(torch.tensor([1.0, 2.0], dtype=torch.float64, device=gpu_device)))
    np.testing.assert_allclose(y, [-1.0, 1.0], rtol=0, atol=1e-12)


def test_two_cell_laplacian_scales_with_D(gpu_device):
    mesh = _mesh(gpu_device)
    phi = torch.tensor([1.0, 2.0], dtype=torch.float64)
    m1 = fvm.laplacian(1.0, phi, mesh=mesh)
    m2 = fvm.laplacian(2.0, phi, mesh=mesh)
    for k in ("lower", "upper", "diag"):
        assert bits_equal(_np(getattr(m2, k)), 2.0 * _np(getattr(m1, k))), k
    # D per cell: linearly interpolated to the face (weights 1/2 here)
    mc = fvm.laplacian(torch.tensor([1.0, 3.0], dtype=torch.float64), phi, mesh=mesh)
    np.testing.assert_allclose(_np(mc.lower), [-2.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(_np(mc.upper), [-2.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(_np(mc.diag), [2.0, 2.0], rtol=0, atol=1e-12)
    same = fvm.laplacian(torch.tensor([1.0, 1.0], dtype=torch.float64), phi, mesh=mesh)
    for k in ("lower", "upper", "diag"):
        np.testing.assert_allclose(_np(getattr(same, k)), _np(getattr(m1, k)), rtol=0, atol=1e-12)
