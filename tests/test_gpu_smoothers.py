"""GPU smoothers bit-exact against the oracle and the reference's own output.

GaussSeidelSmoother: the level-scheduled sweeps must reproduce the sequential
sweeps bit for bit, checked against the reference's output on the golden 4^3
system, and against the oracle on jittered meshes (RCM-ordered: one
workgroup walks the levels; randomly ordered 48^3: one launch per level), on
a topology with faces of both orientations (as agglomerated levels have),
for several omega and sweep counts.  JacobiSmoother: bit-exact against the
oracle; the reference's known answers through the device API."""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import gamg as ogamg
from oracle import ldu as oldu
from pyopenfoam_amd import LduMatrix
from pyopenfoam_amd.solvers import GaussSeidelSmoother, JacobiSmoother
from tests.helpers import bits_equal, make_case, system

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def _matrix(n, own, nei, d, lo, up, dev):
    m = LduMatrix(n, torch.from_numpy(np.asarray(own)), torch.from_numpy(np.asarray(nei)),
                  device=dev)
    m.diag = torch.from_numpy(np.asarray(d, np.float64))
    m.lower = torch.from_numpy(np.asarray(lo, np.float64))
    m.upper = torch.from_numpy(np.asarray(up, np.float64))
    return m


def test_gauss_seidel_reference_golden(gpu_device):
    """The reference GaussSeidelSmoother's own output (2 sweeps, golden 4^3)."""
    g = dict(np.load(G / "sys_4.npz"))
    nfi = g["neighbour"].shape[0]
    own, nei = g["owner"][:nfi], g["neighbour"]
    m = _matrix(len(g["diag"]), own, nei, g["diag"], g["lap_lower"], g["lap_upper"], gpu_device)
    x0 = torch.from_numpy(g["gs_x0"]).to(gpu_device)
    x = GaussSeidelSmoother(m).smooth(x0, torch.from_numpy(g["b"]), 2)
    assert bits_equal(x.cpu().numpy(), g["gs_x"])
    assert bits_equal(x0.cpu().numpy(), g["gs_x0"])  # input untouched


def _gs_case(n, renumber, swap, seed):
    c = make_case(n, jitter=0.2, sigma=1.0, seed=seed, renumber=renumber)
    nfi = c.mesh.n_internal_faces
    d, lo, up = oldu.laplacian(c.gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi],
                               c.geo.cell_volumes, c.owner, c.neighbour)
    own, nei = c.owner.copy(), c.neighbour.copy()
    if swap:
        # flip the orientation of half the faces: lower and upper swap with them
        msk = np.random.default_rng(seed).random(nfi) < 0.5
        own, nei = np.where(msk, nei, own), np.where(msk, own, nei)
        lo, up = np.where(msk, up, lo), np.where(msk, lo, up)
    return c, d, lo, up, own, nei


@pytest.mark.parametrize("n,renumber,swap,sweeps,omega", [
    (8, True, False, 1, 1.0), (16, True, False, 3, 0.8), (16, True, True, 2, 1.0),
    (48, False, False, 2, 1.0), (24, False, True, 1, 1.2)])
def test_gauss_seidel_matches_oracle(gpu_device, n, renumber, swap, sweeps, omega):
    c, d, lo, up, own, nei = _gs_case(n, renumber, swap, seed=n)
    print(f"gs n={n} renumber={renumber} swap={swap} sweeps={sweeps} omega={omega}", flush=True)
    rng = np.random.default_rng(1)
    x0 = rng.standard_normal(c.n_cells)
    m = _matrix(c.n_cells, own, nei, d, lo, up, gpu_device)
    x = GaussSeidelSmoother(m, omega=omega).smooth(torch.from_numpy(x0), torch.from_numpy(c.b),
                                                   sweeps)
    xo = ogamg.gauss_seidel(d, lo, up, own, nei, c.b, x0, sweeps, omega=omega)
    assert bits_equal(x.cpu().numpy(), xo)


@pytest.mark.parametrize("n,nf", [(300, 2000), (200000, 700000)])
def test_gauss_seidel_wide_rows(gpu_device, n, nf):
    """Random graphs whose rows often hold more than 8 couplings (merged
    through the slot map instead of the schedule entry's columns), faces of
    both orientations in random order: 300 cells (one workgroup walks the
    levels) and 200,000 cells (one launch per level)."""
    rng = np.random.default_rng(n)
    own = rng.integers(0, n, nf)
    nei = rng.integers(0, n, nf)
    keep = own != nei
    own, nei = own[keep], nei[keep]
    deg = np.bincount(own, minlength=n) + np.bincount(nei, minlength=n)
    assert deg.max() > 8
    lo = -rng.uniform(0.1, 1.0, own.shape[0])
    up = -rng.uniform(0.1, 1.0, own.shape[0])
    d = np.bincount(own, -lo, n) + np.bincount(nei, -up, n) + 0.5
    b = rng.uniform(-1.0, 1.0, n)
    x0 = rng.standard_normal(n)
    m = _matrix(n, own, nei, d, lo, up, gpu_device)
    x = GaussSeidelSmoother(m, omega=0.9).smooth(torch.from_numpy(x0), torch.from_numpy(b), 2)
    xo = ogamg.gauss_seidel(d, lo, up, own, nei, b, x0, 2, omega=0.9)
    assert bits_equal(x.cpu().numpy(), xo)


@pytest.mark.parametrize("n_iter,omega", [(1, 2.0 / 3.0), (2, 2.0 / 3.0), (5, 1.0)])
def test_jacobi_matches_oracle(gpu_device, n_iter, omega):
    c, d, lo, up = system(16, 0.2, 1.0)
    x0 = np.random.default_rng(2).standard_normal(c.n_cells)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    x = JacobiSmoother(m, omega=omega).smooth(torch.from_numpy(x0), torch.from_numpy(c.b), n_iter)
    xo = ogamg.jacobi(d, lo, up, c.owner, c.neighbour, c.b, x0, n_iter, omega=omega)
    assert bits_equal(x.cpu().numpy(), xo)


def test_smoother_known_answers(gpu_device):
    """Reference tests/unit/solvers/test_smoothers.py: Jacobi on identity and
    diag(2) (:182-206); Gauss-Seidel on a 3-cell SPD chain lowers the
    residual with more sweeps and converges to the solution."""
    e = torch.zeros(0, dtype=torch.int64)
    m = LduMatrix(3, e, e, device=gpu_device)
    m.diag = torch.tensor([1.0, 1.0, 1.0], dtype=torch.float64)
    x = JacobiSmoother(m, omega=1.0).smooth(torch.zeros(3, dtype=torch.float64),
                                            torch.tensor([2.0, 3.0, 4.0], dtype=torch.float64), 1)
    assert x.cpu().tolist() == [2.0, 3.0, 4.0]
    m2 = LduMatrix(2, e, e, device=gpu_device)
    m2.diag = torch.tensor([2.0, 2.0], dtype=torch.float64)
    x = JacobiSmoother(m2, omega=1.0).smooth(torch.zeros(2, dtype=torch.float64),
                                             torch.tensor([4.0, 6.0], dtype=torch.float64), 1)
    assert x.cpu().tolist() == [2.0, 3.0]
    own, nei = torch.tensor([0, 1]), torch.tensor([1, 2])
    A = LduMatrix(3, own, nei, device=gpu_device)
    A.diag = torch.tensor([2.1, 2.1, 1.1], dtype=torch.float64)
    A.lower = torch.tensor([-1.0, -1.0], dtype=torch.float64)
    A.upper = torch.tensor([-1.0, -1.0], dtype=torch.float64)
    b = torch.tensor([1.0, 2.0, 1.0], dtype=torch.float64)
    gs = GaussSeidelSmoother(A)
    res = [(b.to(gpu_device) - A.Ax(gs.smooth(torch.zeros(3, dtype=torch.float64), b, k))).norm()
           .item() for k in (1, 2, 5, 80)]
    assert res[0] > res[1] > res[2] > res[3] and res[3] < 1e-12
