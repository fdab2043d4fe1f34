"""The rest of the reference's linear-solver family on the device, bit-exact
against the oracle: DIC/DILU ``apply_full`` (also against the reference's own
output on the golden 4^3 system), ``SmoothSolver`` (Gauss-Seidel, DIC, DILU),
``DiagonalSolver`` and ``DICGSmoother``."""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import ldu as oldu
from oracle import precond as opre
from oracle import smooth as osm
from pyopenfoam_amd import LduMatrix, create_solver
from pyopenfoam_amd.solvers import (DICGSmoother, DICPreconditioner, DILUPreconditioner,
                                    DiagonalSolver, SmoothSolver, solver_from_dict)
from tests.helpers import bits_equal, system

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def _matrix(n, own, nei, d, lo, up, dev):
    m = LduMatrix(n, torch.from_numpy(np.asarray(own)), torch.from_numpy(np.asarray(nei)),
                  device=dev)
    m.diag = torch.from_numpy(np.asarray(d, np.float64))
    m.lower = torch.from_numpy(np.asarray(lo, np.float64))
    m.upper = torch.from_numpy(np.asarray(up, np.float64))
    return m


def _swapped(c, d, lo, up, seed):
    """Half the faces flipped (owner > neighbour): any face order."""
    msk = np.random.default_rng(seed).random(len(lo)) < 0.5
    own = np.where(msk, c.neighbour, c.owner)
    nei = np.where(msk, c.owner, c.neighbour)
    return own, nei, np.where(msk, up, lo), np.where(msk, lo, up)


def test_apply_full_reference_golden(gpu_device):
    g = dict(np.load(G / "sys_4.npz"))
    nfi = g["neighbour"].shape[0]
    m = _matrix(len(g["diag"]), g["owner"][:nfi], g["neighbour"], g["diag"], g["lap_lower"],
                g["lap_upper"], gpu_device)
    r = torch.from_numpy(g["r"])
    assert bits_equal(DICPreconditioner(m).apply_full(r).cpu().numpy(), g["dic_full"])
    assert bits_equal(DILUPreconditioner(m).apply_full(r).cpu().numpy(), g["dilu_full"])


@pytest.mark.parametrize("n,swap", [(16, False), (24, False), (10, True)])
def test_apply_full_matches_oracle(gpu_device, n, swap):
    """Oriented, owner-sorted faces: cell-level schedules; flipped faces: the
    one-workgroup face-level sweep."""
    c, d, lo, up = system(n, 0.2, 1.0)
    own, nei = c.owner, c.neighbour
    if swap:
        own, nei, lo, up = _swapped(c, d, lo, up, n)
    print(f"apply_full n={n} swap={swap}", flush=True)
    m = _matrix(c.n_cells, own, nei, d, lo, up, gpu_device)
    r = np.random.default_rng(4).standard_normal(c.n_cells)
    rDc = opre.dic_build(d, up, own, nei)
    rDl = opre.dilu_build(d, lo, up, own, nei)
    zc = DICPreconditioner(m).apply_full(torch.from_numpy(r)).cpu().numpy()
    zl = DILUPreconditioner(m).apply_full(torch.from_numpy(r)).cpu().numpy()
    assert bits_equal(zc, opre.dic_apply_full(rDc, up, own, nei, r))
    assert bits_equal(zl, opre.dilu_apply_full(rDl, lo, up, own, nei, r))


@pytest.mark.parametrize("smoother,n_sweeps,omega", [("GaussSeidel", 2, 1.0), ("DIC", 2, 1.0),
                                                     ("DILU", 1, 0.7), ("GaussSeidel", 3, 1.0)])
@pytest.mark.parametrize("n", [8, 20])
def test_smooth_solver_matches_oracle(gpu_device, smoother, n_sweeps, omega, n):
    """x after a fixed number of iterations bit-exact (the monitor cannot stop
    it: tolerance 1e-300, rel_tol 1e-300); the residual history to 1e-9 (its
    norms are the only reductions); 20^3 sweeps by per-level launches."""
    c, d, lo, up = system(n, 0.2, 1.0)
    print(f"smooth {smoother} n={n}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    k = 6
    s = SmoothSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=k, smoother=smoother,
                     n_sweeps=n_sweeps, omega=omega)
    x, it, res = s(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    xo, io, ro, info = osm.smooth_solver(d, lo, up, c.owner, c.neighbour, c.b, np.zeros(c.n_cells),
                                         smoother=smoother, n_sweeps=n_sweeps, omega=omega,
                                         tolerance=1e-300, rel_tol=1e-300, max_iter=k)
    assert it == io == k + 1
    assert bits_equal(x.cpu().numpy(), xo)
    np.testing.assert_allclose(s.last_info.residual_history, info.residual_history, rtol=1e-9)


def test_smooth_solver_stops_like_the_reference(gpu_device):
    """Default rel_tol 0.01: the same iteration at which the monitor stops."""
    c, d, lo, up = system(12, 0.0, 1.0)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    for sm in ("GaussSeidel", "DIC"):
        s = create_solver("smoothSolver", tolerance=1e-12, max_iter=500, smoother=sm)
        x, it, res = s(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
        xo, io, ro, info = osm.smooth_solver(d, lo, up, c.owner, c.neighbour, c.b,
                                             np.zeros(c.n_cells), smoother=sm, tolerance=1e-12,
                                             max_iter=500)
        assert s.last_info.converged == info.converged and it == io, (sm, it, io)
        assert bits_equal(x.cpu().numpy(), xo)
    s = solver_from_dict("smoothSolver", {"smoother": "DILU", "nSweeps": 3, "relTol": 0.05})
    assert isinstance(s, SmoothSolver) and s._n_sweeps == 3 and s._smoother == "DILU"


def test_diagonal_solver(gpu_device):
    c, d, lo, up = system(10, 0.2, 1.0)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    x, it, res = DiagonalSolver()(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    xo, ro = osm.diagonal_solver(d, c.b)
    assert it == 1 and bits_equal(x.cpu().numpy(), xo)
    np.testing.assert_allclose(res, ro, rtol=1e-12, atol=1e-300)
    e = torch.zeros(0, dtype=torch.int64)
    m2 = LduMatrix(2, e, e, device=gpu_device)
    m2.diag = torch.tensor([1e-35, 2.0], dtype=torch.float64)
    x, it, res = create_solver("diagonal")(m2, torch.tensor([1.0, 2.0], dtype=torch.float64),
                                           torch.zeros(2, dtype=torch.float64))
    assert torch.isfinite(x).all() and x[1].item() == 1.0


def test_dicg_smoother(gpu_device):
    c, d, lo, up = system(12, 0.2, 1.0)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    x0 = np.random.default_rng(2).standard_normal(c.n_cells)
    x = DICGSmoother(m, omega=0.9).smooth(torch.from_numpy(x0), torch.from_numpy(c.b), 3)
    rD = opre.dic_build(d, up, c.owner, c.neighbour)
    xo = x0.copy()
    for _ in range(3):
        xo = xo + 0.9 * opre.dic_apply_full(rD, up, c.owner, c.neighbour,
                                            c.b - oldu.ax(d, lo, up, c.owner, c.neighbour, xo))
    assert bits_equal(x.cpu().numpy(), xo)
