"""GPU GAMG against the oracle restatement of the reference GAMG and the
reference's golden maps/histories.  The V-cycle is fixed-order element-wise
arithmetic, so the solution after a given number of cycles is bit-exact; only
the residual norm's reduction tree differs (history to 1e-12)."""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import gamg as ogamg
from pyopenfoam_amd import GAMGSolver, LduMatrix
from tests.helpers import bits_equal, system

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def _matrix(n, own, nei, d, lo, up, dev):
    m = LduMatrix(n, torch.from_numpy(own), torch.from_numpy(nei), device=dev)
    m.diag, m.lower, m.upper = (torch.from_numpy(np.ascontiguousarray(v)) for v in (d, lo, up))
    return m


@pytest.mark.parametrize("n,jit,sig", [(8, 0.0, 0.0), (16, 0.2, 1.0), (24, 0.2, 1.0)])
def test_gamg_hierarchy_and_values(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"gamg hierarchy n={n}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=10)
    hier = solver.hierarchy(m)
    levels, aggs = ogamg.build_hierarchy(ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour))
    assert hier.n_levels == len(levels)
    for k in range(len(levels) - 1):
        f2c, fmap, co, cn = hier.level_maps(k)
        assert np.array_equal(f2c, aggs[k][1])
        assert np.array_equal(co, levels[k + 1].owner) and np.array_equal(cn, levels[k + 1].neighbour)
        dd, ll, uu = solver.coarse_values(m, k + 1)
        assert bits_equal(dd.cpu().numpy(), levels[k + 1].diag)
        assert bits_equal(ll.cpu().numpy(), levels[k + 1].lower)
        assert bits_equal(uu.cpu().numpy(), levels[k + 1].upper)


@pytest.mark.parametrize("n,jit,sig", [(8, 0.0, 0.0), (16, 0.2, 1.0), (32, 0.2, 1.0)])
def test_gamg_solve_bit_exact(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"gamg solve n={n}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    ncyc = 12
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=ncyc)
    x, it, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    xo, io, ro, info = ogamg.gamg(ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour), c.b,
                                  np.zeros(c.n_cells), tolerance=1e-300, rel_tol=1e-300,
                                  max_iter=ncyc)
    assert it == io == ncyc
    assert bits_equal(x.cpu().numpy(), xo)
    np.testing.assert_allclose(solver.last_info.residual_history, info.residual_history, rtol=1e-12)


@pytest.mark.parametrize("name", ["sys_4", "sys_6"])
def test_gamg_golden(gpu_device, name):
    g = dict(np.load(G / f"{name}.npz"))
    nfi = g["neighbour"].shape[0]
    m = _matrix(int(g["n_cells"]), g["owner"][:nfi], g["neighbour"], g["diag"], g["lap_lower"],
                g["lap_upper"], gpu_device)
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=25, min_cells_coarse=4)
    x, it, res = solver(m, torch.from_numpy(g["b"]), torch.zeros(m.n_cells, dtype=torch.float64))
    assert bits_equal(x.cpu().numpy(), g["gamg_x"])
    np.testing.assert_allclose(solver.last_info.residual_history, g["gamg_hist"], rtol=1e-12)
    f2c, *_ = solver.hierarchy(m).level_maps(0)
    assert np.array_equal(f2c, g["gamg_f2c"])


def test_gamg_pair_galerkin_converges(gpu_device):
    """Extension: pair agglomeration + Galerkin coarse diagonal reaches 1e-8
    (the reference-default hierarchy contracts by ~0.99 per cycle)."""
    c, d, lo, up = system(16, 0.2, 1.0)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=400, agglomeration="pair",
                        coarse="galerkin", max_levels=12)
    x, it, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    assert solver.last_info.converged, it
    from oracle import ldu as oldu
    xr = x.cpu().numpy()
    assert np.linalg.norm(c.b - oldu.ax(d, lo, up, c.owner, c.neighbour, xr)) / np.linalg.norm(c.b) < 2e-8


@pytest.mark.parametrize("wg_cells", ["1024", "0"])
@pytest.mark.parametrize("n,jit,sig", [(8, 0.2, 1.0), (24, 0.2, 1.0)])
def test_gamg_gauss_seidel_bit_exact(gpu_device, n, jit, sig, wg_cells, monkeypatch):
    """Extension level_smoother="GaussSeidel": the reference V-cycle with every
    smoothing step a symmetric Gauss-Seidel sweep; bit-exact with the oracle's
    V-cycle over GaussSeidelSmoother after 6 cycles.  The launched levels are
    swept by one workgroup walking their dependency levels (few cells per
    dependency level, the default) or by one launch per dependency level
    (PFM_GS_WG_CELLS=0); the small levels run inside the one-workgroup tail;
    coarse levels carry faces of both orientations."""
    monkeypatch.setenv("PFM_GS_WG_CELLS", wg_cells)
    c, d, lo, up = system(n, jit, sig)
    print(f"gamg gs n={n} wg_cells={wg_cells}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    ncyc = 6
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=ncyc,
                        level_smoother="GaussSeidel")
    x, it, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    xo, io, ro, info = ogamg.gamg(ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour), c.b,
                                  np.zeros(c.n_cells), tolerance=1e-300, rel_tol=1e-300,
                                  max_iter=ncyc, level_smoother="gauss_seidel")
    assert it == io == ncyc
    assert bits_equal(x.cpu().numpy(), xo)
    np.testing.assert_allclose(solver.last_info.residual_history, info.residual_history, rtol=1e-12)


@pytest.mark.parametrize("coarse", ["reference", "galerkin"])
@pytest.mark.parametrize("n,jit,sig", [(16, 0.2, 1.0), (6, 0.2, 1.0)])
def test_gamg_pair_hierarchy_bit_exact(gpu_device, n, jit, sig, coarse):
    """Jacobi V-cycles over a deep pair hierarchy (down to a few cells) against
    the oracle's V-cycle over the oracle's pair maps and coarse matrices (the
    reference's, or with the Galerkin coarse diagonal), bit-exact after 6
    cycles.  At 16^3 one cycle crosses every Jacobi
    execution mode: grid launches (4096, 2048 cells), the one-workgroup tail's
    one-row-per-thread levels (1024 .. 128: registers, x exchanged through
    LDS) and its one-wave levels (64 cells and fewer: registers, cross-lane
    reads); at 6^3 the whole hierarchy is in the tail."""
    c, d, lo, up = system(n, jit, sig)
    print(f"gamg pair n={n} coarse={coarse}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    ncyc = 6
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=ncyc, agglomeration="pair",
                        coarse=coarse, max_levels=20)
    sizes = solver.hierarchy(m).level_sizes()
    x, it, _ = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    lv = ogamg.pair_levels(c.n_cells, c.owner, c.neighbour, n_levels=len(sizes) - 1)
    levels = [ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour)]
    aggs, prev = [], np.arange(c.n_cells)
    for L in lv:
        f2c = np.empty(levels[-1].n, np.int64)
        f2c[prev] = L["cell_map"]
        aggs.append((L["n_cells_coarse"], f2c))
        levels.append(ogamg.coarse_matrix(levels[-1], L["n_cells_coarse"], f2c,
                                          galerkin=coarse == "galerkin")[0])
        prev = L["cell_map"]
    assert [(v.n, len(v.owner)) for v in levels] == [tuple(s) for s in sizes]
    assert min(sizes)[0] <= 8 and len(sizes) >= 4
    xo = np.zeros(c.n_cells)
    for _ in range(ncyc):
        xo = ogamg.v_cycle(levels, aggs, 0, xo, c.b, 2, 2)
    assert it == ncyc
    assert bits_equal(x.cpu().numpy(), xo)


def test_gamg_tail_wide_row_bit_exact(gpu_device):
    """A level too wide for the registers (a hub cell coupled to 21 others:
    22 owner-side slots) takes the one-workgroup block path while the coarser
    levels, where the hub's neighbours have been aggregated away, run one row
    per thread (140, 69 cells) and in one wave (34, 17, 9 cells); bit-exact
    against the oracle's V-cycle after 6 cycles."""
    n = 300
    faces = sorted(set((i, i + 1) for i in range(n - 1)) | set((0, j) for j in range(10, 31)))
    own = np.array([f[0] for f in faces], np.int64)
    nei = np.array([f[1] for f in faces], np.int64)
    rng = np.random.default_rng(3)
    lo = -(0.5 + rng.random(len(faces)))
    up = -(0.5 + rng.random(len(faces)))
    d = np.full(n, 0.1)
    np.add.at(d, own, np.abs(lo))
    np.add.at(d, nei, np.abs(up))
    b = rng.uniform(-1.0, 1.0, n)
    m = _matrix(n, own, nei, d, lo, up, gpu_device)
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=6)
    assert [s[0] for s in solver.hierarchy(m).level_sizes()] == [300, 140, 69, 34, 17, 9]
    x, it, _ = solver(m, torch.from_numpy(b), torch.zeros(n, dtype=torch.float64))
    xo, io, _, info = ogamg.gamg(ogamg.Level(n, d, lo, up, own, nei), b, np.zeros(n),
                                 tolerance=1e-300, rel_tol=1e-300, max_iter=6)
    assert it == io == 6
    assert bits_equal(x.cpu().numpy(), xo)


def test_gamg_non_finite_coefficient_like_the_oracle(gpu_device):
    """A coarse level's first sweep from x = 0 is element-wise only while its
    coefficients are finite; with an infinite coefficient (which the coarse
    levels inherit) the products with zero are formed and NaN/inf propagate
    as in the oracle's V-cycle: same NaN and inf positions, finite entries
    bit-identical."""
    c, d, lo, up = system(16, 0.2, 1.0)
    lo = lo.copy()
    lo[5] = np.inf
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    solver = GAMGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=2)
    assert solver.hierarchy(m).level_sizes()[1][0] > 1024  # a launched coarse level
    x, it, _ = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    with np.errstate(all="ignore"):  # inf * 0 in the oracle, as intended
        xo, io, _, _ = ogamg.gamg(ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour), c.b,
                                  np.zeros(c.n_cells), tolerance=1e-300, rel_tol=1e-300,
                                  max_iter=2)
    xg = x.cpu().numpy()
    assert it == io == 2
    assert np.isnan(xo).any()
    assert np.array_equal(np.isnan(xg), np.isnan(xo)) and np.array_equal(np.isinf(xg), np.isinf(xo))
    fin = np.isfinite(xo)
    assert bits_equal(xg[fin], xo[fin])


def _pair(c, d, lo, up, dev, **kw):
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, dev)
    base = dict(tolerance=1e-300, agglomeration="pair", coarse="galerkin", max_levels=12)
    base.update(kw)
    return m, base


@pytest.mark.parametrize("n", [8, 24])
def test_gamg_mixed_one_cycle_matches_fp64_cycle(gpu_device, n):
    """One refinement step with one inner V-cycle from x0 = 0: x = e, the
    V-cycle applied to r = b.  8^3 (every level FP64: at most 4096 cells):
    bit-exact with one FP64 cycle.  24^3 (level 0 in FP32, an odd number of Jacobi
    sweeps per level so the buffer parity flips): the FP64 cycle to FP32
    rounding (rel 1e-5)."""
    c, d, lo, up = system(n, 0.2, 1.0)
    m, kw = _pair(c, d, lo, up, gpu_device, n_pre_smooth=1)
    b = torch.from_numpy(c.b)
    z = torch.zeros(c.n_cells, dtype=torch.float64)
    x64, *_ = GAMGSolver(rel_tol=1e-300, max_iter=1, **kw)(m, b, z)
    sm = GAMGSolver(rel_tol=1e-300, max_iter=1, precision="mixed", inner_rel_tol=0.999,
                    max_outer=1, **kw)
    xm, *_ = sm(m, b, z)
    assert sm.last_info.inner_iterations == 1
    a, e = x64.cpu().numpy(), xm.cpu().numpy()
    if n == 8:
        assert bits_equal(e, a)
    else:
        assert np.linalg.norm(e - a) <= 1e-5 * np.linalg.norm(a)


def test_gamg_mixed_converges(gpu_device):
    """FP64 refinement around FP32 V-cycles reaches the FP64 tolerance: the
    true residual of the result is below rel_tol times the initial one, and
    the solution agrees with the FP64 GAMG solution to the tolerance's
    order."""
    c, d, lo, up = system(24, 0.2, 1.0)
    m, kw = _pair(c, d, lo, up, gpu_device)
    b = torch.from_numpy(c.b)
    z = torch.zeros(c.n_cells, dtype=torch.float64)
    sm = GAMGSolver(rel_tol=1e-8, max_iter=5000, precision="mixed", inner_rel_tol=1e-3, **kw)
    xm, outer, _ = sm(m, b, z)
    assert sm.last_info.converged
    assert outer >= 3 and sm.last_info.inner_iterations > 0
    from oracle import ldu as oldu
    xr = xm.cpu().numpy()
    true_rel = np.linalg.norm(c.b - oldu.ax(d, lo, up, c.owner, c.neighbour, xr)) / np.linalg.norm(c.b)
    assert true_rel < 1e-8, true_rel
    x64, *_ = GAMGSolver(rel_tol=1e-10, max_iter=5000, **kw)(m, b, z)
    xa = x64.cpu().numpy()
    assert np.linalg.norm(xr - xa) <= 1e-6 * np.linalg.norm(xa)


@pytest.mark.parametrize("n,jit,sig", [(12, 0.0, 1.0), (24, 0.0, 1.0), (16, 0.2, 1.0)])
def test_pcg_gamg_preconditioner_matches_oracle(gpu_device, n, jit, sig):
    """PCG with one reference-hierarchy V-cycle as the preconditioner against
    the oracle PCG with the oracle's V-cycle (bit-exact on its own, see the
    GAMG tests): bit-exact when the oracle sums in the device's order; in the
    reference's orders iterations within one of their spread and the solution
    within ten times it (floor 1e-9)."""
    from oracle import krylov as okry
    from pyopenfoam_amd import PCGSolver
    c, d, lo, up = system(n, jit, sig)
    print(f"pcg-gamg n={n} jitter={jit}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=500, preconditioner="GAMG",
                       gamg_options=dict(agglomeration="greedy", coarse="reference"))
    x, it, _ = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    levels, aggs = ogamg.build_hierarchy(ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour))
    pre = lambda r: ogamg.v_cycle(levels, aggs, 0, np.zeros_like(r), r, 2, 2)
    from oracle import ldu as oldu
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    solve = lambda dot: okry.pcg(mv, pre, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                                 rel_tol=1e-8, max_iter=500, dot=dot)
    # bit-exact against the oracle summing in the device's order
    xd, idev, _, infod = solve(okry.dot_device(c.n_cells, layout="contiguous"))
    xg = x.cpu().numpy()
    assert it == idev and solver.last_info.converged == infod.converged, (it, idev)
    assert bits_equal(np.array(solver.last_info.residual_history), infod.residual_history)
    assert bits_equal(xg, xd)
    # against the reference's own summation orders
    runs = {k: solve(dot) for k, dot in okry.DOTS.items()}
    xo, io, ro, info = runs["blas"]
    its = [r[1] for r in runs.values()]
    assert info.converged and solver.last_info.converged
    assert min(its) - 1 <= it <= max(its) + 1, (it, its)
    np.testing.assert_allclose(solver.last_info.residual_history[:3], info.residual_history[:3],
                               rtol=1e-9)
    nx = np.linalg.norm(xo)
    spread = max(np.linalg.norm(r[0] - xo) for r in runs.values()) / nx
    assert np.linalg.norm(xg - xo) / nx <= max(10 * spread, 1e-9), spread


def _oracle_pair_hierarchy(c, d, lo, up, n_coarse_levels, galerkin):
    lv = ogamg.pair_levels(c.n_cells, c.owner, c.neighbour, n_levels=n_coarse_levels)
    levels = [ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour)]
    aggs, prev = [], np.arange(c.n_cells)
    for L in lv:
        f2c = np.empty(levels[-1].n, np.int64)
        f2c[prev] = L["cell_map"]
        aggs.append((L["n_cells_coarse"], f2c))
        levels.append(ogamg.coarse_matrix(levels[-1], L["n_cells_coarse"], f2c,
                                          galerkin=galerkin)[0])
        prev = L["cell_map"]
    return levels, aggs


@pytest.mark.parametrize("n", [16, 24])
def test_pcg_gamg_pair_galerkin_matches_oracle(gpu_device, n):
    """The bench's PCG-GAMG: PCG preconditioned by one V-cycle over the pair
    hierarchy with the Galerkin coarse diagonal, against the oracle PCG over
    the oracle's V-cycle on the oracle's pair maps and Galerkin coarse
    matrices, summing in the device's order: iterations, history and x bit
    for bit."""
    from oracle import krylov as okry
    from oracle import ldu as oldu
    from pyopenfoam_amd import PCGSolver
    c, d, lo, up = system(n, 0.2, 1.0)
    print(f"pcg-gamg pair/galerkin n={n}", flush=True)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    opts = dict(agglomeration="pair", coarse="galerkin", max_levels=20)
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=500, preconditioner="GAMG",
                       gamg_options=opts)
    x, it, _ = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    sizes = GAMGSolver(**opts).hierarchy(m).level_sizes()
    levels, aggs = _oracle_pair_hierarchy(c, d, lo, up, len(sizes) - 1, True)
    assert [(v.n, len(v.owner)) for v in levels] == [tuple(z) for z in sizes]
    pre = lambda r: ogamg.v_cycle(levels, aggs, 0, np.zeros_like(r), r, 2, 2)
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    xd, idev, _, infod = okry.pcg(mv, pre, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                                  rel_tol=1e-8, max_iter=500,
                                  dot=okry.dot_device(c.n_cells, layout="contiguous"))
    assert infod.converged and solver.last_info.converged
    assert it == idev, (it, idev)
    assert bits_equal(np.array(solver.last_info.residual_history), infod.residual_history)
    assert bits_equal(x.cpu().numpy(), xd)


def test_pcg_gamg_mixed_converges(gpu_device):
    """FP32 finest levels in the V-cycle preconditioner: same stopping rule on
    the FP64 residual, converged to the tolerance, iterations close to the FP64
    preconditioner's."""
    from pyopenfoam_amd import PCGSolver
    c, d, lo, up = system(32, 0.2, 1.0)
    m = _matrix(c.n_cells, c.owner, c.neighbour, d, lo, up, gpu_device)
    b = torch.from_numpy(c.b)
    z = torch.zeros(c.n_cells, dtype=torch.float64)
    s64 = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=1000, preconditioner="GAMG")
    x64, it64, _ = s64(m, b, z)
    s32 = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=1000, preconditioner="GAMG",
                    precision="mixed")
    x32, it32, _ = s32(m, b, z)
    assert s64.last_info.converged and s32.last_info.converged
    assert abs(it32 - it64) <= max(3, it64 // 10), (it32, it64)
    from oracle import ldu as oldu
    for xx in (x64, x32):
        xr = xx.cpu().numpy()
        tr = np.linalg.norm(c.b - oldu.ax(d, lo, up, c.owner, c.neighbour, xr)) / np.linalg.norm(c.b)
        assert tr < 1.5e-8, tr
