"""The oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py) and against the reference's own known answers.

Bit-exact where the reference computation has a fixed operation order
(scatter sums, recurrences, sweeps, graph algorithms); to a stated
tolerance where the reference leaves the order to torch (dot products and
norms) or to Python's set ordering (cell reference points of the geometry).
"""

import math
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import gamg as ogamg
from oracle import krylov as okry
from oracle import ldu as oldu
from oracle import mesh as omesh
from oracle import precond as opre
from pyopenfoam_amd.cases import rcm_permutation
from pyopenfoam_amd.mesh.hexmesh import HexMesh
from tests.helpers import bits_equal

G = Path(__file__).resolve().parent / "golden"


def load(name):
    return dict(np.load(G / f"{name}.npz"))


@pytest.mark.parametrize("name", ["rcm_4", "rcm_6"])
def test_rcm_golden(name):
    g = load(name)
    n = int(g["n_cells"])
    assert np.array_equal(omesh.rcm(n, g["owner"], g["neighbour"]), g["perm"])
    assert np.array_equal(rcm_permutation(n, g["owner"], g["neighbour"]), g["perm"])


@pytest.fixture(params=["sys_4", "sys_6"], scope="module")
def sysg(request):
    g = load(request.param)
    g["nfi"] = g["neighbour"].shape[0]
    g["own"] = g["owner"][: g["nfi"]]
    return g


def test_geometry_golden(sysg):
    g = sysg
    n = int(g["n_cells"])
    if n <= 64:  # loop restatement (slow), small mesh only
        geo = omesh.geometry(g["points"], list(g["faces"]), g["owner"], g["neighbour"], n)
        for k in ("face_centres", "face_areas", "cell_volumes", "cell_centres", "delta"):
            np.testing.assert_allclose(geo[k], g[k], rtol=1e-13, atol=1e-16)
    pg = SimpleNamespace(**omesh.geometry_vec(g["points"], g["faces"], g["owner"], g["neighbour"],
                                              n))
    np.testing.assert_allclose(pg.face_centres, g["face_centres"], rtol=1e-14, atol=1e-16)
    # components that cancel to ~0 are compared against the face area scale
    np.testing.assert_allclose(pg.face_areas, g["face_areas"], rtol=1e-13,
                               atol=1e-13 * np.abs(g["face_areas"]).max())
    np.testing.assert_allclose(pg.cell_volumes, g["cell_volumes"], rtol=1e-13)
    np.testing.assert_allclose(pg.cell_centres, g["cell_centres"], rtol=1e-13)
    np.testing.assert_allclose(pg.delta, g["delta"], rtol=1e-12)
    assert np.all(g["delta"][g["nfi"]:] == 0.0)  # boundary delta is 0 in the reference


def test_dirichlet_delta_golden(sysg):
    """The Dirichlet patch's boundary delta (laplacian_foam.py:341-353): the
    oracle and the host helper the case builder uses both give the reference's
    own values on its geometry; a face of zero area takes the normal
    (1, 1, 1)/sqrt(3) and a face through its cell centre the 1e-30 clamp."""
    from pyopenfoam_amd.mesh.geometry import MeshGeometry, dirichlet_delta
    g = sysg
    fb = g["dir_faces"]
    od = omesh.dirichlet_delta(g["face_centres"], g["face_areas"], g["cell_centres"],
                               g["owner"], fb)
    np.testing.assert_allclose(od, g["dir_delta"], rtol=1e-15)
    mag = oldu.face_mag(g["face_areas"])
    assert bits_equal(mag[fb], g["dir_mag"])
    geo = MeshGeometry(face_centres=g["face_centres"], face_areas=g["face_areas"], mag_sf=mag,
                       cell_volumes=g["cell_volumes"], cell_centres=g["cell_centres"],
                       delta=g["delta"])
    pd = dirichlet_delta(SimpleNamespace(owner=g["owner"]), geo, fb)
    assert bits_equal(pd, od)
    # degenerate faces, closed forms: d = (0.3, -0.2, 0.5) from the cell centre
    fc = np.array([[0.3, -0.2, 0.5], [0.0, 0.0, 0.0], [0.3, -0.2, 0.5]])
    fa = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.0], [0.0, 0.0, 4.0]])
    cc, own, ids = np.zeros((1, 3)), np.zeros(3, np.int64), np.arange(3)
    want = [math.sqrt(3.0) / 0.6, 1e30, 2.0]  # |d.(1,1,1)|/sqrt(3) = 0.6/sqrt(3); clamp; 1/0.5
    od = omesh.dirichlet_delta(fc, fa, cc, own, ids)
    np.testing.assert_allclose(od, want, rtol=1e-15)
    geo = MeshGeometry(face_centres=fc, face_areas=fa, mag_sf=oldu.face_mag(fa),
                       cell_volumes=np.ones(1), cell_centres=cc, delta=np.zeros(3))
    assert bits_equal(dirichlet_delta(SimpleNamespace(owner=own), geo, ids), od)


def test_laplacian_golden_bit_exact(sysg):
    g = sysg
    nfi = g["nfi"]
    mag = oldu.face_mag(g["face_areas"])
    d, lo, up = oldu.laplacian(g["gamma"][:nfi], mag[:nfi], g["delta"][:nfi], g["cell_volumes"],
                               g["own"], g["neighbour"])
    assert bits_equal(lo, g["lap_lower"])
    assert bits_equal(up, g["lap_upper"])
    assert bits_equal(d, g["lap_diag"])
    dd = oldu.dirichlet_diag(d, g["gamma"][g["dir_faces"]], g["dir_mag"], g["dir_delta"],
                             g["cell_volumes"], g["dir_cells"])
    assert bits_equal(dd, g["diag"])


def test_dic_dilu_golden_bit_exact(sysg):
    g = sysg
    d, lo, up = g["diag"], g["lap_lower"], g["lap_upper"]
    own, nei = g["own"], g["neighbour"]
    assert bits_equal(opre.dic_build(d, up, own, nei), g["dic"])
    assert bits_equal(opre.dic_build_py(d, up, own, nei), g["dic"])
    assert bits_equal(opre.dilu_build(d, lo, up, own, nei), g["dilu"])
    assert bits_equal(opre.dilu_build_py(d, lo, up, own, nei), g["dilu"])
    assert bits_equal(opre.dic_apply_full(g["dic"], up, own, nei, g["r"]), g["dic_full"])
    assert bits_equal(opre.dilu_apply_full(g["dilu"], lo, up, own, nei, g["r"]), g["dilu_full"])


@pytest.mark.parametrize("kind", ["pcg_dic", "pcg_none", "bicg_dilu", "bicg_none"])
def test_krylov_golden(sysg, kind):
    g = sysg
    d, lo, up = g["diag"], g["lap_lower"], g["lap_upper"]
    own, nei = g["own"], g["neighbour"]
    mv = lambda v: oldu.ax(d, lo, up, own, nei, v)
    if kind.endswith("dic"):
        pre = lambda r: r / opre.dic_build(d, up, own, nei)
    elif kind.endswith("dilu"):
        pre = lambda r: r / opre.dilu_build(d, lo, up, own, nei)
    else:
        pre = None
    fn = okry.pcg if kind.startswith("pcg") else okry.pbicgstab
    mx = 300 if kind == "bicg_dilu" else 2000
    runs = {k: fn(mv, pre, g["b"], np.zeros(len(d)), tolerance=1e-300, rel_tol=1e-8, max_iter=mx,
                  dot=dot) for k, dot in okry.DOTS.items()}
    ref_h = g[f"{kind}_hist"]
    x, it, res, info = runs["blas"]
    assert bool(g[f"{kind}_converged"]) == info.converged
    # per-entry envelope: 10x the spread of the oracle's own summation orders
    # (PBiCGStab with the reference's DILU amplifies rounding ~1e4x per step)
    k = min(20, len(ref_h), *(len(r[3].residual_history) for r in runs.values()))
    hb = np.array(info.residual_history[:k])
    spread = np.max([np.abs(np.array(r[3].residual_history[:k]) / hb - 1) for r in runs.values()],
                    axis=0)
    spread = np.maximum.accumulate(spread)  # a divergence, once started, only grows
    assert np.all(np.abs(ref_h[:k] / hb - 1) <= np.maximum(10 * spread, 1e-12))
    its = [r[1] for r in runs.values()]
    if info.converged:
        # the reference's count lies inside the oracle's summation-order spread
        assert min(its) - 1 <= len(ref_h) <= max(its) + 1, (len(ref_h), its)
        spread = max(np.linalg.norm(r[0] - x) for r in runs.values()) / np.linalg.norm(x)
        assert np.linalg.norm(g[f"{kind}_x"] - x) / np.linalg.norm(x) <= max(10 * spread, 1e-10)


def test_gamg_golden(sysg):
    g = sysg
    d, lo, up = g["diag"], g["lap_lower"], g["lap_upper"]
    own, nei = g["own"], g["neighbour"]
    n = len(d)
    nagg, f2c = ogamg.aggregate(n, own, nei, max(n // 2, 4))
    assert nagg == int(g["gamg_ncoarse"])
    assert np.array_equal(f2c, g["gamg_f2c"])
    nagg2, f2c2 = ogamg.aggregate_py(n, own, nei, max(n // 2, 4))
    assert nagg2 == nagg and np.array_equal(f2c2, f2c)
    L0 = ogamg.Level(n, d, lo, up, own, nei)
    c, _ = ogamg.coarse_matrix(L0, nagg, f2c)
    assert bits_equal(c.diag, g["gamg_c_diag"])
    assert bits_equal(c.lower, g["gamg_c_lower"])
    assert bits_equal(c.upper, g["gamg_c_upper"])
    assert np.array_equal(c.owner, g["gamg_c_owner"])
    assert np.array_equal(c.neighbour, g["gamg_c_neighbour"])
    x, it, res, info = ogamg.gamg(L0, g["b"], np.zeros(n), tolerance=1e-300, rel_tol=1e-8,
                                  max_iter=25, min_cells_coarse=4)
    # the V-cycle is fixed-order arithmetic: x bit-exact, history to the norm's rounding
    assert bits_equal(x, g["gamg_x"])
    np.testing.assert_allclose(info.residual_history, g["gamg_hist"], rtol=1e-13)


def test_pair_agglomeration_golden(sysg):
    g = sysg
    lv = ogamg.pair_levels(int(g["n_cells"]), g["own"], g["neighbour"], n_levels=5)
    lvp = ogamg.pair_levels_py(int(g["n_cells"]), g["own"], g["neighbour"], n_levels=5)
    assert len(lv) == int(g["pair_nlevels"]) == len(lvp)
    for k, l in enumerate(lv):
        assert np.array_equal(l["cell_map"], g[f"pair_map{k}"])
        assert np.array_equal(lvp[k]["cell_map"], g[f"pair_map{k}"])
        assert [l["n_cells_fine"], l["n_cells_coarse"]] == g[f"pair_n{k}"].tolist()


def test_gauss_seidel_and_adjoint_golden():
    g = load("sys_4")
    nfi = g["neighbour"].shape[0]
    own, nei = g["owner"][:nfi], g["neighbour"]
    d, lo, up = g["diag"], g["lap_lower"], g["lap_upper"]
    x = ogamg.gauss_seidel(d, lo, up, own, nei, g["b"], g["gs_x0"], 2)
    assert bits_equal(x, g["gs_x"])
    # DifferentiableLinearSolve: J = sum(w*x), solves to rel 1e-14
    mv = lambda v: oldu.ax(d, lo, up, own, nei, v)
    rD = opre.dic_build(d, up, own, nei)
    xs, *_ = okry.pcg(mv, lambda r: r / rD, g["b"], np.zeros(len(d)), tolerance=1e-300,
                      rel_tol=1e-14, max_iter=5000)
    mvt = lambda v: oldu.ax(d, up, lo, own, nei, v)
    rDt = opre.dic_build(d, lo, own, nei)
    lam, *_ = okry.pcg(mvt, lambda r: r / rDt, g["adj_w"], np.zeros(len(d)), tolerance=1e-300,
                       rel_tol=1e-14, max_iter=5000)
    gd, gl, gu = oadj.ldu_grads(lam, xs, own, nei)
    np.testing.assert_allclose(xs, g["adj_x"], rtol=1e-10, atol=1e-14)
    for a, b in ((gd, "adj_gdiag"), (gl, "adj_glower"), (gu, "adj_gupper"), (lam, "adj_gb")):
        np.testing.assert_allclose(a, g[b], rtol=1e-8, atol=1e-14 * np.abs(g[b]).max())


def test_laplacian_two_cell_known_answers():
    """The oracle's geometry + fvm.laplacian on the reference tests' two-cell hex
    mesh (two unit cubes stacked in z, tests/unit/discretisation/test_laplacian.py:18-65)
    against that test's assertions (:82-170) with this mesh's values: |S| = delta
    = V = 1 on the shared face, so lower = upper = -1 and diag = [1, 1] (to the
    rounding of the geometry); A 1 = 0 exactly; D per cell [1, 3] interpolates to
    2 on the face (weights 1/2)."""
    pts = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
           [0, 1, 1], [0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2]]
    faces = np.array([[4, 5, 6, 7], [0, 3, 2, 1], [0, 1, 5, 4], [3, 7, 6, 2], [0, 4, 7, 3],
                      [1, 2, 6, 5], [8, 9, 10, 11], [4, 5, 9, 8], [7, 11, 10, 6],
                      [4, 8, 11, 7], [5, 6, 10, 9]])
    own = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1])
    nei = np.array([1])
    geo = omesh.geometry_vec(np.array(pts, np.float64), faces, own, nei, 2)
    np.testing.assert_allclose(geo["cell_volumes"], [1.0, 1.0], rtol=0, atol=1e-14)
    np.testing.assert_allclose(geo["cell_centres"], [[0.5, 0.5, 0.5], [0.5, 0.5, 1.5]],
                               rtol=0, atol=1e-14)
    args = (geo["mag_sf"][:1], geo["delta"][:1], geo["cell_volumes"], own[:1], nei)
    d, lo, up = oldu.laplacian(np.ones(1), *args)
    np.testing.assert_allclose(lo, [-1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(up, [-1.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(d, [1.0, 1.0], rtol=0, atol=1e-12)
    assert oldu.ax(d, lo, up, own[:1], nei, np.ones(2)).tolist() == [0.0, 0.0]
    d2, lo2, up2 = oldu.laplacian(np.full(1, 2.0), *args)
    assert bits_equal(lo2, 2.0 * lo) and bits_equal(up2, 2.0 * up) and bits_equal(d2, 2.0 * d)
    dc, loc, upc = oldu.laplacian(np.array([0.5 * 1.0 + 0.5 * 3.0]), *args)
    np.testing.assert_allclose(loc, [-2.0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(dc, [2.0, 2.0], rtol=0, atol=1e-12)


def test_ax_known_answers():
    # reference tests/unit/core/test_ldu_matrix.py:151-202
    o, q = np.array([0]), np.array([1])
    assert oldu.ax(np.array([4.0, 4.0]), np.array([-1.0]), np.array([-2.0]), o, q,
                   np.array([1.0, 3.0])).tolist() == [1.0, 10.0]
    assert oldu.ax_t(np.array([4.0, 4.0]), np.array([-1.0]), np.array([-2.0]), o, q,
                     np.array([1.0, 3.0])).tolist() == [-2.0, 11.0]
    o3, q3 = np.array([0, 1]), np.array([1, 2])
    assert oldu.ax(np.array([4.0, 6.0, 4.0]), -np.ones(2), -np.ones(2), o3, q3,
                   np.array([1.0, 2.0, 3.0])).tolist() == [2.0, 8.0, 10.0]


def test_ax_matches_dense():
    rng = np.random.default_rng(0)
    g = load("sys_6")
    nfi = g["neighbour"].shape[0]
    own, nei = g["owner"][:nfi], g["neighbour"]
    d, lo, up = g["diag"], g["lap_lower"], g["lap_upper"]
    A = oldu.dense(d, lo, up, own, nei)
    x = rng.standard_normal(len(d))
    np.testing.assert_allclose(oldu.ax(d, lo, up, own, nei, x), A @ x, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(oldu.ax_t(d, lo, up, own, nei, x), A.T @ x, rtol=1e-12, atol=1e-9)


def test_dic_is_exact_cholesky_on_a_chain():
    """On a chain (tridiagonal) incomplete Cholesky has no dropped fill, so the
    DIC diagonal satisfies A = (D + L) D^-1 (D + L^T) exactly for a symmetric
    matrix: a property test independent of the recurrence's code."""
    n = 12
    rng = np.random.default_rng(5)
    own, nei = np.arange(n - 1), np.arange(1, n)
    off = -rng.uniform(0.5, 1.5, n - 1)
    d = np.bincount(own, -off, n) + np.bincount(nei, -off, n) + rng.uniform(0.1, 1.0, n)
    D = opre.dic_build(d, off, own, nei)
    A = oldu.dense(d, off, off, own, nei)
    Lm = np.tril(A, -1)
    M = (np.diag(D) + Lm) @ np.diag(1.0 / D) @ (np.diag(D) + Lm.T)
    np.testing.assert_allclose(M, A, rtol=1e-12, atol=1e-12)


def test_dilu_two_sided_update_known_answer():
    """DILU updates both cells of a face with the values they hold before the
    face (preconditioners.py:258-268): on 2 cells with d = (4, 4),
    lower = -1, upper = -2, both become 4 - 2/4 = 3.5 (a DIC-style one-sided
    recurrence would leave d0 = 4).  On a 3-cell chain the second face sees
    the first face's update of cell 1."""
    o, q = np.array([0]), np.array([1])
    d = np.array([4.0, 4.0])
    assert opre.dilu_build(d, np.array([-1.0]), np.array([-2.0]), o, q).tolist() == [3.5, 3.5]
    assert opre.dic_build(d, np.array([-2.0]), o, q).tolist() == [4.0, 3.0]
    o3, q3 = np.array([0, 1]), np.array([1, 2])
    d3 = np.array([4.0, 6.0, 4.0])
    one = -np.ones(2)
    got = opre.dilu_build(d3, one, one, o3, q3)
    d1 = 6.0 - 1.0 / 4.0
    assert got.tolist() == [4.0 - 1.0 / 6.0, d1 - 1.0 / 4.0, 4.0 - 1.0 / d1]


def test_jacobi_oracle_known_answers():
    """JacobiSmoother known answers of the reference's tests
    (tests/unit/solvers/test_smoothers.py:180-206): identity -> x = b, diag(2)
    -> x = b/2 after one step with omega = 1; plus a 2-cell hand computation
    that fails if the off-diagonal term, omega or the 1/|d| scaling is
    dropped: d = (4, 5), lower = upper = -1, x0 = (1, 2), b = (3, 1),
    omega = 0.5: A x0 = (4*1 - 2, 5*2 - 1) = (2, 9), r = (1, -8),
    x1 = (1 + 0.5/4 * 1, 2 + 0.5/5 * -8) = (1.125, 1.2)."""
    e = np.zeros(0, np.int64)
    x = ogamg.jacobi(np.ones(3), np.zeros(0), np.zeros(0), e, e, np.array([2.0, 3.0, 4.0]),
                     np.zeros(3), 1, omega=1.0)
    assert np.array_equal(x, [2.0, 3.0, 4.0])
    x = ogamg.jacobi(np.full(2, 2.0), np.zeros(0), np.zeros(0), e, e, np.array([4.0, 6.0]),
                     np.zeros(2), 1, omega=1.0)
    assert np.array_equal(x, [2.0, 3.0])
    x = ogamg.jacobi(np.array([4.0, 5.0]), np.array([-1.0]), np.array([-1.0]), np.array([0]),
                     np.array([1]), np.array([3.0, 1.0]), np.array([1.0, 2.0]), 1, omega=0.5)
    np.testing.assert_allclose(x, [1.125, 1.2], rtol=0, atol=1e-15)
    # the exact solution is a fixed point
    A = np.array([[4.0, -1.0], [-1.0, 5.0]])
    xs = np.linalg.solve(A, [3.0, 1.0])
    x = ogamg.jacobi(np.array([4.0, 5.0]), np.array([-1.0]), np.array([-1.0]), np.array([0]),
                     np.array([1]), np.array([3.0, 1.0]), xs, 3, omega=0.7)
    np.testing.assert_allclose(x, xs, rtol=1e-14)


def test_gauss_seidel_v_cycle_oracle_pins():
    """The V-cycle with Gauss-Seidel level smoothing (the device's
    level_smoother="GaussSeidel" is compared with it), on the symmetric 8^3
    system (a hierarchy of 512, 256, 61, 15 and 3 cells): the one-level cycle
    is the coarsest solve (exact), the exact solution is a fixed point (a
    dropped smoother term or a correction from b instead of r moves it), and
    the coarse correction contracts the A-norm error below what the same
    smoothing alone reaches (a dropped or sign-flipped correction does not:
    measured 0.82x, 1.00x and 1.22x of smoothing alone after six cycles)."""
    from tests.helpers import system
    c, d, lo, up = system(8, 0.0, 0.5, 3)
    A = oldu.dense(d, lo, up, c.owner, c.neighbour)
    assert np.array_equal(A, A.T)
    L0 = ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour)
    levels, aggs = ogamg.build_hierarchy(L0, 10, 10)
    assert [lv.n for lv in levels] == [512, 256, 61, 15, 3]
    lc = levels[-1]
    rhs = np.array([1.0, -2.0, 0.5])
    one = ogamg.v_cycle([lc], [], 0, np.zeros(lc.n), rhs, smoother=ogamg.smooth_gs)
    np.testing.assert_allclose(one, np.linalg.solve(
        oldu.dense(lc.diag, lc.lower, lc.upper, lc.owner, lc.neighbour), rhs), rtol=1e-12)
    xs = np.linalg.solve(A, c.b)
    xf = ogamg.v_cycle(levels, aggs, 0, xs, c.b, 2, 2, ogamg.smooth_gs)
    assert np.abs(xf - xs).max() <= 1e-13 * np.abs(xs).max()
    anorm = lambda e: math.sqrt(e @ (A @ e))
    x = np.zeros(c.n_cells)
    y = np.zeros(c.n_cells)
    for _ in range(6):
        x = ogamg.v_cycle(levels, aggs, 0, x, c.b, 2, 2, ogamg.smooth_gs)
        y = ogamg.smooth_gs(L0, y, c.b, 4)
    assert anorm(x - xs) < 0.9 * anorm(y - xs)


def test_smooth_solver_oracle_known_answers():
    """SmoothSolver's Gauss-Seidel sweep (smooth_solver.py:150-180): a 2-cell
    system swept twice by hand (d = (4, 4), l = u = -1, b = (2, 7), x0 = 0:
    (0.5, 1.875) then (0.96875, 1.9921875), exact in binary); identity -> x = b;
    against a dense forward Gauss-Seidel on a jittered 4^3 system (1e-13);
    and the reference tests' identity/2-cell behaviours through the full loop
    (tests/unit/solvers/test_smooth_solver.py:15-60, 86-150)."""
    from oracle import smooth as osm
    one = lambda k: osm.smooth_gs_sweeps(np.array([4.0, 4.0]), np.array([-1.0]), np.array([-1.0]),
                                         np.array([0]), np.array([1]), np.array([2.0, 7.0]),
                                         np.zeros(2), k)
    assert one(1).tolist() == [0.5, 1.875]
    assert one(2).tolist() == [0.96875, 1.9921875]
    e = np.zeros(0, np.int64)
    x = osm.smooth_gs_sweeps(np.ones(3), np.zeros(0), np.zeros(0), e, e, np.array([2.0, 3.0, 4.0]),
                             np.zeros(3), 1)
    assert x.tolist() == [2.0, 3.0, 4.0]
    from tests.helpers import system
    c, d, lo, up = system(4, 0.2, 1.0)
    A = oldu.dense(d, lo, up, c.owner, c.neighbour)
    xs = np.random.default_rng(3).standard_normal(c.n_cells)
    xd = xs.copy()
    for _ in range(3):
        for i in range(c.n_cells):
            xd[i] = (c.b[i] - (A[i] @ xd - A[i, i] * xd[i])) / A[i, i]
    xo = osm.smooth_gs_sweeps(d, lo, up, c.owner, c.neighbour, c.b, xs, 3)
    np.testing.assert_allclose(xo, xd, rtol=1e-13, atol=1e-13 * np.abs(xd).max())
    # full loops: the reference tests' residual bars for DIC and DILU on 2x2 systems
    own, nei = np.array([0]), np.array([1])
    xs, it, res, info = osm.smooth_solver(np.array([4.0, 4.0]), np.array([-1.0]), np.array([-1.0]),
                                          own, nei, np.array([2.0, 7.0]), np.zeros(2),
                                          smoother="DIC", tolerance=1e-10, max_iter=100)
    A2 = np.array([[4.0, -1.0], [-1.0, 4.0]])
    assert info.converged and np.linalg.norm(np.array([2.0, 7.0]) - A2 @ xs) < 1e-4
    xs, it, res, info = osm.smooth_solver(np.array([4.0, 5.0]), np.array([-1.0]), np.array([-2.0]),
                                          own, nei, np.array([3.0, 3.0]), np.zeros(2),
                                          smoother="DILU", omega=0.5, tolerance=1e-10, max_iter=200)
    A2 = np.array([[4.0, -1.0], [-2.0, 5.0]])  # the reference's bars (rel_tol 0.01)
    assert info.converged and np.linalg.norm(np.array([3.0, 3.0]) - A2 @ xs) < 0.1


def test_diagonal_solver_oracle_known_answers():
    """DiagonalSolver (tests/unit/solvers/test_diagonal.py:15-126): x = b/d,
    zero RHS gives zero, a near-zero diagonal is clamped to 1e-30 (finite)."""
    from oracle import smooth as osm
    x, r = osm.diagonal_solver(np.array([2.0, 4.0, 5.0]), np.array([4.0, 8.0, 10.0]))
    assert x.tolist() == [2.0, 2.0, 2.0] and r == 0.0
    x, r = osm.diagonal_solver(np.array([1e-35, 2.0]), np.array([1.0, 2.0]))
    assert np.isfinite(x).all() and x[1] == 1.0 and x[0] == 1.0 / 1e-30


@pytest.mark.parametrize("layout", ["strided", "contiguous", "tiles1", "tiles5", "tiles32"])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 257, 6161, 131_073, 600_001])
def test_dot_device_sums_every_product_once(n, layout):
    """The device-order dot (the device solvers' reduction tree: grid, XCD
    block remap, slice walk, wave butterflies, partials) adds every product
    exactly once: integer-valued products sum exactly, a single non-zero
    product comes back unchanged wherever it sits, and random data lie inside
    the rounding bound of the exactly rounded sum.  The sizes reach a one-slice
    grid, ragged last slices, a grid that is not a multiple of 8 (remap tail)
    and the 2048-workgroup cap with several slices per wave."""
    rng = np.random.default_rng(n)
    if layout.startswith("tiles"):  # the batched solver's order, tiles of T slices
        dot = okry.dot_device(n, tile_slices=int(layout[5:]))
    else:
        dot = okry.dot_device(n, layout=layout)
    a = rng.integers(-1000, 1000, n)
    c = rng.integers(-1000, 1000, n)
    assert dot(a.astype(float), c.astype(float)) == float(np.dot(a, c))
    for i in sorted({0, n // 3, n - 1}):
        e = np.zeros(n)
        e[i] = 3.0
        assert dot(e, np.full(n, 0.5)) == 1.5
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    pr = x * y
    assert abs(dot(x, y) - math.fsum(pr)) <= n * 2.0**-53 * np.abs(pr).sum()


def test_pbicg_oracle_pins():
    """PBiCG (solvers/pbicg.py:107-220): with a symmetric matrix, a symmetric
    diagonal preconditioner and r~0 = r0, BiCG's iterates are CG's (the shadow
    recurrence is the primal one), so the oracle's PBiCG must follow its PCG:
    the same histories to rounding over the first 20 iterations (later ones
    part as rounding grows, as two CG codes do) and the same iteration count
    within 2; on a 2x2 non-symmetric
    system it reaches the exact solution in two iterations (finite
    termination); the identity system converges in one."""
    from tests.helpers import system
    c, d, lo, up = system(6, 0.0, 1.0)
    rD = opre.dic_build(d, up, c.owner, c.neighbour)
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    mvt = lambda v: oldu.ax_t(d, lo, up, c.owner, c.neighbour, v)
    pre = lambda r: r / rD
    xb, ib, rb, infob = okry.pbicg(mv, mvt, pre, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                                   rel_tol=1e-8, max_iter=500)
    xc, ic, rc, infoc = okry.pcg(mv, pre, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                                 rel_tol=1e-8, max_iter=500)
    assert infob.converged and abs(ib - ic) <= 2
    np.testing.assert_allclose(infob.residual_history[:20], infoc.residual_history[:20],
                               rtol=1e-12)
    np.testing.assert_allclose(xb, xc, rtol=0, atol=1e-6 * np.abs(xc).max())
    A = np.array([[4.0, -1.0], [-2.0, 5.0]])
    own, nei = np.array([0]), np.array([1])
    d2, l2, u2 = np.array([4.0, 5.0]), np.array([-1.0]), np.array([-2.0])
    x2, i2, r2, info2 = okry.pbicg(lambda v: oldu.ax(d2, l2, u2, own, nei, v),
                                   lambda v: oldu.ax_t(d2, l2, u2, own, nei, v), None,
                                   np.array([3.0, 3.0]), np.zeros(2), tolerance=1e-12,
                                   rel_tol=1e-14, max_iter=10)
    assert info2.converged and i2 <= 3
    np.testing.assert_allclose(x2, np.linalg.solve(A, [3.0, 3.0]), rtol=1e-12)


@pytest.mark.parametrize("n,jit", [(3, 0.0), (4, 0.3), (5, 0.2)])
def test_geometry_vec_matches_loop_restatement(n, jit):
    """The vectorised geometry against the per-face / per-cell loops (both in
    the reference's order except the centroid weighting, t*(S/4) against
    (t*S)/4, a last-bit difference)."""
    from pyopenfoam_amd.mesh.hexmesh import jitter_points, relabel_cells, structured_hex
    rng = np.random.default_rng(n)
    m = structured_hex(n)
    if jit:
        m = jitter_points(m, n, jit, rng)
    m = relabel_cells(m, rng.permutation(m.n_cells))
    a = omesh.geometry(m.points, list(m.faces), m.owner, m.neighbour, m.n_cells)
    b = omesh.geometry_vec(m.points, m.faces, m.owner, m.neighbour, m.n_cells)
    for k in ("face_centres", "face_areas", "cell_volumes", "cell_centres", "delta"):
        scale = np.abs(a[k]).max()
        np.testing.assert_allclose(b[k], a[k], rtol=1e-13, atol=1e-14 * scale, err_msg=k)
    np.testing.assert_array_equal(b["face_centres"], a["face_centres"])
    assert np.all(b["cell_volumes"] > 0)
    np.testing.assert_allclose(b["cell_volumes"].sum(), 1.0, rtol=1e-12)  # unit cube


def _pe_case(n, jit, seed=0):
    from tests.helpers import make_case
    c = make_case(n, jitter=jit, sigma=0.0, seed=seed)
    nfi = c.mesh.n_internal_faces
    return c, nfi, c.geo


def test_pressure_oracle_matrix_pins():
    """w = 0.5 and A_p = 1 reduce the pressure matrix to the Laplacian with
    D = 1 bit for bit; w = 1 (owner only) to D = 1/|A_P|: a swapped
    interpolation direction or a missing abs fails these."""
    from oracle import pressure as opr
    c, nfi, g = _pe_case(6, 0.2)
    own, nei = c.owner, c.neighbour
    mag, dl, vol = g.mag_sf[:nfi], g.delta[:nfi], g.cell_volumes
    phi = np.zeros(c.mesh.n_faces)
    d0, l0, u0 = oldu.laplacian(np.ones(nfi), mag, dl, vol, own, nei)
    d, lo, up, src = opr.assemble(phi, np.ones(c.n_cells), np.full(c.mesh.n_faces, 0.5), mag, dl,
                                  vol, c.mesh.owner, nei)
    assert bits_equal(lo, l0) and bits_equal(up, u0) and bits_equal(d, d0)
    assert not src.any()
    ap = np.random.default_rng(1).uniform(-3.0, 3.0, c.n_cells)
    d1, l1, u1 = oldu.laplacian(1.0 / np.abs(ap)[own], mag, dl, vol, own, nei)
    d, lo, up, _ = opr.assemble(phi, ap, np.ones(c.mesh.n_faces), mag, dl, vol, c.mesh.owner, nei)
    assert bits_equal(lo, l1) and bits_equal(up, u1) and bits_equal(d, d1)


def test_pressure_oracle_source_pins():
    from oracle import pressure as opr
    c, nfi, g = _pe_case(5, 0.2)
    nfa = c.mesh.n_faces
    own_all, nei = c.mesh.owner, c.neighbour
    args = (np.ones(c.n_cells), np.full(nfa, 0.5), g.mag_sf[:nfi], g.delta[:nfi], g.cell_volumes,
            own_all, nei)
    # a uniform velocity has zero divergence: sum over each closed cell of S.U0
    u0 = np.array([0.3, -1.1, 0.7])
    phi = g.face_areas @ u0
    src = opr.assemble(phi, *args)[3]
    assert np.abs(src).max() <= 1e-12 * np.abs(phi).max() / g.cell_volumes.min()
    # unit flux on one internal face leaves its owner and enters its neighbour
    f = nfi // 3
    phi = np.zeros(nfa)
    phi[f] = 1.0
    src = opr.assemble(phi, *args)[3]
    exp = np.zeros(c.n_cells)
    exp[own_all[f]] = -1.0 / g.cell_volumes[own_all[f]]
    exp[nei[f]] = 1.0 / g.cell_volumes[nei[f]]
    np.testing.assert_array_equal(src, exp)
    # ... and on a boundary face it leaves its owner
    phi = np.zeros(nfa)
    phi[nfa - 2] = 2.0
    src = opr.assemble(phi, *args)[3]
    assert src[own_all[nfa - 2]] == -2.0 / g.cell_volumes[own_all[nfa - 2]]
    assert np.count_nonzero(src) == 1


def test_pressure_oracle_corrections_pins():
    """Uniform orthogonal mesh, p = g.x: the flux correction is -(1/A_p) g.S_f
    on internal faces; the Gauss gradient is g in interior cells, so U =
    HbyA - (alpha/A_p) g there; the limiter caps |U| at 1e4."""
    from oracle import pressure as opr
    c, nfi, geo = _pe_case(6, 0.0)
    own_all, nei = c.mesh.owner, c.neighbour
    w = opr.face_weights(geo.face_centres, geo.cell_centres, own_all, nei)
    np.testing.assert_allclose(w[:nfi], 0.5, rtol=1e-14)
    assert np.all(w[nfi:] == 1.0)
    gvec = np.array([0.4, -1.3, 2.1])
    p = geo.cell_centres @ gvec
    ap = np.full(c.n_cells, 4.0)
    phi = np.random.default_rng(2).normal(size=c.mesh.n_faces)
    out = opr.correct_face_flux(phi, p, ap, w, geo.mag_sf[:nfi], geo.delta[:nfi], own_all, nei)
    corr = -(geo.face_areas[:nfi] @ gvec) / 4.0
    np.testing.assert_allclose(out[:nfi] - phi[:nfi], corr, rtol=1e-9, atol=1e-13)
    assert np.array_equal(out[nfi:], phi[nfi:])
    hb = np.random.default_rng(3).normal(size=(c.n_cells, 3))
    u = opr.correct_velocity(hb, p, ap, w, geo.face_areas, geo.cell_volumes, own_all, nei,
                             alpha=0.5)
    cc = geo.cell_centres
    h = 1.0 / 6
    interior = np.all((cc > h) & (cc < 1 - h), axis=1)
    assert interior.sum() == 64
    np.testing.assert_allclose(u[interior], hb[interior] - 0.5 / 4.0 * gvec, rtol=1e-10,
                               atol=1e-12)
    big = np.zeros((c.n_cells, 3))
    big[:, 0] = 3e4
    u = opr.correct_velocity(big, np.zeros(c.n_cells), ap, w, geo.face_areas, geo.cell_volumes,
                             own_all, nei)
    np.testing.assert_allclose(u[:, 0], 1e4, rtol=1e-15)


def test_pressure_oracle_adjust_phi_pins():
    from oracle import pressure as opr
    rng = np.random.default_rng(4)
    phi = rng.normal(size=40)
    adj, out = opr.adjust_phi(phi, 30)
    assert adj and abs(out[30:].sum()) < 1e-14 and np.array_equal(out[:30], phi[:30])
    phi2 = phi.copy()
    phi2[30:] = 0.0
    assert opr.adjust_phi(phi2, 30)[0] is False
    assert opr.adjust_phi(phi, 30, closed=False)[0] is False
    assert opr.adjust_phi(phi[:30], 30)[0] is False


@pytest.mark.parametrize("n,jit,sig,agglom", [(6, 0.0, 0.0, "pair"), (6, 0.2, 1.0, "pair"),
                                               (7, 0.2, 1.0, "greedy"), (5, 0.0, 1.0, "greedy")])
def test_galerkin_coarse_matrix_is_ptap(n, jit, sig, agglom):
    """``coarse_matrix(galerkin=True)`` against the closed form P^T A P of the
    piecewise-constant prolongation P[i, f2c[i]] = 1: its diagonal on every
    system, and the whole matrix on systems with lower = upper (unjittered
    cubes, where the reference's first-seen face orientation cannot matter).
    The reference coarse diagonal (intra-aggregate couplings dropped) misses
    P^T A P wherever an aggregate holds a face, so a dropped term, a sign or
    lower-only sum fails here."""
    from tests.helpers import system
    c, d, lo, up = system(n, jit, sig)
    lv0 = ogamg.Level(c.n_cells, d, lo, up, c.owner, c.neighbour)
    if agglom == "pair":
        L = ogamg.pair_levels(c.n_cells, c.owner, c.neighbour, n_levels=1)[0]
        nagg, f2c = L["n_cells_coarse"], L["cell_map"].astype(np.int64)
    else:
        nagg, f2c = ogamg.aggregate(c.n_cells, c.owner, c.neighbour, max(c.n_cells // 2, 10))
    A = oldu.dense(d, lo, up, c.owner, c.neighbour)
    P = np.zeros((c.n_cells, nagg))
    P[np.arange(c.n_cells), f2c] = 1.0
    Ac = P.T @ A @ P
    gal, fmap = ogamg.coarse_matrix(lv0, nagg, f2c, galerkin=True)
    ref, _ = ogamg.coarse_matrix(lv0, nagg, f2c)
    scale = np.abs(Ac).max()
    np.testing.assert_allclose(gal.diag, np.diag(Ac), rtol=1e-12, atol=1e-13 * scale)
    assert (fmap < 0).any()
    assert np.abs(ref.diag - np.diag(Ac)).max() > 1e-6 * scale
    if jit == 0.0:
        np.testing.assert_allclose(lo, up, rtol=1e-14)  # volumes equal up to rounding
        np.testing.assert_allclose(oldu.dense(gal.diag, gal.lower, gal.upper, gal.owner,
                                              gal.neighbour), Ac, rtol=1e-12, atol=1e-13 * scale)
    # off-diagonal couplings are the reference's in both forms
    assert bits_equal(gal.lower, ref.lower) and bits_equal(gal.upper, ref.upper)


def test_torch_cpu_pcg_matches_numpy_oracle():
    """The torch-CPU restatement (the benchmark's CPU baseline) against the
    golden-pinned numpy oracle: A x equal to rounding (scatter order); on a
    symmetric system (unjittered: every summation order takes the same
    iterations) the same iteration count within one and the solution within
    1e-9; on a jittered (non-symmetric) one, where CG's iteration count
    depends on the summation order itself, convergence and a true residual at
    the tolerance; a planted sign error in the restatement's product fails."""
    import torch
    from oracle import torch_ref
    from tests.helpers import system
    for jit in (0.0, 0.2):
        c, d, lo, up = system(10, jit, 1.0)
        A = torch_ref.LduCpu(d, lo, up, c.owner, c.neighbour)
        x = np.random.default_rng(2).standard_normal(c.n_cells)
        mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
        np.testing.assert_allclose(A.ax(torch.from_numpy(x)).numpy(), mv(x), rtol=1e-13,
                                   atol=1e-12)
        rD = opre.dic_build(d, up, c.owner, c.neighbour)
        xt, it, _, conv, hist = torch_ref.pcg(A, c.b, np.zeros(c.n_cells), rD, tolerance=1e-300,
                                              rel_tol=1e-8, max_iter=5000)
        runs = [okry.pcg(mv, lambda r: r / rD, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                         rel_tol=1e-8, max_iter=5000, dot=dot) for dot in okry.DOTS.values()]
        its = [r[1] for r in runs]
        assert conv and all(r[3].converged for r in runs)
        np.testing.assert_allclose(hist[:5], runs[0][3].residual_history[:5], rtol=1e-10)
        xo = runs[0][0]
        if jit == 0.0:
            assert max(its) == min(its) and abs(it - its[0]) <= 1, (it, its)
            assert np.linalg.norm(xt.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)
        else:
            assert np.linalg.norm(c.b - mv(xt.numpy())) <= 1.1e-8 * np.linalg.norm(c.b)
        bad = torch_ref.LduCpu(d, -lo, up, c.owner, c.neighbour)
        assert not np.allclose(bad.ax(torch.from_numpy(x)).numpy(), mv(x))
