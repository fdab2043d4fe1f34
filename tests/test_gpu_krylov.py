"""GPU PCG / PBiCGStab against the oracle restatements of the reference solvers.

Bar: the same stopping decision; the first residuals equal to 1e-9; on
symmetric systems iterations within one of the oracle and the solution within
1e-9; on the non-symmetric (jittered) systems, where the reference algorithm
itself depends on the dot-product summation order, iterations inside the
oracle's spread over three summation orders and the solution within ten times
that spread.  The tolerances are written in each test."""

import numpy as np
import pytest
import torch

from oracle import krylov as okry
from oracle import ldu as oldu
from oracle import precond as opre
from pyopenfoam_amd import LduMatrix, PBiCGSTABSolver, PCGSolver, create_solver
from pyopenfoam_amd.solvers import PBiCGSolver
from tests.helpers import bits_equal, system

pytestmark = pytest.mark.gpu


def _matrix(c, d, lo, up, dev):
    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    m.diag = torch.from_numpy(d)
    m.lower = torch.from_numpy(lo)
    m.upper = torch.from_numpy(up)
    return m


def _oracle_spread(solve):
    """Run an oracle solve under the three summation orders of the dot
    products: (iteration range, solutions, per-order infos)."""
    out = {}
    for name, dot in okry.DOTS.items():
        out[name] = solve(dot)
    its = [o[1] for o in out.values()]
    return min(its), max(its), out


@pytest.mark.parametrize("n,jit,sig", [(8, 0.0, 0.0), (16, 0.2, 1.0), (32, 0.0, 0.0), (32, 0.2, 1.0),
                                       (20, 0.3, 1.5)])
@pytest.mark.parametrize("pre", ["DIC", "none"])
def test_pcg_matches_oracle(gpu_device, n, jit, sig, pre):
    """Bit-exact: the oracle PCG summing its dot products in the device's
    order (``dot_device``) gives the same iteration count, residual history
    and solution bits as the device solve, on symmetric and non-symmetric
    systems alike.  Against the reference's own summation orders: on
    symmetric systems (no jitter) every order gives the same iteration count
    and the device is within one of it with the solution within 1e-9;
    jittered meshes make the reference matrix non-symmetric (rows scaled by
    1/V) and CG's iteration count then depends on the order itself (the
    oracle spreads by ~10% across orders), so there only convergence and the
    true residual are checked against them."""
    c, d, lo, up = system(n, jit, sig)
    print(f"pcg n={n} jitter={jit} sigma={sig} precond={pre}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=20000, preconditioner=pre)
    x, iters, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    rD = opre.dic_build(d, up, c.owner, c.neighbour) if pre == "DIC" else None
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    solve = lambda dot: okry.pcg(mv, (lambda r: r / rD) if rD is not None else None, c.b,
                                 np.zeros(c.n_cells), tolerance=1e-300, rel_tol=1e-8,
                                 max_iter=20000, dot=dot)
    xd, idev, rdev, infod = solve(okry.dot_device(c.n_cells))
    xg = x.cpu().numpy()
    assert iters == idev and solver.last_info.converged == infod.converged, (iters, idev)
    assert bits_equal(np.array(solver.last_info.residual_history), infod.residual_history)
    assert bits_equal(xg, xd)
    assert solver.last_info.converged
    true_rel = np.linalg.norm(c.b - mv(xg)) / np.linalg.norm(c.b)
    assert true_rel <= 1.1e-8, true_rel
    if jit == 0.0:
        lo_it, hi_it, runs = _oracle_spread(solve)
        xo, io, ro, info = runs["blas"]
        assert info.converged and hi_it == lo_it
        assert lo_it - 1 <= iters <= hi_it + 1, (iters, lo_it, hi_it)
        h = np.array(solver.last_info.residual_history)
        np.testing.assert_allclose(h[:10], info.residual_history[:10], rtol=1e-9)
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-9


@pytest.mark.parametrize("case", ["one_cell", "diag3", "two_cell"])
@pytest.mark.parametrize("pre", ["none", "DIC"])
def test_pcg_breakdown_after_exact_solve(gpu_device, case, pre):
    """With tolerance 0 the monitor never stops the solve: once CG has solved a
    small system exactly, r = 0, p = 0 and p.Ap = 0 is the reference's
    breakdown (pcg.py:150-152, not converged).  On a diagonal system with the
    diagonal preconditioner that happens at iteration 2, an even iteration, at
    which the device still holds iteration 1's x update (x is updated every
    second iteration); other cases end at odd iterations or run to the cap.
    Iterations, decision and every bit of x as the oracle's."""
    own, nei, lo, up = [], [], [], []
    if case == "one_cell":
        d, b = [4.0], [2.0]
    elif case == "diag3":
        d, b = [2.0, 4.0, 8.0], [1.0, 3.0, 5.0]
    else:
        d, b, own, nei, lo, up = [4.0, 4.0], [2.0, 7.0], [0], [1], [-1.0], [-1.0]
    n = len(d)
    m = LduMatrix(n, torch.tensor(own, dtype=torch.int64), torch.tensor(nei, dtype=torch.int64),
                  device=gpu_device)
    m.diag = torch.tensor(d, dtype=torch.float64)
    if own:
        m.lower = torch.tensor(lo, dtype=torch.float64)
        m.upper = torch.tensor(up, dtype=torch.float64)
    solver = PCGSolver(tolerance=0.0, rel_tol=0.0, max_iter=10, preconditioner=pre)
    x, it, _ = solver(m, torch.tensor(b, dtype=torch.float64), torch.zeros(n, dtype=torch.float64))
    dd, ll, uu = np.array(d), np.array(lo), np.array(up)
    oo, ne = np.array(own, np.int64), np.array(nei, np.int64)
    rD = opre.dic_build(dd, uu, oo, ne) if pre == "DIC" else None
    xo, io, _, info = okry.pcg(lambda v: oldu.ax(dd, ll, uu, oo, ne, v),
                               (lambda r: r / rD) if rD is not None else None, np.array(b),
                               np.zeros(n), tolerance=0.0, rel_tol=0.0, max_iter=10,
                               dot=okry.dot_device(n))
    assert (it, solver.last_info.converged) == (io, info.converged), (it, io)
    assert bits_equal(x.cpu().numpy(), xo)
    if case == "one_cell" or (case == "diag3" and pre == "DIC"):
        # breakdown at iteration 2: the history holds r0 and iteration 1
        assert io == 2 and not info.converged
        np.testing.assert_array_equal(xo, np.array(b) / np.array(d))


def test_pcg_breakdown_spans_every_workgroup(gpu_device):
    """The diag3 breakdown above at 2^21 decoupled cells (d in {2, 4, 8}, so
    z = r / d, A z = r and alpha = 1 are exact): r = 0 after iteration 1 and
    p.Ap = 0 at iteration 2, an even iteration, whose kernel applies the x
    update deferred at iteration 1 in every one of its 2048 workgroups before
    the solve ends.  The breakdown must not stop a workgroup that starts after
    another has seen it (pcg.py:150-156 returns the fully updated x): x and
    the iteration count bit-identical to the oracle."""
    n = 1 << 21
    rng = np.random.default_rng(21)
    d = np.array([2.0, 4.0, 8.0])[rng.integers(0, 3, n)]
    b = rng.uniform(-1.0, 1.0, n)
    e = np.zeros(0, np.int64)
    m = LduMatrix(n, torch.from_numpy(e), torch.from_numpy(e), device=gpu_device)
    m.diag = torch.from_numpy(d)
    solver = PCGSolver(tolerance=0.0, rel_tol=0.0, max_iter=10, preconditioner="DIC")
    x, it, _ = solver(m, torch.from_numpy(b), torch.zeros(n, dtype=torch.float64))
    xo, io, _, info = okry.pcg(lambda v: d * v, lambda r: r / d, b, np.zeros(n), tolerance=0.0,
                               rel_tol=0.0, max_iter=10, dot=okry.dot_device(n))
    assert io == 2 and not info.converged
    assert (it, solver.last_info.converged) == (io, info.converged), (it, io)
    assert solver.last_info.status == 3  # PFM_SOLVE_BREAKDOWN
    assert bits_equal(x.cpu().numpy(), xo)
    assert bits_equal(xo, b / d)


def test_pcg_known_answer_and_semantics(gpu_device):
    # 3-cell SPD chain, reference tests/unit/solvers/conftest.py:10-35 builder
    own, nei = torch.tensor([0, 1]), torch.tensor([1, 2])
    m = LduMatrix(3, own, nei, device=gpu_device)
    f64 = dict(dtype=torch.float64)
    m.lower = torch.tensor([-1.0, -1.0], **f64)
    m.upper = torch.tensor([-1.0, -1.0], **f64)
    m.diag = torch.tensor([1.1, 2.1, 1.1], **f64)
    b = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)
    x, iters, res = PCGSolver(tolerance=1e-12, rel_tol=0.0, max_iter=100)(m, b, torch.zeros(3, dtype=torch.float64))
    dense = np.array([[1.1, -1, 0], [-1, 2.1, -1], [0, -1, 1.1]])
    np.testing.assert_allclose(x.cpu().numpy(), np.linalg.solve(dense, b.numpy()), rtol=1e-10)
    # converged at iteration 0: zero rhs -> r0 = 0 < tolerance
    x0, it0, r0 = PCGSolver()(m, torch.zeros(3, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
    assert it0 == 1 and r0 == 0.0
    # max_iter exhaustion returns max_iter + 1 history entries
    PCGSolver(tolerance=1e-300, rel_tol=1e-300, max_iter=1)(m, b, torch.zeros(3, dtype=torch.float64))
    with pytest.raises(ValueError):
        PCGSolver(preconditioner="bogus")(m, b, torch.zeros(3, dtype=torch.float64))
    with pytest.raises(ValueError):
        create_solver("nope")


def _exact_and_envelope(solver, x, iters, solve, n_cells, k_env):
    """The device solve equals the oracle summing in the device's order, bit
    for bit; against the reference's own summation orders it takes the same
    decision and its first k_env residuals lie inside 10x the oracle's
    spread (floor 1e-9)."""
    xd, idev, rdev, infod = solve(okry.dot_device(n_cells))
    assert iters == idev and solver.last_info.converged == infod.converged, (iters, idev)
    h = np.array(solver.last_info.residual_history)
    assert bits_equal(h, infod.residual_history)
    assert bits_equal(x.cpu().numpy(), xd)
    lo_it, hi_it, runs = _oracle_spread(solve)
    info = runs["blas"][3]
    if all(r[3].converged == info.converged for r in runs.values()):
        # the decision is the same in every order (a stalled BiCGStab can
        # go either way near its breakdown; then there is nothing to compare)
        assert solver.last_info.converged == info.converged
    k = min(k_env, len(h), *(len(r[3].residual_history) for r in runs.values()))
    hb = np.array(info.residual_history[:k])
    spread = np.max([np.abs(np.array(r[3].residual_history[:k]) / hb - 1) for r in runs.values()],
                    axis=0)
    spread = np.maximum.accumulate(spread)  # a divergence, once started, only grows
    assert np.all(np.abs(h[:k] / hb - 1) <= np.maximum(10 * spread, 1e-9)), (h[:k], hb, spread)
    return info


@pytest.mark.parametrize("n,jit,sig", [(8, 0.0, 0.0), (16, 0.2, 1.0), (32, 0.2, 1.0)])
@pytest.mark.parametrize("pre", ["DILU", "none"])
def test_pbicgstab_matches_oracle(gpu_device, n, jit, sig, pre):
    """Bit-exact against the oracle in the device's summation order; the same
    decision as the oracle in the reference's orders (with the reference's
    DILU the omega = t.z/t.t step stalls, without a preconditioner it
    converges) and an early history inside their spread.  The step amplifies
    rounding ~1e4x per iteration, so iteration counts differ across orders
    even on symmetric systems."""
    c, d, lo, up = system(n, jit, sig)
    print(f"pbicgstab n={n} jitter={jit} sigma={sig} precond={pre}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    solver = PBiCGSTABSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=3000, preconditioner=pre)
    x, iters, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    rD = opre.dilu_build(d, lo, up, c.owner, c.neighbour) if pre == "DILU" else None
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    solve = lambda dot: okry.pbicgstab(
        mv, (lambda r: r / rD) if rD is not None else None, c.b, np.zeros(c.n_cells),
        tolerance=1e-300, rel_tol=1e-8, max_iter=3000, dot=dot)
    _exact_and_envelope(solver, x, iters, solve, c.n_cells, 12)


@pytest.mark.parametrize("n,jit,sig", [(8, 0.0, 0.0), (16, 0.2, 1.0), (32, 0.2, 1.0)])
@pytest.mark.parametrize("pre", ["DILU", "DIC", "none"])
def test_pbicg_matches_oracle(gpu_device, n, jit, sig, pre):
    """PBiCG (solvers/pbicg.py:107-220): bit-exact against the oracle in the
    device's summation order; against the reference's orders the same
    decision and an early history inside their spread.  The transposed
    products run on the second packed layout."""
    c, d, lo, up = system(n, jit, sig)
    print(f"pbicg n={n} jitter={jit} sigma={sig} precond={pre}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    solver = PBiCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=3000, preconditioner=pre)
    x, iters, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    rD = {"DILU": lambda: opre.dilu_build(d, lo, up, c.owner, c.neighbour),
          "DIC": lambda: opre.dic_build(d, up, c.owner, c.neighbour),
          "none": lambda: None}[pre]()
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    mvt = lambda v: oldu.ax_t(d, lo, up, c.owner, c.neighbour, v)
    solve = lambda dot: okry.pbicg(
        mv, mvt, (lambda r: r / rD) if rD is not None else None, c.b, np.zeros(c.n_cells),
        tolerance=1e-300, rel_tol=1e-8, max_iter=3000, dot=dot)
    info = _exact_and_envelope(solver, x, iters, solve, c.n_cells, 20)
    if info.converged:
        r = c.b - mv(x.cpu().numpy())
        assert np.sqrt(r @ r) <= 1e-6 * solver.last_info.residual_history[0]


def test_pbicg_from_dict_and_flipped_faces(gpu_device):
    """``solver: PBiCG`` through the dictionary factory; a face set with half the
    faces flipped (owner > neighbour) gives the same solve as the oriented one.
    Unjittered (symmetric) system: the reference's PBiCG carries one search
    direction for both the primal and the shadow recurrences, which is BiCG
    only when A = A^T; on the jittered, non-symmetric systems it stalls (the
    oracle does too, and the parity test above checks that decision)."""
    c, d, lo, up = system(12, 0.0, 1.0)
    m = _matrix(c, d, lo, up, gpu_device)
    s1 = create_solver("PBiCG", tolerance=1e-300, rel_tol=1e-8, max_iter=2000)
    assert isinstance(s1, PBiCGSolver)
    x1, i1, _ = s1(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    assert s1.last_info.converged
    msk = np.random.default_rng(3).random(len(lo)) < 0.5
    own = np.where(msk, c.neighbour, c.owner)
    nei = np.where(msk, c.owner, c.neighbour)
    m2 = LduMatrix(c.n_cells, torch.from_numpy(own), torch.from_numpy(nei), device=gpu_device)
    m2.diag = torch.from_numpy(d)
    m2.lower = torch.from_numpy(np.where(msk, up, lo))
    m2.upper = torch.from_numpy(np.where(msk, lo, up))
    s2 = PBiCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=2000, preconditioner="DILU")
    x2, i2, _ = s2(m2, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    assert s2.last_info.converged and abs(i2 - i1) <= 2
    xa, xb = x1.cpu().numpy(), x2.cpu().numpy()
    assert np.linalg.norm(xa - xb) / np.linalg.norm(xa) < 1e-7


@pytest.mark.parametrize("n,jit,sig", [(16, 0.0, 0.0), (32, 0.0, 1.0), (32, 0.2, 1.0)])
def test_mixed_precision_pcg(gpu_device, n, jit, sig):
    """FP32 inner PCG in FP64 iterative refinement reaches the FP64 stopping
    rule on the true residual; its solution is within 1e-6 of the oracle's
    FP64 PCG solution (both at relative residual 1e-8)."""
    c, d, lo, up = system(n, jit, sig)
    print(f"mixed n={n} jitter={jit} sigma={sig}", flush=True)
    m = _matrix(c, d, lo, up, gpu_device)
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=50000, precision="mixed",
                       inner_rel_tol=1e-3)
    x, it, res = solver(m, torch.from_numpy(c.b), torch.zeros(c.n_cells, dtype=torch.float64))
    info = solver.last_info
    assert info.converged
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    xg = x.cpu().numpy()
    true_rel = np.linalg.norm(c.b - mv(xg)) / np.linalg.norm(c.b)
    assert true_rel < 1e-8
    np.testing.assert_allclose(info.residual_history[-1] / info.residual_history[0], true_rel,
                               rtol=1e-6)
    rD = opre.dic_build(d, up, c.owner, c.neighbour)
    xo, io, _, _ = okry.pcg(mv, lambda r: r / rD, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                            rel_tol=1e-8, max_iter=50000)
    err = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"  inner iterations {info.inner_iterations} (fp64 {io}), outer {it}, err {err:.2e}")
    assert err < 1e-6
