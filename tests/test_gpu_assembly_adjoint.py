"""GPU assembly (fvm.laplacian + Dirichlet term) bit-exact against the oracle
and the reference's golden coefficients; adjoint gradients against the
reference's DifferentiableLinearSolve (golden), the oracle, central finite
differences and the chain rule through the assembly."""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import adjoint as oadj
from oracle import krylov as okry
from oracle import ldu as oldu
from oracle import precond as opre
from pyopenfoam_amd import LduMatrix, PCGSolver
from pyopenfoam_amd.differentiable import DifferentiableLinearSolve, poisson_solve
from pyopenfoam_amd.discretisation.fvm import LaplacianAssemble, assemble
from pyopenfoam_amd.mesh.fv_mesh import FvMesh
from tests.helpers import bits_equal, case, system

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("n,jit,sig", [(4, 0.0, 0.0), (16, 0.2, 1.0), (32, 0.2, 1.0)])
def test_assembly_bit_exact(gpu_device, n, jit, sig):
    c, d, lo, up = system(n, jit, sig)
    print(f"assembly n={n}", flush=True)
    mesh = FvMesh.from_case(c, gpu_device)
    mat = assemble(c, mesh)
    assert bits_equal(mat.lower.cpu().numpy(), lo)
    assert bits_equal(mat.upper.cpu().numpy(), up)
    assert bits_equal(mat.diag.cpu().numpy(), d)


@pytest.mark.parametrize("name", ["sys_4", "sys_6"])
def test_assembly_golden(gpu_device, name):
    """Reference fvm.laplacian coefficients from the reference's own geometry."""
    g = dict(np.load(G / f"{name}.npz"))
    nfi = g["neighbour"].shape[0]
    n = int(g["n_cells"])
    m = LduMatrix(n, torch.from_numpy(g["owner"][:nfi]), torch.from_numpy(g["neighbour"]),
                  device=gpu_device)
    addr = m.addressing
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device)
    mag = oldu.face_mag(g["face_areas"])  # |S| as torch's norm rounds it
    # patch x0 owners are distinct cells, so the per-cell runs are single faces
    cells = torch.from_numpy(g["dir_cells"].astype(np.int32)).to(gpu_device)
    d, lo, up = LaplacianAssemble.apply(t(g["gamma"][:nfi]), t(g["gamma"][g["dir_faces"]]), addr,
                                        t(mag[:nfi]), t(g["delta"][:nfi]), t(g["cell_volumes"]),
                                        cells, t(g["dir_mag"]), t(g["dir_delta"]))
    assert bits_equal(lo.cpu().numpy(), g["lap_lower"])
    assert bits_equal(up.cpu().numpy(), g["lap_upper"])
    assert bits_equal(d.cpu().numpy(), g["diag"])


def test_differentiable_linear_solve_golden(gpu_device):
    g = dict(np.load(G / "sys_4.npz"))
    nfi = g["neighbour"].shape[0]
    own = torch.from_numpy(g["owner"][:nfi])
    nei = torch.from_numpy(g["neighbour"])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device).requires_grad_(True)
    Ad, Al, Au, b = t(g["diag"]), t(g["lap_lower"]), t(g["lap_upper"]), t(g["b"])
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-14, max_iter=5000)
    x = DifferentiableLinearSolve.apply(Ad, Al, Au, b, solver, own, nei,
                                        torch.zeros(int(g["n_cells"]), dtype=torch.float64,
                                                    device=gpu_device), 1e-300, 5000)
    (torch.from_numpy(g["adj_w"]).to(gpu_device) * x).sum().backward()
    np.testing.assert_allclose(x.detach().cpu().numpy(), g["adj_x"], rtol=1e-10, atol=1e-14)
    for grad, key in ((Ad.grad, "adj_gdiag"), (Al.grad, "adj_glower"), (Au.grad, "adj_gupper"),
                      (b.grad, "adj_gb")):
        ref = g[key]
        np.testing.assert_allclose(grad.cpu().numpy(), ref, rtol=1e-8, atol=1e-12 * np.abs(ref).max())


def _dense_system(c, gamma):
    nfi = len(c.neighbour)
    d, lo, up = oldu.laplacian(gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                               c.owner, c.neighbour)
    d = oldu.dirichlet_diag(d, c.dir_gamma, c.dir_mag, c.dir_delta, c.geo.cell_volumes, c.dir_cells)
    return oldu.dense(d, lo, up, c.owner, c.neighbour), d, lo, up


def test_fused_adjoint_matches_chain_rule_and_fd(gpu_device):
    """dJ/dgamma (J = sum w*x) from the fused adjoint kernel, with the forward
    and adjoint PCG solves run to rel 1e-13, against (a) the chain rule of the
    reference gradients -lam x^T through the assembly with dense solves and (b)
    central finite differences (eps = 1e-6) of dense solves on 8 faces."""
    c = case(10, 0.2, 1.0, seed=3)
    dev = gpu_device
    mesh = FvMesh.from_case(c, dev)
    gam = torch.from_numpy(c.gamma).to(dev).requires_grad_(True)
    gamb = torch.from_numpy(c.dir_gamma).to(dev).requires_grad_(True)
    b = torch.from_numpy(c.b).to(dev)
    wn = np.random.default_rng(9).standard_normal(c.n_cells)
    w = torch.from_numpy(wn).to(dev)
    dd = (torch.from_numpy(c.dir_cells.astype(np.int32)).to(dev), torch.from_numpy(c.dir_mag).to(dev),
          torch.from_numpy(c.dir_delta).to(dev))
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-13, max_iter=20000)
    x = poisson_solve(gam, gamb, b, solver, mesh, dd)
    (w * x).sum().backward()
    gg = gam.grad.cpu().numpy()
    gb = gamb.grad.cpu().numpy()
    A, d, lo, up = _dense_system(c, c.gamma)
    xs = np.linalg.solve(A, c.b)
    lam = np.linalg.solve(A.T, wn)
    own, nei = c.owner, c.neighbour
    nfi = len(nei)
    gd, gl, gu = oadj.ldu_grads(lam, xs, own, nei)
    ref, refb = oadj.gamma_grad(gd, gl, gu, c.geo.mag_sf[:nfi], c.geo.delta[:nfi],
                                c.geo.cell_volumes, own, nei, c.dir_cells, c.dir_mag, c.dir_delta)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(gg, ref, rtol=1e-6, atol=1e-8 * scale)
    np.testing.assert_allclose(gb, refb, rtol=1e-6, atol=1e-8 * np.abs(refb).max())
    rng = np.random.default_rng(4)
    for f in rng.choice(nfi, 8, replace=False):
        vals = []
        for s in (+1, -1):
            gp = c.gamma.copy()
            gp[f] += s * 1e-6
            Ap, *_ = _dense_system(c, gp)
            vals.append(float(wn @ np.linalg.solve(Ap, c.b)))
        fd = (vals[0] - vals[1]) / 2e-6
        assert abs(fd - gg[f]) <= 1e-5 * max(abs(fd), 1e-2 * scale), (f, fd, gg[f])


def test_laplacian_backward_chain(gpu_device):
    """Autograd through LaplacianAssemble equals the oracle's chain rule."""
    c = case(8, 0.2, 1.0, seed=5)
    dev = gpu_device
    mesh = FvMesh.from_case(c, dev)
    gam = torch.from_numpy(c.gamma).to(dev).requires_grad_(True)
    gamb = torch.from_numpy(c.dir_gamma).to(dev).requires_grad_(True)
    mat = assemble(c, mesh, gamma=gam, dir_gamma=gamb)
    rng = np.random.default_rng(2)
    wd, wl, wu = (rng.standard_normal(k) for k in (c.n_cells, len(c.neighbour), len(c.neighbour)))
    L = (mat.diag * torch.from_numpy(wd).to(dev)).sum() + (mat.lower * torch.from_numpy(wl).to(dev)).sum() \
        + (mat.upper * torch.from_numpy(wu).to(dev)).sum()
    L.backward()
    nfi = len(c.neighbour)
    ref, refb = oadj.gamma_grad(wd, wl, wu, c.geo.mag_sf[:nfi], c.geo.delta[:nfi],
                                c.geo.cell_volumes, c.owner, c.neighbour, c.dir_cells, c.dir_mag,
                                c.dir_delta)
    np.testing.assert_allclose(gam.grad.cpu().numpy(), ref, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(gamb.grad.cpu().numpy(), refb, rtol=1e-13)


def test_dirichlet_patches_in_reference_order(gpu_device):
    """Several fixedValue patches add to the diagonal one after another,
    ((diag + s_x0) + s_y0) + s_z0 (laplacian_foam.py:318-366), where s_p is the
    patch's own scatter_add; the cells on the cube's edges and corner carry
    faces of two and three patches.  The same faces given as one patch in a
    shuffled order are summed in that order first and then added once.  Both
    bit-exact against the oracle; the two conventions differ on some cells,
    so a kernel that merged patches (or ignored the order) fails one of them."""
    from pyopenfoam_amd.mesh.geometry import dirichlet_delta
    c = case(12, 0.2, 1.0, seed=2)
    dev = gpu_device
    nfi = len(c.neighbour)
    rng = np.random.default_rng(11)
    patches = []
    for name in ("x0", "y0", "z0"):
        p = c.mesh.patch(name)
        faces = np.arange(p["start"], p["start"] + p["size"])
        patches.append((rng.lognormal(0.0, 1.0, faces.shape[0]), c.geo.mag_sf[faces].copy(),
                        dirichlet_delta(c.mesh, c.geo, faces), c.mesh.owner[faces].copy()))
    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    addr = m.addressing
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d0, lo0, up0 = oldu.laplacian(c.gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi],
                                  c.geo.cell_volumes, c.owner, c.neighbour)
    cat = lambda k, order=None: np.concatenate([p[k] for p in patches])[
        slice(None) if order is None else order]

    def gpu(order, sizes):
        cells = t(cat(3, order).astype(np.int32))
        d, lo, up = LaplacianAssemble.apply(t(c.gamma), t(cat(0, order)), addr, t(c.geo.mag_sf[:nfi]),
                                            t(c.geo.delta[:nfi]), t(c.geo.cell_volumes), cells,
                                            t(cat(1, order)), t(cat(2, order)), sizes)
        assert bits_equal(lo.cpu().numpy(), lo0) and bits_equal(up.cpu().numpy(), up0)
        return d.cpu().numpy()

    ref3 = oldu.dirichlet_diag_patches(d0, patches, c.geo.cell_volumes)
    got3 = gpu(None, tuple(len(p[3]) for p in patches))
    assert bits_equal(got3, ref3)
    shuf = rng.permutation(sum(len(p[3]) for p in patches))
    ref1 = oldu.dirichlet_diag(d0, cat(0, shuf), cat(1, shuf), cat(2, shuf), c.geo.cell_volumes,
                               cat(3, shuf))
    got1 = gpu(shuf, None)
    assert bits_equal(got1, ref1)
    assert not bits_equal(ref1, ref3)  # the conventions are distinguishable here


def test_adjoint_gradient_64_faces_fd_and_reference_semantics(gpu_device):
    """The differentiation workload at a size the oracle solves in minutes:
    48^3 jittered mesh (110,592 cells), lognormal(0, 1) face diffusivity as the
    leaf, J = sum(w * x) with w ~ N(0, 1), forward PCG-DIC to rel 1e-10 and the
    hand-written adjoint (A^T lambda = w with the same solver, fused dJ/dgamma
    kernel).  dJ/dgamma on 64 seeded faces against
    (a) central differences with eps = 1e-6 (absolute perturbation of gamma_f),
        each perturbed system solved on the device from the unperturbed
        solution x*: its initial residual b - A' x* is the O(eps) perturbation
        -dA x* (~1e-7), and the solve stops at rel 1e-8 of it or at an
        absolute 1e-15 ||b|| (~2e-13), just above the ~1.3e-14 floor at which
        FP64 forms b - A x on this system; x+ - x- then carries the difference
        to ~1e-6 relative, and
    (b) the reference DifferentiableLinearSolve semantics evaluated by the
        oracle (tests/golden/make_adjoint48.py: x and lambda from the oracle
        PCG-DIC to rel 1e-10, lambda on the swapped LDU from zero,
        differentiable/linear_solver.py:99-151, chained through the assembly),
        plus the gradient's max |.| and 2-norm over every face.
    Tolerances: (a) 1e-4 of |dJ/dgamma_f| plus 1e-6 of the largest gradient,
    (b) 1e-6 of the largest gradient (two solves to 1e-10 in different
    summation orders); a wrong sign, a swapped owner/neighbour or a dropped
    1/V term is off by O(1)."""
    import hashlib
    dev = gpu_device
    g48 = dict(np.load(G / "adjoint48.npz"))
    c, d, lo, up = system(48, 0.2, 1.0, 4)
    h = hashlib.sha1()
    for a in (d, lo, up, c.b):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    assert np.frombuffer(h.digest(), np.uint8).tolist() == g48["digest"].tolist()
    mesh = FvMesh.from_case(c, dev)
    gam = torch.from_numpy(c.gamma).to(dev).requires_grad_(True)
    gamb = torch.from_numpy(c.dir_gamma).to(dev).requires_grad_(True)
    b = torch.from_numpy(c.b).to(dev)
    wn = np.random.default_rng(48).standard_normal(c.n_cells)
    w = torch.from_numpy(wn).to(dev)
    dd = (torch.from_numpy(c.dir_cells.astype(np.int32)).to(dev),
          torch.from_numpy(c.dir_mag).to(dev), torch.from_numpy(c.dir_delta).to(dev))
    solver = PCGSolver(tolerance=1e-300, rel_tol=1e-10, max_iter=50000)
    x = poisson_solve(gam, gamb, b, solver, mesh, dd)
    (w * x).sum().backward()
    assert solver.last_info.converged
    gg = gam.grad.cpu().numpy()
    xs = x.detach()
    scale = float(g48["grad_max"])
    faces = g48["faces"]
    err_b = np.abs(gg[faces] - g48["grad"]).max() / scale
    print(f"adjoint 48^3: forward {solver.last_info.iterations} iterations (oracle "
          f"{g48['iterations'][0]}); vs the reference-semantics gradient: max error "
          f"{err_b:.2e} of max |g|", flush=True)
    assert err_b <= 1e-6, err_b
    np.testing.assert_allclose(np.abs(gg).max(), scale, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(gg), float(g48["grad_norm"]), rtol=1e-6)
    np.testing.assert_allclose(float((w * x).sum().item()), float(g48["J"]), rtol=1e-8)
    # (a) central differences on the same 64 faces
    eps = 1e-6
    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    fd_solver = PCGSolver(tolerance=1e-15 * float(np.linalg.norm(c.b)), rel_tol=1e-8,
                          max_iter=50000)
    worst = 0.0
    for f in faces:
        xs_pm = []
        for sgn in (+1.0, -1.0):
            gp = c.gamma.copy()
            gp[f] += sgn * eps
            _, dp, lp, upp = _dense_system_coeffs(c, gp)
            m.diag, m.lower, m.upper = (torch.from_numpy(v).to(dev) for v in (dp, lp, upp))
            xp, _, _ = fd_solver(m, b, xs)
            assert fd_solver.last_info.converged
            xs_pm.append(xp)
        fd = float((w * (xs_pm[0] - xs_pm[1])).sum().item()) / (2 * eps)
        err = abs(fd - gg[f])
        worst = max(worst, err / (abs(gg[f]) + 1e-2 * scale))
        assert err <= 1e-4 * abs(gg[f]) + 1e-6 * scale, (int(f), fd, gg[f])
    print(f"  vs central differences on 64 faces: worst error {worst:.2e}", flush=True)


def _dense_system_coeffs(c, gamma):
    nfi = len(c.neighbour)
    d, lo, up = oldu.laplacian(gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                               c.owner, c.neighbour)
    d = oldu.dirichlet_diag(d, c.dir_gamma, c.dir_mag, c.dir_delta, c.geo.cell_volumes, c.dir_cells)
    return None, d, lo, up
