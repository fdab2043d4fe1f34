"""The pressure-equation step on the device against the oracle.

Bar: bit-exact (face weights, matrix, source, flux and velocity corrections)
on jittered meshes with A_p of both signs, including |A_p| below the 1e-30
clamp and cells the |U| <= 1e4 limiter rescales; adjust_phi within 1e-12 of
the oracle (the net flux's summation order is the one freedom) with the same
decision; the assembled system solved with the reference cell pinned."""

import numpy as np
import pytest
import torch

from oracle import pressure as opr
from pyopenfoam_amd.mesh.fv_mesh import FvMesh
from pyopenfoam_amd.solvers import (PCGSolver, adjust_phi, assemble_pressure_equation,
                                    correct_face_flux, correct_velocity, solve_pressure_equation)
from tests.helpers import bits_equal, make_case

pytestmark = pytest.mark.gpu


def _setup(n, jit, seed, dev):
    c = make_case(n, jitter=jit, sigma=0.0, seed=seed)
    mesh = FvMesh.from_case(c, dev)
    rng = np.random.default_rng(seed)
    ap = rng.uniform(0.5, 3.0, c.n_cells) * np.where(rng.random(c.n_cells) < 0.3, -1.0, 1.0)
    ap[: 3] = [0.0, 1e-40, -1e-35]  # below the clamp
    phi = rng.normal(size=c.mesh.n_faces)
    p = rng.normal(size=c.n_cells)
    return c, mesh, ap, phi, p, rng


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("n,jit", [(6, 0.2), (13, 0.3), (24, 0.2)])
def test_pressure_step_matches_oracle(gpu_device, n, jit):
    print(f"pressure n={n} jitter={jit}", flush=True)
    c, mesh, ap, phi, p, rng = _setup(n, jit, n, gpu_device)
    g = c.geo
    nfi = c.mesh.n_internal_faces
    own_all, nei = c.mesh.owner, c.neighbour
    w = opr.face_weights(g.face_centres, g.cell_centres, own_all, nei)
    assert bits_equal(mesh.face_weights.cpu().numpy(), w)

    mat = assemble_pressure_equation(_t(phi, gpu_device), _t(ap, gpu_device), mesh)
    d, lo, up, src = opr.assemble(phi, ap, w, g.mag_sf[:nfi], g.delta[:nfi], g.cell_volumes,
                                  own_all, nei)
    assert bits_equal(mat.lower.cpu().numpy(), lo)
    assert bits_equal(mat.upper.cpu().numpy(), up)
    assert bits_equal(mat.diag.cpu().numpy(), d)
    assert bits_equal(mat.source.cpu().numpy(), src)

    out = correct_face_flux(_t(phi, gpu_device), _t(p, gpu_device), _t(ap, gpu_device), mesh)
    ref = opr.correct_face_flux(phi, p, ap, w, g.mag_sf[:nfi], g.delta[:nfi], own_all, nei)
    assert bits_equal(out.cpu().numpy(), ref)

    hb = rng.normal(size=(c.n_cells, 3))
    hb[5] = [2e4, -1e4, 3e3]  # rescaled by the limiter
    u = correct_velocity(None, _t(hb, gpu_device), _t(p, gpu_device), _t(ap, gpu_device), mesh,
                         alpha=0.7)
    uref = opr.correct_velocity(hb, p, ap, w, g.face_areas, g.cell_volumes, own_all, nei,
                                alpha=0.7)
    assert bits_equal(u.cpu().numpy(), uref)
    assert abs(np.linalg.norm(uref[5]) - 1e4) < 1e-9


def test_face_weights_argument_and_errors(gpu_device):
    c, mesh, ap, phi, p, _ = _setup(5, 0.2, 3, gpu_device)
    nfi = c.mesh.n_internal_faces
    w = np.full(c.mesh.n_faces, 1.0)
    mat = assemble_pressure_equation(_t(phi, gpu_device), _t(ap, gpu_device), mesh,
                                     face_weights=_t(w, gpu_device))
    g = c.geo
    d, lo, up, src = opr.assemble(phi, ap, w, g.mag_sf[:nfi], g.delta[:nfi], g.cell_volumes,
                                  c.mesh.owner, c.neighbour)
    assert bits_equal(mat.lower.cpu().numpy(), lo) and bits_equal(mat.diag.cpu().numpy(), d)
    with pytest.raises(ValueError):
        assemble_pressure_equation(_t(phi[:-1], gpu_device), _t(ap, gpu_device), mesh)
    with pytest.raises(ValueError):
        correct_face_flux(_t(phi, gpu_device), _t(p[:-1], gpu_device), _t(ap, gpu_device), mesh)


def test_adjust_phi(gpu_device):
    c, mesh, ap, phi, p, _ = _setup(8, 0.2, 5, gpu_device)
    nfi = c.mesh.n_internal_faces
    t = _t(phi, gpu_device).clone()
    assert adjust_phi(t, mesh) is True
    adj, ref = opr.adjust_phi(phi, nfi)
    assert adj
    out = t.cpu().numpy()
    assert np.array_equal(out[:nfi], phi[:nfi])
    np.testing.assert_allclose(out[nfi:], ref[nfi:], rtol=0, atol=1e-12 * np.abs(phi).max())
    assert abs(out[nfi:].sum()) < 1e-10
    z = phi.copy()
    z[nfi:] = 0.0
    tz = _t(z, gpu_device)
    assert adjust_phi(tz, mesh) is False and np.array_equal(tz.cpu().numpy(), z)
    t2 = _t(phi, gpu_device)
    assert adjust_phi(t2, mesh, closed=False) is False
    assert np.array_equal(t2.cpu().numpy(), phi)


def test_solve_pressure_equation(gpu_device):
    """Pressure step on a closed box: adjust phiHbyA, assemble, pin cell 0,
    PCG-DIC; the solution satisfies the pinned system."""
    c, mesh, _, phi, _, rng = _setup(12, 0.0, 7, gpu_device)
    ap = rng.uniform(1.0, 2.0, c.n_cells)
    t = _t(phi, gpu_device)
    adjust_phi(t, mesh)
    mat = assemble_pressure_equation(t, _t(ap, gpu_device), mesh)
    d0 = mat.diag.cpu().numpy().copy()
    p0 = torch.zeros(c.n_cells, dtype=torch.float64, device=gpu_device)
    p, iters, res = solve_pressure_equation(mat, p0, PCGSolver(tolerance=1e-300, rel_tol=1e-10,
                                                               max_iter=5000),
                                            tolerance=1e-300, max_iter=5000, reference_cell=0)
    assert mat.diag.cpu().numpy()[0] == d0[0] + max(abs(d0[0]), 1.0) * 1e10
    r = (mat.source - mat.Ax(p)).norm().item()
    assert r <= 1e-8 * mat.source.norm().item()
    assert abs(p[0].item()) < 1e-6 * p.abs().max().item()


def _graph(kind, rng):
    """Internal faces (own, nei) and boundary owners of a random topology:
    'shuffled' -- 300 cells, random orientation and order, rows wider than
    four couplings per side; 'shuffled_far' -- 80,000 cells, columns beyond
    the 16-bit offsets; 'sorted_far' -- owner < neighbour sorted by owner,
    columns beyond 65,535; 'sorted_near' -- the same within a narrow band."""
    if kind == "shuffled":
        n, m = 300, 2000
        own, nei = rng.integers(0, n, m), rng.integers(0, n, m)
    elif kind == "shuffled_far":
        n, m = 80000, 240000
        own, nei = rng.integers(0, n, m), rng.integers(0, n, m)
    else:
        n = 80000 if kind == "sorted_far" else 4000
        reach = n if kind == "sorted_far" else 40
        own = np.repeat(np.arange(n), 3)
        nei = own + rng.integers(1, reach, own.shape[0])
    keep = (own != nei) & (nei < n)
    own, nei = own[keep], nei[keep]
    if kind.startswith("sorted"):
        key = np.unique(own.astype(np.int64) * n + nei)
        own, nei = key // n, key % n
    bown = rng.integers(0, n, n // 3)
    return n, own.astype(np.int64), nei.astype(np.int64), bown.astype(np.int64)


@pytest.mark.parametrize("kind", ["shuffled", "shuffled_far", "sorted_far", "sorted_near"])
def test_pressure_kernels_on_random_topologies(gpu_device, kind):
    """Assembly, face-flux and velocity corrections through the C ABI on
    topologies the mesh tests do not reach: faces in any order and
    orientation, rows of more than four couplings per side, columns beyond the
    16-bit offsets, owner-sorted faces with and without them.  Bit-exact
    against the oracle."""
    from pyopenfoam_amd import LduMatrix, _lib
    from pyopenfoam_amd.core.device import current_stream
    print(f"pressure kernels on {kind}", flush=True)
    rng = np.random.default_rng(17)
    n, own, nei, bown = _graph(kind, rng)
    nfi, nb = own.shape[0], bown.shape[0]
    nfa = nfi + nb
    own_all = np.concatenate([own, bown])
    m = LduMatrix(n, torch.from_numpy(own), torch.from_numpy(nei), device=gpu_device)
    info = m.addressing.info
    if kind == "shuffled":
        assert info.max_lower > 4 and info.max_upper > 4
    h = m.addressing.handle
    ap = rng.uniform(0.5, 3.0, n) * np.where(rng.random(n) < 0.3, -1.0, 1.0)
    w = rng.uniform(0.05, 0.95, nfi)
    mag = rng.uniform(0.5, 2.0, nfi)
    dl = rng.uniform(0.5, 2.0, nfi)
    vol = rng.uniform(0.5, 2.0, n)
    phi = rng.normal(size=nfa)
    p = rng.normal(size=n)
    hb = rng.normal(size=(n, 3))
    sf = rng.normal(size=(nfa, 3))
    order = np.argsort(bown, kind="stable").astype(np.int32)
    bptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(bown, minlength=n), out=bptr[1:])
    T = lambda a: _t(a, gpu_device)
    d_bptr, d_bface = T(bptr.astype(np.int32)), T(order)
    d_ap, d_w, d_mag, d_dl, d_vol = T(ap), T(w), T(mag), T(dl), T(vol)
    d_phi, d_p, d_hb, d_sf = T(phi), T(p), T(hb), T(sf)
    L, q, s = _lib.lib(), _lib.ptr, current_stream(gpu_device)
    diag, lower = torch.empty(n, dtype=torch.float64, device=gpu_device), torch.empty(
        nfi, dtype=torch.float64, device=gpu_device)
    upper, src = torch.empty_like(lower), torch.empty_like(diag)
    _lib.check(L.pfm_pressure_assemble(h, nfa, q(d_bptr), q(d_bface), q(d_phi), q(d_ap), q(d_w),
                                       q(d_mag), q(d_dl), q(d_vol), q(diag), q(lower), q(upper),
                                       q(src), s), "assemble")
    flux = torch.empty_like(d_phi)
    _lib.check(L.pfm_correct_face_flux(h, nfa, q(d_phi), q(d_p), q(d_ap), q(d_w), q(d_mag),
                                       q(d_dl), q(flux), s), "flux")
    u = torch.empty_like(d_hb)
    _lib.check(L.pfm_correct_velocity(h, nfa, q(d_bptr), q(d_bface), q(d_hb), q(d_p), q(d_ap),
                                      q(d_w), q(d_sf), q(d_vol), 0.7, q(u), s), "velocity")
    d, lo, up, sr = opr.assemble(phi, ap, w, mag, dl, vol, own_all, nei)
    assert bits_equal(lower.cpu().numpy(), lo) and bits_equal(upper.cpu().numpy(), up)
    assert bits_equal(diag.cpu().numpy(), d) and bits_equal(src.cpu().numpy(), sr)
    assert bits_equal(flux.cpu().numpy(), opr.correct_face_flux(phi, p, ap, w, mag, dl, own_all, nei))
    assert bits_equal(u.cpu().numpy(), opr.correct_velocity(hb, p, ap, w, sf, vol, own_all, nei,
                                                            alpha=0.7))
