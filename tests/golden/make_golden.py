"""Write the golden fixtures of tests/golden/*.npz from the reference itself.

Run in the build container only (it imports pyOpenFOAM from
/root/reference/src; nothing at test time reads the reference):

    python tests/golden/make_golden.py

Only functions that the closed forms and known answers of the tests do not
pin are recorded here:
  * RCM tie-breaking (tools/renumber_mesh.py:163-239)
  * geometry of a jittered mesh (mesh/mesh_geometry.py:35-395)
  * fvm.laplacian on that mesh (discretisation/operators.py:432-518)
  * DIC / DILU factored diagonals (solvers/preconditioners.py:88-133, 228-272)
    and their apply_full sweeps (:153-201, :290-337)
  * PCG / PBiCGStab residual histories (solvers/pcg.py, solvers/pbicgstab.py)
  * GAMG aggregation maps, coarse matrices and V-cycle history (solvers/gamg.py)
  * PairGamgAgglomeration levels (fv_agglomeration/pair_agglomeration.py)
  * GaussSeidelSmoother sweeps (solvers/smoothers.py:92-149)
  * DifferentiableLinearSolve gradients (differentiable/linear_solver.py)
Inputs are small seeded meshes built by pyopenfoam_amd.mesh (structured hex
block, jitter, random cell permutation, RCM, owner<neighbour sort).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, "/root/reference/src")

import torch  # noqa: E402

torch.set_default_dtype(torch.float64)

from pyopenfoam_amd.mesh.hexmesh import jitter_points, relabel_cells, structured_hex  # noqa: E402

from pyfoam.core.ldu_matrix import LduMatrix  # noqa: E402
from pyfoam.differentiable.linear_solver import DifferentiableLinearSolve  # noqa: E402
from pyfoam.discretisation.operators import fvm  # noqa: E402
from pyfoam.fv_agglomeration.pair_agglomeration import PairGamgAgglomeration  # noqa: E402
from pyfoam.mesh.fv_mesh import FvMesh  # noqa: E402
from pyfoam.mesh import mesh_geometry as mg  # noqa: E402
from pyfoam.solvers.gamg import GAMGSolver  # noqa: E402
from pyfoam.solvers.pbicgstab import PBiCGSTABSolver  # noqa: E402
from pyfoam.solvers.pcg import PCGSolver  # noqa: E402
from pyfoam.solvers.preconditioners import DICPreconditioner, DILUPreconditioner  # noqa: E402
from pyfoam.solvers.smoothers import GaussSeidelSmoother  # noqa: E402
from pyfoam.tools.renumber_mesh import _build_adjacency, _reverse_cuthill_mckee  # noqa: E402

I64 = torch.int64


def t(a, dtype=None):
    return torch.as_tensor(np.asarray(a), dtype=dtype)


def permuted(n, seed, jitter=0.0):
    rng = np.random.default_rng(seed)
    m = structured_hex(n)
    if jitter:
        m = jitter_points(m, n, jitter, rng)
    return relabel_cells(m, rng.permutation(m.n_cells)), rng


def ref_rcm(m):
    nfi = m.n_internal_faces
    adj = _build_adjacency(t(m.owner, I64), t(m.neighbour, I64), m.n_cells, nfi)
    return _reverse_cuthill_mckee(adj, m.n_cells).numpy()


def rcm_mesh(n, seed, jitter=0.0):
    m, rng = permuted(n, seed, jitter)
    perm = ref_rcm(m)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.n_cells)
    return relabel_cells(m, inv), rng


def ref_fvmesh(m):
    faces = [t(f, I64) for f in m.faces]
    return FvMesh(t(m.points), faces, t(m.owner, I64), t(m.neighbour, I64), validate=True,
                  device="cpu", dtype=torch.float64)


def ldu(n, own, nei, d, lo, up):
    a = LduMatrix(n, t(own, I64), t(nei, I64), device="cpu", dtype=torch.float64)
    a.diag, a.lower, a.upper = t(d), t(lo), t(up)
    return a


def main():
    out = {}
    # --- RCM ---------------------------------------------------------------
    for n, seed in ((4, 11), (6, 12)):
        m, _ = permuted(n, seed)
        out[f"rcm_{n}"] = dict(owner=m.owner[: m.n_internal_faces], neighbour=m.neighbour,
                               n_cells=m.n_cells, perm=ref_rcm(m))
    # --- jittered 4^3 and 6^3 systems ------------------------------------------
    for n, seed in ((4, 21), (6, 22)):
        m, rng = rcm_mesh(n, seed, jitter=0.2)
        fm = ref_fvmesh(m)
        gamma = rng.lognormal(0.0, 1.0, size=m.n_faces)
        rec = dict(points=m.points, faces=m.faces, owner=m.owner, neighbour=m.neighbour,
                   n_cells=m.n_cells, gamma=gamma,
                   face_centres=fm.face_centres.numpy(), face_areas=fm.face_areas.numpy(),
                   cell_volumes=fm.cell_volumes.numpy(), cell_centres=fm.cell_centres.numpy(),
                   delta=fm.delta_coefficients.numpy())
        phi = torch.zeros(m.n_cells)
        mat = fvm.laplacian(t(gamma), phi, mesh=fm)
        d = mat.diag.numpy().copy()
        lo = mat.lower.numpy().copy()
        up = mat.upper.numpy().copy()
        # make it non-singular: reference laplacianFoam Dirichlet term on patch x0
        p0 = next(p for p in m.patches if p["name"] == "x0")
        fb = np.arange(p0["start"], p0["start"] + p0["size"])
        cells = m.owner[fb]
        S = fm.face_areas.numpy()[fb]
        magS = np.linalg.norm(S, axis=1)
        dP = fm.face_centres.numpy()[fb] - fm.cell_centres.numpy()[cells]
        nh = S / magS[:, None]
        dlt = 1.0 / np.maximum(np.abs((dP * nh).sum(1)), 1e-30)
        coeff = torch.as_tensor(gamma[fb]) * torch.as_tensor(magS) * torch.as_tensor(dlt)
        dd = t(d) + torch.zeros(m.n_cells).scatter_add(0, t(cells, I64),
                                                       coeff / t(fm.cell_volumes.numpy()[cells]))
        rec.update(lap_diag=d, lap_lower=lo, lap_upper=up, dir_faces=fb, dir_cells=cells,
                   dir_mag=magS, dir_delta=dlt, diag=dd.numpy())
        own = m.owner[: m.n_internal_faces]
        nei = m.neighbour
        A = ldu(m.n_cells, own, nei, dd.numpy(), lo, up)
        dic = DICPreconditioner(A)
        dilu = DILUPreconditioner(A)
        r = rng.uniform(-1, 1, m.n_cells)
        rec.update(dic=dic._precond_diag.numpy(), dilu=dilu._precond_diag.numpy(), r=r,
                   dic_full=dic.apply_full(t(r)).numpy(), dilu_full=dilu.apply_full(t(r)).numpy())
        b = rng.uniform(-1, 1, m.n_cells)
        b = b - b.mean()
        rec["b"] = b
        for name, solver in (("pcg_dic", PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=2000)),
                             ("pcg_none", PCGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=2000,
                                                    preconditioner="none")),
                             ("bicg_dilu", PBiCGSTABSolver(tolerance=1e-300, rel_tol=1e-8,
                                                           max_iter=300)),
                             ("bicg_none", PBiCGSTABSolver(tolerance=1e-300, rel_tol=1e-8,
                                                           max_iter=2000, preconditioner="none"))):
            from pyfoam.solvers.residual import ResidualMonitor
            mon = ResidualMonitor(tolerance=1e-300, rel_tol=1e-8)
            x, info = solver._solve(A, t(b), torch.zeros(m.n_cells), mon, solver.max_iter)
            rec[f"{name}_x"] = x.numpy()
            rec[f"{name}_hist"] = np.array(info.residual_history)
            rec[f"{name}_converged"] = info.converged
        if n == 4:
            gs = GaussSeidelSmoother(A)
            x0 = rng.uniform(-1, 1, m.n_cells)
            rec["gs_x0"] = x0
            rec["gs_x"] = gs.smooth(t(x0), t(b), 2).numpy()
            # differentiable solve: J = sum(w * x)
            w = rng.standard_normal(m.n_cells)
            rec["adj_w"] = w
            Ad, Al, Au = (t(v).clone().requires_grad_(True) for v in (dd.numpy(), lo, up))
            bb = t(b).clone().requires_grad_(True)
            solver = PCGSolver(tolerance=1e-300, rel_tol=1e-14, max_iter=5000)
            x = DifferentiableLinearSolve.apply(Ad, Al, Au, bb, solver, t(own, I64), t(nei, I64),
                                                torch.zeros(m.n_cells), 1e-300, 5000)
            (t(w) * x).sum().backward()
            rec.update(adj_x=x.detach().numpy(), adj_gdiag=Ad.grad.numpy(),
                       adj_glower=Al.grad.numpy(), adj_gupper=Au.grad.numpy(),
                       adj_gb=bb.grad.numpy())
        # GAMG: hierarchy maps, coarse matrices, cycle history
        g = GAMGSolver(tolerance=1e-300, rel_tol=1e-8, max_iter=25, min_cells_coarse=4)
        agg = GAMGSolver._build_aggregates(A, max(m.n_cells // 2, 4))
        coarse = GAMGSolver._build_coarse_matrix(A, agg)
        rec.update(gamg_f2c=agg.fine_to_coarse.numpy(), gamg_ncoarse=agg.n_coarse,
                   gamg_c_diag=coarse.diag.numpy(), gamg_c_lower=coarse.lower.numpy(),
                   gamg_c_upper=coarse.upper.numpy(), gamg_c_owner=coarse.owner.numpy(),
                   gamg_c_neighbour=coarse.neighbour.numpy())
        from pyfoam.solvers.residual import ResidualMonitor
        mon = ResidualMonitor(tolerance=1e-300, rel_tol=1e-8)
        xg, info = g._solve(A, t(b), torch.zeros(m.n_cells), mon, 25)
        rec.update(gamg_x=xg.numpy(), gamg_hist=np.array(info.residual_history))
        # pair agglomeration
        pa = PairGamgAgglomeration(m.n_cells, t(own, I64), t(nei, I64))
        levels = pa.build_levels(n_levels=5)
        rec["pair_nlevels"] = len(levels)
        for k, lv in enumerate(levels):
            rec[f"pair_map{k}"] = lv["cell_map"].numpy()
            rec[f"pair_n{k}"] = np.array([lv["n_cells_fine"], lv["n_cells_coarse"]])
        out[f"sys_{n}"] = rec

    for name, rec in out.items():
        np.savez_compressed(HERE / f"{name}.npz", **{k: np.asarray(v) for k, v in rec.items()})
        print("wrote", name, sum(np.asarray(v).nbytes for v in rec.values()), "bytes")


if __name__ == "__main__":
    main()
