"""Write tests/golden/adjoint48.npz: the gradient dJ/dgamma of J = sum(w * x)
on the 48^3 jittered system (lognormal(0, 1), seed 4; w ~ N(0, 1), seed 48)
under the reference's DifferentiableLinearSolve semantics
(differentiable/linear_solver.py:99-151), evaluated by the oracle: x from the
oracle PCG-DIC (solvers/pcg.py:97-187) to rel 1e-10, lambda from the same PCG
on the swapped LDU (A^T) from zero, grad_diag = -lambda x, grad_lower =
-lambda[own] x[nei], grad_upper = -lambda[nei] x[own], chained through
fvm.laplacian (oracle.adjoint.gamma_grad).  Recorded: the 64 seeded faces
(seed 64) and their gradient, the gradient's max |.| and 2-norm over all
faces, J, and a SHA-1 of the system's (diag, lower, upper, b).

Run in the build container (CPU only; two 110,592-cell oracle solves, a few
minutes):

    python tests/golden/make_adjoint48.py
"""

from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import adjoint as oadj  # noqa: E402
from oracle import krylov as okry  # noqa: E402
from oracle import ldu as oldu  # noqa: E402
from oracle import precond as opre  # noqa: E402
from tests.helpers import system  # noqa: E402


def digest(*arrays) -> np.ndarray:
    h = hashlib.sha1()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def main():
    c, d, lo, up = system(48, 0.2, 1.0, 4)
    nfi = len(c.neighbour)
    w = np.random.default_rng(48).standard_normal(c.n_cells)
    rD = opre.dic_build(d, up, c.owner, c.neighbour)
    mv = lambda v: oldu.ax(d, lo, up, c.owner, c.neighbour, v)
    x, ix, _, info = okry.pcg(mv, lambda r: r / rD, c.b, np.zeros(c.n_cells), tolerance=1e-300,
                              rel_tol=1e-10, max_iter=50000)
    print(f"forward: {ix} iterations, converged {info.converged}", flush=True)
    rDt = opre.dic_build(d, lo, c.owner, c.neighbour)
    mvt = lambda v: oldu.ax(d, up, lo, c.owner, c.neighbour, v)
    lam, il, _, infol = okry.pcg(mvt, lambda r: r / rDt, w, np.zeros(c.n_cells),
                                 tolerance=1e-300, rel_tol=1e-10, max_iter=50000)
    print(f"adjoint: {il} iterations, converged {infol.converged}", flush=True)
    gd, gl, gu = oadj.ldu_grads(lam, x, c.owner, c.neighbour)
    g, _ = oadj.gamma_grad(gd, gl, gu, c.geo.mag_sf[:nfi], c.geo.delta[:nfi],
                           c.geo.cell_volumes, c.owner, c.neighbour)
    faces = np.random.default_rng(64).choice(nfi, 64, replace=False)
    np.savez_compressed(HERE / "adjoint48.npz", faces=faces, grad=g[faces],
                        grad_max=np.abs(g).max(), grad_norm=np.linalg.norm(g), J=float(w @ x),
                        iterations=np.array([ix, il]),
                        converged=np.array([info.converged, infol.converged]),
                        digest=digest(d, lo, up, c.b))
    print(f"wrote adjoint48.npz: max |g| {np.abs(g).max():.4e}, |g| {np.linalg.norm(g):.4e}")


if __name__ == "__main__":
    main()
