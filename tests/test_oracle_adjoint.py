"""The oracle's adjoint chain rule (the reference's -lam x^T gradients pushed
through the assembly) against central finite differences of dense solves,
and against the fused closed form -|S| delta (x_P - x_N)(lam_P/V_P - lam_N/V_N)."""

import numpy as np

from oracle import adjoint as oadj
from oracle import ldu as oldu
from tests.helpers import case


def _dense(c, gamma, gamma_b):
    nfi = len(c.neighbour)
    d, lo, up = oldu.laplacian(gamma, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                               c.owner, c.neighbour)
    d = oldu.dirichlet_diag(d, gamma_b, c.dir_mag, c.dir_delta, c.geo.cell_volumes, c.dir_cells)
    return oldu.dense(d, lo, up, c.owner, c.neighbour)


def test_chain_rule_matches_finite_differences_and_closed_form():
    c = case(6, 0.2, 1.0, seed=8)
    rng = np.random.default_rng(1)
    w = rng.standard_normal(c.n_cells)
    A = _dense(c, c.gamma, c.dir_gamma)
    x = np.linalg.solve(A, c.b)
    lam = np.linalg.solve(A.T, w)
    own, nei = c.owner, c.neighbour
    nfi = len(nei)
    gd, gl, gu = oadj.ldu_grads(lam, x, own, nei)
    g, gb = oadj.gamma_grad(gd, gl, gu, c.geo.mag_sf[:nfi], c.geo.delta[:nfi], c.geo.cell_volumes,
                            own, nei, c.dir_cells, c.dir_mag, c.dir_delta)
    V = c.geo.cell_volumes
    sd = c.geo.mag_sf[:nfi] * c.geo.delta[:nfi]
    closed = -sd * (x[own] - x[nei]) * (lam[own] / V[own] - lam[nei] / V[nei])
    np.testing.assert_allclose(g, closed, rtol=1e-9, atol=1e-12 * np.abs(closed).max())
    closed_b = -c.dir_mag * c.dir_delta * lam[c.dir_cells] * x[c.dir_cells] / V[c.dir_cells]
    np.testing.assert_allclose(gb, closed_b, rtol=1e-9)
    eps = 1e-6
    for f in rng.choice(nfi, 10, replace=False):
        vals = []
        for s in (+1, -1):
            gp = c.gamma.copy()
            gp[f] += s * eps
            vals.append(w @ np.linalg.solve(_dense(c, gp, c.dir_gamma), c.b))
        fd = (vals[0] - vals[1]) / (2 * eps)
        assert abs(fd - g[f]) <= 1e-6 * max(abs(fd), 1e-2 * np.abs(g).max())
    for k in range(3):
        gp = c.dir_gamma.copy()
        gp[k] += eps
        gm = c.dir_gamma.copy()
        gm[k] -= eps
        fd = (w @ np.linalg.solve(_dense(c, c.gamma, gp), c.b)
              - w @ np.linalg.solve(_dense(c, c.gamma, gm), c.b)) / (2 * eps)
        assert abs(fd - gb[k]) <= 1e-6 * max(abs(fd), 1e-2 * np.abs(gb).max())
