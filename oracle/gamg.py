"""GAMG solver, pair agglomeration and Gauss-Seidel (oracle; test infrastructure only).

Restates:
  * ``aggregate``      -- GAMGSolver._build_aggregates (solvers/gamg.py:164-239)
  * ``coarse_matrix``  -- GAMGSolver._build_coarse_matrix (:242-333); with
                          ``galerkin=True`` the build's additive extension
                          that keeps the couplings the reference drops
                          inside an aggregate (:297-299) on the coarse
                          diagonal (parity pinned by the dense P^T A P of
                          the piecewise-constant prolongation, see
                          tests/test_oracle_golden.py)
  * ``gamg``           -- GAMGSolver._solve with _v_cycle/_smooth/_restrict/
                          _prolongate (:340-585)
  * ``pair_levels``    -- PairGamgAgglomeration.build_levels/_pair_cells
                          (fv_agglomeration/pair_agglomeration.py:47-131)
  * ``gauss_seidel``   -- GaussSeidelSmoother.smooth (solvers/smoothers.py:92-149)
  * ``jacobi``         -- JacobiSmoother.smooth (solvers/smoothers.py:173-200)
The sequential graph passes and sweeps run in ``c/oracle.c``; the V-cycle
arithmetic is numpy with the reference's op order.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._clib import clib, f64, i64, p
from .krylov import Monitor, _ret
from .ldu import _scatter, ax


@dataclass
class Level:
    n: int
    diag: np.ndarray
    lower: np.ndarray
    upper: np.ndarray
    owner: np.ndarray
    neighbour: np.ndarray


def aggregate(n, owner, neighbour, target):
    o, q = i64(owner), i64(neighbour)
    f2c = np.empty(n, dtype=np.int64)
    nagg = clib().orc_gamg_aggregate(n, o.shape[0], p(o), p(q), int(target), p(f2c))
    return int(nagg), f2c


def aggregate_py(n, owner, neighbour, target):
    adj = [[] for _ in range(n)]
    for f in range(len(owner)):
        a, b = int(owner[f]), int(neighbour[f])
        adj[a].append(b)
        adj[b].append(a)
    f2c = np.full(n, -1, dtype=np.int64)
    nagg = 0
    for cell in range(n):
        if f2c[cell] >= 0:
            continue
        aid = nagg
        nagg += 1
        f2c[cell] = aid
        for nb in adj[cell]:
            if f2c[nb] < 0:
                f2c[nb] = aid
        if nagg >= target:
            for rem in range(n):
                if f2c[rem] < 0:
                    for nb in adj[rem]:
                        if f2c[nb] >= 0:
                            f2c[rem] = f2c[nb]
                            break
                    else:
                        f2c[rem] = 0
            break
    f2c[f2c < 0] = 0
    return nagg, f2c


def coarse_matrix(level: Level, nagg: int, f2c: np.ndarray, galerkin: bool = False):
    """Returns (coarse Level, face_map) with face_map[f] = coarse face or -1.

    The reference (gamg.py:242-333): coarse diag = scatter-sum of the fine
    diagonal; a face between two aggregates adds its lower / upper to the
    coarse face keyed by the aggregate pair (first-seen orientation); a face
    inside one aggregate is dropped (:297-299).  ``galerkin``: such a face
    instead adds lower + upper to its aggregate's diagonal, one face after
    another in face order after the diagonal sum, so the coarse diagonal is
    the diagonal of P^T A P for the piecewise-constant P of ``f2c``."""
    diag = _scatter(level.diag, f2c, nagg)
    nf = level.owner.shape[0]
    o, q, m = i64(level.owner), i64(level.neighbour), i64(f2c)
    lo, up = f64(level.lower), f64(level.upper)
    c_own = np.empty(max(nf, 1), np.int64)
    c_nei = np.empty(max(nf, 1), np.int64)
    c_lo = np.empty(max(nf, 1))
    c_up = np.empty(max(nf, 1))
    fmap = np.empty(max(nf, 1), np.int64)
    ncf = clib().orc_gamg_coarse_faces(nf, p(o), p(q), p(m), p(lo), p(up), p(c_own), p(c_nei),
                                       p(c_lo), p(c_up), p(fmap))
    if galerkin and nf:
        inside = np.nonzero(fmap[:nf] < 0)[0]
        np.add.at(diag, m[o[inside]], lo[inside] + up[inside])  # sequential, face order
    return (Level(nagg, diag, c_lo[:ncf].copy(), c_up[:ncf].copy(), c_own[:ncf].copy(),
                  c_nei[:ncf].copy()), fmap[:nf].copy())


def build_hierarchy(level0: Level, max_levels=10, min_cells_coarse=10):
    levels, aggs = [level0], []
    cur = level0
    for _ in range(max_levels):
        if cur.n <= min_cells_coarse:
            break
        target = max(cur.n // 2, min_cells_coarse)
        nagg, f2c = aggregate(cur.n, cur.owner, cur.neighbour, target)
        if nagg >= cur.n:
            break
        aggs.append((nagg, f2c))
        cur, _ = coarse_matrix(cur, nagg, f2c)
        levels.append(cur)
    return levels, aggs


def _lax(lv: Level, x):
    return ax(lv.diag, lv.lower, lv.upper, lv.owner, lv.neighbour, x)


def smooth(lv: Level, x, b, n_it):
    if n_it <= 0:
        return x
    a = np.abs(lv.diag)
    inv = 1.0 / np.where(a < 1e-30, 1e-30, a)
    for _ in range(n_it):
        r = b - _lax(lv, x)
        x = x + inv * r
    return x


def smooth_gs(lv: Level, x, b, n_it):
    """Level smoothing by symmetric Gauss-Seidel sweeps (the build's
    ``level_smoother="GaussSeidel"`` extension; the reference GAMG uses
    ``smooth``)."""
    if n_it <= 0:
        return x
    return gauss_seidel(lv.diag, lv.lower, lv.upper, lv.owner, lv.neighbour, b, x, n_it)


def v_cycle(levels, aggs, k, x, b, n_pre=2, n_post=2, smoother=smooth):
    lv = levels[k]
    if k == len(levels) - 1:
        return smoother(lv, x, b, 50)
    x = smoother(lv, x, b, n_pre)
    r = b - _lax(lv, x)
    nagg, f2c = aggs[k]
    rc = _scatter(r, f2c, nagg)
    ec = np.zeros_like(rc)
    ec = v_cycle(levels, aggs, k + 1, ec, rc, n_pre, n_post, smoother)
    x = x + ec[f2c]
    x = smoother(lv, x, b, n_post)
    return x


def gamg(level0: Level, b, x0, *, tolerance=1e-6, rel_tol=0.01, max_iter=100, min_iter=0,
         n_pre_smooth=2, n_post_smooth=2, max_levels=10, min_cells_coarse=10,
         level_smoother="jacobi"):
    levels, aggs = build_hierarchy(level0, max_levels, min_cells_coarse)
    mon = Monitor(tolerance, rel_tol, min_iter)
    x = np.array(x0, dtype=np.float64, copy=True)
    converged = False
    sm = smooth_gs if level_smoother == "gauss_seidel" else smooth
    for i in range(1, max_iter + 1):
        x = v_cycle(levels, aggs, 0, x, b, n_pre_smooth, n_post_smooth, sm)
        r = b - _lax(level0, x)
        if mon.update(r, i):
            converged = True
            break
    return _ret(x, mon, converged)


def pair_levels(n_cells, owner, neighbour, n_levels=5):
    """List of dicts {level, n_cells_fine, n_cells_coarse, cell_map} (cumulative maps)."""
    o, q = i64(owner), i64(neighbour)
    nf = q.shape[0]
    levels = []
    n_coarse = n_cells
    cell_map = np.arange(n_cells, dtype=np.int64)
    for level in range(n_levels):
        if n_coarse <= 4:
            break
        out = np.empty(n_cells, np.int64)
        k = clib().orc_pair_level(n_cells, n_coarse, nf, p(o), p(q), p(cell_map), p(out))
        if k == 0:
            break
        if k > n_coarse * 0.9:
            break
        levels.append({"level": level, "n_cells_fine": n_coarse, "n_cells_coarse": int(k),
                       "cell_map": out})
        cell_map = out
        n_coarse = int(k)
    return levels


def pair_levels_py(n_cells, owner, neighbour, n_levels=5):
    nf = len(neighbour)
    levels = []
    n_coarse = n_cells
    cell_map = np.arange(n_cells, dtype=np.int64)
    for level in range(n_levels):
        if n_coarse <= 4:
            break
        paired = np.zeros(n_coarse, bool)
        new_map = cell_map.copy()
        for i in range(nf):
            c0 = int(cell_map[owner[i]])
            c1 = int(cell_map[neighbour[i]])
            if c0 == c1 or paired[c0] or paired[c1]:
                continue
            new_map[new_map == c1] = c0
            paired[c0] = paired[c1] = True
        if paired.sum() == 0:
            break
        uniq = np.unique(new_map)
        remap = {int(v): i for i, v in enumerate(uniq)}
        new_map = np.array([remap[int(v)] for v in new_map], dtype=np.int64)
        k = len(uniq)
        if k > n_coarse * 0.9:
            break
        levels.append({"level": level, "n_cells_fine": n_coarse, "n_cells_coarse": k,
                       "cell_map": new_map})
        cell_map = new_map
        n_coarse = k
    return levels


def gauss_seidel(diag, lower, upper, owner, neighbour, b, x, n_iterations, omega=1.0):
    d, lo, up, bb = f64(diag), f64(lower), f64(upper), f64(b)
    o, q = i64(owner), i64(neighbour)
    xx = np.array(x, dtype=np.float64, copy=True)
    clib().orc_gauss_seidel(d.shape[0], o.shape[0], p(o), p(q), p(d), p(lo), p(up), p(bb),
                            float(omega), int(n_iterations), p(xx))
    return xx


def jacobi(diag, lower, upper, owner, neighbour, b, x, n_iterations, omega=2.0 / 3.0):
    """JacobiSmoother: inv = 1/max(|diag|, 1e-30) once (:178); each iteration
    r = b - A x, x = x + (omega * inv) * r (:196-198, Python's left-to-right
    product)."""
    d = f64(diag)
    inv = 1.0 / np.maximum(np.abs(d), 1e-30)
    xx = np.array(x, dtype=np.float64, copy=True)
    bb = f64(b)
    for _ in range(int(n_iterations)):
        r = bb - ax(d, f64(lower), f64(upper), owner, neighbour, xx)
        xx = xx + (omega * inv) * r
    return xx
