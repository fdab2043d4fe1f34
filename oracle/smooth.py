"""SmoothSolver and DiagonalSolver (oracle; test infrastructure only).

Restates:
  * ``smooth_gs_sweeps`` -- SmoothSolver's Gauss-Seidel sweeps
                            (solvers/smooth_solver.py:150-180), in ``c/oracle.c``
  * ``smooth_solver``    -- SmoothSolver._solve with the GaussSeidel, DIC and
                            DILU smoothers (:96-232)
  * ``diagonal_solver``  -- DiagonalSolver._solve (solvers/diagonal.py:57-85)
"""

from __future__ import annotations

import numpy as np

from ._clib import clib, f64, i64, p
from .krylov import Monitor, _ret, dot_blas
from .ldu import ax
from .precond import dic_apply_full, dic_build, dilu_apply_full, dilu_build


def smooth_gs_sweeps(diag, lower, upper, owner, neighbour, b, x, sweeps):
    d, lo, up, bb = f64(diag), f64(lower), f64(upper), f64(b)
    o, q = i64(owner), i64(neighbour)
    xx = np.array(x, dtype=np.float64, copy=True)
    clib().orc_smooth_gs(d.shape[0], o.shape[0], p(o), p(q), p(d), p(lo), p(up), p(bb),
                         int(sweeps), p(xx))
    return xx


def smooth_solver(diag, lower, upper, owner, neighbour, b, x0, *, smoother="GaussSeidel",
                  n_sweeps=2, omega=1.0, tolerance=1e-6, rel_tol=0.01, max_iter=1000,
                  min_iter=0, dot=dot_blas):
    mon = Monitor(tolerance, rel_tol, min_iter, dot)
    mv = lambda v: ax(diag, lower, upper, owner, neighbour, v)
    x = np.array(x0, dtype=np.float64, copy=True)
    if smoother == "DIC":
        rD = dic_build(diag, upper, owner, neighbour)
        full = lambda r: dic_apply_full(rD, upper, owner, neighbour, r)
    elif smoother == "DILU":
        rD = dilu_build(diag, lower, upper, owner, neighbour)
        full = lambda r: dilu_apply_full(rD, lower, upper, owner, neighbour, r)
    if mon.update(b - mv(x), 0):
        return _ret(x, mon, True)
    for i in range(1, max_iter + 1):
        if smoother == "GaussSeidel":
            x = smooth_gs_sweeps(diag, lower, upper, owner, neighbour, b, x, n_sweeps)
        else:
            for _ in range(n_sweeps):
                x = x + omega * full(b - mv(x))
        if mon.update(b - mv(x), i):
            return _ret(x, mon, True)
    return _ret(x, mon, False)


def diagonal_solver(diag, b):
    d, bb = f64(diag), f64(b)
    x = bb / np.maximum(np.abs(d), 1e-30)
    r = bb - d * x
    return x, float(np.sqrt(dot_blas(r, r)))
