"""DIC / DILU preconditioners (oracle; test infrastructure only).

Restates ``solvers/preconditioners.py``:
  * ``dic_build``  -- DICPreconditioner._build (:88-133)
  * ``dilu_build`` -- DILUPreconditioner._build (:228-272)
  * the diagonal apply (:135-151, :274-288) is r / rD, which callers write inline
  * ``dic_apply_full`` / ``dilu_apply_full`` -- triangular sweeps (:153-201, :290-337)
The ``*_py`` variants are literal Python loops for small systems; the default
variants call the C restatement in ``c/oracle.c`` (same loops, same rounding).
"""

from __future__ import annotations

import numpy as np

from ._clib import clib, f64, i64, p


def dic_build_py(diag, upper, owner, neighbour) -> np.ndarray:
    d = np.array(diag, dtype=np.float64, copy=True)
    n = d.shape[0]
    if len(owner) == 0:
        return np.maximum(np.abs(d), 1e-30)
    lower_adj = [[] for _ in range(n)]
    for f in range(len(owner)):
        lower_adj[int(neighbour[f])].append((int(owner[f]), float(upper[f])))
    for i in range(n):
        for j, a in lower_adj[i]:
            dj = float(d[j])
            if abs(dj) > 1e-30:
                d[i] = d[i] - (a * a) / dj
    a = np.abs(d)
    return np.where(a < 1e-30, 1e-30, a)


def dilu_build_py(diag, lower, upper, owner, neighbour) -> np.ndarray:
    d = np.array(diag, dtype=np.float64, copy=True)
    for f in range(len(owner)):
        pp, nn = int(owner[f]), int(neighbour[f])
        lf, uf = float(lower[f]), float(upper[f])
        dp, dn = float(d[pp]), float(d[nn])
        if abs(dn) > 1e-30:
            d[pp] = d[pp] - (lf * uf) / dn
        if abs(dp) > 1e-30:
            d[nn] = d[nn] - (uf * lf) / dp
    a = np.abs(d)
    return np.where(a < 1e-30, 1e-30, a)


def dic_build(diag, upper, owner, neighbour) -> np.ndarray:
    d, u, o, q = f64(diag), f64(upper), i64(owner), i64(neighbour)
    out = np.empty_like(d)
    if clib().orc_dic_build(d.shape[0], o.shape[0], p(o), p(q), p(d), p(u), p(out)):
        raise MemoryError
    return out


def dilu_build(diag, lower, upper, owner, neighbour) -> np.ndarray:
    d, lo, u, o, q = f64(diag), f64(lower), f64(upper), i64(owner), i64(neighbour)
    out = np.empty_like(d)
    clib().orc_dilu_build(d.shape[0], o.shape[0], p(o), p(q), p(d), p(lo), p(u), p(out))
    return out


def dic_apply_full(rD, upper, owner, neighbour, r) -> np.ndarray:
    rd, u, o, q, rr = f64(rD), f64(upper), i64(owner), i64(neighbour), f64(r)
    z = np.empty_like(rr)
    clib().orc_dic_apply_full(rd.shape[0], o.shape[0], p(o), p(q), p(rd), p(u), p(rr), p(z))
    return z


def dilu_apply_full(rD, lower, upper, owner, neighbour, r) -> np.ndarray:
    rd, lo, u, o, q, rr = f64(rD), f64(lower), f64(upper), i64(owner), i64(neighbour), f64(r)
    z = np.empty_like(rr)
    clib().orc_dilu_apply_full(rd.shape[0], o.shape[0], p(o), p(q), p(rd), p(lo), p(u), p(rr),
                               p(z))
    return z
