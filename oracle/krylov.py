"""Krylov solvers and the residual monitor (oracle; test infrastructure only).

Restates, in numpy float64 with the reference's op order:
  * ``Monitor``   -- ResidualMonitor.update/build_info (solvers/residual.py:112-168)
  * ``pcg``       -- PCGSolver._solve (solvers/pcg.py:97-187)
  * ``pbicgstab`` -- PBiCGSTABSolver._solve (solvers/pbicgstab.py:103-239)
  * ``pbicg``     -- PBiCGSolver._solve (solvers/pbicg.py:107-220); its
                     ``matvec_t`` is the transposed product (oracle.ldu.ax_t);
                     like the reference it updates r~ with A^T p (one search
                     direction), which is BiCG only for symmetric A
``matvec(x)`` is the LDU product (oracle.ldu.ax) and ``precond(r)`` the
diagonal apply r / rD, or None for no preconditioner.  ``dot`` selects the
summation order of the dot products and norms (the reference leaves it to
torch): BLAS (default), exact (correctly rounded) or numpy pairwise; running
all three measures how sensitive a solve is to that order.  The return value
mirrors LinearSolverBase.__call__ (linear_solver.py:88-126) plus the
ConvergenceInfo fields: (x, iterations, final_residual, info).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import math

import numpy as np


@dataclass
class Info:
    converged: bool
    iterations: int
    final_residual: float
    initial_residual: float
    residual_history: list = field(default_factory=list)


def dot_blas(a, b):
    return np.dot(a, b)


def dot_exact(a, b):
    """Correctly rounded dot product (exact sum of the rounded products)."""
    return math.fsum((a * b).tolist())


def dot_pairwise(a, b):
    """numpy's pairwise summation of the rounded products."""
    return np.add.reduce(a * b)


DOTS = {"blas": dot_blas, "exact": dot_exact, "pairwise": dot_pairwise}


def dot_device(n, *, layout="strided", max_red_blocks=2048, block=256, slice_rows=64, xcds=8,
               tile_slices=None):
    """The summation order of the device solvers' fused dot products, so that
    an oracle solve can be compared with a device solve bit for bit.  This is
    one more summation order (the reference leaves the order to torch), set
    by the device kernels' structure:

    * G = min(ceil(n_slices / waves), max_red_blocks) workgroups of 4 waves;
      physical workgroup b works as logical block lb = (b % 8) * (G // 8) +
      b // 8 (b < 8 * (G // 8), else lb = b);
    * wave w of logical block lb walks slices lb*4 + w + k*4G, k = 0, 1, ...
      (``layout="strided"``, the Krylov solvers), or, with ``"contiguous"``
      (the GAMG-preconditioned PCG), slices lb*per + w + 4k below
      min((lb+1)*per, n_slices), per = ceil(n_slices / G); lane l adds the
      product of row 64 s + l (rows < n) to its running sum from 0;
    * each wave is reduced by the butterfly v += v[l ^ off], off = 32 .. 1,
      lane 0's value taken; the 4 waves are added in order (partial of b);
    * the G partials: thread t sums part[t + 256 k], k = 0..7, from 0, then
      the same wave butterfly and in-order wave sum over 256 threads.

    ``tile_slices=T`` gives the order of the batched solver instead (one
    system of a batch): tiles of T consecutive slices in their natural order,
    ceil(n_slices / T) of them (no XCD remap), each tile's four waves walking
    slices tile*T + w + 4k as ``"contiguous"`` does, then the same sum of the
    tile partials.
    """
    waves = block // slice_rows
    nsl = -(-n // slice_rows)
    if tile_slices is not None:
        G = max(1, -(-nsl // tile_slices))
        b = np.arange(G)
        lb = b
        layout = "tiles"
    else:
        G = max(1, min(-(-nsl // waves), max_red_blocks))
        per = G // xcds
        b = np.arange(G)
        lb = np.where(b < per * xcds, (b % xcds) * per + b // xcds, b)
    if layout == "tiles":
        per = tile_slices
        steps = max(1, -(-per // waves))
        s = (lb[:, None, None] * per + np.arange(waves)[None, :, None]
             + np.arange(steps)[None, None, :] * waves)
        end = np.minimum((lb[:, None, None] + 1) * per, nsl) + 0 * s
    elif layout == "strided":
        steps = max(1, -(-nsl // (G * waves)))
        s = (lb[:, None, None] * waves + np.arange(waves)[None, :, None]
             + np.arange(steps)[None, None, :] * (G * waves))    # (G, waves, steps)
        end = np.full(s.shape, nsl)
    elif layout == "contiguous":
        per = -(-nsl // G)
        steps = max(1, -(-per // waves))
        s = (lb[:, None, None] * per + np.arange(waves)[None, :, None]
             + np.arange(steps)[None, None, :] * waves)
        end = np.minimum((lb[:, None, None] + 1) * per, nsl) + 0 * s
    else:
        raise ValueError(layout)
    rows = s[:, :, None, :] * slice_rows + np.arange(slice_rows)[None, None, :, None]
    valid = (s[:, :, None, :] < end[:, :, None, :]) & (rows < n)
    idx = np.where(valid, rows, n)                                 # n -> a 0.0 pad
    lanes = np.arange(slice_rows)
    kthreads = max(max_red_blocks // block, -(-G // block))  # tiles: any count
    tl = np.arange(block)

    def tree(v):  # (..., 64) -> (...,): the wave butterfly, lane 0
        for off in (32, 16, 8, 4, 2, 1):
            v = v + v[..., lanes ^ off]
        return v[..., 0]

    def dot(a, c):
        pr = np.append(np.asarray(a, dtype=np.float64) * np.asarray(c, dtype=np.float64), 0.0)
        vals = pr[idx]                                             # (G, waves, 64, steps)
        acc = np.zeros(vals.shape[:3])
        for k in range(steps):
            acc = acc + vals[..., k]
        wv = tree(acc)                                             # (G, waves)
        part = wv[:, 0]
        for w in range(1, waves):
            part = part + wv[:, w]
        pp = np.zeros(block * kthreads)
        pp[:G] = part
        t = np.zeros(block)
        for k in range(kthreads):
            t = t + pp[tl + k * block]
        wt = tree(t.reshape(waves, slice_rows))
        out = wt[0]
        for w in range(1, waves):
            out = out + wt[w]
        return float(out)

    return dot


class Monitor:
    def __init__(self, tolerance=1e-6, rel_tol=0.01, min_iter=0, dot=dot_blas):
        self.tol, self.rel_tol, self.min_iter = tolerance, rel_tol, min_iter
        self.history: list[float] = []
        self.r0 = None
        self.dot = dot

    def update(self, r: np.ndarray, it: int) -> bool:
        res = float(np.sqrt(self.dot(r, r)))
        self.history.append(res)
        if self.r0 is None:
            self.r0 = res
        if it >= self.min_iter:
            if res < self.tol:
                return True
            if self.r0 > 0:
                if res / self.r0 < self.rel_tol:
                    return True
        return False

    def info(self, converged: bool) -> Info:
        r0 = self.r0 or 0.0
        final = self.history[-1] if self.history else 0.0
        return Info(converged, len(self.history), final, r0, list(self.history))


def pcg(matvec, precond, b, x0, *, tolerance=1e-6, rel_tol=0.01, max_iter=1000, min_iter=0,
        dot=dot_blas):
    mon = Monitor(tolerance, rel_tol, min_iter, dot)
    x = np.array(x0, dtype=np.float64, copy=True)
    r = b - matvec(x)
    if mon.update(r, 0):
        return _ret(x, mon, True)
    z = precond(r) if precond is not None else r.copy()
    p = z.copy()
    rz = dot(r, z)
    for i in range(1, max_iter + 1):
        ap = matvec(p)
        pap = dot(p, ap)
        if abs(float(pap)) < 1e-30:
            return _ret(x, mon, False)
        alpha = rz / pap
        x = x + alpha * p
        r = r - alpha * ap
        if mon.update(r, i):
            return _ret(x, mon, True)
        z = precond(r) if precond is not None else r.copy()
        rz_new = dot(r, z)
        beta = rz_new / rz
        rz = rz_new
        p = z + beta * p
    return _ret(x, mon, False)


def pbicg(matvec, matvec_t, precond, b, x0, *, tolerance=1e-6, rel_tol=0.01, max_iter=1000,
          min_iter=0, dot=dot_blas):
    mon = Monitor(tolerance, rel_tol, min_iter, dot)
    x = np.array(x0, dtype=np.float64, copy=True)
    r = b - matvec(x)
    if mon.update(r, 0):
        return _ret(x, mon, True)
    rt = r.copy()
    ap = (lambda v: precond(v)) if precond is not None else (lambda v: v.copy())
    z = ap(r)
    zt = ap(rt)
    p = z.copy()
    rho = dot(zt, r)
    for i in range(1, max_iter + 1):
        if abs(float(rho)) < 1e-30:
            return _ret(x, mon, False)
        v = matvec(p)
        ztv = dot(zt, v)
        if abs(float(ztv)) < 1e-30:
            return _ret(x, mon, False)
        alpha = rho / ztv
        x = x + alpha * p
        r = r - alpha * v
        if mon.update(r, i):
            return _ret(x, mon, True)
        z = ap(r)
        rt = rt - alpha * matvec_t(p)
        zt = ap(rt)
        rho_new = dot(zt, r)
        beta = rho_new / rho
        rho = rho_new
        p = z + beta * p
    return _ret(x, mon, False)


def pbicgstab(matvec, precond, b, x0, *, tolerance=1e-6, rel_tol=0.01, max_iter=1000, min_iter=0,
              dot=dot_blas):
    mon = Monitor(tolerance, rel_tol, min_iter, dot)
    x = np.array(x0, dtype=np.float64, copy=True)
    r = b - matvec(x)
    if mon.update(r, 0):
        return _ret(x, mon, True)
    rt = r.copy()
    rho = dot(rt, r)
    p = r.copy()
    v = np.zeros_like(r)
    for i in range(1, max_iter + 1):
        if abs(float(rho)) < 1e-30:
            return _ret(x, mon, False)
        y = precond(p) if precond is not None else p.copy()
        v = matvec(y)
        rtv = dot(rt, v)
        if abs(float(rtv)) < 1e-30:
            return _ret(x, mon, False)
        alpha = rho / rtv
        s = r - alpha * v
        if mon.update(s, i):
            x = x + alpha * y
            return _ret(x, mon, True)
        z = precond(s) if precond is not None else s.copy()
        t = matvec(z)
        tt = dot(t, t)
        if abs(float(tt)) < 1e-30:
            x = x + alpha * y
            r = s.copy()
            if mon.update(r, i):
                return _ret(x, mon, True)
            rho_new = dot(rt, r)
            beta = (rho_new / rho) * alpha
            rho = rho_new
            p = r + beta * (p - v)
            continue
        omega = dot(t, z) / tt
        x = x + alpha * y + omega * z
        r = s - omega * t
        if mon.update(r, i):
            return _ret(x, mon, True)
        rho_new = dot(rt, r)
        if abs(float(omega)) < 1e-30:
            return _ret(x, mon, False)
        beta = (rho_new / rho) * (alpha / omega)
        rho = rho_new
        p = r + beta * (p - omega * v)
    return _ret(x, mon, False)


def _ret(x, mon: Monitor, converged: bool):
    info = mon.info(converged)
    return x, info.iterations, info.final_residual, info
