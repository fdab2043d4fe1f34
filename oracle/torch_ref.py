"""The reference's PCG-DIC op sequence in torch on the CPU (oracle; test and
benchmark infrastructure only: the CPU baseline that bench.py times beside the
device solve, never the product path).

Restates, op for op, with torch's CPU kernels and threads:
  * ``ax``  -- LduMatrix.Ax (core/ldu_matrix.py:172-217): y = diag * x, then
               + scatter_add(lower * gather(x, neighbour), owner), then
               + scatter_add(upper * gather(x, owner), neighbour), each
               scatter into a fresh zero tensor (core/backend.py:34-90);
  * ``pcg`` -- PCGSolver._solve (solvers/pcg.py:97-187) with the
               DICPreconditioner.apply z = r / d (preconditioners.py:135-151)
               and ResidualMonitor.update (residual.py:112-145): the norm
               ``torch.norm(r).item()`` and the p.Ap breakdown test
               ``float(pAp)`` synchronise once each per iteration, as there.
The factored diagonal comes from the C restatement of the sequential DIC
recurrence (oracle.precond.dic_build), as the setup line of the baseline.
"""

from __future__ import annotations

import torch


class LduCpu:
    """The reference's LDU storage on the CPU: float64 values, int64 indices."""

    def __init__(self, diag, lower, upper, owner, neighbour):
        t = lambda a, dt: torch.as_tensor(a).to(dtype=dt).contiguous()
        self.diag = t(diag, torch.float64)
        self.lower = t(lower, torch.float64)
        self.upper = t(upper, torch.float64)
        self.owner = t(owner, torch.int64)
        self.neighbour = t(neighbour, torch.int64)
        self.n = int(self.diag.shape[0])

    def ax(self, x: torch.Tensor) -> torch.Tensor:
        y = self.diag * x
        if self.owner.numel():
            x_own = torch.gather(x, 0, self.owner)
            x_nei = torch.gather(x, 0, self.neighbour)
            y = y + torch.zeros(self.n, dtype=x.dtype).scatter_add(0, self.owner,
                                                                   self.lower * x_nei)
            y = y + torch.zeros(self.n, dtype=x.dtype).scatter_add(0, self.neighbour,
                                                                   self.upper * x_own)
        return y


def pcg(A: LduCpu, b, x0, rD, *, tolerance=1e-6, rel_tol=0.01, max_iter=1000, min_iter=0):
    """(x, iterations = len(history), final residual, converged, history)."""
    b = torch.as_tensor(b, dtype=torch.float64)
    x = torch.as_tensor(x0, dtype=torch.float64).clone()
    d = torch.as_tensor(rD, dtype=torch.float64)
    hist: list[float] = []

    def update(r, it):
        res = float(torch.norm(r).item())
        hist.append(res)
        if it >= min_iter:
            if res < tolerance:
                return True
            if hist[0] > 0 and res / hist[0] < rel_tol:
                return True
        return False

    r = b - A.ax(x)
    if update(r, 0):
        return x, len(hist), hist[-1], True, hist
    z = r / d
    p = z.clone()
    rz = torch.dot(r, z)
    for i in range(1, max_iter + 1):
        Ap = A.ax(p)
        pAp = torch.dot(p, Ap)
        if abs(float(pAp)) < 1e-30:
            return x, len(hist), hist[-1], False, hist
        alpha = rz / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        if update(r, i):
            return x, len(hist), hist[-1], True, hist
        z = r / d
        rz_new = torch.dot(r, z)
        beta = rz_new / rz
        rz = rz_new
        p = z + beta * p
    return x, len(hist), hist[-1], False, hist
