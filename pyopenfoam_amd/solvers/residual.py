"""Residual monitor and convergence record (reference ``solvers/residual.py``).

The convergence test itself runs on the device inside every libpfm solver
(the same rule, evaluated by the last workgroup of each reducing kernel);
this host object carries the parameters in, receives the residual history
out, and exposes the reference's ``ConvergenceInfo`` (``:27-55``) and
``ResidualMonitor`` (``:58-168``) interfaces.  ``update`` is kept for host
code that iterates itself: ``||r||_2`` of the given residual, then
``iter >= min_iter and (res < tolerance or (r0 > 0 and res/r0 < rel_tol))``.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass, field

import torch

__all__ = ["ResidualMonitor", "ConvergenceInfo"]

logger = logging.getLogger(__name__)


@dataclass
class ConvergenceInfo:
    converged: bool
    iterations: int
    final_residual: float
    initial_residual: float
    residual_ratio: float
    tolerance: float
    residual_history: list[float] = field(default_factory=list)


class ResidualMonitor:
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, min_iter: int = 0,
                 verbose: bool = False) -> None:
        self._tolerance = tolerance
        self._rel_tol = rel_tol
        self._min_iter = min_iter
        self._verbose = verbose
        self._history: list[float] = []
        self._initial_residual: float | None = None

    @property
    def tolerance(self) -> float:
        return self._tolerance

    @property
    def rel_tol(self) -> float:
        return self._rel_tol

    @property
    def min_iter(self) -> int:
        return self._min_iter

    @property
    def history(self) -> list[float]:
        return self._history

    @property
    def initial_residual(self) -> float | None:
        return self._initial_residual

    def reset(self) -> None:
        self._history.clear()
        self._initial_residual = None

    def _record(self, res_norm: float, iteration: int) -> bool:
        self._history.append(res_norm)
        if self._initial_residual is None:
            self._initial_residual = res_norm
        if self._verbose:
            r0 = self._initial_residual
            ratio = res_norm / r0 if r0 > 0 else 0.0
            logger.info("  Iteration %d: residual = %.6e (ratio = %.6e)", iteration, res_norm, ratio)
        if iteration >= self._min_iter:
            if res_norm < self._tolerance:
                return True
            if self._initial_residual > 0:
                if res_norm / self._initial_residual < self._rel_tol:
                    return True
        return False

    def update(self, residual: torch.Tensor, iteration: int) -> bool:
        return self._record(float(torch.linalg.vector_norm(residual).item()), iteration)

    def load_history(self, history: list[float]) -> None:
        """Adopt a residual history produced on the device."""
        self.reset()
        for k, h in enumerate(history):
            self._history.append(float(h))
            if k == 0:
                self._initial_residual = float(h)
        if self._verbose:
            for k, h in enumerate(self._history):
                logger.info("  Residual %d: %.6e", k, h)

    def build_info(self, converged: bool) -> ConvergenceInfo:
        initial = self._initial_residual or 0.0
        final = self._history[-1] if self._history else 0.0
        ratio = final / initial if initial > 0 else 0.0
        return ConvergenceInfo(converged=converged, iterations=len(self._history),
                               final_residual=final, initial_residual=initial,
                               residual_ratio=ratio, tolerance=self._tolerance,
                               residual_history=list(self._history))
