"""Differentiable linear solves (reference ``src/pyfoam/differentiable``)."""

from .linear_solver import DifferentiableLinearSolve, DifferentiablePoissonSolve, poisson_solve

__all__ = ["DifferentiableLinearSolve", "DifferentiablePoissonSolve", "poisson_solve"]
