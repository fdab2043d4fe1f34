"""Implicit Laplacian assembly (reference ``src/pyfoam/discretisation``)."""

from . import fvm

__all__ = ["fvm"]
