"""Mesh tools (reference ``tools/``): RCM renumbering."""

from .renumber_mesh import RenumberResult, renumber_mesh

__all__ = ["RenumberResult", "renumber_mesh"]
