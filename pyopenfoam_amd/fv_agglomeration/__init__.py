"""GAMG agglomeration (reference ``fv_agglomeration/``): pair agglomeration."""

from .pair_agglomeration import PairGamgAgglomeration

__all__ = ["PairGamgAgglomeration"]
