"""LDU storage and device plumbing (reference ``src/pyfoam/core``)."""

from .device import CFD_DTYPE, INDEX_DTYPE, get_default_dtype, get_device
from .ldu_matrix import LduMatrix

__all__ = ["CFD_DTYPE", "INDEX_DTYPE", "get_default_dtype", "get_device", "LduMatrix"]
