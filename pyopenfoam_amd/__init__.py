"""pyopenfoam_amd: the pressure-equation path of pyOpenFOAM on AMD MI355X.

LDU assembly (``fvm.laplacian``), the LDU SpMV, PCG(DIC) / PBiCGStab(DILU) /
GAMG solvers and the adjoint linear solve, computed by hand-written HIP
kernels for gfx950 in ``libpfm.so`` (C ABI in ``include/pfm.h``) behind the
reference's Python API.
"""

__version__ = "0.1.0"

from .core import LduMatrix  # noqa: E402
from .solvers import (GAMGSolver, PBiCGSTABSolver, PCGSolver, create_solver,  # noqa: E402
                      solver_from_dict)

__all__ = ["LduMatrix", "PCGSolver", "PBiCGSTABSolver", "GAMGSolver", "create_solver",
           "solver_from_dict"]
