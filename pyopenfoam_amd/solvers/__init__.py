"""Linear solvers of the pressure equation (reference ``src/pyfoam/solvers``)."""

from .linear_solver import LinearSolverBase, create_solver, solver_from_dict
from .pbicgstab import PBiCGSTABSolver
from .pcg import PCGSolver
from .gamg import GAMGSolver
from .preconditioners import DICPreconditioner, DILUPreconditioner, Preconditioner
from .residual import ConvergenceInfo, ResidualMonitor
from .smoothers import DICGSmoother, GaussSeidelSmoother, JacobiSmoother, Smoother
from .smooth_solver import SmoothSolver
from .diagonal import DiagonalSolver
from .pbicg import PBiCGSolver
from .pressure_equation import (adjust_phi, assemble_pressure_equation, correct_face_flux,
                                correct_velocity, solve_pressure_equation)

__all__ = ["LinearSolverBase", "create_solver", "solver_from_dict", "PCGSolver",
           "PBiCGSTABSolver", "GAMGSolver", "DICPreconditioner", "DILUPreconditioner",
           "Preconditioner", "ConvergenceInfo", "ResidualMonitor", "GaussSeidelSmoother",
           "JacobiSmoother", "DICGSmoother", "Smoother", "SmoothSolver", "DiagonalSolver", "PBiCGSolver",
           "assemble_pressure_equation", "solve_pressure_equation", "correct_velocity",
           "correct_face_flux", "adjust_phi"]
