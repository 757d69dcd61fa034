from .iterative_lu import PWEForwardModel, PWEIterativeLUSolver, PWEIterativeLUSolverCache
from .ms_solver import MSForwardModelSolver
from .operators import BoundaryType

__all__ = [
    "BoundaryType",
    "MSForwardModelSolver",
    "PWEForwardModel",
    "PWEIterativeLUSolver",
    "PWEIterativeLUSolverCache",
]
