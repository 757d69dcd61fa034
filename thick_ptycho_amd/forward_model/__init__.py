from .full_lu import PWEFullLUSolver
from .iterative_lu import PWEForwardModel, PWEIterativeLUSolver, PWEIterativeLUSolverCache
from .ms_solver import MSForwardModelSolver
from .operators import BoundaryType
from .pint import PWEFullPinTSolver, PWEPetscFullPinTSolver

__all__ = [
    "BoundaryType",
    "MSForwardModelSolver",
    "PWEForwardModel",
    "PWEIterativeLUSolver",
    "PWEIterativeLUSolverCache",
    "PWEFullLUSolver",
    "PWEFullPinTSolver",
    "PWEPetscFullPinTSolver",
]
