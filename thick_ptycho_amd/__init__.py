"""thick_ptycho_amd: MI355X-native kernels for thick-sample ptychography forward models.

Drop-in solver classes (same constructors and methods as bryceshirley/thick_ptycho):
``PWEIterativeLUSolver`` (Crank-Nicolson slice stepping with per-slice LU,
forward / adjoint / reverse) and ``MSForwardModelSolver`` (angular-spectrum
multislice), both running hand-written HIP kernels from ``libtp_hip.so``.
"""

from .forward_model import BoundaryType, MSForwardModelSolver, PWEIterativeLUSolver
from .simulation import SimulationSpace

__all__ = ["BoundaryType", "MSForwardModelSolver", "PWEIterativeLUSolver", "SimulationSpace"]
