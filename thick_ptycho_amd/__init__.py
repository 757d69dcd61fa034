"""thick_ptycho_amd: MI355X-native kernels for thick-sample ptychography forward models.

Drop-in solver classes (same constructors and methods as bryceshirley/thick_ptycho):
``PWEIterativeLUSolver`` (Crank-Nicolson slice stepping with per-slice LU,
forward / adjoint / reverse), ``PWEFullLUSolver`` (the whole space-z system,
by block substitution on the same kernels), ``PWEFullPinTSolver`` / ``PWEPetscFullPinTSolver``
(the whole space-z system by alpha-circulant preconditioned GMRES) and
``MSForwardModelSolver`` (angular-spectrum multislice), all running
hand-written HIP kernels from ``libtp_hip.so``.  Reconstruction drivers live in
``thick_ptycho_amd.reconstruct`` (NLCG, 3PIE, data files) and
``thick_ptycho_amd.recon3d`` (two transverse axes).
"""

from .forward_model import (BoundaryType, MSForwardModelSolver, PWEFullLUSolver, PWEFullPinTSolver,
                            PWEIterativeLUSolver, PWEPetscFullPinTSolver)
from .simulation import SimulationSpace

__all__ = ["BoundaryType", "MSForwardModelSolver", "PWEFullLUSolver", "PWEFullPinTSolver", "PWEIterativeLUSolver",
           "PWEPetscFullPinTSolver", "SimulationSpace"]
