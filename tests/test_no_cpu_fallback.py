"""The product has no CPU path: without a GPU every solver and driver refuses to
run (RuntimeError from the GPU check) instead of computing on the host, and a
missing library is reported as such.  CPU-only test (skipped where a GPU is
visible)."""

import numpy as np
import pytest
import torch

from thick_ptycho_amd import SimulationSpace, _lib
from thick_ptycho_amd import forward_model as fm
from thick_ptycho_amd import reconstruct

pytestmark = pytest.mark.skipif(torch.cuda.is_available(), reason="checks the behaviour without a GPU")


@pytest.mark.parametrize("cls", [fm.PWEIterativeLUSolver, fm.PWEFullLUSolver, fm.PWEFullPinTSolver,
                                 fm.PWEPetscFullPinTSolver, fm.MSForwardModelSolver])
def test_solvers_refuse_without_gpu(cls):
    ss = SimulationSpace(64, 8, 2.0, 4.0, num_probes=2)
    probes = np.ones((1, 2, 64), dtype=complex)
    with pytest.raises(RuntimeError, match="GPU"):
        cls(ss, probes)


def test_drivers_refuse_without_gpu():
    with pytest.raises((RuntimeError, AssertionError)):
        reconstruct.Reconstruction2D(np.ones((2, 64)), np.ones((2, 64), complex), 2 * np.pi, 2.0, 4.0)


def test_missing_library_is_reported(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libtp_hip.so"))
    with pytest.raises(_lib.LibraryMissing, match="no CPU fallback"):
        _lib.load()
