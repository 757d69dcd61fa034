import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """GPU gate: tests marked gpu need a ROCm GPU and the built library; without a GPU they skip."""
    if not gpu_available():
        pytest.skip("no GPU in this environment")
    from thick_ptycho_amd import _lib

    _lib.load()
    return True


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
