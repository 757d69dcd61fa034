"""Host hand-over of the drop-in ``solve()``: complex128 arrays in the reference's
(..., N, nz) layout (base/base_solver.py:180-188) from complex64 device planes.

The copy helper is layout logic only (the GPU runs no part of it on CPU
tensors), so it is checked here against numpy's transpose and widening, with
chunk sizes that split the leading axes unevenly."""

import numpy as np
import pytest
import torch

from thick_ptycho_amd.forward_model import base_solver


@pytest.mark.parametrize("chunk_planes", [1, 2, 5, 1000])
def test_host_complex128_matches_numpy(monkeypatch, chunk_planes):
    nz, n = 7, 5
    monkeypatch.setattr(base_solver, "HOST_CHUNK_BYTES", 16 * nz * n * chunk_planes)
    g = torch.Generator().manual_seed(3)
    x = torch.randn(2, 1, 3, nz, n, dtype=torch.complex64, generator=g)
    view = x.transpose(-1, -2)  # the solvers' (Proj, A, S, N, nz) view of (Proj, A, S, nz, N) planes
    h = base_solver.host_complex128(view)
    assert h.dtype == np.complex128 and h.shape == (2, 1, 3, n, nz) and h.flags["C_CONTIGUOUS"]
    np.testing.assert_array_equal(h, x.numpy().transpose(0, 1, 2, 4, 3).astype(np.complex128))


def test_host_complex128_empty_and_exact_widening():
    assert base_solver.host_complex128(torch.zeros((0, 4, 3), dtype=torch.complex64)).shape == (0, 4, 3)
    v = torch.tensor([[1 + 2j, 2.0**-140 - 3.5j]], dtype=torch.complex64)  # a float32 subnormal survives
    h = base_solver.host_complex128(v)
    assert h[0, 1] == complex(np.float32(2.0**-140), -3.5) and h[0, 0] == 1 + 2j
