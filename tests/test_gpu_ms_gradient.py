"""GPU multislice gradient (forward sweep, loss seed, conj(H) adjoint sweep,
tp_plane_conj_dot accumulation) against the FP64 oracle, which is pinned by
central finite differences (tests/test_oracle_ms.py).

Bars: adjoint stage (FP64 loss seed fed to both sides) <= 1e-4 relative L2;
end to end <= 1e-3 with one transverse axis and <= 5e-3 with two (the
amplitude loss amplifies complex64 exit-wave rounding where |F psi_L| is
small); loss within 1e-5.
"""

import numpy as np
import pytest

from oracle import gradient as og
from oracle import multislice as oms
from tests.cases import rel_l2
from thick_ptycho_amd import gradient, workloads

pytestmark = pytest.mark.gpu
K = 2 * np.pi


@pytest.mark.parametrize("nx", [256, 1024])
def test_ms_gradient_one_axis(gpu, nx):
    print(f"case ms gradient nx={nx}", flush=True)
    rng = np.random.default_rng(nx)
    nz, P, dx, dz = 24, 6, 2.0, 4.0
    n_true = workloads.blob_object_2d(nx, nz, 5, rng)
    n0 = workloads.blob_object_2d(nx, nz, 5, np.random.default_rng(nx + 1))
    probes = workloads.gaussian_probes_1d(nx, np.linspace(nx * 0.2, nx * 0.8, P), nx / 10)
    I = np.abs(np.fft.fft(oms.solve_batched(n_true, probes, dx, K, dz, keep="exit"), axis=-1)) ** 2
    J, g = gradient.ms_amplitude_gradient(n0, probes, np.sqrt(I), dx, K, dz)
    ref, Jr = oms.ms_gradient(n0, probes, dx, K, dz, I)
    got = g.cpu().numpy()
    print(f"loss {J:.6e} vs {Jr:.6e}; end-to-end rel-L2 {rel_l2(got, ref):.2e}")
    assert abs(J - Jr) <= 1e-5 * Jr
    assert rel_l2(got, ref) < 1e-3
    seed = og.amplitude_loss_seed(oms.solve_batched(n0, probes, dx, K, dz, keep="exit"), I)
    _, g2 = gradient.ms_amplitude_gradient(n0, probes, np.sqrt(I), dx, K, dz, seed=seed)
    ref2, _ = oms.ms_gradient(n0, probes, dx, K, dz, I, seed=seed)
    print(f"adjoint stage rel-L2 {rel_l2(g2.cpu().numpy(), ref2):.2e}")
    assert rel_l2(g2.cpu().numpy(), ref2) < 1e-4


def test_ms_gradient_two_axes(gpu):
    rng = np.random.default_rng(9)
    nx, ny, nz, P = 64, 128, 10, 3
    dx, dy, dz = 2.0, 2.0, 4.0
    n_true = workloads.ellipsoid_object_3d(nx, ny, nz, 4, rng, delta=(0, 5e-3), beta=(0, 5e-5))
    n0 = workloads.ellipsoid_object_3d(nx, ny, nz, 4, np.random.default_rng(10), delta=(0, 2e-3), beta=(0, 2e-5))
    probes = workloads.gaussian_probes_2d(nx, ny, [(20, 40), (32, 64), (44, 90)], 12.0)
    ex = oms.solve_batched_2d(n_true, probes, dx, dy, K, dz, keep="exit")
    I = np.abs(np.fft.fft2(ex, axes=(-2, -1))) ** 2
    J, g = gradient.ms_amplitude_gradient_2d(n0, probes, np.sqrt(I), dx, dy, K, dz)
    ref, Jr = oms.ms_gradient_2d(n0, probes, dx, dy, K, dz, I)
    got = g.cpu().numpy()
    seed = og.amplitude_loss_seed_2d(oms.solve_batched_2d(n0, probes, dx, dy, K, dz, keep="exit").reshape(P, -1),
                                     I, nx, ny)
    _, g2 = gradient.ms_amplitude_gradient_2d(n0, probes, np.sqrt(I), dx, dy, K, dz, seed=seed)
    ref2, _ = oms.ms_gradient_2d(n0, probes, dx, dy, K, dz, I, seed=seed.reshape(P, nx, ny))
    print(f"loss {J:.6e} vs {Jr:.6e}; end-to-end rel-L2 {rel_l2(got, ref):.2e}; "
          f"adjoint stage rel-L2 {rel_l2(g2.cpu().numpy(), ref2):.2e}")
    assert abs(J - Jr) <= 1e-5 * Jr
    assert rel_l2(g2.cpu().numpy(), ref2) < 1e-4
    # 8192 detector bins per probe, many where |fft2 psi_L| is ~1e-6 of its peak:
    # the seed's phase there follows complex64 rounding of the exit waves
    assert rel_l2(got, ref) < 5e-3


def test_ms_autograd_backward(gpu):
    """MSExitWaves under torch.autograd against the FP64 multislice gradient."""
    import torch

    rng = np.random.default_rng(31)
    nx, nz, P, dx, dz = 256, 20, 4, 2.0, 4.0
    n_true = workloads.blob_object_2d(nx, nz, 5, rng)
    n0 = workloads.blob_object_2d(nx, nz, 5, np.random.default_rng(32))
    probes = workloads.gaussian_probes_1d(nx, np.linspace(60, 200, P), 25.0)
    I = np.abs(np.fft.fft(oms.solve_batched(n_true, probes, dx, K, dz, keep="exit"), axis=-1)) ** 2
    ref, _ = oms.ms_gradient(n0, probes, dx, K, dz, I)
    n = torch.tensor(n0, dtype=torch.complex128, device="cuda", requires_grad=True)
    ex = gradient.MSExitWaves.apply(n, torch.tensor(probes, device="cuda"), dx, K, dz)
    a = torch.tensor(np.sqrt(I), dtype=torch.float32, device="cuda")
    J = ((torch.fft.fft(ex, dim=-1).abs() - a) ** 2).sum()
    J.backward()
    err = rel_l2(n.grad.cpu().numpy(), ref)
    print(f"autograd multislice gradient rel-L2 {err:.2e}")
    assert err < 1e-3  # measured 1.3e-5
