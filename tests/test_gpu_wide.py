"""GPU parity of the wide-grid tridiagonal kernels (several wavefronts per probe)
against the FP64 oracle.

Widths above 2048 rows run the wide layout; the reference steps such grids in
its own examples (N = 12001 in notebooks/2D_examples/1_generate_simulated_data.ipynb,
N = 7400 in 4.pwe_parallel_in_time_forward.ipynb).  The wide kernels are also
forced at small widths (one and several wavefronts) to cover every mode and
output against the oracle.  Tolerance: relative L2 <= 1e-4 unless stated.
"""

import numpy as np
import pytest
import torch

from oracle import gradient as og
from oracle import pwe
from oracle.gradient import exact_adjoint_states
from tests.cases import random_object, rel_l2
from thick_ptycho_amd import SimulationSpace
from thick_ptycho_amd import gradient as tg
from thick_ptycho_amd.forward_model import PWEIterativeLUSolver, operators, tri_engine

pytestmark = pytest.mark.gpu
TOL = 1e-4
K, DX, DZ = 2 * np.pi, 2.0, 4.0


def _grid(nx, nz):
    return pwe.Grid(dimension=2, nx=nx, nz=nz, dx=DX, dz=DZ, k=K)


def _probes(rng, P, nx):
    x = np.arange(nx)
    centres = np.linspace(nx / 6, 5 * nx / 6, P)
    pr = np.stack([np.exp(-((x - c) ** 2) / (nx / 10) ** 2) for c in centres]).astype(complex)
    return pr + 0.1 * (rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx)))


@pytest.mark.parametrize("nx,nz", [(2049, 16), (4096, 12), (7400, 10), (12001, 8), (16384, 6)])
@pytest.mark.parametrize("mode", ["forward", "adjoint"])
def test_reference_widths_through_drop_in(gpu, nx, nz, mode, rng):
    print(f"case wide drop-in nx={nx} nz={nz} mode={mode}", flush=True)
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = _probes(rng, 3, nx)
    ss = SimulationSpace(nx, nz, DX, DZ, wavelength=2 * np.pi / K, num_probes=3)
    s = PWEIterativeLUSolver(ss, probes[None], bc_type="impedance")
    assert s.waves > 1
    got = s.solve(n=n, mode=mode)[0, 0]
    ref = pwe.solve(_grid(nx, nz), "impedance", n, probes, mode)
    assert rel_l2(got, ref) < TOL


def test_width_limit_raises(gpu):
    with pytest.raises(ValueError, match="16384"):
        operators.layout_for(16385)


@pytest.mark.parametrize("nx,chunk,waves", [(300, 8, 1), (1500, 8, 3), (2048, 16, 2), (4500, 16, 5), (8000, 8, 16)])
@pytest.mark.parametrize("mode", ["forward", "adjoint", "reverse"])
def test_forced_wide_layouts_match_oracle(gpu, nx, chunk, waves, mode, rng):
    print(f"case forced wide nx={nx} chunk={chunk} waves={waves} mode={mode}", flush=True)
    nz = 14
    bc = "dirichlet" if mode == "reverse" else "impedance"
    g = _grid(nx, nz)
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = _probes(rng, 4, nx)
    A, B = operators.step_operators_1d(nx, K, DX, DZ, bc)
    C = torch.from_numpy(pwe.object_contribution(g, n)).to("cuda")
    fac = tri_engine.factor(C, A, B, mode, chunk=chunk, waves=waves)
    assert fac.wide and fac.waves == waves
    u0 = torch.from_numpy(probes).to("cuda", torch.complex64)
    out = torch.empty((4, nz, nx), dtype=torch.complex64, device="cuda")
    ex = torch.empty((4, nx), dtype=torch.complex64, device="cuda")
    ck = torch.empty((4, (nz - 1) // 4 + 1, nx), dtype=torch.complex64, device="cuda")
    tri_engine.sweep([tri_engine.TriScanSpec(fac, u0, out=out, exit=ex, ckpt=ck, ckpt_every=4,
                                             flip=(mode == "adjoint"))])
    ref = pwe.solve(g, bc, n, probes, mode)  # (P, N, nz)
    got = out.cpu().numpy().transpose(0, 2, 1)
    assert rel_l2(got, ref) < TOL
    last = 0 if mode == "adjoint" else nz - 1
    torch.testing.assert_close(ex, out[:, last], rtol=0, atol=0)
    stored = out.flip(1) if mode == "adjoint" else out
    torch.testing.assert_close(ck, stored[:, ::4][:, : ck.shape[1]], rtol=0, atol=0)


@pytest.mark.parametrize("nx,chunk,waves", [(700, 8, 2), (3000, 16, 3)])
def test_wide_exact_adjoint_with_rhs(gpu, nx, chunk, waves, rng):
    """exact_adjoint (per-row sigma) with a seed on the first step's right-hand side."""
    nz = 16
    g = _grid(nx, nz)
    n = random_object(nx, nz, rng, amp=2e-2)
    seed = rng.standard_normal((3, nx)) + 1j * rng.standard_normal((3, nx))
    lam = exact_adjoint_states(g, "impedance", n, seed)
    A, B = operators.step_operators_1d(nx, K, DX, DZ, "impedance")
    C = torch.from_numpy(pwe.object_contribution(g, n)).to("cuda")
    fac = tri_engine.factor(C, A, B, "exact_adjoint", chunk=chunk, waves=waves)
    L = nz - 1
    out = torch.empty((3, nz, nx), dtype=torch.complex64, device="cuda")
    z0 = torch.zeros((3, nx), dtype=torch.complex64, device="cuda")
    rhs = torch.from_numpy(seed).to("cuda", torch.complex64).reshape(3, 1, nx).contiguous()
    tri_engine.sweep([tri_engine.TriScanSpec(fac, z0, out=out, rhs=rhs, rhs_steps=1)])
    got = out.cpu().numpy()  # plane t holds lambda_{L-t+1}
    for t in range(1, L + 1):
        assert rel_l2(got[:, t], lam[:, :, L - t + 1]) < TOL


def test_wide_rhs_block_and_single_probe(gpu, rng):
    """rhs_block columns (forward j-1, adjoint -j) through the drop-in at a wide width."""
    nx, nz = 2600, 10
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = _probes(rng, 2, nx)
    rb = rng.standard_normal((nx, nz - 1)) + 1j * rng.standard_normal((nx, nz - 1))
    ss = SimulationSpace(nx, nz, DX, DZ, wavelength=2 * np.pi / K, num_probes=2)
    s = PWEIterativeLUSolver(ss, probes[None], bc_type="dirichlet")
    g = _grid(nx, nz)
    for mode in ("forward", "adjoint"):
        got = s.solve(n=n, mode=mode, rhs_block=rb)[0, 0]
        ref = pwe.solve(g, "dirichlet", n, probes, mode, rhs_block=rb)
        assert rel_l2(got, ref) < TOL
        one = s._solve_single_probe(scan_idx=1, probe=probes[1], n=n, mode=mode, rhs_block=rb)
        assert rel_l2(one, ref[1]) < TOL


@pytest.mark.parametrize("nx", [1536, 2048, 3000])
def test_gradient_at_wide_widths(gpu, nx):
    """Exact gradient of the exit-wave loss at chunk-32 widths (one probe per
    workgroup with gradient accumulation) and on the wide layout (gradient
    accumulated with global atomics), against the FP64 oracle."""
    print(f"case wide gradient nx={nx}", flush=True)
    rng = np.random.default_rng(nx)
    nz, P = 20, 3
    g = _grid(nx, nz)
    n0 = random_object(nx, nz, rng, amp=5e-3)
    probes = _probes(rng, P, nx)
    d = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    ref, J = og.exact_gradient(g, "dirichlet", n0, probes, None, exit_data=d)
    loss, grad = tg.amplitude_gradient(n0, probes, None, K, DX, DZ, "dirichlet", checkpoint=8, exit_data=d)
    assert abs(loss - J) <= 1e-4 * J
    assert rel_l2(grad.cpu().numpy(), ref) < TOL
