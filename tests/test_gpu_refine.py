"""Refinement kernels against FP64 evaluations on the CPU.

* tp_tri_residual: r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s) of complex64 planes
  (and of a sum of two plane sets), against the same expression in FP64 on the
  CPU (inputs exact in FP64).  The residual of nearly consistent planes is
  ~1e-7 of |u|, so FP32 arithmetic would get it wrong by ~100 %.
* tp_amplitude_seed_refined: the seed of an exit plane given as hi + lo against
  the FP64 seed of the same plane, at a point where the complex64 seed is off.
"""

import numpy as np
import pytest
import torch

from oracle import gradient as og
from oracle import pwe
from tests.cases import random_object, rel_l2
from thick_ptycho_amd import gradient as tg
from thick_ptycho_amd.forward_model import operators, tri_engine

pytestmark = pytest.mark.gpu
K = 2 * np.pi


@pytest.mark.parametrize("bc", ["dirichlet", "impedance"])
def test_residual_matches_fp64(gpu, bc, rng):
    nx, nz, P = 200, 20, 3
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=K)
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, bc, n, "forward"), probes, "forward")  # (P, N, nz)
    u32 = np.ascontiguousarray(u.transpose(0, 2, 1)).astype(np.complex64)  # (P, nz, N): nearly consistent planes
    A0, B0 = pwe.step_operators(g, bc)
    C = pwe.object_contribution(g, n).reshape(nx, nz - 1)
    v = u32.astype(np.complex128)
    ref = np.empty((P, nz - 1, nx), dtype=complex)
    for s in range(1, nz):
        As = A0 - __import__("scipy").sparse.diags(C[:, s - 1])
        w = (v[:, s - 1] + v[:, s]).T
        ref[:, s - 1] = (2 * v[:, s - 1].T - As @ w).T
    A, B = operators.step_operators_1d(nx, K, 2.0, 4.0, bc)
    fac = tri_engine.factor(torch.from_numpy(C).cuda(), A, B, "forward")
    r = tri_engine.residual(fac, torch.from_numpy(u32).cuda()).cpu().numpy()
    scale = np.abs(u).max()
    print(f"{bc}: |r|/|u| = {np.abs(ref).max() / scale:.2e}, rel-L2 {rel_l2(r, ref):.2e}")
    assert rel_l2(r, ref) < 1e-5
    # the iterate as a sum of two plane sets
    half = (u32 / 2).astype(np.complex64)
    rest = (u32.astype(complex) - half.astype(complex)).astype(np.complex64)
    v2 = half.astype(complex) + rest.astype(complex)
    ref2 = np.empty_like(ref)
    for s in range(1, nz):
        As = A0 - __import__("scipy").sparse.diags(C[:, s - 1])
        ref2[:, s - 1] = (2 * v2[:, s - 1].T - As @ (v2[:, s - 1] + v2[:, s]).T).T
    r2 = tri_engine.residual(fac, [torch.from_numpy(half).cuda(), torch.from_numpy(rest).cuda()]).cpu().numpy()
    assert rel_l2(r2, ref2) < 1e-5
    # a view of the later steps takes the same columns
    view = tri_engine.step_slice(fac, 7, 5)
    rv = tri_engine.residual(view, torch.from_numpy(np.ascontiguousarray(u32[:, 7:13])).cuda()).cpu().numpy()
    assert rel_l2(rv, ref[:, 7:12]) < 1e-5


def test_seed_ff_matches_fp64(gpu):
    """An exit plane whose far field spans 8 decades while the data keep a floor
    of 1e-3 of the peak (the ill-conditioning of the amplitude loss): the seed
    of the complex64 plane is set by rounding in the dim bins, the FP64 seed of
    hi + lo matches the FP64 seed to complex64 rounding."""
    rng = np.random.default_rng(5)
    P, n = 4, 1024
    k = np.fft.fftfreq(n) * n
    U = np.exp(-(k[None, :] ** 2) / (2 * 60.0**2) + 2j * np.pi * rng.random((P, n)))
    U = U * np.maximum(1.0, 1e-8 / np.abs(U))  # at least 1e-8 of the peak
    ex = np.fft.ifft(U, axis=-1)
    I = (np.abs(U) * (1 + 0.05 * rng.standard_normal((P, n))) + 1e-3) ** 2
    ref = og.amplitude_loss_seed(ex, I)
    hi = ex.astype(np.complex64)
    lo = (ex - hi.astype(complex)).astype(np.complex64)
    a = torch.from_numpy(np.sqrt(I)).to("cuda", torch.float32)
    s32, _ = tg.amplitude_seed(torch.from_numpy(hi).cuda(), a, 1.0)
    a_hi, a_lo = tg.amplitude_parts(np.sqrt(I), torch.device("cuda"), I.shape)
    sff, loss = tg.amplitude_seed(torch.from_numpy(hi).cuda(), a_hi, 1.0, exit_lo=torch.from_numpy(lo).cuda(),
                                  sqrt_i_lo=a_lo)
    e32, eff = rel_l2(s32.cpu().numpy(), ref), rel_l2(sff.cpu().numpy(), ref)
    print(f"seed rel-L2: complex64 {e32:.2e}, refined {eff:.2e}")
    assert eff < 1e-6
    J = np.sum((np.abs(np.fft.fft(ex, axis=-1)) - np.sqrt(I)) ** 2, axis=-1)
    np.testing.assert_allclose(loss.double().cpu().numpy(), J, rtol=1e-5)
