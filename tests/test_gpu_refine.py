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


# ---------------------------------------------------------------- two transverse axes
def _problem_3d(nx, ny, nz, P, bc, seed):
    from thick_ptycho_amd import workloads

    rng = np.random.default_rng(seed)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n = workloads.ellipsoid_object_3d(nx, ny, nz, 4, rng, delta=(0, 5e-3), beta=(0, 5e-5))
    probes = rng.standard_normal((P, nx * ny)) + 1j * rng.standard_normal((P, nx * ny))
    return g, n, probes


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
def test_band_residual_matches_fp64(gpu, bc):
    """tp_band_residual against the FP64 residual on the assembled two-axis
    operator (pwe.step_operators), for one and two plane sets; the planes of an
    exact FP64 solution held as hi + lo give a residual at FP64 rounding, which a
    wrong coupling, sign or column would not."""
    import scipy.sparse as sps

    from thick_ptycho_amd.forward_model import band_engine as be

    nx, ny, nz, P = 24, 20, 7, 3
    g, n, probes = _problem_3d(nx, ny, nz, P, bc, 11)
    N, L = nx * ny, nz - 1
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, bc, n, "forward"), probes, "forward")  # (P, N, nz)
    hi = np.ascontiguousarray(u.transpose(0, 2, 1)).astype(np.complex64)  # (P, nz, N)
    lo = (u.transpose(0, 2, 1) - hi.astype(complex)).astype(np.complex64)
    A0, _ = pwe.step_operators(g, bc)
    C = pwe.object_contribution(g, n.reshape(N, nz))  # (N, L)

    def ref_res(v):
        out = np.empty((P, L, N), dtype=complex)
        for s in range(1, nz):
            As = A0 - sps.diags(C[:, s - 1])
            out[:, s - 1] = (2 * v[:, s - 1].T - As @ (v[:, s - 1] + v[:, s]).T).T
        return out

    A, B = be.step_operators_3d(nx, ny, K, 2.0, 2.0, 4.0, bc)
    Cd = torch.from_numpy(np.ascontiguousarray(C)).cuda()
    bf = be.BandFactorizer(A, B, "forward", L, 6, torch.device("cuda"))
    r = torch.empty((P, L, N), dtype=torch.complex64, device="cuda")
    d_hi, d_lo = torch.from_numpy(hi).cuda(), torch.from_numpy(lo).cuda()
    got = be.residual(bf, Cd, [d_hi], r).cpu().numpy()
    ref = ref_res(hi.astype(complex))
    print(f"{bc}: |r|/|u| {np.abs(ref).max() / np.abs(u).max():.2e}, rel-L2 {rel_l2(got, ref):.2e}")
    assert rel_l2(got, ref) < 1e-5
    # hi + lo of the FP64 solution: the residual is at FP64 rounding of |u|
    got2 = be.residual(bf, Cd, [d_hi, d_lo], r).cpu().numpy()
    ratio = np.abs(got2).max() / np.abs(u).max()
    print(f"{bc}: residual of the FP64 solution {ratio:.2e} of |u|")
    assert ratio < 1e-13
    # C by step (transposed storage) gives the same residual
    CT = torch.from_numpy(np.ascontiguousarray(C.T)).cuda()
    got3 = be.residual(bf, CT.transpose(0, 1), [d_hi], r).cpu().numpy()
    assert np.array_equal(got3, got)


def test_seed2d_refined_matches_fp64(gpu):
    """FP64 two-axis seed of an exit plane given as hi + lo, with amplitudes hi +
    lo, against the FP64 seed; far field over 8 decades with a data floor (the
    complex64 seed is off there), and a non-square plane."""
    from thick_ptycho_amd import _lib

    rng = np.random.default_rng(9)
    for nx, ny, P in [(64, 128, 3), (256, 256, 2), (32, 16, 5)]:
        kx = np.fft.fftfreq(nx)[:, None] * nx
        ky = np.fft.fftfreq(ny)[None, :] * ny
        U = np.exp(-(kx**2 + ky**2) / (2 * (nx / 10) ** 2))[None] * np.exp(2j * np.pi * rng.random((P, nx, ny)))
        U = U * np.maximum(1.0, 1e-8 / np.abs(U))
        ex = np.fft.ifft2(U, axes=(-2, -1)).reshape(P, nx * ny)
        a = np.abs(U) * (1 + 0.05 * rng.standard_normal((P, nx, ny))) + 1e-3
        ref = og.amplitude_loss_seed_2d(ex, a**2, nx, ny)
        J = [og.amplitude_loss_2d(ex[p:p + 1], a[p:p + 1] ** 2, nx, ny) for p in range(P)]
        hi = ex.astype(np.complex64)
        lo = (ex - hi.astype(complex)).astype(np.complex64)
        a_hi = a.astype(np.float32)
        a_lo = (a - a_hi.astype(float)).astype(np.float32)
        d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()  # noqa: E731
        seed = torch.empty((P, nx * ny), dtype=torch.complex64, device="cuda")
        loss = torch.zeros(P, dtype=torch.float32, device="cuda")
        lib = _lib.load()
        work = torch.empty(int(lib.tp_amplitude_seed2d_refined_work_bytes(nx, ny, P)) // 8 + 1, dtype=torch.float64,
                           device="cuda")
        dhi, dlo, dah, dal = d(hi), d(lo), d(a_hi), d(a_lo)
        _lib.check(lib.tp_amplitude_seed2d_refined(_lib.ptr(dhi), _lib.ptr(dlo), 0, nx * ny, _lib.ptr(dah),
                                                   _lib.ptr(dal), nx, ny, P, 1.0, _lib.ptr(seed), _lib.ptr(loss),
                                                   _lib.ptr(work), _lib.stream_handle()), "seed2d refined")
        err = rel_l2(seed.cpu().numpy(), ref)
        # the complex64 seed of the same plane, for the record
        s32 = torch.empty_like(seed)
        l32 = torch.zeros_like(loss)
        w32 = torch.empty((2, P, nx * ny), dtype=torch.complex64, device="cuda")
        e32 = float("nan")
        if nx >= 64 and ny >= 64:
            _lib.check(lib.tp_amplitude_seed2d(_lib.ptr(dhi), _lib.ptr(dah), nx, ny, P, 1.0, _lib.ptr(s32),
                                               _lib.ptr(l32), _lib.ptr(w32), _lib.stream_handle()), "seed2d")
            e32 = rel_l2(s32.cpu().numpy(), ref)
        print(f"{nx}x{ny}: refined seed rel-L2 {err:.2e} (complex64 seed {e32:.2e})")
        assert err < 1e-6
        np.testing.assert_allclose(loss.double().cpu().numpy(), J, rtol=1e-6)


def _recon_problem(nx, ny, nz, P, seed):
    from thick_ptycho_amd import workloads

    rng = np.random.default_rng(seed)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n_true = workloads.ellipsoid_object_3d(nx, ny, nz, 6, rng, delta=(0, 5e-3), beta=(0, 5e-5))
    side = int(np.sqrt(P))
    probes = workloads.gaussian_probes_2d(nx, ny, workloads.raster_centres(side, nx / (side + 1), nx / 2), nx / 6)
    probes = probes.reshape(P, -1)
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, "dirichlet", n_true, "forward"), probes, keep="exit")
    I = np.abs(np.fft.fft2(u.reshape(P, nx, ny), axes=(-2, -1))) ** 2
    return g, n_true, probes, I


@pytest.mark.parametrize("start", ["free space", "object"])
def test_refined_gradient_3d_matches_oracle(gpu, start):
    """End-to-end 3D amplitude gradient with two refinement steps against the FP64
    exact gradient at 1e-4 (the unrefined complex64 path is ~3e-4 off), from free
    space (the reconstruction's start) and from a perturbed object."""
    from thick_ptycho_amd import recon3d, workloads

    nx, ny, nz, P = 64, 64, 12, 9
    g, n_true, probes, I = _recon_problem(nx, ny, nz, P, 0)
    if start == "free space":
        n0 = np.ones_like(n_true)
    else:
        n0 = workloads.ellipsoid_object_3d(nx, ny, nz, 6, np.random.default_rng(1), delta=(0, 2e-3), beta=(0, 2e-5))
    ref, J = og.exact_gradient_3d(g, "dirichlet", n0, probes, I)
    errs = {}
    for refine in (0, 1, 2):
        rc = recon3d.Reconstruction3D(np.sqrt(I), probes, n0, K, 2.0, 2.0, 4.0, "dirichlet", batch=5, refine=refine)
        rc.gradient()
        got = rc.grad_object().cpu().numpy()
        loss = float(rc.loss.double().sum().item())
        errs[refine] = rel_l2(got, ref)
        print(f"{start}, refine {refine}: gradient rel-L2 {errs[refine]:.2e}, loss rel {abs(loss - J) / J:.2e}")
        if refine == 2:
            assert abs(loss - J) <= 1e-6 * J
    assert errs[2] < 1e-4
    assert errs[2] <= errs[0]
