"""GPU checks of the learned-preconditioned Krylov path (3D step).

* CNN kernels (forward, back-propagated gradient, Adam step) against the CPU
  oracle with the same bfloat16 rounding points; tolerance relative L2 <= 5e-3
  (FP32 vs FP64 accumulation can flip single BF16 roundings);
* BiCGSTAB (none / Jacobi / CNN) against the banded LU solve of the same
  slice (relative L2 <= 2e-6 at tol 1e-7);
* the Krylov sweep against the FP64 oracle (<= 1e-4), the banded-LU fallback
  (forced with max_iter = 0), and the ablation: iterations and time with and
  without the trained network.
"""

import time

import numpy as np
import pytest
import torch

from oracle import learned as ol
from oracle import pwe
from tests.cases import rel_l2
from thick_ptycho_amd import learned as tl
from thick_ptycho_amd import workloads
from thick_ptycho_amd.forward_model import band_engine as be

pytestmark = pytest.mark.gpu
K = 2 * np.pi


def _planes(P, N, rng):
    return (rng.standard_normal((P, N)) + 1j * rng.standard_normal((P, N))).astype(np.complex64)


@pytest.mark.parametrize("nx,ny", [(32, 24), (40, 20)])
def test_cnn_forward_matches_oracle(gpu, nx, ny, rng):
    params = ol.identity_params(rng, 0.05)
    cnn = tl.CNNPreconditioner(params=params)
    v = _planes(3, nx * ny, rng)
    got = cnn.apply(torch.tensor(v, device="cuda"), nx, ny).cpu().numpy()
    ref, _, _ = ol.forward(params.astype(np.complex64).astype(complex), v, nx, ny, bf16_round=True)
    exact, _, _ = ol.forward(params.astype(np.complex64).astype(complex), v, nx, ny, bf16_round=False)
    print(f"cnn forward {nx}x{ny}: vs rounded oracle {rel_l2(got, ref):.2e}, vs exact {rel_l2(got, exact):.2e}")
    assert rel_l2(got, ref) < 5e-3
    assert rel_l2(got, exact) < 3e-2


def test_cnn_gradient_matches_oracle(gpu, rng):
    nx, ny = 32, 16
    params = ol.identity_params(rng, 0.05)
    cnn = tl.CNNPreconditioner(params=params)
    r = _planes(4, nx * ny, rng)
    e = _planes(4, nx * ny, rng)
    loss, grad = cnn.gradient(torch.tensor(r, device="cuda"), torch.tensor(e, device="cuda"), nx, ny)
    rl, rg = ol.loss_and_grad(params.astype(np.complex64).astype(complex), r, e, nx, ny, bf16_round=True)
    g = grad.cpu().numpy()
    print(f"cnn loss {loss:.6e} vs {rl:.6e}; gradient rel-L2 {rel_l2(g, rg):.2e}")
    assert abs(loss - rl) <= 1e-3 * rl
    assert rel_l2(g, rg) < 5e-3


def test_cnn_adam_step_matches_formula(gpu, rng):
    nx, ny = 16, 16
    params = ol.identity_params(rng, 0.05)
    cnn = tl.CNNPreconditioner(params=params)
    r = torch.tensor(_planes(2, nx * ny, rng), device="cuda")
    e = torch.tensor(_planes(2, nx * ny, rng), device="cuda")
    _, g = cnn.gradient(r, e, nx, ny)
    g = g.cpu().numpy().astype(np.complex128)
    cnn.train_step(r, e, nx, ny, lr=1e-3)
    new = cnn.state().astype(np.complex128)
    # first Adam step (bias-corrected moments m = g, v = g^2): w -= lr * g / (|g| + eps) per real component
    upd = g.real / (np.abs(g.real) + 1e-8) + 1j * g.imag / (np.abs(g.imag) + 1e-8)
    exp = params.astype(np.complex64).astype(complex) - 1e-3 * upd
    assert np.allclose(new, exp, atol=2e-6)


def _slice_system(nx, ny, seed, bc="impedance"):
    rng = np.random.default_rng(seed)
    n = workloads.ellipsoid_object_3d(nx, ny, 4, 6, rng)
    return n


@pytest.mark.parametrize("precond", ["none", "jacobi", "cnn"])
def test_bicgstab_matches_band_lu(gpu, precond, rng):
    nx = ny = 48
    n = _slice_system(nx, ny, 5)
    cnn = tl.CNNPreconditioner(seed=1) if precond == "cnn" else None
    ks = tl.KrylovStepper(n, K, 2.0, 2.0, 4.0, "impedance", precond=precond, cnn=cnn, tol=1e-7, max_iter=60)
    b = torch.tensor(_planes(6, nx * ny, rng), device="cuda")
    x = b.clone()
    done = ks.solve_slice(2, b, x).cpu().numpy()
    assert (done > 0).all(), done
    fac = ks._band_slice(2)
    ex = torch.empty_like(b)
    be.sweep([be.BandScanSpec(fac, b, exit=ex)])
    ref = (0.5 * (ex + b)).cpu().numpy()
    err = rel_l2(x.cpu().numpy(), ref)
    print(f"bicgstab {precond}: iterations {done.tolist()}, rel-L2 vs band LU {err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("precond", ["none", "jacobi"])
def test_krylov_sweep_matches_oracle(gpu, precond, rng):
    nx, ny, nz = 40, 32, 12
    n = 1 - (rng.uniform(0, 1e-2, (nx, ny, nz)) + 1j * rng.uniform(0, 1e-3, (nx, ny, nz)))
    probes = _planes(3, nx * ny, rng)
    ks = tl.KrylovStepper(n, K, 2.0, 2.0, 4.0, "dirichlet", precond=precond, tol=1e-7)
    full, stats = ks.sweep(torch.tensor(probes, device="cuda"), keep="all")
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, "dirichlet", n, "forward"), probes.astype(complex))
    got = full.transpose(1, 2).cpu().numpy()
    print(f"krylov sweep {precond}: mean iterations {np.mean(stats.iterations):.2f}, fallbacks {stats.fallback_probes}")
    assert stats.fallback_probes == 0
    assert rel_l2(got, ref) < 1e-4


def test_fallback_to_band_lu(gpu, rng):
    """max_iter = 0: no probe converges, every slice falls back to the banded LU solve."""
    nx, ny, nz = 24, 20, 6
    n = 1 - (rng.uniform(0, 1e-2, (nx, ny, nz)) + 1j * rng.uniform(0, 1e-3, (nx, ny, nz)))
    probes = _planes(2, nx * ny, rng)
    ks = tl.KrylovStepper(n, K, 2.0, 2.0, 4.0, "neumann", precond="none")
    ex, stats = ks.sweep(torch.tensor(probes, device="cuda"), max_iter=0)
    assert stats.fallback_slices == nz - 1 and stats.fallback_probes == 2 * (nz - 1)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, "neumann", n, "forward"), probes.astype(complex), keep="exit")
    assert rel_l2(ex.cpu().numpy(), ref) < 1e-4


@pytest.mark.parametrize("dx", [2.0, 0.25])
def test_learned_preconditioner_ablation(gpu, dx):
    """Config-3 size (128 x 128, the 64 probes of the 8 x 8 raster, 32 slices of the config-3
    object) at the config's sampling (dx = 2) and at a finer one (dx = 0.25 at dz = 4, where
    the step's operator is worse conditioned and the banded LU needs its widest band).  The
    CNN is trained on 20480 banded-LU residual/correction pairs of two held-out objects
    (160 slices x 64 planes).  Then CN stepping by BiCGSTAB to 1e-6 with none / Jacobi / the
    CNN, and the banded LU:
      * iterations: the trained CNN divides the mean BiCGSTAB iterations per slice (1.53 ->
        1.00 at dx = 2, 20.6 -> 7.1 at dx = 0.25) and the cold-start iterations (4 -> 2,
        119 -> 12) by at least the asserted factors;
      * accuracy: every form against the FP64 oracle at 1e-4;
      * time (printed, not asserted: there is no gain to assert): the unpreconditioned
        iteration is the fastest Krylov form at both samplings (8.0 against 16.0 ms at dx = 2,
        65 against 125 ms at dx = 0.25, banded LU 26 / 92 ms), because one CNN application to
        64 planes of 128 x 128 (0.155 ms) costs 1.6 unpreconditioned iterations and BiCGSTAB
        applies it twice per iteration."""
    nx = ny = 128
    sc = workloads.config3(nz=33).scans[0]
    P = sc.probes.shape[0]
    probes = sc.probes.reshape(P, -1)
    u0 = torch.tensor(probes.astype(np.complex64), device="cuda")
    print(f"case dx={dx}", flush=True)
    objs = [workloads.ellipsoid_object_3d(nx, ny, 256, 32, np.random.default_rng(s)) for s in (31, 32)]
    pairs = tl.generate_pairs(objs, K, dx, dx, 4.0, sc.bc, per_slice=64, slices=160, seed=33)
    assert pairs.count == 20480
    cnn = tl.CNNPreconditioner(seed=0)
    l0 = cnn.gradient(pairs.r[:32], pairs.e[:32], nx, ny)[0]
    losses = cnn.train(pairs, steps=1500, batch=32, lr=2e-3, log_every=500)
    del pairs
    print(f"dx {dx}: training loss {l0:.3e} -> {losses[-1]:.3e}")
    assert losses[-1] < 0.05 * l0
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=sc.nz, dx=dx, dy=dx, dz=4.0, k=K)
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, sc.bc, sc.n, "forward"), probes, keep="exit")
    b = torch.tensor(_planes(64, nx * ny, np.random.default_rng(34)), device="cuda")
    res = {}
    for pc in ("none", "jacobi", "cnn"):
        ks = tl.KrylovStepper(sc.n, K, dx, dx, 4.0, sc.bc, precond=pc, cnn=cnn, tol=1e-6, max_iter=400)
        cold = []
        for s_ in (1, 16, 32):
            x = torch.zeros_like(b)
            d = ks.solve_slice(s_, b, x).cpu().numpy()
            cold.append(float(d.mean()) if (d > 0).all() else float("nan"))
        ks.sweep(u0)  # warm-up
        ex, st = ks.sweep(u0)
        err = rel_l2(ex.cpu().numpy(), ref)
        res[pc] = (float(np.mean(st.iterations)), st.seconds, float(np.mean(cold)))
        print(f"dx {dx}: {pc}: stepping {res[pc][0]:.2f} iterations per slice, {st.seconds * 1e3:.1f} ms for "
              f"{sc.nz - 1} slices, cold start {res[pc][2]:.2f} iterations, fallbacks {st.fallback_probes}, "
              f"rel-L2 vs FP64 {err:.2e}")
        assert st.fallback_probes == 0
        assert err < 1e-4, pc
    # the banded LU of the same step
    A, B = be.step_operators_3d(nx, ny, K, dx, dx, 4.0, sc.bc)
    C = tl._contribution(sc.n, K, 4.0, torch.device("cuda"))
    w = be.factor(C, A, B, "forward").w
    ex = torch.empty_like(u0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fac = be.factor(C, A, B, "forward", w=w)
    be.sweep([be.BandScanSpec(fac, u0, exit=ex)])
    torch.cuda.synchronize()
    t_band = time.perf_counter() - t0
    print(f"dx {dx}: banded LU (half-width {w}): factor + sweep {t_band * 1e3:.1f} ms, rel-L2 vs FP64 "
          f"{rel_l2(ex.cpu().numpy(), ref):.2e}")
    gain = {2.0: 0.75, 0.25: 0.4}[dx]
    assert res["cnn"][0] <= gain * res["none"][0]
    assert res["cnn"][2] <= gain * res["none"][2]


def test_training_pairs_from_krylov_match_band(gpu):
    """generate_pairs(solver='krylov') (BiCGSTAB to 1e-7, for steps beyond the band's reach)
    gives the same corrections A_s^{-1} r as the banded-LU pairs for the same r."""
    nx, ny = 48, 40
    obj = workloads.ellipsoid_object_3d(nx, ny, 12, 6, np.random.default_rng(2))
    a = tl.generate_pairs([obj], K, 2.0, 2.0, 4.0, "dirichlet", per_slice=4, slices=5, seed=9)
    b = tl.generate_pairs([obj], K, 2.0, 2.0, 4.0, "dirichlet", per_slice=4, slices=5, seed=9, solver="krylov",
                          tol=1e-7)
    assert torch.equal(a.r, b.r)
    err = rel_l2(b.e.cpu().numpy(), a.e.cpu().numpy())
    print(f"krylov pairs vs band pairs: rel-L2 {err:.2e}")
    assert err < 1e-5
