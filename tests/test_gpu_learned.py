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


def test_learned_preconditioner_ablation(gpu, rng):
    """Train the CNN on banded-LU pairs (objects from seeds 31-32, config-3 generator at
    64 x 64) and compare BiCGSTAB to 1e-6 on a held-out object (seed 30) with no
    preconditioner, Jacobi and the trained network:
      * cold start (x0 = 0, random right-hand sides): iteration counts;
      * CN stepping (x0 = u_{s-1}): iterations, time and accuracy of the sweep."""
    nx = ny = 64
    nz = 24
    objs = [workloads.ellipsoid_object_3d(nx, ny, nz, 12, np.random.default_rng(s)) for s in (31, 32)]
    pairs = tl.generate_pairs(objs, K, 2.0, 2.0, 4.0, "dirichlet", per_slice=16, slices=12, seed=33)
    cnn = tl.CNNPreconditioner(seed=0)
    l0 = cnn.gradient(pairs.r[:16], pairs.e[:16], nx, ny)[0]
    losses = cnn.train(pairs, steps=400, batch=16, lr=2e-3)
    print(f"training loss {l0:.3e} -> {losses[-1]:.3e} on {pairs.count} pairs")
    assert losses[-1] < 0.2 * l0
    test_obj = workloads.ellipsoid_object_3d(nx, ny, nz, 12, np.random.default_rng(30))
    b = torch.tensor(_planes(16, nx * ny, np.random.default_rng(34)), device="cuda")
    cold = {}
    for pc in ("none", "jacobi", "cnn"):
        ks = tl.KrylovStepper(test_obj, K, 2.0, 2.0, 4.0, "dirichlet", precond=pc, cnn=cnn, tol=1e-6, max_iter=60)
        its = []
        for s_ in (1, 8, 16):
            x = torch.zeros_like(b)
            done = ks.solve_slice(s_, b, x).cpu().numpy()
            assert (done > 0).all()
            its.append(done.mean())
        cold[pc] = float(np.mean(its))
        print(f"ablation cold start {pc}: mean BiCGSTAB iterations {cold[pc]:.2f}")
    assert cold["cnn"] < cold["jacobi"] <= cold["none"]
    probes = workloads.gaussian_probes_2d(nx, ny, workloads.raster_centres(4, 10.0, nx / 2), 8.0).reshape(16, -1)
    u0 = torch.tensor(probes.astype(np.complex64), device="cuda")
    res = {}
    for pc in ("none", "jacobi", "cnn"):
        ks = tl.KrylovStepper(test_obj, K, 2.0, 2.0, 4.0, "dirichlet", precond=pc, cnn=cnn, tol=1e-6)
        ks.sweep(u0)  # warm-up
        ex, st = ks.sweep(u0)
        res[pc] = (float(np.mean(st.iterations)), st.seconds, ex.cpu().numpy(), st.fallback_probes)
        print(f"ablation stepping {pc}: mean BiCGSTAB iterations {res[pc][0]:.2f}, sweep {st.seconds * 1e3:.1f} ms, "
              f"fallbacks {st.fallback_probes}")
    assert res["cnn"][0] <= res["none"][0]
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, "dirichlet", test_obj, "forward"), probes, keep="exit")
    for pc in res:
        assert rel_l2(res[pc][2], ref) < 1e-4, pc
