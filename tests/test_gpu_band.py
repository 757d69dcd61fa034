"""GPU parity of the two-transverse-axis (3D object) CN stepping against the FP64 oracle.

The oracle factors each slice with SuperLU (iterative_lu.py:66-133 restated);
the GPU path is the block LU with banded Schur inverses (tp_band_factor) and the
persistent block sweep (tp_band_sweep).  Tolerance: relative L2 <= 1e-4
(complex64 data, FP64 factorisation) unless stated.
"""

import numpy as np
import pytest
import torch

from oracle import pwe
from oracle.gradient import exact_adjoint_states
from tests.cases import rel_l2
from thick_ptycho_amd import SimulationSpace, workloads
from thick_ptycho_amd.forward_model import PWEIterativeLUSolver
from thick_ptycho_amd.forward_model import band_engine as be
from thick_ptycho_amd.gradient import object_contribution

pytestmark = pytest.mark.gpu
TOL = 1e-4
K = 2 * np.pi


def _object(nx, ny, nz, rng, amp=1e-2):
    return 1 - (rng.uniform(0, amp, (nx, ny, nz)) + 1j * rng.uniform(0, amp / 10, (nx, ny, nz)))


def _probes(P, nx, ny, rng):
    return rng.standard_normal((P, nx, ny)) + 1j * rng.standard_normal((P, nx, ny))


def _solver(nx, ny, nz, dx, dy, dz, probes, bc):
    ss = SimulationSpace(nx, nz, dx, dz, wavelength=2 * np.pi / K, ny=ny, dy=dy, num_probes=probes.shape[0])
    return PWEIterativeLUSolver(ss, probes[None], bc_type=bc)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
@pytest.mark.parametrize("mode", ["forward", "adjoint", "reverse"])
@pytest.mark.parametrize("nx,ny,nz", [(24, 20, 10), (65, 9, 7)])
def test_modes_match_oracle(gpu, bc, mode, nx, ny, nz, rng):
    if mode == "reverse" and bc == "impedance":
        nz = 5  # marching back through an absorbing boundary grows ~1e3 per slice
    print(f"case bc={bc} mode={mode} nx={nx} ny={ny} nz={nz}", flush=True)
    n = _object(nx, ny, nz, rng)
    probes = _probes(3, nx, ny, rng)
    got = _solver(nx, ny, nz, 2.0, 3.0, 4.0, probes, bc).solve(n=n, mode=mode)[0, 0]
    assert got.shape == (3, nx, ny, nz)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=3.0, dz=4.0, k=K)
    ref = pwe.solve(g, bc, n, probes, mode)
    assert rel_l2(got.reshape(ref.shape), ref) < TOL


def test_exact_adjoint_states(gpu, rng):
    """exact_adjoint mode against the oracle's lambda_s = A_s^{-H} B_{s+1}^H lambda_{s+1}."""
    nx, ny, nz = 32, 24, 12
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n = _object(nx, ny, nz, rng)
    seed = (rng.standard_normal((4, nx * ny)) + 1j * rng.standard_normal((4, nx * ny)))
    lam = exact_adjoint_states(g, "impedance", n, seed)
    dev = torch.device("cuda")
    A, B = be.step_operators_3d(nx, ny, K, 2.0, 2.0, 4.0, "impedance")
    C = object_contribution(torch.tensor(n.reshape(nx * ny, nz), dtype=torch.complex128, device=dev), K, 4.0)
    fac = be.factor(C, A, B, "exact_adjoint")
    L = nz - 1
    out = torch.empty((4, L + 1, nx * ny), dtype=torch.complex64, device=dev)
    u0 = torch.zeros((4, nx * ny), dtype=torch.complex64, device=dev)
    rhs = torch.tensor(seed, dtype=torch.complex64, device=dev).reshape(4, 1, nx * ny)
    be.sweep([be.BandScanSpec(fac, u0, out=out, rhs=rhs, rhs_steps=1)])
    got = out.cpu().numpy()
    ref = np.stack([lam[:, :, L - t + 1] for t in range(1, L + 1)], axis=1)
    assert rel_l2(got[:, 1:], ref) < TOL


@pytest.mark.parametrize("mode", ["forward", "adjoint"])
@pytest.mark.parametrize("per_probe", [False, True])
def test_rhs_block(gpu, mode, per_probe, rng):
    nx, ny, nz = 20, 16, 9
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n = _object(nx, ny, nz, rng)
    probes = _probes(3, nx, ny, rng)
    N, L = nx * ny, nz - 1
    rb = rng.standard_normal((3, N, L) if per_probe else (N, L)) + 1j * rng.standard_normal((3, N, L) if per_probe
                                                                                           else (N, L))
    s = _solver(nx, ny, nz, 2.0, 2.0, 4.0, probes, "neumann")
    got = s.solve_device(n=n, mode=mode, rhs_block=rb, per_probe_rhs=per_probe)[0, 0]
    got = got.transpose(-1, -2).cpu().numpy()
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, "neumann", n, mode), probes.reshape(3, N), mode, rhs_block=rb,
                            per_probe_rhs=per_probe)
    assert rel_l2(got, ref) < TOL


def test_band_widens_for_slow_decay(gpu, rng):
    """dx = dy = 0.5 at dz = 4 (mu ~ 0.64): the inverse blocks decay slowly, the factor
    widens the band beyond the default and the result still matches."""
    nx, ny, nz = 40, 12, 6
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=0.5, dy=0.5, dz=4.0, k=K)
    n = _object(nx, ny, nz, rng)
    probes = _probes(2, nx, ny, rng)
    s = _solver(nx, ny, nz, 0.5, 0.5, 4.0, probes, "dirichlet")
    got = s.solve(n=n)[0, 0]
    fac = s.projection_cache[0].modes["forward"].factors
    print(f"band half-width {fac.w}, edge ratio {fac.edge:.2e}")
    assert fac.w > 6 and fac.edge <= be.EDGE_TOL
    ref = pwe.solve(g, "dirichlet", n, probes, "forward")
    assert rel_l2(got.reshape(ref.shape), ref) < TOL


def test_two_scans_one_launch(gpu, rng):
    """Scans of different transverse sizes share one persistent launch."""
    dev = torch.device("cuda")
    specs, refs = [], []
    for nx, ny, P in [(48, 32, 3), (20, 36, 5)]:
        nz = 8
        g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
        n = _object(nx, ny, nz, rng)
        probes = _probes(P, nx, ny, rng).reshape(P, -1)
        A, B = be.step_operators_3d(nx, ny, K, 2.0, 2.0, 4.0, "impedance")
        C = object_contribution(torch.tensor(n.reshape(nx * ny, nz), dtype=torch.complex128, device=dev), K, 4.0)
        fac = be.factor(C, A, B, "forward", w=8)
        ex = torch.empty((P, nx * ny), dtype=torch.complex64, device=dev)
        specs.append(be.BandScanSpec(fac, torch.tensor(probes, dtype=torch.complex64, device=dev), exit=ex))
        refs.append(pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", n, "forward"), probes, keep="exit"))
    be.sweep(specs)
    for sp_, ref in zip(specs, refs):
        assert rel_l2(sp_.exit.cpu().numpy(), ref) < TOL


@pytest.fixture(scope="module")
def config3_reference():
    """Config 3 (128 x 128 x 256, 64 probes) and its FP64 oracle exit waves (SuperLU per slice)."""
    sc = workloads.config3().scans[0]
    P = sc.probes.shape[0]
    probes = sc.probes.reshape(P, -1)
    g = pwe.Grid(dimension=3, nx=sc.nx, ny=sc.ny, nz=sc.nz, dx=sc.dx, dy=sc.dy, dz=sc.dz, k=sc.k)
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, sc.bc, sc.n, "forward"), probes, keep="exit")
    return sc, probes, ref


def test_config3_parity(gpu, config3_reference):
    """Config 3 forward: 128 x 128 x 256, 64 probes, exit waves against the FP64 oracle."""
    sc, probes, ref = config3_reference
    P = probes.shape[0]
    ss = SimulationSpace(sc.nx, sc.nz, sc.dx, sc.dz, wavelength=2 * np.pi / sc.k, ny=sc.ny, dy=sc.dy, num_probes=P)
    s = PWEIterativeLUSolver(ss, probes.reshape(1, P, sc.nx, sc.ny), bc_type=sc.bc)
    got = s.solve_device(n=sc.n, keep="exit")[0, 0].cpu().numpy()
    err = rel_l2(got, ref)
    print(f"config 3 exit-wave rel-L2 {err:.2e}, band half-width {s.projection_cache[0].modes['forward'].factors.w}")
    assert err < TOL


def test_config3_learned_krylov_parity(gpu, config3_reference):
    """Config 3 with the learned form: CN stepping by BiCGSTAB to 1e-6 preconditioned by the
    CNN (trained on banded-LU residual/correction pairs of two held-out objects), exit
    waves of all 64 probes after 255 slices against the FP64 oracle (<= 1e-4), with no
    banded-LU fallback."""
    from thick_ptycho_amd import learned as tl

    sc, probes, ref = config3_reference
    P = probes.shape[0]
    objs = [workloads.ellipsoid_object_3d(sc.nx, sc.ny, 64, 32, np.random.default_rng(s)) for s in (31, 32)]
    pairs = tl.generate_pairs(objs, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, per_slice=16, slices=32, seed=33)
    cnn = tl.CNNPreconditioner(seed=0)
    losses = cnn.train(pairs, steps=400, batch=32, lr=2e-3)
    ks = tl.KrylovStepper(sc.n, sc.k, sc.dx, sc.dy, sc.dz, sc.bc, precond="cnn", cnn=cnn, tol=1e-6)
    ex, st = ks.sweep(torch.tensor(probes.astype(np.complex64), device="cuda"))
    err = rel_l2(ex.cpu().numpy(), ref)
    print(f"config 3 CNN-preconditioned BiCGSTAB: exit rel-L2 {err:.2e}, mean iterations "
          f"{np.mean(st.iterations):.2f}, fallbacks {st.fallback_probes}, training loss {losses[-1]:.2e}")
    assert st.fallback_probes == 0
    assert err < TOL


@pytest.mark.parametrize("ppb,schedule", [(2, 0), (4, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("form", ["out_rhs_sigma", "exit_ckpt", "two_scans"])
def test_probes_per_block_matches_one(gpu, ppb, schedule, form, rng):
    """Probe groups sharing each band-coefficient load give the same planes as one
    probe per workgroup (ragged groups: 7 and 5 probes)."""
    dev = torch.device("cuda")
    print(f"case ppb={ppb} schedule={schedule} form={form}", flush=True)

    def build(nx, ny, nz, P, mode, bc="impedance", w=None):
        n = _object(nx, ny, nz, rng)
        A, B = be.step_operators_3d(nx, ny, K, 2.0, 2.0, 4.0, bc)
        C = object_contribution(torch.tensor(n.reshape(nx * ny, nz), dtype=torch.complex128, device=dev), K, 4.0)
        fac = be.factor(C, A, B, mode, w=w)
        u0 = torch.tensor(_probes(P, nx, ny, rng).reshape(P, -1), dtype=torch.complex64, device=dev)
        return fac, u0

    def specs_for():
        if form == "out_rhs_sigma":
            fac, u0 = build(40, 24, 9, 7, "exact_adjoint")
            rhs = torch.randn((7, 1, 40 * 24), dtype=torch.complex64, device=dev)
            return [be.BandScanSpec(fac, u0, out=torch.empty((7, 9, 40 * 24), dtype=torch.complex64, device=dev),
                                    rhs=rhs, rhs_steps=1)]
        if form == "exit_ckpt":
            fac, u0 = build(33, 20, 11, 5, "forward")
            return [be.BandScanSpec(fac, u0, exit=torch.empty((5, 33 * 20), dtype=torch.complex64, device=dev),
                                    ckpt=torch.empty((5, 10 // 4 + 1, 33 * 20), dtype=torch.complex64, device=dev),
                                    ckpt_every=4)]
        out = []
        for nx, ny, P in [(48, 32, 7), (20, 36, 5)]:
            fac, u0 = build(nx, ny, 8, P, "forward", w=8)
            out.append(be.BandScanSpec(fac, u0, out=torch.empty((P, 8, nx * ny), dtype=torch.complex64, device=dev)))
        return out

    specs = specs_for()

    def outputs():
        return [t.clone() for s in specs for t in (s.out, s.exit, s.ckpt) if t is not None]

    be.sweep(specs, probes_per_block=1)
    ref = outputs()
    for s in specs:
        for t in (s.out, s.exit, s.ckpt):
            if t is not None:
                t.fill_(float("nan"))
    if schedule == 1 and specs[0].fac.w > 8:
        pytest.skip("schedule 1 is built for w <= 8")
    be.sweep(specs, probes_per_block=ppb, schedule=schedule)
    for a, b in zip(outputs(), ref):
        assert torch.equal(a, b)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
@pytest.mark.parametrize("mode", ["adjoint", "exact_adjoint"])
def test_adjoint_factors_from_forward(gpu, bc, mode, rng):
    """tp_band_conj_transpose + tp_band_sigma reproduce a direct factorisation of the
    adjoint operators (W'_j = W_j^H, slices reversed)."""
    nx, ny, nz = 36, 28, 10
    dev = torch.device("cuda")
    n = _object(nx, ny, nz, rng)
    A, B = be.step_operators_3d(nx, ny, K, 2.0, 2.0, 4.0, bc)
    C = object_contribution(torch.tensor(n.reshape(nx * ny, nz), dtype=torch.complex128, device=dev), K, 4.0)
    w = be.factor(C, A, B, "forward").w
    fwd = be.BandFactorizer(A, B, "forward", nz - 1, w, dev)
    direct = be.BandFactorizer(A, B, mode, nz - 1, w, dev)
    derived = be.BandFactorizer(A, B, mode, nz - 1, w, dev, source=fwd)
    fwd.run(C)
    a, b = direct.run(C), derived.run(C)
    direct.check()
    derived.check()
    scale = a.W.abs().max()
    assert float((a.W - b.W).abs().max() / scale) < 1e-6
    if a.sigma is not None:
        assert torch.equal(a.sigma, b.sigma)
