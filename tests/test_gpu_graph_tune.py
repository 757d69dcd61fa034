"""HIP-graph replay of the plain form and the batch-shape autotuner (GPU)."""

import numpy as np
import pytest
import torch

from oracle import pwe
from tests.cases import random_object, rel_l2
from thick_ptycho_amd import SimulationSpace, autotune
from thick_ptycho_amd.forward_model import PWEIterativeLUSolver, operators, tri_engine

pytestmark = pytest.mark.gpu
K = 2 * np.pi


def _specs(rng, widths=(200, 96), P=5, nz=20, mode="forward"):
    specs = []
    chunk = operators.chunk_for(max(widths))  # one chunk for every scan of a launch
    for nx in widths:
        n = random_object(nx, nz, rng, amp=1e-2)
        A, B = operators.step_operators_1d(nx, K, 2.0, 4.0, "impedance")
        coef = (K / 1j) * (n - 1)
        C = torch.tensor((4.0 / 2) * (coef[:, :-1] + coef[:, 1:]) / 2, device="cuda")
        fac = tri_engine.factor(C, A, B, mode, chunk=chunk)
        u0 = torch.tensor((rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))).astype(np.complex64),
                          device="cuda")
        out = torch.empty((P, nz, nx), dtype=torch.complex64, device="cuda")
        specs.append(tri_engine.TriScanSpec(fac, u0, out=out))
    return specs


def test_plain_graph_replays_match_eager(gpu, rng):
    specs = _specs(rng)
    tri_engine.sweep_plain(specs)
    eager = [s.out.clone() for s in specs]
    g = tri_engine.PlainGraph(specs)
    for rep in range(3):
        for s in specs:
            s.out.zero_()
        junk = torch.randn(256, 256, device="cuda") @ torch.randn(256, 256, device="cuda")  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for s, e in zip(specs, eager):
            assert torch.equal(s.out, e), rep
        del junk
    # new probes written into the captured input in place
    for s in specs:
        s.u0.copy_(torch.randn_like(s.u0.real).to(torch.complex64) * (1 + 1j))
    g.replay()
    got = [s.out.clone() for s in specs]
    tri_engine.sweep_plain(specs)
    for s, r in zip(specs, got):
        assert torch.equal(s.out, r)


def test_autotune_choice_and_results(gpu, rng):
    specs = _specs(rng, widths=(1024, 640), P=33, nz=12)
    res = autotune.tune_sweep(specs)
    assert res["ppb"] in autotune.candidates(specs)
    assert set(res["ms"]) == {str(c) for c in autotune.candidates(specs)}
    tri_engine.sweep(specs, probes_per_block=res["ppb"])
    a = [s.out.clone() for s in specs]
    tri_engine.sweep(specs, probes_per_block=1)
    for s, x in zip(specs, a):
        assert rel_l2(x.cpu().numpy(), s.out.cpu().numpy()) < 1e-6
    assert autotune.tune_sweep(specs) is res  # cached per shape


def test_solver_auto_probes_per_block(gpu, rng):
    nx, nz = 300, 24
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = rng.standard_normal((6, nx)) + 1j * rng.standard_normal((6, nx))
    ss = SimulationSpace(nx, nz, 2.0, 4.0, num_probes=6)
    s = PWEIterativeLUSolver(ss, probes[None], bc_type="dirichlet", probes_per_block="auto")
    got = s.solve(n=n)[0, 0]
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=K)
    assert rel_l2(got, pwe.solve(g, "dirichlet", n, probes, "forward")) < 1e-4


def test_factor_batch_matches_single(gpu, rng):
    """tp_tri_factor_batch (one launch for several scans and modes) writes the same
    factors as one tp_tri_factor call per job."""
    jobs = []
    for nx in (200, 96, 320):
        n = random_object(nx, 16, rng, amp=1e-2)
        A, B = operators.step_operators_1d(nx, K, 2.0, 4.0, "impedance")
        coef = (K / 1j) * (n - 1)
        C = torch.tensor((4.0 / 2) * (coef[:, :-1] + coef[:, 1:]) / 2, device="cuda")
        for mode in ("forward", "exact_adjoint"):
            jobs.append((C, A, B, mode))
    chunk = operators.chunk_for(320)
    single = [tri_engine.factor(C, A, B, mode, chunk=chunk) for C, A, B, mode in jobs]
    batch = tri_engine.factor_many(jobs, chunk)

    def written(f):  # slots [L][6][64*chunk]; the sigma slot (5) is written only in sigma modes
        c = f.coef.view(f.steps, 6, -1)
        return c if f.sigma else c[:, :5]

    for a, b in zip(single, batch):
        assert torch.equal(written(a), written(b))
