"""Near-carry form of the persistent tridiagonal sweep.

The factor kernel records the coupling reach of its factors (the rows over
which a carry from a neighbouring 64-lane chunk keeps a factor above 2^-32);
below the chunk, the sweep replaces the two 64-lane affine scans by the
neighbouring chunk's local end value and applies it to `reach` rows only, and
derives the multipliers l and q from the stored 1/piv and the operator's
off-diagonals instead of staging them.

Checks: the reach equals the one of a sequential FP64 Thomas factorisation of
the same per-slice operators; near and full carries give the same planes to
FP32 rounding (<= 1e-6 relative L2); both meet the FP64 oracle at 1e-4 where
the coupling is weak (config-1 parameters, dx = 2, dz = 4), intermediate
(dx = 1) and strong enough that the reach covers a whole chunk (dx = 0.25,
which takes the full scans).
"""

import numpy as np
import pytest
import torch

from oracle import pwe
from tests.cases import random_object, rel_l2
from thick_ptycho_amd.forward_model import operators, tri_engine

pytestmark = pytest.mark.gpu
TAU = 2.0 ** -32


def _fp64_reach(g, bc, n, chunk):
    """Largest reach over the slices of the forward mode, from a sequential
    FP64 Thomas factorisation of A0 - diag(C[:, s-1]) split into 64 chunks."""
    A0, _ = pwe.step_operators_1d(g, bc)
    C = pwe.object_contribution(g, n)
    nx, L = C.shape
    sub0, dia0, sup0 = A0.diagonal(-1), A0.diagonal(0), A0.diagonal(1)
    best = 0
    for s in range(L):
        d = dia0 - C[:, s]
        piv = np.empty(nx, complex)
        lmul = np.zeros(nx, complex)
        q = np.zeros(nx, complex)
        piv[0] = d[0]
        for i in range(1, nx):
            lmul[i] = sub0[i - 1] / piv[i - 1]
            piv[i] = d[i] - lmul[i] * sup0[i - 1]
        q[:-1] = -sup0 / piv[:-1]
        for lane in range(64):
            r0 = lane * chunk
            rows = range(r0, min(r0 + chunk, nx))
            gf = 1.0 + 0j
            for m, i in enumerate(rows):
                gf *= -lmul[i]
                if abs(gf.real) + abs(gf.imag) > TAU:
                    best = max(best, m + 1)
            gb = 1.0 + 0j
            for m in range(chunk - 1, -1, -1):
                i = r0 + m
                gb *= q[i] if i < nx else 0.0
                if abs(gb.real) + abs(gb.imag) > TAU:
                    best = max(best, chunk - m)
    return best


def _case(nx, nz, dx, P, rng):
    k = 2 * np.pi
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=dx, dz=4.0, k=k)
    n = random_object(nx, nz, rng, amp=1e-3)
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    return g, n, probes


def _sweep(g, bc, n, probes, mode, chunk, near):
    A, B = operators.step_operators_1d(g.nx, g.k, g.dx, g.dz, bc)
    C = torch.from_numpy(pwe.object_contribution(g, n)).to("cuda")
    fac = tri_engine.factor(C, A, B, mode, chunk=chunk)
    u0 = torch.from_numpy(probes).to("cuda", torch.complex64)
    out = torch.empty((u0.shape[0], g.nz, g.nx), dtype=torch.complex64, device="cuda")
    tri_engine.sweep([tri_engine.TriScanSpec(fac, u0, out=out, near=near)])
    return fac, out


@pytest.mark.parametrize("dx,expect", [(2.0, "near"), (1.0, "near"), (0.25, "full")])
def test_reach_matches_fp64_thomas(gpu, rng, dx, expect):
    print(f"case reach dx={dx}", flush=True)
    g, n, probes = _case(1024, 9, dx, 2, rng)
    fac, _ = _sweep(g, "impedance", n, probes, "forward", 16, True)
    got = int(fac.status[1].item())
    ref = _fp64_reach(g, "impedance", n, 16)
    assert abs(got - ref) <= 1, (got, ref)
    assert (got < 16) == (expect == "near"), got


@pytest.mark.parametrize("dx", [2.0, 1.0, 0.25])
@pytest.mark.parametrize("mode", ["forward", "adjoint", "exact_adjoint"])
@pytest.mark.parametrize("nx,chunk", [(1024, 16), (700, 16), (512, 8), (2048, 32)])
def test_near_and_full_carries_agree(gpu, rng, dx, mode, nx, chunk):
    print(f"case near/full dx={dx} mode={mode} nx={nx} chunk={chunk}", flush=True)
    g, n, probes = _case(nx, 33, dx, 5, rng)
    bc = "impedance"
    _, near = _sweep(g, bc, n, probes, mode, chunk, True)
    _, full = _sweep(g, bc, n, probes, mode, chunk, False)
    assert rel_l2(near.cpu().numpy(), full.cpu().numpy()) < 1e-6
    if mode != "exact_adjoint":
        ref = pwe.solve(g, bc, n, probes, mode)  # (P, nx, nz)
        got = near.cpu().numpy().transpose(0, 2, 1)
        if mode == "adjoint":
            got = got[:, :, ::-1]
        assert rel_l2(got, ref) < 1e-4


def test_shared_status_covers_every_window(gpu, rng):
    """Per-probe factor sets stepped in one sweep share one status: its reach is
    the largest of the windows', so a strongly coupled window keeps the full
    carries for all."""
    k = 2 * np.pi
    jobs, C_list = [], []
    for dx in (2.0, 0.25):
        g = pwe.Grid(dimension=2, nx=512, nz=9, dx=dx, dz=4.0, k=k)
        A, B = operators.step_operators_1d(512, k, dx, 4.0, "impedance")
        C = torch.from_numpy(pwe.object_contribution(g, random_object(512, 9, rng, amp=1e-3))).to("cuda")
        jobs.append((C, A, B, "forward"))
    own = [tri_engine.factor_many([j], 8)[0] for j in jobs]
    shared = tri_engine.new_status("cuda")
    facs = tri_engine.factor_many(jobs, 8, status=shared)
    reach = [int(f.status[1].item()) for f in own]
    assert reach[0] < 8 <= reach[1]
    assert int(shared[1].item()) == max(reach)
    assert all(f.status is shared for f in facs)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
@pytest.mark.parametrize("mode", ["forward", "adjoint", "reverse", "exact_adjoint"])
@pytest.mark.parametrize("nx", [1009, 1010, 1024, 641, 20])
def test_boundary_rows_of_the_derived_multipliers(gpu, rng, bc, mode, nx):
    """The near-carry step derives l and q from 1/piv with the operator's boundary
    off-diagonals (sub at rows 1 and n-1, sup at rows 0 and n-2) placed at every
    position in a lane: n - 1 on a chunk's first row (1009, 641), second row
    (1010), last row (1024), and both ends in the first two lanes (20)."""
    print(f"case boundary rows bc={bc} mode={mode} nx={nx}", flush=True)
    g, n, probes = _case(nx, 9, 2.0, 3, rng)
    if mode == "reverse" and bc == "impedance":
        g = pwe.Grid(dimension=2, nx=nx, nz=5, dx=2.0, dz=4.0, k=g.k)
        n = n[:, :5]
    fac_n, near = _sweep(g, bc, n, probes, mode, 16, True)
    _, full = _sweep(g, bc, n, probes, mode, 16, False)
    assert int(fac_n.status[1].item()) < 16
    assert tri_engine.near_offdiagonals(fac_n.plan, nx) is not None
    assert rel_l2(near.cpu().numpy(), full.cpu().numpy()) < 1e-6
    if mode != "exact_adjoint":
        ref = pwe.solve(g, bc, n, probes, mode)
        got = near.cpu().numpy().transpose(0, 2, 1)
        if mode == "adjoint":
            got = got[:, :, ::-1]
        assert rel_l2(got, ref) < 1e-4


@pytest.mark.parametrize("mode", ["forward", "adjoint"])
def test_wave_mode_at_every_batch_shape(gpu, rng, mode):
    """The near step without sigma or gradient runs in the wave mode: each
    wavefront copies its pivots into an LDS buffer of its own and the
    wavefronts of a workgroup step without barriers.  At 1, 2, 4 and 8 probes
    per workgroup (8 buffers outnumber the shared coefficient block's slots),
    with a scan at half the launch's chunk and probe counts that leave the last
    workgroup part-filled: full fields bitwise equal across the batch shapes,
    within 1e-6 of the full carries and 1e-4 of the FP64 oracle."""
    print(f"case wave mode {mode}", flush=True)
    k, bc = 2 * np.pi, "impedance"
    cases = []
    for nx, P in ((1000, 13), (384, 6)):  # 16 and 8 rows per lane
        g = pwe.Grid(dimension=2, nx=nx, nz=40, dx=2.0, dz=4.0, k=k)
        n = random_object(nx, 40, rng, amp=1e-3)
        probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
        A, B = operators.step_operators_1d(nx, k, 2.0, 4.0, bc)
        C = torch.from_numpy(pwe.object_contribution(g, n)).to("cuda")
        fac = tri_engine.factor(C, A, B, mode, chunk=operators.chunk_for(nx))
        assert int(fac.status[1].item()) < fac.chunk
        u0 = torch.from_numpy(probes).to("cuda", torch.complex64)
        cases.append((g, n, probes, fac, u0))
    assert [c[3].chunk for c in cases] == [16, 8]

    def run(ppb, near=True):
        outs = [torch.empty((c[4].shape[0], c[0].nz, c[0].nx), dtype=torch.complex64, device="cuda") for c in cases]
        tri_engine.sweep([tri_engine.TriScanSpec(c[3], c[4], out=o, near=near) for c, o in zip(cases, outs)],
                         probes_per_block=ppb)
        return [o.cpu().numpy() for o in outs]

    base = run(4)
    for ppb in (1, 2, 8):
        for got, ref in zip(run(ppb), base):
            np.testing.assert_array_equal(got, ref)
    for got, full, (g, n, probes, _, _) in zip(base, run(4, near=False), cases):
        assert rel_l2(got, full) < 1e-6
        ref = pwe.solve(g, bc, n, probes, mode)
        got = got.transpose(0, 2, 1)
        if mode == "adjoint":
            got = got[:, :, ::-1]
        assert rel_l2(got, ref) < 1e-4
