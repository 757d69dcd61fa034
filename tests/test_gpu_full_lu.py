"""Drop-in PWEFullLUSolver (whole space-z system; block forward / backward
substitution on the stepping kernels) against the oracle's SuperLU solve of the
assembled whole-system matrix (pwe.full_lu_solve: lu.solve(b) and
lu.solve(b, trans='H'), full_lu.py:114-172).  Tolerance: relative L2 <= 1e-4."""

import numpy as np
import pytest

from oracle import pwe
from tests.cases import random_object, rel_l2
from thick_ptycho_amd import SimulationSpace, workloads
from thick_ptycho_amd.forward_model import PWEFullLUSolver

pytestmark = pytest.mark.gpu
K = 2 * np.pi


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
@pytest.mark.parametrize("mode", ["forward", "adjoint"])
def test_full_lu_matches_oracle_2d(gpu, bc, mode, rng):
    nx, nz, P = 120, 20, 3
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    ss = SimulationSpace(nx, nz, 2.0, 4.0, wavelength=2 * np.pi / K, num_probes=P)
    s = PWEFullLUSolver(ss, probes[None], bc_type=bc)
    got = s.solve(n=n, mode=mode)[0, 0]
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=K)
    for p in range(P):
        assert rel_l2(got[p], pwe.full_lu_solve(g, bc, n, probes[p], mode)) < 1e-4, p
    one = s._solve_single_probe(scan_idx=1, n=n, probe=probes[1], mode=mode)
    assert rel_l2(one, pwe.full_lu_solve(g, bc, n, probes[1], mode)) < 1e-4


@pytest.mark.parametrize("mode", ["forward", "adjoint"])
def test_full_lu_rhs_block(gpu, mode, rng):
    nx, nz = 64, 12
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = rng.standard_normal((2, nx)) + 0j
    rhs = rng.standard_normal(nx * (nz - 1)) + 1j * rng.standard_normal(nx * (nz - 1))
    ss = SimulationSpace(nx, nz, 2.0, 4.0, wavelength=2 * np.pi / K, num_probes=2)
    got = PWEFullLUSolver(ss, probes[None], bc_type="impedance").solve(n=n, mode=mode, rhs_block=rhs)[0, 0]
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=K)
    ref = pwe.full_lu_solve(g, "impedance", n, probes[0], mode, rhs_block=rhs)
    assert rel_l2(got[0, :, 1:], ref[:, 1:]) < 1e-4


@pytest.mark.parametrize("mode", ["forward", "adjoint"])
def test_full_lu_two_axes(gpu, mode):
    rng = np.random.default_rng(12)
    nx, ny, nz, P = 16, 12, 7, 2
    n = workloads.ellipsoid_object_3d(nx, ny, nz, 3, rng, semi=(2.0, 5.0), delta=(0, 5e-3), beta=(0, 5e-5))
    probes = workloads.gaussian_probes_2d(nx, ny, [(6, 5), (9, 7)], 3.0).reshape(P, -1)
    ss = SimulationSpace(nx, nz, 2.0, 4.0, ny=ny, dy=2.0, wavelength=2 * np.pi / K, num_probes=P)
    got = PWEFullLUSolver(ss, probes.reshape(1, P, nx, ny), bc_type="dirichlet").solve(n=n, mode=mode)[0, 0]
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    got = got.reshape(P, nx * ny, nz)
    for p in range(P):
        ref = pwe.full_lu_solve(g, "dirichlet", n, probes[p], mode)
        assert rel_l2(got[p], ref) < 1e-4, p


@pytest.mark.parametrize("nx,nz", [(37, 2), (1000, 3)])
def test_full_lu_edge_sizes(gpu, nx, nz, rng):
    n = random_object(nx, nz, rng, amp=1e-2)
    probes = rng.standard_normal((2, nx)) + 1j * rng.standard_normal((2, nx))
    ss = SimulationSpace(nx, nz, 2.0, 4.0, wavelength=2 * np.pi / K, num_probes=2)
    s = PWEFullLUSolver(ss, probes[None], bc_type="neumann")
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=K)
    for mode in ("forward", "adjoint"):
        got = s.solve(n=n, mode=mode)[0, 0]
        for p in range(2):
            assert rel_l2(got[p], pwe.full_lu_solve(g, "neumann", n, probes[p], mode)) < 1e-4, (mode, p)
