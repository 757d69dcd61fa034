"""The streaming FP64 3D gradient against SuperLU and the all-planes oracle (CPU)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import gradient as og
from oracle import pwe
from oracle import recon3d as orc
from tests.cases import rel_l2
from thick_ptycho_amd import workloads

K = 2 * np.pi


@pytest.mark.parametrize("bc", ["dirichlet", "neumann", "impedance"])
@pytest.mark.parametrize("herm", [False, True])
def test_block_thomas_matches_superlu(bc, herm, rng):
    nx, ny = 9, 7
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=3, dx=2.0, dy=1.5, dz=4.0, k=K)
    A0, _ = pwe.step_operators(g, bc)
    c = 0.3 * (rng.standard_normal(nx * ny) + 1j * rng.standard_normal(nx * ny))
    A = (A0 - sp.diags(c)).tocsc()
    D, E, F = orc.operator_blocks(A, nx)
    R = rng.standard_normal((nx * ny, 5)) + 1j * rng.standard_normal((nx * ny, 5))
    got = orc.BlockTriLU(D, E, F).solve(R, herm=herm)
    M = A.conj().T.tocsc() if herm else A
    ref = spla.splu(M).solve(R)
    assert rel_l2(got, ref) < 1e-12


def test_forward_exit_matches_superlu_stepping(rng):
    nx, ny, nz, P = 8, 10, 7, 3
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n = workloads.ellipsoid_object_3d(nx, ny, nz, 3, rng, semi=(2, 4), delta=(0, 5e-3), beta=(0, 5e-5))
    probes = rng.standard_normal((P, nx * ny)) + 1j * rng.standard_normal((P, nx * ny))
    ref = pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", n, "forward"), probes, keep="exit")
    assert rel_l2(orc.forward_exit_3d(g, "impedance", n, probes), ref) < 1e-12


def test_operator_blocks_rejects_non_block_tridiagonal():
    M = sp.random(12, 12, density=0.5, random_state=0, format="csr") + sp.eye(12)
    with pytest.raises(ValueError):
        orc.operator_blocks(M, 3)


@pytest.mark.parametrize("seg", [3, 4, 100])
def test_streaming_gradient_matches_all_planes(seg):
    nx, ny, nz, P = 10, 12, 9, 3
    rng = np.random.default_rng(7)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n = workloads.ellipsoid_object_3d(nx, ny, nz, 3, rng, semi=(2, 4), delta=(0, 5e-3), beta=(0, 5e-5))
    probes = rng.standard_normal((P, nx * ny)) + 1j * rng.standard_normal((P, nx * ny))
    I = np.abs(rng.standard_normal((P, nx, ny))) ** 2 * nx * ny
    got, J = orc.gradient_3d(g, "dirichlet", n, probes, I, seg=seg)
    ref, Jr = og.exact_gradient_3d(g, "dirichlet", n, probes, I)
    assert abs(J - Jr) <= 1e-12 * Jr
    assert rel_l2(got, ref) < 1e-11


def test_streaming_reconstruction_matches():
    nx, ny, nz, P = 8, 8, 6, 2
    rng = np.random.default_rng(8)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n_true = workloads.ellipsoid_object_3d(nx, ny, nz, 3, rng, semi=(2, 4), delta=(0, 1e-2), beta=(0, 1e-4))
    probes = rng.standard_normal((P, nx * ny)) + 1j * rng.standard_normal((P, nx * ny))
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, "neumann", n_true, "forward"), probes, keep="exit")
    I = np.abs(np.fft.fft2(u.reshape(P, nx, ny), axes=(-2, -1))) ** 2
    n0 = np.ones((nx, ny, nz), dtype=complex)
    step = orc.lipschitz_step(g, "neumann", n0, probes, I, seg=2)
    n_a, h_a = orc.reconstruct_3d(g, "neumann", n0, probes, I, 3, step, seg=2)
    n_b, h_b = og.reconstruct_3d(g, "neumann", n0, probes, I, 3, step)
    assert h_a[-1] < h_a[0]
    assert np.allclose(h_a, h_b, rtol=1e-11)
    assert rel_l2(n_a - n0, n_b - n0) < 1e-10


def test_amplitude_seed_conditioning_from_free_space():
    """Why complex64 reconstructions from free space cannot follow the FP64 update
    closely: most detector bins of a free-space probe's far field are below 1e-6 of
    its peak while the data there are not, so perturbing the exit waves by 1e-6 of
    their peak (less than 127 complex64 slices accumulate) moves the FP64 gradient
    by ~7e-4 here (64² x 12, 9 probes; 6e-3 for 1e-6 and 1.4e-2 for 1e-5 at 128² x 32
    with 64 probes).  A property of the loss, computed entirely in FP64.""" 
    nx = ny = 64
    nz, P = 12, 9
    rng = np.random.default_rng(40)
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=K)
    n_true = workloads.ellipsoid_object_3d(nx, ny, nz, 6, rng)
    probes = workloads.gaussian_probes_2d(nx, ny, workloads.raster_centres(3, 12.8, 32.0), 8.0).reshape(P, -1)
    ex = orc.forward_exit_3d(g, "dirichlet", n_true, probes)
    I = np.abs(np.fft.fft2(ex.reshape(P, nx, ny), axes=(-2, -1))) ** 2
    n0 = np.ones_like(n_true)
    ex0 = orc.forward_exit_3d(g, "dirichlet", n0, probes)
    U = np.abs(np.fft.fft2(ex0.reshape(P, nx, ny), axes=(-2, -1)))
    assert np.mean(U < 1e-6 * U.max()) > 0.3
    g0, _ = orc.gradient_3d(g, "dirichlet", n0, probes, I, seed=og.amplitude_loss_seed_2d(ex0, I, nx, ny))
    r = np.random.default_rng(1)
    exn = ex0 + 1e-6 * np.abs(ex0).max() * (r.standard_normal(ex0.shape) + 1j * r.standard_normal(ex0.shape))
    g1, _ = orc.gradient_3d(g, "dirichlet", n0, probes, I, seed=og.amplitude_loss_seed_2d(exn, I, nx, ny))
    d = rel_l2(g1, g0)
    print(f"gradient change from a 1e-6 exit-wave perturbation: {d:.2e}")
    assert d > 1e-4
