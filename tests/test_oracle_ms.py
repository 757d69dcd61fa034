"""Pin the FP64 multislice / far-field restatement with closed forms (CPU)."""

import numpy as np
import pytest

from oracle import multislice as oms
from tests.cases import rel_l2


def test_plane_waves_are_eigenvectors():
    """exp(2 pi i m x / N) propagates by H(k_m) per slice in free space."""
    N, nz, dx, dz, k = 128, 6, 0.4, 1.5, 2 * np.pi
    x = np.arange(N)
    for m in (0, 5, 40, -17):
        v = np.exp(2j * np.pi * m * x / N)
        u = oms.solve_single_probe(np.ones((N, nz)), v, dx, k, dz)
        kx = 2 * np.pi * m / (N * dx)
        kz = np.sqrt(max(k * k - kx * kx, 0.0))
        lam = np.exp(1j * (kz - k) * dz)
        for z in range(nz):
            assert np.allclose(u[:, z], v * lam**z, atol=1e-12)


def test_evanescent_band_clipped_not_decaying():
    """kz is clipped to 0 beyond k (quirk Q6): |H| = 1 everywhere."""
    H = oms.propagation_kernel(256, 0.1, 2 * np.pi, 1.0)
    assert np.allclose(np.abs(H), 1.0)
    kx = 2 * np.pi * np.fft.fftfreq(256, 0.1)
    ev = np.abs(kx) > 2 * np.pi
    assert ev.any() and np.allclose(H[ev], np.exp(-1j * 2 * np.pi * 1.0))


def test_constant_object_phase_and_parseval():
    N, nz, dx, dz, k = 64, 5, 1.0, 2.0, 2 * np.pi
    rng = np.random.default_rng(0)
    probe = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    dn = 3e-3 + 1e-4j
    n = np.full((N, nz), 1 + dn)
    u = oms.solve_single_probe(n, probe, dx, k, dz)
    u0 = oms.solve_single_probe(np.ones((N, nz)), probe, dx, k, dz)
    for z in range(nz):
        assert np.allclose(u[:, z], u0[:, z] * np.exp(1j * k * dn * dz) ** z, atol=1e-12)
        # free-space propagation is unitary
        assert np.isclose(np.linalg.norm(u0[:, z]), np.linalg.norm(probe))


def test_batched_equals_single_and_backprop_inverts():
    N, nz, dx, dz, k = 96, 7, 0.7, 1.0, 2 * np.pi
    rng = np.random.default_rng(1)
    n = 1 - 1e-3 * rng.random((N, nz))
    P = rng.standard_normal((3, N)) + 0j
    b = oms.solve_batched(n, P, dx, k, dz)
    for p in range(3):
        assert rel_l2(b[p], oms.solve_single_probe(n, P[p], dx, k, dz)) < 1e-14
    back = oms.backpropagate(oms.backpropagate(b[0][:, 0], dx, k, -dz), dx, k, dz)
    assert rel_l2(back, b[0][:, 0]) < 1e-12


def test_farfield_parseval_and_delta():
    rng = np.random.default_rng(2)
    ex = rng.standard_normal((4, 50)) + 1j * rng.standard_normal((4, 50))
    I = oms.farfield_intensities(ex, 2, 4, 50)
    assert np.allclose(I.sum(axis=1), 50 * np.sum(np.abs(ex) ** 2, axis=1))
    d = np.zeros((1, 50), complex)
    d[0, 0] = 2.0
    assert np.allclose(oms.farfield_intensities(d, 2, 1, 50), 4.0)
    # the literal two-axis restatement transforms across the scan axis (quirk Q9)
    ex3 = rng.standard_normal((3, 16)) + 0j
    q9 = oms.farfield_intensities(ex3, 3, 3, 16)
    assert np.allclose(q9, np.abs(np.fft.fft2(ex3)) ** 2)
    pp = oms.farfield_per_probe_2d(ex3, 4, 4)
    assert np.allclose(pp[1], (np.abs(np.fft.fft2(ex3[1].reshape(4, 4))) ** 2).ravel())


def test_multislice_2d_plane_wave_eigen():
    """exp(i (kx x + ky y)) on the FFT grid is an eigenvector of the two-axis propagator
    with eigenvalue exp(i (kz - k) dz); a constant object multiplies by exp(i k (n - 1) dz)."""
    from oracle.multislice import solve_batched_2d

    nx, ny, nz, dx, dy, dz, k = 32, 16, 6, 0.7, 0.9, 1.3, 2 * np.pi
    mx, my = 5, 3
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    probe = np.exp(2j * np.pi * (mx * x / nx + my * y / ny))
    dn = 1e-3 + 2e-5j
    n = np.full((nx, ny, nz), 1.0 + dn)
    u = solve_batched_2d(n, probe[None], dx, dy, k, dz)[0]
    kx, ky = 2 * np.pi * mx / (nx * dx), 2 * np.pi * my / (ny * dy)
    lam = np.exp(1j * (np.sqrt(k**2 - kx**2 - ky**2) - k) * dz) * np.exp(1j * k * dn * dz)
    for z in range(nz):
        assert np.allclose(u[:, :, z], probe * lam**z, atol=1e-12)


def test_multislice_2d_unitary_in_free_space():
    from oracle.multislice import solve_batched_2d

    rng = np.random.default_rng(0)
    nx, ny, nz = 16, 32, 5
    probes = rng.standard_normal((2, nx, ny)) + 1j * rng.standard_normal((2, nx, ny))
    u = solve_batched_2d(np.ones((nx, ny, nz)), probes, 2.0, 2.0, 2 * np.pi, 4.0, keep="exit")
    assert np.allclose(np.linalg.norm(u.reshape(2, -1), axis=1), np.linalg.norm(probes.reshape(2, -1), axis=1))


@pytest.mark.parametrize("axes", [1, 2])
def test_ms_gradient_finite_differences(axes):
    """The multislice amplitude-loss gradient against central differences of the
    loss in FP64 (real and imaginary parts of n at seeded entries)."""
    rng = np.random.default_rng(5)
    k, dz = 2 * np.pi, 4.0
    if axes == 1:
        nx, nz, P = 48, 9, 3
        shape = (nx, nz)
        probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
        I = np.abs(rng.standard_normal((P, nx))) ** 2 * nx

        def run(n):
            return oms.ms_gradient(n, probes, 2.0, k, dz, I)
    else:
        nx, ny, nz, P = 16, 12, 7, 2
        shape = (nx, ny, nz)
        probes = rng.standard_normal((P, nx, ny)) + 1j * rng.standard_normal((P, nx, ny))
        I = np.abs(rng.standard_normal((P, nx, ny))) ** 2 * nx * ny

        def run(n):
            return oms.ms_gradient_2d(n, probes, 2.0, 1.5, k, dz, I)
    n = 1 - (rng.uniform(0, 1e-2, shape) + 1j * rng.uniform(0, 1e-3, shape))
    grad, _ = run(n)
    h = 1e-6
    for _ in range(12):
        idx = tuple(int(rng.integers(0, s)) for s in shape)
        for part in (1.0, 1j):
            e = np.zeros(shape, dtype=complex)
            e[idx] = h * part
            fd = (run(n + e)[1] - run(n - e)[1]) / (2 * h)
            an = grad[idx].real if part == 1.0 else grad[idx].imag
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (idx, part, fd, an)
    assert np.all(grad[..., -1] == 0)
