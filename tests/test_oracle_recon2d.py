"""Pin the FP64 restatement of the reference's NLCG step and object filters (CPU).

* gaussian_lowpass_phase: a cosine mode of integer frequency (my, mx) is an
  eigenfunction with eigenvalue (w(my, mx) + w(-my, -mx)) / 2, w the weight of
  frequency bin (ky, kx): exp(-(q_x^2 + q_y^2) / (2 sigma^2)) with
  q(k) = -1/2 + ((k - n//2) mod n) / n, i.e. fftshift of the linspace grid
  (exp(-((mx/nx)^2 + (my/ny)^2) / (2 sigma^2)) for even sizes; at odd sizes
  the grid has no zero and even the DC term is damped -- the reference's own
  filter); a transposed axis or a wrong shift changes it; a wide sigma leaves
  any input unchanged.
* tv_denoise_amp: a constant array is a fixed point; one iteration on a delta
  matches the closed form of the dual step (the divergence of the first
  dual field); a small lam changes the input by O(lam^2) (the dual field
  is O(lam)); it lowers the total variation of a noisy piecewise-constant input.
* compute_alphak: the forward solve with a right-hand side equals the
  whole-system SuperLU solve of the same block system (probe 0), the
  perturbation grad_A u equals its closed form -g0 (u_j + u_{j-1}), and
  alpha scales as 1/c for d -> c d.
"""

import numpy as np
import pytest

from oracle import gradient as og
from oracle import pwe
from oracle import recon2d as orc


def _cos_mode(ny, nx, my, mx):
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    return np.cos(2 * np.pi * (my * y / ny + mx * x / nx))


def _bin_weight(ky, kx, ny, nx, sigma):
    qy = -0.5 + ((ky - ny // 2) % ny) / ny
    qx = -0.5 + ((kx - nx // 2) % nx) / nx
    return np.exp(-(qx**2 + qy**2) / (2 * sigma**2))


@pytest.mark.parametrize("ny,nx", [(40, 24), (33, 17)])
def test_lowpass_eigenmodes(ny, nx):
    sigma = 0.07
    dc = _bin_weight(0, 0, ny, nx, sigma)
    assert np.allclose(orc.gaussian_lowpass_phase(np.full((ny, nx), 0.3), sigma), 0.3 * dc, atol=1e-14)
    if ny % 2 == 0 and nx % 2 == 0:
        assert dc == 1.0
    for my, mx in [(1, 0), (0, 2), (3, 5), (-2, 4)]:
        f = _cos_mode(ny, nx, my, mx)
        lam = 0.5 * (_bin_weight(my, mx, ny, nx, sigma) + _bin_weight(-my, -mx, ny, nx, sigma))
        if ny % 2 == 0 and nx % 2 == 0:
            assert lam == pytest.approx(np.exp(-((mx / nx) ** 2 + (my / ny) ** 2) / (2 * sigma**2)))
        assert np.allclose(orc.gaussian_lowpass_phase(f, sigma), lam * f, atol=1e-12)
    r = np.random.default_rng(0).standard_normal((ny, nx))
    assert np.allclose(orc.gaussian_lowpass_phase(r, 1e4), r, atol=1e-8)


def test_tv_fixed_point_and_single_step():
    c = np.full((12, 9), 0.8)
    assert np.allclose(orc.tv_denoise_amp(c, lam=0.05), c, atol=1e-15)
    # one iteration from p = 0 on a delta of height h at (i, j): uxy = u, the dual
    # field is nonzero where a forward difference sees the delta, and the output
    # is u - lam div(p) with div(p)[y, x] = px[y, x-1] - px[y, x] + py[y-1, x] - py[y, x]
    lam, eps, h = 0.1, 1e-8, 2.0
    u = np.zeros((7, 8))
    i, j = 3, 4
    u[i, j] = h
    out = orc.tv_denoise_amp(u, lam=lam, n_iter=1, eps=eps)
    px = np.zeros_like(u)
    py = np.zeros_like(u)
    for (y, x), (dx, dy) in {(i, j): (-h, -h), (i, j - 1): (h, 0.0), (i - 1, j): (0.0, h)}.items():
        den = 1 + np.sqrt(dx * dx + dy * dy + eps) / lam
        px[y, x] = 0.25 * dx / den
        py[y, x] = 0.25 * dy / den
    exp = u.copy()
    for y in range(u.shape[0]):
        for x in range(u.shape[1]):
            div = px[y, x - 1] - px[y, x] + py[y - 1, x] - py[y, x]
            exp[y, x] -= lam * div
    assert np.allclose(out, exp, atol=1e-14)


def test_tv_small_lambda_and_variation():
    rng = np.random.default_rng(3)
    u = np.kron(rng.random((4, 3)), np.ones((6, 8))) + 0.05 * rng.standard_normal((24, 24))
    small = orc.tv_denoise_amp(u, lam=1e-4)
    assert np.max(np.abs(small - u)) < 1e-7

    def tv(a):
        return np.sum(np.abs(np.diff(a, axis=0))) + np.sum(np.abs(np.diff(a, axis=1)))

    assert tv(orc.tv_denoise_amp(u, lam=0.2)) < 0.8 * tv(u)


def _case(seed=0, nx=30, nz=8, P=3):
    rng = np.random.default_rng(seed)
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=2 * np.pi)
    n = 1 - 5e-3 * (rng.random((nx, nz)) + 0.1j * rng.random((nx, nz)))
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    return g, n, probes, rng


def test_forward_with_rhs_equals_whole_system_solve():
    g, n, probes, rng = _case()
    N, L = g.nx, g.nz - 1
    R = rng.standard_normal((N, L)) + 1j * rng.standard_normal((N, L))
    fwd = pwe.build_lu_cache(g, "dirichlet", n, "forward")
    step = pwe.solve_single_probe(g, fwd, np.zeros(N), "forward", rhs_block=R)
    whole = pwe.full_lu_solve(g, "dirichlet", n, np.zeros(N), rhs_block=R.T.ravel())
    assert np.allclose(step, whole, rtol=0, atol=1e-10 * np.abs(whole).max())


def test_perturbation_closed_form_and_alpha_scaling():
    g, n, probes, rng = _case(seed=2)
    N, L = g.nx, g.nz - 1
    fwd = pwe.build_lu_cache(g, "dirichlet", n, "forward")
    u = pwe.solve_batched(g, fwd, probes, "forward")
    uk = og.to_vector_form(u)
    grad_A = -og.gradient_operator(g, n)
    g0 = g.dz * g.k / 2j
    for p in range(probes.shape[0]):
        closed = np.stack([-g0 * (u[p, :, j] + (u[p, :, j - 1] if j > 1 else 0)) for j in range(1, L + 1)])
        assert np.allclose(grad_A @ uk[p], closed.ravel(), atol=1e-12)
    grad = rng.standard_normal(N * L)
    d = rng.standard_normal(N * L)
    a1 = orc.reference_alphak(g, fwd, probes, uk, grad_A, grad, d)
    a2 = orc.reference_alphak(g, fwd, probes, uk, grad_A, grad, 3.0 * d)
    assert a1 != 0 and a2 == pytest.approx(a1 / 3.0, rel=1e-10)


def test_reference_loop_runs_and_records():
    g, n, probes, rng = _case(seed=4, nx=40, nz=10, P=4)
    fwd = pwe.build_lu_cache(g, "dirichlet", n, "forward")
    I = np.abs(np.fft.fft(pwe.solve_batched(g, fwd, probes, keep="exit"), axis=-1)) ** 2
    n0 = np.ones_like(n)
    nk, res, steps = orc.reference_reconstruct(g, "dirichlet", n0, probes, I, 3, tv_lambda_amp=1e-3,
                                               low_pass_sigma_phase=0.2)
    assert len(res) == 3 and len(steps) == 3 and np.all(np.isfinite(nk))
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, "dirichlet", n0, "forward"), probes, keep="exit")
    e = u - og.phase_retrieval_model(u, I)
    assert res[0] == pytest.approx(np.sqrt(np.mean(np.abs(e) ** 2)), rel=1e-12)
