"""Pin the FP64 gradients (CPU).

``exact_gradient`` is checked against central finite differences of the loss
(h = 1e-6 in FP64) on random entries of n, real and imaginary parts.  The
reference-form gradient (``reference_gradient``) has no independent golden
values -- the reference suite never tests it and importing the reference was
not used.  It is pinned where the object is uniform along z (its adjoint then
coincides with the exact one, so it equals a closed form in the exact adjoint
states), besides the structural checks: its adjoint states equal the restated
reference adjoint sweep, and its error term is the exact loss seed scaled by
1/N (base_reconstructor.py:171-207).
"""

import numpy as np
import pytest

from oracle import gradient as og
from oracle import pwe


def _problem(seed=0, nx=24, nz=10, P=3):
    rng = np.random.default_rng(seed)
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=2 * np.pi)
    n_true = 1 - 5e-3 * (rng.random((nx, nz)) + 0.1j * rng.random((nx, nz)))
    x = np.arange(nx)
    probes = np.stack([np.exp(-((x - c) ** 2) / 16.0) for c in np.linspace(5, nx - 5, P)]).astype(complex)
    fwd = pwe.build_lu_cache(g, "dirichlet", n_true, "forward")
    I = np.abs(np.fft.fft(pwe.solve_batched(g, fwd, probes, keep="exit"), axis=-1)) ** 2
    n0 = 1 - 2e-3 * (rng.random((nx, nz)) + 0.1j * rng.random((nx, nz)))
    return g, n0, probes, I, rng


def _loss(g, n, probes, I):
    fwd = pwe.build_lu_cache(g, "dirichlet", n, "forward")
    return og.amplitude_loss(pwe.solve_batched(g, fwd, probes, keep="exit"), I)


def test_exact_gradient_matches_central_differences():
    g, n0, probes, I, rng = _problem()
    grad, J = og.exact_gradient(g, "dirichlet", n0, probes, I)
    assert np.isclose(J, _loss(g, n0, probes, I))
    h = 1e-6
    for _ in range(12):
        ix, iz = rng.integers(0, g.nx), rng.integers(0, g.nz)
        for direction, part in ((1.0, "re"), (1j, "im")):
            npl = n0.copy()
            nmi = n0.copy()
            npl[ix, iz] += h * direction
            nmi[ix, iz] -= h * direction
            fd = (_loss(g, npl, probes, I) - _loss(g, nmi, probes, I)) / (2 * h)
            an = grad[ix, iz].real if part == "re" else grad[ix, iz].imag
            assert abs(fd - an) <= 1e-5 * max(1e-12, np.max(np.abs(grad))) + 1e-4 * abs(fd), (ix, iz, part, fd, an)


def test_reference_gradient_structure():
    g, n0, probes, I, _ = _problem(seed=3)
    gr, gi = og.reference_gradient(g, "dirichlet", n0, probes, I)
    N, L = g.nx, g.nz - 1
    assert gr.shape == (N * L,) and gi.shape == (N * L,)
    # error term = (1/N) * exact loss seed
    fwd = pwe.build_lu_cache(g, "dirichlet", n0, "forward")
    ex = pwe.solve_batched(g, fwd, probes, keep="exit")
    e = ex - og.phase_retrieval_model(ex, I)
    assert np.allclose(e, og.amplitude_loss_seed(ex, I) / N)
    # the reference gradient equals the closed form with its own adjoint states;
    # the loop passes grad_Ak = -get_gradient(n), so g = -g0 (u_s + u_{s-1})
    adj = pwe.build_lu_cache(g, "dirichlet", n0, "adjoint")
    u = pwe.solve_batched(g, fwd, probes, "forward")
    g0 = -g.dz * g.k / 2j
    gr2 = np.zeros((L, N))
    gi2 = np.zeros((L, N))
    for p in range(probes.shape[0]):
        rb = np.zeros((N, L), complex)
        rb[:, -1] = e[p]
        lam = pwe.solve_single_probe(g, adj, None, "adjoint", rhs_block=rb)  # flipped: column k = v_{L-k}
        for s in range(L):  # block s pairs g_s with v_{L-s}
            gs = g0 * (u[p, :, s + 1] + (u[p, :, s] if s >= 1 else 0))
            gr2[s] -= (np.conj(gs) * lam[:, s]).real
            gi2[s] -= (np.conj(1j * gs) * lam[:, s]).real
    assert np.allclose(gr, gr2.ravel()) and np.allclose(gi, gi2.ravel())


def test_exact_gradient_3d_central_differences():
    """Two transverse axes: the oracle's exact gradient of the 2D far-field amplitude
    misfit against FP64 central differences (h = 1e-6) at random entries."""
    from oracle.gradient import amplitude_loss_2d, exact_gradient_3d

    rng = np.random.default_rng(8)
    nx, ny, nz = 8, 6, 5
    g = pwe.Grid(dimension=3, nx=nx, ny=ny, nz=nz, dx=2.0, dy=2.0, dz=4.0, k=2 * np.pi)
    n_true = 1 - (rng.uniform(0, 5e-3, (nx, ny, nz)) + 1j * rng.uniform(0, 5e-5, (nx, ny, nz)))
    n0 = 1 - (rng.uniform(0, 2e-3, (nx, ny, nz)) + 1j * rng.uniform(0, 2e-5, (nx, ny, nz)))
    probes = rng.standard_normal((3, nx * ny)) + 1j * rng.standard_normal((3, nx * ny))
    u = pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", n_true, "forward"), probes, keep="exit")
    I = np.abs(np.fft.fft2(u.reshape(3, nx, ny), axes=(-2, -1))) ** 2
    grad, J = exact_gradient_3d(g, "impedance", n0, probes, I)

    def loss(n):
        ex = pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", n, "forward"), probes, keep="exit")
        return amplitude_loss_2d(ex, I, nx, ny)

    assert abs(loss(n0) - J) <= 1e-12 * J
    h = 1e-6
    scale = np.abs(grad).max()
    for i in range(16):
        idx = (rng.integers(nx), rng.integers(ny), rng.integers(nz))
        d = 1.0 if i % 2 == 0 else 1j
        npl, nmi = n0.copy(), n0.copy()
        npl[idx] += h * d
        nmi[idx] -= h * d
        fd = (loss(npl) - loss(nmi)) / (2 * h)
        an = grad[idx].real if i % 2 == 0 else grad[idx].imag
        assert abs(fd - an) <= 1e-5 * scale + 1e-4 * abs(fd), (idx, fd, an)


def test_simulated_data_file_round_trip(tmp_path):
    """The reference's simulated_data.npz keys (generate_data.py:193-197)."""
    from thick_ptycho_amd import reconstruct

    rng = np.random.default_rng(0)
    I = rng.random((4, 16))
    ex = rng.standard_normal((4, 16)) + 1j * rng.standard_normal((4, 16))
    path = str(tmp_path / "simulated_data.npz")
    reconstruct.save_simulated_data(path, I, ex)
    with np.load(path, allow_pickle=False) as f:
        assert sorted(f.files) == ["exit_waves", "intensity"]
    back = reconstruct.load_simulated_data(path)
    assert np.array_equal(back["intensity"], I) and np.array_equal(back["exit_waves"], ex)


@pytest.mark.parametrize("g_new, g_old, beta", [
    ((0.5, 0.0), (1.0, 0.0), 0.0),   # PR = -0.25 clipped at zero
    ((1.0, 1.0), (1.0, 0.0), 1.0),   # PR = 1 below FR = 2
    ((2.0, 0.0), (1.0, 0.0), 2.0),   # PR = 2 below FR = 4
    ((-1.0, 0.0), (1.0, 0.0), 1.0),  # PR = 2 capped by FR = 1
    ((1.0, 0.0), (0.0, 0.0), 0.0),   # zero previous gradient
])
def test_nlcg_beta_rule(g_new, g_old, beta):
    """min(FR, max(PR, 0)) as compute_betak (pwe_reconstructor.py:221-228)."""
    import torch

    from thick_ptycho_amd import reconstruct

    got = reconstruct.beta_rule(torch.tensor(g_new, dtype=torch.float64), torch.tensor(g_old, dtype=torch.float64))
    assert got == pytest.approx(beta)


def test_reference_gradient_equals_exact_adjoint_form_when_uniform_in_z(rng):
    """Pin of the restated compute_grad_least_squares: when the object does not
    vary along z the reference-form adjoint (quirk Q3: slice L+1-j's object term
    on both sides) coincides with the exact adjoint, so the reference gradient
    (grad_Ak = -get_gradient(n), pwe_reconstructor.py:257) must equal
    conj(eo) sum_p conj(u_s + u_{s-1}) lambda_s / N with lambda from
    the exact adjoint (the path pinned by finite differences above), eo =
    (dz/2)(k/i), the u_0 term absent at s = 1 and the 1/N of its exit-wave
    error.  A wrong flip, sign, conjugate or seed scale fails it; with an object
    varying along z the two differ (the quirk itself)."""
    nx, nz, P = 40, 9, 3
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=2 * np.pi)
    prof = 1 - (5e-3 + 5e-5j) * np.exp(-((np.arange(nx) - 20) / 6.0) ** 2)
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    I = np.abs(rng.standard_normal((P, nx))) ** 2 * nx
    eo = (g.dz / 2) * (g.k / 1j)

    def both(n):
        gr, gi = og.reference_gradient(g, "impedance", n, probes, I)
        ref = (gr + 1j * gi).reshape(nz - 1, nx).T
        u = pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", n, "forward"), probes, "forward")
        lam = og.exact_adjoint_states(g, "impedance", n, og.amplitude_loss_seed(u[:, :, -1], I))
        pred = np.stack([np.conj(eo) * np.sum(np.conj(u[:, :, s] + (u[:, :, s - 1] if s > 1 else 0))
                                                * lam[:, :, s], axis=0) / nx for s in range(1, nz)], axis=1)
        return ref, pred

    ref, pred = both(np.repeat(prof[:, None], nz, axis=1))
    assert np.linalg.norm(ref - pred) <= 1e-12 * np.linalg.norm(pred)
    zvar = np.repeat(prof[:, None], nz, axis=1) * (1 + 2e-3 * np.arange(nz))[None, :]
    ref, pred = both(zvar)
    assert np.linalg.norm(ref - pred) > 1e-6 * np.linalg.norm(pred)


def test_exit_wave_loss_gradient_finite_differences():
    """exact_gradient with exit_data: J = sum |u_L - d|^2, central differences at
    random entries (real and imaginary parts of n)."""
    nx, nz, P = 24, 7, 2
    rng = np.random.default_rng(5)
    g = pwe.Grid(dimension=2, nx=nx, nz=nz, dx=2.0, dz=4.0, k=2 * np.pi)
    n = 1 - (rng.uniform(0, 1e-2, (nx, nz)) + 1j * rng.uniform(0, 1e-4, (nx, nz)))
    probes = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    d = rng.standard_normal((P, nx)) + 1j * rng.standard_normal((P, nx))
    grad, J = og.exact_gradient(g, "impedance", n, probes, None, exit_data=d)

    def loss(m):
        u = pwe.solve_batched(g, pwe.build_lu_cache(g, "impedance", m, "forward"), probes, keep="exit")
        return float(np.sum(np.abs(u - d) ** 2))

    assert J == pytest.approx(loss(n), rel=1e-12)
    h = 1e-6
    for i, j in [(3, 2), (12, 5), (20, 0), (7, 6)]:
        for unit, part in ((1.0, np.real), (1j, np.imag)):
            e = np.zeros_like(n)
            e[i, j] = unit * h
            fd = (loss(n + e) - loss(n - e)) / (2 * h)
            assert fd == pytest.approx(part(grad[i, j]), rel=1e-5, abs=1e-9 * abs(J))
