"""Pins the CPU restatement of the learned preconditioner and BiCGSTAB (oracle/learned.py)
against library routines and finite differences."""

import numpy as np
import pytest
import scipy.signal
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import learned as ol
from thick_ptycho_amd.forward_model import band_engine as be


def test_conv3x3_matches_correlate2d():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 7, 3))
    W = rng.standard_normal((9, 4, 3))
    got = ol.conv3x3(x, W)
    for p in range(2):
        for o in range(4):
            ref = sum(scipy.signal.correlate2d(x[p, :, :, i], W[:, o, i].reshape(3, 3), mode="same")
                      for i in range(3))
            assert np.allclose(got[p, :, :, o], ref, atol=1e-12)


def test_structured_network_is_complex_convolution():
    """The real network with [[a,-b],[b,a]] blocks equals the complex network."""
    rng = np.random.default_rng(1)
    params = ol.identity_params(rng, 0.3)
    nx, ny = 9, 6
    v = rng.standard_normal((2, ny * nx)) + 1j * rng.standard_normal((2, ny * nx))
    got, _, s = ol.forward(params, v, nx, ny, bf16_round=False)
    x = (v / s[:, None]).reshape(2, ny, nx, 1)
    for l in range(4):
        w = ol.layer_params(params, l)
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        y = np.zeros(x.shape[:3] + (w.shape[1],), dtype=complex)
        for tap in range(9):
            dy, dx = divmod(tap, 3)
            y += np.einsum("pyxi,oi->pyxo", xp[:, dy:dy + ny, dx:dx + nx, :], w[tap])
        x = y
    ref = x[..., 0].reshape(2, -1) * s[:, None]
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_loss_gradient_central_differences():
    rng = np.random.default_rng(2)
    params = ol.identity_params(rng, 0.1)
    nx, ny = 8, 5
    v = rng.standard_normal((3, ny * nx)) + 1j * rng.standard_normal((3, ny * nx))
    tgt = rng.standard_normal((3, ny * nx)) + 1j * rng.standard_normal((3, ny * nx))
    _, g = ol.loss_and_grad(params, v, tgt, nx, ny, bf16_round=False)
    h = 1e-6
    for q in rng.choice(ol.NPARAM, 16, replace=False):
        for part in (1.0, 1j):
            pp, pm = params.copy(), params.copy()
            pp[q] += h * part
            pm[q] -= h * part
            fd = (ol.loss_and_grad(pp, v, tgt, nx, ny, False)[0] - ol.loss_and_grad(pm, v, tgt, nx, ny, False)[0]) / (2 * h)
            an = g[q].real if part == 1.0 else g[q].imag
            assert abs(fd - an) <= 1e-6 * max(1.0, np.abs(g).max()), (q, part, fd, an)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], dtype=np.float32)
    got = ol.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0  # halfway, rounds to even mantissa
    assert got[2] == 1.0 + 2 ** -7
    assert got[3] == -2.5
    assert abs(got[4] - 3.140625) < 1e-12


def _operator(nx=12, ny=10, seed=3):
    rng = np.random.default_rng(seed)
    A, _ = be.step_operators_3d(nx, ny, 2 * np.pi, 2.0, 2.0, 4.0, "impedance")
    c = 1e-2 * (rng.standard_normal(nx * ny) + 1j * rng.standard_normal(nx * ny))
    return (A.to_sparse() - sp.diags(c)).tocsr(), rng


@pytest.mark.parametrize("precond", [None, "jacobi", "network"])
def test_bicgstab_solves(precond):
    M, rng = _operator()
    b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    if precond == "jacobi":
        Mf = lambda z: z / M.diagonal()
    elif precond == "network":
        params = ol.identity_params(rng, 0.02)
        Mf = lambda z: ol.forward(params, z[None], 12, 10, bf16_round=False)[0][0]
    else:
        Mf = None
    x, it, hist = ol.bicgstab(M, b, b.copy(), Mf, tol=1e-10, maxiter=200)
    assert it > 0
    ref = spla.spsolve(M.tocsc(), b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-8


def test_bicgstab_residual_history_matches_scipy():
    M, rng = _operator(seed=4)
    b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    x0 = np.zeros_like(b)
    _, _, hist = ol.bicgstab(M, b, x0, None, tol=1e-12, maxiter=6)
    seen = []
    spla.bicgstab(M, b, x0=x0, rtol=1e-12, maxiter=6,
                  callback=lambda xk: seen.append(np.linalg.norm(b - M @ xk) / np.linalg.norm(b)))
    n = min(len(seen), len(hist) - 1, 5)
    assert n >= 3
    assert np.allclose(hist[1:n + 1], seen[:n], rtol=1e-6)
