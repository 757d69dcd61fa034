"""The allocation-free TV entry point (tp_tv_denoise_noalloc) against tp_tv_denoise.

Both run the same kernels; the allocation-free form takes its dual field and
work plane from the library's per-GPU device array instead of a caller buffer.
Two calls in a row, the second on a smaller plane (the array keeps the first
call's contents, so the second call checks that it clears what must start at
zero), are compared bitwise with tp_tv_denoise on caller-allocated work, and the
second also against the FP64 oracle's restatement of the reference's
tv_denoise_amp.  A plane above the array's size is refused before any launch.
"""

import numpy as np
import pytest
import torch

from oracle import recon2d as orc
from tests.cases import rel_l2
from thick_ptycho_amd import _lib

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21  # TP_TV_NOALLOC_MAX_POINTS


def _plane(ny, nx, seed):
    rng = np.random.default_rng(seed)
    n = 1 - (rng.uniform(0, 1e-2, (ny, nx)) + 1j * rng.uniform(0, 1e-3, (ny, nx)))
    return np.abs(n)


def _with_work(A, lam, n_iter, eps):
    lib = _lib.load()
    u = torch.from_numpy(A).cuda().contiguous()
    ny, nx = A.shape
    work = torch.empty(3 * ny * nx, dtype=torch.float64, device="cuda")
    _lib.check(lib.tp_tv_denoise(_lib.ptr(u), ny, nx, lam, n_iter, eps, _lib.ptr(work), _lib.stream_handle()),
               "tp_tv_denoise")
    return u.cpu().numpy()


def _no_alloc(A, lam, n_iter, eps):
    lib = _lib.load()
    u = torch.from_numpy(A).cuda().contiguous()
    ny, nx = A.shape
    _lib.check(lib.tp_tv_denoise_noalloc(_lib.ptr(u), ny, nx, lam, n_iter, eps, _lib.stream_handle()),
               "tp_tv_denoise_noalloc")
    return u.cpu().numpy()


def test_no_alloc_matches_work_buffer_form(gpu):
    lam, n_iter, eps = 2e-3, 30, 1e-12
    # a large plane first, then a smaller one that reuses the stale array
    for ny, nx, seed in [(1024, 600, 1), (257, 130, 2)]:
        print(f"tp_tv_denoise_noalloc {ny}x{nx} seed {seed}", flush=True)
        A = _plane(ny, nx, seed)
        got = _no_alloc(A, lam, n_iter, eps)
        ref = _with_work(A, lam, n_iter, eps)
        assert np.array_equal(got, ref), f"{ny}x{nx}: max diff {np.abs(got - ref).max():.3e}"
    oracle = orc.tv_denoise_amp(A, lam=lam, n_iter=n_iter, eps=eps)
    assert rel_l2(got - A, oracle - A) < 1e-10


def test_no_alloc_at_the_size_limit(gpu):
    ny, nx = 2048, MAX_POINTS // 2048  # exactly the limit
    A = _plane(ny, nx, 3)
    got = _no_alloc(A, 1e-3, 5, 1e-12)
    ref = _with_work(A, 1e-3, 5, 1e-12)
    assert np.array_equal(got, ref)


def test_no_alloc_refuses_planes_above_the_limit(gpu):
    lib = _lib.load()
    u = torch.zeros(8, dtype=torch.float64, device="cuda")  # never touched: the size check comes first
    rc = lib.tp_tv_denoise_noalloc(_lib.ptr(u), 2048, MAX_POINTS // 2048 + 1, 1e-3, 5, 1e-12, _lib.stream_handle())
    assert rc == _lib.TP_ERR_INVALID
    assert b"exceed" in lib.tp_last_error()
    with pytest.raises(ValueError, match="exceed"):
        _lib.check(rc, "tp_tv_denoise_noalloc")
