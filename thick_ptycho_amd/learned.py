"""Learned-preconditioned Krylov stepping for 3D objects (two transverse axes).

Each CN step u_s = 2 A_s^{-1} u_{s-1} - u_{s-1} solves A_s x = u_{s-1} for all
probes with BiCGSTAB (``tp_bicgstab_solve``), preconditioned by nothing, by
Jacobi or by ``CNNPreconditioner``: a complex-linear 4-layer 3x3 CNN (BF16
weights on MFMA, FP32 accumulation) trained on the GPU on residual/correction
pairs (r, A_s^{-1} r) produced by the banded LU solver.  Probes whose
iteration breaks down or does not reach the tolerance are re-solved on that
slice with the banded LU (``tp_band_factor`` / ``tp_band_sweep``), which also
serves as the reference for the Krylov results in the tests.

Nothing here has a counterpart in the reference (its 3D solver factors every
slice with SuperLU, iterative_lu.py:66-197); it is an alternative solver for
the same step, selected by the caller.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .forward_model import band_engine as be

LAYERS = ((16, 1), (16, 16), (16, 16), (1, 16))  # (out, in) complex channels per layer


def _dev(device):
    return device or torch.device("cuda", torch.cuda.current_device())


class CNNPreconditioner:
    """Complex parameters (complex64 [count] on the GPU) and their BF16 expansion."""

    def __init__(self, device=None, seed: int = 0, noise: float = 1e-2, params: Optional[np.ndarray] = None):
        self.device = _dev(device)
        lib = _lib.load()
        self.count = int(lib.tp_cnn_param_count())
        if params is None:
            params = self.identity_params(np.random.default_rng(seed), noise)
        self.params = torch.tensor(np.asarray(params, dtype=np.complex64), device=self.device)
        self.wbuf = torch.zeros(int(lib.tp_cnn_weight_bytes()) // 2, dtype=torch.bfloat16, device=self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.history: List[float] = []
        self.expand()

    @staticmethod
    def identity_params(rng: np.random.Generator, noise: float) -> np.ndarray:
        """Centre tap passing complex channel 0 through every layer (the identity map) plus noise."""
        parts = []
        for oc, ic in LAYERS:
            w = noise * (rng.standard_normal((9, oc, ic)) + 1j * rng.standard_normal((9, oc, ic)))
            w[4, 0, 0] += 1.0
            parts.append(w.reshape(-1))
        return np.concatenate(parts)

    def expand(self) -> None:
        _lib.check(_lib.load().tp_cnn_expand(_lib.ptr(self.params), _lib.ptr(self.wbuf), _lib.stream_handle()),
                   "tp_cnn_expand")

    def _work(self, P, nx, ny, train):
        nb = int(_lib.load().tp_cnn_workspace_bytes(P, nx, ny, int(train)))
        return torch.empty(nb, dtype=torch.uint8, device=self.device)

    def apply(self, v: torch.Tensor, nx: int, ny: int) -> torch.Tensor:
        """M(v) for complex64 planes (P, ny*nx)."""
        v = v.contiguous()
        P = v.shape[0]
        out = torch.empty_like(v)
        work = self._work(P, nx, ny, False)
        _lib.check(_lib.load().tp_cnn_apply(_lib.ptr(self.wbuf), _lib.ptr(v), _lib.ptr(out), P, nx, ny,
                                            _lib.ptr(work), _lib.stream_handle()), "tp_cnn_apply")
        return out

    def _train_call(self, r, e, nx, ny, step, lr, grad=None):
        P = r.shape[0]
        work = self._work(P, nx, ny, True)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        _lib.check(
            _lib.load().tp_cnn_train_step(_lib.ptr(self.params), _lib.ptr(self.wbuf), _lib.ptr(self.m),
                                          _lib.ptr(self.v), _lib.ptr(r), _lib.ptr(e), P, nx, ny, step, lr, 0.9,
                                          0.999, 1e-8, _lib.ptr(work), _lib.ptr(loss), _lib.ptr(grad),
                                          _lib.stream_handle()),
            "tp_cnn_train_step",
        )
        return loss

    def gradient(self, r: torch.Tensor, e: torch.Tensor, nx: int, ny: int):
        """(loss, complex gradient dL/da + i dL/db) without updating."""
        grad = torch.empty_like(self.params)
        loss = self._train_call(r.contiguous(), e.contiguous(), nx, ny, 0, 0.0, grad)
        return float(loss.item()), grad

    def train_step(self, r: torch.Tensor, e: torch.Tensor, nx: int, ny: int, lr: float = 1e-3) -> torch.Tensor:
        """One Adam step on the batch (r, e); returns the loss before the step (device scalar)."""
        self.step += 1
        return self._train_call(r.contiguous(), e.contiguous(), nx, ny, self.step, lr)

    def train(self, pairs: "PairSet", steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0,
              log_every: int = 50) -> List[float]:
        g = torch.Generator(device="cpu").manual_seed(seed)
        losses = []
        for k in range(steps):
            idx = torch.randint(0, pairs.count, (batch,), generator=g).to(self.device)
            l = self.train_step(pairs.r[idx], pairs.e[idx], pairs.nx, pairs.ny, lr)
            if (k + 1) % log_every == 0 or k + 1 == steps:
                losses.append(float(l.item()))
        self.history += losses
        return losses

    def state(self) -> np.ndarray:
        return self.params.cpu().numpy()


@dataclass
class PairSet:
    r: torch.Tensor  # complex64 (count, N) residual planes
    e: torch.Tensor  # complex64 (count, N) corrections A^{-1} r
    nx: int
    ny: int

    @property
    def count(self) -> int:
        return self.r.shape[0]


def generate_pairs(objects, k: float, dx: float, dy: float, dz: float, bc: str, per_slice: int = 8,
                   slices: Optional[int] = None, seed: int = 31, device=None, solver: str = "band",
                   tol: float = 1e-6) -> PairSet:
    """Residual/correction pairs from the solver itself: for each object (nx, ny, nz)
    and each of ``slices`` slices (evenly spaced), ``per_slice`` random complex planes r
    and e = A_s^{-1} r.  ``solver="band"``: one banded-LU step (u_1 = 2 A^{-1} r - r,
    e = (u_1 + r)/2); ``solver="krylov"``: unpreconditioned BiCGSTAB to ``tol``, for
    steps whose inverse blocks decay too slowly for the band (e.g. dx = 0.125 at dz = 4)."""
    if solver not in ("band", "krylov"):
        raise ValueError("solver must be 'band' or 'krylov'")
    dev = _dev(device)
    rng = np.random.default_rng(seed)
    rs, es = [], []
    nx = ny = None
    for n in objects:
        nx, ny, nz = n.shape
        N = nx * ny
        L = nz - 1
        picks = np.linspace(0, L - 1, slices or L).round().astype(int)
        if solver == "band":
            A, B = be.step_operators_3d(nx, ny, k, dx, dy, dz, bc)
            C = _contribution(n, k, dz, dev)
            fac = be.factor(C, A, B, "forward")
        else:
            ks = KrylovStepper(n, k, dx, dy, dz, bc, precond="none", tol=tol, max_iter=2000, device=dev)
        for s0 in picks:
            r = rng.standard_normal((per_slice, N)) + 1j * rng.standard_normal((per_slice, N))
            r = torch.tensor(r.astype(np.complex64), device=dev)
            if solver == "band":
                ex = torch.empty_like(r)
                be.sweep([be.BandScanSpec(_slice_factors(fac, int(s0)), r, exit=ex)])
                e = 0.5 * (ex + r)
            else:
                e = torch.zeros_like(r)
                done = ks.solve_slice(int(s0) + 1, r, e).cpu().numpy()
                if (done <= 0).any():
                    raise RuntimeError(f"BiCGSTAB did not reach {tol:g} on slice {int(s0) + 1} for training pairs")
            rs.append(r)
            es.append(e)
    return PairSet(torch.cat(rs), torch.cat(es), nx, ny)


def _contribution(n: np.ndarray, k: float, dz: float, dev) -> torch.Tensor:
    nx, ny, nz = n.shape
    nt = torch.tensor(np.asarray(n).reshape(nx * ny, nz), dtype=torch.complex128, device=dev)
    coef = (k / 1j) * (nt - 1)
    return ((dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2).contiguous()


def _slice_factors(fac: be.BandFactors, s0: int, count: int = 1) -> be.BandFactors:
    """View of the factors of steps s0+1 .. s0+count."""
    per = fac.W.numel() // fac.steps
    return be.BandFactors(fac.W[s0 * per:(s0 + count) * per], fac.cpl, None, fac.nx, fac.ny, count, fac.w,
                          fac.mode, fac.b_sign, fac.edge, keep=[fac])


PRECONDITIONERS = {"none": 0, "jacobi": 1, "cnn": 2}


@dataclass
class KrylovStats:
    iterations: List[float] = field(default_factory=list)  # mean over probes, per slice
    max_iterations: List[int] = field(default_factory=list)
    fallback_probes: int = 0
    fallback_slices: int = 0
    seconds: float = 0.0


class KrylovStepper:
    """Forward CN stepping of one 3D object with BiCGSTAB per slice."""

    def __init__(self, n: np.ndarray, k: float, dx: float, dy: float, dz: float, bc: str, precond: str = "cnn",
                 cnn: Optional[CNNPreconditioner] = None, tol: float = 1e-6, max_iter: int = 40, device=None):
        self.dev = _dev(device)
        nx, ny, nz = np.asarray(n).shape
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny
        self.k, self.dz = k, dz
        self.A, self.B = be.step_operators_3d(nx, ny, k, dx, dy, dz, bc)
        self.C = _contribution(n, k, dz, self.dev)
        self.tri64 = torch.tensor(self.A.tri, dtype=torch.complex128, device=self.dev)
        self.tri32 = self.tri64.to(torch.complex64)
        self.cpl32 = torch.tensor(np.stack([self.A.f, self.A.e], axis=1).astype(np.complex64), device=self.dev)
        if precond not in PRECONDITIONERS:
            raise ValueError(f"precond must be one of {sorted(PRECONDITIONERS)}")
        if precond == "cnn" and cnn is None:
            raise ValueError("precond='cnn' needs a CNNPreconditioner")
        self.precond, self.cnn, self.tol, self.max_iter = precond, cnn, tol, max_iter
        self._band = None
        self._bufs = {}

    def _band_slice(self, s: int) -> be.BandFactors:
        """Banded factors of step s alone (the fallback), factored on first use."""
        if self._band is None:
            self._band = {}
        if s not in self._band:
            col = self.C[:, s - 1:s].contiguous()
            self._band[s] = be.factor(col, self.A, self.B, "forward")
        return self._band[s]

    def _buffers(self, P: int, mi: int):
        """Per-(probe count, iteration cap) buffers, allocated once (graph capture replays them)."""
        key = (P, mi)
        if key not in self._bufs:
            lib = _lib.load()
            pc = PRECONDITIONERS[self.precond]
            self._bufs[key] = (
                torch.empty(self.N, dtype=torch.complex64, device=self.dev),
                torch.empty(int(lib.tp_bicgstab_workspace_bytes(P, self.nx, self.ny, max(mi, 1), pc)),
                            dtype=torch.uint8, device=self.dev),
                torch.zeros(P, dtype=torch.int32, device=self.dev),
            )
        return self._bufs[key]

    def solve_slice(self, s: int, b: torch.Tensor, x: torch.Tensor, max_iter: Optional[int] = None,
                    check_every: int = 1) -> torch.Tensor:
        """BiCGSTAB for A_s x = b (x holds the initial guess); returns done flags (int [P],
        the iterations taken, or <= 0 where the probe did not converge), a buffer reused by
        the next call.  ``check_every`` = 0 never reads the flags on the host: all
        ``max_iter`` iterations are launched (converged probes skip their work), so the
        call can be captured into a HIP graph."""
        lib = _lib.load()
        P = b.shape[0]
        mi = self.max_iter if max_iter is None else max_iter
        diag, work, done = self._buffers(P, mi)
        sh = _lib.stream_handle()
        _lib.check(lib.tp_krylov_diag(_lib.ptr(self.C), self.nx, self.ny, self.nz - 1, _lib.ptr(self.tri64), -1.0, 0,
                                      s - 1, _lib.ptr(diag), sh), "tp_krylov_diag")
        pc = PRECONDITIONERS[self.precond]
        _lib.check(
            lib.tp_bicgstab_solve(_lib.ptr(diag), _lib.ptr(self.tri32), _lib.ptr(self.cpl32), self.nx, self.ny, P,
                                  _lib.ptr(b), _lib.ptr(x), pc, _lib.ptr(self.cnn.wbuf) if pc == 2 else 0, self.tol,
                                  mi, check_every, _lib.ptr(work), _lib.ptr(done), sh),
            "tp_bicgstab_solve",
        )
        return done

    def sweep(self, u0: torch.Tensor, keep: str = "exit", max_iter: Optional[int] = None,
              fallback: bool = True) -> tuple:
        """Forward sweep of the probes u0 (complex64 (P, N)); returns (planes, KrylovStats) with
        planes (P, N) for keep='exit' or (P, nz, N) for keep='all'."""
        lib = _lib.load()
        P = u0.shape[0]
        stats = KrylovStats()
        t0 = time.perf_counter()
        u = u0.contiguous().clone()
        full = None
        if keep == "all":
            full = torch.empty((P, self.nz, self.N), dtype=torch.complex64, device=self.dev)
            full[:, 0] = u
        x = torch.empty_like(u)
        for s in range(1, self.nz):
            x.copy_(u)
            done = self.solve_slice(s, u, x, max_iter)
            d = done.cpu().numpy().copy()
            bad = np.nonzero(d <= 0)[0]
            good = d[d > 0]
            stats.iterations.append(float(good.mean()) if good.size else float("nan"))
            stats.max_iterations.append(int(good.max()) if good.size else 0)
            un = torch.empty_like(u)
            _lib.check(lib.tp_cn_finish(P * self.N, _lib.ptr(x), _lib.ptr(u), _lib.ptr(un), 2.0,
                                        _lib.stream_handle()), "tp_cn_finish")
            if bad.size:
                if not fallback:
                    raise RuntimeError(f"BiCGSTAB did not converge on slice {s} for {bad.size} probes")
                stats.fallback_probes += int(bad.size)
                stats.fallback_slices += 1
                idx = torch.tensor(bad, device=self.dev)
                ub = u[idx].contiguous()
                eb = torch.empty_like(ub)
                be.sweep([be.BandScanSpec(self._band_slice(s), ub, exit=eb)])
                un[idx] = eb
            u = un
            if full is not None:
                full[:, s] = u
        torch.cuda.synchronize()
        stats.seconds = time.perf_counter() - t0
        return (full if full is not None else u), stats

    def sweep_device(self, u0: torch.Tensor, out: torch.Tensor, done_log: torch.Tensor,
                     max_iter: Optional[int] = None) -> None:
        """Forward sweep without host synchronisation (capturable into a HIP graph):
        every slice runs ``max_iter`` BiCGSTAB iterations with the convergence flags kept
        on the device (converged probes skip their work, the CNN included), the planes go
        to ``out`` (P, nz, N) complex64 and each slice's flags to ``done_log`` (nz - 1, P)
        int32.  No banded-LU fallback: the caller reads ``done_log`` afterwards (a probe
        with a flag <= 0 did not converge on that slice) and recomputes with ``sweep``."""
        lib = _lib.load()
        P = u0.shape[0]
        assert out.shape == (P, self.nz, self.N) and out.dtype == torch.complex64 and out.is_contiguous()
        assert done_log.shape == (self.nz - 1, P) and done_log.dtype == torch.int32
        out[:, 0].copy_(u0)
        x = self._bufs.setdefault(("x", P), torch.empty((P, self.N), dtype=torch.complex64, device=self.dev))
        for s in range(1, self.nz):
            u = out[:, s - 1]
            x.copy_(u)
            ucont = self._bufs.setdefault(("u", P), torch.empty((P, self.N), dtype=torch.complex64, device=self.dev))
            ucont.copy_(u)
            done = self.solve_slice(s, ucont, x, max_iter, check_every=0)
            done_log[s - 1].copy_(done)
            nxt = self._bufs.setdefault(("n", P), torch.empty((P, self.N), dtype=torch.complex64, device=self.dev))
            _lib.check(lib.tp_cn_finish(P * self.N, _lib.ptr(x), _lib.ptr(ucont), _lib.ptr(nxt), 2.0,
                                        _lib.stream_handle()), "tp_cn_finish")
            out[:, s].copy_(nxt)
