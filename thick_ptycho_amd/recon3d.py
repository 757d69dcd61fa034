"""Iterative reconstruction of a 3D object from far-field amplitudes on the GPU.

Each iteration evaluates J(n) = sum_p sum (|fft2(u_L^p)| - sqrt(I^p))^2 and its
exact gradient (forward CN sweep of every probe, loss seed from the 2D far
field, exact-adjoint sweep, per-slice accumulation
G[s-1] = sum_p conj(lambda_s)(u_s + u_{s-1})), then takes the steepest-descent
step n <- n - alpha grad J with alpha = 0.5 / L, L a secant estimate of the
gradient's Lipschitz constant.  The step operators are factored per iteration
with the banded block LU at a fixed band half-width, so an iteration contains no
host synchronisation and can be captured once into a HIP graph and replayed.

The reference reconstructs with NLCG (reconstruction/pwe_reconstructor.py:230-368,
out of scope here); this loop exercises the same forward / adjoint / gradient
kernels for two transverse axes.

With ``forward="krylov"`` the forward planes come from the learned-preconditioned
BiCGSTAB stepper instead (``learned.KrylovStepper`` on the current object); the
adjoint and the gradient stay on the banded LU.  Captured into a HIP graph, each
slice runs a fixed number of BiCGSTAB iterations (``krylov_graph_iters``) with the
convergence flags kept on the device; after each replay the flags are read once,
and an iteration in which a probe missed the tolerance on some slice is undone and
recomputed eagerly, with the banded LU of that slice for the probes that miss it
(the fallback).
"""

from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from . import _lib
from .forward_model import band_engine as be


class Reconstruction3D:
    def __init__(self, sqrt_i, probes, n0, k: float, dx: float, dy: float, dz: float, bc: str = "dirichlet",
                 batch: Optional[int] = None, w: Optional[int] = None, probes_per_block: int = 1, device=None,
                 forward: str = "band", cnn=None, krylov_tol: float = 1e-6, exit_data=None, refine: int = 0,
                 krylov_graph_iters: int = 4):
        """``sqrt_i`` (P, nx, ny): measured far-field amplitudes (float64 amplitudes are
        kept as float32 high and low parts); or ``sqrt_i=None`` with ``exit_data``
        (P, nx, ny) complex exit waves and the loss sum |u_L - exit_data|^2 (the
        reference's non phase-retrieval data; iterations then run eagerly).

        ``refine`` (0, 1 or 2; banded forward stepping with amplitude data): refinement
        steps of the exit plane before the loss seed, each an FP64 residual of the
        recurrence (tp_band_residual) and a complex64 correction sweep of the same
        factors, with the seed then formed in FP64 (tp_amplitude_seed2d_refined).
        Where the model's far field lies many orders of magnitude below its peak the
        seed's phase follows the exit plane's rounding; two steps bring the exit
        plane to FP64 accuracy, at the cost of two residual passes and two more band
        sweeps per gradient."""
        dev = device or torch.device("cuda", torch.cuda.current_device())
        self.dev = dev
        n0 = np.asarray(n0)
        nx, ny, nz = n0.shape
        self.nx, self.ny, self.nz, self.N, self.L = nx, ny, nz, nx * ny, nz - 1
        self.k, self.dz = k, dz
        self.ppb = probes_per_block
        self.schedule = 0
        P = probes.shape[0]
        self.P = P
        self.batch = batch or P
        self.n = torch.tensor(n0, dtype=torch.complex128, device=dev)
        pr = torch.as_tensor(probes).to(dev).reshape(P, -1)
        self.probes = pr.to(torch.complex64).contiguous()
        # complex128 probes: their complex64 rounding error starts the first refinement sweep
        self.probes_lo = (pr - self.probes.to(pr.dtype)).to(torch.complex64).contiguous() \
            if pr.dtype == torch.complex128 else None
        if (sqrt_i is None) == (exit_data is None):
            raise ValueError("give either sqrt_i (intensity data) or exit_data (exit waves)")
        self.sqrt_i_lo = None
        if sqrt_i is None:
            self.sqrt_i = None
        else:
            a = torch.as_tensor(sqrt_i).to(dev).reshape(P, nx, ny)
            self.sqrt_i = a.to(torch.float32).contiguous()
            if a.dtype == torch.float64:
                self.sqrt_i_lo = (a - self.sqrt_i.double()).to(torch.float32).contiguous()
        self.exit_data = None if exit_data is None else \
            torch.as_tensor(exit_data).to(dev, torch.complex64).reshape(P, nx * ny).contiguous()
        A, B = be.step_operators_3d(nx, ny, k, dx, dy, dz, bc)
        self.C = torch.empty((self.N, self.L), dtype=torch.complex128, device=dev)
        self._contribution()
        w = w or be.factor(self.C, A, B, "forward").w  # band half-width chosen once, at n0
        self.w = w
        self.fwd = be.BandFactorizer(A, B, "forward", self.L, w, dev)
        self.adj = be.BandFactorizer(A, B, "exact_adjoint", self.L, w, dev, source=self.fwd)
        Pb = self.batch
        self.u = torch.empty((Pb, self.L + 1, self.N), dtype=torch.complex64, device=dev)
        self.lam = torch.empty_like(self.u)
        self.zero = torch.zeros((Pb, self.N), dtype=torch.complex64, device=dev)
        self.seed = torch.empty((Pb, 1, self.N), dtype=torch.complex64, device=dev)
        self.exit = torch.empty((Pb, self.N), dtype=torch.complex64, device=dev)
        self.work = torch.empty((2, Pb, self.N), dtype=torch.complex64, device=dev)
        self.loss = torch.zeros(P, dtype=torch.float32, device=dev)
        self.G = torch.zeros((self.L, self.N), dtype=torch.complex64, device=dev)
        self.grad = torch.zeros((self.N, nz), dtype=torch.complex128, device=dev)
        self.alpha = torch.zeros(1, dtype=torch.float64, device=dev)
        self.graph = None
        if forward not in ("band", "krylov"):
            raise ValueError("forward must be 'band' or 'krylov'")
        if refine not in (0, 1, 2):
            raise ValueError("refine must be 0, 1 or 2")
        if refine and (forward != "band" or self.sqrt_i is None):
            raise ValueError("refinement is built for the banded forward stepping with amplitude data")
        self.refine = refine
        if refine:
            lib = _lib.load()
            self.CT = torch.empty((self.L, self.N), dtype=torch.complex128, device=dev)  # C by step: coalesced reads
            self.corr = torch.empty_like(self.u) if refine == 2 else None
            self.xparts = torch.empty((3, Pb, self.N), dtype=torch.complex64, device=dev)
            self.sweep_work = torch.empty((Pb, 2, self.N), dtype=torch.complex64, device=dev)
            nbytes = int(lib.tp_amplitude_seed2d_refined_work_bytes(nx, ny, Pb))
            self.seed_work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=dev)
        self.forward = forward
        self.krylov = None
        self.krylov_stats = []
        if forward == "krylov":
            from . import learned

            self.krylov = learned.KrylovStepper(n0, k, dx, dy, dz, bc, precond="cnn" if cnn is not None else "none",
                                                cnn=cnn, tol=krylov_tol, device=dev)
            self.krylov.C = self.C  # the stepper follows the object being reconstructed
            self.krylov_graph_iters = krylov_graph_iters
            nbatch = (P + self.batch - 1) // self.batch
            self.krylov_done = torch.ones((nbatch, self.L, self.batch), dtype=torch.int32, device=dev)
            self._krylov_device = False  # sweep_device (inside a graph capture) or sweep with fallback
            self.krylov_redone = 0

    def _contribution(self) -> None:
        coef = (self.k / 1j) * (self.n.reshape(self.N, self.nz) - 1)
        self.C.copy_((self.dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2)

    def gradient(self, seed: Optional[torch.Tensor] = None) -> None:
        """self.loss (per probe) and self.grad ((N, nz), dJ/dRe n + i dJ/dIm n) at self.n;
        asynchronous.  ``seed`` ((P, N) complex64 on the device) replaces the loss
        seed dJ/d conj(u_L) in the adjoint sweep when given."""
        lib = _lib.load()
        sh = _lib.stream_handle
        self._contribution()
        fwd = self.fwd.run(self.C)
        adj = self.adj.run(self.C)
        if self.krylov is not None:
            self.krylov._band = None  # fallback factors of the previous object are stale
        self.G.zero_()
        self.loss.zero_()
        for b0 in range(0, self.P, self.batch):
            b1 = min(b0 + self.batch, self.P)
            nb = b1 - b0
            u, lam = self.u[:nb], self.lam[:nb]
            if self.krylov is not None and self._krylov_device:
                self.krylov.sweep_device(self.probes[b0:b1], u, self.krylov_done[b0 // self.batch, :, :nb],
                                         max_iter=self.krylov_graph_iters)
            elif self.krylov is not None:
                planes, st = self.krylov.sweep(self.probes[b0:b1], keep="all")
                u.copy_(planes)
                self.krylov_stats.append(st)
            else:
                be.sweep([be.BandScanSpec(fwd, self.probes[b0:b1], out=u)], self.ppb, schedule=self.schedule)
            ex = self.exit[:nb]
            ex.copy_(u[:, self.L])
            if self.refine and seed is None:
                self._refined_seed(b0, b1, u)
            elif self.exit_data is not None:  # J = sum |u_L - d|^2: seed u_L - d
                torch.sub(ex, self.exit_data[b0:b1], out=self.seed[:nb, 0])
                self.loss[b0:b1] = (self.seed[:nb, 0].abs() ** 2).sum(dim=1)
            else:
                _lib.check(lib.tp_amplitude_seed2d(_lib.ptr(ex), _lib.ptr(self.sqrt_i[b0:b1]), self.nx, self.ny, nb,
                                                   1.0, _lib.ptr(self.seed[:nb]), _lib.ptr(self.loss[b0:b1]),
                                                   _lib.ptr(self.work), sh()), "tp_amplitude_seed2d")
            rhs = self.seed[:nb] if seed is None else seed[b0:b1].reshape(nb, 1, self.N)
            be.sweep([be.BandScanSpec(adj, self.zero[:nb], out=lam, rhs=rhs, rhs_steps=1)], self.ppb,
                     schedule=self.schedule)
            _lib.check(lib.tp_grad3d_accumulate(_lib.ptr(u), _lib.ptr(lam), nb, self.L, self.N, _lib.ptr(self.G), sh()),
                       "tp_grad3d_accumulate")
        Gt = self.G.transpose(0, 1).to(torch.complex128)
        self.grad.zero_()
        self.grad[:, :-1] += Gt
        self.grad[:, 1:] += Gt
        self.grad.mul_((self.dz / 4) * (self.k / 1j)).conj_physical_().mul_(2)

    def _refined_seed(self, b0: int, b1: int, u: torch.Tensor) -> None:
        """Loss seed and per-probe loss of the refined exit plane u_L + e1_L (+ e2_L):
        FP64 residuals of the forward recurrence at the iterate, complex64 correction
        sweeps of the forward factors (the first started from the probes' complex64
        rounding error when they were given in complex128), the seed in FP64.  The residuals use lam as scratch (it is written by the adjoint sweep
        afterwards)."""
        lib = _lib.load()
        nb, L = b1 - b0, self.L
        fac = self.fwd.fac
        self.CT.copy_(self.C.transpose(0, 1))
        Cv = self.CT.transpose(0, 1)
        r = self.lam[:nb, :L]
        X = self.xparts[:, :nb]
        X[0].copy_(u[:, L])
        e0 = self.zero[:nb] if self.probes_lo is None else self.probes_lo[b0:b1]
        be.residual(self.fwd, Cv, [u], r)
        if self.refine == 1:
            be.sweep([be.BandScanSpec(fac, e0, exit=X[1], rhs=r, rhs_steps=L, work=self.sweep_work)],
                     self.ppb, schedule=self.schedule)
            parts = (X[0], X[1], None)
        else:
            e1 = self.corr[:nb]
            be.sweep([be.BandScanSpec(fac, e0, out=e1, rhs=r, rhs_steps=L)], self.ppb, schedule=self.schedule)
            be.residual(self.fwd, Cv, [u, e1], r)
            X[1].copy_(e1[:, L])
            be.sweep([be.BandScanSpec(fac, self.zero[:nb], exit=X[2], rhs=r, rhs_steps=L, work=self.sweep_work)],
                     self.ppb, schedule=self.schedule)
            parts = (X[0], X[1], X[2])
        lo = None if self.sqrt_i_lo is None else self.sqrt_i_lo[b0:b1]
        _lib.check(lib.tp_amplitude_seed2d_refined(*[_lib.ptr(x) for x in parts], self.N, _lib.ptr(self.sqrt_i[b0:b1]),
                                                   _lib.ptr(lo), self.nx, self.ny, nb, 1.0, _lib.ptr(self.seed[:nb]),
                                                   _lib.ptr(self.loss[b0:b1]), _lib.ptr(self.seed_work),
                                                   _lib.stream_handle()), "tp_amplitude_seed2d_refined")

    def grad_object(self) -> torch.Tensor:
        return self.grad.reshape(self.nx, self.ny, self.nz)

    def _update(self) -> None:
        self.n.sub_(self.alpha * self.grad.reshape(self.nx, self.ny, self.nz))

    def lipschitz(self, h: float = 1e-5) -> float:
        """Secant estimate ||g(n - d) - g(n)|| / ||d|| along d = h g / max|g|."""
        self.gradient()
        g0 = self.grad.clone()
        d = (h / g0.abs().max()) * g0.reshape(self.nx, self.ny, self.nz)
        n0 = self.n.clone()
        self.n.sub_(d)
        self.gradient()
        Lc = float(torch.linalg.vector_norm(self.grad - g0) / torch.linalg.vector_norm(d))
        self.n.copy_(n0)
        return Lc

    def tune(self, candidates=None) -> dict:
        """Pick the band sweeps' shape (probes per workgroup, schedule) by timing one
        gradient evaluation with each; returns {"ppb/schedule": ms}."""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        times = {}
        for shape in candidates or be.sweep_shapes(self.w):
            self.ppb, self.schedule = shape
            self.gradient()
            best = float("inf")
            for _ in range(2):  # the better of two evaluations
                e0.record()
                self.gradient()
                e1.record()
                torch.cuda.synchronize()
                best = min(best, e0.elapsed_time(e1))
            times[shape] = best
        self.ppb, self.schedule = min(times, key=times.get)
        self.graph = None
        return {f"{p}/{q}": t for (p, q), t in times.items()}

    def capture(self) -> None:
        """Capture one iteration (gradient + update) into a HIP graph."""
        self._krylov_device = self.krylov is not None
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        n_save = self.n.clone()
        with torch.cuda.stream(side):  # first launches outside the capture
            self.gradient()
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.gradient()
            self._update()
        self.n.copy_(n_save)  # the capture does not execute; keep n as it was
        self._krylov_device = False

    def run(self, iters: int, alpha: Optional[float] = None, graph: bool = True, check_band: bool = True) -> dict:
        """Steepest descent; returns {'loss': [J before each iteration], 'alpha': ...}."""
        if alpha is None:
            alpha = 0.5 / self.lipschitz()
        self.alpha.fill_(alpha)
        graph = graph and self.exit_data is None  # torch reductions stay out of graphs
        if graph and self.graph is None:
            self.capture()
        hist: List[float] = []
        n_prev = torch.empty_like(self.n) if (graph and self.krylov is not None) else None
        for _ in range(iters):
            if graph:
                if n_prev is not None:
                    n_prev.copy_(self.n)
                    self.krylov_done.fill_(1)
                self.graph.replay()
                if n_prev is not None and int(self.krylov_done.min().item()) <= 0:
                    # a probe missed the tolerance within the captured iterations on some
                    # slice: undo the step and redo it eagerly with the banded-LU fallback
                    self.n.copy_(n_prev)
                    self.krylov_redone += 1
                    self.gradient()
                    self._update()
            else:
                self.gradient()
                self._update()
            hist.append(float(self.loss.double().sum().item()))
        if check_band:
            self.fwd.check()
            self.adj.check()
        return {"loss": hist, "alpha": alpha}

    def reset(self, n) -> None:
        """Restart from the object ``n`` (nx, ny, nz); the captured graph stays valid."""
        self.n.copy_(torch.as_tensor(np.asarray(n), dtype=torch.complex128))

    def object(self) -> np.ndarray:
        return self.n.cpu().numpy()


def simulate_exit_3d(n, probes, k: float, dx: float, dy: float, dz: float, bc: str = "dirichlet", refine: int = 2,
                     batch: Optional[int] = None, device=None) -> torch.Tensor:
    """Exit planes of banded-LU forward stepping of every probe through ``n``
    (nx, ny, nz), refined: complex64 (1 + refine, P, nx*ny) on the GPU whose sum is
    u_L to FP64 accuracy after two refinement steps (the complex64 plane, then
    the exit parts of the correction sweeps driven by FP64 residuals of the
    recurrence, tp_band_residual; complex128 probes start the first correction
    from their complex64 rounding error).  Keeps every plane of a batch of probes."""
    if refine not in (0, 1, 2):
        raise ValueError("refine must be 0, 1 or 2")
    dev = device or torch.device("cuda", torch.cuda.current_device())
    n = np.asarray(n)
    nx, ny, nz = n.shape
    N, L = nx * ny, nz - 1
    P = probes.shape[0]
    batch = batch or P
    coef = (k / 1j) * (n.reshape(N, nz) - 1)
    C = torch.tensor((dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2, dtype=torch.complex128, device=dev)
    A, B = be.step_operators_3d(nx, ny, k, dx, dy, dz, bc)
    w = be.factor(C, A, B, "forward").w
    bf = be.BandFactorizer(A, B, "forward", L, w, dev)
    fac = bf.run(C)
    bf.check()
    pr = torch.as_tensor(probes).to(dev).reshape(P, N)
    u0 = pr.to(torch.complex64).contiguous()
    lo = (pr - u0.to(pr.dtype)).to(torch.complex64).contiguous() if pr.dtype == torch.complex128 else None
    out = torch.empty((1 + refine, P, N), dtype=torch.complex64, device=dev)
    Pb = min(batch, P)
    u = torch.empty((Pb, L + 1, N), dtype=torch.complex64, device=dev)
    e1 = torch.empty_like(u) if refine == 2 else None
    r = torch.empty((Pb, L, N), dtype=torch.complex64, device=dev) if refine else None
    zero = torch.zeros((Pb, N), dtype=torch.complex64, device=dev)
    for b0 in range(0, P, batch):
        b1 = min(b0 + batch, P)
        nb = b1 - b0
        be.sweep([be.BandScanSpec(fac, u0[b0:b1], out=u[:nb])])
        out[0, b0:b1].copy_(u[:nb, L])
        if refine:
            e0 = zero[:nb] if lo is None else lo[b0:b1]
            be.residual(bf, C, [u[:nb]], r[:nb])
            if refine == 1:
                be.sweep([be.BandScanSpec(fac, e0, exit=out[1, b0:b1], rhs=r[:nb], rhs_steps=L)])
            else:
                be.sweep([be.BandScanSpec(fac, e0, out=e1[:nb], rhs=r[:nb], rhs_steps=L)])
                out[1, b0:b1].copy_(e1[:nb, L])
                be.residual(bf, C, [u[:nb], e1[:nb]], r[:nb])
                be.sweep([be.BandScanSpec(fac, zero[:nb], exit=out[2, b0:b1], rhs=r[:nb], rhs_steps=L)])
    return out


def simulate_intensity_3d(n, probes, k: float, dx: float, dy: float, dz: float, bc: str = "dirichlet",
                          batch: Optional[int] = None, device=None) -> torch.Tensor:
    """|fft2(u_L)|^2 per probe, float32 (P, nx, ny) on the GPU: banded-LU forward
    stepping of every probe through ``n`` (nx, ny, nz), exit planes only."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    n = np.asarray(n)
    nx, ny, nz = n.shape
    N = nx * ny
    P = probes.shape[0]
    batch = batch or P
    coef = (k / 1j) * (n.reshape(N, nz) - 1)
    C = torch.tensor((dz / 2) * (coef[:, :-1] + coef[:, 1:]) / 2, dtype=torch.complex128, device=dev)
    A, B = be.step_operators_3d(nx, ny, k, dx, dy, dz, bc)
    fac = be.factor(C, A, B, "forward")
    u0 = torch.as_tensor(probes).to(dev, torch.complex64).reshape(P, N).contiguous()
    ex = torch.empty((P, N), dtype=torch.complex64, device=dev)
    for b0 in range(0, P, batch):
        b1 = min(b0 + batch, P)
        be.sweep([be.BandScanSpec(fac, u0[b0:b1], exit=ex[b0:b1])])
    from .forward_model import ms_engine

    return ms_engine.farfield_2d(ex, nx, ny)
