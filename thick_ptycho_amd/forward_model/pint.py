"""Drop-in ``PWEFullPinTSolver``: the whole space-z system of one transverse axis
solved at once by right-preconditioned GMRES with the alpha-circulant block
preconditioner, every probe of a projection in one batch on the GPU.

Mirrors thick_ptycho/forward_model/pwe/solvers/full_pint.py:71-275 and
.../solvers/utils/_pint_utils.py:27-199.  The system is

    (A - diag C_j) u_j - (B + diag C_j) u_{j-1} = b_j,   j = 1..L,  u_0 = 0,

with b_1 = B0 u_probe (forward_model.py:155-193) or ``rhs_block``; adjoint mode
uses (A^H, B^H) and conj(C) with its slices reversed (full_pint.py:148-160).
Its solution is the reference's whole-system solve (the first block drops
C_1 u_probe, like full_lu.py), not the slice stepping.

GPU design:

* M x is ``tp_pint_matvec`` (FP32 inside the Krylov cycles, FP64 for the
  residual between cycles: each restart is one step of iterative refinement,
  so the solution reaches the FP64 tolerance although the cycles run in FP32);
* P_alpha^{-1} = Gamma^{-1} F blockdiag(A - omega_k B)^{-1} F^{-1} Gamma: the two
  DFTs along z with Gamma and the block scalings folded in are dense L x L
  matrices applied as one batched complex GEMM each (rocBLAS through
  torch.matmul); the L shifted tridiagonal blocks are factored once per mode
  with ``tp_tri_factor`` (A - omega B = (1 + omega)(A - (2 omega / (1 + omega)) I)
  since A + B = 2I) and solved by ``tp_tri_sweep`` with one wavefront per
  (probe, frequency): a one-step sweep from u = 0 with the transformed vector
  as its right-hand side, each frequency its own factor set (coef_pstride);
* Gram-Schmidt (classical, applied twice): ``tp_pint_dots`` forms every
  projection of the new vector (and its norm) in one pass over the basis,
  ``tp_pint_combine`` subtracts them in another; the small Hessenberg
  least-squares problems by Givens rotations in FP64 on the host.

Differences from the reference, same solution: the blocks of the preconditioner
follow the mode's own pair (the reference keeps the forward free-space pair in
adjoint mode, where GMRES then needs many more iterations or stalls); GMRES runs
on all probes of a projection together instead of one probe at a time; the
restart residual is exact in FP64.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, fields
from typing import Optional

import numpy as np
import torch

from .. import _lib
from . import tri_engine
from .base_solver import BasePWESolver, host_complex128
from .iterative_lu import PWEForwardModel
from .operators import BoundaryType, Tridiag, layout_for


def pint_matvec(ops: torch.Tensor, Ct: torch.Tensor, x: torch.Tensor, b: Optional[torch.Tensor] = None,
                out: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """y = M x (or b - M x) for x (P, L, N); complex64 or complex128 throughout.
    Ct is (L, N) for every probe or (P, L, N), one object term per probe."""
    P, L, N = x.shape
    fp64 = x.dtype == torch.complex128
    for t in (ops, Ct, x, b, out):
        if t is not None:
            assert t.is_cuda and t.dtype == x.dtype and t.is_contiguous()
    assert ops.numel() == 6 * N and Ct.shape in ((L, N), (P, L, N)) and (b is None or b.shape == x.shape)
    y = out if out is not None else torch.empty_like(x)
    assert y.shape == x.shape
    c_pstride = L * N if Ct.dim() == 3 else 0
    _lib.check(_lib.load().tp_pint_matvec(_lib.ptr(ops), _lib.ptr(Ct), c_pstride, _lib.ptr(x), _lib.ptr(b),
                                          _lib.ptr(y), N, L, P, int(fp64), _lib.stream_handle(stream)),
               "tp_pint_matvec")
    return y


def _ops_tensor(A: Tridiag, B: Tridiag, dtype, dev) -> torch.Tensor:
    arr = np.concatenate([A.sub, A.diag, A.sup, B.sub, B.diag, B.sup]).astype(np.complex128)
    return torch.from_numpy(arr).to(dev, dtype)


class PiTPreconditioner:
    """P_alpha^{-1} (_pint_utils.py:62-199) for batches (P, L, N) on the GPU."""

    def __init__(self, A: Tridiag, B: Tridiag, N: int, L: int, alpha: float, device, chunk: Optional[int] = None):
        self.A, self.B, self.N, self.L, self.alpha = A, B, N, L, alpha
        self.chunk, self.waves = (chunk, 1) if chunk else layout_for(N)
        gamma = alpha ** (np.arange(L) / L)
        omega = np.exp(2j * np.pi * np.arange(L) / L) * alpha ** (1.0 / L)
        E = np.exp(2j * np.pi * np.outer(np.arange(L), np.arange(L)) / L) / np.sqrt(L)  # ifft(norm='ortho')
        self.Wf = torch.from_numpy((E * gamma[None, :]) / (1 + omega)[:, None]).to(device, torch.complex64)
        self.Wb = torch.from_numpy(np.conj(E) / gamma[:, None]).to(device, torch.complex64)
        shift = 2 * omega / (1 + omega)
        Ccols = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(shift[None, :], (N, L)))).to(device)
        self.fac = tri_engine.factor(Ccols, A, B, "forward", chunk=self.chunk,
                                     waves=self.waves if self.waves > 1 else None)
        self.fac.check()
        f = self.fac
        self._one = tri_engine.TriFactors(f.coef, N, 1, self.chunk, "forward", f.b_sign, f.status, False, keep=[f],
                                          waves=f.waves, wide=f.wide, plan=f.plan)
        self._zero = torch.zeros((1, N), dtype=torch.complex64, device=device).expand(L, N)

    def block_solve(self, Xh: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """out[p, k] = (A - (2 w_k / (1 + w_k)) I)^{-1} Xh[p, k] for every probe p and frequency k."""
        P = Xh.shape[0]
        per = tri_engine.coef_elems(1, self.chunk, self.waves)
        specs = [tri_engine.TriScanSpec(self._one, self._zero, exit=out[p], rhs=Xh[p].unsqueeze(1), rhs_steps=1,
                                        coef_pstride=per) for p in range(P)]
        for i in range(0, P, _lib.TP_MAX_SCANS):
            tri_engine.sweep(specs[i:i + _lib.TP_MAX_SCANS], probes_per_block=1)
        return out

    def apply(self, X: torch.Tensor) -> torch.Tensor:
        """P_alpha^{-1} X for X complex64 (P, L, N)."""
        Xh = torch.matmul(self.Wf, X)
        Z = torch.empty_like(Xh)
        self.block_solve(Xh, Z)
        return torch.matmul(self.Wb, Z)


MAX_RESTART = 32  # rows of one tp_pint_dots call


def arnoldi_dots(V: torch.Tensor, nrows: int, w: torch.Tensor, stream=None):
    """(h (P, nrows) = conj(V[:, :nrows]) . w, ||w||^2 (P,)) with V (P, R, len), w (P, len)
    complex64 (tp_pint_dots, then a sum over the blocks' partials)."""
    P, _, LN = V.shape
    assert V.stride(-1) == 1 and w.stride(-1) == 1 and w.shape == (P, LN) and nrows <= MAX_RESTART
    nblk = int(max(1, min(-(-1024 // max(P, 1)), -(-LN // 256), 65535)))
    part = torch.empty((P, nblk, MAX_RESTART + 1), dtype=torch.complex64, device=V.device)
    _lib.check(_lib.load().tp_pint_dots(_lib.ptr(V), V.stride(0), V.stride(1), nrows, _lib.ptr(w), w.stride(0), LN, P,
                                        nblk, _lib.ptr(part), _lib.stream_handle(stream)), "tp_pint_dots")
    tot = part.sum(dim=1)
    return tot[:, :nrows], tot[:, nrows].real


def arnoldi_combine(V: torch.Tensor, nrows: int, c: torch.Tensor, w: Optional[torch.Tensor], out: torch.Tensor,
                    stream=None) -> torch.Tensor:
    """out = w - sum_i V[:, i] c[:, i] (w None: zero) for V (P, R, len), c (P, >= nrows)."""
    P, _, LN = V.shape
    c = c.to(torch.complex64).contiguous()
    assert out.shape == (P, LN) and out.stride(-1) == 1 and c.shape[0] == P and c.shape[1] >= nrows
    _lib.check(_lib.load().tp_pint_combine(_lib.ptr(V), V.stride(0), V.stride(1), nrows, _lib.ptr(c), c.stride(0),
                                           _lib.ptr(w), 0 if w is None else w.stride(0), _lib.ptr(out), out.stride(0),
                                           LN, P, _lib.stream_handle(stream)), "tp_pint_combine")
    return out


def _givens(a: np.ndarray, b: np.ndarray):
    """(c, s, r) with [[conj c, conj s], [-s, c]] [a; b] = [r; 0], elementwise."""
    r = np.sqrt(np.abs(a) ** 2 + np.abs(b) ** 2)
    safe = np.where(r > 0, r, 1.0)
    c = np.where(r > 0, a / safe, 1.0)
    s = np.where(r > 0, b / safe, 0.0)
    return c, s, r


@dataclass
class GMRESResult:
    x: torch.Tensor          # complex128 (P, L, N)
    converged: np.ndarray    # bool (P,)
    iterations: int          # inner iterations of the batch (each one M P^{-1} product)
    cycles: int
    residual: np.ndarray     # final ||b - M x|| per probe (FP64)


def gmres(ops32, ops64, Ct32, Ct64, pre: PiTPreconditioner, b: torch.Tensor, x0: Optional[torch.Tensor] = None,
          atol: float = 1e-8, rtol: float = 1e-5, restart: int = 20, maxiter: int = 100, inner_rtol: float = 1e-5,
          log=None) -> GMRESResult:
    """Restarted right-preconditioned GMRES on every probe of b (P, L, N) complex128.

    Stops a probe when ||b - M x|| <= max(atol, rtol ||b||) (scipy's rule) or after
    ``maxiter`` cycles.  Each cycle solves M d = r (r the FP64 residual) by at most
    ``restart`` FP32 iterations, to max(tolerance, inner_rtol ||r||)."""
    P, L, N = b.shape
    dev = b.device
    LN = L * N
    if not 1 <= restart <= MAX_RESTART:
        raise ValueError(f"restart must be between 1 and {MAX_RESTART}")
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    bnorm = torch.linalg.vector_norm(b.reshape(P, -1), dim=1).cpu().numpy()
    tol = np.maximum(atol, rtol * bnorm)
    V = torch.empty((P, restart + 1, LN), dtype=torch.complex64, device=dev)
    r64 = torch.empty_like(b)
    its = cycles = 0
    beta = bnorm
    done = np.zeros(P, dtype=bool)
    for cycles in range(1, maxiter + 1):
        pint_matvec(ops64, Ct64, x, b, out=r64)
        beta = torch.linalg.vector_norm(r64.reshape(P, -1), dim=1).cpu().numpy()
        done = beta <= tol
        if log:
            log(f"cycle {cycles}: max relative residual {np.max(beta / np.maximum(bnorm, 1e-300)):.3e}")
        if done.all():
            cycles -= 1
            break
        target = np.maximum(tol, inner_rtol * beta)
        inv = torch.from_numpy(np.where(beta > 0, 1.0 / np.maximum(beta, 1e-300), 0.0)).to(dev)
        V[:, 0] = (r64.reshape(P, LN) * inv[:, None]).to(torch.complex64)
        H = np.zeros((P, restart + 1, restart), dtype=complex)
        cs = np.zeros((P, restart), dtype=complex)
        sn = np.zeros((P, restart), dtype=complex)
        g = np.zeros((P, restart + 1), dtype=complex)
        g[:, 0] = beta
        kp = np.where(done, 0, -1)  # iterations each probe uses (-1: still running)
        k = 0
        for j in range(restart):
            w = pint_matvec(ops32, Ct32, pre.apply(V[:, j].view(P, L, N))).reshape(P, LN)
            h1, _ = arnoldi_dots(V, j + 1, w)  # classical Gram-Schmidt, applied twice
            arnoldi_combine(V, j + 1, h1, w, w)
            h2, _ = arnoldi_dots(V, j + 1, w)
            arnoldi_combine(V, j + 1, h2, w, w)
            _, nrm2 = arnoldi_dots(V, 0, w)
            hn = torch.sqrt(nrm2)
            scale = torch.where(hn > 1e-30, 1.0 / torch.clamp(hn, min=1e-30), torch.zeros_like(hn))
            V[:, j + 1] = w * scale[:, None]
            hc = torch.cat([h1 + h2, hn.to(torch.complex64)[:, None]], dim=1).cpu().numpy()
            col = hc.astype(complex)
            for i in range(j):
                a0, b0 = col[:, i].copy(), col[:, i + 1].copy()
                col[:, i] = np.conj(cs[:, i]) * a0 + np.conj(sn[:, i]) * b0
                col[:, i + 1] = -sn[:, i] * a0 + cs[:, i] * b0
            cs[:, j], sn[:, j], col[:, j] = _givens(col[:, j], col[:, j + 1])
            col[:, j + 1] = 0
            H[:, : j + 2, j] = col
            g[:, j + 1] = -sn[:, j] * g[:, j]
            g[:, j] = np.conj(cs[:, j]) * g[:, j]
            its += 1
            k = j + 1
            newly = (kp < 0) & (np.abs(g[:, j + 1]) <= target)
            kp[newly] = k
            if (kp >= 0).all():
                break
        kp[kp < 0] = k
        kmax = int(kp.max())
        if kmax == 0:
            continue
        y = np.zeros((P, kmax), dtype=complex)
        for p in range(P):  # back substitution R y = g on the probe's own k
            kk = int(kp[p])
            for i in range(kk - 1, -1, -1):
                y[p, i] = (g[p, i] - H[p, i, i + 1:kk] @ y[p, i + 1:kk]) / H[p, i, i]
        yt = torch.from_numpy((-y).astype(np.complex64)).to(dev)
        d = arnoldi_combine(V, kmax, yt, None, torch.empty((P, LN), dtype=torch.complex64, device=dev))
        x += pre.apply(d.view(P, L, N)).to(torch.complex128)
    else:
        pint_matvec(ops64, Ct64, x, b, out=r64)
        beta = torch.linalg.vector_norm(r64.reshape(P, -1), dim=1).cpu().numpy()
        done = beta <= tol
    return GMRESResult(x, done, its, cycles, beta)


@dataclass
class PWEFullPinTSolverCache:
    """Per (projection, mode): the device operator of the whole system."""

    cached_n: Optional[np.ndarray] = None
    ops: Optional[tuple] = None       # (ops32, ops64)
    Ct: Optional[tuple] = None        # (Ct32, Ct64), (L, N)
    pre: Optional[PiTPreconditioner] = None
    b: Optional[np.ndarray] = None

    def reset(self, n=None):
        for f in fields(self):
            setattr(self, f.name, None)
        self.cached_n = n


class PWEFullPinTSolver(BasePWESolver):
    """Whole-system PWE solver with the PiT preconditioner (one transverse axis)."""

    solver_cache_class = PWEFullPinTSolverCache

    def __init__(self, simulation_space, ptycho_probes, bc_type=BoundaryType.IMPEDANCE, results_dir="",
                 use_logging=False, verbose=False, log=None, alpha=1e-2, atol=1e-8, test_bcs=None, device=None,
                 rtol: float = 1e-5, restart: int = 20, maxiter: int = 100, warm_start: bool = True):
        super().__init__(simulation_space, ptycho_probes, bc_type=bc_type, results_dir=results_dir,
                         use_logging=use_logging, verbose=verbose, log=log, test_bcs=test_bcs, device=device)
        if simulation_space.dimension != 2:
            raise NotImplementedError("the parallel-in-time solver is built for one transverse axis")
        if test_bcs is not None:
            raise NotImplementedError("test_bcs right-hand sides are not built for the parallel-in-time solver")
        if not 1 <= restart <= MAX_RESTART:
            raise ValueError(f"restart must be between 1 and {MAX_RESTART}")
        self.pwe_finite_differences = PWEForwardModel(simulation_space, bc_type=bc_type)
        self.alpha, self.atol, self.rtol = alpha, atol, rtol
        self.restart, self.maxiter, self.warm_start = restart, maxiter, warm_start
        self._pre = {}
        self._u0_cache = {}
        self.last_result: Optional[GMRESResult] = None

    def _pair(self, mode: str):
        fm = self.pwe_finite_differences
        if mode == "adjoint":
            return fm.A.conj_transpose(), fm.B.conj_transpose()
        return fm.A, fm.B  # 'reverse' runs the forward system, as in the reference

    def preconditioner(self, mode: str) -> PiTPreconditioner:
        key = "adjoint" if mode == "adjoint" else "forward"
        if key not in self._pre:
            A, B = self._pair(mode)
            self._pre[key] = PiTPreconditioner(A, B, self.block_size, self.nz - 1, self.alpha, self.device)
        return self._pre[key]

    def _object_rows(self, n, mode: str, scan_idx: int) -> np.ndarray:
        """C_j as rows (L, N) for ``mode`` (full_pint.py:133-160)."""
        C = self.simulation_space.create_object_contribution(n=n, scan_index=scan_idx).reshape(-1, self.nz - 1)
        if mode == "adjoint":
            C = np.flip(np.conj(C), axis=1)
        return np.ascontiguousarray(C.T).astype(np.complex128)

    def _construct_solve_cache(self, n=None, mode: str = "forward", scan_idx: int = 0, proj_idx: int = 0) -> None:
        t0 = time.time()
        A, B = self._pair(mode)
        cache = self.projection_cache[proj_idx].modes[mode]
        cache.ops = (_ops_tensor(A, B, torch.complex64, self.device), _ops_tensor(A, B, torch.complex128, self.device))
        Ct64 = torch.from_numpy(self._object_rows(n, mode, scan_idx)).to(self.device)
        cache.Ct = (Ct64.to(torch.complex64), Ct64)
        cache.pre = self.preconditioner(mode)
        cache.b = np.zeros(self.block_size, dtype=complex)
        cache.cached_n = n
        self._timed(f"prepared the whole-system operator ({mode})", t0)

    def _rhs(self, probes: torch.Tensor, rhs_block) -> torch.Tensor:
        """(P, L, N) complex128: B0 u_probe in the first block (forward_model.py:155-193)
        or the given rhs_block ((N, L) or its (N*L,) full-system vector) for every probe."""
        P, N = probes.shape
        L = self.nz - 1
        if rhs_block is not None:
            r = np.asarray(rhs_block, dtype=complex)
            r = r.reshape(L, N) if r.ndim == 1 else r.reshape(N, L).T
            return torch.from_numpy(np.ascontiguousarray(r)).to(self.device).expand(P, L, N).contiguous()
        fm = self.pwe_finite_differences
        b = torch.zeros((P, L, N), dtype=torch.complex128, device=self.device)
        B0 = fm.B
        d = torch.from_numpy(B0.diag).to(self.device)
        lo = torch.from_numpy(B0.sub).to(self.device)
        up = torch.from_numpy(B0.sup).to(self.device)
        u = probes.to(torch.complex128)
        b[:, 0] = d * u
        b[:, 0, 1:] += lo[1:] * u[:, :-1]
        b[:, 0, :-1] += up[:-1] * u[:, 1:]
        return b

    def _solve_batch(self, cache, mode: str, probes: torch.Tensor, rhs_block, proj_idx: int,
                     Ct: Optional[tuple] = None) -> torch.Tensor:
        b = self._rhs(probes, rhs_block)
        Ct = Ct or cache.Ct
        key = (proj_idx, mode)
        x0 = self._u0_cache.get(key) if self.warm_start else None
        if x0 is not None and x0.shape != b.shape:
            x0 = None
        t0 = time.time()
        res = gmres(cache.ops[0], cache.ops[1], Ct[0], Ct[1], cache.pre, b, x0=x0, atol=self.atol,
                    rtol=self.rtol, restart=self.restart, maxiter=self.maxiter,
                    log=self._log if self.verbose else None)
        self.last_result = res
        self._u0_cache[key] = res.x
        if self.verbose:
            state = "converged" if res.converged.all() else f"{int((~res.converged).sum())} probes not converged"
            self._timed(f"PiT GMRES {state}: {res.iterations} iterations, {res.cycles} cycles", t0)
        return res.x

    def _probe_tensor(self, probes) -> torch.Tensor:
        t = probes.to(self.device) if isinstance(probes, torch.Tensor) else torch.from_numpy(np.asarray(probes))
        return t.to(self.device, torch.complex128).reshape(-1, self.block_size)

    def _solve_single_probe_impl(self, scan_idx: int = 0, proj_idx: int = 0, probe=None, mode: str = "forward",
                                 rhs_block=None) -> np.ndarray:
        N = self.block_size
        pr = np.zeros(N, dtype=complex) if probe is None else np.asarray(probe, dtype=complex).reshape(N)
        cache = self.projection_cache[proj_idx].modes[mode]
        x = self._solve_batch(cache, mode, self._probe_tensor(pr[None]), rhs_block, proj_idx)
        u = x[0].transpose(0, 1).cpu().numpy()
        return np.concatenate([pr.reshape(N, 1), u], axis=1)

    def solve_device(self, n=None, mode: str = "forward", rhs_block=None, probes=None) -> torch.Tensor:
        """complex128 (Proj, A, S, nz, N) on the GPU, plane 0 the probe; every probe of
        a projection goes through one batched GMRES."""
        assert mode in {"forward", "adjoint", "reverse"}, f"Invalid mode: {mode!r}"
        if n is None:
            n = self.simulation_space.refractive_index_empty
        pr = self._probe_tensor(self.probes if probes is None else probes)
        A, S, N = self.num_angles, self.num_probes, self.block_size
        out = torch.empty((self.num_projections, A, S, self.nz, N), dtype=torch.complex128, device=self.device)
        for p in range(self.num_projections):
            cache = self._get_or_construct_cache(n=n, mode=mode, proj_idx=p)
            Ct = None
            if self.solve_reduced_domain:  # every scan on its own window: one object term per probe
                n_p = self.get_projected_obj(n=n, proj_idx=p)
                rows = np.stack([self._object_rows(n_p, mode, s) for s in range(S)])
                Ct64 = torch.from_numpy(rows).to(self.device).expand(A, S, *rows.shape[1:]).reshape(A * S, *rows.shape[1:])
                Ct = (Ct64.to(torch.complex64).contiguous(), Ct64.contiguous())
            x = self._solve_batch(cache, mode, pr, rhs_block, p, Ct)
            out[p, :, :, 0] = pr.reshape(A, S, N)
            out[p, :, :, 1:] = x.reshape(A, S, self.nz - 1, N)
        return out

    def solve(self, n=None, mode: str = "forward", rhs_block=None, probes=None, **kwargs) -> np.ndarray:
        """Reference-compatible solve (base/base_solver.py:84-154): complex128
        (Proj, A, S, nx, nz) on the host."""
        u = self.solve_device(n=n, mode=mode, rhs_block=rhs_block, probes=probes)
        u = host_complex128(u.transpose(-1, -2))
        return u.reshape(self.num_projections, self.num_angles, self.num_probes, *self.effective_shape)


class PWEPetscFullPinTSolver(PWEFullPinTSolver):
    """The PETSc form of the same solve (full_pit_petsc.py:104-327): KSP GMRES with
    right preconditioning, the unpreconditioned residual norm and an absolute
    tolerance only (ksp.setTolerances(atol, rtol=1e-30, max_it=20), :259-278).
    Here rtol is 0, the first cycle runs at most 20 FP32 iterations and at most
    two FP64-refinement cycles follow."""

    def __init__(self, simulation_space, ptycho_probes, bc_type=BoundaryType.IMPEDANCE, results_dir="",
                 use_logging=False, verbose=False, log=None, alpha=1e-2, atol=1e-8, test_bcs=None, device=None):
        super().__init__(simulation_space, ptycho_probes, bc_type=bc_type, results_dir=results_dir,
                         use_logging=use_logging, verbose=verbose, log=log, alpha=alpha, atol=atol,
                         test_bcs=test_bcs, device=device, rtol=0.0, restart=20, maxiter=3)
