// Krylov solve of the two-transverse-axis CN step: BiCGSTAB (van der Vorst,
// SIAM J. Sci. Stat. Comput. 13 (1992) 631-644) with right preconditioning,
// for all probes of a slice at once.
//
//   r = b - A x, rh = r, p = r
//   loop: ph = M p; v = A ph; alpha = rho / (rh, v); s = r - alpha v
//         sh = M s; t = A sh; omega = (t, s) / (t, t)
//         x += alpha ph + omega sh; r = s - omega t
//         rho' = (rh, r); beta = (rho' / rho)(alpha / omega); p = r + beta (p - omega v)
//
// Vectors are complex64; every inner product is accumulated in FP64 (per-
// iteration accumulators, so the scalars of iteration k are recomputed by each
// kernel from them and no scalar-only launches are needed).  A probe stops
// updating once ||r|| <= tol ||b|| (or ||s|| <= tol ||b|| at the half step);
// a zero denominator marks it failed.  M: identity, Jacobi (diag(A)^{-1}) or the
// learned CNN of tp_learn.hip.
//
// The operator is the block-tridiagonal A(s) of tp_band_factor: per slice the
// diagonal d = tri_diag[type(j)] + lhs_sign * c(s) (tp_krylov_diag), off-
// diagonals from the base operator.
#include <vector>

#include "tp_common.h"

namespace tp {

// defined in tp_learn.hip
int cnn_apply_internal(const void* wb, const float2* in, float2* out, int P, int nx, int ny, void* work,
                       hipStream_t st, const int* skip);
size_t cnn_work_bytes_internal(int P, int nx, int ny);

struct KOp {
  const float2* diag;  // [N] this slice's diagonal
  const float2* tri;   // [3][3][nx] base operator (sub, diag, sup) per block type
  const float2* cpl;   // [ny][2] (f_j, e_j)
  int nx, ny;
};

__device__ __forceinline__ float2 apply_A(const KOp& op, const float2* __restrict__ x, int64_t f, int j, int i) {
  const int type = j == 0 ? 0 : (j == op.ny - 1 ? 2 : 1);
  const float2* tr = op.tri + (size_t)type * 3 * op.nx;
  float2 y = cmul(op.diag[f], x[f]);
  if (i > 0) y = cfma(tr[i], x[f - 1], y);
  if (i < op.nx - 1) y = cfma(tr[2 * op.nx + i], x[f + 1], y);
  if (j > 0) y = cfma(op.cpl[2 * j], x[f - op.nx], y);
  if (j < op.ny - 1) y = cfma(op.cpl[2 * j + 1], x[f + op.nx], y);
  return y;
}

// block sum of up to 4 doubles -> atomics
template <int K>
__device__ __forceinline__ void block_acc(double (&v)[K], double* dst) {
  __shared__ double red[4][K];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w][threadIdx.x];
    atomicAdd(&dst[threadIdx.x], s);
  }
}

// per-probe accumulators of one iteration
enum { A_RV_RE, A_RV_IM, A_SS, A_TS_RE, A_TS_IM, A_TT, A_RR, A_RHR_RE, A_RHR_IM, A_NACC };
// per-probe state: [0..1] rho0 (complex), [2] ||b||^2
enum { S_RHO_RE, S_RHO_IM, S_BB, S_NS };

struct KState {
  double* acc;   // [max_iter][P][A_NACC]
  double* st;    // [P][S_NS]
  int* done;     // [P]: 0 running, k+1 converged after iteration k, -1 failed
  int P;
  int64_t N;
  double tol2;
};

__device__ __forceinline__ double2 dc(double re, double im) { return make_double2(re, im); }
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) { return zmul(a, zinv(b)); }

// rho of iteration k (rho_0 = (rh, r0))
__device__ __forceinline__ double2 rho_of(const KState& s, int k, int p) {
  if (k == 0) return dc(s.st[p * S_NS + S_RHO_RE], s.st[p * S_NS + S_RHO_IM]);
  const double* a = s.acc + ((size_t)(k - 1) * s.P + p) * A_NACC;
  return dc(a[A_RHR_RE], a[A_RHR_IM]);
}

__device__ __forceinline__ double2 alpha_of(const KState& s, int k, int p) {
  const double* a = s.acc + ((size_t)k * s.P + p) * A_NACC;
  return zdiv(rho_of(s, k, p), dc(a[A_RV_RE], a[A_RV_IM]));
}

__device__ __forceinline__ double2 omega_of(const KState& s, int k, int p) {
  const double* a = s.acc + ((size_t)k * s.P + p) * A_NACC;
  return zdiv(dc(a[A_TS_RE], a[A_TS_IM]), dc(a[A_TT], 0.0));
}

__global__ void k_init(KOp op, KState s, const float2* __restrict__ b, const float2* __restrict__ x,
                       float2* __restrict__ r, float2* __restrict__ rh, float2* __restrict__ p) {
  const int pr = blockIdx.y;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (f < s.N) {
    const int j = (int)(f / op.nx), i = (int)(f - (int64_t)j * op.nx);
    const size_t o = (size_t)pr * s.N;
    const float2 ax = apply_A(op, x + o, f, j, i);
    const float2 bb = b[o + f];
    const float2 rr = csub(bb, ax);
    r[o + f] = rr;
    rh[o + f] = rr;
    p[o + f] = rr;
    v[0] = (double)rr.x * rr.x + (double)rr.y * rr.y;  // (rh, r) = ||r||^2
    v[2] = (double)bb.x * bb.x + (double)bb.y * bb.y;
  }
  block_acc<3>(v, s.st + (size_t)pr * S_NS);
}

// v = A ph; acc (rh, v)
__global__ void k_spmv_v(KOp op, KState s, int k, const float2* __restrict__ ph, float2* __restrict__ vv,
                         const float2* __restrict__ rh) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (f < s.N) {
    const int j = (int)(f / op.nx), i = (int)(f - (int64_t)j * op.nx);
    const size_t o = (size_t)pr * s.N;
    const float2 y = apply_A(op, ph + o, f, j, i);
    vv[o + f] = y;
    const float2 h = rh[o + f];
    v[0] = (double)h.x * y.x + (double)h.y * y.y;  // conj(h) * y
    v[1] = (double)h.x * y.y - (double)h.y * y.x;
  }
  block_acc<2>(v, s.acc + ((size_t)k * s.P + pr) * A_NACC + A_RV_RE);
}

// s = r - alpha v; acc ||s||^2
__global__ void k_s(KState s, int k, const float2* __restrict__ r, const float2* __restrict__ vv,
                    float2* __restrict__ sv) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double2 al = alpha_of(s, k, pr);
  const float2 a = make_float2((float)al.x, (float)al.y);
  double v[1] = {0.0};
  if (f < s.N) {
    const size_t o = (size_t)pr * s.N + f;
    const float2 x = cfms(a, vv[o], r[o]);
    sv[o] = x;
    v[0] = (double)x.x * x.x + (double)x.y * x.y;
  }
  block_acc<1>(v, s.acc + ((size_t)k * s.P + pr) * A_NACC + A_SS);
}

// t = A sh; acc (t, s), (t, t).  Skipped when ||s|| already meets the tolerance.
__global__ void k_spmv_t(KOp op, KState s, int k, const float2* __restrict__ sh, float2* __restrict__ t,
                         const float2* __restrict__ sv) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const double* a = s.acc + ((size_t)k * s.P + pr) * A_NACC;
  if (a[A_SS] <= s.tol2 * s.st[pr * S_NS + S_BB]) return;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (f < s.N) {
    const int j = (int)(f / op.nx), i = (int)(f - (int64_t)j * op.nx);
    const size_t o = (size_t)pr * s.N;
    const float2 y = apply_A(op, sh + o, f, j, i);
    t[o + f] = y;
    const float2 z = sv[o + f];
    v[0] = (double)y.x * z.x + (double)y.y * z.y;  // conj(t) * s
    v[1] = (double)y.x * z.y - (double)y.y * z.x;
    v[2] = (double)y.x * y.x + (double)y.y * y.y;
  }
  block_acc<3>(v, s.acc + ((size_t)k * s.P + pr) * A_NACC + A_TS_RE);
}

// x += alpha ph + omega sh; r = s - omega t; acc ||r||^2, (rh, r)
__global__ void k_xr(KState s, int k, float2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ ph,
                     const float2* __restrict__ sh, const float2* __restrict__ sv, const float2* __restrict__ t,
                     const float2* __restrict__ rh) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const double* a = s.acc + ((size_t)k * s.P + pr) * A_NACC;
  const bool half = a[A_SS] <= s.tol2 * s.st[pr * S_NS + S_BB];
  const double2 al = alpha_of(s, k, pr);
  const double2 om = half ? dc(0.0, 0.0) : omega_of(s, k, pr);
  const float2 fa = make_float2((float)al.x, (float)al.y), fo = make_float2((float)om.x, (float)om.y);
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (f < s.N) {
    const size_t o = (size_t)pr * s.N + f;
    float2 xx = cfma(fa, ph[o], x[o]);
    float2 rr = sv[o];
    if (!half) {
      xx = cfma(fo, sh[o], xx);
      rr = cfms(fo, t[o], rr);
    }
    x[o] = xx;
    r[o] = rr;
    const float2 h = rh[o];
    v[0] = (double)rr.x * rr.x + (double)rr.y * rr.y;
    v[1] = (double)h.x * rr.x + (double)h.y * rr.y;
    v[2] = (double)h.x * rr.y - (double)h.y * rr.x;
  }
  block_acc<3>(v, s.acc + ((size_t)k * s.P + pr) * A_NACC + A_RR);
}

// convergence / breakdown bookkeeping, one thread per probe
__global__ void k_status(KState s, int k) {
  const int pr = blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= s.P || s.done[pr] != 0) return;
  const double* a = s.acc + ((size_t)k * s.P + pr) * A_NACC;
  const double bb = s.st[pr * S_NS + S_BB];
  const double2 rho = rho_of(s, k, pr);
  const bool half = a[A_SS] <= s.tol2 * bb;
  if (half || a[A_RR] <= s.tol2 * bb) {
    s.done[pr] = k + 1;
    return;
  }
  const bool bad = (a[A_RV_RE] == 0.0 && a[A_RV_IM] == 0.0) || a[A_TT] == 0.0 ||
                   (a[A_RHR_RE] == 0.0 && a[A_RHR_IM] == 0.0) || (rho.x == 0.0 && rho.y == 0.0) ||
                   !(a[A_RR] == a[A_RR]);
  if (bad) s.done[pr] = -1;
}

// p = r + beta (p - omega v), beta = (rho_{k+1} / rho_k)(alpha / omega)
__global__ void k_p(KState s, int k, float2* __restrict__ p, const float2* __restrict__ r,
                    const float2* __restrict__ vv) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const double2 al = alpha_of(s, k, pr), om = omega_of(s, k, pr);
  const double2 beta = zmul(zdiv(rho_of(s, k + 1, pr), rho_of(s, k, pr)), zdiv(al, om));
  const float2 fb = make_float2((float)beta.x, (float)beta.y), fo = make_float2((float)om.x, (float)om.y);
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < s.N) {
    const size_t o = (size_t)pr * s.N + f;
    p[o] = cfma(fb, cfms(fo, vv[o], p[o]), r[o]);
  }
}

__global__ void k_jacobi(KState s, const float2* __restrict__ diag, const float2* __restrict__ in,
                         float2* __restrict__ out) {
  const int pr = blockIdx.y;
  if (s.done[pr] != 0) return;
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < s.N) {
    const float2 d = diag[f];
    const float inv = 1.f / (d.x * d.x + d.y * d.y);
    const size_t o = (size_t)pr * s.N + f;
    out[o] = cmul(in[o], make_float2(d.x * inv, -d.y * inv));
  }
}

__global__ void k_krylov_diag(const double2* __restrict__ C, int nx, int ny, int ncol, const double2* __restrict__ tri,
                              double lsign, int conj_c, int col, float2* __restrict__ diag) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = (int64_t)nx * ny;
  if (f >= N) return;
  const int j = (int)(f / nx), i = (int)(f - (int64_t)j * nx);
  const int type = j == 0 ? 0 : (j == ny - 1 ? 2 : 1);
  double2 d = tri[((size_t)type * 3 + 1) * nx + i];
  if (col >= 0 && col < ncol) {
    double2 c = C[(size_t)f * ncol + col];
    if (conj_c) c = zconj(c);
    d = zadd(d, zscale(c, lsign));
  }
  diag[f] = to_f2(d);
}

// u_s = 2 x - u_{s-1} (or x - u_{s-1} with sigma form)
__global__ void k_cn_finish(int64_t total, const float2* __restrict__ x, const float2* __restrict__ uprev,
                            float2* __restrict__ out, float xs) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < total) out[f] = csub(cscale(x[f], xs), uprev[f]);
}

__global__ void k_fill_zero_d(double* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}
__global__ void k_fill_zero_i(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

struct KWork {
  float2 *r, *rh, *p, *v, *s, *t, *ph, *sh;
  double *acc, *st;
  int* done;
  void* cnn;
};

static size_t krylov_layout(int P, int64_t N, int max_iter, int precond, int nx, int ny, KWork* w, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += al256(bytes);
    return q;
  };
  const size_t vb = (size_t)P * N * sizeof(float2);
  KWork k;
  k.r = (float2*)take(vb);
  k.rh = (float2*)take(vb);
  k.p = (float2*)take(vb);
  k.v = (float2*)take(vb);
  k.s = (float2*)take(vb);
  k.t = (float2*)take(vb);
  k.ph = precond ? (float2*)take(vb) : nullptr;
  k.sh = precond ? (float2*)take(vb) : nullptr;
  k.acc = (double*)take((size_t)max_iter * P * A_NACC * sizeof(double));
  k.st = (double*)take((size_t)P * S_NS * sizeof(double));
  k.done = (int*)take((size_t)P * sizeof(int));
  k.cnn = precond == 2 ? (void*)take(cnn_work_bytes_internal(P, nx, ny)) : nullptr;
  if (w) *w = k;
  return off;
}

}  // namespace tp

using namespace tp;

extern "C" int tp_krylov_diag(const void* d_C, int nx, int ny, int ncol, const void* d_tri, double lhs_sign,
                              int conj_c, int col, void* d_diag, void* stream) {
  if (!d_C || !d_tri || !d_diag || nx < 1 || ny < 1 || ncol < 1) {
    set_error("tp_krylov_diag: invalid arguments");
    return TP_ERR_INVALID;
  }
  const int64_t N = (int64_t)nx * ny;
  hipLaunchKernelGGL(k_krylov_diag, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)d_C, nx, ny, ncol, (const double2*)d_tri, lhs_sign, conj_c, col, (float2*)d_diag);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int64_t tp_bicgstab_workspace_bytes(int P, int nx, int ny, int max_iter, int precond) {
  return (int64_t)krylov_layout(P, (int64_t)nx * ny, max_iter, precond, nx, ny, nullptr, nullptr);
}

extern "C" int tp_bicgstab_solve(const void* d_diag, const void* d_tri32, const void* d_cpl32, int nx, int ny, int P,
                                 const void* d_b, void* d_x, int precond, const void* d_cnn_w, double tol,
                                 int max_iter, int check_every, void* d_work, int* d_done, void* stream) {
  if (!d_diag || !d_tri32 || !d_cpl32 || !d_b || !d_x || !d_work || nx < 1 || ny < 1 || P < 1 || max_iter < 0 ||
      precond < 0 || precond > 2 || (precond == 2 && !d_cnn_w)) {
    set_error("tp_bicgstab_solve: invalid arguments (P=%d, nx=%d, ny=%d, max_iter=%d, precond=%d)", P, nx, ny,
              max_iter, precond);
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = (int64_t)nx * ny;
  KWork w;
  krylov_layout(P, N, max_iter > 0 ? max_iter : 1, precond, nx, ny, &w, (char*)d_work);
  KOp op;
  op.diag = (const float2*)d_diag;
  op.tri = (const float2*)d_tri32;
  op.cpl = (const float2*)d_cpl32;
  op.nx = nx;
  op.ny = ny;
  KState s;
  s.acc = w.acc;
  s.st = w.st;
  s.done = w.done;
  s.P = P;
  s.N = N;
  s.tol2 = tol * tol;
  hipLaunchKernelGGL(k_fill_zero_d, dim3(256), dim3(256), 0, st, w.acc, (size_t)(max_iter > 0 ? max_iter : 1) * P * A_NACC);
  hipLaunchKernelGGL(k_fill_zero_d, dim3(1), dim3(256), 0, st, w.st, (size_t)P * S_NS);
  hipLaunchKernelGGL(k_fill_zero_i, dim3((P + 255) / 256), dim3(256), 0, st, w.done, P);
  const dim3 grid((unsigned)((N + 255) / 256), (unsigned)P), blk(256);
  float2* x = (float2*)d_x;
  hipLaunchKernelGGL(k_init, grid, blk, 0, st, op, s, (const float2*)d_b, (const float2*)x, w.r, w.rh, w.p);
  float2* ph = precond ? w.ph : w.p;
  float2* sh = precond ? w.sh : w.s;
  std::vector<int> h_done(check_every > 0 ? P : 0);
  int rc = TP_OK;
  for (int k = 0; k < max_iter; ++k) {
    if (precond == 1) hipLaunchKernelGGL(k_jacobi, grid, blk, 0, st, s, op.diag, (const float2*)w.p, ph);
    if (precond == 2) {
      rc = cnn_apply_internal(d_cnn_w, w.p, ph, P, nx, ny, w.cnn, st, w.done);
      if (rc) break;
    }
    hipLaunchKernelGGL(k_spmv_v, grid, blk, 0, st, op, s, k, (const float2*)ph, w.v, (const float2*)w.rh);
    hipLaunchKernelGGL(k_s, grid, blk, 0, st, s, k, (const float2*)w.r, (const float2*)w.v, w.s);
    if (precond == 1) hipLaunchKernelGGL(k_jacobi, grid, blk, 0, st, s, op.diag, (const float2*)w.s, sh);
    if (precond == 2) {
      rc = cnn_apply_internal(d_cnn_w, w.s, sh, P, nx, ny, w.cnn, st, w.done);
      if (rc) break;
    }
    hipLaunchKernelGGL(k_spmv_t, grid, blk, 0, st, op, s, k, (const float2*)sh, w.t, (const float2*)w.s);
    hipLaunchKernelGGL(k_xr, grid, blk, 0, st, s, k, x, w.r, (const float2*)ph, (const float2*)sh, (const float2*)w.s,
                       (const float2*)w.t, (const float2*)w.rh);
    if (k + 1 < max_iter) {
      // p update needs rho_{k+1}: computed before the status so that converged probes skip it
      hipLaunchKernelGGL(k_p, grid, blk, 0, st, s, k, w.p, (const float2*)w.r, (const float2*)w.v);
    }
    hipLaunchKernelGGL(k_status, dim3((P + 255) / 256), dim3(256), 0, st, s, k);
    if (check_every > 0 && ((k + 1) % check_every == 0 || k + 1 == max_iter)) {
      hipError_t e = hipMemcpyAsync(h_done.data(), w.done, (size_t)P * sizeof(int), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) {
        rc = hip_status(e, "tp_bicgstab_solve: status read");
        break;
      }
      bool all = true;
      for (int q = 0; q < P; ++q) all = all && h_done[q] != 0;
      if (all) break;
    }
  }
  if (rc) return rc;
  if (d_done) TP_CHECK_HIP(hipMemcpyAsync(d_done, w.done, (size_t)P * sizeof(int), hipMemcpyDeviceToDevice, st));
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_cn_finish(int64_t total, const void* d_x, const void* d_uprev, void* d_out, double xscale,
                            void* stream) {
  if (!d_x || !d_uprev || !d_out || total < 0) {
    set_error("tp_cn_finish: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (total == 0) return TP_OK;
  hipLaunchKernelGGL(k_cn_finish, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     total, (const float2*)d_x, (const float2*)d_uprev, (float2*)d_out, (float)xscale);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
