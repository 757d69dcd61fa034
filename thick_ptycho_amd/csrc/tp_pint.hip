// Whole space-z system of the CN stepping, applied matrix-free (the operator of
// the parallel-in-time solve):
//
//   (M x)_j = (A - diag C_j) x_j - (B + diag C_j) x_{j-1}     (x_{-1} = 0)
//
// for slices j = 0..L-1 of every probe, A and B tridiagonal N x N (the step
// operators, or their conjugate transposes in adjoint mode).  Vectors are
// (P, L, N) with the transverse index fastest, so one slice of one probe is a
// contiguous row and a thread per (probe, slice, row) reads its three-point
// stencils from the same and the previous row of the wavefront's neighbours.
// HBM-bound: per element it reads x once (the neighbours and the previous slice
// come from L1/L2) and C (from L2 when the probes share it) and writes y.  FP64 for the residual of the
// refinement loop, FP32 inside the Krylov cycles.
#include "tp_common.h"

namespace tp {
namespace {

__device__ __forceinline__ float2 zero_of(float2) { return make_float2(0.f, 0.f); }
__device__ __forceinline__ double2 zero_of(double2) { return make_double2(0.0, 0.0); }
__device__ __forceinline__ float2 mac(float2 a, float2 b, float2 c) { return cfma(a, b, c); }
__device__ __forceinline__ double2 mac(double2 a, double2 b, double2 c) { return zadd(zmul(a, b), c); }
__device__ __forceinline__ float2 sub2(float2 a, float2 b) { return csub(a, b); }
__device__ __forceinline__ double2 sub2(double2 a, double2 b) { return zsub(a, b); }
__device__ __forceinline__ float2 add2(float2 a, float2 b) { return cadd(a, b); }
__device__ __forceinline__ double2 add2(double2 a, double2 b) { return zadd(a, b); }

// ops: A sub, A diag, A sup, B sub, B diag, B sup, each N values (sub[0] and
// sup[N-1] are not read).
// C: (L, N) shared by all probes (c_pstride 0) or one (L, N) block per probe
// (c_pstride = L * N: reduced-domain windows).
template <typename T>
__global__ __launch_bounds__(256) void pint_matvec_kernel(const T* __restrict__ ops, const T* __restrict__ C,
                                                          int64_t c_pstride, const T* __restrict__ x, const T* b, T* y,
                                                          int N, int L, int64_t total) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= total) return;
  const int i = (int)(t % N);
  const int64_t row = t / N;  // p * L + j
  const int j = (int)(row % L);
  const T* xj = x + row * N;
  const T c = C[(row / L) * c_pstride + (int64_t)j * N + i];
  const T xc = xj[i];
  T acc = mac(ops[N + i], xc, zero_of(xc));
  if (i > 0) acc = mac(ops[i], xj[i - 1], acc);
  if (i < N - 1) acc = mac(ops[2 * N + i], xj[i + 1], acc);
  T s = xc;  // diag C_j multiplies x_j and x_{j-1}
  if (j > 0) {
    const T* xp = xj - N;
    const T pc = xp[i];
    T q = mac(ops[4 * N + i], pc, zero_of(pc));
    if (i > 0) q = mac(ops[3 * N + i], xp[i - 1], q);
    if (i < N - 1) q = mac(ops[5 * N + i], xp[i + 1], q);
    acc = sub2(acc, q);
    s = add2(s, pc);
  }
  acc = sub2(acc, mac(c, s, zero_of(s)));
  y[t] = b ? sub2(b[t], acc) : acc;
}

// ---------------------------------------------------------------- Krylov basis
// Arnoldi's projections for a batch of probes: V is (P, rows, len) complex64
// (row i of probe p at p * v_pstride + i * v_rstride), w is (P, len).
constexpr int PINT_MAX_ROWS = 32;

// partial[(p * nblk + blk) * (PINT_MAX_ROWS + 1) + i] = sum over the block's share
// of e of conj(V[p, i, e]) w[p, e] for i < nrows, and sum |w[p, e]|^2 (in .x) in
// slot nrows.  Each wavefront reads 64 consecutive e of every row (coalesced);
// the per-block sums are reduced over blocks by the caller.
__global__ __launch_bounds__(256) void pint_dots_kernel(const float2* __restrict__ V, int64_t v_pstride,
                                                        int64_t v_rstride, int nrows, const float2* __restrict__ w,
                                                        int64_t w_pstride, int64_t len, int nblk,
                                                        float2* __restrict__ partial) {
  const int p = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x;
  const float2* wp = w + (int64_t)p * w_pstride;
  const float2* Vp = V + (int64_t)p * v_pstride;
  float2 acc[PINT_MAX_ROWS];
#pragma unroll
  for (int i = 0; i < PINT_MAX_ROWS; ++i) acc[i] = make_float2(0.f, 0.f);
  float nrm = 0.f;
  const int64_t per = ((len + nblk - 1) / nblk + 255) / 256 * 256;
  const int64_t e0 = (int64_t)blk * per;
  const int64_t e1 = e0 + per < len ? e0 + per : len;
  for (int64_t e = e0 + tid; e < e1; e += 256) {
    const float2 x = wp[e];
    nrm = fmaf(x.x, x.x, fmaf(x.y, x.y, nrm));
#pragma unroll
    for (int i = 0; i < PINT_MAX_ROWS; ++i) {
      if (i < nrows) {
        const float2 v = Vp[(int64_t)i * v_rstride + e];
        acc[i] = cfma(cconj(v), x, acc[i]);
      }
    }
  }
  __shared__ float red[4][2 * PINT_MAX_ROWS + 1];
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int i = 0; i <= PINT_MAX_ROWS; ++i) {
    if (i > nrows) break;
    const int ii = i < PINT_MAX_ROWS ? i : PINT_MAX_ROWS - 1;  // slot PINT_MAX_ROWS holds only the norm
    float a = i < nrows ? acc[ii].x : nrm;
    float b = i < nrows ? acc[ii].y : 0.f;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (lane == 0) {
      red[wave][2 * i] = a;
      if (i < PINT_MAX_ROWS) red[wave][2 * i + 1] = b;
    }
  }
  __syncthreads();
  if (tid <= nrows) {
    float a = 0.f, b = 0.f;
    for (int q = 0; q < 4; ++q) {
      a += red[q][2 * tid];
      if (tid < PINT_MAX_ROWS) b += red[q][2 * tid + 1];
    }
    partial[((int64_t)p * nblk + blk) * (PINT_MAX_ROWS + 1) + tid] = make_float2(a, b);
  }
}

// out[p, e] = w[p, e] - sum_{i < nrows} V[p, i, e] c[p, i]  (w null: 0; w may alias out).
__global__ __launch_bounds__(256) void pint_combine_kernel(const float2* __restrict__ V, int64_t v_pstride,
                                                           int64_t v_rstride, int nrows, const float2* __restrict__ c,
                                                           int c_pstride, const float2* w, int64_t w_pstride,
                                                           float2* out, int64_t out_pstride, int64_t len) {
  const int p = blockIdx.y;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= len) return;
  const float2* Vp = V + (int64_t)p * v_pstride + e;
  const float2* cp = c + (int64_t)p * c_pstride;
  float2 acc = w ? w[(int64_t)p * w_pstride + e] : make_float2(0.f, 0.f);
  for (int i = 0; i < nrows; ++i) acc = cfms(Vp[(int64_t)i * v_rstride], cp[i], acc);
  out[(int64_t)p * out_pstride + e] = acc;
}

}  // namespace
}  // namespace tp

using namespace tp;

extern "C" int tp_pint_dots(const void* d_V, int64_t v_pstride, int64_t v_rstride, int nrows, const void* d_w,
                            int64_t w_pstride, int64_t len, int P, int nblk, void* d_partial, void* stream) {
  if (!d_w || !d_partial || nrows < 0 || nrows > PINT_MAX_ROWS || (nrows > 0 && !d_V) || len < 1 || P < 0 ||
      nblk < 1 || nblk > 65535 || P > 65535) {
    set_error("tp_pint_dots: invalid arguments (nrows = %d of at most %d, len = %lld, P = %d, nblk = %d)", nrows,
              PINT_MAX_ROWS, (long long)len, P, nblk);
    return TP_ERR_INVALID;
  }
  if (P == 0) return TP_OK;
  hipLaunchKernelGGL(pint_dots_kernel, dim3(nblk, P), dim3(256), 0, (hipStream_t)stream, (const float2*)d_V,
                     v_pstride, v_rstride, nrows, (const float2*)d_w, w_pstride, len, nblk, (float2*)d_partial);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_pint_combine(const void* d_V, int64_t v_pstride, int64_t v_rstride, int nrows, const void* d_c,
                               int c_pstride, const void* d_w, int64_t w_pstride, void* d_out, int64_t out_pstride,
                               int64_t len, int P, void* stream) {
  if (!d_out || nrows < 0 || (nrows > 0 && (!d_V || !d_c)) || len < 1 || P < 0 || P > 65535 ||
      (len + 255) / 256 > 0x7fffffff) {
    set_error("tp_pint_combine: invalid arguments (nrows = %d, len = %lld, P = %d)", nrows, (long long)len, P);
    return TP_ERR_INVALID;
  }
  if (P == 0) return TP_OK;
  hipLaunchKernelGGL(pint_combine_kernel, dim3((unsigned)((len + 255) / 256), P), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)d_V, v_pstride, v_rstride, nrows, (const float2*)d_c, c_pstride,
                     (const float2*)d_w, w_pstride, (float2*)d_out, out_pstride, len);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_pint_matvec(const void* d_ops, const void* d_C, int64_t c_pstride, const void* d_x, const void* d_b,
                              void* d_y, int N, int L, int P, int fp64, void* stream) {
  if (!d_ops || !d_C || !d_x || !d_y || N < 2 || L < 1 || P < 0 || (fp64 != 0 && fp64 != 1) ||
      (c_pstride != 0 && c_pstride < (int64_t)L * N)) {
    set_error("tp_pint_matvec: invalid arguments (N = %d, L = %d, P = %d, fp64 = %d)", N, L, P, fp64);
    return TP_ERR_INVALID;
  }
  if (d_x == d_y) {  // b may alias y: each thread reads b[t] before writing y[t]
    set_error("tp_pint_matvec: x and y must not alias (y is written while neighbours of x are read)");
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)P * L * N;
  if (total == 0) return TP_OK;
  const dim3 grid((unsigned)((total + 255) / 256));
  if (fp64)
    hipLaunchKernelGGL(pint_matvec_kernel<double2>, grid, dim3(256), 0, (hipStream_t)stream, (const double2*)d_ops,
                       (const double2*)d_C, c_pstride, (const double2*)d_x, (const double2*)d_b, (double2*)d_y, N, L,
                       total);
  else
    hipLaunchKernelGGL(pint_matvec_kernel<float2>, grid, dim3(256), 0, (hipStream_t)stream, (const float2*)d_ops,
                       (const float2*)d_C, c_pstride, (const float2*)d_x, (const float2*)d_b, (float2*)d_y, N, L,
                       total);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
