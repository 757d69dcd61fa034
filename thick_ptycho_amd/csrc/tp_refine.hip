// Iterative refinement of the forward CN sweep, and the loss seed of a refined
// exit plane.
//
// The complex64 sweep gives planes u^_s with A_s u^_s = B_s u^_{s-1} + rounding.
// With B_s = 2I - A_s (sigma = 2) the residual of the FP64 recurrence at an
// iterate v is
//     r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s),
// a small difference of O(|v|) terms: tp_tri_residual forms it in FP64 from the
// iterate held as a sum of up to three complex64 plane sets (u^, and the
// corrections of earlier refinement steps) and the FP64 operator.  A complex64
// sweep of the same factors on r, started from the probe's own complex64
// rounding error (first step) or zero, gives the next correction; each step
// multiplies the exit-plane error by about the complex64 sweep's relative
// error (2e-5 after 511 slices), so two steps reach FP64 level.
//
// tp_amplitude_seed_refined forms the amplitude-misfit seed
// scale * sum_k W_k exp(+2 pi i j k / n), W = (|U| - a) U / |U|, U = fft(x), from
// an exit plane x given as up to three complex64 parts and amplitudes a given
// as float32 hi + lo, with both transforms in FP64 (radix-2 in LDS, n a power
// of two up to 8192).  The amplitude loss needs this: where the model's far
// field is many orders of magnitude below its peak and the data are not, the
// seed takes the phase of U there, and errors of 1e-15 of the peak already
// move the gradient by ~1e-5.
#include "tp_common.h"

namespace tp {

// ------------------------------------------------------------------ residual
struct ResidualArgs {
  const float2* u[3];  // parts of the iterate (u[1], u[2] may be null)
  int64_t u_pstride, u_zstride;
  const double2* C;
  const double2* base;
  float2* r;
  int64_t r_pstride, r_sstride;
  double lsign;
  int probes, n, steps, ncol, conj_c, lcol0, cstep, s_first;
};

__device__ __forceinline__ double2 plane_sum(const ResidualArgs& a, int64_t off) {
  const float2 x = a.u[0][off];
  double2 v = make_double2(x.x, x.y);
#pragma unroll
  for (int k = 1; k < 3; ++k)
    if (a.u[k]) {
      const float2 y = a.u[k][off];
      v = zadd(v, make_double2(y.x, y.y));
    }
  return v;
}

__global__ void tri_residual_kernel(const ResidualArgs a) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_p = (int64_t)a.steps * a.n;
  if (id >= per_p * a.probes) return;
  const int p = (int)(id / per_p);
  const int64_t rem = id % per_p;
  const int t = (int)(rem / a.n) + 1;  // local step 1..steps
  const int i = (int)(rem % a.n);
  const int n = a.n;
  const int64_t prev = (int64_t)p * a.u_pstride + (int64_t)(t - 1) * a.u_zstride;  // v_{s-1}
  const int64_t cur = prev + a.u_zstride;                                          // v_s
  auto w_at = [&](int j) { return zadd(plane_sum(a, prev + j), plane_sum(a, cur + j)); };
  const int col = a.lcol0 + a.cstep * (a.s_first + t - 1);
  double2 c = (col >= 0 && col < a.ncol) ? a.C[(size_t)i * a.ncol + col] : make_double2(0.0, 0.0);
  if (a.conj_c) c = zconj(c);
  const double2 d = zadd(a.base[n + i], zscale(c, a.lsign));
  double2 acc = zmul(d, w_at(i));
  if (i > 0) acc = zadd(acc, zmul(a.base[i], w_at(i - 1)));
  if (i < n - 1) acc = zadd(acc, zmul(a.base[2 * n + i], w_at(i + 1)));
  const double2 r = zsub(zscale(plane_sum(a, prev + i), 2.0), acc);
  a.r[(size_t)p * a.r_pstride + (int64_t)(t - 1) * a.r_sstride + i] = to_f2(r);
}

// ------------------------------------------------------------------ FP64 FFT in LDS
__device__ __forceinline__ int rf_rev(int i, int lg) { return (int)(__brev((unsigned)i) >> (32 - lg)); }

// tab[e] = (cos, sin)(2 pi e / M), e <= M/8
__device__ __forceinline__ void rf_build_table(double2* tab, int M, int tid, int T) {
  for (int e = tid; e <= (M >> 3); e += T) {
    double s, c;
    sincospi(2.0 * (double)e / (double)M, &s, &c);
    tab[e] = make_double2(c, s);
  }
}

// exp(sign 2 pi i e / M), 0 <= e < M/2, from the octant table
__device__ __forceinline__ double2 rf_twiddle(const double2* tab, int e, int M, double sign) {
  const int q = M >> 2, o = M >> 3;
  const bool hi = e >= q;
  const int r = hi ? e - q : e;
  const double2 t = (r <= o) ? tab[r] : make_double2(tab[q - r].y, tab[q - r].x);
  const double c = hi ? -t.y : t.x;
  const double s = hi ? t.x : t.y;
  return make_double2(c, sign * s);
}

__device__ __forceinline__ void rf_dit(double2* buf, const double2* tab, int M, int lg, int tid, int T) {
  for (int s = 0; s < lg; ++s) {
    const int h = 1 << s, shift = lg - 1 - s;
    for (int b = tid; b < (M >> 1); b += T) {
      const int r = b & (h - 1);
      const int j = ((b >> s) << (s + 1)) | r;
      const double2 w = rf_twiddle(tab, r << shift, M, -1.0);
      const double2 a0 = buf[j];
      const double2 a1 = zmul(buf[j + h], w);
      buf[j] = zadd(a0, a1);
      buf[j + h] = zsub(a0, a1);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void rf_dif(double2* buf, const double2* tab, int M, int lg, int tid, int T) {
  for (int s = lg - 1; s >= 0; --s) {
    const int h = 1 << s, shift = lg - 1 - s;
    for (int b = tid; b < (M >> 1); b += T) {
      const int r = b & (h - 1);
      const int j = ((b >> s) << (s + 1)) | r;
      const double2 w = rf_twiddle(tab, r << shift, M, 1.0);
      const double2 a0 = buf[j];
      const double2 a1 = buf[j + h];
      buf[j] = zadd(a0, a1);
      buf[j + h] = zmul(zsub(a0, a1), w);
    }
    __syncthreads();
  }
}

struct SeedArgs {
  const float2* x[3];  // parts of the exit plane (x[1], x[2] may be null)
  int64_t ld;
  const float* amp;
  const float* amp_lo;  // optional low part of the amplitudes
  float2* seed;
  float* loss;
  double scale;
  int n, lg, probes;
};

__global__ __launch_bounds__(1024) void amplitude_seed_refined_kernel(const SeedArgs a) {
  extern __shared__ double2 rsm[];
  const int n = a.n, lg = a.lg, T = blockDim.x, tid = threadIdx.x;
  double2* buf = rsm;
  double2* tab = rsm + n;
  double* red = reinterpret_cast<double*>(tab + (n >> 3) + 1);
  const int p = blockIdx.x;
  rf_build_table(tab, n, tid, T);
  for (int j = tid; j < n; j += T) {
    double2 v = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (a.x[k]) {
        const float2 y = a.x[k][(size_t)p * a.ld + j];
        v = zadd(v, make_double2(y.x, y.y));
      }
    buf[rf_rev(j, lg)] = v;
  }
  __syncthreads();
  rf_dit(buf, tab, n, lg, tid, T);
  const float* am = a.amp + (size_t)p * n;
  const float* al = a.amp_lo ? a.amp_lo + (size_t)p * n : nullptr;
  double part = 0.0;
  for (int k = tid; k < n; k += T) {
    const double2 U = buf[k];
    const double mag = sqrt(U.x * U.x + U.y * U.y);
    const double dlt = mag - ((double)am[k] + (al ? (double)al[k] : 0.0));
    part += dlt * dlt;
    buf[k] = mag > 0.0 ? zscale(U, a.scale * dlt / mag) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  rf_dif(buf, tab, n, lg, tid, T);
  float2* sd = a.seed + (size_t)p * n;
  for (int j = tid; j < n; j += T) sd[j] = to_f2(buf[rf_rev(j, lg)]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < (T >> 6); ++w) t += red[w];
    a.loss[p] = (float)t;
  }
}

// ------------------------------------------------------------------ two transverse axes
// r_s = 2 v_{s-1} - A_s (v_{s-1} + v_s) for the block-tridiagonal operator of
// two transverse axes: unknown (j, i) at j * nx + i, diagonal block of type
// 0 (j = 0), 2 (j = ny - 1) or 1 with the tridiagonal (sub, diag, sup) of tri,
// couplings f_j to block j - 1 and e_j to block j + 1, plus lsign * c_s on the
// diagonal.  Grid: x over the N unknowns, y over probes, z over steps, so the
// probes of one step run together and share that step's column of C in L2.
struct BandResidualArgs {
  const float2* u[3];
  int64_t u_pstride, u_zstride;
  const double2* C;
  int64_t c_istride, c_cstride;
  const double2* tri;  // [3][3][nx]
  const double2* cpl;  // [ny][2]: (f_j, e_j)
  float2* r;
  int64_t r_pstride, r_sstride;
  double lsign;
  int nx, ny, ncol, conj_c, lcol0, cstep, s_first;
};

__device__ __forceinline__ double2 band_plane_sum(const BandResidualArgs& a, int64_t off) {
  const float2 x = a.u[0][off];
  double2 v = make_double2(x.x, x.y);
#pragma unroll
  for (int k = 1; k < 3; ++k)
    if (a.u[k]) {
      const float2 y = a.u[k][off];
      v = zadd(v, make_double2(y.x, y.y));
    }
  return v;
}

__global__ __launch_bounds__(256) void band_residual_kernel(const BandResidualArgs a) {
  const int nx = a.nx, ny = a.ny;
  const int64_t N = (int64_t)nx * ny;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N) return;
  const int p = blockIdx.y;
  const int t = blockIdx.z + 1;  // local step 1..steps
  const int j = (int)(idx / nx), i = (int)(idx % nx);
  const int64_t prev = (int64_t)p * a.u_pstride + (int64_t)(t - 1) * a.u_zstride;
  const int64_t cur = prev + a.u_zstride;
  auto w_at = [&](int64_t q) { return zadd(band_plane_sum(a, prev + q), band_plane_sum(a, cur + q)); };
  const int col = a.lcol0 + a.cstep * (a.s_first + t - 1);
  double2 c = (col >= 0 && col < a.ncol) ? a.C[idx * a.c_istride + (int64_t)col * a.c_cstride] : make_double2(0.0, 0.0);
  if (a.conj_c) c = zconj(c);
  const int type = j == 0 ? 0 : (j == ny - 1 ? 2 : 1);
  const double2* tb = a.tri + (size_t)type * 3 * nx;
  const double2 d = zadd(tb[nx + i], zscale(c, a.lsign));
  double2 acc = zmul(d, w_at(idx));
  if (i > 0) acc = zadd(acc, zmul(tb[i], w_at(idx - 1)));
  if (i < nx - 1) acc = zadd(acc, zmul(tb[2 * nx + i], w_at(idx + 1)));
  if (j > 0) acc = zadd(acc, zmul(a.cpl[2 * j], w_at(idx - nx)));
  if (j < ny - 1) acc = zadd(acc, zmul(a.cpl[2 * j + 1], w_at(idx + nx)));
  const double2 r = zsub(zscale(band_plane_sum(a, prev + idx), 2.0), acc);
  a.r[(size_t)p * a.r_pstride + (int64_t)(t - 1) * a.r_sstride + idx] = to_f2(r);
}

// ------------------------------------------------------------------ FP64 2D seed
// The amplitude-misfit seed of a refined exit plane with two transverse axes:
// U = fft2(x) of the plane x (nx rows of ny values, x = x0 + x1 + x2),
// W = (|U| - a) U / |U|, seed = scale * (unnormalised inverse fft2 of W), loss =
// sum (|U| - a)^2, all in FP64, in three passes through an FP64 work plane:
// rows (along y), columns (along x: forward, seed weights, inverse), rows
// (inverse along y).  A workgroup takes `nseq` sequences of length M side by
// side in LDS.
__device__ void rf_dit_multi(double2* buf, const double2* tab, int M, int lg, int nseq, int tid, int T) {
  const int half = M >> 1;
  for (int s = 0; s < lg; ++s) {
    const int h = 1 << s, shift = lg - 1 - s;
    for (int b = tid; b < nseq * half; b += T) {
      const int q = b / half, bb = b - q * half;
      const int r = bb & (h - 1);
      const int j = q * M + (((bb >> s) << (s + 1)) | r);
      const double2 w = rf_twiddle(tab, r << shift, M, -1.0);
      const double2 a0 = buf[j];
      const double2 a1 = zmul(buf[j + h], w);
      buf[j] = zadd(a0, a1);
      buf[j + h] = zsub(a0, a1);
    }
    __syncthreads();
  }
}

__device__ void rf_dif_multi(double2* buf, const double2* tab, int M, int lg, int nseq, int tid, int T) {
  const int half = M >> 1;
  for (int s = lg - 1; s >= 0; --s) {
    const int h = 1 << s, shift = lg - 1 - s;
    for (int b = tid; b < nseq * half; b += T) {
      const int q = b / half, bb = b - q * half;
      const int r = bb & (h - 1);
      const int j = q * M + (((bb >> s) << (s + 1)) | r);
      const double2 w = rf_twiddle(tab, r << shift, M, 1.0);
      const double2 a0 = buf[j];
      const double2 a1 = buf[j + h];
      buf[j] = zadd(a0, a1);
      buf[j + h] = zmul(zsub(a0, a1), w);
    }
    __syncthreads();
  }
}

struct Seed2dArgs {
  const float2* x[3];
  int64_t ld;  // probe stride of the parts (elements)
  const float* amp;
  const float* amp_lo;
  float2* seed;
  float* loss;
  double2* work;    // [probes][nx][ny]
  double* partial;  // [probes][ny / cols]
  double scale;
  int nx, ny, lgx, lgy, probes, rows, cols;  // rows: rows per row-pass workgroup; cols: columns per column workgroup
};

constexpr int RF2_THREADS = 256;

// pass 1: x rows, FFT along y (bit-reversed placement, DIT), natural order to work
__global__ __launch_bounds__(RF2_THREADS) void seed2d_rows_fwd_kernel(const Seed2dArgs a) {
  extern __shared__ double2 rsm[];
  const int ny = a.ny, R = a.rows, tid = threadIdx.x, T = blockDim.x;
  double2* buf = rsm;
  double2* tab = rsm + (size_t)R * ny;
  const int groups = a.nx / R;
  const int p = blockIdx.x / groups, x0 = (blockIdx.x % groups) * R;
  rf_build_table(tab, ny, tid, T);
  for (int l = tid; l < R * ny; l += T) {
    const int q = l / ny, y = l - q * ny;
    const int64_t off = (int64_t)p * a.ld + (int64_t)(x0 + q) * ny + y;
    double2 v = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (a.x[k]) {
        const float2 e = a.x[k][off];
        v = zadd(v, make_double2(e.x, e.y));
      }
    buf[q * ny + rf_rev(y, a.lgy)] = v;
  }
  __syncthreads();
  rf_dit_multi(buf, tab, ny, a.lgy, R, tid, T);
  double2* wk = a.work + ((size_t)p * a.nx + x0) * ny;
  for (int l = tid; l < R * ny; l += T) wk[l] = buf[l];
}

// pass 2: `cols` columns, FFT along x, weights and loss, inverse along x, back to work
__global__ __launch_bounds__(RF2_THREADS) void seed2d_cols_kernel(const Seed2dArgs a) {
  extern __shared__ double2 rsm[];
  const int nx = a.nx, ny = a.ny, CB = a.cols, tid = threadIdx.x, T = blockDim.x;
  double2* buf = rsm;  // [CB][nx]
  double2* tab = rsm + (size_t)CB * nx;
  double* red = reinterpret_cast<double*>(tab + (nx >> 3) + 1);
  const int groups = ny / CB;
  const int p = blockIdx.x / groups, g = blockIdx.x % groups, y0 = g * CB;
  rf_build_table(tab, nx, tid, T);
  double2* wk = a.work + (size_t)p * nx * ny;
  for (int l = tid; l < nx * CB; l += T) {  // consecutive threads read consecutive y
    const int x = l / CB, c = l - x * CB;
    buf[c * nx + rf_rev(x, a.lgx)] = wk[(size_t)x * ny + y0 + c];
  }
  __syncthreads();
  rf_dit_multi(buf, tab, nx, a.lgx, CB, tid, T);
  const float* am = a.amp + (size_t)p * nx * ny;
  const float* al = a.amp_lo ? a.amp_lo + (size_t)p * nx * ny : nullptr;
  double part = 0.0;
  for (int l = tid; l < nx * CB; l += T) {
    const int c = l / nx, x = l - c * nx;
    const size_t k = (size_t)x * ny + y0 + c;
    const double2 U = buf[l];
    const double mag = sqrt(U.x * U.x + U.y * U.y);
    const double dlt = mag - ((double)am[k] + (al ? (double)al[k] : 0.0));
    part += dlt * dlt;
    buf[l] = mag > 0.0 ? zscale(U, a.scale * dlt / mag) : make_double2(0.0, 0.0);
  }
  __syncthreads();
  rf_dif_multi(buf, tab, nx, a.lgx, CB, tid, T);
  for (int l = tid; l < nx * CB; l += T) {
    const int x = l / CB, c = l - x * CB;
    wk[(size_t)x * ny + y0 + c] = buf[c * nx + rf_rev(x, a.lgx)];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < (T >> 6); ++w) t += red[w];
    a.partial[(size_t)p * groups + g] = t;
  }
}

// pass 3: rows, inverse FFT along y (DIF on natural order), seed in complex64;
// the first workgroup of each probe sums the loss partials in a fixed order
__global__ __launch_bounds__(RF2_THREADS) void seed2d_rows_inv_kernel(const Seed2dArgs a) {
  extern __shared__ double2 rsm[];
  const int ny = a.ny, R = a.rows, tid = threadIdx.x, T = blockDim.x;
  double2* buf = rsm;
  double2* tab = rsm + (size_t)R * ny;
  const int groups = a.nx / R;
  const int p = blockIdx.x / groups, x0 = (blockIdx.x % groups) * R;
  rf_build_table(tab, ny, tid, T);
  const double2* wk = a.work + ((size_t)p * a.nx + x0) * ny;
  for (int l = tid; l < R * ny; l += T) buf[l] = wk[l];
  __syncthreads();
  rf_dif_multi(buf, tab, ny, a.lgy, R, tid, T);
  float2* sd = a.seed + (size_t)p * a.nx * ny + (size_t)x0 * ny;
  for (int l = tid; l < R * ny; l += T) {
    const int q = l / ny, y = l - q * ny;
    sd[l] = to_f2(buf[q * ny + rf_rev(y, a.lgy)]);
  }
  if (x0 == 0 && tid == 0) {
    const int cg = ny / a.cols;
    double t = 0.0;
    for (int g = 0; g < cg; ++g) t += a.partial[(size_t)p * cg + g];
    a.loss[p] = (float)t;
  }
}

}  // namespace tp

using namespace tp;

extern "C" int tp_band_residual(const void* d_u, const void* d_u1, const void* d_u2, int64_t u_pstride,
                                int64_t u_zstride, int probes, int nx, int ny, int steps, const void* d_C,
                                int64_t c_istride, int64_t c_cstride, int ncol, const void* d_tri, const void* d_cpl,
                                double lhs_sign, int conj_c, int lhs_col0, int col_step, int s_first, void* d_r,
                                int64_t r_pstride, int64_t r_sstride, void* stream) {
  if (!d_u || !d_C || !d_tri || !d_cpl || !d_r || probes < 0 || nx < 2 || ny < 2 || steps < 1 || ncol < 1 ||
      probes > 65535 || steps > 65535) {
    set_error("tp_band_residual: invalid arguments (nx=%d, ny=%d, steps=%d, probes=%d)", nx, ny, steps, probes);
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  BandResidualArgs a;
  a.u[0] = (const float2*)d_u;
  a.u[1] = (const float2*)d_u1;
  a.u[2] = (const float2*)d_u2;
  a.u_pstride = u_pstride;
  a.u_zstride = u_zstride;
  a.C = (const double2*)d_C;
  a.c_istride = c_istride;
  a.c_cstride = c_cstride;
  a.tri = (const double2*)d_tri;
  a.cpl = (const double2*)d_cpl;
  a.r = (float2*)d_r;
  a.r_pstride = r_pstride;
  a.r_sstride = r_sstride;
  a.lsign = lhs_sign;
  a.nx = nx;
  a.ny = ny;
  a.ncol = ncol;
  a.conj_c = conj_c;
  a.lcol0 = lhs_col0;
  a.cstep = col_step;
  a.s_first = s_first;
  const int64_t N = (int64_t)nx * ny;
  hipLaunchKernelGGL(band_residual_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)probes, (unsigned)steps),
                     dim3(256), 0, (hipStream_t)stream, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

static int pow2_log(int n) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  return (1 << lg) == n ? lg : -1;
}

extern "C" size_t tp_amplitude_seed2d_refined_work_bytes(int nx, int ny, int probes) {
  if (nx < 1 || ny < 1 || probes < 0) return 0;
  return (size_t)probes * nx * ny * sizeof(double2) + (size_t)probes * ny * sizeof(double);
}

extern "C" int tp_amplitude_seed2d_refined(const void* d_x0, const void* d_x1, const void* d_x2, int64_t ld,
                                           const void* d_sqrt_i, const void* d_sqrt_i_lo, int nx, int ny, int probes,
                                           double scale, void* d_seed, void* d_loss, void* d_work, void* stream) {
  const int lgx = pow2_log(nx), lgy = pow2_log(ny);
  if (lgx < 3 || lgy < 3 || nx > 2048 || ny > 2048) {
    set_error("tp_amplitude_seed2d_refined: plane %d x %d not supported (powers of two from 8 to 2048)", nx, ny);
    return TP_ERR_UNSUPPORTED;
  }
  if (!d_x0 || !d_sqrt_i || !d_seed || !d_loss || !d_work || probes < 0 || ld < (int64_t)nx * ny) {
    set_error("tp_amplitude_seed2d_refined: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  Seed2dArgs a;
  a.x[0] = (const float2*)d_x0;
  a.x[1] = (const float2*)d_x1;
  a.x[2] = (const float2*)d_x2;
  a.ld = ld;
  a.amp = (const float*)d_sqrt_i;
  a.amp_lo = (const float*)d_sqrt_i_lo;
  a.seed = (float2*)d_seed;
  a.loss = (float*)d_loss;
  a.work = (double2*)d_work;
  a.partial = (double*)((double2*)d_work + (size_t)probes * nx * ny);
  a.scale = scale;
  a.nx = nx;
  a.ny = ny;
  a.lgx = lgx;
  a.lgy = lgy;
  a.probes = probes;
  // about 2048 points (32 KiB of FP64 values) per workgroup
  a.rows = ny >= 2048 ? 1 : (2048 / ny < nx ? 2048 / ny : nx);
  a.cols = nx >= 2048 ? 1 : (2048 / nx < ny ? 2048 / nx : ny);
  hipStream_t st = (hipStream_t)stream;
  const size_t lds_rows = ((size_t)a.rows * ny + (ny >> 3) + 1) * sizeof(double2);
  const size_t lds_cols = ((size_t)a.cols * nx + (nx >> 3) + 1) * sizeof(double2) + 4 * sizeof(double);
  hipLaunchKernelGGL(seed2d_rows_fwd_kernel, dim3((unsigned)(probes * (nx / a.rows))), dim3(RF2_THREADS), lds_rows, st,
                     a);
  hipLaunchKernelGGL(seed2d_cols_kernel, dim3((unsigned)(probes * (ny / a.cols))), dim3(RF2_THREADS), lds_cols, st, a);
  hipLaunchKernelGGL(seed2d_rows_inv_kernel, dim3((unsigned)(probes * (nx / a.rows))), dim3(RF2_THREADS), lds_rows, st,
                     a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tri_residual(const void* d_u, const void* d_u1, const void* d_u2, int64_t u_pstride,
                               int64_t u_zstride, int probes, int n, int steps, const void* d_C, int ncol,
                               const void* d_base, double lhs_sign, int conj_c, int lhs_col0, int col_step,
                               int s_first, void* d_r, int64_t r_pstride, int64_t r_sstride, void* stream) {
  if (!d_u || !d_C || !d_base || !d_r || probes < 0 || n < 1 || steps < 1 || ncol < 1) {
    set_error("tp_tri_residual: invalid arguments (n=%d, steps=%d, probes=%d)", n, steps, probes);
    return TP_ERR_INVALID;
  }
  ResidualArgs a;
  a.u[0] = (const float2*)d_u;
  a.u[1] = (const float2*)d_u1;
  a.u[2] = (const float2*)d_u2;
  a.u_pstride = u_pstride;
  a.u_zstride = u_zstride;
  a.C = (const double2*)d_C;
  a.base = (const double2*)d_base;
  a.r = (float2*)d_r;
  a.r_pstride = r_pstride;
  a.r_sstride = r_sstride;
  a.lsign = lhs_sign;
  a.probes = probes;
  a.n = n;
  a.steps = steps;
  a.ncol = ncol;
  a.conj_c = conj_c;
  a.lcol0 = lhs_col0;
  a.cstep = col_step;
  a.s_first = s_first;
  const int64_t total = (int64_t)probes * steps * n;
  if (total == 0) return TP_OK;
  hipLaunchKernelGGL(tri_residual_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_amplitude_seed_refined(const void* d_x0, const void* d_x1, const void* d_x2, int64_t ld,
                                         const void* d_sqrt_i, const void* d_sqrt_i_lo, int n, int probes,
                                         double scale, void* d_seed, void* d_loss, void* stream) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  if (n < 64 || n > 8192 || (1 << lg) != n) {
    set_error("tp_amplitude_seed_refined: n = %d must be a power of two from 64 to 8192", n);
    return TP_ERR_UNSUPPORTED;
  }
  if (!d_x0 || !d_sqrt_i || !d_seed || !d_loss || probes < 0 || ld < n) {
    set_error("tp_amplitude_seed_refined: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  SeedArgs a;
  a.x[0] = (const float2*)d_x0;
  a.x[1] = (const float2*)d_x1;
  a.x[2] = (const float2*)d_x2;
  a.ld = ld;
  a.amp = (const float*)d_sqrt_i;
  a.amp_lo = (const float*)d_sqrt_i_lo;
  a.seed = (float2*)d_seed;
  a.loss = (float*)d_loss;
  a.scale = scale;
  a.n = n;
  a.lg = lg;
  a.probes = probes;
  const int T = n / 4 < 64 ? 64 : (n / 4 > 1024 ? 1024 : n / 4);
  const size_t lds = ((size_t)n + (n >> 3) + 1) * sizeof(double2) + 16 * sizeof(double);
  if (lds > 65536)
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)amplitude_seed_refined_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(amplitude_seed_refined_kernel, dim3(probes), dim3(T), lds, (hipStream_t)stream, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
