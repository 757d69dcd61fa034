// Angular-spectrum multislice propagation and far-field intensities.
//
// One wavefront owns one probe's field of n = 64 * R points (R = 1..32); lane l
// keeps x[l + 64 m] (m < R) in registers.  Each slice applies
//     psi <- ifft(fft(psi * T_z) * H)
// with a Stockham radix FFT: the first stage (radix R) runs on the register
// layout, later stages exchange through a per-wavefront LDS ping-pong buffer,
// and the last stage returns the natural-order spectrum to registers so that
// H (pre-scaled by 1/n) is applied in registers and the inverse transform uses
// the same schedule with conjugate twiddles.  Only T_z is read per slice and
// only psi_{z+1} is written.
#include "tp_fft.h"

namespace tp {
struct MsArgs {
  tp_ms_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans;
};

constexpr int MS_PPB = 4;

__device__ __forceinline__ int ms_find_scan(const MsArgs& a, int blk) {
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  return sc;
}

// Field planes stream out with non-temporal stores: they are not read back by
// the sweep, and ordinary stores push the transmission and propagator out of
// the caches (split form at 512 x 1024 x 1023: 1.80e8 -> 2.51e8 steps/s).
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
}

template <int R>
__device__ __forceinline__ void ms_store(const tp_ms_scan& d, int p, int z, const float2* x, int lane) {
  if (d.out) {
    float2* o = (float2*)d.out + (size_t)p * d.out_pstride + (int64_t)z * d.out_zstride;
#pragma unroll
    for (int m = 0; m < R; ++m) store_nt(o + lane + 64 * m, x[m]);
  }
  if (d.exitp && z == d.steps) {
    float2* o = (float2*)d.exitp + (size_t)p * d.exit_pstride;
#pragma unroll
    for (int m = 0; m < R; ++m) o[lane + 64 * m] = x[m];
  }
}

// One slice: x <- ifft(fft(x * T) * H)  (H already holds the 1/n factor).
template <int R>
__device__ __forceinline__ void ms_slice(float2* x, const float2* T, const float2* H, float2* a, float2* b,
                                         const float2* tw, int lane) {
#pragma unroll
  for (int m = 0; m < R; ++m) x[m] = cmul(x[m], T[m]);
  fft_any<R, -1>(x, a, b, tw, lane);
#pragma unroll
  for (int m = 0; m < R; ++m) x[m] = cmul(x[m], H[m]);
  fft_any<R, +1>(x, a, b, tw, lane);
}

// PPB probes (one per wavefront) per workgroup: 4 when the launch has enough
// probes to give every CU a workgroup, else 1 so that the waves spread over all CUs.
template <int R, int PPB>
__global__ __launch_bounds__(64 * PPB) void ms_sweep_kernel(const MsArgs a) {
  constexpr int N = 64 * R;
  __shared__ float2 tw[N];
  __shared__ float2 buf[PPB][2][ms_bufn<R>()];
  const int blk = blockIdx.x;
  const int sc = ms_find_scan(a, blk);
  const tp_ms_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * PPB + wave;
  build_twiddles(tw, N);
  __syncthreads();
  if (p >= d.probes) return;
  float2 H[R], T[R], Tn[R], x[R];
  const float2* Hg = (const float2*)d.H;
  const float2* Tg = (const float2*)d.T;
  const float2* u0 = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    H[m] = Hg[spec_index<R>(lane, m)];
    if (d.conj_h) H[m] = cconj(H[m]);
    x[m] = u0[lane + 64 * m];
    T[m] = Tg[lane + 64 * m];
  }
  ms_store<R>(d, p, 0, x, lane);
  for (int z = 0; z < d.steps; ++z) {
    if (z + 1 < d.steps) {
#pragma unroll
      for (int m = 0; m < R; ++m) Tn[m] = Tg[(size_t)(z + 1) * N + lane + 64 * m];
    }
    ms_slice<R>(x, T, H, buf[wave][0], buf[wave][1], tw, lane);
    ms_store<R>(d, p, z + 1, x, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) T[m] = Tn[m];
  }
}

// Split form (n = 64 R, R in {8, 16}): S wavefronts share one probe, so that a
// launch of P probes keeps S P wavefronts in flight (the one-wavefront form
// leaves SIMDs idle and is bound by the latency of its LDS stages).  With
// M = n / S, n = n2 + M n1 and k = k1 + S k2,
//   X[k1 + S k2] = sum_n2 W_M^{n2 k2} [ W_n^{n2 k1} sum_n1 W_S^{n1 k1} x[n2 + M n1] ],
// so an S-point DFT over x[n2 + M n1] (n1 = 0..S-1), the twiddle W_n^{n2 k1}
// and a wave-level M-point FFT of the k1 = w sub-sequence give X[w + S k2] at
// lane/register k2 = l + 64 m of wavefront w; the inverse runs the steps
// backwards with conjugate twiddles.
//
// Between slices the field stays in the layout of the S-point DFTs: wavefront w
// keeps x[n2 + M n1] for its own n2 (n2 = w M/S + l + 64 t), which is exactly
// what its inverse S-point DFT produces and what the next slice's forward one
// takes.  So the transmission multiply, the plane stores (S runs of 64
// consecutive points per t, coalesced) and the next forward DFT need no
// exchange, and a slice crosses the wavefronts twice through LDS (two
// workgroup barriers) instead of four.
template <int S, int RM>
struct CrossTw {
  float2 fwd[RM / S][S];  // W_n^{n2 k}, n2 = w M/S + lane + 64 t
  float2 inv[RM];         // W_n^{(lane + 64 m) w}, conjugated when used
};

template <int S, int RM>
__device__ __forceinline__ void cross_tw_load(CrossTw<S, RM>& c, const float2* twN, int w, int lane) {
  constexpr int M = 64 * RM, N = S * M, QL = M / S / 64;
#pragma unroll
  for (int t = 0; t < QL; ++t) {
    const int n2 = w * (M / S) + lane + 64 * t;
#pragma unroll
    for (int k = 1; k < S; ++k) c.fwd[t][k] = twN[(n2 * k) & (N - 1)];
  }
#pragma unroll
  for (int m = 0; m < RM; ++m) c.inv[m] = twN[((lane + 64 * m) * w) & (N - 1)];
}

template <int R, int S>
__global__ __launch_bounds__(64 * S) void ms_split_kernel(const MsArgs a) {
  constexpr int N = 64 * R, M = N / S, RM = R / S, QL = M / S / 64;
  static_assert(RM == 4 && M % (64 * S) == 0, "split form: 256-point wave transforms");
  __shared__ float2 twN[N];
  __shared__ float2 twM[M];
  __shared__ float2 ta[N], tb[N];
  __shared__ float2 buf[S][2][ms_bufn<RM>()];
  const int blk = blockIdx.x;  // one probe per workgroup
  const int sc = ms_find_scan(a, blk);
  const tp_ms_scan& d = a.scans[sc];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = blk - a.blk_start[sc];
  build_twiddles(twN, N);
  build_twiddles(twM, M);
  __syncthreads();
  CrossTw<S, RM> ctw;
  cross_tw_load<S, RM>(ctw, twN, w, lane);
  Tw256 wtw;
  tw256_load(wtw, twM, lane);
  // xs[t][r] = x[r M + n2(t)], n2(t) = w M/S + lane + 64 t
  auto pos = [&](int t, int r) { return r * M + w * (M / S) + lane + 64 * t; };
  float2 H[RM], x[RM], xs[QL][S], T[QL][S], Tn[QL][S];
  const float2* Hg = (const float2*)d.H;
  const float2* Tg = (const float2*)d.T;
  const float2* u0 = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
#pragma unroll
  for (int m = 0; m < RM; ++m) {
    H[m] = Hg[w + S * (lane + 64 * m)];
    if (d.conj_h) H[m] = cconj(H[m]);
  }
#pragma unroll
  for (int t = 0; t < QL; ++t)
#pragma unroll
    for (int r = 0; r < S; ++r) {
      xs[t][r] = u0[pos(t, r)];
      T[t][r] = Tg[pos(t, r)];
    }
  float2* outp = d.out ? (float2*)d.out + (size_t)p * d.out_pstride : nullptr;
  if (outp) {
#pragma unroll
    for (int t = 0; t < QL; ++t)
#pragma unroll
      for (int r = 0; r < S; ++r) outp[pos(t, r)] = xs[t][r];
  }
  for (int z = 0; z < d.steps; ++z) {
    if (z + 1 < d.steps) {
#pragma unroll
      for (int t = 0; t < QL; ++t)
#pragma unroll
        for (int r = 0; r < S; ++r) Tn[t][r] = Tg[(size_t)(z + 1) * N + pos(t, r)];
    }
    // transmission, forward S-point DFTs and twiddles, into the k1 = w sub-sequences
#pragma unroll
    for (int t = 0; t < QL; ++t) {
      float2 v[S];
#pragma unroll
      for (int r = 0; r < S; ++r) v[r] = cmul(xs[t][r], T[t][r]);
      dft_reg<S, -1>(v);
#pragma unroll
      for (int k = 1; k < S; ++k) v[k] = cmul(v[k], ctw.fwd[t][k]);
#pragma unroll
      for (int k = 0; k < S; ++k) tb[pos(t, k)] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < RM; ++m) x[m] = tb[w * M + lane + 64 * m];
    fft_wave4_rt<-1>(x, buf[w][0], buf[w][1], wtw, lane);
#pragma unroll
    for (int m = 0; m < RM; ++m) x[m] = cmul(x[m], H[m]);
    fft_wave4_rt<+1>(x, buf[w][0], buf[w][1], wtw, lane);
#pragma unroll
    for (int m = 0; m < RM; ++m) ta[w * M + lane + 64 * m] = cmul(x[m], cconj(ctw.inv[m]));
    __syncthreads();
    // inverse S-point DFTs back to x[r M + n2]
#pragma unroll
    for (int t = 0; t < QL; ++t) {
      float2 v[S];
#pragma unroll
      for (int r = 0; r < S; ++r) v[r] = ta[pos(t, r)];
      dft_reg<S, +1>(v);
#pragma unroll
      for (int k = 0; k < S; ++k) xs[t][k] = v[k];
    }
    if (outp) {
      float2* o = outp + (int64_t)(z + 1) * d.out_zstride;
#pragma unroll
      for (int t = 0; t < QL; ++t)
#pragma unroll
        for (int r = 0; r < S; ++r) store_nt(o + pos(t, r), xs[t][r]);
    }
#pragma unroll
    for (int t = 0; t < QL; ++t)
#pragma unroll
      for (int r = 0; r < S; ++r) T[t][r] = Tn[t][r];
  }
  if (d.exitp) {
    float2* o = (float2*)d.exitp + (size_t)p * d.exit_pstride;
#pragma unroll
    for (int t = 0; t < QL; ++t)
#pragma unroll
      for (int r = 0; r < S; ++r) o[pos(t, r)] = xs[t][r];
  }
}

template <int R>
__global__ __launch_bounds__(64 * MS_PPB) void ms_step_kernel(const MsArgs a, int z) {
  constexpr int N = 64 * R;
  __shared__ float2 tw[N];
  __shared__ float2 buf[MS_PPB][2][ms_bufn<R>()];
  const int blk = blockIdx.x;
  const int sc = ms_find_scan(a, blk);
  const tp_ms_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * MS_PPB + wave;
  build_twiddles(tw, N);
  __syncthreads();
  if (p >= d.probes) return;
  float2 H[R], T[R], x[R];
  const float2* src = (z == 0) ? (const float2*)d.u0 + (size_t)p * d.u0_pstride
                               : (const float2*)d.out + (size_t)p * d.out_pstride + (int64_t)z * d.out_zstride;
  const float2* Hg = (const float2*)d.H;
  const float2* Tg = (const float2*)d.T + (size_t)z * N;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    H[m] = Hg[spec_index<R>(lane, m)];
    if (d.conj_h) H[m] = cconj(H[m]);
    x[m] = src[lane + 64 * m];
    T[m] = Tg[lane + 64 * m];
  }
  if (z == 0) ms_store<R>(d, p, 0, x, lane);
  ms_slice<R>(x, T, H, buf[wave][0], buf[wave][1], tw, lane);
  ms_store<R>(d, p, z + 1, x, lane);
}

template <int R>
__global__ __launch_bounds__(64 * MS_PPB) void farfield_kernel(const float2* __restrict__ ex, int64_t ld, int probes,
                                                               float* __restrict__ out) {
  constexpr int N = 64 * R;
  __shared__ float2 tw[N];
  __shared__ float2 buf[MS_PPB][2][ms_bufn<R>()];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * MS_PPB + wave;
  build_twiddles(tw, N);
  __syncthreads();
  if (p >= probes) return;
  float2 x[R];
#pragma unroll
  for (int m = 0; m < R; ++m) x[m] = ex[(size_t)p * ld + lane + 64 * m];
  fft_any<R, -1>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
  for (int m = 0; m < R; ++m) out[(size_t)p * N + spec_index<R>(lane, m)] = fmaf(x[m].x, x[m].x, x[m].y * x[m].y);
}

// Amplitude-misfit seed: U = fft(exit); loss = sum (|U| - a)^2;
// seed = scale * unnormalised-ifft((|U| - a) U / |U|).
template <int R>
__global__ __launch_bounds__(64 * MS_PPB) void amplitude_seed_kernel(const float2* __restrict__ ex,
                                                                     const float* __restrict__ amp, int probes,
                                                                     float scale, float2* __restrict__ seed,
                                                                     float* __restrict__ loss) {
  constexpr int N = 64 * R;
  __shared__ float2 tw[N];
  __shared__ float2 buf[MS_PPB][2][ms_bufn<R>()];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * MS_PPB + wave;
  build_twiddles(tw, N);
  __syncthreads();
  if (p >= probes) return;
  float2 x[R];
#pragma unroll
  for (int m = 0; m < R; ++m) x[m] = ex[(size_t)p * N + lane + 64 * m];
  fft_any<R, -1>(x, buf[wave][0], buf[wave][1], tw, lane);
  float part = 0.f;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const float a = amp[(size_t)p * N + spec_index<R>(lane, m)];
    const float mag = sqrtf(fmaf(x[m].x, x[m].x, x[m].y * x[m].y));
    const float dlt = mag - a;
    part = fmaf(dlt, dlt, part);
    const float w = mag > 0.f ? scale * dlt / mag : 0.f;
    x[m] = cscale(x[m], w);
  }
  fft_any<R, +1>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
  for (int m = 0; m < R; ++m) seed[(size_t)p * N + lane + 64 * m] = x[m];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if (lane == 0) loss[p] = part;
}

__global__ void ms_transmission_kernel(const double2* __restrict__ nidx, int n, int nz, double k, double dz,
                                       double nmed, float2* __restrict__ T) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n * (nz - 1);
  if (id >= total) return;
  const int z = (int)(id / n);
  const int x = (int)(id % n);
  const double2 v = nidx[(size_t)x * nz + z];
  // exp(i k (n - nmed) dz) = exp(-k dz Im n) * (cos, sin)(k dz (Re n - nmed))
  const double amp = exp(-k * dz * v.y);
  double s, c;
  sincos(k * dz * (v.x - nmed), &s, &c);
  T[id] = make_float2((float)(amp * c), (float)(amp * s));
}

__global__ void ms_kernel_kernel(int n, double dx, double k, double dz, float2* __restrict__ H) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  // numpy.fft.fftfreq ordering
  const int f = (x < (n + 1) / 2) ? x : x - n;
  const double kx = 2.0 * M_PI * (double)f / ((double)n * dx);
  const double inside = k * k - kx * kx;
  const double kz = inside > 0.0 ? sqrt(inside) : 0.0;
  double s, c;
  sincos((kz - k) * dz, &s, &c);
  H[x] = make_float2((float)(c / n), (float)(s / n));
}

static int ms_radix(int n, int& r) {
  if (n >= 64 && n <= 2048 && n % 64 == 0) {
    r = n / 64;
    return TP_OK;
  }
  set_error("multislice/far-field length n = %d not supported (multiples of 64 up to 2048)", n);
  return TP_ERR_UNSUPPORTED;
}

#define TP_MS_R_CASES(X) \
  X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
  X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

static int ms_build(const tp_ms_scan* scans, int nscans, MsArgs& a, int& nblk, int& r, int ppb = MS_PPB) {
  if (nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  memset(&a, 0, sizeof(a));
  nblk = 0;
  r = 0;
  for (int i = 0; i < nscans; ++i) {
    const tp_ms_scan& d = scans[i];
    int ri = 0;
    int rc = ms_radix(d.n, ri);
    if (rc) return rc;
    if (i == 0) r = ri;
    if (ri != r) {
      set_error("all scans of one multislice launch must share n (scan %d has %d)", i, d.n);
      return TP_ERR_INVALID;
    }
    if (d.steps < 1 || d.probes < 0 || !d.T || !d.H || !d.u0) {
      set_error("scan %d: steps >= 1, T, H and u0 are required", i);
      return TP_ERR_INVALID;
    }
    a.scans[i] = d;
    a.blk_start[i] = nblk;
    nblk += (d.probes + ppb - 1) / ppb;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  return TP_OK;
}

}  // namespace tp

using namespace tp;

extern "C" int tp_ms_transmission(const void* d_n, int n, int nz, double k, double dz, double n_medium, void* d_T,
                                  void* stream) {
  if (!d_n || !d_T || n < 1 || nz < 2) {
    set_error("tp_ms_transmission: invalid arguments (n=%d, nz=%d)", n, nz);
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)n * (nz - 1);
  hipLaunchKernelGGL(ms_transmission_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)d_n, n, nz, k, dz, n_medium, (float2*)d_T);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_ms_kernel(int n, double dx, double k, double dz, void* d_H, void* stream) {
  if (!d_H || n < 1) {
    set_error("tp_ms_kernel: invalid arguments");
    return TP_ERR_INVALID;
  }
  hipLaunchKernelGGL(ms_kernel_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, dx, k, dz,
                     (float2*)d_H);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_ms_sweep(const tp_ms_scan* scans, int nscans, void* stream) {
  MsArgs a;
  int nblk = 0, r = 0;
  int rc = ms_build(scans, nscans, a, nblk, r);
  if (rc) return rc;
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  if (r == 8 || r == 16) {  // split form: one probe per workgroup of S wavefronts
    rc = ms_build(scans, nscans, a, nblk, r, 1);
    if (rc) return rc;
    if (r == 8)
      hipLaunchKernelGGL((ms_split_kernel<8, 2>), dim3(nblk), dim3(128), 0, st, a);
    else
      hipLaunchKernelGGL((ms_split_kernel<16, 4>), dim3(nblk), dim3(256), 0, st, a);
    TP_CHECK_HIP(hipGetLastError());
    return TP_OK;
  }
  const bool spread = nblk < 256;  // fewer 4-probe workgroups than CUs
  if (spread) {
    rc = ms_build(scans, nscans, a, nblk, r, 1);
    if (rc) return rc;
  }
  switch (r) {
#define TP_X(RV)                                                                                         \
  case RV:                                                                                              \
    if (spread)                                                                                         \
      hipLaunchKernelGGL((ms_sweep_kernel<RV, 1>), dim3(nblk), dim3(64), 0, st, a);                     \
    else                                                                                                \
      hipLaunchKernelGGL((ms_sweep_kernel<RV, MS_PPB>), dim3(nblk), dim3(64 * MS_PPB), 0, st, a);       \
    break;
    TP_MS_R_CASES(TP_X)
#undef TP_X
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_ms_step(const tp_ms_scan* scans, int nscans, int z, void* stream) {
  MsArgs a;
  int nblk = 0, r = 0;
  int rc = ms_build(scans, nscans, a, nblk, r);
  if (rc) return rc;
  for (int i = 0; i < nscans; ++i) {
    if (!scans[i].out || z < 0 || z >= scans[i].steps) {
      set_error("tp_ms_step: scan %d needs out != NULL and 0 <= z < steps (z=%d)", i, z);
      return TP_ERR_INVALID;
    }
  }
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (r) {
#define TP_X(RV) \
  case RV: hipLaunchKernelGGL((ms_step_kernel<RV>), dim3(nblk), dim3(64 * MS_PPB), 0, st, a, z); break;
    TP_MS_R_CASES(TP_X)
#undef TP_X
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_amplitude_seed(const void* d_exit, const void* d_sqrt_i, int n, int probes, double scale,
                                 void* d_seed, void* d_loss, void* stream) {
  int r = 0;
  int rc = ms_radix(n, r);
  if (rc) return rc;
  if (!d_exit || !d_sqrt_i || !d_seed || !d_loss || probes < 0) {
    set_error("tp_amplitude_seed: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  const dim3 grid((probes + MS_PPB - 1) / MS_PPB), block(64 * MS_PPB);
  hipStream_t st = (hipStream_t)stream;
  const float2* ex = (const float2*)d_exit;
  const float* a = (const float*)d_sqrt_i;
  float2* sd = (float2*)d_seed;
  float* ls = (float*)d_loss;
  const float sc = (float)scale;
  switch (r) {
#define TP_X(RV) \
  case RV: hipLaunchKernelGGL((amplitude_seed_kernel<RV>), grid, block, 0, st, ex, a, probes, sc, sd, ls); break;
    TP_MS_R_CASES(TP_X)
#undef TP_X
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_farfield(const void* d_exit, int64_t ld, int n, int probes, void* d_intensity, void* stream) {
  int r = 0;
  int rc = ms_radix(n, r);
  if (rc) return rc;
  if (!d_exit || !d_intensity || probes < 0 || ld < n) {
    set_error("tp_farfield: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  const dim3 grid((probes + MS_PPB - 1) / MS_PPB), block(64 * MS_PPB);
  hipStream_t st = (hipStream_t)stream;
  const float2* ex = (const float2*)d_exit;
  float* o = (float*)d_intensity;
  switch (r) {
#define TP_X(RV) \
  case RV: hipLaunchKernelGGL((farfield_kernel<RV>), grid, block, 0, st, ex, ld, probes, o); break;
    TP_MS_R_CASES(TP_X)
#undef TP_X
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
