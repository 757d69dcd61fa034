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
#include "tp_common.h"

namespace tp {

// cos(2 pi j / 32), j = 0..31
__device__ constexpr float kCos32[32] = {
    1.0f,                0.98078528040323043f,  0.92387953251128674f,  0.83146961230254524f,
    0.70710678118654757f, 0.55557023301960229f, 0.38268343236508984f,  0.19509032201612833f,
    0.0f,                -0.19509032201612819f, -0.38268343236508973f, -0.55557023301960196f,
    -0.70710678118654746f, -0.83146961230254535f, -0.92387953251128674f, -0.98078528040323043f,
    -1.0f,               -0.98078528040323043f, -0.92387953251128685f, -0.83146961230254546f,
    -0.70710678118654768f, -0.55557023301960218f, -0.38268343236509034f, -0.19509032201612866f,
    0.0f,                0.1950903220161283f,   0.38268343236509f,     0.55557023301960184f,
    0.70710678118654735f, 0.83146961230254524f, 0.92387953251128652f,  0.98078528040323043f};

// exp(SIGN * 2 pi i j / 32)
template <int SIGN>
__device__ __forceinline__ float2 tw32(int j) {
  const float c = kCos32[j & 31];
  const float s = kCos32[(j + 24) & 31];  // sin(2 pi j/32) = cos(2 pi (j-8)/32)
  return make_float2(c, SIGN > 0 ? s : -s);
}

template <int R>
__host__ __device__ constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1, k = R >> 1; b < R; b <<= 1, k >>= 1)
    if (i & b) r |= k;
  return r;
}

// In-register DFT of size R (power of two <= 32), natural order in and out:
// X[k] = sum_j v[j] exp(SIGN 2 pi i j k / R).
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(float2* v) {
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bitrev<R>(i);
    if (j > i) {
      const float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        const float2 a = v[i + k];
        float2 b = v[i + k + len / 2];
        if (k == 0) {
        } else if (4 * k == len) {
          // multiply by SIGN * i
          b = SIGN > 0 ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
        } else {
          b = cmul(b, tw32<SIGN>(k * (32 / len)));
        }
        v[i + k] = cadd(a, b);
        v[i + k + len / 2] = csub(a, b);
      }
    }
  }
}

// Twiddle W_N^e = exp(-2 pi i e / N) from the LDS table, conjugated for SIGN > 0.
template <int SIGN>
__device__ __forceinline__ float2 twid(const float2* tw, int e) {
  const float2 w = tw[e];
  return SIGN > 0 ? cconj(w) : w;
}

// Stockham stage (Govindaraju et al., SC'08 formulation): sub-transforms of
// size NS are merged by radix RS.  Butterfly j reads src[j + r*N/RS] and writes
// dst[(j/NS)*NS*RS + j%NS + r*NS].
template <int N, int RS, int NS, int SIGN>
__device__ __forceinline__ void stage_lds(const float2* src, float2* dst, const float2* tw, int lane) {
  constexpr int B = N / RS;
  constexpr int PER = B / 64;
  static_assert(PER >= 1, "stage needs at least 64 butterflies");
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int j = lane + 64 * t;
    float2 v[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) v[r] = src[j + r * B];
    const int jm = j & (NS - 1);
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < RS; ++r) v[r] = cmul(v[r], twid<SIGN>(tw, jm * r * (N / (NS * RS))));
    }
    dft_reg<RS, SIGN>(v);
    const int base = (j / NS) * NS * RS + jm;
#pragma unroll
    for (int r = 0; r < RS; ++r) dst[base + r * NS] = v[r];
  }
}

// Last stage: NS * RS == N, results returned to registers in natural order
// (lane l, register m holds X[l + 64 m]).
template <int N, int RS, int SIGN>
__device__ __forceinline__ void stage_to_reg(const float2* src, float2* x, const float2* tw, int lane) {
  constexpr int B = N / RS;
  constexpr int NS = B;
  constexpr int PER = B / 64;
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int j = lane + 64 * t;
    float2 v[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) v[r] = src[j + r * B];
#pragma unroll
    for (int r = 1; r < RS; ++r) v[r] = cmul(v[r], twid<SIGN>(tw, j * r));
    dft_reg<RS, SIGN>(v);
#pragma unroll
    for (int r = 0; r < RS; ++r) x[t + r * PER] = v[r];
  }
}

// First stage from the register layout (NS = 1, no twiddles): lane l holds
// x[l + 64 m]; writes dst[l*R + r].
template <int R, int SIGN>
__device__ __forceinline__ void stage_from_reg(float2* x, float2* dst, int lane) {
  dft_reg<R, SIGN>(x);
#pragma unroll
  for (int r = 0; r < R; ++r) dst[lane * R + r] = x[r];
}

// Full FFT of n = 64 R points, register layout in and out; buffers a, b are
// per-wavefront LDS scratch of n points each.
template <int R, int SIGN>
__device__ __forceinline__ void fft_wave(float2* x, float2* a, float2* b, const float2* tw, int lane) {
  constexpr int N = 64 * R;
  if constexpr (R == 4) {  // 4 * 4 * 4 * 4
    stage_from_reg<4, SIGN>(x, a, lane);
    stage_lds<N, 4, 4, SIGN>(a, b, tw, lane);
    stage_lds<N, 4, 16, SIGN>(b, a, tw, lane);
    stage_to_reg<N, 4, SIGN>(a, x, tw, lane);
  } else if constexpr (R == 8) {  // 8 * 8 * 8
    stage_from_reg<8, SIGN>(x, a, lane);
    stage_lds<N, 8, 8, SIGN>(a, b, tw, lane);
    stage_to_reg<N, 8, SIGN>(b, x, tw, lane);
  } else if constexpr (R == 16) {  // 16 * 16 * 4
    stage_from_reg<16, SIGN>(x, a, lane);
    stage_lds<N, 16, 16, SIGN>(a, b, tw, lane);
    stage_to_reg<N, 4, SIGN>(b, x, tw, lane);
  } else {  // R == 32: 32 * 8 * 8
    stage_from_reg<32, SIGN>(x, a, lane);
    stage_lds<N, 8, 32, SIGN>(a, b, tw, lane);
    stage_to_reg<N, 8, SIGN>(b, x, tw, lane);
  }
}

// Twiddle table W_N^e, e < N, built per workgroup in FP64.
__device__ __forceinline__ void build_twiddles(float2* tw, int n) {
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    double s, c;
    sincospi(-2.0 * (double)e / (double)n, &s, &c);
    tw[e] = make_float2((float)c, (float)s);
  }
}

// ---------------------------------------------------------------- any n = 64 R
// For R not in {4, 8, 16, 32}: n = l + 64 m, k = k1 + R k2 gives
//   X[k1 + R k2] = sum_l w64^{l k2} [ wN^{l k1} sum_m x[l + 64 m] wR^{m k1} ]
// an R-point DFT in registers (lane l, register m), the twiddle wN^{l k1}, then a
// 64-point radix-2 transform across the lanes with xor shuffles (decimation in
// frequency: natural in, bit-reversed out).  The spectrum is left with lane L,
// register k1 holding X[k1 + R bitrev6(L)]; the inverse runs the steps backwards
// (decimation in time: bit-reversed in, natural out) and returns natural order.
template <int R>
constexpr bool ms_pow2() {
  return R == 4 || R == 8 || R == 16 || R == 32;
}

__device__ __forceinline__ int bitrev6(int x) { return (int)(__builtin_bitreverse32((unsigned)x) >> 26); }

// frequency held by lane `lane`, register k1 after the forward transform
template <int R>
__device__ __forceinline__ int spec_index(int lane, int k1) {
  if constexpr (ms_pow2<R>())
    return lane + 64 * k1;
  else
    return k1 + R * bitrev6(lane);
}

__device__ __forceinline__ float2 shfl_xor2(float2 v, int m) {
  return make_float2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

// y[k] = sum_m x[m] exp(SIGN 2 pi i m k / R), wR^j = W_N^{64 j} from the table
template <int R, int SIGN>
__device__ __forceinline__ void dft_small(float2* x, const float2* tw) {
  float2 wr[R];
#pragma unroll
  for (int j = 0; j < R; ++j) wr[j] = twid<SIGN>(tw, 64 * j);
  float2 y[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    float2 acc = x[0];
#pragma unroll
    for (int m = 1; m < R; ++m) acc = cfma(x[m], wr[(m * k) % R], acc);
    y[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) x[k] = y[k];
}

template <int R>
__device__ __forceinline__ void fft_gen_fwd(float2* x, const float2* tw, int lane) {
  constexpr int N = 64 * R;
  dft_small<R, -1>(x, tw);
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) x[k1] = cmul(x[k1], twid<-1>(tw, (lane * k1) % N));
#pragma unroll
  for (int h = 32; h >= 1; h >>= 1) {
    const float2 w = twid<-1>(tw, (lane & (h - 1)) * (32 / h) * R);
    const bool bot = (lane & h) != 0;
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) {
      const float2 p = shfl_xor2(x[k1], h);
      x[k1] = bot ? cmul(csub(p, x[k1]), w) : cadd(x[k1], p);
    }
  }
}

template <int R>
__device__ __forceinline__ void fft_gen_inv(float2* x, const float2* tw, int lane) {
  constexpr int N = 64 * R;
#pragma unroll
  for (int h = 1; h <= 32; h <<= 1) {
    const float2 w = twid<+1>(tw, (lane & (h - 1)) * (32 / h) * R);
    const bool bot = (lane & h) != 0;
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) {
      const float2 p = shfl_xor2(x[k1], h);
      x[k1] = bot ? csub(p, cmul(w, x[k1])) : cfma(w, p, x[k1]);
    }
  }
#pragma unroll
  for (int k1 = 1; k1 < R; ++k1) x[k1] = cmul(x[k1], twid<+1>(tw, (lane * k1) % N));
  dft_small<R, +1>(x, tw);
}

// forward (SIGN -1) / inverse (SIGN +1, unnormalised) transform of one wavefront's field
template <int R, int SIGN>
__device__ __forceinline__ void fft_any(float2* x, float2* a, float2* b, const float2* tw, int lane) {
  if constexpr (ms_pow2<R>()) {
    fft_wave<R, SIGN>(x, a, b, tw, lane);
  } else if constexpr (SIGN < 0) {
    fft_gen_fwd<R>(x, tw, lane);
  } else {
    fft_gen_inv<R>(x, tw, lane);
  }
}

// LDS ping-pong buffers are needed only by the power-of-two Stockham path
template <int R>
constexpr int ms_bufn() {
  return ms_pow2<R>() ? 64 * R : 1;
}

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

template <int R>
__device__ __forceinline__ void ms_store(const tp_ms_scan& d, int p, int z, const float2* x, int lane) {
  if (d.out) {
    float2* o = (float2*)d.out + (size_t)p * d.out_pstride + (int64_t)z * d.out_zstride;
#pragma unroll
    for (int m = 0; m < R; ++m) o[lane + 64 * m] = x[m];
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

template <int R>
__global__ __launch_bounds__(64 * MS_PPB) void ms_sweep_kernel(const MsArgs a) {
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

static int ms_build(const tp_ms_scan* scans, int nscans, MsArgs& a, int& nblk, int& r) {
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
    nblk += (d.probes + MS_PPB - 1) / MS_PPB;
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
  switch (r) {
#define TP_X(RV) \
  case RV: hipLaunchKernelGGL((ms_sweep_kernel<RV>), dim3(nblk), dim3(64 * MS_PPB), 0, st, a); break;
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
