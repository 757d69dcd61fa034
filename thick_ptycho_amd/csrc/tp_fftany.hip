// Spectral operations at any length n: multislice propagation (a circular
// convolution of length n), far-field intensities and the amplitude-misfit
// seed (DFTs of length n), for the widths the one-wavefront kernels
// (n = 64 r, r <= 32, tp_ms.hip) do not hold -- the reference's 2D
// reconstruction notebook runs n = 751, a prime.
//
// Every operation is a circular convolution on a grid of F points (F a power
// of two, F >= 2n - 1), one workgroup per probe:
//     c = IFFT_F( FFT_F(pad(a)) * S ),
// FFT_F a radix-2 decimation-in-time transform in LDS (bit-reversed in, natural
// out), IFFT_F the decimation-in-frequency inverse (natural in, bit-reversed
// out, unnormalised), S a spectrum with 1/F folded in.  The inverse leaves
// y_i at the slot the forward transform reads x_i from, so a sweep keeps its
// state in place from slice to slice.
//  * multislice: ifft_n(H fft_n(x)) is the circular convolution of x with
//    h = ifft_n(H); with x zero-padded to F points it is the linear convolution
//    with h extended periodically to (-n, n), exact for F >= 2n - 1.
//    S = FFT_F of that extension / F; back-propagation with conj(H) uses conj(S).
//  * DFT of length n (Bluestein 1970): X_k = conj(w_k) sum_j (x_j conj(w_j)) w_{k-j},
//    chirp w_m = exp(i pi m^2 / n); S = FFT_F of w on (-n, n) / F.  The
//    unnormalised inverse DFT uses the conjugate chirps, whose spectrum is conj(S).
// At most 16384 points (128 KiB) fit in LDS.  Above that (F = 32768, n up to
// 16384) the transform takes one radix-2 frequency split F = 2M: the padded
// input's upper half is zero and only the first n <= M outputs are kept, so
//     c_i = Y0_i + conj(W_F^i) Y1_i,  Y0 = IFFT_M(FFT_M(a) S_even),
//     Y1 = IFFT_M(FFT_M(W_F^. a) S_odd),
// two LDS transforms per convolution with Y0 and the input staged in a
// per-probe workspace of F points.
//
// Set-up (FP64 direct sums on the device, cached by the caller): the chirp,
// h = ifft_n(H) and the spectra.
#include "tp_common.h"

namespace tp {

constexpr int FA_LDS_MAX = 16384;  // points of the LDS transform
constexpr int FA_F_MAX = 32768;    // largest convolution grid (n <= 16384)

__device__ __forceinline__ int fa_rev(int i, int lg) { return (int)(__brev((unsigned)i) >> (32 - lg)); }

// tab[i] = (cos, sin)(2 pi i / M) for i <= M/8 (FP64 sincospi, stored complex64).
__device__ __forceinline__ void fa_build_table(float2* tab, int M, int tid, int T) {
  for (int i = tid; i <= (M >> 3); i += T) {
    double s, c;
    sincospi(2.0 * (double)i / (double)M, &s, &c);
    tab[i] = make_float2((float)c, (float)s);
  }
}

// exp(sign * 2 pi i * e / M) for 0 <= e < M/2, from the octant table.
__device__ __forceinline__ float2 fa_twiddle(const float2* tab, int e, int M, float sign) {
  const int q = M >> 2, o = M >> 3;
  const bool hi = e >= q;
  const int r = hi ? e - q : e;
  const float2 t = (r <= o) ? tab[r] : make_float2(tab[q - r].y, tab[q - r].x);
  const float c = hi ? -t.y : t.x;
  const float s = hi ? t.x : t.y;
  return make_float2(c, sign * s);
}

// In-place decimation-in-time FFT (bit-reversed input, natural output), sign -1.
__device__ __forceinline__ void fa_dit(float2* buf, const float2* tab, int M, int lg, int tid, int T) {
  for (int s = 0; s < lg; ++s) {
    const int h = 1 << s;
    const int shift = lg - 1 - s;
    for (int b = tid; b < (M >> 1); b += T) {
      const int r = b & (h - 1);
      const int j = ((b >> s) << (s + 1)) | r;
      const float2 w = fa_twiddle(tab, r << shift, M, -1.f);
      const float2 a0 = buf[j];
      const float2 a1 = cmul(buf[j + h], w);
      buf[j] = cadd(a0, a1);
      buf[j + h] = csub(a0, a1);
    }
    __syncthreads();
  }
}

// In-place decimation-in-frequency inverse FFT (natural input, bit-reversed
// output, unnormalised), sign +1.
__device__ __forceinline__ void fa_dif(float2* buf, const float2* tab, int M, int lg, int tid, int T) {
  for (int s = lg - 1; s >= 0; --s) {
    const int h = 1 << s;
    const int shift = lg - 1 - s;
    for (int b = tid; b < (M >> 1); b += T) {
      const int r = b & (h - 1);
      const int j = ((b >> s) << (s + 1)) | r;
      const float2 w = fa_twiddle(tab, r << shift, M, +1.f);
      const float2 a0 = buf[j];
      const float2 a1 = buf[j + h];
      buf[j] = cadd(a0, a1);
      buf[j + h] = cmul(csub(a0, a1), w);
    }
    __syncthreads();
  }
}

// c = conv(a, S) on F points, a_i = pre(i) for i < n (0 beyond), and post(i, c_i)
// for i < n.  Non-split: F = M, pre may read the LDS slot of its own i (the
// state of a sweep).  Split: F = 2M, `work` holds F points for this probe.
template <bool SPLIT, typename Pre, typename Post>
__device__ __forceinline__ void fa_convolve(float2* buf, const float2* tab, int M, int lg, int n, const float2* S,
                                            bool conj_s, float2* work, const Pre& pre, const Post& post, int tid,
                                            int T) {
  if constexpr (!SPLIT) {
    for (int i = tid; i < M; i += T) {
      const int r = fa_rev(i, lg);
      buf[r] = (i < n) ? pre(i) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fa_dit(buf, tab, M, lg, tid, T);
    for (int k = tid; k < M; k += T) {
      float2 g = S[k];
      if (conj_s) g = cconj(g);
      buf[k] = cmul(buf[k], g);
    }
    __syncthreads();
    fa_dif(buf, tab, M, lg, tid, T);
    for (int i = tid; i < n; i += T) post(i, buf[fa_rev(i, lg)]);
    __syncthreads();
  } else {
    // W_F^i = W_M^{i/2} (i even) or W_M^{(i-1)/2} W_F^1 (i odd), 2i < F
    double s1, c1;
    sincospi(-1.0 / (double)M, &s1, &c1);
    const float2 wf1 = make_float2((float)c1, (float)s1);
    float2* y0 = work;      // IFFT_M of the even half, bit-reversed slots
    float2* stage = work + M;  // a_i, i < n
    for (int i = tid; i < M; i += T) {
      const float2 a = (i < n) ? pre(i) : make_float2(0.f, 0.f);
      stage[i] = a;
      buf[fa_rev(i, lg)] = a;
    }
    __syncthreads();
    fa_dit(buf, tab, M, lg, tid, T);
    for (int k = tid; k < M; k += T) {
      float2 g = S[2 * k];
      if (conj_s) g = cconj(g);
      buf[k] = cmul(buf[k], g);
    }
    __syncthreads();
    fa_dif(buf, tab, M, lg, tid, T);
    for (int i = tid; i < M; i += T) y0[i] = buf[i];
    __syncthreads();
    for (int i = tid; i < M; i += T) {
      float2 v = make_float2(0.f, 0.f);
      if (i < n) {
        float2 w = fa_twiddle(tab, i >> 1, M, -1.f);
        if (i & 1) w = cmul(w, wf1);
        v = cmul(stage[i], w);
      }
      buf[fa_rev(i, lg)] = v;
    }
    __syncthreads();
    fa_dit(buf, tab, M, lg, tid, T);
    for (int k = tid; k < M; k += T) {
      float2 g = S[2 * k + 1];
      if (conj_s) g = cconj(g);
      buf[k] = cmul(buf[k], g);
    }
    __syncthreads();
    fa_dif(buf, tab, M, lg, tid, T);
    for (int i = tid; i < n; i += T) {
      float2 w = fa_twiddle(tab, i >> 1, M, -1.f);
      if (i & 1) w = cmul(w, wf1);
      const int r = fa_rev(i, lg);
      post(i, cfma(cconj(w), buf[r], y0[r]));
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ multislice
struct FaMsArgs {
  tp_ms_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans;
  int M, lg;  // LDS transform
  float2* work;
  int64_t work_pstride;
};

template <bool SPLIT>
__global__ __launch_bounds__(1024) void fa_ms_sweep_kernel(const FaMsArgs a) {
  extern __shared__ float2 fsm[];
  const int M = a.M, lg = a.lg, T = blockDim.x, tid = threadIdx.x;
  float2* buf = fsm;
  float2* tab = fsm + M;
  const int blk = blockIdx.x;
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  const tp_ms_scan& d = a.scans[sc];
  const int p = blk - a.blk_start[sc];  // one probe per workgroup, every probe exists
  const int n = d.n;
  fa_build_table(tab, M, tid, T);
  float2* work = SPLIT ? a.work + (size_t)blk * a.work_pstride : nullptr;
  float2* state = SPLIT ? work + M : nullptr;  // the staging half holds the state between slices
  const float2* u0 = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
  float2* outp = d.out ? (float2*)d.out + (size_t)p * d.out_pstride : nullptr;
  for (int i = tid; i < M; i += T) {
    const float2 v = (i < n) ? u0[i] : make_float2(0.f, 0.f);
    if (SPLIT) {
      state[i] = v;
    } else {
      buf[fa_rev(i, lg)] = v;
    }
    if (outp && i < n) outp[i] = v;
  }
  __syncthreads();
  const float2* S = (const float2*)d.H;
  for (int z = 0; z < d.steps; ++z) {
    const float2* Tz = (const float2*)d.T + (size_t)z * n;
    float2* o = outp ? outp + (int64_t)(z + 1) * d.out_zstride : nullptr;
    auto pre = [&](int i) { return cmul(SPLIT ? state[i] : buf[fa_rev(i, lg)], Tz[i]); };
    auto post = [&](int i, float2 c) {
      if (SPLIT) state[i] = c;
      if (o) o[i] = c;
    };
    fa_convolve<SPLIT>(buf, tab, M, lg, n, S, d.conj_h != 0, work, pre, post, tid, T);
  }
  if (d.exitp) {
    float2* e = (float2*)d.exitp + (size_t)p * d.exit_pstride;
    for (int i = tid; i < n; i += T) e[i] = SPLIT ? state[i] : buf[fa_rev(i, lg)];
  }
}

// ------------------------------------------------------------------ DFT of length n
struct FaDftArgs {
  const float2* x;
  int64_t ld;
  const float* amp;  // sqrt of the measured intensities (seed only)
  const float2* chirp;
  const float2* S;
  float* inten;  // far field
  float2* seed;  // seed
  float* loss;
  float scale;
  int n, probes, M, lg;
  float2* work;
};

// MODE 0: far field |X_k|^2.  MODE 1: amplitude seed (two Bluestein DFTs).
template <int MODE, bool SPLIT>
__global__ __launch_bounds__(1024) void fa_dft_kernel(const FaDftArgs a) {
  extern __shared__ float2 fsm[];
  const int M = a.M, lg = a.lg, T = blockDim.x, tid = threadIdx.x;
  float2* buf = fsm;
  float2* tab = fsm + M;
  float* red = reinterpret_cast<float*>(tab + (M >> 3) + 1);
  const int p = blockIdx.x;
  const int n = a.n;
  fa_build_table(tab, M, tid, T);
  __syncthreads();
  float2* work = SPLIT ? a.work + (size_t)p * (2 * M) : nullptr;
  const float2* x = a.x + (size_t)p * a.ld;
  const float2* w = a.chirp;
  auto pre = [&](int i) { return cmul(x[i], cconj(w[i])); };
  if (MODE == 0) {
    float* o = a.inten + (size_t)p * n;
    auto post = [&](int i, float2 c) { o[i] = fmaf(c.x, c.x, c.y * c.y); };
    fa_convolve<SPLIT>(buf, tab, M, lg, n, a.S, false, work, pre, post, tid, T);
  } else {
    float2* sd = a.seed + (size_t)p * n;
    const float* am = a.amp + (size_t)p * n;
    float part = 0.f;
    // U_k = conj(w_k) c_k; V_k = scale (|U_k| - a_k) U_k / |U_k|; keep V_k w_k in the output row
    auto post1 = [&](int k, float2 c) {
      const float2 U = cmul(cconj(w[k]), c);
      const float mag = sqrtf(fmaf(U.x, U.x, U.y * U.y));
      const float dlt = mag - am[k];
      part = fmaf(dlt, dlt, part);
      const float f = mag > 0.f ? a.scale * dlt / mag : 0.f;
      sd[k] = cmul(cscale(U, f), w[k]);
    };
    fa_convolve<SPLIT>(buf, tab, M, lg, n, a.S, false, work, pre, post1, tid, T);
    // seed_j = sum_k V_k exp(+2 pi i j k / n) = w_j sum_k (V_k w_k) conj(w_{j-k})
    auto pre2 = [&](int k) { return sd[k]; };
    auto post2 = [&](int j, float2 c) { sd[j] = cmul(w[j], c); };
    fa_convolve<SPLIT>(buf, tab, M, lg, n, a.S, true, work, pre2, post2, tid, T);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
      float t = 0.f;
      for (int i = 0; i < (T >> 6); ++i) t += red[i];
      a.loss[p] = t;
    }
  }
}

// ------------------------------------------------------------------ strided DFT
// X_k = sum_j x_j exp(-+2 pi i j k / n) of `count` sequences x_j = in[t*in_dist +
// j*in_stride] into out[t*out_dist + k*out_stride] (complex64, or |X_k|^2 in
// float32 with `power`); the inverse (unnormalised) as conj(DFT(conj(x))).
struct FaStridedArgs {
  const float2* in;
  float2* out;
  float* pout;
  int64_t in_stride, in_dist, out_stride, out_dist;
  const float2* chirp;
  const float2* S;
  int n, count, inverse, M, lg;
  float2* work;
};

template <bool SPLIT>
__global__ __launch_bounds__(1024) void fa_dft_strided_kernel(const FaStridedArgs a) {
  extern __shared__ float2 fsm[];
  const int M = a.M, lg = a.lg, T = blockDim.x, tid = threadIdx.x;
  float2* buf = fsm;
  float2* tab = fsm + M;
  const int t = blockIdx.x;
  fa_build_table(tab, M, tid, T);
  __syncthreads();
  float2* work = SPLIT ? a.work + (size_t)t * (2 * M) : nullptr;
  const float2* x = a.in + t * a.in_dist;
  const float2* w = a.chirp;
  const bool inv = a.inverse != 0;
  auto pre = [&](int j) {
    const float2 v = x[j * a.in_stride];
    return cmul(inv ? cconj(v) : v, cconj(w[j]));
  };
  auto post = [&](int k, float2 c) {
    if (a.pout) {
      a.pout[t * a.out_dist + k * a.out_stride] = fmaf(c.x, c.x, c.y * c.y);
    } else {
      const float2 X = cmul(cconj(w[k]), c);
      a.out[t * a.out_dist + k * a.out_stride] = inv ? cconj(X) : X;
    }
  };
  fa_convolve<SPLIT>(buf, tab, M, lg, a.n, a.S, false, work, pre, post, tid, T);
}

// x[t*dist + k1*B + j2] *= exp(sign 2 pi i j2 k1 / (A B)), the twiddles between
// the two passes of a four-step DFT of length A B.
__global__ void fa_twiddle_kernel(float2* __restrict__ x, int A, int B, int count, int64_t dist, int sign) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)A * B;
  if (id >= per * count) return;
  const int64_t t = id / per;
  const int64_t r = id % per;
  const int64_t k1 = r / B, j2 = r % B;
  const int64_t m = (j2 * k1) % per;
  double s, c;
  sincospi(2.0 * (double)m / (double)per, &s, &c);
  float2* p = x + t * dist + r;
  *p = cmul(*p, make_float2((float)c, (float)(sign * s)));
}

// ------------------------------------------------------------------ set-up (FP64)
// H_q = exp(i (kz_q - k) dz), kz = sqrt(max(k^2 - kx^2, 0)), kx = 2 pi fftfreq(n, dx)
__global__ void fa_prop_kernel(int n, double dx, double k, double dz, double2* __restrict__ H) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int f = (q < (n + 1) / 2) ? q : q - n;
  const double kx = 2.0 * M_PI * (double)f / ((double)n * dx);
  const double inside = k * k - kx * kx;
  const double kz = inside > 0.0 ? sqrt(inside) : 0.0;
  double s, c;
  sincos((kz - k) * dz, &s, &c);
  H[q] = make_double2(c, s);
}

// h_j = (1/n) sum_q H_q exp(+2 pi i j q / n): the phase advances by a fixed
// rotation per term and is recomputed exactly every 32 terms.
__global__ void fa_idft_kernel(int n, const double2* __restrict__ H, double2* __restrict__ h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double rs, rc;
  sincospi(2.0 * (double)j / (double)n, &rs, &rc);
  const double2 rot = make_double2(rc, rs);
  double2 acc = make_double2(0.0, 0.0), e = make_double2(1.0, 0.0);
  for (int q = 0; q < n; ++q) {
    if ((q & 31) == 0) {
      const long long m = ((long long)j * q) % n;
      double s, c;
      sincospi(2.0 * (double)m / (double)n, &s, &c);
      e = make_double2(c, s);
    }
    acc = zadd(acc, zmul(H[q], e));
    e = zmul(e, rot);
  }
  h[j] = zscale(acc, 1.0 / n);
}

// chirp w_m = exp(i pi (m^2 mod 2n) / n), m < n: complex128 and complex64
__global__ void fa_chirp_kernel(int n, double2* __restrict__ w64, float2* __restrict__ w32) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const long long r = ((long long)m * m) % (2LL * n);
  double s, c;
  sincospi((double)r / (double)n, &s, &c);
  w64[m] = make_double2(c, s);
  w32[m] = make_float2((float)c, (float)s);
}

// S_k = (1/F) sum_{d=-(n-1)}^{n-1} f(d) exp(-2 pi i d k / F), f(d) = src[d mod n]
// (periodic, MODE 0) or src[|d|] (even, MODE 1).
template <int MODE>
__global__ void fa_spectrum_kernel(int n, int F, const double2* __restrict__ src, float2* __restrict__ S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F) return;
  double rs, rc;
  sincospi(-2.0 * (double)k / (double)F, &rs, &rc);
  const double2 rot = make_double2(rc, rs);
  double2 acc = make_double2(0.0, 0.0), e = make_double2(1.0, 0.0);
  int t = 0;
  for (int d = -(n - 1); d <= n - 1; ++d, ++t) {
    if ((t & 31) == 0) {
      long long m = ((long long)d * k) % F;
      if (m < 0) m += F;
      double s, c;
      sincospi(-2.0 * (double)m / (double)F, &s, &c);
      e = make_double2(c, s);
    }
    const int idx = MODE == 0 ? (d < 0 ? d + n : d) : (d < 0 ? -d : d);
    acc = zadd(acc, zmul(src[idx], e));
    e = zmul(e, rot);
  }
  S[k] = to_f2(zscale(acc, 1.0 / F));
}

// ------------------------------------------------------------------ host side
static int fa_grid(int n, int& F, int& M, int& lg) {
  if (n < 1 || n > FA_F_MAX / 2) {
    set_error("spectral length n = %d outside [1, %d]", n, FA_F_MAX / 2);
    return TP_ERR_UNSUPPORTED;
  }
  F = 64;
  while (F < 2 * n - 1) F <<= 1;
  M = F > FA_LDS_MAX ? F / 2 : F;
  lg = 0;
  while ((1 << lg) < M) ++lg;
  return TP_OK;
}

static int fa_threads(int M) { return M / 8 < 64 ? 64 : (M / 8 > 1024 ? 1024 : M / 8); }

static size_t fa_lds(int M) { return ((size_t)M + (M >> 3) + 1) * sizeof(float2) + 16 * sizeof(float); }

template <typename K>
static int fa_set_lds(K kernel, size_t lds) {
  if (lds > 65536)
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return TP_OK;
}

}  // namespace tp

using namespace tp;

extern "C" int tp_spectral_len(int n) {
  int F = 0, M = 0, lg = 0;
  if (fa_grid(n, F, M, lg)) return TP_ERR_UNSUPPORTED;
  return F;
}

extern "C" size_t tp_spectral_setup_bytes(int n) { return (size_t)2 * (n > 0 ? n : 0) * sizeof(double2); }

extern "C" size_t tp_spectral_work_bytes(int n, int probes) {
  int F = 0, M = 0, lg = 0;
  if (fa_grid(n, F, M, lg) || F <= FA_LDS_MAX || probes <= 0) return 0;
  return (size_t)probes * F * sizeof(float2);
}

extern "C" int tp_ms_kernel_any(int n, double dx, double k, double dz, void* d_S, void* d_setup, void* stream) {
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(n, F, M, lg);
  if (rc) return rc;
  if (!d_S || !d_setup) {
    set_error("tp_ms_kernel_any: d_S and d_setup are required");
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  double2* H = (double2*)d_setup;
  double2* h = H + n;
  hipLaunchKernelGGL(fa_prop_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, dx, k, dz, H);
  hipLaunchKernelGGL(fa_idft_kernel, dim3((n + 127) / 128), dim3(128), 0, st, n, (const double2*)H, h);
  hipLaunchKernelGGL((fa_spectrum_kernel<0>), dim3((F + 127) / 128), dim3(128), 0, st, n, F, (const double2*)h,
                     (float2*)d_S);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_chirp_tables(int n, void* d_chirp, void* d_S, void* d_setup, void* stream) {
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(n, F, M, lg);
  if (rc) return rc;
  if (!d_chirp || !d_S || !d_setup) {
    set_error("tp_chirp_tables: d_chirp, d_S and d_setup are required");
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  double2* w64 = (double2*)d_setup;
  hipLaunchKernelGGL(fa_chirp_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, w64, (float2*)d_chirp);
  hipLaunchKernelGGL((fa_spectrum_kernel<1>), dim3((F + 127) / 128), dim3(128), 0, st, n, F, (const double2*)w64,
                     (float2*)d_S);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_ms_sweep_any(const tp_ms_scan* scans, int nscans, void* d_work, void* stream) {
  if (!scans || nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(scans[0].n, F, M, lg);
  if (rc) return rc;
  FaMsArgs a;
  memset(&a, 0, sizeof(a));
  int nblk = 0;
  for (int i = 0; i < nscans; ++i) {
    const tp_ms_scan& d = scans[i];
    if (d.n != scans[0].n) {
      set_error("all scans of one multislice launch must share n (scan %d has %d)", i, d.n);
      return TP_ERR_INVALID;
    }
    if (d.steps < 1 || d.probes < 0 || !d.T || !d.H || !d.u0) {
      set_error("scan %d: steps >= 1, T, H and u0 are required", i);
      return TP_ERR_INVALID;
    }
    a.scans[i] = d;
    a.blk_start[i] = nblk;
    nblk += d.probes;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  a.M = M;
  a.lg = lg;
  const bool split = F > M;
  if (split && !d_work) {
    set_error("tp_ms_sweep_any: n = %d needs a workspace of tp_spectral_work_bytes(n, probes) bytes", scans[0].n);
    return TP_ERR_INVALID;
  }
  a.work = (float2*)d_work;
  a.work_pstride = F;
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = fa_lds(M);
  const int T = fa_threads(M);
  if (split) {
    rc = fa_set_lds(fa_ms_sweep_kernel<true>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_ms_sweep_kernel<true>), dim3(nblk), dim3(T), lds, st, a);
  } else {
    rc = fa_set_lds(fa_ms_sweep_kernel<false>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_ms_sweep_kernel<false>), dim3(nblk), dim3(T), lds, st, a);
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

template <int MODE>
static int fa_dft_launch(FaDftArgs& a, int F, int M, hipStream_t st) {
  const size_t lds = fa_lds(M);
  const int T = fa_threads(M);
  int rc;
  if (F > M) {
    rc = fa_set_lds(fa_dft_kernel<MODE, true>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_dft_kernel<MODE, true>), dim3(a.probes), dim3(T), lds, st, a);
  } else {
    rc = fa_set_lds(fa_dft_kernel<MODE, false>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_dft_kernel<MODE, false>), dim3(a.probes), dim3(T), lds, st, a);
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_farfield_any(const void* d_exit, int64_t ld, int n, int probes, const void* d_chirp,
                               const void* d_S, void* d_intensity, void* d_work, void* stream) {
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(n, F, M, lg);
  if (rc) return rc;
  if (!d_exit || !d_chirp || !d_S || !d_intensity || probes < 0 || ld < n || (F > M && !d_work)) {
    set_error("tp_farfield_any: invalid arguments (n=%d, probes=%d; a workspace is needed above n = %d)", n, probes,
              FA_LDS_MAX / 2);
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  FaDftArgs a;
  memset(&a, 0, sizeof(a));
  a.x = (const float2*)d_exit;
  a.ld = ld;
  a.chirp = (const float2*)d_chirp;
  a.S = (const float2*)d_S;
  a.inten = (float*)d_intensity;
  a.n = n;
  a.probes = probes;
  a.M = M;
  a.lg = lg;
  a.work = (float2*)d_work;
  return fa_dft_launch<0>(a, F, M, (hipStream_t)stream);
}

extern "C" int tp_amplitude_seed_any(const void* d_exit, const void* d_sqrt_i, int n, int probes, double scale,
                                     const void* d_chirp, const void* d_S, void* d_seed, void* d_loss, void* d_work,
                                     void* stream) {
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(n, F, M, lg);
  if (rc) return rc;
  if (!d_exit || !d_sqrt_i || !d_chirp || !d_S || !d_seed || !d_loss || probes < 0 || (F > M && !d_work)) {
    set_error("tp_amplitude_seed_any: invalid arguments (n=%d, probes=%d; a workspace is needed above n = %d)", n,
              probes, FA_LDS_MAX / 2);
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  FaDftArgs a;
  memset(&a, 0, sizeof(a));
  a.x = (const float2*)d_exit;
  a.ld = n;
  a.amp = (const float*)d_sqrt_i;
  a.chirp = (const float2*)d_chirp;
  a.S = (const float2*)d_S;
  a.seed = (float2*)d_seed;
  a.loss = (float*)d_loss;
  a.scale = (float)scale;
  a.n = n;
  a.probes = probes;
  a.M = M;
  a.lg = lg;
  a.work = (float2*)d_work;
  return fa_dft_launch<1>(a, F, M, (hipStream_t)stream);
}

extern "C" int tp_dft_any(const void* d_in, int64_t in_stride, int64_t in_dist, void* d_out, int64_t out_stride,
                          int64_t out_dist, int n, int count, int inverse, int power, const void* d_chirp,
                          const void* d_S, void* d_work, void* stream) {
  int F = 0, M = 0, lg = 0;
  int rc = fa_grid(n, F, M, lg);
  if (rc) return rc;
  if (!d_in || !d_out || !d_chirp || !d_S || count < 0 || (F > M && !d_work)) {
    set_error("tp_dft_any: invalid arguments (n=%d, count=%d; a workspace is needed above n = %d)", n, count,
              FA_LDS_MAX / 2);
    return TP_ERR_INVALID;
  }
  if (count == 0) return TP_OK;
  FaStridedArgs a;
  a.in = (const float2*)d_in;
  a.out = power ? nullptr : (float2*)d_out;
  a.pout = power ? (float*)d_out : nullptr;
  a.in_stride = in_stride;
  a.in_dist = in_dist;
  a.out_stride = out_stride;
  a.out_dist = out_dist;
  a.chirp = (const float2*)d_chirp;
  a.S = (const float2*)d_S;
  a.n = n;
  a.count = count;
  a.inverse = inverse;
  a.M = M;
  a.lg = lg;
  a.work = (float2*)d_work;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = fa_lds(M);
  const int T = fa_threads(M);
  if (F > M) {
    rc = fa_set_lds(fa_dft_strided_kernel<true>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_dft_strided_kernel<true>), dim3(count), dim3(T), lds, st, a);
  } else {
    rc = fa_set_lds(fa_dft_strided_kernel<false>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((fa_dft_strided_kernel<false>), dim3(count), dim3(T), lds, st, a);
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_twiddle_mul(void* d_x, int A, int B, int count, int64_t dist, int sign, void* stream) {
  if (!d_x || A < 1 || B < 1 || count < 0 || (sign != 1 && sign != -1) || dist < (int64_t)A * B) {
    set_error("tp_twiddle_mul: invalid arguments (A=%d, B=%d, count=%d, sign=%d)", A, B, count, sign);
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)A * B * count;
  if (total == 0) return TP_OK;
  hipLaunchKernelGGL(fa_twiddle_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (float2*)d_x, A, B, count, dist, sign);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
