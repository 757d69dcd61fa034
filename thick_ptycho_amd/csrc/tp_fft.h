// Wave-level complex FFT of n = 64 R points (R = 1 .. 32) held in registers:
// lane l keeps x[l + 64 m], m < R.  R in {4, 8, 16, 32}: Stockham radix stages
// through two per-wavefront LDS buffers, natural order in and out.  Other R: an
// R-point DFT in registers and a 64-point radix-2 transform across lanes with
// xor shuffles; the spectrum is left in the order spec_index<R>() describes.
// The twiddle table W_n^e = exp(-2 pi i e / n), e < n, is built in LDS by
// build_twiddles (FP64 sincospi, stored complex64).
#pragma once
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

// LDS position of element i of a Stockham buffer: one padding slot after every
// 16 complex values (128 bytes), so that the strided writes of the register
// stage (lane*R + r) and of the merge stages fall into distinct banks.
__device__ __forceinline__ constexpr int lp(int i) { return i + (i >> 4); }
// Wider padding (four slots per 16 values) for the buffer that receives the
// radix-4, NS = 4 stage of the 256-point transform: its writes
// (j/4)*16 + j%4 + 4r would meet in 4-way bank conflicts with lp.
__device__ __forceinline__ constexpr int lp4(int i) { return i + ((i >> 4) << 2); }
template <int PAD>
__device__ __forceinline__ constexpr int lpad(int i) { return PAD == 4 ? lp4(i) : lp(i); }

// Stockham stage (Govindaraju et al., SC'08 formulation): sub-transforms of
// size NS are merged by radix RS.  Butterfly j reads src[j + r*N/RS] and writes
// dst[(j/NS)*NS*RS + j%NS + r*NS].
template <int N, int RS, int NS, int SIGN, int PIN = 1, int POUT = 1>
__device__ __forceinline__ void stage_lds(const float2* src, float2* dst, const float2* tw, int lane) {
  constexpr int B = N / RS;
  constexpr int PER = B / 64;
  static_assert(PER >= 1, "stage needs at least 64 butterflies");
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int j = lane + 64 * t;
    float2 v[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) v[r] = src[lpad<PIN>(j + r * B)];
    const int jm = j & (NS - 1);
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < RS; ++r) v[r] = cmul(v[r], twid<SIGN>(tw, jm * r * (N / (NS * RS))));
    }
    dft_reg<RS, SIGN>(v);
    const int base = (j / NS) * NS * RS + jm;
#pragma unroll
    for (int r = 0; r < RS; ++r) dst[lpad<POUT>(base + r * NS)] = v[r];
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
    for (int r = 0; r < RS; ++r) v[r] = src[lp(j + r * B)];
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
  for (int r = 0; r < R; ++r) dst[lp(lane * R + r)] = x[r];
}

// Full FFT of n = 64 R points, register layout in and out; buffers a, b are
// per-wavefront LDS scratch of ms_bufn<R>() points each (n plus padding).
// W4: the wider padding of buffer b for R = 4 (conflict-free, 25 % more LDS);
// kernels whose occupancy is bound by LDS keep the narrow one.
template <int R, int SIGN, bool W4 = true>
__device__ __forceinline__ void fft_wave(float2* x, float2* a, float2* b, const float2* tw, int lane) {
  constexpr int N = 64 * R;
  constexpr int PB = W4 ? 4 : 1;
  if constexpr (R == 4) {  // 4 * 4 * 4 * 4
    stage_from_reg<4, SIGN>(x, a, lane);
    stage_lds<N, 4, 4, SIGN, 1, PB>(a, b, tw, lane);
    stage_lds<N, 4, 16, SIGN, PB, 1>(b, a, tw, lane);
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

// The 256-point wave transform (fft_wave<4>) with its twiddles in registers.
// Every call reads the same powers of W_256 (per lane: three for each of the
// two LDS stages and three for the last), so a sweep loads them once
// (tw256_load) instead of once per transform from the LDS table, whose lookups
// at lane strides of 2 and 3 bank-conflict.  Forward powers are kept; the
// inverse uses their conjugates.
struct Tw256 {
  float2 s1[3], s2[3], s3[3];
};

__device__ __forceinline__ void tw256_load(Tw256& t, const float2* tw, int lane) {
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    t.s1[r - 1] = tw[(lane & 3) * r * 16];  // stage NS = 4:  W_256^{(j mod 4) r 16}
    t.s2[r - 1] = tw[(lane & 15) * r * 4];  // stage NS = 16: W_256^{(j mod 16) r 4}
    t.s3[r - 1] = tw[lane * r];             // last stage:    W_256^{j r}
  }
}

template <int SIGN>
__device__ __forceinline__ float2 tw_sign(float2 w) {
  return SIGN > 0 ? cconj(w) : w;
}

// Paddings of the two buffers of fft_wave4_rt: one slot per 32 values in a, four
// in b (gfx950's LDS serves 64 banks of 4 bytes; these keep each half-wavefront's
// 8-byte accesses of every stage on distinct banks, where the 16-value paddings
// of fft_wave, sized for 32 banks, leave two-way conflicts).
__device__ __forceinline__ constexpr int lpa(int i) { return i + (i >> 5); }
__device__ __forceinline__ constexpr int lpb(int i) { return i + ((i >> 5) << 2); }

// fft_wave<4, SIGN, true> with the twiddles of `t`: the same stages and
// buffers (at least ms_bufn<4>() points each), natural order in and out.
template <int SIGN>
__device__ __forceinline__ void fft_wave4_rt(float2* x, float2* a, float2* b, const Tw256& t, int lane) {
  const int j = lane;
  dft_reg<4, SIGN>(x);
#pragma unroll
  for (int r = 0; r < 4; ++r) a[lpa(j * 4 + r)] = x[r];
  {  // radix 4, NS = 4: a -> b
    float2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = a[lpa(j + r * 64)];
#pragma unroll
    for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], tw_sign<SIGN>(t.s1[r - 1]));
    dft_reg<4, SIGN>(v);
    const int base = (j >> 2) * 16 + (j & 3);
#pragma unroll
    for (int r = 0; r < 4; ++r) b[lpb(base + r * 4)] = v[r];
  }
  {  // radix 4, NS = 16: b -> a
    float2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = b[lpb(j + r * 64)];
#pragma unroll
    for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], tw_sign<SIGN>(t.s2[r - 1]));
    dft_reg<4, SIGN>(v);
    const int base = (j >> 4) * 64 + (j & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) a[lpa(base + r * 16)] = v[r];
  }
  {  // last stage: a -> registers, natural order
    float2 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = a[lpa(j + r * 64)];
#pragma unroll
    for (int r = 1; r < 4; ++r) v[r] = cmul(v[r], tw_sign<SIGN>(t.s3[r - 1]));
    dft_reg<4, SIGN>(v);
#pragma unroll
    for (int r = 0; r < 4; ++r) x[r] = v[r];
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
template <int R, int SIGN, bool W4 = true>
__device__ __forceinline__ void fft_any(float2* x, float2* a, float2* b, const float2* tw, int lane) {
  if constexpr (ms_pow2<R>()) {
    fft_wave<R, SIGN, W4>(x, a, b, tw, lane);
  } else if constexpr (SIGN < 0) {
    fft_gen_fwd<R>(x, tw, lane);
  } else {
    fft_gen_inv<R>(x, tw, lane);
  }
}

// LDS ping-pong buffers are needed only by the power-of-two Stockham path
template <int R, bool W4 = true>
constexpr int ms_bufn() {  // lp4(64 R - 1) + 1 for R = 4 with W4, lp(64 R - 1) + 1 otherwise
  return ms_pow2<R>() ? (R == 4 && W4 ? 64 * R + 16 * R : 64 * R + 4 * R) : 1;
}


}  // namespace tp
