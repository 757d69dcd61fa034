// Learned preconditioner for the two-transverse-axis CN step: a complex-linear
// 4-layer 3x3 CNN on the (ny, nx) plane, 1 -> 16 -> 16 -> 16 -> 1 complex
// channels (2 -> 32 -> 32 -> 32 -> 2 real channels), no biases and no
// activations, so that M is a fixed linear operator as BiCGSTAB requires.
// Each complex weight w = a + ib acts on (re, im) channel pairs as the real
// block [[a, -b], [b, a]].
//
// Convolutions run as implicit GEMMs on v_mfma_f32_16x16x32_bf16 (BF16
// activations and weights, FP32 accumulation): M = 16 output pixels along x,
// N = 16 output channels, K = 32 input channels of one tap (or all 9 taps x 2
// channels of the first layer).  Lane layout of the instruction (checked on
// gfx950): A[l%16][(l/16)*8 + t], B[(l/16)*8 + t][l%16], C[(l/16)*4 + t][l%16].
// A workgroup of 4 wavefronts computes an 8 x 16 output tile (2 rows per
// wavefront) from an LDS copy of the 10 x 18 input tile; the back-propagated
// input gradient uses the same kernel with flipped taps and swapped channels.
// The weight gradient is an MFMA GEMM over pixels (M = output channels, N =
// input channels, K = pixels) from channel-major LDS copies of the tile.
//
// Activations are NHWC: pixel (y, x) of image p at ((p*H + y)*W + x)*C, with
// y the block index and x the index inside a block (flat index y*W + x of the
// CN state).
#include "tp_common.h"

namespace tp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TH = 8, TW = 16;           // output tile
constexpr int HR = TH + 2, HC = TW + 2;  // input tile with halo

// BF16 weight buffer layout (elements)
constexpr int OFF_W1 = 0;                   // layer 1 forward, packed [32 o][32 k], k = tap*2 + c (< 18)
constexpr int OFF_W2 = OFF_W1 + 32 * 32;    // layer 2 forward [9][32 o][32 i]
constexpr int OFF_W3 = OFF_W2 + 9 * 32 * 32;
constexpr int OFF_W4 = OFF_W3 + 9 * 32 * 32;  // layer 4 forward [9][16 o (2 used)][32 i]
constexpr int OFF_W4T = OFF_W4 + 9 * 16 * 32;  // layer 4 transposed, packed [32 i][32 k], k = tap'*2 + o
constexpr int OFF_W3T = OFF_W4T + 32 * 32;     // layer 3 transposed [9][32 i][32 o]
constexpr int OFF_W2T = OFF_W3T + 9 * 32 * 32;
constexpr int W_TOTAL = OFF_W2T + 9 * 32 * 32;

// complex parameter layout: layer l has [9][OC][IC] complex weights
__host__ __device__ constexpr int layer_oc(int l) { return l == 3 ? 1 : 16; }
__host__ __device__ constexpr int layer_ic(int l) { return l == 0 ? 1 : 16; }
__host__ __device__ constexpr int layer_off(int l) {
  return l == 0 ? 0 : (l == 1 ? 144 : (l == 2 ? 144 + 2304 : 144 + 2 * 2304));
}
constexpr int NPARAM = 144 + 2 * 2304 + 144;
// real gradient layout: layer l has [9][2*OC][2*IC] floats
__host__ __device__ constexpr int grad_off(int l) { return 4 * layer_off(l); }
constexpr int NGRAD = 4 * NPARAM;

struct ConvArgs {
  const bf16* in;
  const bf16* w;
  bf16* out;
  float* outc;
  const float* scale;
  const int* skip;  // optional per-image flags: images with a nonzero flag are not computed
  int P, H, W, tiles_x, tiles_y;
};

template <int CIN, int COUT, bool FINAL>
__global__ __launch_bounds__(256) void conv3x3_kernel(const ConvArgs a) {
  constexpr int CINP = CIN == 32 ? 40 : 2;  // LDS pixel stride (bf16)
  constexpr int NT = (COUT + 15) / 16;
  constexpr int WROWS = NT * 16;
  constexpr int WSZ = CIN == 32 ? 9 * WROWS * 32 : WROWS * 32;
  __shared__ __align__(16) bf16 in_t[HR * HC * CINP];
  __shared__ __align__(16) bf16 w_l[WSZ];
  __shared__ __align__(16) bf16 o_t[FINAL ? 8 : TH * TW * 32];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kg = lane >> 4;
  for (int k = tid * 8; k < WSZ; k += 256 * 8) *(bf16x8*)&w_l[k] = *(const bf16x8*)&a.w[k];
  const int per_img = a.tiles_y * a.tiles_x;
  const int ntiles = a.P * per_img;
  const bf16 bz = (bf16)0.f;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int img = tile / per_img;
    if (a.skip && a.skip[img] != 0) continue;  // uniform over the workgroup
    const int rem = tile - img * per_img;
    const int y0 = (rem / a.tiles_x) * TH, x0 = (rem % a.tiles_x) * TW;
    __syncthreads();
    if (CIN == 32) {
      for (int c = tid; c < HR * HC * 4; c += 256) {
        const int px = c >> 2, q = c & 3;
        const int r = px / HC, cc = px - r * HC;
        const int y = y0 + r - 1, x = x0 + cc - 1;
        bf16x8 v = {bz, bz, bz, bz, bz, bz, bz, bz};
        if (y >= 0 && y < a.H && x >= 0 && x < a.W)
          v = *(const bf16x8*)&a.in[(((size_t)img * a.H + y) * a.W + x) * 32 + q * 8];
        *(bf16x8*)&in_t[px * CINP + q * 8] = v;
      }
    } else {
      for (int px = tid; px < HR * HC; px += 256) {
        const int r = px / HC, cc = px - r * HC;
        const int y = y0 + r - 1, x = x0 + cc - 1;
        bf16 v0 = bz, v1 = bz;
        if (y >= 0 && y < a.H && x >= 0 && x < a.W) {
          const size_t o = (((size_t)img * a.H + y) * a.W + x) * 2;
          v0 = a.in[o];
          v1 = a.in[o + 1];
        }
        in_t[px * 2] = v0;
        in_t[px * 2 + 1] = v1;
      }
    }
    __syncthreads();
    f32x4 acc[2][NT];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (CIN == 32) {
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3, dx = tap % 3;
        bf16x8 A[2];
#pragma unroll
        for (int m = 0; m < 2; ++m)
          A[m] = *(const bf16x8*)&in_t[(((wave * 2 + m) + dy) * HC + col + dx) * CINP + kg * 8];
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          const bf16x8 B = *(const bf16x8*)&w_l[(tap * WROWS + n * 16 + col) * 32 + kg * 8];
#pragma unroll
          for (int m = 0; m < 2; ++m) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][n], 0, 0, 0);
        }
      }
    } else {
      bf16x8 A[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const int k = kg * 8 + t;
          bf16 v = bz;
          if (k < 18) {
            const int tap = k >> 1, ch = k & 1;
            v = in_t[(((wave * 2 + m) + tap / 3) * HC + col + tap % 3) * 2 + ch];
          }
          A[m][t] = v;
        }
      }
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const bf16x8 B = *(const bf16x8*)&w_l[(n * 16 + col) * 32 + kg * 8];
#pragma unroll
        for (int m = 0; m < 2; ++m) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[m], B, acc[m][n], 0, 0, 0);
      }
    }
    if (FINAL) {  // channels 0 (re) and 1 (im) of N-tile 0, scaled back, to complex64
      if (col < 2) {
        const float sc = a.scale[img];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int y = y0 + wave * 2 + m, x = x0 + kg * 4 + t;
            if (y < a.H && x < a.W) a.outc[(((size_t)img * a.H + y) * a.W + x) * 2 + col] = acc[m][0][t] * sc;
          }
      }
    } else {
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            o_t[((wave * 2 + m) * TW + kg * 4 + t) * 32 + n * 16 + col] = (bf16)acc[m][n][t];
      __syncthreads();
      for (int c = tid; c < TH * TW * 4; c += 256) {
        const int pix = c >> 2, q = c & 3;
        const int y = y0 + pix / TW, x = x0 + pix % TW;
        if (y < a.H && x < a.W)
          *(bf16x8*)&a.out[(((size_t)img * a.H + y) * a.W + x) * 32 + q * 8] = *(const bf16x8*)&o_t[pix * 32 + q * 8];
      }
    }
  }
}

struct WgradArgs {
  const bf16* in;    // NHWC, CIN channels
  const bf16* dout;  // NHWC, COUT channels
  float* g;          // [9][COUT][CIN], accumulated
  int P, H, W, tiles_x, tiles_y;
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(const WgradArgs a) {
  constexpr int NTO = (COUT + 15) / 16, NTI = (CIN + 15) / 16;
  constexpr int CO = NTO * 16, CI = NTI * 16;
  constexpr int NB = 9 * NTO * NTI;
  constexpr int MAXB = (NB + 3) / 4;
  constexpr int NPIX = TH * TW;
  __shared__ __align__(16) bf16 dT[CO * NPIX];          // [o][pixel]
  __shared__ __align__(16) bf16 inT[3 * CI * HR * TW];  // [dx][i][R * 16 + c] = in(R, c + dx)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kg = lane >> 4;
  const bf16 bz = (bf16)0.f;
  f32x4 acc[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int per_img = a.tiles_y * a.tiles_x;
  const int ntiles = a.P * per_img;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int img = tile / per_img;
    const int rem = tile - img * per_img;
    const int y0 = (rem / a.tiles_x) * TH, x0 = (rem % a.tiles_x) * TW;
    __syncthreads();
    for (int e = tid; e < CO * NPIX; e += 256) {
      const int o = e / NPIX, pix = e - o * NPIX;
      const int y = y0 + pix / TW, x = x0 + pix % TW;
      bf16 v = bz;
      if (o < COUT && y < a.H && x < a.W) v = a.dout[(((size_t)img * a.H + y) * a.W + x) * COUT + o];
      dT[e] = v;
    }
    for (int e = tid; e < CI * HR * HC; e += 256) {
      const int i = e / (HR * HC), px = e - i * (HR * HC);
      const int R = px / HC, cc = px - R * HC;
      const int y = y0 + R - 1, x = x0 + cc - 1;
      bf16 v = bz;
      if (i < CIN && y >= 0 && y < a.H && x >= 0 && x < a.W) v = a.in[(((size_t)img * a.H + y) * a.W + x) * CIN + i];
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int c = cc - dx;
        if (c >= 0 && c < TW) inT[(dx * CI + i) * (HR * TW) + R * TW + c] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < MAXB; ++bi) {
      const int b = wave + 4 * bi;
      if (b < NB) {
        const int tap = b / (NTO * NTI), r2 = b - tap * (NTO * NTI);
        const int ot = r2 / NTI, it = r2 - ot * NTI;
        const int dy = tap / 3, dx = tap % 3;
#pragma unroll
        for (int kk = 0; kk < NPIX / 32; ++kk) {
          const int k0 = kk * 32 + kg * 8;  // 8 consecutive pixels of one tile row
          const int r = k0 / TW, c = k0 % TW;
          const bf16x8 A = *(const bf16x8*)&dT[(ot * 16 + col) * NPIX + k0];
          const bf16x8 B = *(const bf16x8*)&inT[(dx * CI + it * 16 + col) * (HR * TW) + (r + dy) * TW + c];
          acc[bi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[bi], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int bi = 0; bi < MAXB; ++bi) {
    const int b = wave + 4 * bi;
    if (b < NB) {
      const int tap = b / (NTO * NTI), r2 = b - tap * (NTO * NTI);
      const int ot = r2 / NTI, it = r2 - ot * NTI;
      const int i = it * 16 + col;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int o = ot * 16 + kg * 4 + t;
        if (o < COUT && i < CIN) atomicAdd(&a.g[(tap * COUT + o) * CIN + i], acc[bi][t]);
      }
    }
  }
}

// workgroup sum (256 threads) -> one FP64 atomic per workgroup
__device__ __forceinline__ void wg_atomic_add(double s, double* dst) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(dst, red[0] + red[1] + red[2] + red[3]);
}

// complex64 planes -> BF16 NHWC (re, im) divided by the plane's RMS (kept in scale)
__global__ __launch_bounds__(256) void cnn_norm_kernel(const float2* __restrict__ v, int64_t N,
                                                       double* __restrict__ nrm2, const int* __restrict__ skip) {
  const int p = blockIdx.y;
  if (skip && skip[p] != 0) return;  // uniform over the workgroup
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 x = v[(size_t)p * N + i];
    s += (double)x.x * x.x + (double)x.y * x.y;
  }
  wg_atomic_add(s, &nrm2[p]);
}

__global__ void cnn_pack_kernel(const float2* __restrict__ v, int64_t N, const double* __restrict__ nrm2,
                                bf16* __restrict__ x0, float* __restrict__ scale, const int* __restrict__ skip) {
  const int p = blockIdx.y;
  if (skip && skip[p] != 0) return;
  const double ms = nrm2[p] / (double)N;
  const float s = ms > 0.0 ? (float)sqrt(ms) : 1.f;
  const float inv = 1.f / s;
  if (blockIdx.x == 0 && threadIdx.x == 0) scale[p] = s;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 x = v[(size_t)p * N + i];
    x0[((size_t)p * N + i) * 2] = (bf16)(x.x * inv);
    x0[((size_t)p * N + i) * 2 + 1] = (bf16)(x.y * inv);
  }
}

// out = M(v) (already multiplied back by s): with d = (out - tgt)/s, the raw network
// output minus tgt/s, loss += sum |d|^2 / (P N) and dout = 2 d / (P N) as BF16 NHWC (2 channels)
__global__ __launch_bounds__(256) void cnn_loss_kernel(const float2* __restrict__ out, const float2* __restrict__ tgt,
                                                       int64_t N, int P, const float* __restrict__ scale,
                                                       bf16* __restrict__ dout, double* __restrict__ loss) {
  const int p = blockIdx.y;
  const float inv = 1.f / scale[p];
  const double norm = 1.0 / ((double)P * (double)N);
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const size_t o = (size_t)p * N + i;
    const float2 d = make_float2((out[o].x - tgt[o].x) * inv, (out[o].y - tgt[o].y) * inv);
    s += (double)d.x * d.x + (double)d.y * d.y;
    dout[o * 2] = (bf16)(2.f * (float)norm * d.x);
    dout[o * 2 + 1] = (bf16)(2.f * (float)norm * d.y);
  }
  wg_atomic_add(s * norm, loss);
}

// real weight (o_r, i_r) of complex weight w: [[a, -b], [b, a]]
__device__ __forceinline__ float real_entry(float2 w, int al, int be) {
  return al == be ? w.x : (al == 1 ? w.y : -w.y);
}

__device__ __forceinline__ void expand_one(const float2 w, int l, int tap, int oc, int ic, bf16* wb) {
#pragma unroll
  for (int al = 0; al < 2; ++al)
#pragma unroll
    for (int be = 0; be < 2; ++be) {
      const bf16 v = (bf16)real_entry(w, al, be);
      const int o = 2 * oc + al, i = 2 * ic + be;
      if (l == 0) {
        wb[OFF_W1 + o * 32 + tap * 2 + i] = v;
      } else if (l == 3) {
        wb[OFF_W4 + (tap * 16 + o) * 32 + i] = v;
        wb[OFF_W4T + i * 32 + (8 - tap) * 2 + o] = v;
      } else {
        const int fw = l == 1 ? OFF_W2 : OFF_W3, tw = l == 1 ? OFF_W2T : OFF_W3T;
        wb[fw + (tap * 32 + o) * 32 + i] = v;
        wb[tw + ((8 - tap) * 32 + i) * 32 + o] = v;
      }
    }
}

__device__ __forceinline__ void param_index(int q, int& l, int& tap, int& oc, int& ic) {
  l = q < layer_off(1) ? 0 : (q < layer_off(2) ? 1 : (q < layer_off(3) ? 2 : 3));
  const int r = q - layer_off(l);
  const int OC = layer_oc(l), IC = layer_ic(l);
  tap = r / (OC * IC);
  const int r2 = r - tap * OC * IC;
  oc = r2 / IC;
  ic = r2 - oc * IC;
}

__global__ void cnn_expand_kernel(const float2* __restrict__ params, bf16* __restrict__ wb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  int l, tap, oc, ic;
  param_index(q, l, tap, oc, ic);
  expand_one(params[q], l, tap, oc, ic, wb);
}

// Adam on (a, b) of every complex weight from the real gradients, then re-expand.
__global__ void cnn_adam_kernel(float2* __restrict__ params, const float* __restrict__ gr, float2* __restrict__ m,
                                float2* __restrict__ v, float2* __restrict__ gout, int step, float lr, float b1,
                                float b2, float eps, bf16* __restrict__ wb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  int l, tap, oc, ic;
  param_index(q, l, tap, oc, ic);
  const int OC2 = 2 * layer_oc(l), IC2 = 2 * layer_ic(l);
  const float* G = gr + grad_off(l) + (size_t)tap * OC2 * IC2;
  const int o = 2 * oc, i = 2 * ic;
  const float ga = G[o * IC2 + i] + G[(o + 1) * IC2 + i + 1];
  const float gb = G[(o + 1) * IC2 + i] - G[o * IC2 + i + 1];
  if (gout) gout[q] = make_float2(ga, gb);
  if (step <= 0) return;
  float2 mm = m[q], vv = v[q], w = params[q];
  mm.x = b1 * mm.x + (1.f - b1) * ga;
  mm.y = b1 * mm.y + (1.f - b1) * gb;
  vv.x = b2 * vv.x + (1.f - b2) * ga * ga;
  vv.y = b2 * vv.y + (1.f - b2) * gb * gb;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  w.x -= lr * (mm.x / c1) / (sqrtf(vv.x / c2) + eps);
  w.y -= lr * (mm.y / c1) / (sqrtf(vv.y / c2) + eps);
  m[q] = mm;
  v[q] = vv;
  params[q] = w;
  expand_one(w, l, tap, oc, ic, wb);
}

// ------------------------------------------------------------------ host helpers
struct CnnWork {
  bf16 *x0, *x1, *x2, *x3, *d;  // activations (x0: 2 channels, others 32) and gradient scratch (32)
  bf16* d2;                      // second gradient buffer (32 channels)
  bf16* dl;                      // loss gradient (2 channels)
  float2* out;                   // network output (training)
  float* scale;
  double* nrm2;
  float* gr;                     // real weight gradients
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t cnn_work_layout(int P, int64_t N, int train, CnnWork* w, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align256(bytes);
    return p;
  };
  const size_t PN = (size_t)P * N;
  char* x0 = take(PN * 2 * sizeof(bf16));
  char* x1 = take(PN * 32 * sizeof(bf16));
  char* x2 = take(PN * 32 * sizeof(bf16));
  char* x3 = train ? take(PN * 32 * sizeof(bf16)) : nullptr;
  char* d = train ? take(PN * 32 * sizeof(bf16)) : nullptr;
  char* d2 = train ? take(PN * 32 * sizeof(bf16)) : nullptr;
  char* dl = train ? take(PN * 2 * sizeof(bf16)) : nullptr;
  char* out = train ? take(PN * sizeof(float2)) : nullptr;
  char* scale = take((size_t)P * sizeof(float));
  char* nrm2 = take((size_t)P * sizeof(double));
  char* gr = train ? take((size_t)NGRAD * sizeof(float)) : nullptr;
  if (w) {
    w->x0 = (bf16*)x0;
    w->x1 = (bf16*)x1;
    w->x2 = (bf16*)x2;
    w->x3 = (bf16*)(train ? x3 : x1);  // inference reuses x1 for layer 3's output
    w->d = (bf16*)d;
    w->d2 = (bf16*)d2;
    w->dl = (bf16*)dl;
    w->out = (float2*)out;
    w->scale = (float*)scale;
    w->nrm2 = (double*)nrm2;
    w->gr = (float*)gr;
  }
  return off;
}

template <int CIN, int COUT, bool FINAL>
static void launch_conv(const bf16* in, const bf16* w, bf16* out, float* outc, const float* scale, int P, int nx,
                        int ny, hipStream_t st, const int* skip = nullptr) {
  ConvArgs a;
  a.skip = skip;
  a.in = in;
  a.w = w;
  a.out = out;
  a.outc = outc;
  a.scale = scale;
  a.P = P;
  a.H = ny;
  a.W = nx;
  a.tiles_x = (nx + TW - 1) / TW;
  a.tiles_y = (ny + TH - 1) / TH;
  const int ntiles = P * a.tiles_x * a.tiles_y;
  const int grid = ntiles < 2048 ? ntiles : 2048;
  hipLaunchKernelGGL((conv3x3_kernel<CIN, COUT, FINAL>), dim3(grid), dim3(256), 0, st, a);
}

template <int CIN, int COUT>
static void launch_wgrad(const bf16* in, const bf16* dout, float* g, int P, int nx, int ny, hipStream_t st) {
  WgradArgs a;
  a.in = in;
  a.dout = dout;
  a.g = g;
  a.P = P;
  a.H = ny;
  a.W = nx;
  a.tiles_x = (nx + TW - 1) / TW;
  a.tiles_y = (ny + TH - 1) / TH;
  const int ntiles = P * a.tiles_x * a.tiles_y;
  const int grid = ntiles < 256 ? ntiles : 256;
  hipLaunchKernelGGL((conv_wgrad_kernel<CIN, COUT>), dim3(grid), dim3(256), 0, st, a);
}

__global__ void zero_bytes_kernel(uint32_t* p, size_t words) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

static void zero_words(void* p, size_t bytes, hipStream_t st) {
  const size_t words = bytes / 4;
  const int grid = (int)((words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024);
  hipLaunchKernelGGL(zero_bytes_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, (uint32_t*)p, words);
}

// forward pass through the four layers; x3 holds layer 3's output, the result goes to out (complex64)
// `skip` (optional, per plane): planes with a nonzero flag are left unwritten
static void cnn_forward(const bf16* wb, const float2* in, float2* out, int P, int nx, int ny, const CnnWork& w,
                        hipStream_t st, const int* skip = nullptr) {
  const int64_t N = (int64_t)nx * ny;
  zero_words(w.nrm2, (size_t)P * sizeof(double), st);
  const dim3 g(N / 256 + 1 < 16 ? (unsigned)(N / 256 + 1) : 16u, (unsigned)P);
  hipLaunchKernelGGL(cnn_norm_kernel, g, dim3(256), 0, st, in, N, w.nrm2, skip);
  hipLaunchKernelGGL(cnn_pack_kernel, g, dim3(256), 0, st, in, N, w.nrm2, w.x0, w.scale, skip);
  launch_conv<2, 32, false>(w.x0, wb + OFF_W1, w.x1, nullptr, nullptr, P, nx, ny, st, skip);
  launch_conv<32, 32, false>(w.x1, wb + OFF_W2, w.x2, nullptr, nullptr, P, nx, ny, st, skip);
  launch_conv<32, 32, false>(w.x2, wb + OFF_W3, w.x3, nullptr, nullptr, P, nx, ny, st, skip);
  launch_conv<32, 2, true>(w.x3, wb + OFF_W4, nullptr, (float*)out, w.scale, P, nx, ny, st, skip);
}

static int check_cnn_dims(int P, int nx, int ny) {
  if (P < 1 || nx < 1 || ny < 1) {
    set_error("CNN: invalid sizes (P=%d, nx=%d, ny=%d)", P, nx, ny);
    return TP_ERR_INVALID;
  }
  return TP_OK;
}

// entry points for the Krylov solver in the same library
int cnn_apply_internal(const void* wb, const float2* in, float2* out, int P, int nx, int ny, void* work,
                       hipStream_t st, const int* skip) {
  CnnWork w;
  cnn_work_layout(P, (int64_t)nx * ny, 0, &w, (char*)work);
  cnn_forward((const bf16*)wb, in, out, P, nx, ny, w, st, skip);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

size_t cnn_work_bytes_internal(int P, int nx, int ny) {
  return cnn_work_layout(P, (int64_t)nx * ny, 0, nullptr, nullptr);
}

}  // namespace tp

using namespace tp;

extern "C" int tp_cnn_param_count(void) { return NPARAM; }
extern "C" int64_t tp_cnn_weight_bytes(void) { return (int64_t)W_TOTAL * sizeof(bf16); }
extern "C" int64_t tp_cnn_workspace_bytes(int P, int nx, int ny, int train) {
  return (int64_t)cnn_work_layout(P, (int64_t)nx * ny, train, nullptr, nullptr);
}

extern "C" int tp_cnn_expand(const void* d_params, void* d_wbuf, void* stream) {
  if (!d_params || !d_wbuf) {
    set_error("tp_cnn_expand: null pointer");
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  zero_words(d_wbuf, (size_t)W_TOTAL * sizeof(bf16), st);
  hipLaunchKernelGGL(cnn_expand_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, (const float2*)d_params,
                     (bf16*)d_wbuf);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_cnn_apply(const void* d_wbuf, const void* d_in, void* d_out, int P, int nx, int ny, void* d_work,
                            void* stream) {
  int rc = check_cnn_dims(P, nx, ny);
  if (rc) return rc;
  if (!d_wbuf || !d_in || !d_out || !d_work) {
    set_error("tp_cnn_apply: null pointer");
    return TP_ERR_INVALID;
  }
  CnnWork w;
  cnn_work_layout(P, (int64_t)nx * ny, 0, &w, (char*)d_work);
  cnn_forward((const bf16*)d_wbuf, (const float2*)d_in, (float2*)d_out, P, nx, ny, w, (hipStream_t)stream);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_cnn_train_step(void* d_params, void* d_wbuf, void* d_m, void* d_v, const void* d_in,
                                 const void* d_target, int P, int nx, int ny, int step, double lr, double beta1,
                                 double beta2, double eps, void* d_work, double* d_loss, void* d_grad, void* stream) {
  int rc = check_cnn_dims(P, nx, ny);
  if (rc) return rc;
  if (!d_params || !d_wbuf || !d_in || !d_target || !d_work || !d_loss || (step > 0 && (!d_m || !d_v))) {
    set_error("tp_cnn_train_step: null pointer");
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = (int64_t)nx * ny;
  CnnWork w;
  cnn_work_layout(P, N, 1, &w, (char*)d_work);
  const bf16* wb = (const bf16*)d_wbuf;
  cnn_forward(wb, (const float2*)d_in, w.out, P, nx, ny, w, st);
  zero_words(d_loss, sizeof(double), st);
  const dim3 g(N / 256 + 1 < 16 ? (unsigned)(N / 256 + 1) : 16u, (unsigned)P);
  hipLaunchKernelGGL(cnn_loss_kernel, g, dim3(256), 0, st, w.out, (const float2*)d_target, N, P, w.scale, w.dl,
                     d_loss);
  zero_words(w.gr, (size_t)NGRAD * sizeof(float), st);
  // layer 4: dW4 from (x3, dl); d = dL/dx3
  launch_wgrad<32, 2>(w.x3, w.dl, w.gr + grad_off(3), P, nx, ny, st);
  launch_conv<2, 32, false>(w.dl, wb + OFF_W4T, w.d, nullptr, nullptr, P, nx, ny, st);
  // layer 3
  launch_wgrad<32, 32>(w.x2, w.d, w.gr + grad_off(2), P, nx, ny, st);
  launch_conv<32, 32, false>(w.d, wb + OFF_W3T, w.d2, nullptr, nullptr, P, nx, ny, st);
  // layer 2
  launch_wgrad<32, 32>(w.x1, w.d2, w.gr + grad_off(1), P, nx, ny, st);
  launch_conv<32, 32, false>(w.d2, wb + OFF_W2T, w.d, nullptr, nullptr, P, nx, ny, st);
  // layer 1
  launch_wgrad<2, 32>(w.x0, w.d, w.gr + grad_off(0), P, nx, ny, st);
  hipLaunchKernelGGL(cnn_adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, st, (float2*)d_params, w.gr,
                     (float2*)d_m, (float2*)d_v, (float2*)d_grad, step, (float)lr, (float)beta1, (float)beta2,
                     (float)eps, (bf16*)d_wbuf);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
