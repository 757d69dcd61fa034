// Two-transverse-axis multislice (2D FFT per slice) and per-probe 2D far field.
//
// Planes are complex64 (A rows x B columns), element (a, b) at a*B + b (a the
// first array axis, b the second; the flat index of the CN state).  A 2D
// transform is two batched row passes, each writing its result transposed
// through an LDS tile so that the next pass again reads contiguous rows:
//
//   multislice slice:  pass 1  rows b:  y = FFT_b(psi * T)             -> [kb][a]
//                      pass 2  rows a:  y = IFFT_a(H * FFT_a(.))       -> [a][kb]
//                      pass 3  rows kb: psi' = IFFT_b(.)               -> [a][b]
//   far field:         pass 1 as above (no T), pass 2: |FFT_a(.)|^2    -> [ka][kb]
//
// H = exp(i (kz - k) dz) / (A B), kz = sqrt(max(k^2 - ka^2 - kb^2, 0)) (the
// one-axis rule of ms_solver.py:58-72 with both transverse frequencies; the
// reference has no two-axis multislice).  One wavefront transforms one row
// (tp_fft.h, narrow LDS padding: these kernels' occupancy is LDS-bound); a
// workgroup of 4 wavefronts handles RPW consecutive rows.  Plane
// sides: powers of two from 64 to 512.
#include "tp_fft.h"

namespace tp {

constexpr int M2_WAVES = 4;

template <int R>
constexpr int m2_rows() {  // rows per workgroup: twiddles + FFT buffers + tile stay below 64 KiB of LDS
  return R <= 4 ? 16 : 4;
}

struct Ms2dPass {
  const float2* in;   // plane p at in + p * in_pstride, [rows][L]
  float2* out;        // transposed [L][rows] (passes 1, 2) or natural [rows][L] (pass 3), plane p at out + p * out_pstride
  float* inten;       // far field output [P][L][rows] (pass 2 of the far field)
  const float2* T;    // [rows][L] transmission (pass 1 of a slice) or NULL
  const float2* H;    // pass 2: H[row][k_this_axis] ([rows][L])
  int P, rows, L;     // rows per plane and row length (L = 64 R)
  int conj_h;
  int64_t in_pstride, out_pstride;
};

// LDS tile rows are L + 1 values apart, so that the column reads of the
// transposed store (consecutive r, fixed k) fall into distinct banks.
__host__ __device__ constexpr int tile_ld(int L) { return L + 1; }

// Transposed store of the workgroup's tile: out[p][k][row0 + r] = tile[r][k]
__device__ __forceinline__ void store_tile_t(const float2* tile, float2* out, int64_t pstride, int p, int rows, int L,
                                             int row0, int nr, int tid, int nt) {
  float2* o = out + (size_t)p * pstride;
  for (int idx = tid; idx < nr * L; idx += nt) {
    const int k = idx / nr, r = idx - k * nr;
    o[(size_t)k * rows + row0 + r] = tile[r * tile_ld(L) + k];
  }
}

// pass 1: y = FFT(row * T), stored transposed with natural frequency index
template <int R>
__global__ __launch_bounds__(64 * M2_WAVES) void ms2d_pass1_kernel(const Ms2dPass a) {
  constexpr int L = 64 * R;
  constexpr int RPW = m2_rows<R>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[M2_WAVES][2][ms_bufn<R, false>()];
  __shared__ float2 tile[RPW * tile_ld(L)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  build_twiddles(tw, L);
  const int groups = (a.rows + RPW - 1) / RPW;
  const int p = blockIdx.x / groups, row0 = (blockIdx.x % groups) * RPW;
  const int nr = min(RPW, a.rows - row0);
  __syncthreads();
  for (int r = wave; r < nr; r += M2_WAVES) {
    const float2* src = a.in + (size_t)p * a.in_pstride + (size_t)(row0 + r) * L;
    const float2* t = a.T ? a.T + (size_t)(row0 + r) * L : nullptr;
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      x[m] = src[lane + 64 * m];
      if (t) x[m] = cmul(x[m], t[lane + 64 * m]);
    }
    fft_any<R, -1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) tile[r * tile_ld(L) + spec_index<R>(lane, m)] = x[m];
  }
  __syncthreads();
  store_tile_t(tile, a.out, a.out_pstride, p, a.rows, L, row0, nr, threadIdx.x, blockDim.x);
}

// pass 2: rows indexed by the other axis' frequency kr (row), along this axis:
//   multislice: y = IFFT(H[k][kr] * FFT(row)) stored transposed (natural index)
//   far field:  |FFT(row)|^2 stored transposed into inten
template <int R, bool FAR>
__global__ __launch_bounds__(64 * M2_WAVES) void ms2d_pass2_kernel(const Ms2dPass a) {
  constexpr int L = 64 * R;
  constexpr int RPW = m2_rows<R>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[M2_WAVES][2][ms_bufn<R, false>()];
  __shared__ float2 tile[RPW * tile_ld(L)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  build_twiddles(tw, L);
  const int groups = (a.rows + RPW - 1) / RPW;
  const int p = blockIdx.x / groups, row0 = (blockIdx.x % groups) * RPW;
  const int nr = min(RPW, a.rows - row0);
  __syncthreads();
  for (int r = wave; r < nr; r += M2_WAVES) {
    const int kr = row0 + r;
    const float2* src = a.in + (size_t)p * a.in_pstride + (size_t)kr * L;
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = src[lane + 64 * m];
    fft_any<R, -1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
    if (FAR) {
      float* ti = reinterpret_cast<float*>(tile);
#pragma unroll
      for (int m = 0; m < R; ++m) ti[r * tile_ld(L) + spec_index<R>(lane, m)] = fmaf(x[m].x, x[m].x, x[m].y * x[m].y);
    } else {
      // H is [kr][this-axis frequency]: one row per pass-2 row, coalesced
#pragma unroll
      for (int m = 0; m < R; ++m) {
        float2 h = a.H[(size_t)kr * L + spec_index<R>(lane, m)];
        if (a.conj_h) h = cconj(h);
        x[m] = cmul(x[m], h);
      }
      fft_any<R, +1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
      for (int m = 0; m < R; ++m) tile[r * tile_ld(L) + lane + 64 * m] = x[m];
    }
  }
  __syncthreads();
  if (FAR) {
    const float* ti = reinterpret_cast<const float*>(tile);
    float* o = a.inten + (size_t)p * a.rows * L;
    for (int idx = threadIdx.x; idx < nr * L; idx += blockDim.x) {
      const int k = idx / nr, r = idx - k * nr;
      o[(size_t)k * a.rows + row0 + r] = ti[r * tile_ld(L) + k];
    }
  } else {
    store_tile_t(tile, a.out, a.out_pstride, p, a.rows, L, row0, nr, threadIdx.x, blockDim.x);
  }
}

// pass 2 of the amplitude seed: rows kr (the other axis' frequency); U = FFT(row),
// loss[p] += sum (|U| - a)^2, y = IFFT(scale (|U| - a) U / |U|) stored transposed.
// amp: [P][L][rows] (a at frequency (k_this, kr)).
template <int R>
__global__ __launch_bounds__(64 * M2_WAVES) void ms2d_seed_kernel(const Ms2dPass a, const float* __restrict__ amp,
                                                                  float scale, float* __restrict__ loss) {
  constexpr int L = 64 * R;
  constexpr int RPW = m2_rows<R>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[M2_WAVES][2][ms_bufn<R, false>()];
  __shared__ float2 tile[RPW * tile_ld(L)];
  __shared__ float red[M2_WAVES];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  build_twiddles(tw, L);
  const int groups = (a.rows + RPW - 1) / RPW;
  const int p = blockIdx.x / groups, row0 = (blockIdx.x % groups) * RPW;
  const int nr = min(RPW, a.rows - row0);
  __syncthreads();
  float part = 0.f;
  for (int r = wave; r < nr; r += M2_WAVES) {
    const int kr = row0 + r;
    const float2* src = a.in + (size_t)p * a.in_pstride + (size_t)kr * L;
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = src[lane + 64 * m];
    fft_any<R, -1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const float av = amp[(size_t)p * a.rows * L + (size_t)spec_index<R>(lane, m) * a.rows + kr];
      const float mag = sqrtf(fmaf(x[m].x, x[m].x, x[m].y * x[m].y));
      const float d = mag - av;
      part = fmaf(d, d, part);
      x[m] = cscale(x[m], mag > 0.f ? scale * d / mag : 0.f);
    }
    fft_any<R, +1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) tile[r * tile_ld(L) + lane + 64 * m] = x[m];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if (lane == 0) red[wave] = part;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&loss[p], red[0] + red[1] + red[2] + red[3]);
  store_tile_t(tile, a.out, a.out_pstride, p, a.rows, L, row0, nr, threadIdx.x, blockDim.x);
}

// G[s-1][f] += sum_p conj(lam_s[p][f]) (u_s[p][f] + u_{s-1}[p][f]), s = 1..L, with
// u_s = u[p][s] (full forward field) and lam_s = lam[p][L - s + 1] (full exact-adjoint
// field in sweep order).  One thread per (s, f); FP32 sum over the probes.
__global__ void grad3d_kernel(const float2* __restrict__ u, const float2* __restrict__ lam, int P, int L, int64_t N,
                              float2* __restrict__ G) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y + 1;
  if (f >= N) return;
  const int64_t ps = (int64_t)(L + 1) * N;
  float2 acc = make_float2(0.f, 0.f);
  for (int p = 0; p < P; ++p) {
    const float2 a = u[(size_t)p * ps + (size_t)s * N + f];
    const float2 b = u[(size_t)p * ps + (size_t)(s - 1) * N + f];
    const float2 l = lam[(size_t)p * ps + (size_t)(L - s + 1) * N + f];
    acc = cfma(cconj(l), cadd(a, b), acc);
  }
  float2* g = G + (size_t)(s - 1) * N + f;
  *g = cadd(*g, acc);
}

// G[z][f] += sum_p conj(a[p][a_z0 + a_zs z][f]) b[p][b_z0 + b_zs z][f]  (planes of N values)
__global__ void plane_conj_dot_kernel(const float2* __restrict__ a, int64_t a_ps, int a_z0, int a_zs,
                                      const float2* __restrict__ b, int64_t b_ps, int b_z0, int b_zs, int P,
                                      int64_t N, float2* __restrict__ G) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int z = blockIdx.y;
  if (f >= N) return;
  const float2* pa = a + (int64_t)(a_z0 + a_zs * z) * N + f;
  const float2* pb = b + (int64_t)(b_z0 + b_zs * z) * N + f;
  float2 acc = make_float2(0.f, 0.f);
  for (int p = 0; p < P; ++p) acc = cfma(cconj(pa[(size_t)p * a_ps]), pb[(size_t)p * b_ps], acc);
  float2* g = G + (size_t)z * N + f;
  *g = cadd(*g, acc);
}

// pass 3: rows hold a spectrum in natural order; gather it into the FFT's
// spectrum order, inverse transform, store natural (not transposed)
template <int R>
__global__ __launch_bounds__(64 * M2_WAVES) void ms2d_pass3_kernel(const Ms2dPass a) {
  constexpr int L = 64 * R;
  constexpr int RPW = m2_rows<R>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[M2_WAVES][2][ms_bufn<R, false>()];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  build_twiddles(tw, L);
  const int groups = (a.rows + RPW - 1) / RPW;
  const int p = blockIdx.x / groups, row0 = (blockIdx.x % groups) * RPW;
  const int nr = min(RPW, a.rows - row0);
  __syncthreads();
  for (int r = wave; r < nr; r += M2_WAVES) {
    const float2* src = a.in + (size_t)p * a.in_pstride + (size_t)(row0 + r) * L;
    float2* dst = a.out + (size_t)p * a.out_pstride + (size_t)(row0 + r) * L;
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = src[spec_index<R>(lane, m)];
    fft_any<R, +1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) dst[lane + 64 * m] = x[m];
  }
}

// pass 3 of slice z fused with pass 1 of slice z + 1 (both run along the rows a):
// psi_{z+1} = IFFT_b(spectrum row), optionally stored (natural order, to
// `store`), then y = FFT_b(psi_{z+1} * T_{z+1}) stored transposed for the next
// pass 2.  The plane between the two never goes through HBM unless it is kept.
template <int R>
__global__ __launch_bounds__(64 * M2_WAVES) void ms2d_pass31_kernel(const Ms2dPass a, float2* store,
                                                                   int64_t store_pstride) {
  constexpr int L = 64 * R;
  constexpr int RPW = m2_rows<R>();
  __shared__ float2 tw[L];
  __shared__ float2 buf[M2_WAVES][2][ms_bufn<R, false>()];
  __shared__ float2 tile[RPW * tile_ld(L)];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  build_twiddles(tw, L);
  const int groups = (a.rows + RPW - 1) / RPW;
  const int p = blockIdx.x / groups, row0 = (blockIdx.x % groups) * RPW;
  const int nr = min(RPW, a.rows - row0);
  __syncthreads();
  for (int r = wave; r < nr; r += M2_WAVES) {
    const float2* src = a.in + (size_t)p * a.in_pstride + (size_t)(row0 + r) * L;
    const float2* t = a.T + (size_t)(row0 + r) * L;
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = src[spec_index<R>(lane, m)];
    fft_any<R, +1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
    if (store) {
      float2* dst = store + (size_t)p * store_pstride + (size_t)(row0 + r) * L;
#pragma unroll
      for (int m = 0; m < R; ++m) dst[lane + 64 * m] = x[m];
    }
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = cmul(x[m], t[lane + 64 * m]);
    fft_any<R, -1, false>(x, buf[wave][0], buf[wave][1], tw, lane);
#pragma unroll
    for (int m = 0; m < R; ++m) tile[r * tile_ld(L) + spec_index<R>(lane, m)] = x[m];
  }
  __syncthreads();
  store_tile_t(tile, a.out, a.out_pstride, p, a.rows, L, row0, nr, threadIdx.x, blockDim.x);
}

__global__ void ms2d_kernel_kernel(int A, int B, double da, double db, double k, double dz, float2* __restrict__ H) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)A * B) return;
  const int ia = (int)(id / B), ib = (int)(id % B);
  const int fa = ia < (A + 1) / 2 ? ia : ia - A;  // numpy.fft.fftfreq ordering
  const int fb = ib < (B + 1) / 2 ? ib : ib - B;
  const double ka = 2.0 * M_PI * fa / (A * da), kb = 2.0 * M_PI * fb / (B * db);
  const double inside = k * k - ka * ka - kb * kb;
  const double kz = inside > 0.0 ? sqrt(inside) : 0.0;
  double s, c;
  sincos((kz - k) * dz, &s, &c);
  const double sc = 1.0 / ((double)A * B);
  H[(size_t)ib * A + ia] = make_float2((float)(c * sc), (float)(s * sc));  // [kb][ka]
}

// dst[p][i] = src[p][i] for P planes of N values with the given plane strides
__global__ void ms2d_copy_kernel(const float2* __restrict__ src, int64_t sps, float2* __restrict__ dst, int64_t dps,
                                 int64_t N) {
  const int p = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
    dst[(size_t)p * dps + i] = src[(size_t)p * sps + i];
}

static bool m2_size_ok(int n) { return n >= 64 && n <= 512 && (n & (n - 1)) == 0; }

template <int R>
static void launch_p1(const Ms2dPass& a, hipStream_t st) {
  const int groups = (a.rows + m2_rows<R>() - 1) / m2_rows<R>();
  hipLaunchKernelGGL((ms2d_pass1_kernel<R>), dim3(a.P * groups), dim3(64 * M2_WAVES), 0, st, a);
}
template <int R, bool FAR>
static void launch_p2(const Ms2dPass& a, hipStream_t st) {
  const int groups = (a.rows + m2_rows<R>() - 1) / m2_rows<R>();
  hipLaunchKernelGGL((ms2d_pass2_kernel<R, FAR>), dim3(a.P * groups), dim3(64 * M2_WAVES), 0, st, a);
}
template <int R>
static void launch_p3(const Ms2dPass& a, hipStream_t st) {
  const int groups = (a.rows + m2_rows<R>() - 1) / m2_rows<R>();
  hipLaunchKernelGGL((ms2d_pass3_kernel<R>), dim3(a.P * groups), dim3(64 * M2_WAVES), 0, st, a);
}
template <int R>
static void launch_p31(const Ms2dPass& a, float2* store, int64_t store_pstride, hipStream_t st) {
  const int groups = (a.rows + m2_rows<R>() - 1) / m2_rows<R>();
  hipLaunchKernelGGL((ms2d_pass31_kernel<R>), dim3(a.P * groups), dim3(64 * M2_WAVES), 0, st, a, store,
                     store_pstride);
}

#define TP_M2_DISPATCH(L, CALL)                      \
  switch ((L) / 64) {                               \
    case 1: { constexpr int RV = 1; CALL; } break;   \
    case 2: { constexpr int RV = 2; CALL; } break;   \
    case 4: { constexpr int RV = 4; CALL; } break;   \
    case 8: { constexpr int RV = 8; CALL; } break;   \
  }

}  // namespace tp

using namespace tp;

extern "C" int tp_ms2d_kernel(int nx, int ny, double dx, double dy, double k, double dz, void* d_H, void* stream) {
  if (!d_H || nx < 1 || ny < 1) {
    set_error("tp_ms2d_kernel: invalid arguments");
    return TP_ERR_INVALID;
  }
  const int64_t n = (int64_t)nx * ny;
  hipLaunchKernelGGL(ms2d_kernel_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, nx, ny,
                     dx, dy, k, dz, (float2*)d_H);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_ms2d_sweep(const void* d_T, const void* d_H, const void* d_u0, void* d_out, void* d_exit, int nx,
                             int ny, int steps, int probes, int conj_h, void* d_work, void* stream) {
  if (!m2_size_ok(nx) || !m2_size_ok(ny)) {
    set_error("tp_ms2d_sweep: plane %d x %d not supported (powers of two from 64 to 512)", nx, ny);
    return TP_ERR_UNSUPPORTED;
  }
  if (!d_T || !d_H || !d_u0 || !d_work || steps < 1 || probes < 0 || (!d_out && !d_exit)) {
    set_error("tp_ms2d_sweep: invalid arguments (steps=%d, probes=%d)", steps, probes);
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = (int64_t)nx * ny, PN = (int64_t)probes * N;
  float2* w1 = (float2*)d_work;  // transposed intermediate, plane p at w1 + p N ([ny][nx])
  float2* w2 = w1 + PN;          // spectrum along y, [nx][ny]
  float2* psi = w2 + PN;         // last plane when no full field is written
  float2* out = (float2*)d_out;
  const int64_t ops = (int64_t)(steps + 1) * N;  // full-field plane stride of one probe
  const dim3 cg(64, (unsigned)probes);
  if (out) hipLaunchKernelGGL(ms2d_copy_kernel, cg, dim3(256), 0, st, (const float2*)d_u0, N, out, ops, N);
  // Two passes per slice: pass 2 (along x) and pass 3 of the slice fused with
  // pass 1 of the next (both along y); pass 1 runs alone before the first slice
  // and pass 3 alone after the last.
  Ms2dPass a1{(const float2*)d_u0, w1, nullptr, (const float2*)d_T, nullptr, probes, nx, ny, 0, N, N};
  TP_M2_DISPATCH(ny, launch_p1<RV>(a1, st));
  for (int z = 0; z < steps; ++z) {
    Ms2dPass a2{w1, w2, nullptr, nullptr, (const float2*)d_H, probes, ny, nx, conj_h, N, N};
    TP_M2_DISPATCH(nx, (launch_p2<RV, false>(a2, st)));
    if (z + 1 < steps) {
      Ms2dPass a31{w2, w1, nullptr, (const float2*)d_T + (size_t)(z + 1) * N, nullptr, probes, nx, ny, 0, N, N};
      float2* keep = out ? out + (size_t)(z + 1) * N : nullptr;
      TP_M2_DISPATCH(ny, launch_p31<RV>(a31, keep, ops, st));
    } else {
      float2* last = out ? out + (size_t)(z + 1) * N : psi;
      Ms2dPass a3{w2, last, nullptr, nullptr, nullptr, probes, nx, ny, 0, N, out ? ops : N};
      TP_M2_DISPATCH(ny, launch_p3<RV>(a3, st));
      if (d_exit)
        hipLaunchKernelGGL(ms2d_copy_kernel, cg, dim3(256), 0, st, (const float2*)last, out ? ops : N,
                           (float2*)d_exit, N, N);
    }
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_farfield2d(const void* d_exit, int nx, int ny, int probes, void* d_intensity, void* d_work,
                             void* stream) {
  if (!m2_size_ok(nx) || !m2_size_ok(ny)) {
    set_error("tp_farfield2d: plane %d x %d not supported (powers of two from 64 to 512)", nx, ny);
    return TP_ERR_UNSUPPORTED;
  }
  if (!d_exit || !d_intensity || !d_work || probes < 0) {
    set_error("tp_farfield2d: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = (int64_t)nx * ny;
  Ms2dPass a1{(const float2*)d_exit, (float2*)d_work, nullptr, nullptr, nullptr, probes, nx, ny, 0, N, N};
  TP_M2_DISPATCH(ny, launch_p1<RV>(a1, st));
  Ms2dPass a2{(const float2*)d_work, nullptr, (float*)d_intensity, nullptr, nullptr, probes, ny, nx, 0, N, N};
  TP_M2_DISPATCH(nx, (launch_p2<RV, true>(a2, st)));
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_amplitude_seed2d(const void* d_exit, const void* d_sqrt_i, int nx, int ny, int probes, double scale,
                                   void* d_seed, void* d_loss, void* d_work, void* stream) {
  if (!m2_size_ok(nx) || !m2_size_ok(ny)) {
    set_error("tp_amplitude_seed2d: plane %d x %d not supported (powers of two from 64 to 512)", nx, ny);
    return TP_ERR_UNSUPPORTED;
  }
  if (!d_exit || !d_sqrt_i || !d_seed || !d_loss || !d_work || probes < 0) {
    set_error("tp_amplitude_seed2d: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = (int64_t)nx * ny, PN = (int64_t)probes * N;
  float2* w1 = (float2*)d_work;
  float2* w2 = w1 + PN;
  Ms2dPass a1{(const float2*)d_exit, w1, nullptr, nullptr, nullptr, probes, nx, ny, 0, N, N};
  TP_M2_DISPATCH(ny, launch_p1<RV>(a1, st));
  Ms2dPass a2{w1, w2, nullptr, nullptr, nullptr, probes, ny, nx, 0, N, N};
  switch (nx / 64) {
#define TP_SEED_CASE(RV)                                                                                     \
  case RV: {                                                                                                \
    const int groups = (ny + m2_rows<RV>() - 1) / m2_rows<RV>();                                              \
    hipLaunchKernelGGL((ms2d_seed_kernel<RV>), dim3(probes * groups), dim3(64 * M2_WAVES), 0, st, a2,        \
                       (const float*)d_sqrt_i, (float)scale, (float*)d_loss);                                \
  } break;
    TP_SEED_CASE(1)
    TP_SEED_CASE(2)
    TP_SEED_CASE(4)
    TP_SEED_CASE(8)
#undef TP_SEED_CASE
  }
  Ms2dPass a3{w2, (float2*)d_seed, nullptr, nullptr, nullptr, probes, nx, ny, 0, N, N};
  TP_M2_DISPATCH(ny, launch_p3<RV>(a3, st));
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_plane_conj_dot(const void* d_a, int64_t a_pstride, int a_z0, int a_zstep, const void* d_b,
                                 int64_t b_pstride, int b_z0, int b_zstep, int probes, int nplanes, int64_t N,
                                 void* d_G, void* stream) {
  if (!d_a || !d_b || !d_G || probes < 0 || nplanes < 1 || N < 1 || nplanes > 65535) {
    set_error("tp_plane_conj_dot: invalid arguments (probes=%d, nplanes=%d)", probes, nplanes);
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  hipLaunchKernelGGL(plane_conj_dot_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)nplanes), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)d_a, a_pstride, a_z0, a_zstep, (const float2*)d_b,
                     b_pstride, b_z0, b_zstep, probes, N, (float2*)d_G);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_grad3d_accumulate(const void* d_u, const void* d_lam, int probes, int steps, int64_t N, void* d_G,
                                    void* stream) {
  if (!d_u || !d_lam || !d_G || probes < 0 || steps < 1 || N < 1) {
    set_error("tp_grad3d_accumulate: invalid arguments");
    return TP_ERR_INVALID;
  }
  if (probes == 0) return TP_OK;
  hipLaunchKernelGGL(grad3d_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)steps), dim3(256), 0,
                     (hipStream_t)stream, (const float2*)d_u, (const float2*)d_lam, probes, steps, N, (float2*)d_G);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
