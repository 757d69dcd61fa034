// Crank–Nicolson slice stepping on wide transverse grids (one transverse axis):
// one probe per workgroup of W wavefronts, for widths beyond the 64 x 32 rows
// one wavefront holds (the reference steps N = 12001 and N = 7400 grids).
//
// The rows are split over T = 64 W lanes of CH consecutive rows (lane g of the
// workgroup owns rows [g CH, (g+1) CH)).  Per-step coefficients are stored
// lane-interleaved, row i at (i % CH) T + i / CH, so the m-th coefficient of
// every lane is one coalesced load; with W = 1 this is the layout of the
// one-wavefront kernels (tp_tri.hip).  A step is the same arithmetic as there:
//     u_s = A_s^{-1}(sigma u_{s-1} + b_s) - u_{s-1},
// chunk-local triangular recurrences in registers, then the chunk carries.  A
// carry is an affine map y -> A y + B per lane; the maps are composed by a DPP
// scan inside each wavefront and then across the W wavefronts through LDS
// (each wavefront folds the totals of the wavefronts before it, at most 15
// complex multiply-adds), so a step costs two workgroup barriers, one per
// triangular sweep.  The coefficients come from global memory: one step of a
// 16384-row grid is 640 KiB, beyond LDS; the probes of a scan share them in L2.
//
// Factor: one workgroup of T lanes per step; the FP64 Moebius pivot product of
// tp_tri.hip with its prefix scan continued over the wavefronts through LDS.
#include "tp_scan.h"

namespace tp {

constexpr int WIDE_MAX_WAVES = 16;

// ------------------------------------------------------------------ factor
template <int CH>
__global__ __launch_bounds__(1024) void triw_factor_kernel(const FactorArgs f) {
  extern __shared__ Mob shm[];  // [W][2][64] wavefront scans, then [W] wavefront totals
  const int T = blockDim.x, W = T >> 6;
  const int g = threadIdx.x, wave = g >> 6, lane = g & 63;
  const int s = blockIdx.x;  // grid = steps: every workgroup reaches the barrier below
  Mob* scan0 = shm + wave * 128;
  Mob* tot = shm + 128 * W;
  const int n = f.n;
  const int NPW = T * CH;
  const int row0 = g * CH;
  const double2 zero = make_double2(0.0, 0.0);
  const double2 one = make_double2(1.0, 0.0);
  const double2* lsub = f.base;
  const double2* ldia = f.base + n;
  const double2* lsup = f.base + 2 * n;
  const double2* rdia = f.base + 4 * n;
  const int lc = f.lcol0 + f.cstep * s;
  const int rc = f.rcol0 + f.cstep * s;
  const bool lok = lc >= 0 && lc < f.ncol;
  const bool rok = rc >= 0 && rc < f.ncol;
  auto lhs_diag = [&](int i) {
    double2 c = lok ? f.C[(size_t)i * f.ncol + lc] : zero;
    if (f.conj_c) c = zconj(c);
    return zadd(ldia[i], zscale(c, f.lsign));
  };

  // chunk product P = M_last ... M_first, M_i = [[d_i, -a_i c_{i-1}], [1, 0]]
  Mob P = mob_identity();
  for (int m = 0; m < CH; ++m) {
    const int i = row0 + m;
    if (i < n) {
      const double2 ac = (i > 0) ? zmul(lsub[i], lsup[i - 1]) : zero;
      const Mob M = {lhs_diag(i), zscale(ac, -1.0), one, zero};
      P = mob_mul(M, P);
      if ((m & 3) == 3) P = mob_normalise(P);
    }
  }
  P = mob_normalise(P);
  // inclusive prefix product inside the wavefront (Hillis-Steele through LDS)
  int buf = 0;
  scan0[lane] = P;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int dist = 1; dist < 64; dist <<= 1) {
    const Mob own = scan0[buf * 64 + lane];
    Mob res = own;
    if (lane >= dist) res = mob_normalise(mob_mul(own, scan0[buf * 64 + lane - dist]));
    __builtin_amdgcn_wave_barrier();
    scan0[(buf ^ 1) * 64 + lane] = res;
    __builtin_amdgcn_wave_barrier();
    buf ^= 1;
  }
  if (lane == 63) tot[wave] = scan0[buf * 64 + 63];
  __syncthreads();
  // product of the totals of the wavefronts before this one
  Mob G = mob_identity();
  for (int v = 0; v < wave; ++v) G = mob_normalise(mob_mul(tot[v], G));
  // state before this chunk: (theta_{start-1}, theta_{start-2}) = Q (1, 0)^T
  double2 piv_prev = one;
  if (g > 0) {
    const Mob Q = (lane == 0) ? G : mob_normalise(mob_mul(scan0[buf * 64 + lane - 1], G));
    piv_prev = zmul(Q.a, zinv(Q.c));
  }

  float2* o = f.coef + (size_t)s * A_COUNT * NPW;
  double2 q[CH];
  double2 gf = one;
  bool bad = false;
  int reach = 0;  // coupling reach of this lane's chunk (tp_scan.h)
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    const int i = row0 + m;
    const int pi = m * T + g;
    if (i < n) {
      const double2 dm = lhs_diag(i);
      const double2 l = (i == 0) ? zero : zmul(lsub[i], zinv(piv_prev));
      const double2 piv = (i == 0) ? dm : zsub(dm, zmul(l, lsup[i - 1]));
      if (piv.x == 0.0 && piv.y == 0.0) bad = true;
      const double2 pinv = zinv(piv);
      q[m] = (i == n - 1) ? zero : zscale(zmul(lsup[i], pinv), -1.0);
      gf = zmul(gf, zscale(l, -1.0));
      if (zabs1(gf) > kCarryTau) reach = m + 1;
      o[A_L * NPW + pi] = to_f2(l);
      o[A_GF * NPW + pi] = to_f2(gf);
      o[A_PINV * NPW + pi] = to_f2(pinv);
      o[A_Q * NPW + pi] = to_f2(q[m]);
      if (f.with_sigma) {
        double2 cr = rok ? f.C[(size_t)i * f.ncol + rc] : zero;
        if (f.conj_c) cr = zconj(cr);
        o[A_SG * NPW + pi] = to_f2(zadd(dm, zadd(rdia[i], zscale(cr, f.rsign))));
      }
      piv_prev = piv;
    } else {
      q[m] = zero;
      o[A_L * NPW + pi] = make_float2(0.f, 0.f);
      o[A_GF * NPW + pi] = make_float2(0.f, 0.f);
      o[A_PINV * NPW + pi] = make_float2(0.f, 0.f);
      o[A_Q * NPW + pi] = make_float2(0.f, 0.f);
      if (f.with_sigma) o[A_SG * NPW + pi] = make_float2(0.f, 0.f);
    }
  }
  double2 gb = one;
#pragma unroll
  for (int m = CH - 1; m >= 0; --m) {
    gb = zmul(gb, q[m]);
    if (zabs1(gb) > kCarryTau) reach = max(reach, CH - m);
    o[A_GB * NPW + m * T + g] = to_f2(gb);
  }
  if (f.status) {
    if (bad) f.status[0] = 1;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) reach = max(reach, __shfl_xor(reach, off));
    if (lane == 0) atomicMax(f.status + 1, reach);
  }
}

// ------------------------------------------------------------------ sweep
// Coefficient loads through a buffer resource: the lane's byte offset is one
// VGPR shared by every load and the array/row offset (A * n_pad + m * T) * 8
// is a uniform SGPR value, so the 5-6 x CH loads of a step need no per-load
// 64-bit address registers.
struct CoefRows {
  __amdgpu_buffer_rsrc_t rsrc;
  int voff;  // lane byte offset

  __device__ __forceinline__ float2 operator()(int uniform_elem) const {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rsrc, voff, uniform_elem * 8, 0));
  }
};

struct WideArgs {
  tp_tri_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans;
};

// This wavefront's rows [wrow, wrow + 64 CH) of one plane, from the LDS tile
// (row wrow + j at tile_index(j)) as 16-byte stores of row pairs.
template <int CH>
__device__ __forceinline__ void wide_store(float2* o, const float2* tile, int wrow, int n, int lane) {
  const bool aligned = (reinterpret_cast<uintptr_t>(o) & 15) == 0;
#pragma unroll
  for (int k = 0; k < CH / 2; ++k) {
    const int j = 2 * (lane + 64 * k);
    const int r = wrow + j;
    if (aligned && r + 1 < n) {  // planes stream out: non-temporal stores
      __builtin_nontemporal_store(*reinterpret_cast<const f4v*>(tile + tile_index(j)), reinterpret_cast<f4v*>(o + r));
    } else {
      if (r < n) o[r] = tile[tile_index(j)];
      if (r + 1 < n) o[r + 1] = tile[tile_index(j + 1)];
    }
  }
}

// Plane s (held in the wavefront's tile) to every output that takes it.
template <int CH>
__device__ __forceinline__ void wide_plane(const tp_tri_scan& d, int p, int s, const float2* tile, int wrow, int lane) {
  const int n = d.n;
  if (d.out)
    wide_store<CH>((float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * s) * d.out_zstride,
                   tile, wrow, n, lane);
  if (d.ckpt && (s % d.ckpt_every) == 0)
    wide_store<CH>((float2*)d.ckpt + (size_t)p * d.ckpt_pstride + (int64_t)(s / d.ckpt_every) * d.ckpt_zstride, tile,
                   wrow, n, lane);
  if (d.exitp && s == d.steps) wide_store<CH>((float2*)d.exitp + (size_t)p * d.exit_pstride, tile, wrow, n, lane);
}

// The state plane lives in the wavefront's LDS tile, not in registers: a step
// reads u_{s-1} from it to form the right-hand side, solves in registers (y),
// and reads it again for u_s = 2x - u_{s-1}, writing u_s in place.  The tile is
// what the plane stores read anyway, and keeping u out of the registers leaves
// a 1024-lane workgroup (128 VGPRs per lane) room for the 16-row chunk.
template <int CH, bool SIGMA, bool GRAD>
__global__ __launch_bounds__(1024) void triw_sweep_kernel(const WideArgs a) {
  constexpr int RW = 64 * CH;  // rows per wavefront
  constexpr int TILE = RW + RW / 8;
  extern __shared__ f4v smw[];
  const int T = blockDim.x, W = T >> 6;
  const int g = threadIdx.x, wave = g >> 6, lane = g & 63;
  const int blk = blockIdx.x;
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  const tp_tri_scan& d = a.scans[sc];
  const int p = blk - a.blk_start[sc];  // one probe per workgroup, every probe exists
  const int n = d.n;
  const int L = d.steps;
  const int NPW = T * CH;
  const int wrow = wave * RW;
  const int row0 = g * CH;
  const float2* coef = (const float2*)d.coef + (size_t)p * d.coef_pstride;
  float2* tile = reinterpret_cast<float2*>(smw) + wave * TILE;
  float2* xup = reinterpret_cast<float2*>(smw) + W * TILE;  // per wavefront (A, B) of the forward carry
  float2* xdn = xup + 2 * W;                                 // and of the backward carry
  const float2 zero = make_float2(0.f, 0.f), one = make_float2(1.f, 0.f);

  {
    const float2* u0 = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int j = lane + 64 * k;
      tile[tile_index(j)] = (wrow + j < n) ? u0[wrow + j] : zero;
    }
    __builtin_amdgcn_wave_barrier();
  }
  float2* mine = tile + tile_index(lane * CH);  // this lane's CH rows (contiguous: a chunk never straddles a pad)

  for (int s = 1; s <= L; ++s) {
    wide_plane<CH>(d, p, s - 1, tile, wrow, lane);
    CoefRows cf;  // this step's coefficient arrays; 0x00020000: 32-bit data format (gfx9 raw buffer)
    cf.rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)(coef + (size_t)(s - 1) * A_COUNT * NPW), (short)0,
                                                A_COUNT * NPW * (int)sizeof(float2), 0x00020000);
    cf.voff = g * (int)sizeof(float2);
    const bool has_rhs = d.rhs != nullptr && (d.rhs_steps == 0 || s <= d.rhs_steps);
    const float2* rb = has_rhs ? (const float2*)d.rhs + (size_t)p * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride
                               : nullptr;
    float2 y[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      float2 r = mine[m];
      if (SIGMA) r = cmul(cf(A_SG * NPW + m * T), r);
      if (has_rhs) {
        const float2 b = (row0 + m < n) ? rb[row0 + m] : zero;
        r = SIGMA ? cadd(r, b) : make_float2(fmaf(0.5f, b.x, r.x), fmaf(0.5f, b.y, r.y));
      }
      y[m] = r;
    }
#pragma unroll
    for (int m = 1; m < CH; ++m) y[m] = cfms(cf(A_L * NPW + m * T), y[m - 1], y[m]);
    {  // forward carry: y entering lane g
      float2 A = cf(A_GF * NPW + (CH - 1) * T);
      float2 B = y[CH - 1];
      scan_affine_up(A, B);
      if (lane == 63) {
        xup[2 * wave] = A;
        xup[2 * wave + 1] = B;
      }
      __syncthreads();
      float2 Y = zero;
      for (int v = 0; v < wave; ++v) Y = cfma(xup[2 * v], Y, xup[2 * v + 1]);
      const float2 yin = cfma(from_prev_lane(A, one), Y, from_prev_lane(B));
#pragma unroll
      for (int m = 0; m < CH; ++m) y[m] = cfma(cf(A_GF * NPW + m * T), yin, y[m]);
    }
#pragma unroll
    for (int m = 0; m < CH; ++m) y[m] = cmul(y[m], cf(A_PINV * NPW + m * T));
#pragma unroll
    for (int m = CH - 2; m >= 0; --m) y[m] = cfma(cf(A_Q * NPW + m * T), y[m + 1], y[m]);
    {  // backward carry: x entering lane g from above
      float2 A = cf(A_GB * NPW);
      float2 B = y[0];
      scan_affine_down(A, B, lane);
      if (lane == 0) {
        xdn[2 * wave] = A;
        xdn[2 * wave + 1] = B;
      }
      __syncthreads();
      float2 X = zero;
      for (int v = W - 1; v > wave; --v) X = cfma(xdn[2 * v], X, xdn[2 * v + 1]);
      const float2 xin = cfma(from_next_lane(A, one), X, from_next_lane(B));
#pragma unroll
      for (int m = 0; m < CH; ++m) {
        const float2 x = cfma(cf(A_GB * NPW + m * T), xin, y[m]);
        const float2 old = mine[m];
        y[m] = SIGMA ? csub(x, old) : make_float2(fmaf(2.f, x.x, -old.x), fmaf(2.f, x.y, -old.y));
        mine[m] = y[m];
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (GRAD) {  // grad[row] += conj(lambda_s) (a + b), one atomic per row and component
      const int grow = d.grad_z0 + d.grad_zdir * (s - 1);
      const bool skip_b = d.grad_skip_row0 && grow == 0;
      const float2* ga = (const float2*)d.gfwd + (size_t)p * d.gfwd_pstride +
                         (int64_t)(d.gfwd_z0 + d.gfwd_zdir * (s - 1)) * d.gfwd_zstride;
      const float2* gb = (const float2*)d.gfwd + (size_t)p * d.gfwd_pstride +
                         (int64_t)(d.gfwd_z0 + d.gfwd_zdir * s) * d.gfwd_zstride;
      float* gr = (float*)((float2*)d.grad + (int64_t)grow * d.grad_zstride);
#pragma unroll
      for (int m = 0; m < CH; ++m) {
        const int i = row0 + m;
        if (i < n) {
          const float2 w = skip_b ? ga[i] : cadd(ga[i], gb[i]);
          const float2 c = cmul(cconj(y[m]), w);
          unsafeAtomicAdd(gr + 2 * i, c.x);
          unsafeAtomicAdd(gr + 2 * i + 1, c.y);
        }
      }
    }
  }
  wide_plane<CH>(d, p, L, tile, wrow, lane);
}

// ------------------------------------------------------------------ host side
static int check_wide(int chunk, int waves) {
  if (chunk != 8 && chunk != 16) {
    set_error("wide tridiagonal kernels take chunk 8 or 16 (got %d)", chunk);
    return TP_ERR_INVALID;
  }
  if (waves < 1 || waves > WIDE_MAX_WAVES) {
    set_error("wide tridiagonal kernels take 1 to %d wavefronts per probe (got %d)", WIDE_MAX_WAVES, waves);
    return TP_ERR_INVALID;
  }
  return TP_OK;
}

template <int CH>
static size_t wide_sweep_lds(int waves) {
  return ((size_t)waves * (64 * CH + 8 * CH) + 4 * (size_t)waves) * sizeof(float2);
}

template <int CH, bool SIGMA, bool GRAD>
static int launch_wide(const WideArgs& a, int nblk, int waves, hipStream_t st) {
  const size_t lds = wide_sweep_lds<CH>(waves);
  if (lds > 65536) {
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)triw_sweep_kernel<CH, SIGMA, GRAD>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL((triw_sweep_kernel<CH, SIGMA, GRAD>), dim3(nblk), dim3(64 * waves), lds, st, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

template <int CH>
static int launch_wide_modes(const WideArgs& a, int nblk, int waves, bool sigma, bool grad, hipStream_t st) {
  if (sigma) return grad ? launch_wide<CH, true, true>(a, nblk, waves, st) : launch_wide<CH, true, false>(a, nblk, waves, st);
  return grad ? launch_wide<CH, false, true>(a, nblk, waves, st) : launch_wide<CH, false, false>(a, nblk, waves, st);
}

}  // namespace tp

using namespace tp;

extern "C" size_t tp_tri_coef_bytes_wide(int n, int steps, int chunk, int waves) {
  (void)n;
  return (size_t)steps * A_COUNT * 64 * (size_t)waves * chunk * sizeof(float2);
}

extern "C" int tp_tri_factor_wide(const void* d_C, int n, int ncol, const void* d_base, double lhs_sign,
                                  double rhs_sign, int conj_c, int lhs_col0, int rhs_col0, int col_step, int steps,
                                  int chunk, int waves, int with_sigma, void* d_coef, int* d_status, void* stream) {
  int rc = check_wide(chunk, waves);
  if (rc) return rc;
  if (n < 1 || n > 64 * waves * chunk || steps < 1 || ncol < 1 || !d_C || !d_base || !d_coef) {
    set_error("tp_tri_factor_wide: invalid sizes or null pointer (n=%d, chunk=%d, waves=%d, steps=%d, ncol=%d)", n,
              chunk, waves, steps, ncol);
    return TP_ERR_INVALID;
  }
  FactorArgs f;
  f.C = (const double2*)d_C;
  f.base = (const double2*)d_base;
  f.n = n;
  f.ncol = ncol;
  f.steps = steps;
  f.lsign = lhs_sign;
  f.rsign = rhs_sign;
  f.conj_c = conj_c;
  f.lcol0 = lhs_col0;
  f.rcol0 = rhs_col0;
  f.cstep = col_step;
  f.with_sigma = with_sigma;
  f.coef = (float2*)d_coef;
  f.status = d_status;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = (size_t)129 * waves * sizeof(Mob);
  if (chunk == 8) {
    if (lds > 65536)
      TP_CHECK_HIP(hipFuncSetAttribute((const void*)triw_factor_kernel<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds));
    hipLaunchKernelGGL((triw_factor_kernel<8>), dim3(steps), dim3(64 * waves), lds, st, f);
  } else {
    if (lds > 65536)
      TP_CHECK_HIP(hipFuncSetAttribute((const void*)triw_factor_kernel<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds));
    hipLaunchKernelGGL((triw_factor_kernel<16>), dim3(steps), dim3(64 * waves), lds, st, f);
  }
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tri_sweep_wide(const tp_tri_scan* scans, int nscans, int chunk, int waves, void* stream) {
  int rc = check_wide(chunk, waves);
  if (rc) return rc;
  if (!scans || nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  WideArgs a;
  memset(&a, 0, sizeof(a));
  int nblk = 0;
  const bool sigma = scans[0].sigma != 0;
  const bool grad = scans[0].grad != nullptr;
  for (int i = 0; i < nscans; ++i) {
    const tp_tri_scan& d = scans[i];
    if (d.chunk != 0 && d.chunk != chunk) {
      set_error("scan %d: chunk %d differs from the launch's chunk %d (wide sweeps take one chunk)", i, d.chunk, chunk);
      return TP_ERR_INVALID;
    }
    if (d.n < 1 || d.n > 64 * waves * chunk) {
      set_error("scan %d: n = %d outside [1, 64*waves*chunk = %d]", i, d.n, 64 * waves * chunk);
      return TP_ERR_INVALID;
    }
    if (d.steps < 1 || d.probes < 0 || !d.coef || !d.u0) {
      set_error("scan %d: steps >= 1, probes >= 0, coef and u0 are required", i);
      return TP_ERR_INVALID;
    }
    if ((d.sigma != 0) != sigma) {
      set_error("scan %d: all scans of one launch must share the sigma flag", i);
      return TP_ERR_INVALID;
    }
    if (d.ckpt && d.ckpt_every < 1) {
      set_error("scan %d: ckpt_every must be >= 1 when ckpt is set", i);
      return TP_ERR_INVALID;
    }
    if (d.coef_pstride < 0 || (reinterpret_cast<uintptr_t>(d.coef) & 7)) {
      set_error("scan %d: coef must be 8-byte aligned and coef_pstride >= 0", i);
      return TP_ERR_INVALID;
    }
    if ((d.grad != nullptr) != grad || (grad && d.gfwd == nullptr)) {
      set_error("scan %d: gradient accumulation needs grad and gfwd, on every scan of the launch", i);
      return TP_ERR_INVALID;
    }
    a.scans[i] = d;
    a.blk_start[i] = nblk;
    nblk += d.probes;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  if (chunk == 8) return launch_wide_modes<8>(a, nblk, waves, sigma, grad, st);
  return launch_wide_modes<16>(a, nblk, waves, sigma, grad, st);
}
