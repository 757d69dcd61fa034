// Crank–Nicolson slice stepping with per-slice tridiagonal LU (one transverse axis).
//
// Every propagation mode has the form   A_s u_s = B_s u_{s-1} + b_s   with a
// tridiagonal A_s shared by all probes of a scan and B_s = diag(sigma_s) - A_s
// (the CN pair A0 = I - K, B0 = I + K sums to 2I, so sigma = 2 except for the
// exact adjoint, where the RHS uses the next slice's object term).  Hence
//     u_s = A_s^{-1}(sigma u_{s-1} + b_s) - u_{s-1}
// and no mat-vec is needed: one tridiagonal solve per probe and step.
//
// Factor (one wavefront per step, FP64): the pivot recurrence
// piv_i = d_i - a_i c_{i-1} / piv_{i-1} is a product of 2x2 Moebius matrices
// [[d_i, -a_i c_{i-1}], [1, 0]]; each lane forms the product over its chunk of
// rows, an exclusive prefix product across the 64 lanes (normalised, FP64)
// gives every chunk its incoming pivot, then each lane runs the recurrence over
// its rows.  Per row and step it stores complex64, in the lane-interleaved
// layout row i -> (i % CH) * 64 + i / CH:
//   l    forward multiplier  a_i / piv_{i-1}          (y_i = r_i - l_i y_{i-1})
//   gf   prod_{m = chunk start .. i} (-l_m)           (carry into the chunk)
//   pinv 1 / piv_i
//   q    -c_i / piv_i                                 (x_i = y_i pinv_i + q_i x_{i+1})
//   gb   prod_{m = i .. chunk end} q_m                (carry from the next chunk)
//   sg   sigma_i (only when the mode needs it)
//
// Sweep: one probe per wavefront, lane l owns rows [l*CH, (l+1)*CH).  Both
// triangular recurrences run chunk-locally in registers, the chunk carries are
// combined by wavefront scans of affine maps (DPP row shifts), and a register
// pass applies them.  The persistent kernel walks every step in one launch
// with the state in registers; the plain kernel advances one step per launch.
#include "tp_scan.h"

namespace tp {

// ------------------------------------------------------------------ factor
constexpr int FACTOR_WAVES = 4;

template <int CH>
__device__ __forceinline__ void tri_factor_body(const FactorArgs& f, int blk, Mob (*sh)[2][64]) {
  constexpr int NP = 64 * CH;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int s = blk * FACTOR_WAVES + wave;
  if (s >= f.steps) return;  // whole wavefront; no workgroup barrier below
  const int n = f.n;
  const int row0 = lane * CH;
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

  double2 d[CH], ac[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    const int i = row0 + m;
    d[m] = one;
    ac[m] = zero;
    if (i < n) {
      double2 c = lok ? f.C[(size_t)i * f.ncol + lc] : zero;
      if (f.conj_c) c = zconj(c);
      d[m] = zadd(ldia[i], zscale(c, f.lsign));
      if (i > 0) ac[m] = zmul(lsub[i], lsup[i - 1]);
    }
  }
  // chunk product P = M_last ... M_first, M_i = [[d_i, -a_i c_{i-1}], [1, 0]]
  Mob P = {one, zero, zero, one};
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    if (row0 + m < n) {
      Mob M = {d[m], zscale(ac[m], -1.0), one, zero};
      P = mob_mul(M, P);
      if ((m & 3) == 3) P = mob_normalise(P);
    }
  }
  P = mob_normalise(P);
  // inclusive prefix product across lanes (Hillis-Steele through LDS):
  // Q_l = P_l * P_{l-1} * ... * P_0
  int buf = 0;
  sh[wave][0][lane] = P;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int dist = 1; dist < 64; dist <<= 1) {
    const Mob own = sh[wave][buf][lane];
    Mob res = own;
    if (lane >= dist) res = mob_normalise(mob_mul(own, sh[wave][buf][lane - dist]));
    __builtin_amdgcn_wave_barrier();
    sh[wave][buf ^ 1][lane] = res;
    __builtin_amdgcn_wave_barrier();
    buf ^= 1;
  }
  // state before this chunk: (theta_{start-1}, theta_{start-2}) = Q_{l-1} (1, 0)^T
  double2 piv_prev = one;
  if (lane > 0) {
    const Mob Q = sh[wave][buf][lane - 1];
    piv_prev = zmul(Q.a, zinv(Q.c));
  }

  float2* o = f.coef + (size_t)s * A_COUNT * NP;
  double2 q[CH];
  double2 gf = one;
  bool bad = false;
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    const int i = row0 + m;
    const int pi = m * 64 + lane;
    if (i < n) {
      const double2 l = (i == 0) ? zero : zmul(lsub[i], zinv(piv_prev));
      const double2 piv = (i == 0) ? d[m] : zsub(d[m], zmul(l, lsup[i - 1]));
      if (piv.x == 0.0 && piv.y == 0.0) bad = true;
      const double2 pinv = zinv(piv);
      q[m] = (i == n - 1) ? zero : zscale(zmul(lsup[i], pinv), -1.0);
      gf = zmul(gf, zscale(l, -1.0));
      o[A_L * NP + pi] = to_f2(l);
      o[A_GF * NP + pi] = to_f2(gf);
      o[A_PINV * NP + pi] = to_f2(pinv);
      o[A_Q * NP + pi] = to_f2(q[m]);
      if (f.with_sigma) {
        double2 cr = rok ? f.C[(size_t)i * f.ncol + rc] : zero;
        if (f.conj_c) cr = zconj(cr);
        o[A_SG * NP + pi] = to_f2(zadd(d[m], zadd(rdia[i], zscale(cr, f.rsign))));
      }
      piv_prev = piv;
    } else {
      q[m] = zero;
      o[A_L * NP + pi] = make_float2(0.f, 0.f);
      o[A_GF * NP + pi] = make_float2(0.f, 0.f);
      o[A_PINV * NP + pi] = make_float2(0.f, 0.f);
      o[A_Q * NP + pi] = make_float2(0.f, 0.f);
      if (f.with_sigma) o[A_SG * NP + pi] = make_float2(0.f, 0.f);
    }
  }
  double2 gb = one;
#pragma unroll
  for (int m = CH - 1; m >= 0; --m) {
    gb = zmul(gb, q[m]);
    o[A_GB * NP + m * 64 + lane] = to_f2(gb);
  }
  if (bad && f.status) f.status[0] = 1;
}

template <int CH>
__global__ __launch_bounds__(64 * FACTOR_WAVES) void tri_factor_kernel(const FactorArgs f) {
  __shared__ Mob sh[FACTOR_WAVES][2][64];
  tri_factor_body<CH>(f, blockIdx.x, sh);
}

// Several independent factorisations (scans, directions) in one launch: each
// alone leaves most CUs waiting on one long dependency chain per wavefront.
struct FactorBatch {
  FactorArgs jobs[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int njobs;
};

template <int CH>
__global__ __launch_bounds__(64 * FACTOR_WAVES) void tri_factor_batch_kernel(const FactorBatch b) {
  __shared__ Mob sh[FACTOR_WAVES][2][64];
  int j = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < b.njobs && (int)blockIdx.x >= b.blk_start[i]) j = i;
  tri_factor_body<CH>(b.jobs[j], blockIdx.x - b.blk_start[j], sh);
}

// One CN step for the CH rows of this lane:  u <- A^{-1}(sigma u + b) - u.
// `cf` points at the step's coefficient arrays (LDS or global, lane-interleaved).
// Without SIGMA (sigma = 2):  r = u + b/2 and u <- 2 x - u.
// With SIGMA:                 r = sigma u + b and u <- x - u.
struct NoHook {
  __device__ __forceinline__ void operator()(int) const {}
};

// Direct form: every coefficient is read where it is used (the register
// budget of CH = 32 holds no second bank).
template <int CH, bool HAS_RHS, bool SIGMA, typename Ptr, typename Hook>
__device__ __forceinline__ void tri_step_lane_direct(float2 (&u)[CH], Ptr cf, const float2 (&rhs)[CH], int lane,
                                                     const Hook& hook) {
  constexpr int NP = 64 * CH;
  float2 y[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    float2 r = u[m];
    if (SIGMA) r = cmul(cf[A_SG * NP + m * 64 + lane], r);
    if (HAS_RHS) r = SIGMA ? cadd(r, rhs[m]) : make_float2(fmaf(0.5f, rhs[m].x, r.x), fmaf(0.5f, rhs[m].y, r.y));
    y[m] = r;
  }
  hook(0);
#pragma unroll
  for (int m = 1; m < CH; ++m) y[m] = cfms(cf[A_L * NP + m * 64 + lane], y[m - 1], y[m]);
  hook(1);
  {
    float2 A = cf[A_GF * NP + (CH - 1) * 64 + lane];
    float2 B = y[CH - 1];
    scan_affine_up(A, B);
    const float2 yin = from_prev_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) y[m] = cfma(cf[A_GF * NP + m * 64 + lane], yin, y[m]);
  }
#pragma unroll
  for (int m = 0; m < CH; ++m) y[m] = cmul(y[m], cf[A_PINV * NP + m * 64 + lane]);
#pragma unroll
  for (int m = CH - 2; m >= 0; --m) y[m] = cfma(cf[A_Q * NP + m * 64 + lane], y[m + 1], y[m]);
  hook(2);
  {
    float2 A = cf[A_GB * NP + lane];
    float2 B = y[0];
    scan_affine_down(A, B, lane);
    const float2 xin = from_next_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const float2 x = cfma(cf[A_GB * NP + m * 64 + lane], xin, y[m]);
      u[m] = SIGMA ? csub(x, u[m]) : make_float2(fmaf(2.f, x.x, -u[m].x), fmaf(2.f, x.y, -u[m].y));
    }
  }
  hook(3);
}

// Staged form (CH <= 16): the coefficients of each phase are read into
// registers one phase ahead of their use (two banks of CH values), and
// scheduling barriers keep the compiler from sinking the reads next to their
// uses, so a phase waits on reads issued a phase earlier instead of on one read
// every few instructions.  This pays where a SIMD holds one wavefront and
// nothing else covers the waits; with two wavefronts per SIMD the direct form's
// freer schedule is faster.
template <int CH, bool HAS_RHS, bool SIGMA, bool STAGED = true, typename Ptr, typename Hook = NoHook>
__device__ __forceinline__ void tri_step_lane(float2 (&u)[CH], Ptr cf, const float2 (&rhs)[CH], int lane,
                                              const Hook& hook = Hook()) {
  if constexpr (CH > 16 || !STAGED) {
    tri_step_lane_direct<CH, HAS_RHS, SIGMA>(u, cf, rhs, lane, hook);
    return;
  }
  constexpr int NP = 64 * CH;
  float2 y[CH], c1[CH], c2[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    c1[m] = cf[A_L * NP + m * 64 + lane];
    c2[m] = cf[A_GF * NP + m * 64 + lane];
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    float2 r = u[m];
    if (SIGMA) r = cmul(cf[A_SG * NP + m * 64 + lane], r);
    if (HAS_RHS) r = SIGMA ? cadd(r, rhs[m]) : make_float2(fmaf(0.5f, rhs[m].x, r.x), fmaf(0.5f, rhs[m].y, r.y));
    y[m] = r;
  }
  hook(0);
#pragma unroll
  for (int m = 1; m < CH; ++m) y[m] = cfms(c1[m], y[m - 1], y[m]);
  hook(1);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < CH; ++m) c1[m] = cf[A_PINV * NP + m * 64 + lane];
  __builtin_amdgcn_sched_barrier(0);
  {
    float2 A = c2[CH - 1];
    float2 B = y[CH - 1];
    scan_affine_up(A, B);
    const float2 yin = from_prev_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) y[m] = cfma(c2[m], yin, y[m]);
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < CH; ++m) c2[m] = cf[A_Q * NP + m * 64 + lane];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < CH; ++m) y[m] = cmul(y[m], c1[m]);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < CH; ++m) c1[m] = cf[A_GB * NP + m * 64 + lane];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = CH - 2; m >= 0; --m) y[m] = cfma(c2[m], y[m + 1], y[m]);
  hook(2);
  {
    float2 A = c1[0];
    float2 B = y[0];
    scan_affine_down(A, B, lane);
    const float2 xin = from_next_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const float2 x = cfma(c1[m], xin, y[m]);
      u[m] = SIGMA ? csub(x, u[m]) : make_float2(fmaf(2.f, x.x, -u[m].x), fmaf(2.f, x.y, -u[m].y));
    }
  }
  hook(3);
}

struct SweepArgs {
  tp_tri_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans;
  int ppb;
  int nblk;
};

__device__ __forceinline__ int find_scan(const SweepArgs& a, int blk) {
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  return sc;
}

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of logical blocks so that the probes of one scan (which share
// every step's coefficients) meet in as few L2 caches as possible.
__device__ __forceinline__ int xcd_block(int hw, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = hw & 7, j = hw >> 3;
  return x * q + min(x, r) + j;
}

template <int CH>
__device__ __forceinline__ void store_rows(float2* o, const float2 (&u)[CH], int row0, int n) {
  if (row0 + CH <= n) {
#pragma unroll
    for (int m = 0; m < CH; ++m) o[row0 + m] = u[m];
  } else {
#pragma unroll
    for (int m = 0; m < CH; ++m)
      if (row0 + m < n) o[row0 + m] = u[m];
  }
}

template <int CH>
__device__ __forceinline__ void store_plane(const tp_tri_scan& d, int p, int s, const float2 (&u)[CH], int row0) {
  const int n = d.n;
  if (d.out)
    store_rows<CH>((float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * s) * d.out_zstride,
                   u, row0, n);
  if (d.ckpt && (s % d.ckpt_every) == 0)
    store_rows<CH>((float2*)d.ckpt + (size_t)p * d.ckpt_pstride + (int64_t)(s / d.ckpt_every) * d.ckpt_zstride, u,
                   row0, n);
  if (d.exitp && s == d.steps) store_rows<CH>((float2*)d.exitp + (size_t)p * d.exit_pstride, u, row0, n);
}

// Coalesced plane stores: the wavefront's plane goes through an LDS tile
// (row i at i + 2*(i/16): two pad words per 16 rows against bank conflicts,
// row pairs stay 16-byte aligned) and leaves as 16-byte stores, each
// instruction writing 64 consecutive row pairs (1 KiB).
template <int CH>
__device__ __forceinline__ void tile_put(float2* tile, const float2 (&u)[CH], int row0) {
#pragma unroll
  for (int m = 0; m < CH; ++m) tile[tile_index(row0 + m)] = u[m];
  __builtin_amdgcn_wave_barrier();
}

template <int CH>
struct PlaneWriter {
  float2* dst[3];
  const float2* tile;
  int n, lane;

  __device__ __forceinline__ void init(const tp_tri_scan& d, int p, int s, bool act, const float2* t, int ln) {
    tile = t;
    n = d.n;
    lane = ln;
    dst[0] = (act && d.out)
                 ? (float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * s) * d.out_zstride
                 : nullptr;
    dst[1] = (act && d.ckpt && (s % d.ckpt_every) == 0)
                 ? (float2*)d.ckpt + (size_t)p * d.ckpt_pstride + (int64_t)(s / d.ckpt_every) * d.ckpt_zstride
                 : nullptr;
    dst[2] = (act && d.exitp && s == d.steps) ? (float2*)d.exitp + (size_t)p * d.exit_pstride : nullptr;
  }

  // group g of 4: row pairs j = 2 (lane + 64 k) for k in [g*CH/8, (g+1)*CH/8),
  // each read from the tile just before its store
  __device__ __forceinline__ void operator()(int g) const {
    constexpr int KP = CH / 2 > 0 ? CH / 2 : 1;  // pair-instructions per plane
    const int k0 = (g * KP) / 4, k1 = ((g + 1) * KP) / 4;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      float2* o = dst[t];
      if (o == nullptr) continue;
      const bool aligned = (reinterpret_cast<uintptr_t>(o) & 15) == 0;
      for (int k = k0; k < k1; ++k) {
        const int j = 2 * (lane + 64 * k);
        if (aligned && j + 1 < n) {
          typedef float f4 __attribute__((ext_vector_type(4)));
          *reinterpret_cast<f4*>(o + j) = *reinterpret_cast<const f4*>(tile + tile_index(j));
        } else {
          if (j < n) o[j] = tile[tile_index(j)];
          if (j + 1 < n) o[j + 1] = tile[tile_index(j + 1)];
        }
      }
    }
  }
  __device__ __forceinline__ void all() const {
    (*this)(0);
    (*this)(1);
    (*this)(2);
    (*this)(3);
  }

  // Batched form: every LDS read of the plane is issued before the first store,
  // so the stores wait on the LDS once rather than once each (faster where one
  // wavefront per SIMD leaves the waits exposed, slower in full launches).
  __device__ __forceinline__ void all_batched() const {
    typedef float f4 __attribute__((ext_vector_type(4)));
    constexpr int NP = 64 * CH;
    constexpr int KP = CH / 2 > 0 ? CH / 2 : 1;  // pair-instructions per plane
    if (dst[0] == nullptr && dst[1] == nullptr && dst[2] == nullptr) return;
    f4 v[KP];
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int j = 2 * (lane + 64 * k);
      v[k] = (CH >= 2 || j < NP) ? *reinterpret_cast<const f4*>(tile + tile_index(j)) : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      float2* o = dst[t];
      if (o == nullptr) continue;
      const bool aligned = (reinterpret_cast<uintptr_t>(o) & 15) == 0;
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const int j = 2 * (lane + 64 * k);
        if (aligned && j + 1 < n) {
          *reinterpret_cast<f4*>(o + j) = v[k];
        } else {
          if (j < n) o[j] = make_float2(v[k].x, v[k].y);
          if (j + 1 < n) o[j + 1] = make_float2(v[k].z, v[k].w);
        }
      }
    }
  }
};

// ------------------------------------------------------------------ persistent sweep

// One workgroup = PPB wavefronts = PPB probes of one scan walking all steps in
// lockstep.  The step's coefficients are staged in LDS and shared by the PPB
// wavefronts; the next step's block is fetched into registers while the
// current step computes, and the plane computed in step s is written at the
// start of step s+1 (before that fetch) so the stores drain behind the math.
template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED>
__device__ __forceinline__ void tri_sweep_body(const SweepArgs& a, const int blk, const int sc) {
  constexpr int NP = 64 * CH;
  constexpr int NT = 64 * PPB;
  constexpr int NA = SIGMA ? 6 : 5;
  constexpr int NV = NA * NP / 2;             // 16-byte vectors staged per step
  constexpr int STRIDE_V = A_COUNT * NP / 2;  // 16-byte vectors per step in memory
  constexpr int PER = (NV + NT - 1) / NT;
  extern __shared__ f4v smem_v[];
  const float2* cf_l = reinterpret_cast<const float2*>(smem_v);
  constexpr int TILE = NP + NP / 8;  // per-wavefront store tile (float2)

  const tp_tri_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * PPB + wave;
  const bool act = p < d.probes;
  const int n = d.n;
  const int L = d.steps;
  const int row0 = lane * CH;
  // per-probe factors (one probe per workgroup) or the scan's shared factors
  const f4v* coef4 = reinterpret_cast<const f4v*>(d.coef) + (PPB == 1 ? (size_t)(act ? p : 0) * (d.coef_pstride / 2) : 0);
  float2* tile = reinterpret_cast<float2*>(smem_v) + NA * NP + wave * TILE;
  // gradient row accumulated over the workgroup's probes (tile_index layout)
  float2* gbuf = reinterpret_cast<float2*>(smem_v) + NA * NP + PPB * TILE;

  f4v st[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = threadIdx.x + k * NT;
    if (NV % NT == 0 || idx < NV) st[k] = coef4[idx];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int idx = threadIdx.x + k * NT;
    if (NV % NT == 0 || idx < NV) smem_v[idx] = st[k];
  }

  // The initial plane arrives through the LDS tile: coalesced, and the state
  // registers never hold a pending global load when the step loop starts
  // (otherwise the first use of u in every iteration waits on vmcnt(0)).
  float2 u[CH];
  {
    const float2* u0 = (const float2*)d.u0 + (size_t)(act ? p : 0) * d.u0_pstride;
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int j = lane + 64 * k;
      tile[tile_index(j)] = (act && j < n) ? u0[j] : make_float2(0.f, 0.f);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int m = 0; m < CH; ++m) u[m] = tile[tile_index(row0 + m)];
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  float2 ga[CH], gb[CH];
  if (GRAD) {
    for (int idx = threadIdx.x; idx < TILE; idx += NT) gbuf[idx] = make_float2(0.f, 0.f);
    const float2* g = (const float2*)d.gfwd + (size_t)(act ? p : 0) * d.gfwd_pstride + (int64_t)d.gfwd_z0 * d.gfwd_zstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) ga[m] = (act && row0 + m < n) ? g[row0 + m] : make_float2(0.f, 0.f);
    __syncthreads();
  }

  float2 rb[CH];
  for (int s = 1; s <= L; ++s) {
    // Plane s-1 is written first, then the next step's coefficients are
    // fetched: vmcnt retires in order, so the wait for that fetch at the end
    // of the step only covers stores that have had a whole step to drain.
    {
      tile_put<CH>(tile, u, row0);
      PlaneWriter<CH> w;
      w.init(d, p, s - 1, act, tile, lane);
      if constexpr (STAGED) w.all_batched(); else w.all();
    }
    if (s < L) {
      const f4v* src = coef4 + (size_t)s * STRIDE_V;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        if (NV % NT == 0 || idx < NV) st[k] = src[idx];
      }
    }
    if (GRAD) {  // plane b = u_{s-1} of this step (prefetched before the math)
      const float2* g =
          (const float2*)d.gfwd + (size_t)(act ? p : 0) * d.gfwd_pstride + (int64_t)(d.gfwd_z0 + d.gfwd_zdir * s) * d.gfwd_zstride;
#pragma unroll
      for (int m = 0; m < CH; ++m) gb[m] = (act && row0 + m < n) ? g[row0 + m] : make_float2(0.f, 0.f);
    }
    if (d.rhs != nullptr && (d.rhs_steps == 0 || s <= d.rhs_steps)) {
      const float2* r = (const float2*)d.rhs + (size_t)(act ? p : 0) * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
      for (int m = 0; m < CH; ++m) rb[m] = (act && row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
      tri_step_lane<CH, true, SIGMA, STAGED>(u, cf_l, rb, lane);
    } else {
      tri_step_lane<CH, false, SIGMA, STAGED>(u, cf_l, rb, lane);
    }
    int grow = 0;
    if (GRAD) {
      grow = d.grad_z0 + d.grad_zdir * (s - 1);
      const bool skip_b = d.grad_skip_row0 && grow == 0;
      if (act) {
#pragma unroll
        for (int m = 0; m < CH; ++m) {
          const float2 w = skip_b ? ga[m] : cadd(ga[m], gb[m]);
          const float2 c = cmul(cconj(u[m]), w);
          float2* slot = gbuf + tile_index(row0 + m);
          atomicAdd(&slot->x, c.x);
          atomicAdd(&slot->y, c.y);
        }
      }
#pragma unroll
      for (int m = 0; m < CH; ++m) ga[m] = gb[m];
    }
    __syncthreads();
    if (GRAD) {  // one global atomic per row and component per workgroup
      float* gr = (float*)((float2*)d.grad + (int64_t)grow * d.grad_zstride);
      for (int i = threadIdx.x; i < n; i += NT) {
        const float2 v = gbuf[tile_index(i)];
        unsafeAtomicAdd(gr + 2 * i, v.x);
        unsafeAtomicAdd(gr + 2 * i + 1, v.y);
        gbuf[tile_index(i)] = make_float2(0.f, 0.f);
      }
    }
    if (s < L) {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        if (NV % NT == 0 || idx < NV) smem_v[idx] = st[k];
      }
    }
    __syncthreads();
  }
  tile_put<CH>(tile, u, row0);
  {
    PlaneWriter<CH> w;
    w.init(d, p, L, act, tile, lane);
    if constexpr (STAGED) w.all_batched(); else w.all();
  }
}

// A workgroup's probes all belong to one scan, so the choice between the
// launch's chunk and half of it (tp_tri_scan.chunk) is uniform per workgroup;
// the LDS is sized for the launch's chunk and the narrow body uses a prefix.
template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED>
__global__ __launch_bounds__(64 * PPB) void tri_sweep_kernel(const SweepArgs a) {
  // Blocks in launch order: the dispatcher spreads consecutive workgroups over
  // the XCDs and CUs, so the first workgroups of the launch (the first scans)
  // get a CU each before any CU takes a second one.  Grouping each XCD's blocks
  // into one contiguous range instead (coefficient reuse in one L2) measured
  // 3-4 % slower on the config-1 launch (4.07-4.21 ms against 3.95-4.08 ms):
  // it stacks two wide-scan workgroups on the same CUs.
  const int blk = (int)blockIdx.x;
  const int sc = find_scan(a, blk);
  if constexpr (!GRAD && CH >= 16) {
    if (a.scans[sc].chunk == CH / 2) {
      tri_sweep_body<CH / 2, PPB, SIGMA, GRAD, STAGED>(a, blk, sc);
      return;
    }
  }
  tri_sweep_body<CH, PPB, SIGMA, GRAD, STAGED>(a, blk, sc);
}

// ------------------------------------------------------------------ plain one-step
// One launch per step; plane s-1 is read from memory and plane s written, both
// through the per-wavefront LDS tile so that global accesses are coalesced.
template <int CH, bool SIGMA>
__global__ __launch_bounds__(256) void tri_step_kernel(const SweepArgs a, int s) {
  constexpr int NP = 64 * CH;
  constexpr int TILE = NP + NP / 8;
  extern __shared__ float2 tiles[];
  const int blk = xcd_block(blockIdx.x, a.nblk);
  const int sc = find_scan(a, blk);
  const tp_tri_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * a.ppb + wave;
  if (wave >= a.ppb || p >= d.probes) return;  // whole wavefront; no workgroup barrier below
  const int n = d.n;
  const int row0 = lane * CH;
  float2* tile = tiles + wave * TILE;
  const float2* src;
  if (s == 1)
    src = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
  else
    src = (const float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * (s - 1)) * d.out_zstride;
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    const int j = lane + 64 * k;
    tile[tile_index(j)] = (j < n) ? src[j] : make_float2(0.f, 0.f);
  }
  __builtin_amdgcn_wave_barrier();
  float2 u[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m) u[m] = tile[tile_index(row0 + m)];
  __builtin_amdgcn_wave_barrier();
  if (s == 1) {
    PlaneWriter<CH> w0;
    w0.init(d, p, 0, true, tile, lane);
    w0.all_batched();
    __builtin_amdgcn_wave_barrier();
  }
  float2 rb[CH];
  const float2* cf = (const float2*)d.coef + (size_t)(s - 1) * A_COUNT * NP;
  if (d.rhs != nullptr && (d.rhs_steps == 0 || s <= d.rhs_steps)) {
    const float2* r = (const float2*)d.rhs + (size_t)p * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) rb[m] = (row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
    tri_step_lane<CH, true, SIGMA>(u, cf, rb, lane);
  } else {
    tri_step_lane<CH, false, SIGMA>(u, cf, rb, lane);
  }
  tile_put<CH>(tile, u, row0);
  PlaneWriter<CH> w;
  w.init(d, p, s, true, tile, lane);
  w.all_batched();
}

template <int CH, bool SIGMA>
static int launch_step(const SweepArgs& a, int s, hipStream_t st) {
  constexpr int NP = 64 * CH;
  const size_t lds = (size_t)4 * (NP + NP / 8) * sizeof(float2);
  if (lds > 65536) {
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)tri_step_kernel<CH, SIGMA>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL((tri_step_kernel<CH, SIGMA>), dim3(a.nblk), dim3(256), lds, st, a, s);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

// ------------------------------------------------------------------ host side
static int check_chunk(int chunk) {
  switch (chunk) {
    case 1: case 2: case 4: case 8: case 16: case 32: return TP_OK;
    default:
      set_error("chunk must be 1, 2, 4, 8, 16 or 32 (got %d)", chunk);
      return TP_ERR_INVALID;
  }
}

static int build_args(const tp_tri_scan* scans, int nscans, int chunk, int ppb, SweepArgs& a, int& sigma,
                      bool allow_half_chunk) {
  if (nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  memset(&a, 0, sizeof(a));
  int nblk = 0;
  sigma = scans[0].sigma;
  for (int i = 0; i < nscans; ++i) {
    const tp_tri_scan& d = scans[i];
    const int sch = d.chunk == 0 ? chunk : d.chunk;
    if (sch != chunk && !(allow_half_chunk && 2 * sch == chunk && chunk >= 16 && d.grad == nullptr)) {
      set_error("scan %d: chunk %d with a launch chunk of %d: a scan may take half the launch's chunk (16 or 32) "
                "in the persistent sweep without gradient accumulation", i, sch, chunk);
      return TP_ERR_INVALID;
    }
    if (d.n < 1 || d.n > 64 * sch) {
      set_error("scan %d: n = %d outside [1, 64*chunk = %d]", i, d.n, 64 * sch);
      return TP_ERR_INVALID;
    }
    if (d.steps < 1 || d.probes < 0 || !d.coef || !d.u0) {
      set_error("scan %d: steps >= 1, probes >= 0, coef and u0 are required", i);
      return TP_ERR_INVALID;
    }
    if (d.sigma != sigma) {
      set_error("scan %d: all scans of one launch must share the sigma flag", i);
      return TP_ERR_INVALID;
    }
    if (d.ckpt && d.ckpt_every < 1) {
      set_error("scan %d: ckpt_every must be >= 1 when ckpt is set", i);
      return TP_ERR_INVALID;
    }
    if (reinterpret_cast<uintptr_t>(d.coef) & 15) {
      set_error("scan %d: coef must be 16-byte aligned", i);
      return TP_ERR_INVALID;
    }
    if (d.coef_pstride != 0 && (ppb != 1 || d.grad != nullptr || (d.coef_pstride & 1) || d.coef_pstride < 0)) {
      set_error("scan %d: per-probe factors (coef_pstride = %lld) need probes_per_block = 1, no gradient "
                "accumulation and an even, positive stride", i, (long long)d.coef_pstride);
      return TP_ERR_INVALID;
    }
    if ((d.grad != nullptr) != (scans[0].grad != nullptr) || ((d.grad != nullptr) && d.gfwd == nullptr)) {
      set_error("scan %d: gradient accumulation needs grad and gfwd, on every scan of the launch", i);
      return TP_ERR_INVALID;
    }
    a.scans[i] = d;
    a.blk_start[i] = nblk;
    nblk += (d.probes + ppb - 1) / ppb;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  a.ppb = ppb;
  a.nblk = nblk;
  return TP_OK;
}

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED>
static int launch_sweep_form(const SweepArgs& a, size_t lds, hipStream_t st) {
  if (lds > 65536) {
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)tri_sweep_kernel<CH, PPB, SIGMA, GRAD, STAGED>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL((tri_sweep_kernel<CH, PPB, SIGMA, GRAD, STAGED>), dim3(a.nblk), dim3(64 * PPB), lds, st, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

// Launches of at most one workgroup per compute unit leave one wavefront per
// SIMD and take the staged step; fuller launches take the direct one.
template <int CH, int PPB, bool SIGMA, bool GRAD>
static int launch_sweep(const SweepArgs& a, hipStream_t st) {
  constexpr int NP = 64 * CH;
  const size_t lds =
      ((size_t)(SIGMA ? 6 : 5) * NP + (size_t)(PPB + (GRAD ? 1 : 0)) * (NP + NP / 8)) * sizeof(float2);
  if (lds > 163840) {
    set_error("chunk %d with %d probes per block needs %zu bytes of LDS (> 160 KiB); use fewer probes per block",
              CH, PPB, lds);
    return TP_ERR_INVALID;
  }
  if constexpr (CH <= 16) {
    if (a.nblk <= device_cus()) return launch_sweep_form<CH, PPB, SIGMA, GRAD, true>(a, lds, st);
  }
  return launch_sweep_form<CH, PPB, SIGMA, GRAD, false>(a, lds, st);
}

template <int CH, bool SIGMA>
static int sweep_ppb(const SweepArgs& a, int ppb, bool grad, hipStream_t st) {
  if (grad) {  // gradient accumulation is built for 1 and 4 probes per workgroup
    if (ppb == 1) return launch_sweep<CH, 1, SIGMA, true>(a, st);
    if (ppb == 4) return launch_sweep<CH, 4, SIGMA, true>(a, st);
    set_error("gradient sweeps need probes_per_block 1 or 4 (got %d)", ppb);
    return TP_ERR_INVALID;
  }
  switch (ppb) {
    case 1: return launch_sweep<CH, 1, SIGMA, false>(a, st);
    case 2: return launch_sweep<CH, 2, SIGMA, false>(a, st);
    case 4: return launch_sweep<CH, 4, SIGMA, false>(a, st);
    case 8: return launch_sweep<CH, 8, SIGMA, false>(a, st);
    default:
      set_error("probes_per_block must be 1, 2, 4 or 8 (got %d)", ppb);
      return TP_ERR_INVALID;
  }
}

template <int CH>
static int sweep_sigma(const SweepArgs& a, int ppb, int sigma, bool grad, hipStream_t st) {
  return sigma ? sweep_ppb<CH, true>(a, ppb, grad, st) : sweep_ppb<CH, false>(a, ppb, grad, st);
}

}  // namespace tp

using namespace tp;

extern "C" size_t tp_tri_coef_bytes(int n, int steps, int chunk) {
  (void)n;
  return (size_t)steps * A_COUNT * 64 * chunk * sizeof(float2);
}

extern "C" int tp_tri_factor(const void* d_C, int n, int ncol, const void* d_base, double lhs_sign, double rhs_sign,
                             int conj_c, int lhs_col0, int rhs_col0, int col_step, int steps, int chunk,
                             int with_sigma, void* d_coef, int* d_status, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  if (n < 1 || n > 64 * chunk || steps < 1 || ncol < 1 || !d_C || !d_base || !d_coef) {
    set_error("tp_tri_factor: invalid sizes or null pointer (n=%d, chunk=%d, steps=%d, ncol=%d)", n, chunk, steps,
              ncol);
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
  const dim3 grid((steps + FACTOR_WAVES - 1) / FACTOR_WAVES), block(64 * FACTOR_WAVES);
#define TP_FACTOR_CASE(CH) \
  case CH: hipLaunchKernelGGL((tri_factor_kernel<CH>), grid, block, 0, st, f); break;
  switch (chunk) {
    TP_FACTOR_CASE(1)
    TP_FACTOR_CASE(2)
    TP_FACTOR_CASE(4)
    TP_FACTOR_CASE(8)
    TP_FACTOR_CASE(16)
    TP_FACTOR_CASE(32)
  }
#undef TP_FACTOR_CASE
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tri_factor_batch(const tp_tri_factor_job* jobs, int njobs, int chunk, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  if (!jobs || njobs < 1 || njobs > TP_MAX_SCANS) {
    set_error("tp_tri_factor_batch: njobs must be in [1, %d] (got %d)", TP_MAX_SCANS, njobs);
    return TP_ERR_INVALID;
  }
  FactorBatch b;
  memset(&b, 0, sizeof(b));
  int nblk = 0;
  for (int k = 0; k < njobs; ++k) {
    const tp_tri_factor_job& q = jobs[k];
    if (q.n < 1 || q.n > 64 * chunk || q.steps < 1 || q.ncol < 1 || !q.C || !q.base || !q.coef) {
      set_error("tp_tri_factor_batch: job %d: invalid sizes or null pointer (n=%d, chunk=%d, steps=%d)", k, q.n,
                chunk, q.steps);
      return TP_ERR_INVALID;
    }
    FactorArgs& f = b.jobs[k];
    f.C = (const double2*)q.C;
    f.base = (const double2*)q.base;
    f.n = q.n;
    f.ncol = q.ncol;
    f.steps = q.steps;
    f.lsign = q.lhs_sign;
    f.rsign = q.rhs_sign;
    f.conj_c = q.conj_c;
    f.lcol0 = q.lhs_col0;
    f.rcol0 = q.rhs_col0;
    f.cstep = q.col_step;
    f.with_sigma = q.with_sigma;
    f.coef = (float2*)q.coef;
    f.status = q.status;
    b.blk_start[k] = nblk;
    nblk += (q.steps + FACTOR_WAVES - 1) / FACTOR_WAVES;
  }
  b.blk_start[njobs] = nblk;
  b.njobs = njobs;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(nblk), block(64 * FACTOR_WAVES);
#define TP_FACTOR_CASE(CH) \
  case CH: hipLaunchKernelGGL((tri_factor_batch_kernel<CH>), grid, block, 0, st, b); break;
  switch (chunk) {
    TP_FACTOR_CASE(1)
    TP_FACTOR_CASE(2)
    TP_FACTOR_CASE(4)
    TP_FACTOR_CASE(8)
    TP_FACTOR_CASE(16)
    TP_FACTOR_CASE(32)
  }
#undef TP_FACTOR_CASE
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tri_sweep(const tp_tri_scan* scans, int nscans, int chunk, int probes_per_block, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  SweepArgs a;
  int sigma = 0;
  rc = build_args(scans, nscans, chunk, probes_per_block, a, sigma, true);
  if (rc) return rc;
  if (a.nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool grad = scans[0].grad != nullptr;
  switch (chunk) {
    case 1: return sweep_sigma<1>(a, probes_per_block, sigma, grad, st);
    case 2: return sweep_sigma<2>(a, probes_per_block, sigma, grad, st);
    case 4: return sweep_sigma<4>(a, probes_per_block, sigma, grad, st);
    case 8: return sweep_sigma<8>(a, probes_per_block, sigma, grad, st);
    case 16: return sweep_sigma<16>(a, probes_per_block, sigma, grad, st);
    case 32: return sweep_sigma<32>(a, probes_per_block, sigma, grad, st);
  }
  return TP_ERR_INVALID;
}

extern "C" int tp_tri_step(const tp_tri_scan* scans, int nscans, int chunk, int s, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  SweepArgs a;
  int sigma = 0;
  const int ppb = 4;
  rc = build_args(scans, nscans, chunk, ppb, a, sigma, false);
  if (rc) return rc;
  for (int i = 0; i < nscans; ++i) {
    if (!scans[i].out || s < 1 || s > scans[i].steps || scans[i].grad) {
      set_error("tp_tri_step: scan %d needs out != NULL, 1 <= s <= steps (s=%d) and no gradient", i, s);
      return TP_ERR_INVALID;
    }
  }
  if (a.nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
#define TP_STEP_CASE(CH) \
  case CH: return sigma ? launch_step<CH, true>(a, s, st) : launch_step<CH, false>(a, s, st);
  switch (chunk) {
    TP_STEP_CASE(1)
    TP_STEP_CASE(2)
    TP_STEP_CASE(4)
    TP_STEP_CASE(8)
    TP_STEP_CASE(16)
    TP_STEP_CASE(32)
  }
#undef TP_STEP_CASE
  return TP_ERR_INVALID;
}
