// One CN step per lane, plane stores and the launch arguments shared by the
// persistent sweep (tp_tri_sweep.h, one translation unit per chunk) and the
// plain one-step kernel (tp_tri.hip).
#pragma once
#include "tp_scan.h"

namespace tp {

// One CN step for the CH rows of this lane:  u <- A^{-1}(sigma u + b) - u.
// `cf` points at the step's coefficient arrays (LDS or global, lane-interleaved).
// Without SIGMA (sigma = 2):  r = u + b/2 and u <- 2 x - u.
// With SIGMA:                 r = sigma u + b and u <- x - u.
struct NoHook {
  __device__ __forceinline__ void operator()(int) const {}
};

// Rows of a chunk taken or skipped together by the near-carry form.
template <int CH>
constexpr int kNearGroup = CH < 4 ? CH : 4;
constexpr int kNearMinChunk = 8;

// The near-carry step works from the pivots alone: with the LHS off-diagonals
// i * sub_i and i * sup_i (purely imaginary, uniform on interior rows,
// tp_tri_scan.offd), forward elimination is z_m = pinv_m (b_m - i sub_m z_{m-1})
// and back substitution x_m = z_m - pinv_m (i sup_m x_{m+1}).  Each array has
// one boundary value at either end (rows 1 and n-1 of sub, 0 and n-2 of sup); a
// lane keeps the values of its own rows: the first lane's, and at the uniform
// rows mn = (n-1) % CH, mq = (n-2) % CH those of the lanes holding rows n-1 and
// n-2.
struct Offd {
  float sub, sup;        // interior rows (imaginary parts)
  float sub_x1, sup_x0;  // this lane at m = 1 (sub) / m = 0 (sup)
  float sub_xn, sup_xn;  // this lane at m = mn (sub) / m = mq (sup)
  int mn, mq;
};

__device__ __forceinline__ Offd make_offd(const tp_tri_scan& d, int chunk, int lane) {
  const float* v = d.offd;  // (re, im) pairs; the real parts are zero
  Offd o;
  o.sub = v[1];
  o.sup = v[3];
  const float sub1 = v[5], subn = v[7], sup0 = v[9], supn = v[11];
  const int n = d.n;
  o.mn = (n - 1) % chunk;
  o.mq = n >= 2 ? (n - 2) % chunk : -1;
  o.sub_x1 = lane == 0 ? sub1 : o.sub;
  o.sup_x0 = lane == 0 ? sup0 : o.sup;
  o.sub_xn = lane == (n - 1) / chunk ? subn : (o.mn == 1 ? o.sub_x1 : o.sub);
  o.sup_xn = (n >= 2 && lane == (n - 2) / chunk) ? supn : (o.mq == 0 ? o.sup_x0 : o.sup);
  return o;
}

// imaginary parts of sub and sup at this lane's row m (m a compile-time index)
__device__ __forceinline__ float offd_sub(const Offd& o, int m) {
  float a = (m == 1) ? o.sub_x1 : o.sub;
  if (m == o.mn) a = o.sub_xn;
  return a;
}
__device__ __forceinline__ float offd_sup(const Offd& o, int m) {
  float c = (m == 0) ? o.sup_x0 : o.sup;
  if (m == o.mq) c = o.sup_xn;
  return c;
}
// c - i b z and i b z (b real)
__device__ __forceinline__ float2 cfms_i(float b, float2 z, float2 c) {
  return make_float2(fmaf(b, z.y, c.x), fmaf(-b, z.x, c.y));
}
__device__ __forceinline__ float2 cmul_i(float b, float2 z) { return make_float2(-b * z.y, b * z.x); }

// Near-carry phases shared by both step forms.  Forward: elimination within
// the chunk, then the carry of the previous chunk's last z: its increment
// dz_{-1} = z_in goes through dz_m = pinv_m (-i sub_m dz_{m-1}) onto the first
// `reach` rows (rounded to groups of G).  Backward: substitution within the
// chunk, then the next chunk's first x through dx_m = pinv_m (-i sup_m
// dx_{m+1}) onto the last `reach` rows.
template <int CH, int G>
__device__ __forceinline__ void near_forward(float2 (&y)[CH], const float2 (&pv)[CH], const Offd& od, int reach) {
  y[0] = cmul(pv[0], y[0]);
#pragma unroll
  for (int m = 1; m < CH; ++m) y[m] = cmul(pv[m], cfms_i(offd_sub(od, m), y[m - 1], y[m]));
  float2 c = from_prev_lane(y[CH - 1]);
#pragma unroll
  for (int g = 0; g < CH; g += G) {
    if (g < reach) {
#pragma unroll
      for (int m = g; m < g + G; ++m) {
        c = cmul(pv[m], cmul_i(-offd_sub(od, m), c));
        y[m] = cadd(y[m], c);
      }
    }
  }
}
template <int CH, int G>
__device__ __forceinline__ void near_backward(float2 (&y)[CH], const float2 (&pv)[CH], const Offd& od, int reach) {
#pragma unroll
  for (int m = CH - 2; m >= 0; --m) y[m] = cfms(pv[m], cmul_i(offd_sup(od, m), y[m + 1]), y[m]);
  float2 d = from_next_lane(y[0]);
#pragma unroll
  for (int g = CH - G; g >= 0; g -= G) {
    if (g + G > CH - reach) {
#pragma unroll
      for (int m = g + G - 1; m >= g; --m) {
        d = cmul(pv[m], cmul_i(-offd_sup(od, m), d));
        y[m] = cadd(y[m], d);
      }
    }
  }
}

// The near-carry step once y holds the right-hand side and pv the pivots.
template <int CH, bool SIGMA>
__device__ __forceinline__ void near_solve(float2 (&u)[CH], float2 (&y)[CH], const float2 (&pv)[CH], const Offd& od,
                                           int reach) {
  constexpr int G = kNearGroup<CH>;
  near_forward<CH, G>(y, pv, od, reach);
  near_backward<CH, G>(y, pv, od, reach);
#pragma unroll
  for (int m = 0; m < CH; ++m)
    u[m] = SIGMA ? csub(y[m], u[m]) : make_float2(fmaf(2.f, y[m].x, -u[m].x), fmaf(2.f, y[m].y, -u[m].y));
}

// Near-carry step without sigma, pivots already in registers.
template <int CH, bool HAS_RHS>
__device__ __forceinline__ void near_step_regs(float2 (&u)[CH], const float2 (&pv)[CH], const float2 (&rhs)[CH],
                                               const Offd& od, int reach) {
  float2 y[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m)
    y[m] = HAS_RHS ? make_float2(fmaf(0.5f, rhs[m].x, u[m].x), fmaf(0.5f, rhs[m].y, u[m].y)) : u[m];
  near_solve<CH, false>(u, y, pv, od, reach);
}

// Direct form: every coefficient is read where it is used (the register
// budget of CH = 32 holds no second bank).

template <int CH, bool HAS_RHS, bool SIGMA, bool NEAR, typename Ptr, typename Hook>
__device__ __forceinline__ void tri_step_lane_direct(float2 (&u)[CH], Ptr cf, const float2 (&rhs)[CH], int lane,
                                                     int reach, const Offd& od, const Hook& hook) {
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
  if constexpr (NEAR) {  // pinv of every row gives l, q and the carries
    float2 pv[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) pv[m] = cf[A_PINV * NP + m * 64 + lane];
    near_solve<CH, SIGMA>(u, y, pv, od, reach);
    hook(3);
    return;
  }
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
template <int CH, bool HAS_RHS, bool SIGMA, bool STAGED = true, bool NEAR = false, typename Ptr,
          typename Hook = NoHook>
__device__ __forceinline__ void tri_step_lane(float2 (&u)[CH], Ptr cf, const float2 (&rhs)[CH], int lane,
                                              int reach = CH, const Offd& od = Offd{}, const Hook& hook = Hook()) {
  if constexpr (CH > 16 || !STAGED) {
    tri_step_lane_direct<CH, HAS_RHS, SIGMA, NEAR>(u, cf, rhs, lane, reach, od, hook);
    return;
  }
  constexpr int NP = 64 * CH;
  float2 y[CH], c1[CH], c2[CH];
  // near: c1 holds pinv for the whole step (l, q and the carries derive from it)
#pragma unroll
  for (int m = 0; m < CH; ++m) c1[m] = cf[(NEAR ? A_PINV : A_L) * NP + m * 64 + lane];
  if constexpr (!NEAR) {
#pragma unroll
    for (int m = 0; m < CH; ++m) c2[m] = cf[A_GF * NP + m * 64 + lane];
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
  if constexpr (NEAR) {
    near_solve<CH, SIGMA>(u, y, c1, od, reach);
    hook(3);
    return;
  }
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
  int nblk;         // logical workgroups
  int nblk_hw;      // workgroups launched (persistent sweep: XCD-aware order)
  int cus_per_xcd;  // 0: workgroups in launch order
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

// Destinations of plane s of probe p in the plain one-step form.
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
  __device__ __forceinline__ void all_batched() const;
};

// Output planes of one probe in a persistent sweep: the destinations of plane
// z = 0, 1, ..., steps in turn (full field, checkpoints, exit plane), kept as
// wave-uniform pointers advanced per plane rather than recomputed from the
// scan's strides every step.
struct PlaneCursor {
  float2* out;
  float2* ck;
  float2* ex;
  int64_t out_step, ck_step;
  int ck_every, ck_left, steps;

  __device__ __forceinline__ void init(const tp_tri_scan& d, int p, bool act) {
    out = (act && d.out) ? (float2*)d.out + (size_t)p * d.out_pstride + (int64_t)d.out_z0 * d.out_zstride : nullptr;
    out_step = (int64_t)d.out_zdir * d.out_zstride;
    ck = (act && d.ckpt) ? (float2*)d.ckpt + (size_t)p * d.ckpt_pstride : nullptr;
    ck_step = d.ckpt_zstride;
    ck_every = d.ckpt_every;
    ck_left = 0;
    ex = (act && d.exitp) ? (float2*)d.exitp + (size_t)p * d.exit_pstride : nullptr;
    steps = d.steps;
  }
  // destinations of plane z (called for z = 0, 1, ... in order)
  __device__ __forceinline__ void next(int z, float2* (&dst)[3]) {
    dst[0] = out;
    if (out) out += out_step;
    dst[1] = nullptr;
    if (ck) {
      if (ck_left == 0) {
        dst[1] = ck;
        ck += ck_step;
        ck_left = ck_every;
      }
      --ck_left;
    }
    dst[2] = (z == steps) ? ex : nullptr;
  }
};

// Stores the plane held in the wavefront's tile to up to three destinations:
// every LDS read of the plane first (one wait), then 16-byte stores of 64
// consecutive row pairs per instruction.  At 16-byte aligned destinations the
// first n/128 instructions are whole and take no per-lane bounds checks; the
// destinations go through one copy of the store code.  The stores are
// non-temporal: the planes stream out without displacing the per-step
// coefficient blocks from L2 (config-1 launch 3.32 -> 3.23 ms).
template <int CH>
__device__ __forceinline__ void write_plane(const float2* tile, float2* const (&dst)[3], int n, int lane) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  constexpr int NP = 64 * CH;
  constexpr int KP = CH / 2 > 0 ? CH / 2 : 1;  // pair-instructions per plane
  float2 *o0 = dst[0], *o1 = dst[1], *o2 = dst[2];
  if (o0 == nullptr && o1 == nullptr && o2 == nullptr) return;
  f4 v[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int j = 2 * (lane + 64 * k);
    v[k] = (CH >= 2 || j < NP) ? *reinterpret_cast<const f4*>(tile + tile_index(j)) : f4{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll 1
  for (int t = 0; t < 3; ++t) {
    float2* o = o0;
    o0 = o1;
    o1 = o2;
    o2 = nullptr;
    if (o == nullptr) continue;
    const bool aligned = (reinterpret_cast<uintptr_t>(o) & 15) == 0;
    const int kfull = (aligned && CH >= 2) ? (n >> 7) : 0;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int j = 2 * (lane + 64 * k);
      if (k < kfull) {
        __builtin_nontemporal_store(v[k], reinterpret_cast<f4*>(o + j));
      } else if (aligned && j + 1 < n) {
        __builtin_nontemporal_store(v[k], reinterpret_cast<f4*>(o + j));
      } else {
        if (j < n) o[j] = make_float2(v[k].x, v[k].y);
        if (j + 1 < n) o[j + 1] = make_float2(v[k].z, v[k].w);
      }
    }
  }
}

template <int CH>
__device__ __forceinline__ void PlaneWriter<CH>::all_batched() const {
  write_plane<CH>(tile, dst, n, lane);
}

// Compute units of the current device (cached per device).
inline int device_cus() {
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

// Logical workgroup of hardware workgroup h.  Workgroups are dispatched to the
// 8 XCDs round-robin (h % 8), and within an XCD the j-th (j = h / 8) takes
// compute unit j % C (C units per XCD) once every unit has one.  The logical
// workgroups (scans in launch order, widest first) are dealt out in rounds of
// 8 C: in each round every XCD takes a contiguous run of them, so that a
// scan's workgroups, which read the same per-step coefficient blocks, share
// few L2 caches, and each unit takes at most one workgroup per round.  The
// last round may leave slots empty (-1).  C = 0: launch order.
__device__ __forceinline__ int sweep_block(const SweepArgs& a, int h) {
  const int C = a.cus_per_xcd;
  if (C == 0) return h;
  const int x = h & 7, j = h >> 3;
  const int r = j / C, i = j - r * C;
  const int first = r * 8 * C;
  const int nr = min(a.nblk - first, 8 * C);
  const int q = nr >> 3, rem = nr & 7;
  const int cnt = q + (x < rem ? 1 : 0);
  if (i >= cnt) return -1;
  return first + x * q + min(x, rem) + i;
}

// Host side of sweep_block: with C compute units per XCD (device_cus() / 8),
// the launch takes 8 * (full rounds * C + the last round's largest run) slots.
inline void sweep_xcd_order(SweepArgs& a) {
  a.nblk_hw = a.nblk;
  a.cus_per_xcd = 0;
  const int cus = device_cus();
  if (cus >= 8 && cus % 8 == 0 && a.nblk > 0) {
    const int C = cus / 8;
    const int rounds = (a.nblk + 8 * C - 1) / (8 * C);
    const int last = a.nblk - (rounds - 1) * 8 * C;
    a.cus_per_xcd = C;
    a.nblk_hw = 8 * ((rounds - 1) * C + (last + 7) / 8);
  }
}

// The persistent sweep of one chunk (tp_tri_sweep_c<chunk>.hip).
int tri_sweep_launch_c1(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);
int tri_sweep_launch_c2(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);
int tri_sweep_launch_c4(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);
int tri_sweep_launch_c8(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);
int tri_sweep_launch_c16(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);
int tri_sweep_launch_c32(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st);

}  // namespace tp
