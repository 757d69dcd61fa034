// Persistent batched sweep of the tridiagonal CN stepping (tp_tri_sweep): the
// kernel templates and their launchers.  Each chunk's kernels are instantiated
// in their own translation unit (tp_tri_sweep_c<chunk>.hip) so that they build
// in parallel.
#pragma once
#include "tp_tri_step.h"

namespace tp {

// ------------------------------------------------------------------ persistent sweep

// Whether staging vector k of a step (threads' 16-byte vectors k*NT .. k*NT+NT-1)
// holds rows a near-carry step reads: pinv and sigma, but none of l, q, gf and
// gb, which the step forms from pinv and the operator's off-diagonals (uniform
// per k when a vector block of NT lies inside one array, else everything is
// staged).
template <int CH, int NT>
__device__ __forceinline__ bool stage_needed(int k) {
  constexpr int VPA = 32 * CH;  // 16-byte vectors per array
  if constexpr (NT > VPA) {
    return true;
  } else {
    const int arr = (k * NT) / VPA;
    return arr == A_PINV || arr == A_SG;
  }
}

// Blocks of 64 * CH values ahead of the plane tiles in LDS: the staged
// coefficient arrays, or one pivot buffer per wavefront in the wave mode.
template <int PPB, bool SIGMA, bool GRAD>
constexpr int coef_slots() {
  constexpr int na = SIGMA ? 6 : 5;
  return (!SIGMA && !GRAD && PPB > na) ? PPB : na;
}

template <int CH, int PPB, bool SIGMA, bool GRAD>
constexpr size_t sweep_body_lds() {
  constexpr int NP = 64 * CH, TILE = NP + NP / 8;
  return ((size_t)coef_slots<PPB, SIGMA, GRAD>() * NP + (size_t)(PPB + (GRAD ? 1 : 0)) * TILE) * sizeof(float2);
}

// s_waitcnt operands (gfx9 encoding): every counter at zero; LDS only.
constexpr int kWaitAll = 0, kWaitLds = 0xC07F;

// One wavefront copies NP pivots (8 * NP bytes) from global memory into its
// LDS buffer without passing through registers: 16 bytes per lane and
// instruction, 1 KiB per instruction.
template <int NP>
__device__ __forceinline__ void pivots_to_lds(const float2* src, float2* buf, int lane) {
  static_assert(NP % 128 == 0, "whole 1 KiB blocks");
  const f4v* s = reinterpret_cast<const f4v*>(src) + lane;
#pragma unroll
  for (int i = 0; i < NP / 128; ++i)
    __builtin_amdgcn_global_load_lds((const void*)(s + 64 * i),
                                     (__attribute__((address_space(3))) void*)(buf + 128 * i), 16, 0, 0);
}

// One workgroup = PPB wavefronts = PPB probes of one scan walking all steps in
// lockstep.  The step's coefficients are staged in LDS and shared by the PPB
// wavefronts; the next step's block is fetched into registers while the
// current step computes.  Each step first writes the plane of the previous
// step, then issues that fetch: vmcnt retires in order, so the wait for the
// fetch at the end of the step only covers stores that have had the whole step
// to drain.  (Measured on the config-1 launch: issuing the fetch before the
// stores, 4.09 ms; two coefficient blocks with one barrier per step and each
// wavefront's tile in its share of the idle block, 3.9-4.3 ms; this form,
// 3.5 ms.)  RHS: scans may carry an extra right-hand side (scans without one
// take b = 0).
//
// Wave mode: a near-carry step without sigma or gradient reads only the
// pivots.  Each wavefront copies the next step's into an LDS buffer of its own
// after its plane stores (the same one-step lead and in-order wait as the
// shared block), and the wavefronts of a workgroup step without workgroup
// barriers.
template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED, bool NEAR, bool RHS>
__device__ __forceinline__ void tri_sweep_body(const SweepArgs& a, const int blk, const int sc, const int reach) {
  constexpr bool WAVE = NEAR && !SIGMA && !GRAD;
  constexpr int NP = 64 * CH;
  constexpr int NT = 64 * PPB;
  constexpr int NA = SIGMA ? 6 : 5;
  constexpr int NV = NA * NP / 2;             // 16-byte vectors staged per step
  constexpr int STRIDE_V = A_COUNT * NP / 2;  // 16-byte vectors per step in memory
  constexpr int TILE = NP + NP / 8;           // per-wavefront store tile (float2)
  constexpr int PER = (NV + NT - 1) / NT;
  constexpr int NS = coef_slots<PPB, SIGMA, GRAD>();
  extern __shared__ f4v smem_v[];
  float2* const smem = reinterpret_cast<float2*>(smem_v);

  const tp_tri_scan& d = a.scans[sc];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * PPB + wave;
  const bool act = p < d.probes;
  PlaneCursor cur;
  cur.init(d, p, act);
  const Offd od = NEAR ? make_offd(d, CH, lane) : Offd{};
  const int n = d.n;
  const int L = d.steps;
  const int row0 = lane * CH;
  // per-probe factors (one probe per workgroup) or the scan's shared factors
  const f4v* coef4 = reinterpret_cast<const f4v*>(d.coef) + (PPB == 1 ? (size_t)(act ? p : 0) * (d.coef_pstride / 2) : 0);
  // LDS: the step's coefficient block, the PPB plane tiles, the gradient row
  const float2* const cf = smem;
  float2* const tile = smem + NS * NP + wave * TILE;
  float2* const gbuf = smem + NS * NP + PPB * TILE;
  float2* const pbuf = smem + wave * NP;  // wave mode: this wavefront's pivots
  // staging vector k of this thread, and whether the step reads it
  auto sneed = [&](int k) -> bool {
    const bool in = NV % NT == 0 || (int)threadIdx.x + k * NT < NV;
    if constexpr (NEAR) return in && stage_needed<CH, NT>(k);
    return in;
  };

  f4v st[PER];
  if constexpr (WAVE) {
    pivots_to_lds<NP>(reinterpret_cast<const float2*>(coef4) + A_PINV * NP, pbuf, lane);
  } else {
#pragma unroll
    for (int k = 0; k < PER; ++k)
      if (sneed(k)) st[k] = coef4[threadIdx.x + k * NT];
#pragma unroll
    for (int k = 0; k < PER; ++k)
      if (sneed(k)) smem_v[threadIdx.x + k * NT] = st[k];
  }

  // The initial plane arrives through the tile: coalesced, and the state
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

  float2 ga[CH], gb[CH];
  if (GRAD) {
    for (int idx = threadIdx.x; idx < TILE; idx += NT) gbuf[idx] = make_float2(0.f, 0.f);
    const float2* g = (const float2*)d.gfwd + (size_t)(act ? p : 0) * d.gfwd_pstride + (int64_t)d.gfwd_z0 * d.gfwd_zstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) ga[m] = (act && row0 + m < n) ? g[row0 + m] : make_float2(0.f, 0.f);
  }
  __syncthreads();

  float2 rb[CH];
  for (int s = 1; s <= L; ++s) {
    float2 pv[WAVE ? CH : 1];
    if constexpr (WAVE) {
      // this step's pivots, copied a step ago behind the previous plane's
      // stores: the wait covers stores that have had the whole step to drain
      __builtin_amdgcn_s_waitcnt(kWaitAll);
      asm volatile("" ::: "memory");
#pragma unroll
      for (int m = 0; m < CH; ++m) pv[m] = pbuf[m * 64 + lane];
    }
    {
      tile_put<CH>(tile, u, row0);
      float2* dst[3];
      cur.next(s - 1, dst);
      write_plane<CH>(tile, dst, n, lane);
    }
    if constexpr (WAVE) {
      if (s < L) {  // the pivot reads above are done before the copy lands
        __builtin_amdgcn_s_waitcnt(kWaitLds);
        asm volatile("" ::: "memory");
        pivots_to_lds<NP>(reinterpret_cast<const float2*>(coef4 + (size_t)s * STRIDE_V) + A_PINV * NP, pbuf, lane);
      }
    }
    if (!WAVE && s < L) {
      const f4v* src = coef4 + (size_t)s * STRIDE_V;
#pragma unroll
      for (int k = 0; k < PER; ++k)
        if (sneed(k)) st[k] = src[threadIdx.x + k * NT];
    }
    if (GRAD) {  // plane b = u_{s-1} of this step
      const float2* g =
          (const float2*)d.gfwd + (size_t)(act ? p : 0) * d.gfwd_pstride + (int64_t)(d.gfwd_z0 + d.gfwd_zdir * s) * d.gfwd_zstride;
#pragma unroll
      for (int m = 0; m < CH; ++m) gb[m] = (act && row0 + m < n) ? g[row0 + m] : make_float2(0.f, 0.f);
    }
    if constexpr (RHS) {
      if (d.rhs != nullptr && (d.rhs_steps == 0 || s <= d.rhs_steps)) {
        const float2* r = (const float2*)d.rhs + (size_t)(act ? p : 0) * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
        for (int m = 0; m < CH; ++m) rb[m] = (act && row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
      } else {
#pragma unroll
        for (int m = 0; m < CH; ++m) rb[m] = make_float2(0.f, 0.f);
      }
    }
    if constexpr (WAVE)
      near_step_regs<CH, RHS>(u, pv, rb, od, reach);
    else
      tri_step_lane<CH, RHS, SIGMA, STAGED, NEAR>(u, cf, rb, lane, reach, od);
    if constexpr (WAVE) {  // the tile is this wavefront's own: the next step's writes follow its reads
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    const int grow = d.grad_z0 + d.grad_zdir * (s - 1);
    if (GRAD) {
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
      for (int k = 0; k < PER; ++k)
        if (sneed(k)) smem_v[threadIdx.x + k * NT] = st[k];
    }
    __syncthreads();
  }
  tile_put<CH>(tile, u, row0);
  {
    float2* dst[3];
    cur.next(L, dst);
    write_plane<CH>(tile, dst, n, lane);
  }
}

// The body for a workgroup's scan: its chunk (the launch's or half of it) and
// the near-carry or full step.
template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED, bool RHS>
__device__ __forceinline__ void tri_sweep_dispatch(const SweepArgs& a, const int blk, const int sc, const int reach) {
  const tp_tri_scan& d = a.scans[sc];
  // the near-carry step takes rows in groups of kNearGroup: round the reach up
  // (a carry on a few more rows is exact, just not needed)
  if constexpr (!GRAD && CH >= 16) {
    if (d.chunk == CH / 2) {
      constexpr int G = kNearGroup<CH / 2>;
      if (reach < CH / 2)
        tri_sweep_body<CH / 2, PPB, SIGMA, GRAD, STAGED, true, RHS>(a, blk, sc, (reach + G - 1) / G * G);
      else
        tri_sweep_body<CH / 2, PPB, SIGMA, GRAD, STAGED, false, RHS>(a, blk, sc, CH / 2);
      return;
    }
  }
  if constexpr (CH >= kNearMinChunk) {
    constexpr int G = kNearGroup<CH>;
    if (reach < CH) {
      tri_sweep_body<CH, PPB, SIGMA, GRAD, STAGED, true, RHS>(a, blk, sc, (reach + G - 1) / G * G);
      return;
    }
  }
  tri_sweep_body<CH, PPB, SIGMA, GRAD, STAGED, false, RHS>(a, blk, sc, CH);
}

// A workgroup's probes all belong to one scan, so the choice between the
// launch's chunk and half of it (tp_tri_scan.chunk) is uniform per workgroup;
// the LDS is sized for the launch's chunk and the narrow body uses a prefix.
// So is the choice of the near-carry step (tp_tri_scan.reach below the scan's
// chunk), built for chunks of 8 rows and more: shorter chunks pass a carry
// through several lanes at the coupling of the CN operators.
template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED, bool RHS>
__global__ __launch_bounds__(64 * PPB) void tri_sweep_kernel(const SweepArgs a) {
  const int blk = sweep_block(a, (int)blockIdx.x);
  if (blk < 0) return;  // padding slot of the XCD-aware order (whole workgroup)
  const int sc = find_scan(a, blk);
  const tp_tri_scan& d = a.scans[sc];
  // the near-carry step needs the factors' reach and the LHS off-diagonals
  const int reach = (d.reach && d.has_offd) ? __builtin_amdgcn_readfirstlane(*d.reach) : 64;
  tri_sweep_dispatch<CH, PPB, SIGMA, GRAD, STAGED, RHS>(a, blk, sc, reach);
}

template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED, bool RHS>
static int launch_sweep_form(const SweepArgs& a, size_t lds, hipStream_t st) {
  if (lds > 65536) {
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)tri_sweep_kernel<CH, PPB, SIGMA, GRAD, STAGED, RHS>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL((tri_sweep_kernel<CH, PPB, SIGMA, GRAD, STAGED, RHS>), dim3(a.nblk_hw), dim3(64 * PPB), lds,
                     st, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

template <int CH, int PPB, bool SIGMA, bool GRAD, bool STAGED>
static int launch_sweep_rhs(const SweepArgs& a, size_t lds, bool rhs, hipStream_t st) {
  return rhs ? launch_sweep_form<CH, PPB, SIGMA, GRAD, STAGED, true>(a, lds, st)
             : launch_sweep_form<CH, PPB, SIGMA, GRAD, STAGED, false>(a, lds, st);
}

// Launches of at most one workgroup per compute unit leave one wavefront per
// SIMD and take the staged step; fuller launches take the direct one.  The
// LDS covers the launch's chunk and the half-chunk body of narrow scans.
template <int CH, int PPB, bool SIGMA, bool GRAD>
static int launch_sweep(const SweepArgs& a, bool rhs, hipStream_t st) {
  size_t lds = sweep_body_lds<CH, PPB, SIGMA, GRAD>();
  if constexpr (!GRAD && CH >= 16) lds = lds > sweep_body_lds<CH / 2, PPB, SIGMA, GRAD>() ? lds : sweep_body_lds<CH / 2, PPB, SIGMA, GRAD>();
  if (lds > 163840) {
    set_error("chunk %d with %d probes per block needs %zu bytes of LDS (> 160 KiB); use fewer probes per block",
              CH, PPB, lds);
    return TP_ERR_INVALID;
  }
  if constexpr (CH <= 16) {
    if (a.nblk <= device_cus()) return launch_sweep_rhs<CH, PPB, SIGMA, GRAD, true>(a, lds, rhs, st);
  }
  return launch_sweep_rhs<CH, PPB, SIGMA, GRAD, false>(a, lds, rhs, st);
}

template <int CH, bool SIGMA>
static int sweep_ppb(const SweepArgs& a, int ppb, bool grad, bool rhs, hipStream_t st) {
  if (grad) {  // gradient accumulation is built for 1 and 4 probes per workgroup
    if (ppb == 1) return launch_sweep<CH, 1, SIGMA, true>(a, rhs, st);
    if (ppb == 4) return launch_sweep<CH, 4, SIGMA, true>(a, rhs, st);
    set_error("gradient sweeps need probes_per_block 1 or 4 (got %d)", ppb);
    return TP_ERR_INVALID;
  }
  switch (ppb) {
    case 1: return launch_sweep<CH, 1, SIGMA, false>(a, rhs, st);
    case 2: return launch_sweep<CH, 2, SIGMA, false>(a, rhs, st);
    case 4: return launch_sweep<CH, 4, SIGMA, false>(a, rhs, st);
    case 8: return launch_sweep<CH, 8, SIGMA, false>(a, rhs, st);
    default:
      set_error("probes_per_block must be 1, 2, 4 or 8 (got %d)", ppb);
      return TP_ERR_INVALID;
  }
}

template <int CH>
static int sweep_sigma(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st) {
  return sigma ? sweep_ppb<CH, true>(a, ppb, grad, rhs, st) : sweep_ppb<CH, false>(a, ppb, grad, rhs, st);
}

}  // namespace tp
