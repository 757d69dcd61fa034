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
#include "tp_tri_step.h"

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

  // l, gf, q and gb go out with non-temporal stores: a near-carry sweep reads
  // only pinv (and sigma), which then keep their place in the caches (a sweep
  // whose pivots were pushed out to HBM by other writes is measured 15 %
  // slower on the config-1 launch)
  float2* o = f.coef + (size_t)s * A_COUNT * NP;
  auto put_nt = [](float2* p, float2 v) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    f2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
  };
  double2 q[CH];
  double2 gf = one;
  bool bad = false;
  int reach = 0;  // coupling reach of this lane's chunk (see below)
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
      if (zabs1(gf) > kCarryTau) reach = m + 1;
      put_nt(o + A_L * NP + pi, to_f2(l));
      put_nt(o + A_GF * NP + pi, to_f2(gf));
      o[A_PINV * NP + pi] = to_f2(pinv);
      put_nt(o + A_Q * NP + pi, to_f2(q[m]));
      if (f.with_sigma) {
        double2 cr = rok ? f.C[(size_t)i * f.ncol + rc] : zero;
        if (f.conj_c) cr = zconj(cr);
        o[A_SG * NP + pi] = to_f2(zadd(d[m], zadd(rdia[i], zscale(cr, f.rsign))));
      }
      piv_prev = piv;
    } else {
      q[m] = zero;
      put_nt(o + A_L * NP + pi, make_float2(0.f, 0.f));
      put_nt(o + A_GF * NP + pi, make_float2(0.f, 0.f));
      o[A_PINV * NP + pi] = make_float2(0.f, 0.f);
      put_nt(o + A_Q * NP + pi, make_float2(0.f, 0.f));
      if (f.with_sigma) o[A_SG * NP + pi] = make_float2(0.f, 0.f);
    }
  }
  double2 gb = one;
#pragma unroll
  for (int m = CH - 1; m >= 0; --m) {
    gb = zmul(gb, q[m]);
    if (zabs1(gb) > kCarryTau) reach = max(reach, CH - m);
    put_nt(o + A_GB * NP + m * 64 + lane, to_f2(gb));
  }
  if (f.status) {
    if (bad) f.status[0] = 1;
    // one atomic per wavefront: the largest reach over its lanes
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) reach = max(reach, __shfl_xor(reach, off));
    if (lane == 0) atomicMax(f.status + 1, reach);
  }
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
  const bool has_rhs = d.rhs != nullptr && (d.rhs_steps == 0 || s <= d.rhs_steps);
  if (has_rhs) {
    const float2* r = (const float2*)d.rhs + (size_t)p * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) rb[m] = (row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
  }
  // near-carry step as in the persistent sweep (the factors' reach and the LHS
  // off-diagonals): only the pivots are read, no 64-lane scans
  bool near = false;
  if constexpr (CH >= kNearMinChunk) {
    const int reach = (d.reach && d.has_offd) ? __builtin_amdgcn_readfirstlane(*d.reach) : CH;
    if (reach < CH) {
      near = true;
      constexpr int G = kNearGroup<CH>;
      const int rr = (reach + G - 1) / G * G;
      const Offd od = make_offd(d, CH, lane);
      if (has_rhs)
        tri_step_lane<CH, true, SIGMA, true, true>(u, cf, rb, lane, rr, od);
      else
        tri_step_lane<CH, false, SIGMA, true, true>(u, cf, rb, lane, rr, od);
    }
  }
  if (!near) {
    if (has_rhs)
      tri_step_lane<CH, true, SIGMA>(u, cf, rb, lane);
    else
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
                      bool allow_half_chunk, bool widest_first = false) {
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
  if (widest_first) {
    // Workgroups in launch order take one compute unit each before any takes a
    // second: the widest scans first, so that a unit that takes two pairs a
    // wide workgroup with a narrower one (stable order among equal widths).
    tp_tri_scan sorted[TP_MAX_SCANS];
    int m = 0;
    for (int w = 64 * 32; w >= 1 && m < nscans; --w)
      for (int i = 0; i < nscans; ++i)
        if (a.scans[i].n == w) sorted[m++] = a.scans[i];
    nblk = 0;
    for (int i = 0; i < nscans; ++i) {
      a.scans[i] = sorted[i];
      a.blk_start[i] = nblk;
      nblk += (sorted[i].probes + ppb - 1) / ppb;
    }
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  a.ppb = ppb;
  a.nblk = nblk;
  a.nblk_hw = nblk;
  a.cus_per_xcd = 0;
  if (widest_first) sweep_xcd_order(a);
  return TP_OK;
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
  rc = build_args(scans, nscans, chunk, probes_per_block, a, sigma, true, true);
  if (rc) return rc;
  if (a.nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool grad = scans[0].grad != nullptr;
  bool rhs = false;
  for (int i = 0; i < nscans; ++i) rhs = rhs || scans[i].rhs != nullptr;
  switch (chunk) {
    case 1: return tri_sweep_launch_c1(a, probes_per_block, sigma, grad, rhs, st);
    case 2: return tri_sweep_launch_c2(a, probes_per_block, sigma, grad, rhs, st);
    case 4: return tri_sweep_launch_c4(a, probes_per_block, sigma, grad, rhs, st);
    case 8: return tri_sweep_launch_c8(a, probes_per_block, sigma, grad, rhs, st);
    case 16: return tri_sweep_launch_c16(a, probes_per_block, sigma, grad, rhs, st);
    case 32: return tri_sweep_launch_c32(a, probes_per_block, sigma, grad, rhs, st);
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
