// Crank–Nicolson slice stepping with per-slice tridiagonal LU (one transverse axis).
//
// Per step s the systems are   lhs(s) u_s = rhs(s) u_{s-1} (+ b_s)   with
// lhs/rhs tridiagonal and shared by every probe of a scan.  The factor kernel
// turns lhs(s) into an LU (FP64 pivots, no pivot search: the CN operators are
// strictly diagonally dominant) and stores, per row i and step s, complex64:
//   bd   rhs diagonal
//   l    forward multiplier  l_i = lhs_sub_i / piv_{i-1}      (y_i = r_i - l_i y_{i-1})
//   gf   prod_{m = chunk start .. i} (-l_m)                   (carry into the chunk)
//   pinv 1 / piv_i
//   q    -lhs_sup_i / piv_i                                  (x_i = y_i pinv_i + q_i x_{i+1})
//   gb   prod_{m = i .. chunk end} q_m                        (carry from the next chunk)
//
// The sweep assigns one probe to one wavefront; lane l owns the CHUNK
// consecutive rows [l*CHUNK, (l+1)*CHUNK).  Both triangular recurrences run
// chunk-locally in registers, the chunk carries are combined by a
// wavefront-wide scan of affine maps (y -> g*y + c), and a final register pass
// applies the carries.  No row is ever re-read from HBM within a step.
#include "tp_common.h"

namespace tp {

enum { A_BD = 0, A_L = 1, A_GF = 2, A_PINV = 3, A_Q = 4, A_GB = 5, A_COUNT = 6 };

__host__ __device__ constexpr int perm_index(int i, int chunk) { return (i % chunk) * 64 + i / chunk; }

// ------------------------------------------------------------------ factor
template <int CH>
__global__ __launch_bounds__(64) void tri_factor_kernel(const double2* __restrict__ C, int n, int ncol,
                                                        const double2* __restrict__ base, double lsign,
                                                        double rsign, int conj_c, int lcol0, int rcol0,
                                                        int cstep, int steps, float2* __restrict__ coef,
                                                        int* status) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= steps) return;
  constexpr int NP = 64 * CH;
  float2* o = coef + (size_t)s * A_COUNT * NP;
  const double2* lsub = base;
  const double2* ldia = base + n;
  const double2* lsup = base + 2 * n;
  const double2* rdia = base + 4 * n;
  const int lc = lcol0 + cstep * s;
  const int rc = rcol0 + cstep * s;
  const bool lok = lc >= 0 && lc < ncol;
  const bool rok = rc >= 0 && rc < ncol;
  const double2 zero = make_double2(0.0, 0.0);
  double2 pinv_prev = zero;
  bool bad = false;
  for (int t = 0; t < 64; ++t) {
    double2 qbuf[CH];
    double2 gf = make_double2(1.0, 0.0);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const int i = t * CH + m;
      const int pi = m * 64 + t;
      if (i < n) {
        double2 c = lok ? C[(size_t)i * ncol + lc] : zero;
        double2 cr = rok ? C[(size_t)i * ncol + rc] : zero;
        if (conj_c) {
          c = zconj(c);
          cr = zconj(cr);
        }
        const double2 d = zadd(ldia[i], zscale(c, lsign));
        const double2 l = (i == 0) ? zero : zmul(lsub[i], pinv_prev);
        const double2 piv = (i == 0) ? d : zsub(d, zmul(l, lsup[i - 1]));
        if (piv.x == 0.0 && piv.y == 0.0) bad = true;
        const double2 pinv = zinv(piv);
        const double2 q = (i == n - 1) ? zero : zscale(zmul(lsup[i], pinv), -1.0);
        const double2 bd = zadd(rdia[i], zscale(cr, rsign));
        gf = zmul(gf, zscale(l, -1.0));
        o[A_BD * NP + pi] = to_f2(bd);
        o[A_L * NP + pi] = to_f2(l);
        o[A_GF * NP + pi] = to_f2(gf);
        o[A_PINV * NP + pi] = to_f2(pinv);
        o[A_Q * NP + pi] = to_f2(q);
        qbuf[m] = q;
        pinv_prev = pinv;
      } else {
#pragma unroll
        for (int a = 0; a < A_COUNT; ++a) o[a * NP + pi] = make_float2(0.f, 0.f);
        qbuf[m] = zero;
      }
    }
    double2 gb = make_double2(1.0, 0.0);
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) {
      gb = zmul(gb, qbuf[m]);
      o[A_GB * NP + m * 64 + t] = to_f2(gb);
    }
  }
  if (bad && status) status[0] = 1;
}

__global__ void tri_offdiag_kernel(const double2* __restrict__ base, int n, int chunk, float2* __restrict__ offd) {
  const int np = 64 * chunk;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int pi = perm_index(i, chunk);
  float2 e = make_float2(0.f, 0.f), f = make_float2(0.f, 0.f);
  if (i < n) {
    if (i > 0) e = to_f2(base[3 * n + i]);
    if (i < n - 1) f = to_f2(base[5 * n + i]);
  }
  offd[pi] = e;
  offd[np + pi] = f;
}

// ------------------------------------------------------------------ wave scans
// Inclusive scans over the 64 lanes of affine maps f_l(y) = A_l y + B_l
// (complex A, B), built from DPP row shifts.  UP composes towards higher lanes
// (lane l ends with f_l o ... o f_0), DOWN towards lower lanes (lane l ends
// with f_l o ... o f_63).  Lanes whose DPP source is outside the row keep the
// identity map passed as the `old` operand.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f(float old, float src) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float2 dpp_c(float2 old, float2 src) {
  return make_float2(dpp_f<CTRL, ROW_MASK>(old.x, src.x), dpp_f<CTRL, ROW_MASK>(old.y, src.y));
}
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ void affine_step(float2& A, float2& B) {
  const float2 Ap = dpp_c<CTRL, ROW_MASK>(make_float2(1.f, 0.f), A);
  const float2 Bp = dpp_c<CTRL, ROW_MASK>(make_float2(0.f, 0.f), B);
  B = cfma(A, Bp, B);
  A = cmul(A, Ap);
}
// DPP controls (GFX9 encoding)
constexpr int DPP_ROW_SHL1 = 0x101, DPP_ROW_SHL2 = 0x102, DPP_ROW_SHL4 = 0x104, DPP_ROW_SHL8 = 0x108;
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

__device__ __forceinline__ void scan_affine_up(float2& A, float2& B, int lane) {
  (void)lane;
  affine_step<DPP_ROW_SHR1>(A, B);
  affine_step<DPP_ROW_SHR2>(A, B);
  affine_step<DPP_ROW_SHR4>(A, B);
  affine_step<DPP_ROW_SHR8>(A, B);
  affine_step<DPP_ROW_BCAST15, 0xA>(A, B);
  affine_step<DPP_ROW_BCAST31, 0xC>(A, B);
}

__device__ __forceinline__ float2 readlane_c(float2 v, int l) {
  return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                     __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l)));
}

__device__ __forceinline__ void scan_affine_down(float2& A, float2& B, int lane) {
  affine_step<DPP_ROW_SHL1>(A, B);
  affine_step<DPP_ROW_SHL2>(A, B);
  affine_step<DPP_ROW_SHL4>(A, B);
  affine_step<DPP_ROW_SHL8>(A, B);
  // lanes 16, 32, 48 now hold the totals of rows 1, 2, 3
  const float2 A1 = readlane_c(A, 16), B1 = readlane_c(B, 16);
  const float2 A2 = readlane_c(A, 32), B2 = readlane_c(B, 32);
  const float2 A3 = readlane_c(A, 48), B3 = readlane_c(B, 48);
  const float2 A23 = cmul(A2, A3), B23 = cfma(A2, B3, B2);
  const float2 A123 = cmul(A1, A23), B123 = cfma(A1, B23, B1);
  const int row = lane >> 4;
  float2 TA = make_float2(1.f, 0.f), TB = make_float2(0.f, 0.f);
  if (row == 0) { TA = A123; TB = B123; }
  if (row == 1) { TA = A23; TB = B23; }
  if (row == 2) { TA = A3; TB = B3; }
  B = cfma(A, TB, B);
  A = cmul(A, TA);
}

// previous lane's value (lane 0 gets 0) / next lane's value (lane 63 gets 0)
__device__ __forceinline__ float2 from_prev_lane(float2 v) { return dpp_c<DPP_WAVE_SHR1>(make_float2(0.f, 0.f), v); }
__device__ __forceinline__ float2 from_next_lane(float2 v) { return dpp_c<DPP_WAVE_SHL1>(make_float2(0.f, 0.f), v); }

// One CN step for the CH rows of this lane.  `cf` points at the step's six
// permuted coefficient arrays, `od` at the two permuted off-diagonal arrays
// (LDS or global).  `rhs` (natural order rows of this lane) may be null.
template <int CH, bool HAS_RHS, typename Ptr>
__device__ __forceinline__ void tri_step_lane(float2 (&u)[CH], Ptr cf, Ptr od, const float2 (&rhs)[CH], int row0,
                                              int n, int lane) {
  constexpr int NP = 64 * CH;
  const float2 left = from_prev_lane(u[CH - 1]);
  const float2 right = from_next_lane(u[0]);
  float2 y[CH];
#pragma unroll
  for (int m = 0; m < CH; ++m) {
    const int k = m * 64 + lane;
    const float2 um = (m == 0) ? left : u[m - 1];
    const float2 up = (m == CH - 1) ? right : u[m + 1];
    float2 r = cmul(cf[A_BD * NP + k], u[m]);
    r = cfma(od[k], um, r);
    r = cfma(od[NP + k], up, r);
    if (HAS_RHS) r = cadd(r, rhs[m]);
    y[m] = r;
  }
#pragma unroll
  for (int m = 1; m < CH; ++m) y[m] = cfms(cf[A_L * NP + m * 64 + lane], y[m - 1], y[m]);
  {
    float2 A = cf[A_GF * NP + (CH - 1) * 64 + lane];
    float2 B = y[CH - 1];
    scan_affine_up(A, B, lane);
    const float2 yin = from_prev_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) y[m] = cfma(cf[A_GF * NP + m * 64 + lane], yin, y[m]);
  }
#pragma unroll
  for (int m = 0; m < CH; ++m) y[m] = cmul(y[m], cf[A_PINV * NP + m * 64 + lane]);
#pragma unroll
  for (int m = CH - 2; m >= 0; --m) y[m] = cfma(cf[A_Q * NP + m * 64 + lane], y[m + 1], y[m]);
  {
    float2 A = cf[A_GB * NP + lane];
    float2 B = y[0];
    scan_affine_down(A, B, lane);
    const float2 xin = from_next_lane(B);
#pragma unroll
    for (int m = 0; m < CH; ++m) u[m] = cfma(cf[A_GB * NP + m * 64 + lane], xin, y[m]);
  }
}

struct SweepArgs {
  tp_tri_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans;
  int ppb;
};

__device__ __forceinline__ int find_scan(const SweepArgs& a, int blk) {
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  return sc;
}

// ------------------------------------------------------------------ persistent sweep
typedef float f4v __attribute__((ext_vector_type(4)));

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
__device__ __forceinline__ void store_plane_fast(const tp_tri_scan& d, int p, int s, const float2 (&u)[CH], int row0) {
  const int n = d.n;
  if (d.out) store_rows<CH>((float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * s) * d.out_zstride, u, row0, n);
  if (d.ckpt && (s % d.ckpt_every) == 0)
    store_rows<CH>((float2*)d.ckpt + (size_t)p * d.ckpt_pstride + (int64_t)(s / d.ckpt_every) * d.ckpt_zstride, u, row0, n);
  if (d.exitp && s == d.steps) store_rows<CH>((float2*)d.exitp + (size_t)p * d.exit_pstride, u, row0, n);
}

// One workgroup = PPB wavefronts = PPB probes of one scan walking all steps in
// lockstep.  The step's coefficients (6 * 64*CH complex64) are staged in LDS
// and shared by the PPB wavefronts; the next step's block is fetched into
// registers while the current step computes.  The plane computed in step s is
// written at the start of step s+1 so that the stores drain behind the math.
template <int CH, int PPB>
__global__ __launch_bounds__(64 * PPB) void tri_sweep_kernel(const SweepArgs a) {
  constexpr int NP = 64 * CH;
  constexpr int NT = 64 * PPB;
  constexpr int NV = A_COUNT * NP / 2;  // 16-byte vectors per step
  constexpr int PER = (NV + NT - 1) / NT;
  extern __shared__ f4v smem_v[];
  float2* smem = reinterpret_cast<float2*>(smem_v);
  const float2* cf_l = smem;
  const float2* od_l = smem + A_COUNT * NP;

  const int blk = blockIdx.x;
  const int sc = find_scan(a, blk);
  const tp_tri_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * PPB + wave;
  const bool act = p < d.probes;
  const int n = d.n;
  const int L = d.steps;
  const int row0 = lane * CH;
  const f4v* coef4 = reinterpret_cast<const f4v*>(d.coef);

  {
    const f4v* od4 = reinterpret_cast<const f4v*>(d.offd);
    for (int idx = threadIdx.x; idx < NP; idx += NT) smem_v[NV + idx] = od4[idx];
  }
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

  float2 u[CH];
  {
    const float2* u0 = (const float2*)d.u0 + (size_t)(act ? p : 0) * d.u0_pstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) u[m] = (act && row0 + m < n) ? u0[row0 + m] : make_float2(0.f, 0.f);
  }
  __syncthreads();

  float2 rb[CH];
  for (int s = 1; s <= L; ++s) {
    if (act) store_plane_fast<CH>(d, p, s - 1, u, row0);
    if (s < L) {
      const f4v* src = coef4 + (size_t)s * NV;
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        if (NV % NT == 0 || idx < NV) st[k] = src[idx];
      }
    }
    if (d.rhs != nullptr) {
      const float2* r = (const float2*)d.rhs + (size_t)(act ? p : 0) * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
      for (int m = 0; m < CH; ++m) rb[m] = (act && row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
      tri_step_lane<CH, true>(u, cf_l, od_l, rb, row0, n, lane);
    } else {
      tri_step_lane<CH, false>(u, cf_l, od_l, rb, row0, n, lane);
    }
    __syncthreads();
    if (s < L) {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * NT;
        if (NV % NT == 0 || idx < NV) smem_v[idx] = st[k];
      }
    }
    __syncthreads();
  }
  if (act) store_plane_fast<CH>(d, p, L, u, row0);
}

// ------------------------------------------------------------------ plain one-step
template <int CH>
__global__ __launch_bounds__(256) void tri_step_kernel(const SweepArgs a, int s) {
  constexpr int NP = 64 * CH;
  const int blk = blockIdx.x;
  const int sc = find_scan(a, blk);
  const tp_tri_scan& d = a.scans[sc];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int p = (blk - a.blk_start[sc]) * a.ppb + wave;
  if (wave >= a.ppb || p >= d.probes) return;
  const int n = d.n;
  const int row0 = lane * CH;
  float2 u[CH];
  const float2* src;
  if (s == 1)
    src = (const float2*)d.u0 + (size_t)p * d.u0_pstride;
  else
    src = (const float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * (s - 1)) * d.out_zstride;
#pragma unroll
  for (int m = 0; m < CH; ++m) u[m] = (row0 + m < n) ? src[row0 + m] : make_float2(0.f, 0.f);
  if (s == 1) store_plane_fast<CH>(d, p, 0, u, row0);
  float2 rb[CH];
  const float2* cf = (const float2*)d.coef + (size_t)(s - 1) * A_COUNT * NP;
  if (d.rhs != nullptr) {
    const float2* r = (const float2*)d.rhs + (size_t)p * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride;
#pragma unroll
    for (int m = 0; m < CH; ++m) rb[m] = (row0 + m < n) ? r[row0 + m] : make_float2(0.f, 0.f);
    tri_step_lane<CH, true>(u, cf, (const float2*)d.offd, rb, row0, n, lane);
  } else {
    tri_step_lane<CH, false>(u, cf, (const float2*)d.offd, rb, row0, n, lane);
  }
  store_plane_fast<CH>(d, p, s, u, row0);
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

static int build_args(const tp_tri_scan* scans, int nscans, int chunk, int ppb, SweepArgs& a, int& nblk) {
  if (nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  memset(&a, 0, sizeof(a));
  nblk = 0;
  for (int i = 0; i < nscans; ++i) {
    const tp_tri_scan& d = scans[i];
    if (d.n < 1 || d.n > 64 * chunk) {
      set_error("scan %d: n = %d outside [1, 64*chunk = %d]", i, d.n, 64 * chunk);
      return TP_ERR_INVALID;
    }
    if (d.steps < 1 || d.probes < 0 || !d.coef || !d.offd || !d.u0) {
      set_error("scan %d: steps >= 1, probes >= 0, coef, offd and u0 are required", i);
      return TP_ERR_INVALID;
    }
    if (d.ckpt && d.ckpt_every < 1) {
      set_error("scan %d: ckpt_every must be >= 1 when ckpt is set", i);
      return TP_ERR_INVALID;
    }
    if ((reinterpret_cast<uintptr_t>(d.coef) & 15) || (reinterpret_cast<uintptr_t>(d.offd) & 15)) {
      set_error("scan %d: coef and offd must be 16-byte aligned", i);
      return TP_ERR_INVALID;
    }
    a.scans[i] = d;
    a.blk_start[i] = nblk;
    nblk += (d.probes + ppb - 1) / ppb;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  a.ppb = ppb;
  return TP_OK;
}

template <int CH, int PPB>
static int launch_sweep(const SweepArgs& a, int nblk, hipStream_t st) {
  constexpr int NP = 64 * CH;
  const size_t lds = (size_t)(A_COUNT + 2) * NP * sizeof(float2);
  if (lds > 65536) {
    TP_CHECK_HIP(hipFuncSetAttribute((const void*)tri_sweep_kernel<CH, PPB>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  hipLaunchKernelGGL((tri_sweep_kernel<CH, PPB>), dim3(nblk), dim3(64 * PPB), lds, st, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

template <int CH>
static int sweep_ppb(const SweepArgs& a, int nblk, int ppb, hipStream_t st) {
  switch (ppb) {
    case 1: return launch_sweep<CH, 1>(a, nblk, st);
    case 2: return launch_sweep<CH, 2>(a, nblk, st);
    case 4: return launch_sweep<CH, 4>(a, nblk, st);
    case 8: return launch_sweep<CH, 8>(a, nblk, st);
    default:
      set_error("probes_per_block must be 1, 2, 4 or 8 (got %d)", ppb);
      return TP_ERR_INVALID;
  }
}

}  // namespace tp

using namespace tp;

extern "C" size_t tp_tri_coef_bytes(int n, int steps, int chunk) {
  (void)n;
  return (size_t)steps * A_COUNT * 64 * chunk * sizeof(float2);
}

extern "C" int tp_tri_factor(const void* d_C, int n, int ncol, const void* d_base, double lhs_sign, double rhs_sign,
                             int conj_c, int lhs_col0, int rhs_col0, int col_step, int steps, int chunk, void* d_coef,
                             int* d_status, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  if (n < 1 || n > 64 * chunk || steps < 1 || ncol < 1 || !d_C || !d_base || !d_coef) {
    set_error("tp_tri_factor: invalid sizes or null pointer (n=%d, chunk=%d, steps=%d, ncol=%d)", n, chunk, steps,
              ncol);
    return TP_ERR_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((steps + 63) / 64), block(64);
  const double2* C = (const double2*)d_C;
  const double2* B = (const double2*)d_base;
  float2* o = (float2*)d_coef;
#define TP_FACTOR_CASE(CH)                                                                                  \
  case CH:                                                                                                  \
    hipLaunchKernelGGL((tri_factor_kernel<CH>), grid, block, 0, st, C, n, ncol, B, lhs_sign, rhs_sign, conj_c, \
                       lhs_col0, rhs_col0, col_step, steps, o, d_status);                                   \
    break;
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

extern "C" int tp_tri_offdiag(const void* d_base, int n, int chunk, void* d_offd, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  if (n < 1 || n > 64 * chunk || !d_base || !d_offd) {
    set_error("tp_tri_offdiag: invalid arguments (n=%d, chunk=%d)", n, chunk);
    return TP_ERR_INVALID;
  }
  const int np = 64 * chunk;
  hipLaunchKernelGGL(tri_offdiag_kernel, dim3((np + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)d_base, n, chunk, (float2*)d_offd);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tri_sweep(const tp_tri_scan* scans, int nscans, int chunk, int probes_per_block, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  SweepArgs a;
  int nblk = 0;
  rc = build_args(scans, nscans, chunk, probes_per_block, a, nblk);
  if (rc) return rc;
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (chunk) {
    case 1: return sweep_ppb<1>(a, nblk, probes_per_block, st);
    case 2: return sweep_ppb<2>(a, nblk, probes_per_block, st);
    case 4: return sweep_ppb<4>(a, nblk, probes_per_block, st);
    case 8: return sweep_ppb<8>(a, nblk, probes_per_block, st);
    case 16: return sweep_ppb<16>(a, nblk, probes_per_block, st);
    case 32: return sweep_ppb<32>(a, nblk, probes_per_block, st);
  }
  return TP_ERR_INVALID;
}

extern "C" int tp_tri_step(const tp_tri_scan* scans, int nscans, int chunk, int s, void* stream) {
  int rc = check_chunk(chunk);
  if (rc) return rc;
  SweepArgs a;
  int nblk = 0;
  const int ppb = 4;
  rc = build_args(scans, nscans, chunk, ppb, a, nblk);
  if (rc) return rc;
  for (int i = 0; i < nscans; ++i) {
    if (!scans[i].out || s < 1 || s > scans[i].steps) {
      set_error("tp_tri_step: scan %d needs out != NULL and 1 <= s <= steps (s=%d)", i, s);
      return TP_ERR_INVALID;
    }
  }
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
#define TP_STEP_CASE(CH)                                                                  \
  case CH:                                                                                \
    hipLaunchKernelGGL((tri_step_kernel<CH>), dim3(nblk), dim3(64 * ppb), 0, st, a, s); \
    break;
  switch (chunk) {
    TP_STEP_CASE(1)
    TP_STEP_CASE(2)
    TP_STEP_CASE(4)
    TP_STEP_CASE(8)
    TP_STEP_CASE(16)
    TP_STEP_CASE(32)
  }
#undef TP_STEP_CASE
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
