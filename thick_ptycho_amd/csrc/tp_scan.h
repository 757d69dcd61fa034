// Wavefront scans and factor types shared by the tridiagonal stepping kernels
// (one wavefront per probe in tp_tri.hip, several in tp_triw.hip).
#pragma once
#include "tp_common.h"

namespace tp {

// Per-row coefficient arrays of one step, in this order.
enum { A_L = 0, A_GF = 1, A_PINV = 2, A_Q = 3, A_GB = 4, A_SG = 5, A_COUNT = 6 };

// Coupling reach.  The carry into a chunk from the previous one enters row m
// with the factor gf_m = prod_{k <= m} (-l_k), and the carry from the next one
// with gb_m = prod_{k >= m} q_k.  The factor kernels record, per factor set,
// the largest number of rows of a chunk (from its first row forward, or its
// last row backward) whose factor exceeds kCarryTau in |re| + |im|.  When that
// reach is below the chunk, every chunk's total product is below kCarryTau,
// so the carry into chunk l is the neighbouring chunk's local end value alone
// (the terms dropped pass through at least one whole chunk) and only the
// first/last `reach` rows take it.  kCarryTau = 2^-32 is 256 times below the
// FP32 unit roundoff: the dropped terms are smaller than the rounding of one
// addition on those rows.
constexpr double kCarryTau = 0x1p-32;

// Arguments of one per-step factorisation job (see tp_tri_factor).
struct FactorArgs {
  const double2* C;
  const double2* base;
  int n, ncol, steps;
  double lsign, rsign;
  int conj_c, lcol0, rcol0, cstep, with_sigma;
  float2* coef;
  int* status;
};

struct Mob {  // 2x2 complex matrix [[a, b], [c, d]]
  double2 a, b, c, d;
};

__device__ __forceinline__ Mob mob_mul(const Mob& x, const Mob& y) {  // x * y
  Mob r;
  r.a = zadd(zmul(x.a, y.a), zmul(x.b, y.c));
  r.b = zadd(zmul(x.a, y.b), zmul(x.b, y.d));
  r.c = zadd(zmul(x.c, y.a), zmul(x.d, y.c));
  r.d = zadd(zmul(x.c, y.b), zmul(x.d, y.d));
  return r;
}
__device__ __forceinline__ double zabs1(double2 v) { return fabs(v.x) + fabs(v.y); }
__device__ __forceinline__ Mob mob_normalise(Mob m) {
  const double s = fmax(fmax(zabs1(m.a), zabs1(m.b)), fmax(zabs1(m.c), zabs1(m.d)));
  if (s > 0.0) {
    const double r = 1.0 / s;
    m.a = zscale(m.a, r);
    m.b = zscale(m.b, r);
    m.c = zscale(m.c, r);
    m.d = zscale(m.d, r);
  }
  return m;
}
__device__ __forceinline__ Mob mob_identity() {
  const double2 zero = make_double2(0.0, 0.0), one = make_double2(1.0, 0.0);
  return Mob{one, zero, zero, one};
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

__device__ __forceinline__ void scan_affine_up(float2& A, float2& B) {
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

// previous lane's value (lane 0 gets `old`) / next lane's value (lane 63 gets `old`)
__device__ __forceinline__ float2 from_prev_lane(float2 v, float2 old = make_float2(0.f, 0.f)) {
  return dpp_c<DPP_WAVE_SHR1>(old, v);
}
__device__ __forceinline__ float2 from_next_lane(float2 v, float2 old = make_float2(0.f, 0.f)) {
  return dpp_c<DPP_WAVE_SHL1>(old, v);
}

// Plane tiles in LDS: row i at i + 2*(i/16) (two pad words per 16 rows against
// bank conflicts; row pairs stay 16-byte aligned).
__device__ __forceinline__ int tile_index(int i) { return i + 2 * (i >> 4); }

typedef float f4v __attribute__((ext_vector_type(4)));

}  // namespace tp
