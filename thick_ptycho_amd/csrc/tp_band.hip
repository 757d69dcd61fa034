// CN slice stepping with two transverse axes: the per-slice operator is block
// tridiagonal (ny blocks of nx x nx),
//
//     A = [ A_00  e_0 I                      ]
//         [ f_1 I A_11  e_1 I                ]     A_jj tridiagonal (x stencil + object term),
//         [        ...    ...    ...         ]     couplings scaled identities (y stencil),
//         [               f_{ny-1} I  A_{ny-1,ny-1} ]
//
// which is the reference's pentadiagonal-in-band matrix with bandwidth nx
// (operator_matrices.py:75-126, boundary_conditions.py:96-130).
//
// Factor (once per slice, FP64): block LU, S_0 = A_00, S_j = A_jj - f_j e_{j-1} S_{j-1}^{-1}.
// The Schur inverses W_j = S_j^{-1} of these diagonally dominant blocks decay
// geometrically away from the diagonal; each W_j is kept on a band of
// half-width w chosen by the caller.  The band of W_j is computed exactly from
// the banded LDU factorisation of S_j (no fill inside the band) with the
// Takahashi recurrences
//     Z = D^{-1} L^{-1} + (I - U) Z,   Z = U^{-1} D^{-1} + Z (I - L),
// which need only band entries of Z.  The largest |W_j| entry at distance w
// (relative to the largest diagonal entry) is reported per slice so the caller
// can widen the band until what is dropped lies below single-precision
// rounding.  W_j is stored as complex64 diagonals in pairs (16-byte loads):
// W_j(i, i+d-w) at W[s][j] + ((d/2) * nx + i) * 2 + d % 2, d = 0 .. 2w (+1 zero).
//
// Sweep (per probe, persistent over slices): with A + B = 2I (or diag(sigma)),
//     u_s = 2 A^{-1} u_{s-1} - u_{s-1}      (sigma: A^{-1}(sigma u_{s-1}) - u_{s-1})
// and A^{-1} r by block forward elimination / back substitution:
//     y_j = r_j - f_j z_{j-1},  z_j = W_j y_j            (j = 0 .. ny-1)
//     x_{ny-1} = z_{ny-1},      x_j = z_j - e_j W_j x_{j+1}
// One workgroup per probe, one transverse row per thread; every global access
// of a thread is to its own row of each block, so only the band mat-vec
// exchanges data (through LDS).
#include <algorithm>

#include "tp_common.h"

namespace tp {

struct BandFactorArgs {
  const double2* C;    // (N, ncol) row-major, N = nx * ny
  const double2* tri;  // [3][3][nx]: block type (first, interior, last) x (sub, diag, sup)
  const double2* cpl;  // [ny][2]: (f_j, e_j)
  int nx, ny, ncol, steps, w;
  double lsign, rsign;
  int conj_c, lcol0, rcol0, cstep, with_sigma;
  float2* W;
  float2* sigma;
  float* edge;
  int* status;
};

__device__ __forceinline__ double zabs(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

// sigma(t) = 2 + lsign C(t, lc) + rsign C(t, rc) (columns outside the object dropped)
__device__ __forceinline__ float2 band_sigma_at(const double2* C, int ncol, int64_t t, int lc, bool lok, int rc, bool rok,
                                                double lsign, double rsign, int conj_c) {
  double2 v = make_double2(2.0, 0.0);
  if (lok) {
    double2 c = C[t * ncol + lc];
    if (conj_c) c = zconj(c);
    v = zadd(v, zscale(c, lsign));
  }
  if (rok) {
    double2 c = C[t * ncol + rc];
    if (conj_c) c = zconj(c);
    v = zadd(v, zscale(c, rsign));
  }
  return to_f2(v);
}

// DPP move of a double (two 32-bit halves) with a row-local control
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over lanes [0, NL) of the wavefront (other lanes must hold 0), uniform
// result: within each row of 16 lanes quad_perm [1,0,3,2], [2,3,0,1], then
// row_half_mirror and row_mirror pair every lane with the other quads / half,
// leaving the row sum in all 16 lanes; rows are then added with readlane.
template <int NL>
__device__ __forceinline__ double lanes_sum(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  double t = readlane_d(v, 0);
  if (NL > 16) t += readlane_d(v, 16);
  if (NL > 32) t += readlane_d(v, 32);
  if (NL > 48) t += readlane_d(v, 48);
  return t;
}

template <int W>
__global__ __launch_bounds__(64) void band_factor_kernel(const BandFactorArgs f) {
  constexpr int BW = 2 * W + 1;
  constexpr int UPD = (W * W + 63) / 64;  // LDU updates per lane
  extern __shared__ double2 S[];          // band of the current block: S[i * BW + d] = M(i, i + d - W)
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  const int nx = f.nx, ny = f.ny;
  const int lc = f.lcol0 + f.cstep * s;
  const int rc = f.rcol0 + f.cstep * s;
  const bool lok = lc >= 0 && lc < f.ncol;
  const bool rok = rc >= 0 && rc < f.ncol;
  const double2 zero = make_double2(0.0, 0.0);
  double edge = 0.0, dmax = 0.0;
  bool singular = false;

  double2* Dg = S + nx * BW;  // [3][nx]: sub, diagonal (+ object term), super of A_jj
  for (int j = 0; j < ny; ++j) {
    const int type = (j == 0) ? 0 : (j == ny - 1 ? 2 : 1);
    const double2* tri = f.tri + (size_t)type * 3 * nx;
    const double2 fe = (j > 0) ? zmul(f.cpl[2 * j], f.cpl[2 * (j - 1) + 1]) : zero;
    const double2 mfe = make_double2(-fe.x, -fe.y);
    // stage A_jj's three diagonals in LDS: all of a lane's global loads are
    // issued together instead of one dependent load per band entry
    for (int i0 = 0; i0 < nx; i0 += 256) {
      double2 a[4], b[4], c[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = min(i0 + lane + 64 * q, nx - 1);
        a[q] = tri[i];
        b[q] = tri[nx + i];
        c[q] = tri[2 * nx + i];
        if (lok) {
          double2 cc = f.C[(size_t)(j * nx + i) * f.ncol + lc];
          if (f.conj_c) cc = zconj(cc);
          b[q] = zadd(b[q], zscale(cc, f.lsign));
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + lane + 64 * q;
        if (i < nx) {
          Dg[i] = a[q];
          Dg[nx + i] = b[q];
          Dg[2 * nx + i] = c[q];
        }
      }
    }
    __syncthreads();
    // S_j = A_jj - f_j e_{j-1} W_{j-1}  (W_{j-1} is still in S)
    for (int t = lane; t < nx * BW; t += 64) {
      const int i = t / BW, d = t - i * BW, col = i + d - W;
      double2 v = zero;
      if (col >= 0 && col < nx) {
        if (j > 0) v = zmul(mfe, S[t]);
        if (d >= W - 1 && d <= W + 1) v = zadd(v, Dg[(d - W + 1) * nx + i]);
      }
      S[t] = v;
    }
    __syncthreads();

    // banded LDU, no pivoting (the blocks are strictly diagonally dominant).
    // Step k updates the trailing window with the unscaled column / row k; the
    // scaling of column / row k into L and U is done during step k + 1, which
    // touches neither, so one barrier per step suffices.
    double2 inv_prev = zero;
    int m_prev = 0;
    for (int k = 0; k <= nx; ++k) {
      const int m = k < nx ? min(W, nx - 1 - k) : 0;
      double2 inv = zero;
      if (k < nx) {
        const double2 piv = S[k * BW + W];
        if (piv.x == 0.0 && piv.y == 0.0) singular = true;
        inv = zinv(piv);
        if (lane == 0) Dg[k] = inv;  // 1 / D_k for the Takahashi diagonal (Dg is free after the build)
#pragma unroll
        for (int u = 0; u < UPD; ++u) {
          const int t = lane + 64 * u;
          const int a = t / W, b = t - a * W;
          if (t < W * W && a < m && b < m) {
            const int i = k + 1 + a, c = k + 1 + b;
            const double2 sik = S[i * BW + (k - i + W)];
            const double2 skc = S[k * BW + (c - k + W)];
            const int o = i * BW + (c - i + W);
            S[o] = zsub(S[o], zmul(zmul(sik, skc), inv));
          }
        }
      }
      if (k > 0) {
        const int kp = k - 1;
        if (lane < m_prev) {
          const int o = (kp + 1 + lane) * BW + (W - 1 - lane);  // L(kp+1+lane, kp)
          S[o] = zmul(S[o], inv_prev);
        } else if (lane >= W && lane - W < m_prev) {
          const int o = kp * BW + (W + 1 + lane - W);  // U(kp, kp+1+lane-W)
          S[o] = zmul(S[o], inv_prev);
        }
      }
      __syncthreads();
      inv_prev = inv;
      m_prev = m;
    }

    // Takahashi: band of Z = S^{-1}, rows from the bottom.  Lanes [0, W): upper
    // entries Z(i, i+1+lane); lanes [W, 2W): lower entries Z(i+1+lane-W, i).
    for (int i = nx - 1; i >= 0; --i) {
      const int m = min(W, nx - 1 - i);
      const bool up = lane < W, lo = lane >= W && lane < 2 * W;
      const int c = i + 1 + (up ? lane : lane - W);
      const bool act = (up || lo) && c - i - 1 < m;
      double2 acc = zero;
#pragma unroll
      for (int q = 0; q < W; ++q) {
        if (act && q < m) {
          const int r = i + 1 + q;
          // upper: U(i, r) Z(r, c); lower: Z(c, r) L(r, i)
          const double2 x = up ? S[i * BW + (r - i + W)] : S[c * BW + (r - c + W)];
          const double2 y = up ? S[r * BW + (c - r + W)] : S[r * BW + (i - r + W)];
          acc = zadd(acc, zmul(x, y));
        }
      }
      const double2 val = make_double2(-acc.x, -acc.y);
      double2 prod = zero;
      if (act && lo) prod = zmul(S[i * BW + (c - i + W)], val);  // U(i, c) Z(c, i)
      const double px = lanes_sum<2 * W>(prod.x), py = lanes_sum<2 * W>(prod.y);
      const double2 dinv = Dg[i];
      __syncthreads();
      if (act) {
        if (up) {
          S[i * BW + (c - i + W)] = val;
        } else {
          S[c * BW + (i - c + W)] = val;
        }
        if (c - i == W) edge = fmax(edge, zabs(val));
      }
      if (lane == 0) {
        const double2 z = zsub(dinv, make_double2(px, py));
        S[i * BW + W] = z;
        dmax = fmax(dmax, zabs(z));
      }
      __syncthreads();
    }

    // pair layout: (d, i) at ((d/2) * nx + i) * 2 + d % 2, one padding diagonal
    float2* Wout = f.W + ((size_t)s * ny + j) * (BW + 1) * nx;
    for (int t = lane; t < nx * (BW + 1); t += 64) {
      const int pr = t >> 1, dp = pr / nx, i = pr - dp * nx, d = 2 * dp + (t & 1);
      Wout[t] = d < BW ? to_f2(S[i * BW + d]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }

  if (f.with_sigma) {  // sigma = lhs diagonal + rhs diagonal = 2 + object terms of both slices
    const int N = nx * ny;
    float2* sg = f.sigma + (size_t)s * N;
    for (int t = lane; t < N; t += 64)
      sg[t] = band_sigma_at(f.C, f.ncol, t, lc, lok, rc, rok, f.lsign, f.rsign, f.conj_c);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    edge = fmax(edge, __shfl_xor(edge, off, 64));
    dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
  }
  if (lane == 0) {
    f.edge[s] = (float)(dmax > 0.0 ? edge / dmax : 0.0);
    if (singular) atomicOr(f.status, 1);
  }
}

// ------------------------------------------------------------------ sweep
struct BandSweepArgs {
  tp_band_scan scans[TP_MAX_SCANS];
  int blk_start[TP_MAX_SCANS + 1];
  int nscans, nblk;
};

__device__ __forceinline__ int band_find_scan(const BandSweepArgs& a, int blk) {
  int sc = 0;
#pragma unroll
  for (int i = 1; i < TP_MAX_SCANS; ++i)
    if (i < a.nscans && blk >= a.blk_start[i]) sc = i;
  return sc;
}

// Workgroups are dispatched round-robin over the 8 XCDs; consecutive logical
// blocks (probes of one scan, which read the same band coefficients) are kept on
// one XCD's L2.
__device__ __forceinline__ int band_xcd_block(int hw, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = hw & 7, j = hw >> 3;
  return x * q + min(x, r) + j;
}

__device__ __forceinline__ float2* band_plane(const tp_band_scan& d, int p, int s, int64_t N) {
  if (d.out) return (float2*)d.out + (size_t)p * d.out_pstride + (int64_t)(d.out_z0 + d.out_zdir * s) * d.out_zstride;
  return (float2*)d.work + (size_t)p * d.work_pstride + (int64_t)(s & 1) * N;
}

typedef float f4v __attribute__((ext_vector_type(4)));

// One thread's row of one block: the 2W+1 band coefficients as W+1 pairs
// (16-byte loads; the pair layout puts (d, i) at ((d/2) * nx + i) * 2 + d % 2).
template <int W>
struct BandRow {
  f4v p[W + 1];
};

template <int W>
__device__ __forceinline__ void load_row(BandRow<W>& r, const float2* blk, int nx, int i, bool ok) {
  const f4v* b4 = reinterpret_cast<const f4v*>(blk);
#pragma unroll
  for (int q = 0; q <= W; ++q) r.p[q] = ok ? b4[(size_t)q * nx + i] : f4v{0.f, 0.f, 0.f, 0.f};
}

// z = sum_d W(i, i+d-W) v(i+d-W), v staged in xb at offset W with zero pads.
template <int W>
__device__ __forceinline__ float2 row_mv(const BandRow<W>& r, const float2* xb, int i) {
  float2 z = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q <= W; ++q) {
    z = cfma(make_float2(r.p[q].x, r.p[q].y), xb[i + 2 * q], z);
    if (2 * q + 1 < 2 * W + 1) z = cfma(make_float2(r.p[q].z, r.p[q].w), xb[i + 2 * q + 1], z);
  }
  return z;
}

// Blocks are processed in groups of G; the next group's operands are fetched
// into a second register bank while the current group computes, so a step
// waits only on loads issued a whole group earlier.  A workgroup steps PB
// probes together: each thread's row of W_j is loaded once and applied to all
// PB probes, which divides the band-coefficient traffic (the dominant HBM
// stream when many probes share a slice) by PB.  G keeps two banks of G rows
// within the register budget and one group's loads below the 63 outstanding
// vector-memory operations the wait counter tracks.
// OCC selects the other schedule: one block per bank and at most 256 VGPRs, so
// that two workgroups share a CU and hide each other's latency instead.
template <int W, int PB, bool OCC = false>
struct BandGroup {
  static constexpr int R = (W + 1) * 4 + PB * 4 + 4;  // VGPRs of one block of one bank
  static constexpr int GP = 100 / R;
  static constexpr int G = OCC ? 1
                               : (PB == 1 ? (W <= 4 ? 4 : (W <= 6 ? 3 : (W <= 12 ? 2 : 1)))
                                          : (GP < 1 ? 1 : (GP > 4 ? 4 : GP)));
};

template <int W, int G, int PB>
struct FwdBank {
  BandRow<W> w[G];
  float2 u[G][PB], b[G][PB], sg[G], f[G];
};

template <int W, int G, int PB>
struct BwdBank {
  BandRow<W> w[G];
  float2 z[G][PB], u[G][PB], e[G];
};

struct BandCtx {
  const float2* Ws;  // this slice's band coefficients
  const float2* prev;  // probe k's previous plane at prev + k * prev_ps
  float2* cur;
  const float2* sg;
  const float2* rb;
  const float2* cpl;
  int64_t prev_ps, cur_ps, rb_ps;
  int nx, ny, i, ic;  // ic = min(i, nx - 1): always-valid row for loads
  int np;             // probes of this workgroup (1 .. PB)
  bool act;
};

// Prefetch loads are unconditional from clamped indices (row ic = min(i, nx-1),
// block clamped into [0, ny-1], probe clamped into [0, np-1]): values of rows
// >= nx or of absent probes are never stored, and blocks outside the range are
// never used.
template <int W, int G, int PB, bool SIGMA, bool RHS>
__device__ __forceinline__ void fwd_load(FwdBank<W, G, PB>& bk, int j0, const BandCtx& c) {
  const f4v* w4 = reinterpret_cast<const f4v*>(c.Ws);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = min(j0 + q, c.ny - 1);
    const uint32_t o = (uint32_t)(j * c.nx + c.ic);
    const uint32_t ow = (uint32_t)(j * (W + 1) * c.nx + c.ic);
#pragma unroll
    for (int t = 0; t <= W; ++t) bk.w[q].p[t] = w4[ow + t * c.nx];
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int kk = min(k, c.np - 1);
      bk.u[q][k] = c.prev[kk * c.prev_ps + o];
      if (RHS) bk.b[q][k] = c.rb[kk * c.rb_ps + o];
    }
    if (SIGMA) bk.sg[q] = c.sg[o];
    bk.f[q] = c.cpl[2 * j];
  }
}

template <int W, int G, int PB>
__device__ __forceinline__ void bwd_load(BwdBank<W, G, PB>& bk, int j0, const BandCtx& c) {
  const f4v* w4 = reinterpret_cast<const f4v*>(c.Ws);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = max(j0 - q, 0);
    const uint32_t o = (uint32_t)(j * c.nx + c.ic);
    const uint32_t ow = (uint32_t)(j * (W + 1) * c.nx + c.ic);
#pragma unroll
    for (int t = 0; t <= W; ++t) bk.w[q].p[t] = w4[ow + t * c.nx];
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const int kk = min(k, c.np - 1);
      bk.z[q][k] = c.cur[kk * c.cur_ps + o];
      bk.u[q][k] = c.prev[kk * c.prev_ps + o];
    }
    bk.e[q] = c.cpl[2 * j + 1];
  }
}

template <int W, int G, int PB, bool SIGMA, bool RHS>
__device__ __forceinline__ void fwd_run(const FwdBank<W, G, PB>& bk, int j0, const BandCtx& c, float2* xbuf, int XS,
                                        int& xsel, float2 (&zprev)[PB]) {
  const float2 zero = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = j0 + q;
    if (j >= c.ny) break;
    float2* xb = xbuf + xsel * PB * XS;
    xsel ^= 1;
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      float2 y;
      if (SIGMA) {
        y = cmul(bk.sg[q], bk.u[q][k]);
        if (RHS) y = cadd(y, bk.b[q][k]);
      } else {
        y = RHS ? cfma(make_float2(0.5f, 0.f), bk.b[q][k], bk.u[q][k]) : bk.u[q][k];
      }
      y = cfms(bk.f[q], zprev[k], y);  // f_0 = 0
      xb[k * XS + W + c.i] = c.act ? y : zero;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const float2 z = row_mv<W>(bk.w[q], xb + k * XS, c.i);
      if (c.act && k < c.np) c.cur[k * c.cur_ps + (uint32_t)(j * c.nx + c.i)] = z;
      zprev[k] = z;
    }
  }
}

template <int W, int G, int PB, bool SIGMA>
__device__ __forceinline__ void bwd_run(const BwdBank<W, G, PB>& bk, int j0, const BandCtx& c, float2* xbuf, int XS,
                                        int& xsel, float2 (&xnext)[PB]) {
  const float2 zero = make_float2(0.f, 0.f);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = j0 - q;
    if (j < 0) break;
    float2 x[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) x[k] = bk.z[q][k];
    if (j < c.ny - 1) {
      float2* xb = xbuf + xsel * PB * XS;
      xsel ^= 1;
#pragma unroll
      for (int k = 0; k < PB; ++k) xb[k * XS + W + c.i] = c.act ? xnext[k] : zero;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PB; ++k) x[k] = cfms(bk.e[q], row_mv<W>(bk.w[q], xb + k * XS, c.i), bk.z[q][k]);
    }
#pragma unroll
    for (int k = 0; k < PB; ++k) {
      const float2 un = SIGMA ? csub(x[k], bk.u[q][k]) : csub(cscale(x[k], 2.f), bk.u[q][k]);
      if (c.act && k < c.np) c.cur[k * c.cur_ps + (uint32_t)(j * c.nx + c.i)] = un;
      xnext[k] = x[k];
    }
  }
}

// Each thread copies its own rows of a plane (it wrote them; no other thread
// touches them), so no synchronisation is needed.
__device__ __forceinline__ void copy_rows(float2* dst, const float2* src, int nx, int ny, int i) {
  for (int j = 0; j < ny; ++j) dst[j * nx + i] = src[j * nx + i];
}

__device__ __forceinline__ int64_t band_plane_stride(const tp_band_scan& d) {
  return d.out ? d.out_pstride : d.work_pstride;
}

template <int W, int PB, bool SIGMA, bool RHS, bool OCC>
__global__ __launch_bounds__(256, OCC ? 2 : 1) void band_sweep_kernel(const BandSweepArgs a) {
  constexpr int G = BandGroup<W, PB, OCC>::G;
  extern __shared__ float2 xbuf[];  // two exchange buffers of PB x (NT + 2W)
  const int NT = blockDim.x;
  const int blk = band_xcd_block(blockIdx.x, a.nblk);
  const int sc = band_find_scan(a, blk);
  const tp_band_scan d = a.scans[sc];
  const int p0 = (blk - a.blk_start[sc]) * PB;
  const int np = min(PB, d.probes - p0);
  const int i = threadIdx.x;
  const int nx = d.nx, ny = d.ny, L = d.steps;
  const int64_t N = (int64_t)nx * ny;
  const bool act = i < nx;
  const int XS = NT + 2 * W;
  for (int t = i; t < 2 * PB * XS; t += NT) xbuf[t] = make_float2(0.f, 0.f);
  __syncthreads();
  int xsel = 0;
  const float2 zero = make_float2(0.f, 0.f);

  BandCtx c;
  c.cpl = (const float2*)d.cpl;
  c.nx = nx;
  c.ny = ny;
  c.i = i;
  c.ic = min(i, nx - 1);
  c.act = act;
  c.np = np;
  c.cur_ps = band_plane_stride(d);
  c.rb_ps = d.rhs_pstride;
  const size_t slice_w = (size_t)ny * (2 * W + 2) * nx;
  const float2* u0 = (const float2*)d.u0 + (size_t)p0 * d.u0_pstride;
  if (act) {
    for (int k = 0; k < np; ++k) {
      const float2* src = u0 + k * d.u0_pstride;
      if (d.out) copy_rows(band_plane(d, p0 + k, 0, N), src, nx, ny, i);
      if (d.ckpt) copy_rows((float2*)d.ckpt + (size_t)(p0 + k) * d.ckpt_pstride, src, nx, ny, i);
    }
  }

  for (int s = 1; s <= L; ++s) {
    c.prev = (s == 1) ? u0 : band_plane(d, p0, s - 1, N);
    c.prev_ps = (s == 1) ? d.u0_pstride : c.cur_ps;
    c.cur = band_plane(d, p0, s, N);
    c.Ws = (const float2*)d.W + (size_t)(s - 1) * slice_w;
    c.sg = SIGMA ? (const float2*)d.sigma + (size_t)(s - 1) * N : nullptr;
    c.rb = RHS ? (const float2*)d.rhs + (size_t)p0 * d.rhs_pstride + (int64_t)(s - 1) * d.rhs_sstride : nullptr;
    const bool with_b = RHS && (d.rhs_steps == 0 || s <= d.rhs_steps);

    // forward elimination: y_j = r_j - f_j z_{j-1}, z_j = W_j y_j (z kept in the output plane)
    {
      float2 zprev[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) zprev[k] = zero;
      FwdBank<W, G, PB> A, B;
      if (with_b) {
        fwd_load<W, G, PB, SIGMA, RHS>(A, 0, c);
        for (int j0 = 0; j0 < ny; j0 += 2 * G) {
          fwd_load<W, G, PB, SIGMA, RHS>(B, j0 + G, c);
          fwd_run<W, G, PB, SIGMA, RHS>(A, j0, c, xbuf, XS, xsel, zprev);
          fwd_load<W, G, PB, SIGMA, RHS>(A, j0 + 2 * G, c);
          fwd_run<W, G, PB, SIGMA, RHS>(B, j0 + G, c, xbuf, XS, xsel, zprev);
        }
      } else {
        fwd_load<W, G, PB, SIGMA, false>(A, 0, c);
        for (int j0 = 0; j0 < ny; j0 += 2 * G) {
          fwd_load<W, G, PB, SIGMA, false>(B, j0 + G, c);
          fwd_run<W, G, PB, SIGMA, false>(A, j0, c, xbuf, XS, xsel, zprev);
          fwd_load<W, G, PB, SIGMA, false>(A, j0 + 2 * G, c);
          fwd_run<W, G, PB, SIGMA, false>(B, j0 + G, c, xbuf, XS, xsel, zprev);
        }
      }
    }

    // back substitution: x_j = z_j - e_j W_j x_{j+1}; u_s = 2x - u_{s-1} (sigma: x - u_{s-1})
    {
      float2 xnext[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k) xnext[k] = zero;
      BwdBank<W, G, PB> A, B;
      bwd_load<W, G, PB>(A, ny - 1, c);
      for (int j0 = ny - 1; j0 >= 0; j0 -= 2 * G) {
        bwd_load<W, G, PB>(B, j0 - G, c);
        bwd_run<W, G, PB, SIGMA>(A, j0, c, xbuf, XS, xsel, xnext);
        bwd_load<W, G, PB>(A, j0 - 2 * G, c);
        bwd_run<W, G, PB, SIGMA>(B, j0 - G, c, xbuf, XS, xsel, xnext);
      }
    }

    if (act) {
      for (int k = 0; k < np; ++k) {
        const float2* src = c.cur + k * c.cur_ps;
        if (d.exitp && s == L) copy_rows((float2*)d.exitp + (size_t)(p0 + k) * d.exit_pstride, src, nx, ny, i);
        if (d.ckpt && (s % d.ckpt_every) == 0)
          copy_rows((float2*)d.ckpt + (size_t)(p0 + k) * d.ckpt_pstride + (int64_t)(s / d.ckpt_every) * d.ckpt_zstride,
                    src, nx, ny, i);
      }
    }
  }
}

template <int W, int PB, bool SIGMA, bool OCC>
static int launch_band_sweep(const BandSweepArgs& a, int nt, bool rhs, hipStream_t st) {
  const size_t lds = (size_t)2 * PB * (nt + 2 * W) * sizeof(float2);
  if (rhs)
    hipLaunchKernelGGL((band_sweep_kernel<W, PB, SIGMA, true, OCC>), dim3(a.nblk), dim3(nt), lds, st, a);
  else
    hipLaunchKernelGGL((band_sweep_kernel<W, PB, SIGMA, false, OCC>), dim3(a.nblk), dim3(nt), lds, st, a);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

// schedule 0: deep prefetch (one workgroup per CU); 1: two workgroups per CU
// (built for w <= 8 and 1 or 2 probes per workgroup)
template <int W, bool SIGMA>
static int launch_band_sweep_pb(const BandSweepArgs& a, int nt, bool rhs, int pb, int schedule, hipStream_t st) {
  if (schedule == 1) {
    if constexpr (W <= 8) {
      if (pb == 1) return launch_band_sweep<W, 1, SIGMA, true>(a, nt, rhs, st);
      if (pb == 2) return launch_band_sweep<W, 2, SIGMA, true>(a, nt, rhs, st);
    }
    set_error("tp_band_sweep_ex: schedule 1 is built for w <= 8 and 1 or 2 probes per workgroup (w=%d, %d)", W, pb);
    return TP_ERR_UNSUPPORTED;
  }
  switch (pb) {
    case 1:
      return launch_band_sweep<W, 1, SIGMA, false>(a, nt, rhs, st);
    case 2:
      return launch_band_sweep<W, 2, SIGMA, false>(a, nt, rhs, st);
    case 4:
      return launch_band_sweep<W, 4, SIGMA, false>(a, nt, rhs, st);
  }
  set_error("tp_band_sweep: probes_per_block must be 1, 2 or 4 (got %d)", pb);
  return TP_ERR_INVALID;
}

static bool band_w_supported(int w) {
  return w == 4 || w == 6 || w == 8 || w == 12 || w == 16 || w == 24 || w == 32;
}

// ------------------------------------------------------------------ factors of A^H from those of A
// A^H has blocks A_jj^H and couplings conj(e_{j-1}) (below) and conj(f_{j+1})
// (above), so its block recurrence gives S'_j = (A_jj - f_j e_{j-1} W_{j-1})^H =
// S_j^H and W'_j = W_j^H: W'_j(i, i+d-w) = conj(W_j(i+d-w, i)), which is band
// index 2w - d of row i + d - w.  With `reverse`, slot a of the output holds
// slice steps-1-a of the input (the adjoint sweep walks the slices backwards).
__global__ __launch_bounds__(256) void band_conj_transpose_kernel(const float2* __restrict__ src,
                                                                  float2* __restrict__ dst, int nx, int ny, int steps,
                                                                  int w, int reverse) {
  const int BW = 2 * w + 1;
  const int PW = (BW + 1) * nx;  // complex values per block in the pair layout
  const int blk = blockIdx.x;    // slot * ny + j
  const int slot = blk / ny, j = blk - slot * ny;
  const int sslot = reverse ? steps - 1 - slot : slot;
  const float2* S = src + ((size_t)sslot * ny + j) * PW;
  float2* D = dst + (size_t)blk * PW;
  for (int t = threadIdx.x; t < PW; t += blockDim.x) {
    const int pr = t >> 1, dp = pr / nx, i = pr - dp * nx, d = 2 * dp + (t & 1);
    const int r = i + d - w;
    float2 v = make_float2(0.f, 0.f);
    if (d < BW && r >= 0 && r < nx) {
      const int dd = 2 * w - d;
      const float2 x = S[((dd >> 1) * nx + r) * 2 + (dd & 1)];
      v = make_float2(x.x, -x.y);
    }
    D[t] = v;
  }
}

__global__ __launch_bounds__(256) void band_sigma_kernel(const double2* C, int64_t N, int ncol, double lsign,
                                                         double rsign, int conj_c, int lcol0, int rcol0, int cstep,
                                                         float2* sigma) {
  const int s = blockIdx.y;
  const int lc = lcol0 + cstep * s, rc = rcol0 + cstep * s;
  const bool lok = lc >= 0 && lc < ncol, rok = rc >= 0 && rc < ncol;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += (int64_t)gridDim.x * blockDim.x)
    sigma[(size_t)s * N + t] = band_sigma_at(C, ncol, t, lc, lok, rc, rok, lsign, rsign, conj_c);
}

}  // namespace tp

using namespace tp;

extern "C" int64_t tp_band_coef_bytes(int nx, int ny, int steps, int w) {
  return (int64_t)steps * ny * (2 * w + 2) * nx * (int64_t)sizeof(float2);
}

extern "C" int tp_band_factor(const void* d_C, int nx, int ny, int ncol, const void* d_tri, const void* d_cpl,
                              double lhs_sign, double rhs_sign, int conj_c, int lhs_col0, int rhs_col0, int col_step,
                              int steps, int w, int with_sigma, void* d_W, void* d_sigma, float* d_edge,
                              int* d_status, void* stream) {
  if (!d_C || !d_tri || !d_cpl || !d_W || !d_edge || !d_status || (with_sigma && !d_sigma)) {
    set_error("tp_band_factor: null pointer");
    return TP_ERR_INVALID;
  }
  if (nx < 2 || ny < 2 || steps < 1 || ncol < 1 || w < 1 || w > 32 || w >= nx) {
    set_error("tp_band_factor: invalid sizes (nx=%d, ny=%d, steps=%d, ncol=%d, w=%d; need nx, ny >= 2, "
              "w in {4, 6, 8, 12, 16, 24, 32}, w < nx)", nx, ny, steps, ncol, w);
    return TP_ERR_INVALID;
  }
  const size_t lds = (size_t)nx * (2 * w + 4) * sizeof(double2);  // band + three staged diagonals
  if (lds > 163840) {
    set_error("tp_band_factor: band of %d rows x %d diagonals needs %zu bytes of LDS (> 160 KiB)", nx, 2 * w + 1, lds);
    return TP_ERR_UNSUPPORTED;
  }
  BandFactorArgs f;
  f.C = (const double2*)d_C;
  f.tri = (const double2*)d_tri;
  f.cpl = (const double2*)d_cpl;
  f.nx = nx;
  f.ny = ny;
  f.ncol = ncol;
  f.steps = steps;
  f.w = w;
  f.lsign = lhs_sign;
  f.rsign = rhs_sign;
  f.conj_c = conj_c;
  f.lcol0 = lhs_col0;
  f.rcol0 = rhs_col0;
  f.cstep = col_step;
  f.with_sigma = with_sigma;
  f.W = (float2*)d_W;
  f.sigma = (float2*)d_sigma;
  f.edge = d_edge;
  f.status = d_status;
  hipStream_t st = (hipStream_t)stream;
#define TP_FACTOR_W(WV)                                                                                \
  case WV:                                                                                            \
    if (lds > 65536)                                                                                  \
      TP_CHECK_HIP(hipFuncSetAttribute((const void*)band_factor_kernel<WV>,                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));         \
    hipLaunchKernelGGL((band_factor_kernel<WV>), dim3(steps), dim3(64), lds, st, f);                  \
    break;
  switch (w) {
    TP_FACTOR_W(4)
    TP_FACTOR_W(6)
    TP_FACTOR_W(8)
    TP_FACTOR_W(12)
    TP_FACTOR_W(16)
    TP_FACTOR_W(24)
    TP_FACTOR_W(32)
    default:
      set_error("tp_band_factor: band half-width %d not built (4, 6, 8, 12, 16, 24, 32)", w);
      return TP_ERR_UNSUPPORTED;
  }
#undef TP_FACTOR_W
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_band_conj_transpose(const void* d_W, int nx, int ny, int steps, int w, int reverse_slices,
                                      void* d_WH, void* stream) {
  if (!d_W || !d_WH || d_W == d_WH) {
    set_error("tp_band_conj_transpose: need distinct input and output buffers");
    return TP_ERR_INVALID;
  }
  if (nx < 2 || ny < 2 || steps < 1 || !band_w_supported(w) || w >= nx) {
    set_error("tp_band_conj_transpose: invalid sizes (nx=%d, ny=%d, steps=%d, w=%d)", nx, ny, steps, w);
    return TP_ERR_INVALID;
  }
  hipLaunchKernelGGL(band_conj_transpose_kernel, dim3((unsigned)(steps * ny)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)d_W, (float2*)d_WH, nx, ny, steps, w, reverse_slices);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_band_sigma(const void* d_C, int nx, int ny, int ncol, double lhs_sign, double rhs_sign, int conj_c,
                             int lhs_col0, int rhs_col0, int col_step, int steps, void* d_sigma, void* stream) {
  if (!d_C || !d_sigma || nx < 1 || ny < 1 || ncol < 1 || steps < 1) {
    set_error("tp_band_sigma: invalid arguments (nx=%d, ny=%d, ncol=%d, steps=%d)", nx, ny, ncol, steps);
    return TP_ERR_INVALID;
  }
  const int64_t N = (int64_t)nx * ny;
  const unsigned gx = (unsigned)std::min<int64_t>((N + 255) / 256, 1024);
  hipLaunchKernelGGL(band_sigma_kernel, dim3(gx, (unsigned)steps), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)d_C, N, ncol, lhs_sign, rhs_sign, conj_c, lhs_col0, rhs_col0, col_step,
                     (float2*)d_sigma);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_band_sweep_ex(const tp_band_scan* scans, int nscans, int probes_per_block, int schedule,
                                void* stream) {
  if (schedule != 0 && schedule != 1) {
    set_error("tp_band_sweep_ex: schedule must be 0 or 1 (got %d)", schedule);
    return TP_ERR_INVALID;
  }
  if (!scans || nscans < 1 || nscans > TP_MAX_SCANS) {
    set_error("tp_band_sweep: nscans must be in [1, %d] (got %d)", TP_MAX_SCANS, nscans);
    return TP_ERR_INVALID;
  }
  const int pb = probes_per_block;
  if (pb != 1 && pb != 2 && pb != 4) {
    set_error("tp_band_sweep: probes_per_block must be 1, 2 or 4 (got %d)", pb);
    return TP_ERR_INVALID;
  }
  BandSweepArgs a;
  memset(&a, 0, sizeof(a));
  const int w = scans[0].w;
  const bool sigma = scans[0].sigma != nullptr;
  bool rhs = false;
  int nt = 64, nblk = 0;
  for (int k = 0; k < nscans; ++k) {
    const tp_band_scan& d = scans[k];
    if (d.w != w || (d.sigma != nullptr) != sigma) {
      set_error("tp_band_sweep: scan %d: all scans of one launch must share the band width and the sigma form", k);
      return TP_ERR_INVALID;
    }
    if (!band_w_supported(d.w)) {
      set_error("tp_band_sweep: band half-width %d not built (4, 6, 8, 12, 16, 24, 32)", d.w);
      return TP_ERR_UNSUPPORTED;
    }
    if (d.nx < 2 || d.nx > 256 || d.ny < 2 || d.steps < 1 || d.probes < 0 || !d.W || !d.cpl || !d.u0) {
      set_error("tp_band_sweep: scan %d: need 2 <= nx <= 256, ny >= 2, steps >= 1, W, cpl and u0 (nx=%d, ny=%d)", k,
                d.nx, d.ny);
      return TP_ERR_INVALID;
    }
    if (!d.out && !d.work) {
      set_error("tp_band_sweep: scan %d: without full-field output a work buffer of two planes per probe is required",
                k);
      return TP_ERR_INVALID;
    }
    if (d.ckpt && d.ckpt_every < 1) {
      set_error("tp_band_sweep: scan %d: ckpt_every must be >= 1", k);
      return TP_ERR_INVALID;
    }
    nt = std::max(nt, (d.nx + 63) / 64 * 64);
    rhs = rhs || d.rhs != nullptr;
    a.scans[k] = d;
    a.blk_start[k] = nblk;
    nblk += (d.probes + pb - 1) / pb;
  }
  a.blk_start[nscans] = nblk;
  a.nscans = nscans;
  a.nblk = nblk;
  if (nblk == 0) return TP_OK;
  hipStream_t st = (hipStream_t)stream;
#define TP_BAND_CASE(WV)                                                                           \
  case WV:                                                                                        \
    return sigma ? launch_band_sweep_pb<WV, true>(a, nt, rhs, pb, schedule, st)                     \
                 : launch_band_sweep_pb<WV, false>(a, nt, rhs, pb, schedule, st);
  switch (w) {
    TP_BAND_CASE(4)
    TP_BAND_CASE(6)
    TP_BAND_CASE(8)
    TP_BAND_CASE(12)
    TP_BAND_CASE(16)
    TP_BAND_CASE(24)
    TP_BAND_CASE(32)
  }
#undef TP_BAND_CASE
  return TP_ERR_UNSUPPORTED;
}

extern "C" int tp_band_sweep(const tp_band_scan* scans, int nscans, int probes_per_block, void* stream) {
  return tp_band_sweep_ex(scans, nscans, probes_per_block, 0, stream);
}
