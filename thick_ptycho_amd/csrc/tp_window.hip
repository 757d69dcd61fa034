// Gradient accumulation of reduced-domain (per-scan window) solves.
//
// Scan s steps through rows [a_s, a_s + W) of the object.  With its forward
// planes u^s_j (j = 0..L, full fields) and its exact-adjoint planes in stepping
// order (plane k holds lambda^s_{L-k+1}), the contribution of slice j to column
// x of the full-width accumulator is
//   G[j-1][x] = sum over scans whose window holds x of
//               conj(lambda^s_j[x - a_s]) (u^s_j[x - a_s] + u^s_{j-1}[x - a_s]).
// One thread per (slice, column) walks the scans, so overlapping windows sum
// without atomics and every output is written once; consecutive threads read
// consecutive rows of each scan's planes.
#include "tp_common.h"

namespace tp {
namespace {

__global__ __launch_bounds__(256) void window_grad_kernel(const float2* __restrict__ u, const float2* __restrict__ lam,
                                                          const int* __restrict__ win, int S, int L, int W, int nx,
                                                          float2* __restrict__ G) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)L * nx) return;
  const int x = (int)(t % nx);
  const int j = (int)(t / nx) + 1;  // slice 1..L
  float2 acc = make_float2(0.f, 0.f);
  for (int s = 0; s < S; ++s) {
    const int i = x - win[s];
    if (i < 0 || i >= W) continue;
    const int64_t base = (int64_t)s * (L + 1);
    const float2 uj = u[(base + j) * W + i];
    const float2 up = u[(base + j - 1) * W + i];
    const float2 l = lam[(base + (L - j + 1)) * W + i];
    acc = cfma(cconj(l), cadd(uj, up), acc);
  }
  G[(int64_t)(j - 1) * nx + x] = acc;
}

}  // namespace
}  // namespace tp

using namespace tp;

extern "C" int tp_window_grad_accumulate(const void* d_u, const void* d_lam, const int* d_win, int S, int L, int W,
                                         int nx, void* d_G, void* stream) {
  if (!d_u || !d_lam || !d_win || !d_G || S < 1 || L < 1 || W < 1 || nx < W) {
    set_error("tp_window_grad_accumulate: invalid arguments (S = %d, L = %d, W = %d, nx = %d)", S, L, W, nx);
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)L * nx;
  hipLaunchKernelGGL(window_grad_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)d_u, (const float2*)d_lam, d_win, S, L, W, nx, (float2*)d_G);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
