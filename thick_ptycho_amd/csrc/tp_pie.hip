// Multislice 3PIE object update (one scan position at a time).
//
// For one scan the reference back-propagates the corrected exit wave through
// z = nz-1 .. 0 as x <- conj(T_z) P^H x and updates slice z of the refractive
// index from dpsi_z = x - psi_z.  The transmission T_z of that step is taken
// before slice z is updated and later steps only read slices not yet updated,
// so the whole back-propagation can run first on the unchanged object (one
// multislice sweep with conj(H) and the per-step factors made here) and every
// slice's update afterwards in parallel:
//   pie_tadj_kernel:   Tadj[0] = 1, Tadj[j] = conj(T_{nz-j}), j = 1..nz-1, so the
//                      sweep's plane k is P^H applied to conj(T_{nz-k+1}) (plane k-1);
//   pie_update_kernel: one workgroup per slice z, x = conj(T_z) back[nz - z],
//                      n_z += step conj(psi_z) (x - psi_z) / (max |psi_z|^2 + eps) / (i k dz).
// n is complex128 (rows x nz, row stride ld, the window's first row at d_n);
// psi (nz planes) and back (nz + 1 planes) are complex64 with rows of N.
#include "tp_common.h"

namespace tp {
namespace {

__device__ __forceinline__ double2 transmission_d(double2 v, double k, double dz, double nmed) {
  const double amp = exp(-k * dz * v.y);
  double s, c;
  sincos(k * dz * (v.x - nmed), &s, &c);
  return make_double2(amp * c, amp * s);
}

__global__ __launch_bounds__(256) void pie_tadj_kernel(const double2* __restrict__ n, int64_t ld, int N, int nz,
                                                       double k, double dz, double nmed, float2* __restrict__ Tadj) {
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= (int64_t)N * nz) return;
  const int j = (int)(id / N), x = (int)(id % N);
  if (j == 0) {
    Tadj[id] = make_float2(1.f, 0.f);
    return;
  }
  const double2 t = transmission_d(n[(int64_t)x * ld + (nz - j)], k, dz, nmed);
  Tadj[id] = make_float2((float)t.x, (float)-t.y);
}

__global__ __launch_bounds__(256) void pie_update_kernel(double2* __restrict__ n, int64_t ld, int N, int nz,
                                                         const float2* __restrict__ psi,
                                                         const float2* __restrict__ back, double k, double dz,
                                                         double nmed, double step, double eps) {
  const int z = blockIdx.x, tid = threadIdx.x;
  const float2* pz = psi + (int64_t)z * N;
  const float2* bz = back + (int64_t)(nz - z) * N;
  float m = 0.f;
  for (int x = tid; x < N; x += 256) {
    const float2 p = pz[x];
    m = fmaxf(m, fmaf(p.x, p.x, p.y * p.y));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float wmax[4];
  if ((tid & 63) == 0) wmax[tid >> 6] = m;
  __syncthreads();
  m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  // step / (max + eps) / (i k dz) = -i step / ((max + eps) k dz)
  const double f = step / (((double)m + eps) * k * dz);
  for (int x = tid; x < N; x += 256) {
    double2* nv = n + (int64_t)x * ld + z;
    const double2 v = *nv;
    const double2 t = transmission_d(v, k, dz, nmed);
    const float2 b = bz[x];
    const float2 p = pz[x];
    // x_back = conj(T) b, dpsi = x_back - p, g = conj(p) dpsi
    const double bx = t.x * b.x + t.y * b.y, by = t.x * b.y - t.y * b.x;
    const double dx_ = bx - p.x, dy_ = by - p.y;
    const double gx = p.x * dx_ + p.y * dy_, gy = p.x * dy_ - p.y * dx_;
    *nv = make_double2(v.x + f * gy, v.y - f * gx);  // + (-i) f g
  }
}

}  // namespace
}  // namespace tp

using namespace tp;

extern "C" int tp_pie_adjoint_transmission(const void* d_n, int64_t ld, int N, int nz, double k, double dz,
                                           double n_medium, void* d_Tadj, void* stream) {
  if (!d_n || !d_Tadj || N < 1 || nz < 1 || ld < nz) {
    set_error("tp_pie_adjoint_transmission: invalid arguments (N = %d, nz = %d, ld = %lld)", N, nz, (long long)ld);
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)N * nz;
  hipLaunchKernelGGL(pie_tadj_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)d_n, ld, N, nz, k, dz, n_medium, (float2*)d_Tadj);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_pie_update(void* d_n, int64_t ld, int N, int nz, const void* d_psi, const void* d_back, double k,
                             double dz, double n_medium, double step, double eps, void* stream) {
  if (!d_n || !d_psi || !d_back || N < 1 || nz < 1 || ld < nz || !(k * dz != 0.0)) {
    set_error("tp_pie_update: invalid arguments (N = %d, nz = %d, ld = %lld)", N, nz, (long long)ld);
    return TP_ERR_INVALID;
  }
  hipLaunchKernelGGL(pie_update_kernel, dim3(nz), dim3(256), 0, (hipStream_t)stream, (double2*)d_n, ld, N, nz,
                     (const float2*)d_psi, (const float2*)d_back, k, dz, n_medium, step, eps);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
