// Object filters of the one-axis reconstruction loop, on the complex128 object
// n (ny rows of nx values, row-major; for the (x, z) object ny = x, nx = z):
//   tp_object_polar:   A = |n|, Phi = arg(n)                 (FP64)
//   tp_tv_denoise:     Chambolle-type TV iterations on A with periodic
//                      differences and tau = 1/4, in two passes per iteration
//                      (divergence of the dual field, then the dual update)
//   tp_lowpass_weight: spectrum *= exp(-(qx^2 + qy^2) / (2 sigma^2)) with q the
//                      fftshift-ordered linspace(-1/2, 1/2, n, endpoint=False)
//   tp_object_join:    n = A exp(i Phi)
// Reference: ReconstructorPWE.tv_denoise_amp / gaussian_lowpass_phase /
// global_object_update (thick_ptycho/reconstruction/pwe_reconstructor.py:370-419).
#include <mutex>

#include "tp_common.h"

namespace tp {

__global__ void polar_kernel(const double2* __restrict__ n, int64_t total, double* __restrict__ A,
                             double* __restrict__ Phi) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double2 v = n[i];
  A[i] = hypot(v.x, v.y);
  Phi[i] = atan2(v.y, v.x);
}

__global__ void join_kernel(const double* __restrict__ A, const double* __restrict__ Phi, int64_t total,
                            double2* __restrict__ n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  double s, c;
  sincos(Phi[i], &s, &c);
  n[i] = make_double2(A[i] * c, A[i] * s);
}

// uxy = u - lam * (px[y, x-1] - px[y, x] + py[y-1, x] - py[y, x])  (np.roll by +1)
__global__ void tv_div_kernel(const double* __restrict__ u, const double* __restrict__ px,
                              const double* __restrict__ py, double lam, int ny, int nx, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)ny * nx) return;
  const int y = (int)(i / nx), x = (int)(i % nx);
  const int xm = x == 0 ? nx - 1 : x - 1, ym = y == 0 ? ny - 1 : y - 1;
  const double div = px[(int64_t)y * nx + xm] - px[i] + py[(int64_t)ym * nx + x] - py[i];
  out[i] = u[i] - lam * div;
}

// dx = uxy[y, x+1] - uxy[y, x], dy = uxy[y+1, x] - uxy[y, x]  (np.roll by -1)
__global__ void tv_dual_kernel(const double* __restrict__ uxy, double lam, double eps, int ny, int nx,
                               double* __restrict__ px, double* __restrict__ py) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)ny * nx) return;
  const int y = (int)(i / nx), x = (int)(i % nx);
  const int xp = x == nx - 1 ? 0 : x + 1, yp = y == ny - 1 ? 0 : y + 1;
  const double dx = uxy[(int64_t)y * nx + xp] - uxy[i];
  const double dy = uxy[(int64_t)yp * nx + x] - uxy[i];
  const double denom = 1.0 + sqrt(dx * dx + dy * dy + eps) / lam;
  px[i] = (px[i] + 0.25 * dx) / denom;
  py[i] = (py[i] + 0.25 * dy) / denom;
}

__global__ void lowpass_weight_kernel(float2* __restrict__ X, int ny, int nx, double sigma) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)ny * nx) return;
  const int y = (int)(i / nx), x = (int)(i % nx);
  // fftshift(W)[y, x] = W[(y - ny/2) mod ny, (x - nx/2) mod nx], q_j = -1/2 + j/n
  const int jy = (y - ny / 2 + ny) % ny, jx = (x - nx / 2 + nx) % nx;
  const double qy = -0.5 + (double)jy / ny, qx = -0.5 + (double)jx / nx;
  const double w = exp(-(qx * qx + qy * qy) / (2.0 * sigma * sigma));
  X[i] = make_float2((float)(X[i].x * w), (float)(X[i].y * w));
}

__global__ void fill_zero_kernel(double* __restrict__ x, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) x[i] = 0.0;
}

__global__ void copy_kernel(const double* __restrict__ src, double* __restrict__ dst, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) dst[i] = src[i];
}

static dim3 grid_for(int64_t total) { return dim3((unsigned)((total + 255) / 256)); }

}  // namespace tp

using namespace tp;

extern "C" int tp_object_polar(const void* d_n, int ny, int nx, void* d_A, void* d_Phi, void* stream) {
  if (!d_n || !d_A || !d_Phi || ny < 1 || nx < 1) {
    set_error("tp_object_polar: invalid arguments");
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)ny * nx;
  hipLaunchKernelGGL(polar_kernel, grid_for(total), dim3(256), 0, (hipStream_t)stream, (const double2*)d_n, total,
                     (double*)d_A, (double*)d_Phi);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_object_join(const void* d_A, const void* d_Phi, int ny, int nx, void* d_n, void* stream) {
  if (!d_n || !d_A || !d_Phi || ny < 1 || nx < 1) {
    set_error("tp_object_join: invalid arguments");
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)ny * nx;
  hipLaunchKernelGGL(join_kernel, grid_for(total), dim3(256), 0, (hipStream_t)stream, (const double*)d_A,
                     (const double*)d_Phi, total, (double2*)d_n);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

// The TV iterations on u with the dual field and the work plane in `work`
// (3 * ny * nx float64); shared by tp_tv_denoise and tp_tv_denoise_noalloc.
static int tv_run(double* u, int ny, int nx, double lam, int n_iter, double eps, double* work, hipStream_t st) {
  const int64_t total = (int64_t)ny * nx;
  double* px = work;
  double* py = px + total;
  double* uxy = py + total;
  const dim3 g = grid_for(total);
  // fill kernels rather than hipMemsetAsync / hipMemcpyAsync: the calls stay capturable into HIP graphs
  hipLaunchKernelGGL(fill_zero_kernel, grid_for(2 * total), dim3(256), 0, st, px, 2 * total);
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(tv_div_kernel, g, dim3(256), 0, st, (const double*)u, (const double*)px, (const double*)py,
                       lam, ny, nx, uxy);
    hipLaunchKernelGGL(tv_dual_kernel, g, dim3(256), 0, st, (const double*)uxy, lam, eps, ny, nx, px, py);
  }
  hipLaunchKernelGGL(tv_div_kernel, g, dim3(256), 0, st, (const double*)u, (const double*)px, (const double*)py, lam,
                     ny, nx, uxy);
  hipLaunchKernelGGL(copy_kernel, g, dim3(256), 0, st, (const double*)uxy, u, total);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}

extern "C" int tp_tv_denoise(void* d_u, int ny, int nx, double lam, int n_iter, double eps, void* d_work,
                             void* stream) {
  if (!d_u || !d_work || ny < 1 || nx < 1 || n_iter < 0 || !(lam > 0.0)) {
    set_error("tp_tv_denoise: invalid arguments (ny=%d, nx=%d, lam=%g, n_iter=%d)", ny, nx, lam, n_iter);
    return TP_ERR_INVALID;
  }
  return tv_run((double*)d_u, ny, nx, lam, n_iter, eps, (double*)d_work, (hipStream_t)stream);
}

// Work space of tp_tv_denoise_noalloc: one copy per GPU, held until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host it is
// only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[3 * (size_t)TP_TV_NOALLOC_MAX_POINTS];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[TP_NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per device.
static int tv_scratch(double** out) {
  int dev = -1;
  TP_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= TP_NOALLOC_MAX_DEVICES) {
    set_error("tp_tv_denoise_noalloc: device %d beyond the %d this library tracks", dev, TP_NOALLOC_MAX_DEVICES);
    return TP_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    TP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error("tp_tv_denoise_noalloc: no device address for the scratch array on device %d", dev);
      return TP_ERR_HIP;
    }
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return TP_OK;
}

extern "C" int tp_tv_denoise_noalloc(void* d_u, int ny, int nx, double lam, int n_iter, double eps, void* stream) {
  if (!d_u || ny < 1 || nx < 1 || n_iter < 0 || !(lam > 0.0)) {
    set_error("tp_tv_denoise_noalloc: invalid arguments (ny=%d, nx=%d, lam=%g, n_iter=%d)", ny, nx, lam, n_iter);
    return TP_ERR_INVALID;
  }
  if ((int64_t)ny * nx > (int64_t)TP_TV_NOALLOC_MAX_POINTS) {
    set_error("tp_tv_denoise_noalloc: %d x %d = %lld points exceed the built-in work space of %d points; "
              "use tp_tv_denoise with a caller work buffer",
              ny, nx, (long long)ny * nx, TP_TV_NOALLOC_MAX_POINTS);
    return TP_ERR_INVALID;
  }
  double* work = nullptr;
  const int rc = tv_scratch(&work);
  if (rc != TP_OK) return rc;
  return tv_run((double*)d_u, ny, nx, lam, n_iter, eps, work, (hipStream_t)stream);
}

extern "C" int tp_lowpass_weight(void* d_X, int ny, int nx, double sigma, void* stream) {
  if (!d_X || ny < 1 || nx < 1 || !(sigma > 0.0)) {
    set_error("tp_lowpass_weight: invalid arguments");
    return TP_ERR_INVALID;
  }
  const int64_t total = (int64_t)ny * nx;
  hipLaunchKernelGGL(lowpass_weight_kernel, grid_for(total), dim3(256), 0, (hipStream_t)stream, (float2*)d_X, ny, nx,
                     sigma);
  TP_CHECK_HIP(hipGetLastError());
  return TP_OK;
}
