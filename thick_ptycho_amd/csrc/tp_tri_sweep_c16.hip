// Persistent sweep kernels for chunks of 16 rows per lane.
#include "tp_tri_sweep.h"

namespace tp {

int tri_sweep_launch_c16(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st) {
  return sweep_sigma<16>(a, ppb, sigma, grad, rhs, st);
}

}  // namespace tp
