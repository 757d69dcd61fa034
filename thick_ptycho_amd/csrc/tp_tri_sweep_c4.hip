// Persistent sweep kernels for chunks of 4 rows per lane.
#include "tp_tri_sweep.h"

namespace tp {

int tri_sweep_launch_c4(const SweepArgs& a, int ppb, int sigma, bool grad, bool rhs, hipStream_t st) {
  return sweep_sigma<4>(a, ppb, sigma, grad, rhs, st);
}

}  // namespace tp
