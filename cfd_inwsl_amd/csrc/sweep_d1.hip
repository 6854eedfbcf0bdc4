// Sweep kernels for 1D Euler (nv = 3).
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_sweep_d1(int ax, int sch, int flx, bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  (void)ax;
  return launch_sch<1, 0>(sch, flx, cw, a, grid, s);
}
}  // namespace cfd
