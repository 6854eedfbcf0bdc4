// Sweep kernels for 2D Euler (nv = 4).
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_sweep_d2(int ax, int sch, int flx, bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  if (ax == 0) return launch_sch<2, 0>(sch, flx, cw, a, grid, s);
  return launch_sch<2, 1>(sch, flx, cw, a, grid, s);
}
}  // namespace cfd
