// Sweep kernels for 3D Euler (nv = 5).
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_sweep_d3(int ax, int sch, int flx, bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  if (ax == 0) return launch_sch<3, 0>(sch, flx, cw, a, grid, s);
  if (ax == 1) return launch_sch<3, 1>(sch, flx, cw, a, grid, s);
  return launch_sch<3, 2>(sch, flx, cw, a, grid, s);
}
}  // namespace cfd
