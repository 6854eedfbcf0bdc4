// Sweep kernels for 3D Euler, direction 2.
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_d3_a2(int sch, int flx, bool cw, int tile, const SweepArgs* a, const BatchArgs* b, dim3 grid,
                          hipStream_t s) {
  return launch_sch<3, 2>(sch, flx, cw, tile, a, b, grid, s);
}
}  // namespace cfd
