// Sweep kernels for 2D Euler, direction 0.
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_d2_a0(int sch, int flx, bool cw, int tile, const SweepArgs* a, const BatchArgs* b, dim3 grid,
                          hipStream_t s) {
  return launch_sch<2, 0>(sch, flx, cw, tile, a, b, grid, s);
}
}  // namespace cfd
