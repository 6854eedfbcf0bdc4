// Sweep kernels for 1D Euler, direction 0.
#define CFD_WITH_FP32 1
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_d1_a0(int sch, int flx, bool cw, int tile, bool fp32, const SweepArgs* a, const BatchArgs* b,
                          dim3 grid, hipStream_t s) {
  return launch_sch<1, 0>(sch, flx, cw, tile, fp32, a, b, grid, s);
}
}  // namespace cfd
