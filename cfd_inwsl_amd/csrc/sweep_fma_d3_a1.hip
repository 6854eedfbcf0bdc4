// FMA-contracted FP64 sweep kernels for 3D Euler, direction 1 (built with
// -ffp-contract=fast: a*b + c of the path evaluated with one rounding; one
// Newton step fewer in divisions and square roots).
#define CFD_WITH_FP32 0
#define CFD_FMA_MODE 1
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_fma_d3_a1(int sch, int flx, bool cw, int tile, const SweepArgs* a, dim3 grid, hipStream_t s) {
  return launch_fma_sch<3, 1>(sch, flx, cw, tile, a, grid, s);
}
}  // namespace cfd
