// Recording sweeps of the adjoint's recomputed forward (sweep_kernel_rec):
// 2D FP64 characteristic-wise HLLC with WENO-JS or a TENO-family scheme, the
// tile shapes the planner uses (x: 32x8, 16x16; y: 16x16, 8x32).
#define CFD_WITH_FP32 0
#include "sweep_inst.hpp"

namespace cfd {
namespace {
template <int AX, int TL> hipError_t rec_sch(int sch, const SweepArgs* a, dim3 grid, hipStream_t s) {
  switch (sch) {
  case SCH_JS: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_JS, TL>), grid, dim3(256), 0, s, *a); break;
  case SCH_TENO: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_TENO, TL>), grid, dim3(256), 0, s, *a); break;
  case SCH_TENOS: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_TENOS, TL>), grid, dim3(256), 0, s, *a); break;
  case SCH_TENOA: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_TENOA, TL>), grid, dim3(256), 0, s, *a); break;
  case SCH_TENOSA: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_TENOSA, TL>), grid, dim3(256), 0, s, *a); break;
  case SCH_WHF: hipLaunchKernelGGL((sweep_kernel_rec<AX, SCH_WHF, TL>), grid, dim3(256), 0, s, *a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
}  // namespace

hipError_t launch_sweep_rec(int ax, int sch, int tile, const SweepArgs* a, dim3 grid, hipStream_t s) {
  if (ax == 0) {
    if (tile == 0) return rec_sch<0, 0>(sch, a, grid, s);
    if (tile == 1) return rec_sch<0, 1>(sch, a, grid, s);
  } else if (ax == 1) {
    if (tile == 1) return rec_sch<1, 1>(sch, a, grid, s);
    if (tile == 2) return rec_sch<1, 2>(sch, a, grid, s);
  }
  return hipErrorInvalidValue;
}
}  // namespace cfd
