// Launch table of the sweep kernel for one dimension D: one instantiation per
// (axis, interpolation scheme, Riemann flux, projection).
#pragma once
#include "sweep.hpp"

namespace cfd {

template <int D, int AX, int SCH, int FLX, bool CW>
static hipError_t launch_one(const SweepArgs& a, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((sweep_kernel<D, AX, SCH, FLX, CW>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int D, int AX, int SCH, int FLX>
static hipError_t launch_cw(bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  return cw ? launch_one<D, AX, SCH, FLX, true>(a, grid, s) : launch_one<D, AX, SCH, FLX, false>(a, grid, s);
}

template <int D, int AX, int SCH>
static hipError_t launch_flx(int flx, bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  return flx == FLX_HLLC ? launch_cw<D, AX, SCH, FLX_HLLC>(cw, a, grid, s)
                         : launch_cw<D, AX, SCH, FLX_ROE>(cw, a, grid, s);
}

template <int D, int AX>
static hipError_t launch_sch(int sch, int flx, bool cw, const SweepArgs& a, dim3 grid, hipStream_t s) {
  switch (sch) {
  case SCH_JS: return launch_flx<D, AX, SCH_JS>(flx, cw, a, grid, s);
  case SCH_Z: return launch_flx<D, AX, SCH_Z>(flx, cw, a, grid, s);
  case SCH_TENO: return launch_flx<D, AX, SCH_TENO>(flx, cw, a, grid, s);
  default: return launch_flx<D, AX, SCH_TENOS>(flx, cw, a, grid, s);
  }
}

}  // namespace cfd
