// Run-time dispatch of the sweep launch tables.
#define CFD_WITH_FP32 0
#include "sweep_inst.hpp"

namespace cfd {
hipError_t launch_d1_a0(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);
hipError_t launch_d2_a0(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);
hipError_t launch_d2_a1(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);
hipError_t launch_d3_a0(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);
hipError_t launch_d3_a1(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);
hipError_t launch_d3_a2(int, int, bool, int, bool, const SweepArgs*, const BatchArgs*, dim3, hipStream_t);

hipError_t launch_fma_d1_a0(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);
hipError_t launch_fma_d2_a0(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);
hipError_t launch_fma_d2_a1(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);
hipError_t launch_fma_d3_a0(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);
hipError_t launch_fma_d3_a1(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);
hipError_t launch_fma_d3_a2(int, int, bool, int, const SweepArgs*, dim3, hipStream_t);

hipError_t launch_sweep_fma(int D, int ax, int sch, int flx, bool cw, int tile, const SweepArgs* a, dim3 grid,
                            hipStream_t s) {
  if (D == 1) return launch_fma_d1_a0(sch, flx, cw, tile, a, grid, s);
  if (D == 2) return ax == 0 ? launch_fma_d2_a0(sch, flx, cw, tile, a, grid, s) : launch_fma_d2_a1(sch, flx, cw, tile, a, grid, s);
  if (ax == 0) return launch_fma_d3_a0(sch, flx, cw, tile, a, grid, s);
  if (ax == 1) return launch_fma_d3_a1(sch, flx, cw, tile, a, grid, s);
  return launch_fma_d3_a2(sch, flx, cw, tile, a, grid, s);
}

hipError_t launch_sweep_dispatch(int D, int ax, int sch, int flx, bool cw, int tile, bool fp32, const SweepArgs* a,
                                 const BatchArgs* b, dim3 grid, hipStream_t s) {
  if (D == 1) return launch_d1_a0(sch, flx, cw, tile, fp32, a, b, grid, s);
  if (D == 2) return ax == 0 ? launch_d2_a0(sch, flx, cw, tile, fp32, a, b, grid, s) : launch_d2_a1(sch, flx, cw, tile, fp32, a, b, grid, s);
  if (ax == 0) return launch_d3_a0(sch, flx, cw, tile, fp32, a, b, grid, s);
  if (ax == 1) return launch_d3_a1(sch, flx, cw, tile, fp32, a, b, grid, s);
  return launch_d3_a2(sch, flx, cw, tile, fp32, a, b, grid, s);
}
}  // namespace cfd
