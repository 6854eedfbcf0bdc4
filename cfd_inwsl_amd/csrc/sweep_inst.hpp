// Launch tables of the sweep kernels for one (D, AX): one instantiation per
// (interpolation scheme, Riemann flux, projection, tile shape), for the
// single-zone kernel and the persistent batched kernel.
#pragma once
#include "sweep.hpp"

namespace cfd {

// Tile shapes compiled per axis: x sweeps {0: 32x8, 1: 16x16},
// y/z sweeps {1: 16x16, 2: 8x32} (see Tile<>).
template <int AX> struct TilePair;
template <> struct TilePair<0> { static constexpr int A = 0, B = 1; };
template <> struct TilePair<1> { static constexpr int A = 1, B = 2; };
template <> struct TilePair<2> { static constexpr int A = 1, B = 2; };

template <int D, int AX, int SCH, int FLX, bool CW, int TL>
static hipError_t launch_one(bool fp32, const SweepArgs* a, const BatchArgs* b, dim3 grid, hipStream_t s) {
  if constexpr (SCH == SCH_WHF) {  // single-zone FP64 kernels only
    if (b || fp32) return hipErrorInvalidValue;
    hipLaunchKernelGGL((sweep_kernel<D, AX, SCH, FLX, CW, TL, double>), grid, dim3(256), 0, s, *a);
  } else if (b) {
    if (fp32) {
#if CFD_WITH_FP32
      hipLaunchKernelGGL((sweep_batch_kernel<D, AX, SCH, FLX, CW, TL, float>), grid, dim3(256), 0, s, *b);
#else
      return hipErrorInvalidValue;
#endif
    } else {
      hipLaunchKernelGGL((sweep_batch_kernel<D, AX, SCH, FLX, CW, TL, double>), grid, dim3(256), 0, s, *b);
    }
  } else if (fp32) {
#if CFD_WITH_FP32
    hipLaunchKernelGGL((sweep_kernel<D, AX, SCH, FLX, CW, TL, float>), grid, dim3(256), 0, s, *a);
#else
    return hipErrorInvalidValue;
#endif
  } else {
    hipLaunchKernelGGL((sweep_kernel<D, AX, SCH, FLX, CW, TL, double>), grid, dim3(256), 0, s, *a);
  }
  return hipGetLastError();
}

template <int D, int AX, int SCH, int FLX, bool CW>
static hipError_t launch_tile(int tile, bool fp32, const SweepArgs* a, const BatchArgs* b, dim3 grid, hipStream_t s) {
  if (tile == TilePair<AX>::A) return launch_one<D, AX, SCH, FLX, CW, TilePair<AX>::A>(fp32, a, b, grid, s);
  if (tile == TilePair<AX>::B) return launch_one<D, AX, SCH, FLX, CW, TilePair<AX>::B>(fp32, a, b, grid, s);
  return hipErrorInvalidValue;
}

template <int D, int AX, int SCH, int FLX>
static hipError_t launch_cw(bool cw, int tile, bool fp32, const SweepArgs* a, const BatchArgs* b, dim3 grid,
                            hipStream_t s) {
  return cw ? launch_tile<D, AX, SCH, FLX, true>(tile, fp32, a, b, grid, s)
            : launch_tile<D, AX, SCH, FLX, false>(tile, fp32, a, b, grid, s);
}

template <int D, int AX, int SCH>
static hipError_t launch_flx(int flx, bool cw, int tile, bool fp32, const SweepArgs* a, const BatchArgs* b, dim3 grid,
                             hipStream_t s) {
  return flx == FLX_HLLC ? launch_cw<D, AX, SCH, FLX_HLLC>(cw, tile, fp32, a, b, grid, s)
                         : launch_cw<D, AX, SCH, FLX_ROE>(cw, tile, fp32, a, b, grid, s);
}

template <int D, int AX>
static hipError_t launch_sch(int sch, int flx, bool cw, int tile, bool fp32, const SweepArgs* a, const BatchArgs* b,
                             dim3 grid, hipStream_t s) {
  switch (sch) {
  case SCH_JS: return launch_flx<D, AX, SCH_JS>(flx, cw, tile, fp32, a, b, grid, s);
  case SCH_Z: return launch_flx<D, AX, SCH_Z>(flx, cw, tile, fp32, a, b, grid, s);
  case SCH_TENO: return launch_flx<D, AX, SCH_TENO>(flx, cw, tile, fp32, a, b, grid, s);
  case SCH_TENOA: return launch_flx<D, AX, SCH_TENOA>(flx, cw, tile, fp32, a, b, grid, s);
  case SCH_TENOSA: return launch_flx<D, AX, SCH_TENOSA>(flx, cw, tile, fp32, a, b, grid, s);
  case SCH_WHF: return launch_flx<D, AX, SCH_WHF>(flx, cw, tile, fp32, a, b, grid, s);
  default: return launch_flx<D, AX, SCH_TENOS>(flx, cw, tile, fp32, a, b, grid, s);
  }
}

// Launch tables of the FMA-contracted FP64 kernels (single zones): the
// schemes of the benchmark configurations -- WENO-JS, WENO-Z, TENO5 (with its
// cut-off constants), TENO-S -- with HLLC or Roe, characteristic- or
// component-wise.  Instantiated only by the sweep_fma_d*.hip units.
template <int D, int AX, int SCH, int FLX, bool CW>
static hipError_t launch_fma_one(int tile, const SweepArgs* a, dim3 grid, hipStream_t s) {
  if (tile == TilePair<AX>::A)
    hipLaunchKernelGGL((sweep_kernel_fma<D, AX, SCH, FLX, CW, TilePair<AX>::A>), grid, dim3(256), 0, s, *a);
  else if (tile == TilePair<AX>::B)
    hipLaunchKernelGGL((sweep_kernel_fma<D, AX, SCH, FLX, CW, TilePair<AX>::B>), grid, dim3(256), 0, s, *a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
template <int D, int AX, int SCH>
static hipError_t launch_fma_flx(int flx, bool cw, int tile, const SweepArgs* a, dim3 grid, hipStream_t s) {
  if (flx == FLX_HLLC)
    return cw ? launch_fma_one<D, AX, SCH, FLX_HLLC, true>(tile, a, grid, s)
              : launch_fma_one<D, AX, SCH, FLX_HLLC, false>(tile, a, grid, s);
  return cw ? launch_fma_one<D, AX, SCH, FLX_ROE, true>(tile, a, grid, s)
            : launch_fma_one<D, AX, SCH, FLX_ROE, false>(tile, a, grid, s);
}
template <int D, int AX>
static hipError_t launch_fma_sch(int sch, int flx, bool cw, int tile, const SweepArgs* a, dim3 grid, hipStream_t s) {
  switch (sch) {
  case SCH_JS: return launch_fma_flx<D, AX, SCH_JS>(flx, cw, tile, a, grid, s);
  case SCH_Z: return launch_fma_flx<D, AX, SCH_Z>(flx, cw, tile, a, grid, s);
  case SCH_TENO: return launch_fma_flx<D, AX, SCH_TENO>(flx, cw, tile, a, grid, s);
  case SCH_TENOS: return launch_fma_flx<D, AX, SCH_TENOS>(flx, cw, tile, a, grid, s);
  default: return hipErrorInvalidValue;
  }
}
inline bool fma_scheme_supported(int sch) { return sch == SCH_JS || sch == SCH_Z || sch == SCH_TENO || sch == SCH_TENOS; }

// Entry points defined in sweep_d<D>_<axis>.hip
// fp32: FP32 storage / arithmetic kernels (not compiled for 1D).
hipError_t launch_sweep_dispatch(int D, int ax, int sch, int flx, bool cw, int tile, bool fp32, const SweepArgs* a,
                                 const BatchArgs* b, dim3 grid, hipStream_t s);
// the adjoint's recording 2D sweeps (sweep_rec.hip; SweepArgs::rec)
hipError_t launch_sweep_rec(int ax, int sch, int tile, const SweepArgs* a, dim3 grid, hipStream_t s);
// the FMA-contracted FP64 kernels (sweep_fma_d<D>_<axis>.hip)
hipError_t launch_sweep_fma(int D, int ax, int sch, int flx, bool cw, int tile, const SweepArgs* a, dim3 grid,
                            hipStream_t s);

}  // namespace cfd
