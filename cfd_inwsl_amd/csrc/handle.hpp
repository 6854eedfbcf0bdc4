// Internal state of a cfd_handle, shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/congfd.h"

struct AdjointStore;
namespace cfd {
struct SweepAux;
}

struct cfd_handle {
  cfd_config cfg;
  int D, NV;
  long N;
  int sch, flx;
  bool cw;
  double ct;
  double* U0;            // FP64 state at the start of the step
  void *Ua, *Ub, *R;     // stage states and rhs: element type of the sweeps
  float* U0f;            // FP32 shadow of U0 (FP32 precision only)
  bool fp32;
  bool fma;              // CFD_PRECISION_FP64_FMA: the FMA-contracted FP64 sweep kernels
  size_t esz;            // bytes per element of Ua / Ub / R
  // device scalars: [0] dt, [1..3] max wave speed per direction (bits), [4] time
  // elapsed in the current call, [5] max |rhs| (bits), [6] absolute time,
  // [7] time at the start of the step; cfd_steps_to_output: [8] time since the
  // last output, [9] output reached, [10] halt the queued steps, [11] steps done
  double* scal;
  const double* halt;    // scal + 10 while cfd_steps_to_output queues steps, else null
  double* steplog;       // (t, dt, rhoLinf) per step of cfd_steps_to_output
  int steplog_cap;
  cfd::SweepAux* aux;    // device: per direction, the sweeps' rarely used arguments
  bool fuse_wave;        // stage 3's last sweep writes the next step's CFL maxima per workgroup
  double* wpart;         // those maxima: D x wpart_blocks
  long wpart_cap;        // doubles allocated
  long wpart_blocks;     // workgroups of the last launch that wrote them
  double* wave_scratch;  // non-null during cfd_rk3_steps_cfl_noalloc: the maxima go there instead
  double* partial;       // block partial sums for totals
  int npartial;
  int keep_rhs;
  int tile[3];     // tile shape per direction
  int seg[3];      // segment length per direction
  AdjointStore* adj;     // checkpoints and work buffers of the adjoint (lazily allocated)
  hipStream_t stream;
};

int cfd_fail(int code, const std::string& msg);
// rec (2D FP64 char-wise HLLC only): the direction's interface record, written
// by the recording sweep (SweepArgs::rec); rec[ax] for rk_stage
hipError_t launch_sweep(cfd_handle* h, int ax, const void* U, int flags, int stage, const double* U0, void* Uout,
                        int tile, int seg, void* Ushadow, void* rec = nullptr);
int rk_stage(cfd_handle* h, int stage, const void* Ucur, const double* U0, void* Uout, void* Ushadow,
             void* const* rec = nullptr);
const void* sweep_state(cfd_handle* h);
void cfd_adjoint_release(cfd_handle* h);
