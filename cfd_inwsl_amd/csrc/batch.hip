// Batches of independent zones of uneven sizes (C ABI: cfd_batch_*).
//
// All zones share pooled state buffers (each zone keeps the single-block
// structure-of-arrays layout at its own offset).  One RK3 step is:
//   zero the per-zone wave-speed maxima and work-queue heads,
//   batched wave-speed reduction -> per-zone dt,
//   3 stages x D directions of the persistent sweep kernel: each launch
//   drains a work list covering every zone (largest zones first),
// captured once as a hipGraph and replayed per step (CFD_BATCH_GRAPH), or
// launched eagerly (CFD_BATCH_PERSISTENT), or, as a baseline, with the
// single-zone kernels launched zone after zone (CFD_BATCH_PER_ZONE).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/congfd.h"
#include "launch_plan.hpp"
#include "sweep_inst.hpp"

using namespace cfd;

// error slot and scheme table of the single-zone entry points (capi.hip)
int cfd_set_error_from_batch(int code, const char* msg);
int cfd_select_scheme(int inter, int dim, int* sch, double* ct);
int cfd_check_faces(int dim, const int* n, const int* bc, std::string* msg);

// Time bookkeeping of BlockSolver::RK3_SSP + solve (blockSolver.cpp:62,74,86,
// 123), as the single-zone clock: t + dt - dt/2 - dt/2 + dt, operation by operation.
__device__ __forceinline__ double zone_step_time(double t, double dt) { return ((t + dt) - dt / 2 - dt / 2) + dt; }

#define BCHK(expr)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return cfd_set_error_from_batch(CFD_E_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

struct cfd_batch {
  cfd_config common;
  int D, NV, nz;
  int sch, flx;
  bool cw;
  double ct;
  std::vector<ZoneDesc> hz;       // host copy of the zone table
  long total;                     // pooled elements per buffer (sum nv*N)
  // CFD_PRECISION_FP32: the sweeps read and write FP32 pools (Ua, Ub, R and
  // the shadow U0f of U0); U0 and the stage combination stay FP64.  Ub is
  // FP64-sized either way: it doubles as the host-transfer staging buffer.
  bool fp32;
  size_t esz;                     // bytes per element of Ua, R (sizeof(double) or sizeof(float))
  double *U0, *Ub;
  void *Ua, *R;
  float* U0f;
  ZoneDesc* dz;
  double* zscal;                  // per zone: [0] dt, [1..3] lambda bits, [4] t (stride 8)
  double* dt;                     // per-zone dt (contiguous, for the sweep kernels)
  WorkItem* items[3];
  int nitems[3];
  int tile[3];
  unsigned* counters;             // 3 stages x 3 directions
  int* wblk_zone;                 // wave-speed block -> zone
  long* wblk_start;               // wave-speed block -> first cell of the zone slab
  int nwblk;
  int persist_blocks;
  hipGraphExec_t graph;
  bool graph_ready;
  double graph_cfl;
  hipStream_t stream;
};

// ---- kernels -------------------------------------------------------------------
__global__ void kb_reset(double* zscal, int nz, unsigned* counters) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nz) {
    unsigned long long* b = reinterpret_cast<unsigned long long*>(zscal + 8 * i);
    b[1] = 0ull;
    b[2] = 0ull;
    b[3] = 0ull;
  }
  if (i < 9) counters[i] = 0u;
}

constexpr int WS_CELLS = 2048;  // cells per wave-speed block

template <int D>
__global__ __launch_bounds__(256) void kb_wave_speed(const double* __restrict__ U, const ZoneDesc* zones,
                                                      const int* wblk_zone, const long* wblk_start, double* zscal) {
  constexpr int NV = D + 2;
  const int z = wblk_zone[blockIdx.x];
  const ZoneDesc& zd = zones[z];
  const long c0 = wblk_start[blockIdx.x];
  const long c1 = min(c0 + (long)WS_CELLS, zd.N);
  double m[3] = {0.0, 0.0, 0.0};
  for (long i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
    double c[NV], p[NV];
    for (int v = 0; v < NV; v++) c[v] = U[zd.off + v * zd.N + i];
    cons_to_prim<D>(c, p);
    double cs = sqrt(CFD_GAMMA * p[NV - 1] / p[0]);
    for (int d = 0; d < D; d++) {
      double l = (cs + fabs(p[1 + d])) / zd.h[d];
      m[d] = m[d] < l ? l : m[d];
    }
  }
  __shared__ double red[3][256];
  for (int d = 0; d < D; d++) red[d][threadIdx.x] = m[d];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int d = 0; d < D; d++)
        if (red[d][threadIdx.x] < red[d][threadIdx.x + s]) red[d][threadIdx.x] = red[d][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    unsigned long long* b = reinterpret_cast<unsigned long long*>(zscal + 8 * z);
    for (int d = 0; d < D; d++) atomicMax(&b[1 + d], (unsigned long long)__double_as_longlong(red[d][0]));
  }
}

// dt = cfl / ((l0 + l1) + l2) per zone (BlockSolver::getTimeIntervalExplicit)
template <int D> __global__ void kb_finish_dt(double* zscal, double* dt, int nz, double cfl) {
  int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nz) return;
  double lam = 0;
  for (int d = 0; d < D; d++) lam += zscal[8 * z + 1 + d];
  double v = cfl / lam;
  zscal[8 * z] = v;
  zscal[8 * z + 4] = zone_step_time(zscal[8 * z + 4], v);
  dt[z] = v;
}

__global__ void kb_aos_to_soa(const double* __restrict__ in, double* __restrict__ out, long N, int NV) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    for (int v = 0; v < NV; v++) out[v * N + i] = in[i * NV + v];
}
__global__ void kb_to_f32(const double* __restrict__ in, float* __restrict__ out, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}
__global__ void kb_soa_to_aos(const double* __restrict__ in, double* __restrict__ out, long N, int NV) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    for (int v = 0; v < NV; v++) out[i * NV + v] = in[v * N + i];
}

// ---- host helpers ------------------------------------------------------------------
static int build_items(cfd_batch* b) {
  for (int ax = 0; ax < b->D; ax++) {
    std::vector<int> order(b->nz);
    for (int z = 0; z < b->nz; z++) order[z] = z;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->hz[x].N > b->hz[y].N; });
    std::vector<WorkItem> items;
    for (int z : order) {
      const ZoneDesc& zd = b->hz[z];
      const long nb = plan_pencil_blocks(zd.n, ax, b->tile[ax]);
      const int seg = zd.seg_len[ax];
      const int nseg = (zd.n[ax] + seg - 1) / seg;
      for (int s = 0; s < nseg; s++)
        for (long x = 0; x < nb; x++) items.push_back({z, (int)x, s});
    }
    if (b->items[ax]) (void)hipFree(b->items[ax]);
    BCHK(hipMalloc(&b->items[ax], sizeof(WorkItem) * items.size()));
    BCHK(hipMemcpy(b->items[ax], items.data(), sizeof(WorkItem) * items.size(), hipMemcpyHostToDevice));
    b->nitems[ax] = (int)items.size();
  }
  return 0;
}

// Segment lengths of every zone for the current tile choice (fewest lane
// slots; in a batch the zones together fill the chip, so no occupancy term),
// then the zone table and work lists.
static int plan_zones(cfd_batch* b) {
  for (auto& zd : b->hz)
    for (int ax = 0; ax < b->D; ax++) {
      const int ca = tile_ca(b->tile[ax]), len = zd.n[ax];
      auto slots = [&](int seg) {
        long chunks = 0;
        for (int c0 = 0; c0 < len; c0 += seg) chunks += ((len - c0 < seg ? len - c0 : seg) + 3 + ca - 1) / ca;
        return chunks;
      };
      int best = len;
      long bc = slots(len);
      for (int k = 1; k * ca - 3 < len; k++) {
        int seg = k * ca - 3;
        if (seg < 16) continue;
        long c = slots(seg);
        if (c < bc) { bc = c; best = seg; }
      }
      zd.seg_len[ax] = best;
    }
  BCHK(hipMemcpy(b->dz, b->hz.data(), sizeof(ZoneDesc) * b->nz, hipMemcpyHostToDevice));
  if (b->graph_ready) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph_ready = false;
  }
  return build_items(b);
}

static int batch_step_persistent(cfd_batch* b, double cfl) {
  hipLaunchKernelGGL(kb_reset, dim3((b->nz + 255) / 256 + 1), dim3(256), 0, b->stream, b->zscal, b->nz, b->counters);
  switch (b->D) {
  case 1:
    hipLaunchKernelGGL(kb_wave_speed<1>, dim3(b->nwblk), dim3(256), 0, b->stream, b->U0, b->dz, b->wblk_zone, b->wblk_start, b->zscal);
    hipLaunchKernelGGL(kb_finish_dt<1>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl);
    break;
  case 2:
    hipLaunchKernelGGL(kb_wave_speed<2>, dim3(b->nwblk), dim3(256), 0, b->stream, b->U0, b->dz, b->wblk_zone, b->wblk_start, b->zscal);
    hipLaunchKernelGGL(kb_finish_dt<2>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl);
    break;
  default:
    hipLaunchKernelGGL(kb_wave_speed<3>, dim3(b->nwblk), dim3(256), 0, b->stream, b->U0, b->dz, b->wblk_zone, b->wblk_start, b->zscal);
    hipLaunchKernelGGL(kb_finish_dt<3>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl);
    break;
  }
  BCHK(hipGetLastError());
  const void* ucur[3] = {b->fp32 ? (const void*)b->U0f : (const void*)b->U0, b->Ua, b->Ub};
  void* uout[3] = {b->Ua, b->Ub, b->U0};
  for (int st = 0; st < 3; st++)
    for (int ax = 0; ax < b->D; ax++) {
      BatchArgs a;
      std::memset(&a, 0, sizeof(a));
      a.U = ucur[st];
      a.R = b->R;
      a.U0 = b->U0;
      a.Uout = uout[st];
      a.Ushadow = (b->fp32 && st == 2) ? (void*)b->U0f : nullptr;
      a.dt = b->dt;
      a.zones = b->dz;
      a.items = b->items[ax];
      a.nitems = b->nitems[ax];
      a.counter = b->counters + 3 * st + ax;
      a.flags = (ax > 0 ? SW_READ_R : 0) | (ax < b->D - 1 ? SW_WRITE_R : SW_RK);
      a.stage = st + 1;
      a.ct = b->ct;
      int grid = std::min(b->persist_blocks, b->nitems[ax]);
      BCHK(launch_sweep_dispatch(b->D, ax, b->sch, b->flx, b->cw, b->tile[ax], b->fp32, nullptr, &a, dim3(grid), b->stream));
    }
  return 0;
}

// baseline: the single-zone kernels, zone after zone
static int batch_step_per_zone(cfd_batch* b, double cfl) {
  hipLaunchKernelGGL(kb_reset, dim3((b->nz + 255) / 256 + 1), dim3(256), 0, b->stream, b->zscal, b->nz, b->counters);
  BCHK(hipGetLastError());
  // the wave-speed table is laid out zone after zone; launch each zone's slice
  {
    std::vector<int> zfirst(b->nz + 1, 0);
    long acc = 0;
    for (int z = 0; z < b->nz; z++) {
      zfirst[z] = (int)acc;
      acc += (b->hz[z].N + WS_CELLS - 1) / WS_CELLS;
    }
    zfirst[b->nz] = (int)acc;
    for (int z = 0; z < b->nz; z++) {
      int nblk = zfirst[z + 1] - zfirst[z];
      const int* zt = b->wblk_zone + zfirst[z];
      const long* zs = b->wblk_start + zfirst[z];
      switch (b->D) {
      case 1: hipLaunchKernelGGL(kb_wave_speed<1>, dim3(nblk), dim3(256), 0, b->stream, b->U0, b->dz, zt, zs, b->zscal); break;
      case 2: hipLaunchKernelGGL(kb_wave_speed<2>, dim3(nblk), dim3(256), 0, b->stream, b->U0, b->dz, zt, zs, b->zscal); break;
      default: hipLaunchKernelGGL(kb_wave_speed<3>, dim3(nblk), dim3(256), 0, b->stream, b->U0, b->dz, zt, zs, b->zscal); break;
      }
    }
  }
  switch (b->D) {
  case 1: hipLaunchKernelGGL(kb_finish_dt<1>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl); break;
  case 2: hipLaunchKernelGGL(kb_finish_dt<2>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl); break;
  default: hipLaunchKernelGGL(kb_finish_dt<3>, dim3((b->nz + 63) / 64), dim3(64), 0, b->stream, b->zscal, b->dt, b->nz, cfl); break;
  }
  BCHK(hipGetLastError());
  // element offsets into pools of the sweeps' type (T) and of FP64
  auto at = [&](const void* pool, bool t_typed, long off) -> void* {
    const size_t e = t_typed ? b->esz : sizeof(double);
    return static_cast<char*>(const_cast<void*>(pool)) + e * off;
  };
  const void* ucur[3] = {b->fp32 ? (const void*)b->U0f : (const void*)b->U0, b->Ua, b->Ub};
  void* uout[3] = {b->Ua, b->Ub, b->U0};
  for (int z = 0; z < b->nz; z++) {
    const ZoneDesc& zd = b->hz[z];
    for (int st = 0; st < 3; st++)
      for (int ax = 0; ax < b->D; ax++) {
        SweepArgs a;
        std::memset(&a, 0, sizeof(a));
        a.U = at(ucur[st], true, zd.off);
        a.R = at(b->R, true, zd.off);
        a.U0 = b->U0 + zd.off;
        a.Uout = at(uout[st], st != 2, zd.off);
        a.Ushadow = (b->fp32 && st == 2) ? at(b->U0f, true, zd.off) : nullptr;
        a.dtp = b->dt + z;
        a.N = zd.N;
        for (int i = 0; i < 3; i++) a.n[i] = zd.n[i];
        a.flags = (ax > 0 ? SW_READ_R : 0) | (ax < b->D - 1 ? SW_WRITE_R : SW_RK);
        a.stage = st + 1;
        a.seg_len = zd.seg_len[ax];
        a.h = zd.h[ax];
        a.ct = b->ct;
        a.bc_lo = zd.bc[2 * ax];
        a.bc_hi = zd.bc[2 * ax + 1];
        a.aux = &b->dz[z].aux[ax];
        long nb = plan_pencil_blocks(zd.n, ax, b->tile[ax]);
        int nseg = (zd.n[ax] + zd.seg_len[ax] - 1) / zd.seg_len[ax];
        BCHK(launch_sweep_dispatch(b->D, ax, b->sch, b->flx, b->cw, b->tile[ax], b->fp32, &a, nullptr,
                                   dim3((unsigned)nb, (unsigned)nseg), b->stream));
      }
  }
  return 0;
}

static int batch_capture(cfd_batch* b, double cfl) {
  if (b->graph_ready && b->graph_cfl == cfl) return 0;
  if (b->graph_ready) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph_ready = false;
  }
  hipGraph_t g;
  BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_step_persistent(b, cfl);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) return rc;
  BCHK(e);
  BCHK(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0));
  (void)hipGraphDestroy(g);
  b->graph_ready = true;
  b->graph_cfl = cfl;
  return 0;
}

extern "C" {

int cfd_batch_create(const cfd_config* c, int nzones, const cfd_zone* zones, cfd_batch** out) {
  if (!c || !zones || !out || nzones < 1) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  *out = nullptr;
  // validate the common part through the single-zone path's rules
  if (c->dim < 1 || c->dim > 3) return cfd_set_error_from_batch(CFD_E_ARG, "dim must be 1, 2 or 3");
  if (c->diff_method != CFD_DIFF_MND6) return cfd_set_error_from_batch(CFD_E_UNSUPP, "only DiffMethod MND6 (3) is implemented");
  if (c->flux_method != CFD_FLUX_HLLC && c->flux_method != CFD_FLUX_ROE)
    return cfd_set_error_from_batch(CFD_E_ARG, "flux_method must be HLLC (0) or ROE (1)");
  if (c->recon != CFD_RECON_CHAR && c->recon != CFD_RECON_COMPONENT)
    return cfd_set_error_from_batch(CFD_E_ARG, "recon must be CFD_RECON_CHAR or CFD_RECON_COMPONENT");
  int sch;
  double ct;
  if (c->precision != CFD_PRECISION_FP64 && c->precision != CFD_PRECISION_FP32)
    return cfd_set_error_from_batch(CFD_E_UNSUPP, "batches run in CFD_PRECISION_FP64 or CFD_PRECISION_FP32");
  if (c->gravity[0] != 0.0 || c->gravity[1] != 0.0 || c->gravity[2] != 0.0)
    return cfd_set_error_from_batch(CFD_E_UNSUPP, "batches take no source term");
  if (cfd_select_scheme(c->inter_method, c->dim, &sch, &ct))
    return cfd_set_error_from_batch(CFD_E_UNSUPP, "inter_method is not implemented");
  if (sch == SCH_WHF)
    return cfd_set_error_from_batch(CFD_E_UNSUPP, "this inter_method runs on single zones only");
  cfd_batch* b = new cfd_batch();
  b->common = *c;
  b->D = c->dim;
  b->NV = c->dim + 2;
  b->nz = nzones;
  b->sch = sch;
  b->ct = ct;
  b->flx = c->flux_method;
  b->cw = c->recon == CFD_RECON_CHAR;
  b->fp32 = c->precision == CFD_PRECISION_FP32;
  b->esz = b->fp32 ? sizeof(float) : sizeof(double);
  b->tile[0] = 0;
  b->tile[1] = 1;
  b->tile[2] = 1;
  b->persist_blocks = 256 * 4;
  long off = 0;
  for (int z = 0; z < nzones; z++) {
    ZoneDesc zd;
    std::memset(&zd, 0, sizeof(zd));
    for (int d = 0; d < 3; d++) {
      zd.n[d] = zones[z].n[d];
      zd.h[d] = zones[z].h[d];
      if (zd.n[d] < 1 || (d >= c->dim && zd.n[d] != 1) || (d < c->dim && !(zd.h[d] > 0))) {
        delete b;
        return cfd_set_error_from_batch(CFD_E_ARG, ("zone " + std::to_string(z) + ": bad n/h").c_str());
      }
      zd.seg_len[d] = zd.n[d];
    }
    {
      std::string msg;
      if (cfd_check_faces(c->dim, zd.n, zones[z].bc, &msg)) {
        delete b;
        return cfd_set_error_from_batch(CFD_E_ARG, ("zone " + std::to_string(z) + ": " + msg).c_str());
      }
    }
    for (int d = 0; d < 3; d++) fill_research_consts(zd.aux[d].rc);
    for (int f = 0; f < 6; f++) {
      if (zones[z].bc[f] == BC_DMR_TOP || zones[z].bc[f] == BC_DMR_WALL) {
        delete b;
        return cfd_set_error_from_batch(CFD_E_UNSUPP, "batches take no time- or position-dependent boundaries");
      }
      zd.bc[f] = zones[z].bc[f];
      for (int v = 0; v < 5; v++) {
        if (f % 2 == 0) zd.aux[f / 2].prim_lo[v] = zones[z].bc_prim[f][v];
        else zd.aux[f / 2].prim_hi[v] = zones[z].bc_prim[f][v];
      }
    }
    zd.N = (long)zd.n[0] * zd.n[1] * zd.n[2];
    zd.off = off;
    off += zd.N * b->NV;
    b->hz.push_back(zd);
  }
  b->total = off;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipMalloc(&b->U0, sizeof(double) * off);
  if (e == hipSuccess) e = hipMalloc(&b->Ua, b->esz * off);
  if (e == hipSuccess) e = hipMalloc(&b->Ub, sizeof(double) * off);
  if (e == hipSuccess) e = hipMalloc(&b->R, b->esz * off);
  if (e == hipSuccess && b->fp32) e = hipMalloc(&b->U0f, sizeof(float) * off);
  if (e == hipSuccess) e = hipMalloc(&b->dz, sizeof(ZoneDesc) * nzones);
  if (e == hipSuccess) e = hipMalloc(&b->zscal, sizeof(double) * 8 * nzones);
  if (e == hipSuccess) e = hipMalloc(&b->dt, sizeof(double) * nzones);
  if (e == hipSuccess) e = hipMalloc(&b->counters, sizeof(unsigned) * 9);
  if (e == hipSuccess) e = hipStreamCreate(&b->stream);
  // wave-speed block table (zone after zone)
  std::vector<int> wz;
  std::vector<long> ws;
  for (int z = 0; z < nzones; z++)
    for (long s0 = 0; s0 < b->hz[z].N; s0 += WS_CELLS) {
      wz.push_back(z);
      ws.push_back(s0);
    }
  b->nwblk = (int)wz.size();
  if (e == hipSuccess) e = hipMalloc(&b->wblk_zone, sizeof(int) * wz.size());
  if (e == hipSuccess) e = hipMalloc(&b->wblk_start, sizeof(long) * ws.size());
  if (e == hipSuccess) e = hipMemcpy(b->wblk_zone, wz.data(), sizeof(int) * wz.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->wblk_start, ws.data(), sizeof(long) * ws.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->zscal, 0, sizeof(double) * 8 * nzones);
  if (e != hipSuccess) {
    std::string m = std::string("batch allocation: ") + hipGetErrorString(e);
    cfd_batch_destroy(b);
    return cfd_set_error_from_batch(CFD_E_HIP, m.c_str());
  }
  int rc = plan_zones(b);
  if (rc) {
    cfd_batch_destroy(b);
    return rc;
  }
  *out = b;
  return CFD_OK;
}

int cfd_batch_destroy(cfd_batch* b) {
  if (!b) return CFD_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph_ready) (void)hipGraphExecDestroy(b->graph);
  (void)hipFree(b->U0);
  (void)hipFree(b->Ua);
  (void)hipFree(b->Ub);
  (void)hipFree(b->R);
  if (b->U0f) (void)hipFree(b->U0f);
  (void)hipFree(b->dz);
  (void)hipFree(b->zscal);
  (void)hipFree(b->dt);
  (void)hipFree(b->counters);
  (void)hipFree(b->wblk_zone);
  (void)hipFree(b->wblk_start);
  for (int ax = 0; ax < 3; ax++)
    if (b->items[ax]) (void)hipFree(b->items[ax]);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CFD_OK;
}

int cfd_batch_set_state(cfd_batch* b, int zone, const double* cons) {
  if (!b || !cons || zone < 0 || zone >= b->nz) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  const ZoneDesc& zd = b->hz[zone];
  BCHK(hipMemcpyAsync(b->Ub + zd.off, cons, sizeof(double) * zd.N * b->NV, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(kb_aos_to_soa, dim3((unsigned)((zd.N + 255) / 256)), dim3(256), 0, b->stream, b->Ub + zd.off,
                     b->U0 + zd.off, zd.N, b->NV);
  if (b->fp32)
    hipLaunchKernelGGL(kb_to_f32, dim3((unsigned)((zd.N * b->NV + 255) / 256)), dim3(256), 0, b->stream,
                       b->U0 + zd.off, b->U0f + zd.off, zd.N * b->NV);
  BCHK(hipGetLastError());
  BCHK(hipStreamSynchronize(b->stream));
  return CFD_OK;
}

int cfd_batch_get_state(cfd_batch* b, int zone, double* cons) {
  if (!b || !cons || zone < 0 || zone >= b->nz) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  const ZoneDesc& zd = b->hz[zone];
  hipLaunchKernelGGL(kb_soa_to_aos, dim3((unsigned)((zd.N + 255) / 256)), dim3(256), 0, b->stream, b->U0 + zd.off,
                     b->Ub + zd.off, zd.N, b->NV);
  BCHK(hipGetLastError());
  BCHK(hipMemcpyAsync(cons, b->Ub + zd.off, sizeof(double) * zd.N * b->NV, hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  return CFD_OK;
}

int cfd_batch_steps_cfl(cfd_batch* b, double cfl, int nsteps, int mode) {
  if (!b || !(cfl > 0) || nsteps < 0) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  int rc;
  if (mode == CFD_BATCH_GRAPH) {
    if ((rc = batch_capture(b, cfl))) return rc;
    for (int s = 0; s < nsteps; s++) BCHK(hipGraphLaunch(b->graph, b->stream));
    return CFD_OK;
  }
  for (int s = 0; s < nsteps; s++) {
    rc = mode == CFD_BATCH_PER_ZONE ? batch_step_per_zone(b, cfl) : batch_step_persistent(b, cfl);
    if (rc) return rc;
  }
  return CFD_OK;
}

int cfd_batch_times(cfd_batch* b, double* t_out) {
  if (!b || !t_out) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  std::vector<double> s((size_t)8 * b->nz);
  BCHK(hipMemcpyAsync(s.data(), b->zscal, sizeof(double) * s.size(), hipMemcpyDeviceToHost, b->stream));
  BCHK(hipStreamSynchronize(b->stream));
  for (int z = 0; z < b->nz; z++) t_out[z] = s[8 * z + 4];
  return CFD_OK;
}

int cfd_batch_autotune(cfd_batch* b, int reps) {
  if (!b || reps < 1) return cfd_set_error_from_batch(CFD_E_ARG, "bad argument");
  // state and zone clocks are restored afterwards
  std::vector<double> save((size_t)b->total), zsave((size_t)8 * b->nz);
  BCHK(hipMemcpy(save.data(), b->U0, sizeof(double) * b->total, hipMemcpyDeviceToHost));
  BCHK(hipMemcpy(zsave.data(), b->zscal, sizeof(double) * zsave.size(), hipMemcpyDeviceToHost));
  hipEvent_t e0, e1;
  BCHK(hipEventCreate(&e0));
  BCHK(hipEventCreate(&e1));
  const int tiles_x[2] = {0, 1}, tiles_yz[2] = {1, 2};
  const int grids[3] = {512, 1024, 2048};
  double best = 1e30;
  int bt[3] = {b->tile[0], b->tile[1], b->tile[2]}, bg = b->persist_blocks;
  for (int tx = 0; tx < 2; tx++)
    for (int ty = 0; ty < (b->D > 1 ? 2 : 1); ty++)
      for (int g = 0; g < 3; g++) {
        b->tile[0] = tiles_x[tx];
        b->tile[1] = b->tile[2] = tiles_yz[ty];
        b->persist_blocks = grids[g];
        int rc = plan_zones(b);
        if (rc) return rc;
        if ((rc = batch_step_persistent(b, 0.5))) return rc;
        BCHK(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; r++)
          if ((rc = batch_step_persistent(b, 0.5))) return rc;
        BCHK(hipEventRecord(e1, b->stream));
        BCHK(hipEventSynchronize(e1));
        float ms = 0;
        BCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          bt[0] = b->tile[0];
          bt[1] = b->tile[1];
          bt[2] = b->tile[2];
          bg = b->persist_blocks;
        }
        BCHK(hipMemcpy(b->U0, save.data(), sizeof(double) * b->total, hipMemcpyHostToDevice));
      }
  if (b->fp32) {  // the shadow of the restored state
    hipLaunchKernelGGL(kb_to_f32, dim3((unsigned)((b->total + 255) / 256)), dim3(256), 0, b->stream, b->U0, b->U0f,
                       b->total);
    BCHK(hipGetLastError());
    BCHK(hipStreamSynchronize(b->stream));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int d = 0; d < 3; d++) b->tile[d] = bt[d];
  b->persist_blocks = bg;
  int rc = plan_zones(b);
  if (rc) return rc;
  BCHK(hipMemcpy(b->zscal, zsave.data(), sizeof(double) * zsave.size(), hipMemcpyHostToDevice));
  return CFD_OK;
}

int cfd_batch_synchronize(cfd_batch* b) {
  if (!b) return cfd_set_error_from_batch(CFD_E_ARG, "null batch");
  BCHK(hipStreamSynchronize(b->stream));
  return CFD_OK;
}

}  // extern "C"
