// Direction-sweep kernel: one pass of SpDistributor::rhsSolve along axis AX
// (zone_FD/src/src/sp_distributor.cpp:9-78) fused with the ghost fill
// (OneDBnd::update, zone_FD/src/src/oneDBnd.cpp:31-114), the per-stage
// consToPrim (zone_FD/src/src/equation.cpp:8-99), the MND6 flux derivative
// (MidNodeAndNodeDif6::solve, zone_FD/src/src/differ.cpp:54-103) and, on the
// last direction, the SSP-RK3 stage update (BlockSolver::RK3_SSP,
// zone_FD/src/src/blockSolver.cpp:43-87).
//
// Work decomposition (MI355X, wave64):
//   * a workgroup of 256 lanes owns CP pencils and a segment [c0, c1) of
//     their cells along AX, and marches along AX in chunks of CA interfaces;
//   * lane (a, p) loads cell m0+1+a of pencil p, converts it to primitive
//     variables once and keeps (prim, rebuilt cons) in an LDS ring buffer;
//   * lane (a, p) then computes interface m0+a (characteristic projection,
//     nonlinear interpolation, Riemann flux) from six ring cells, writes the
//     flux to an LDS flux ring, and finally updates cell m0-3+a.
//   For AX = 0 the lanes run along x (CA = 32, CP = 8), for AX = 1, 2 they run
//   across x-adjacent pencils (CA = 16, CP = 16): every global access is a
//   coalesced row segment.  A segment computes n+3 interfaces for n cells
//   (the reference's per-pencil count), so redundancy is 3 interfaces per
//   segment.
#pragma once
#include "euler_point.hpp"

namespace cfd {

enum BcCode {
  BC_PERIODIC = 1,
  BC_SYMMETRY1D = 2,
  BC_DIRICHLET = 3,
  BC_OUTFLOW = 7,
  BC_SYMMETRYX = 10,
  BC_SYMMETRYY = 11,
  BC_SYMMETRYZ = 101,
  BC_MIRROR_COPY = 102,
  BC_DMR_TOP = 8,     // DoubleMachUp (oneDBnd.cpp:87-108), y-high face of 2D
  BC_DMR_WALL = 103,  // double-Mach wall wiring (initializer.cpp:916-934), y-low face of 2D
};

// flags of a sweep launch
// SW_WAVE (last sweep of stage 3): the CFL maxima of the new state are
// written per workgroup to wave[d * nblocks + block], so the next step needs
// no separate pass over the state (a small kernel reduces the partials)
enum SweepFlags { SW_READ_R = 1, SW_WRITE_R = 2, SW_RK = 4, SW_SOURCE = 8, SW_WAVE = 16 };

// Buffers are typed by the kernel's compute type T (double or float): U, R
// and the stage outputs hold T; U0 (the state at the start of the step) is
// always FP64.  With T = float the stage-3 output is the FP64 state U0 plus
// its FP32 shadow `Ushadow` read by the next step's sweeps.
// Arguments used only at boundaries, by sources or by the fused CFL maxima,
// kept in device memory behind one pointer so that they do not occupy scalar
// registers for the whole kernel (the FP64 sweeps are at the SGPR limit).
struct SweepAux {
  double prim_lo[5], prim_hi[5];  // Dirichlet states of the two faces (primitive)
  // position- and time-dependent boundaries (2D y sweeps): cell centres at
  // x0 + (i + 0.5) hx in every direction (Initializer::initUniformBlock,
  // initializer.cpp:672-677)
  double x0, hx;
  double grav[3];  // SW_SOURCE: gravity (SourceTerm::calGravitySource)
  ResearchConsts rc;  // constants of the research TENO variants
};

struct SweepArgs {
  const void* U;     // state whose RHS is evaluated (SoA planes of N), T
  void* R;           // RHS accumulator (SoA), T
  const double* U0;  // RK: state at the start of the step (FP64)
  void* Uout;        // RK: stage output (T; FP64 at stage 3 when T = float)
  void* Ushadow;     // RK, T = float, stage 3: FP32 copy of the new state
  const double* dtp; // RK: device time step
  long N;            // cells per variable plane
  int n[3];          // cells per direction
  int flags;
  int stage;         // 1, 2, 3
  int seg_len;       // cells per segment along AX
  double h;          // spacing along AX
  double ct;         // TENO cut-off constant
  int bc_lo, bc_hi;
  const SweepAux* aux;  // boundary states, origin, gravity, spacings (time: dtp[6], dtp[7])
  // when non-null and *halt != 0 the launch does nothing (queued steps after
  // an output time in cfd_steps_to_output)
  const double* halt;
  // SW_WAVE: the per-workgroup maxima of c + |u_d| (D x blocks)
  double* wave;
  // sweep_kernel_rec (2D FP64 characteristic-wise HLLC, the adjoint's
  // recomputed forward): per interface of the pencil (n + 3 of them, k + 1 =
  // the interface between cells k - 1 and k) the characteristic half-node
  // values QL[4], QR[4] (8 planes of rec_plane doubles) and one word of branch
  // bits (after them): TENO flags (bits 6k..6k+2 left, 6k+3..6k+5 right value
  // of wave k), positivity fallback left (24) / right (25), HLLC branch bits
  // (26..31: clip, left shock, right shock, region << 3).  Plane index
  // (rec_index): x sweep pencil * (n + 3) + k + 1; y sweep pencil pairs, i.e.
  // ((pencil / 2) * (n + 3) + k + 1) * 2 + pencil % 2, so that the y sweep's
  // rows of x-adjacent pencils write 64-byte runs and the adjoint's whole-pencil
  // VJP (two x-adjacent pencils per workgroup) reads contiguous runs.
  void* rec;
  long rec_plane;
};

// doubles per record plane of direction ax (the y planes hold an even number
// of pencils)
inline long rec_plane_of(const int* n, int ax) {
  return (long)(n[ax] + 3) * (ax == 0 ? (long)n[1] : (long)((n[0] + 1) & ~1));
}
// plane index of interface m (0 .. n+2) of pencil gp
template <int AX> __host__ __device__ __forceinline__ long rec_index(long gp, int m, int n) {
  return AX == 0 ? gp * (n + 3) + m : (((gp >> 1) * (n + 3) + m) << 1) + (gp & 1);
}

// Tile shapes: CA interfaces per chunk along the sweep axis x CP pencils.
// For AX = 0 lanes run along the pencil (CA >= 16 keeps loads coalesced);
// for AX > 0 lanes run across x-adjacent pencils (CP >= 16).
template <int T> struct Tile;
template <> struct Tile<0> { static constexpr int CA = 32, CP = 8; };
template <> struct Tile<1> { static constexpr int CA = 16, CP = 16; };
template <> struct Tile<2> { static constexpr int CA = 8, CP = 32; };
__host__ __device__ constexpr int tile_ca(int t) { return t == 0 ? 32 : (t == 1 ? 16 : 8); }
__host__ __device__ constexpr int tile_cp(int t) { return t == 0 ? 8 : (t == 1 ? 16 : 32); }

// Prefetch (template flag PF of sweep_block): the cell phase's global
// operands (R, U^n, U^(k)) are loaded when the chunk starts and held in
// registers across the interface phase.  At two waves per SIMD (the FP64
// kernels) loads issued in the cell phase itself stall it (3D 256^3 y/z
// sweeps 1.47/1.62 -> 1.27/1.31 ms); where the extra registers would cost
// occupancy (the FP32 kernels, the bounded kernels; in the persistent batch
// kernel every sweep but the 2D FP64 y sweep) the operands are loaded in the
// cell phase.
template <class T> struct SweepPrefetch { static constexpr bool value = sizeof(T) == sizeof(double); };
// Waves per SIMD the compiler must leave room for (launch bounds), from
// same-box A/B runs.  Without a bound the FP64 characteristic-wise kernels
// that need ~256 registers take a few accumulation registers on top and run
// one wave per SIMD; bounded to two they run 2-40 % faster:
// * every FP64 x sweep of 2D / 3D (the first sweep of a stage reads neither
//   R nor U^n, so it has nothing to prefetch and no prefetch registers):
//   2D WENO-JS + HLLC 0.317 -> 0.233 ms at 2048^2, 3D TENO-S 1.72 -> 1.06 ms
//   at 256^3; the 2D Roe x sweep spills 76 registers at three waves and is
//   7 % slower there;
// * the 3D WENO-JS / WENO-Z / TENO-A / TENO5 y and z sweeps (17-41 % at
//   256^3; TENO5 keeps the prefetch, 15 % faster than without it), and the
//   research variants' kernels.
// The other y / z sweeps fit two waves with the prefetch (TENO-S 3D loses
// 20-30 % in y / z when bounded).  The FP32 3D TENO-family kernels are
// bounded to three waves without the prefetch (x: 168 VGPRs, no spills;
// y / z then take ~120 and run four): 3D 256^3 TENO5 x sweep 0.91 -> 0.74 ms,
// +8 % per step; the FP32 WENO-JS / WENO-Z x sweeps would spill there.
template <int D, int AX, int SCH, int FLX, class T> struct SweepOcc {
  static constexpr bool dbl = sizeof(T) == sizeof(double);
  static constexpr int waves = !dbl ? (D == 3 && SCH != SCH_JS && SCH != SCH_Z ? 3 : 1)
                               : (D >= 2 && AX == 0) ? 2
                               : (D == 3 && (SCH == SCH_JS || SCH == SCH_Z || SCH == SCH_TENOA || SCH == SCH_TENO)) ? 2
                               : (D >= 2 && (SCH == SCH_TENOSA || SCH == SCH_WHF)) ? 2
                                                                                  : 1;
  // the cell phase's operands are prefetched where the registers allow it
  // (never in the first sweep of a 2D / 3D stage: nothing to prefetch)
  static constexpr bool prefetch = (D == 1 || AX != 0) && (waves == 1 || (D == 3 && dbl && SCH == SCH_TENO));
};

// One workgroup's share of a sweep: pencil group bx, segment by.
// x mod R for 0 <= x < 2R: min(x, x - R) as unsigned (x - R wraps above x
// when x < R) -- two VALU instructions where a modulo takes a multiply-high
// sequence
template <int R> __device__ __forceinline__ int ring_wrap(int x) {
  return (int)min((unsigned)x, (unsigned)(x - R));
}
// c mod R of a non-negative-or-small signed (wave-uniform) position
template <int R> __device__ __forceinline__ int ring_mod(int c) { return ((c % R) + R) % R; }

// Schemes whose half-node value is one of the eight stencil choices of a
// cut-off flag (teno_pick of cut_flag: what inter5 returns for them).
template <int SCH> struct TenoFamily {
  static constexpr bool value = SCH == SCH_TENO || SCH == SCH_TENOS || SCH == SCH_TENOA || SCH == SCH_TENOSA ||
                                SCH == SCH_WHF;
};

template <int D, int AX, int SCH, int FLX, bool CHARW, int TILE, class T, bool PF, bool REC = false>
__device__ __forceinline__ void sweep_block(const SweepArgs& a, const long bx, const int by) {
  static_assert(!REC || (D == 2 && CHARW && FLX == FLX_HLLC), "recording sweeps: 2D char-wise HLLC");
  constexpr int NV = D + 2;
  const T* __restrict__ Ut = static_cast<const T*>(a.U);
  T* Rt = static_cast<T*>(a.R);
  const T ct = T(a.ct);
  constexpr int CA = Tile<TILE>::CA, CP = Tile<TILE>::CP;
  constexpr int RC = CA + 5;  // cell ring slots
  constexpr int FR = CA + 3;  // flux ring slots
  constexpr int NCV = CHARW ? 2 * NV : NV;  // per cell: prim [+ rebuilt cons]
  // LDS layout: [var][slot][pencil] for AX>0 (pencil fastest = lane fastest),
  //             [var][pencil][slot] for AX=0 (slot fastest = lane fastest).
  __shared__ T cell[NCV * RC * CP];
  __shared__ T flux[NV * FR * CP];
  // 3D: physical (node) flux of every ring cell, computed once when the cell
  // is loaded (the cell phase needs it at i-1 and i+1 of each cell): one
  // node flux per cell instead of two (+1.3 % at 256^3).  1D/2D compute both
  // in the cell phase, where the extra LDS ring costs the 2D kernels more in
  // register spills than it saves (-4 % at 2048^2, same-box A/B).
  constexpr bool NF_LDS = D == 3;
  __shared__ T nflux[NF_LDS ? NV * RC * CP : 1];
  auto cidx = [](int v, int slot, int p) -> int {
    return AX == 0 ? (v * CP + p) * RC + slot : (v * RC + slot) * CP + p;
  };
  auto fidx = [](int v, int slot, int p) -> int {
    return AX == 0 ? (v * CP + p) * FR + slot : (v * FR + slot) * CP + p;
  };

  const int lane = threadIdx.x;
  const int la = AX == 0 ? (lane & (CA - 1)) : (lane / CP);
  const int lp = AX == 0 ? (lane / CA) : (lane & (CP - 1));

  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int n = a.n[AX];
  const long sa = AX == 0 ? 1 : (AX == 1 ? (long)nx : (long)nx * ny);
  const long npencil = AX == 0 ? (long)ny * nz : (AX == 1 ? (long)nx * nz : (long)nx * ny);
  const long gp = bx * CP + lp;
  const bool pvalid = gp < npencil;
  long base = 0;
  if (pvalid) {
    if (AX == 0) base = gp * nx;
    else if (AX == 1) base = (gp % nx) + (gp / nx) * (long)nx * ny;
    else base = gp;
  }
  const int c0 = by * a.seg_len;
  const int c1 = min(c0 + a.seg_len, n);

  // ---- load one cell (position g along AX, may be a ghost) into the ring ----
  // fetch_cell: the cell's raw values (the conservative state read from HBM,
  // or the primitive state of a synthesised ghost: raw_prim); commit_cell:
  // primitive conversion, symmetry flip and the ring write.
  auto fetch_cell = [&](int g, T* pr, bool& raw_prim, int& flip) {
    flip = -1;
    bool dir = false, exact = false;
    int src = g;
    if (g < 0 || g >= n) {
      const bool lo = g < 0;
      int bc = lo ? a.bc_lo : a.bc_hi;
      if (D == 2 && AX == 1 && (bc == BC_DMR_WALL || bc == BC_DMR_TOP)) {
        // the pencil's cell-centre x (2D y sweep: pencil index = column)
        const double x0 = a.aux->x0, hx = a.aux->hx;
        const double x = x0 + ((int)gp + 0.5) * hx;
        if (bc == BC_DMR_WALL) {
          bc = x < 1.0 / 6.0 ? BC_DIRICHLET : BC_SYMMETRYY;
        } else {
          // ghost k = g - n at y = -dh/2 + k dh from the face, time of the stage
          // (RK3_SSP: t, t + dt, t + dt - dt/2; blockSolver.cpp:62,74)
          const double dh = (x0 + ((n - 1) + 0.5) * hx) - (x0 + ((n - 2) + 0.5) * hx);
          const double y = -dh / 2 + (g - n) * dh;
          double t = a.dtp[6];
          if (a.stage >= 1) {
            t = a.dtp[7];
            if (a.stage >= 2) t = t + a.dtp[0];
            if (a.stage == 3) t = t - a.dtp[0] / 2;
          }
          const double gt = 1.0 / 6.0 + sqrt(3.0) / 3.0 * (1.0 + 20 * t);
          const bool post = y > sqrt(3.0) * (x - gt);
          for (int v = 0; v < NV; v++)
            pr[v] = T(post ? a.aux->prim_hi[v] : (v == 0 ? 1.4 : (v == NV - 1 ? 1.0 : 0.0)));
          dir = true;
          exact = true;
        }
      }
      switch (bc) {
      case BC_PERIODIC: src = ((g % n) + n) % n; break;
      case BC_OUTFLOW: src = lo ? 0 : n - 1; break;
      case BC_DIRICHLET: dir = true; break;
      case BC_SYMMETRY1D:
      case BC_SYMMETRYX: src = lo ? -1 - g : 2 * n - 1 - g; flip = 1; break;
      case BC_SYMMETRYY: src = lo ? -1 - g : 2 * n - 1 - g; flip = 2; break;
      case BC_SYMMETRYZ: src = lo ? -1 - g : 2 * n - 1 - g; flip = 3; break;
      default: src = lo ? -1 - g : 2 * n - 1 - g; break;  // mirror copy
      }
      if (dir && !exact)
        for (int v = 0; v < NV; v++) pr[v] = T(lo ? a.aux->prim_lo[v] : a.aux->prim_hi[v]);
    }
    raw_prim = dir;
    if (!dir) {
      const long idx = base + (long)src * sa;
      for (int v = 0; v < NV; v++) pr[v] = Ut[v * a.N + idx];
    }
  };
  auto commit_cell = [&](int slot, const T* raw, bool raw_prim, int flip) {
    T pr[NV];
    if (raw_prim) {
      for (int v = 0; v < NV; v++) pr[v] = raw[v];
    } else {
      cons_to_prim<D>(raw, pr);
      for (int v = 1; v < NV - 1; v++)
        if (v == flip) pr[v] = -pr[v];
    }
    for (int v = 0; v < NV; v++) cell[cidx(v, slot, lp)] = pr[v];
    if (NF_LDS) {
      T nf[NV];
      node_flux<D, AX>(pr, nf);
      for (int v = 0; v < NV; v++) nflux[cidx(v, slot, lp)] = nf[v];
    }
    if (CHARW) {
      T cc[NV];
      prim_to_cons<D>(pr, cc);
      for (int v = 0; v < NV; v++) cell[cidx(NV + v, slot, lp)] = cc[v];
    }
  };
  // ring slot of cell g: ring_wrap(base + la) with a wave-uniform base
  auto load_cell = [&](int g, int slot) {
    T raw[NV];
    bool raw_prim;
    int flip;
    fetch_cell(g, raw, raw_prim, flip);
    commit_cell(slot, raw, raw_prim, flip);
  };

  // preload cells c0-4 .. c0
  if (la < 5 && pvalid) load_cell(c0 - 4 + la, ring_wrap<RC>(ring_mod<RC>(c0 - 4) + la));

  const double w0 = 3.0 / 2.0, w1 = -3.0 / 10.0, w2 = 1.0 / 30.0;
  const int m_end = c1 + 3;  // interfaces [c0, c1+3)
  // SW_WAVE: this lane's CFL maxima, kept in LDS (not in registers across the
  // interface phase)
  __shared__ double wmax[3][256];
  if (a.flags & SW_WAVE)
    for (int d = 0; d < D; d++) wmax[d][threadIdx.x] = 0.0;
  for (int m0 = c0; m0 < m_end; m0 += CA) {
    // wave-uniform ring bases of this chunk: the cell slot of cell m0-4
    // (interface windows m0-4.., the cell phase's neighbours, the cells
    // loaded from m0+1 = m0-4+5) and the flux slot of interface m0-3 (the
    // cell phase's interfaces m0-3.., the interfaces m0.. computed)
    const int bI = ring_mod<RC>(m0 - 4), bF3 = ring_mod<FR>(m0 - 3);
    // 1. load cells m0+1 .. m0+CA (up to c1+3)
    {
      const int g = m0 + 1 + la;
      if (pvalid && g <= c1 + 3) load_cell(g, ring_wrap<RC>(bI + 5 + la));
    }
    // operands of this chunk's cell phase (cell m0-3+la), in flight during phase 2
    T pf_r[NV];
    double pf_u0[NV], pf_uc[NV];
    if (PF) {
      const int i = m0 - 3 + la;
      if (pvalid && i >= c0 && i < c1) {
        const long idx = base + (long)i * sa;
        for (int v = 0; v < NV; v++) {
          pf_r[v] = (a.flags & SW_READ_R) ? Rt[v * a.N + idx] : T(0.0);
          pf_u0[v] = (a.flags & SW_RK) ? a.U0[v * a.N + idx] : 0.0;
          pf_uc[v] = ((a.flags & SW_RK) && a.stage != 1) ? (double)Ut[v * a.N + idx] : 0.0;
        }
      }
    }
    __syncthreads();
    // 2. interface m = m0 + la: between cells m-1 and m... (cells m-4 .. m+1)
    {
      const int m = m0 + la;
      if (pvalid && m < m_end) {
        // window cells m-4+j, j = 0..5; the interface lies between m-2 and m-1
        int sl[6];
        for (int j = 0; j < 6; j++) sl[j] = ring_wrap<RC>(bI + la + j);
        T outL[NV], outR[NV];
        if (!CHARW) {
          for (int v = 0; v < NV; v++) {
            T q0 = cell[cidx(v, sl[0], lp)], q1 = cell[cidx(v, sl[1], lp)], q2 = cell[cidx(v, sl[2], lp)];
            T q3 = cell[cidx(v, sl[3], lp)], q4 = cell[cidx(v, sl[4], lp)], q5 = cell[cidx(v, sl[5], lp)];
            outL[v] = inter5<SCH>(q0, q1, q2, q3, q4, ct, &a.aux->rc);
            outR[v] = inter5<SCH>(q5, q4, q3, q2, q1, ct, &a.aux->rc);
          }
        } else {
          T pL[NV], pR[NV];
          for (int v = 0; v < NV; v++) {
            pL[v] = cell[cidx(v, sl[2], lp)];
            pR[v] = cell[cidx(v, sl[3], lp)];
          }
          if (D == 1) {
            RoeSys1<T> es;
            es.build(pL, pR);
            T QL[3], QR[3];
            for (int k = 0; k < 3; k++) {
              T ch[6];
              for (int j = 0; j < 6; j++) {
                T c0v = cell[cidx(NV + 0, sl[j], lp)], c1v = cell[cidx(NV + 1, sl[j], lp)],
                  c2v = cell[cidx(NV + 2, sl[j], lp)];
                ch[j] = c0v * es.L[3 * k] + c1v * es.L[3 * k + 1] + c2v * es.L[3 * k + 2];
              }
              QL[k] = inter5<SCH>(ch[0], ch[1], ch[2], ch[3], ch[4], ct, &a.aux->rc);
              QR[k] = inter5<SCH>(ch[5], ch[4], ch[3], ch[2], ch[1], ct, &a.aux->rc);
            }
            T cL[3], cR[3];
            for (int i = 0; i < 3; i++) {
              cL[i] = QL[0] * es.R[3 * i] + QL[1] * es.R[3 * i + 1] + QL[2] * es.R[3 * i + 2];
              cR[i] = QR[0] * es.R[3 * i] + QR[1] * es.R[3 * i + 1] + QR[2] * es.R[3 * i + 2];
            }
            T consL[3], consR[3];
            prim_to_cons<1>(pL, consL);
            prim_to_cons<1>(pR, consR);
            positive_limit(consL, cL, 0);
            positive_limit(consR, cR, 0);
            positive_limit(consL, cL, 1);
            positive_limit(consR, cR, 1);
            eig_cons_to_prim<1>(cL, outL);
            eig_cons_to_prim<1>(cR, outR);
          } else {
            RoeSys<D, AX, T> es;
            es.build(pL, pR);
            T cw[6][NV];
            for (int j = 0; j < 6; j++)
              for (int v = 0; v < NV; v++) cw[j][v] = cell[cidx(NV + v, sl[j], lp)];
            T QL[NV], QR[NV];
            unsigned rbits = 0u;  // REC: branch bits of this interface
            for (int k = 0; k < NV; k++) {
              T l[NV];
              if (k != 0 && !(D == 3 && k == 4)) es.lrow(k, l);
              T ch[6];
              for (int j = 0; j < 6; j++) ch[j] = es.proj(k, l, cw[j]);
              if (REC && TenoFamily<SCH>::value) {
                // inter5's own evaluation for these schemes, flags kept
                const unsigned fl = cut_flag<SCH>(ch[0], ch[1], ch[2], ch[3], ch[4], ct, &a.aux->rc);
                const unsigned fr = cut_flag<SCH>(ch[5], ch[4], ch[3], ch[2], ch[1], ct, &a.aux->rc);
                QL[k] = teno_pick(fl, ch[0], ch[1], ch[2], ch[3], ch[4]);
                QR[k] = teno_pick(fr, ch[5], ch[4], ch[3], ch[2], ch[1]);
                rbits |= (fl | (fr << 3)) << (6 * k);
              } else {
                QL[k] = inter5<SCH>(ch[0], ch[1], ch[2], ch[3], ch[4], ct, &a.aux->rc);
                QR[k] = inter5<SCH>(ch[5], ch[4], ch[3], ch[2], ch[1], ct, &a.aux->rc);
              }
            }
            long ri = 0;
            if (REC) {  // the half-node values, stored as soon as they exist
              ri = rec_index<AX>(gp, m, n);
              double* rd = static_cast<double*>(a.rec);
              for (int k = 0; k < NV; k++) {
                rd[k * a.rec_plane + ri] = QL[k];
                rd[(NV + k) * a.rec_plane + ri] = QR[k];
              }
            }
            // charToCons in the summation order (ch3, ch4, ch2, ch1[, ch5])
            T cL[NV], cR[NV];
            es.back(QL, QR, cL, cR);
            eig_cons_to_prim<D>(cL, outL);
            eig_cons_to_prim<D>(cR, outR);
            // positivity fallback (reconstructor5order.hpp:301-302)
            const bool fbL = outL[0] < 0 || outL[NV - 1] < 0;
            const bool fbR = outR[0] < 0 || outR[NV - 1] < 0;
            if (fbL)
              for (int v = 0; v < NV; v++) outL[v] = pL[v];
            if (fbR) {
              outR[0] = pR[0];
              for (int v = 1; v < NV - 1; v++) outR[v] = pL[v];
              outR[NV - 1] = pR[NV - 1];
            }
            if (REC) {
              rbits |= (fbL ? 1u << 24 : 0u) | (fbR ? 1u << 25 : 0u);
              int hb = 0;
              T F[NV];
              hllc_md<D, AX>(outL, outR, F, &hb);
              reinterpret_cast<unsigned*>(static_cast<double*>(a.rec) + 2 * NV * a.rec_plane)[ri] =
                  rbits | ((unsigned)hb << 26);
              const int fs = ring_wrap<FR>(bF3 + 3 + la);
              for (int v = 0; v < NV; v++) flux[fidx(v, fs, lp)] = F[v];
            }
          }
        }
        if (!REC) {
          T F[NV];
          if (D == 1) {
            if (FLX == FLX_HLLC) hllc1d(outL, outR, F);
            else roe1d(outL, outR, F);
          } else {
            if (FLX == FLX_HLLC) hllc_md<D, AX>(outL, outR, F);
            else roe_md<D, AX>(outL, outR, F);
          }
          const int fs = ring_wrap<FR>(bF3 + 3 + la);
          for (int v = 0; v < NV; v++) flux[fidx(v, fs, lp)] = F[v];
        }
      }
    }
    __syncthreads();
    // 3. cell i = m0-3+la uses interfaces i .. i+3 and node fluxes at i-1, i+1
    {
      const int i = m0 - 3 + la;
      if (pvalid && i >= c0 && i < c1) {
        T fm[NV], fp[NV];
        const int sm = ring_wrap<RC>(bI + la), sp = ring_wrap<RC>(bI + la + 2);
        if (NF_LDS) {
          for (int v = 0; v < NV; v++) {
            fm[v] = nflux[cidx(v, sm, lp)];
            fp[v] = nflux[cidx(v, sp, lp)];
          }
        } else {
          T pm[NV], pp[NV];
          for (int v = 0; v < NV; v++) {
            pm[v] = cell[cidx(v, sm, lp)];
            pp[v] = cell[cidx(v, sp, lp)];
          }
          node_flux<D, AX>(pm, fm);
          node_flux<D, AX>(pp, fp);
        }
        const int f0 = ring_wrap<FR>(bF3 + la), f1 = ring_wrap<FR>(bF3 + la + 1);
        const int f2 = ring_wrap<FR>(bF3 + la + 2), f3 = ring_wrap<FR>(bF3 + la + 3);
        const long idx = base + (long)i * sa;
        double dtv = 0.0;
        if (a.flags & SW_RK) dtv = *a.dtp;
        const T hh = T(a.h);
        // gravity source of the cell (SourceTerm::calGravitySource,
        // sourceTerm.cpp:3-20, for an acceleration vector), subtracted from
        // the rhs after every direction's flux derivative
        T src[NV];
        const bool source = AX == D - 1 && (a.flags & SW_SOURCE);
        if (source) {
          const int si = ring_wrap<RC>(bI + la + 1);
          T pc[NV];
          for (int v = 0; v < NV; v++) pc[v] = cell[cidx(v, si, lp)];
          const double* grav = a.aux->grav;
          T ug = pc[1] * T(grav[0]);
          for (int d = 1; d < D; d++) ug = ug + pc[1 + d] * T(grav[d]);
          src[0] = T(0.0);
          for (int d = 0; d < D; d++) src[1 + d] = pc[0] * T(grav[d]);
          src[NV - 1] = pc[0] * ug;
        }
        double unv[NV];
        for (int v = 0; v < NV; v++) {
          const T H0 = flux[fidx(v, f0, lp)], H1 = flux[fidx(v, f1, lp)];
          const T H2 = flux[fidx(v, f2, lp)], H3 = flux[fidx(v, f3, lp)];
          const T term = fdv(T(w1) * (fp[v] - fm[v]) + T(w0) * (H2 - H1) + T(w2) * (H3 - H0), hh);
          const T rold = PF ? pf_r[v] : ((a.flags & SW_READ_R) ? Rt[v * a.N + idx] : T(0.0));
          T r = rold + term;
          if (source && v > 0) r = r - src[v];
          if (a.flags & SW_WRITE_R) Rt[v * a.N + idx] = r;
          if (a.flags & SW_RK) {
            // stage combination in FP64 (BlockSolver::RK3_SSP, blockSolver.cpp:60,72,84)
            const double u0 = PF ? pf_u0[v] : a.U0[v * a.N + idx];
            const double rd = (double)r;
            double un;
            if (a.stage == 1) {
              un = u0 - dtv * rd;
            } else {
              const double uc = PF ? pf_uc[v] : (double)Ut[v * a.N + idx];
              if (a.stage == 2) un = 0.75 * u0 - 0.25 * dtv * rd + 0.25 * uc;
              else un = 1.0 / 3.0 * u0 - 2.0 / 3.0 * dtv * rd + 2.0 / 3.0 * uc;
            }
            if (sizeof(T) == sizeof(double) || a.stage != 3) {
              static_cast<T*>(a.Uout)[v * a.N + idx] = (T)un;
            } else {
              static_cast<double*>(a.Uout)[v * a.N + idx] = un;
              static_cast<T*>(a.Ushadow)[v * a.N + idx] = (T)un;
            }
            unv[v] = un;
          }
        }
        if (a.flags & SW_WAVE) {
          // BlockSolver::getTimeIntervalExplicit's maxima (blockSolver.cpp:325-338)
          // of the new state, the expressions of k_wave_speed
          double p[NV];
          cons_to_prim<D>(unv, p);
          // max_i (c + |u_d|)_i / h_d = (max_i (c + |u_d|)_i) / h_d exactly (a
          // correctly rounded division by a positive constant is monotone), so
          // the division is left to the reduction of the partials
          const double cs = fsq(fdv(CFD_GAMMA * p[NV - 1], p[0]));
          for (int d = 0; d < D; d++) {
            const double l = cs + fabs(p[1 + d]);
            const double m = wmax[d][threadIdx.x];
            wmax[d][threadIdx.x] = m < l ? l : m;
          }
        }
      }
    }
    __syncthreads();
  }
  if (a.flags & SW_WAVE) {
    // workgroup maxima: wave reductions, then the four waves through LDS
    __shared__ double wred[4][3];
    for (int d = 0; d < D; d++) {
      double m = wmax[d][threadIdx.x];
      for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = m < o ? o : m;
      }
      if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6][d] = m;
    }
    __syncthreads();
    if (threadIdx.x < D) {
      double m = wred[0][threadIdx.x];
      for (int w = 1; w < 4; w++) m = m < wred[w][threadIdx.x] ? wred[w][threadIdx.x] : m;
      const long nb = (long)gridDim.x * gridDim.y;
      a.wave[threadIdx.x * nb + (long)blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
  }
}

template <int D, int AX, int SCH, int FLX, bool CHARW, int TILE, class T>
__global__ __launch_bounds__(256, (SweepOcc<D, AX, SCH, FLX, T>::waves)) void sweep_kernel(SweepArgs a) {
  if (a.halt && *a.halt != 0.0) return;
  sweep_block<D, AX, SCH, FLX, CHARW, TILE, T, SweepPrefetch<T>::value && SweepOcc<D, AX, SCH, FLX, T>::prefetch>(
      a, blockIdx.x, blockIdx.y);
}

// The same kernel for the FMA-contracted FP64 mode (CFD_PRECISION_FP64_FMA):
// instantiated only in the translation units compiled with
// -ffp-contract=fast (sweep_fma_d*.hip), where every a*b + c of the path may
// be evaluated with one rounding; its own name keeps those kernels apart from
// the exact ones.
template <int D, int AX, int SCH, int FLX, bool CHARW, int TILE>
__global__ __launch_bounds__(256, (SweepOcc<D, AX, SCH, FLX, double>::waves)) void sweep_kernel_fma(SweepArgs a) {
  if (a.halt && *a.halt != 0.0) return;
  sweep_block<D, AX, SCH, FLX, CHARW, TILE, double, SweepOcc<D, AX, SCH, FLX, double>::prefetch>(a, blockIdx.x,
                                                                                              blockIdx.y);
}

// The adjoint's recomputed forward: the 2D FP64 characteristic-wise HLLC
// sweep that also writes each interface's half-node values and branch bits
// (SweepArgs::rec); its results are the plain kernel's, bit for bit.
template <int AX, int SCH, int TILE>
__global__ __launch_bounds__(256, (SweepOcc<2, AX, SCH, FLX_HLLC, double>::waves)) void sweep_kernel_rec(SweepArgs a) {
  if (a.halt && *a.halt != 0.0) return;
  sweep_block<2, AX, SCH, FLX_HLLC, true, TILE, double, SweepOcc<2, AX, SCH, FLX_HLLC, double>::prefetch, true>(
      a, blockIdx.x, blockIdx.y);
}

// ---- batched zones ------------------------------------------------------------
// Independent structured zones of uneven sizes share pooled buffers; every
// zone keeps the single-block layout (nv planes of its own cell count) at
// element offset `off` of each pool.
struct ZoneDesc {
  long off;          // element offset of the zone in every pool
  long N;            // cells
  int n[3];
  int seg_len[3];    // segment length per direction
  double h[3];
  int bc[6];
  SweepAux aux[3];   // per direction: the faces' Dirichlet states (origin, gravity, spacings zero)
};

struct WorkItem {
  int zone;
  int bx;  // pencil group
  int by;  // segment
};

// Pools are typed like SweepArgs' buffers: U, R and the stage outputs hold
// the kernel's compute type T; U0 is FP64; with T = float the stage-3 output
// is the FP64 pool U0 plus its FP32 shadow Ushadow.
struct BatchArgs {
  const void* U;
  void* R;
  const double* U0;
  void* Uout;
  void* Ushadow;
  const double* dt;        // per-zone time step
  const ZoneDesc* zones;
  const WorkItem* items;   // work list of this direction, largest zones first
  int nitems;
  unsigned* counter;       // work-queue head of this launch (zeroed before the launch)
  int flags;
  int stage;
  double ct;
};

// Persistent kernel: each workgroup pulls work items from the queue until it
// is empty, so one launch covers every zone of the batch.  Two waves per
// SIMD; the 2D FP64 y sweep prefetches its cell-phase operands (96 uneven
// zones, WCNS5 + HLLC: 7160 -> 8145 Mcell-stage/s, same bits).
template <int D, int AX, int SCH, int FLX, bool CHARW, int TILE, class T>
__global__ __launch_bounds__(256, 2) void sweep_batch_kernel(BatchArgs b) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int)atomicAdd(b.counter, 1u);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= b.nitems) break;
    const WorkItem w = b.items[item];
    const int z = __builtin_amdgcn_readfirstlane(w.zone);
    const ZoneDesc& zd = b.zones[z];
    SweepArgs a;
    a.U = static_cast<const T*>(b.U) + zd.off;
    a.R = static_cast<T*>(b.R) + zd.off;
    a.U0 = b.U0 + zd.off;
    if (sizeof(T) == sizeof(double) || b.stage != 3) {
      a.Uout = static_cast<T*>(b.Uout) + zd.off;
      a.Ushadow = nullptr;
    } else {
      a.Uout = static_cast<double*>(b.Uout) + zd.off;
      a.Ushadow = static_cast<T*>(b.Ushadow) + zd.off;
    }
    a.halt = nullptr;
    a.wave = nullptr;
    a.dtp = b.dt + z;
    a.N = zd.N;
    a.n[0] = zd.n[0];
    a.n[1] = zd.n[1];
    a.n[2] = zd.n[2];
    a.flags = b.flags;
    a.stage = b.stage;
    a.seg_len = zd.seg_len[AX];
    a.h = zd.h[AX];
    a.ct = b.ct;
    a.bc_lo = zd.bc[2 * AX];
    a.bc_hi = zd.bc[2 * AX + 1];
    a.aux = &zd.aux[AX];  // batches take neither position-dependent boundaries nor sources
    sweep_block<D, AX, SCH, FLX, CHARW, TILE, T, (D == 2 && AX != 0 && sizeof(T) == sizeof(double))>(
        a, __builtin_amdgcn_readfirstlane(w.bx), __builtin_amdgcn_readfirstlane(w.by));
  }
}

}  // namespace cfd
