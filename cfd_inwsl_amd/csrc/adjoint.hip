// Discrete adjoint of the 2D right-hand side and of the SSP-RK3 step for the
// characteristic-wise WCNS5 (WENO-JS) or TENO-family (TENO5, TENO-S, TENO-A,
// the whf_* research variants) + HLLC + MND6 configurations, FP64.  The TENO
// family's half-node value is the stencil its cut-off flag selects, so its
// reverse pass is that stencil with the flag the forward took.
//
// Forward map (per direction, per pencil of n cells, G_k the interface flux
// between cells k-1 and k built from cells k-3 .. k+2, f_j the node flux):
//   rhs_i += (w1 (f_{i+1} - f_{i-1}) + w0 (G_{i+1} - G_i) + w2 (G_{i+2} - G_{i-1})) / h
// (SpDistributor::rhsSolve, zone_FD/src/src/sp_distributor.cpp:9-78, with
// MidNodeAndNodeDif6, zone_FD/src/src/differ.cpp:54-103).  Its transpose:
//   lambda_G_k = (w0 (mu_{k-1} - mu_k) + w2 (mu_{k-2} - mu_{k+1})) / h
//   lambda_f_j = w1 (mu_{j-1} - mu_{j+1}) / h              (mu = 0 off the pencil)
// then the interface / node-flux vector-Jacobian products give primitive
// cotangents on cells -4 .. n+3, ghosts fold back onto their source cells,
// and the conservative cotangent is (d prim / d cons)^T of the sum.
//
// Kernel: one workgroup owns CP pencils x one segment of cells and marches
// along the pencil in chunks of CA interfaces.  Per chunk: load cells
// (primitive from U, scaled cotangent) into LDS rings; each lane takes one
// interface through the reverse pass and stages its six cell cotangents in
// LDS (two staging buffers, this chunk's and the previous chunk's); after one
// barrier each lane finalises one cell whose six interfaces are done, summing
// their staged contributions in a fixed interface order (no two lanes write
// one cell and the summation order is fixed: results are deterministic).  Ghost-cell cotangents go to a small
// per-pencil buffer; the x sweep's are folded onto their source cells by the
// y sweep's cell phase (stream order makes them visible), the y sweep's by a
// small kernel after it.  A direction periodic on both faces has no ghosts:
// its pencils are marched as rings (indices mod n; the interfaces the forward
// evaluates twice across the seam, with ghost copies, are the same function
// of the same cells, so the ring's single interface carries their summed
// cotangent).
//
// RK3 (BlockSolver::RK3_SSP, zone_FD/src/src/blockSolver.cpp:43-87):
//   u1 = u0 - dt R(u0); u2 = 3/4 u0 + 1/4 u1 - 1/4 dt R(u1);
//   u3 = 1/3 u0 + 2/3 u2 - 2/3 dt R(u2)
// reversed with l3 = dJ/du3:
//   l2 = 2/3 l3 - 2/3 dt R'(u2)^T l3;  l1 = 1/4 l2 - 1/4 dt R'(u1)^T l2;
//   l0 = 1/3 l3 + 3/4 l2 + l1 - dt R'(u0)^T l1.
// Time steps are fixed (dt is a constant of the map, not a function of u).
// The forward pass keeps the state every `checkpoint_every` steps; the
// backward pass recomputes each segment's stage states from its checkpoint
// with the forward sweep kernels, then walks the segment backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>

#include "adjoint_point.hpp"
#include "euler_point.hpp"
#include "handle.hpp"
#include "sweep.hpp"

namespace {

// Workgroup tile of the VJP sweep per direction: CA interfaces per chunk x CP
// pencils (256 lanes), and the waves per SIMD it is compiled for.
#ifndef CFD_ADJ_CA0
#define CFD_ADJ_CA0 32
#define CFD_ADJ_CP0 8
#endif
#ifndef CFD_ADJ_CA1
#define CFD_ADJ_CA1 16
#define CFD_ADJ_CP1 16
#endif
#ifndef CFD_ADJ_OCC
#define CFD_ADJ_OCC 1
#endif

template <int AX> struct VjpTile {
  static constexpr int CA = AX == 0 ? CFD_ADJ_CA0 : CFD_ADJ_CA1;
  static constexpr int CP = AX == 0 ? CFD_ADJ_CP0 : CFD_ADJ_CP1;
};
// whole ring pencils per workgroup (FULL): two pencils x 128 interfaces
constexpr int kFullCA = 128, kFullCP = 2;

// OUT = base + coef * (d rhs_AX / d U)^T MU, with base = OUT, or (init) the
// combination sum_i b_i X_i of up to three SoA arrays (null X_i: absent).
struct VjpInit {
  const double* X[3];
  double b[3];
};

struct VjpArgs {
  const double* U;   // linearisation state, SoA planes of N
  const double* MU;  // cotangent of the rhs, SoA
  double* OUT;       // result, SoA (may alias an X of init)
  VjpInit init;
  int use_init;
  double* GH;        // ghost cotangents of this direction [pencil][8][4]
  const double* GHX; // y sweep: the x sweep's ghost cotangents, folded in here
  // y sweep: per x ghost slot q (4 bits at 4q) its source column as c < 4 ? c
  // : c - nx + 8 (15: fixed state), and (2 bits at 2q) the velocity
  // component it negates (0: none)
  unsigned gx_src, gx_flip;
  double coef;
  double ct;         // cut-off constant / law of the TENO-family scheme
  const cfd::ResearchConsts* rc;  // constants of the research variants (device)
  long N;
  int n[2];
  int seg_len;
  double h;
  int bc_lo, bc_hi;
  // both faces of this direction periodic: the pencil is marched as a ring
  // (cells and interfaces indexed mod n), so no ghost cell and no ghost
  // cotangent exists -- every interface's cotangents land on real cells
  int ring;
  const double* bc_prim;  // Dirichlet ghost states of this direction: low face [0..3], high [4..7]
  // k_rhs_vjp<..., REC = true>: the recording sweep's record of this state
  // and direction (SweepArgs::rec: QL, QR planes, then the branch words)
  const double* rec;
  long rec_plane;
};

// Forward record of one interface in the forward sweep's own arithmetic
// (the characteristic-wise path of sweep_block, operation for operation), so
// that every branch the reverse pass follows -- positivity fallback, HLLC
// pressure clipping, shock factors, wave region -- is the branch the solver
// took, also where a predicate is tied at rounding level (pressure
// equilibrium makes p* - p_L vanish up to rounding).
template <int AX, class IP, int SCH, class W>
__device__ __forceinline__ void iface_forward_exact(const W& w, cfdadj::IfaceFwd& f, double ct,
                                                    const cfd::ResearchConsts* rc) {
  double pL[4], pR[4];
  for (int v = 0; v < 4; v++) {
    pL[v] = w(2, v);
    pR[v] = w(3, v);
  }
  cfd::RoeSys<2, AX, double> es;
  es.build(pL, pR);
  f.tflag = 0u;
  for (int k = 0; k < 4; k++) {
    double l[4];
    if (k != 0) es.lrow(k, l);
    double ch[6];
    for (int j = 0; j < 6; j++) {
      double c[4];
      for (int v = 0; v < 4; v++) c[v] = w.cons(j, v);
      ch[j] = es.proj(k, l, c);
    }
    if constexpr (IP::kLinear) {
      const unsigned fl = cfd::cut_flag<SCH>(ch[0], ch[1], ch[2], ch[3], ch[4], ct, rc);
      const unsigned fr = cfd::cut_flag<SCH>(ch[5], ch[4], ch[3], ch[2], ch[1], ct, rc);
      f.QL[k] = cfd::teno_pick(fl, ch[0], ch[1], ch[2], ch[3], ch[4]);
      f.QR[k] = cfd::teno_pick(fr, ch[5], ch[4], ch[3], ch[2], ch[1]);
      f.tflag |= (fl | (fr << 3)) << (6 * k);
    } else {
      f.QL[k] = cfd::inter5<cfd::SCH_JS>(ch[0], ch[1], ch[2], ch[3], ch[4], 0.0);
      f.QR[k] = cfd::inter5<cfd::SCH_JS>(ch[5], ch[4], ch[3], ch[2], ch[1], 0.0);
    }
  }
  // charToCons in the summation order (ch3, ch4, ch2, ch1)
  es.back(f.QL, f.QR, f.cL, f.cR);
  cfd::eig_cons_to_prim<2>(f.cL, f.oL);
  cfd::eig_cons_to_prim<2>(f.cR, f.oR);
  f.fbL = f.oL[0] < 0 || f.oL[3] < 0;
  f.fbR = f.oR[0] < 0 || f.oR[3] < 0;
  if (f.fbL)
    for (int v = 0; v < 4; v++) f.oL[v] = pL[v];
  if (f.fbR) {
    f.oR[0] = pR[0];
    f.oR[1] = pL[1];
    f.oR[2] = pL[2];
    f.oR[3] = pR[3];
  }
  f.hb = cfdadj::hllc_branch_bits<AX>(f.oL, f.oR);
}

// REC: the interface's half-node values and branch bits come from the
// recording sweep of the same state (iface_from_record) instead of an exact
// re-evaluation of the interface.
//
// FULL (a ring direction whose pencils are each marched by one workgroup,
// seg_len = n): the n interfaces are evaluated once each, in chunks from
// interface 0.  Cells -3 .. 1 of the first chunk (n-3 .. n-1, 0, 1) still miss
// interfaces the last chunk evaluates: their partial sums are carried in LDS
// and completed after the last chunk, so no interface is evaluated twice
// (a segment of L cells otherwise evaluates L + 5).
template <int AX, int CA, int CP, class IP, int SCH, bool REC, bool FULL>
__global__ __launch_bounds__(CA* CP, CFD_ADJ_OCC) void k_rhs_vjp(VjpArgs a) {
  constexpr int RC = CA + 6;  // ring slots: cells k0-4 .. k0+CA+1
  __shared__ double wc[8 * RC * CP];  // primitive state, then the conservative state rebuilt from it
  __shared__ double mc[4 * RC * CP];  // coef * cotangent
  // six cell cotangents per interface, for this chunk and the one before
  // (double-buffered by chunk parity: a cell's six interfaces span both)
  __shared__ double st[2 * 24 * CA * CP];
  __shared__ double carry[FULL ? 20 * CP : 1];  // FULL: cells -3 .. 1 after the first chunk
  auto idx = [](int v, int slot, int p) -> int {
    return AX == 0 ? (v * CP + p) * RC + slot : (v * RC + slot) * CP + p;
  };
  // (cell j of the interface, variable) x interface x pencil, laid out in lane order
  auto sidx = [](int jv, int kk, int p) -> int { return jv * CA * CP + (AX == 0 ? p * CA + kk : kk * CP + p); };
  auto slot_of = [](int c) -> int { return (c + 8 * RC) % RC; };

  const int lane = threadIdx.x;
  const int la = AX == 0 ? (lane % CA) : (lane / CP);
  const int lp = AX == 0 ? (lane / CA) : (lane % CP);
  const int nx = a.n[0], ny = a.n[1];
  const int n = a.n[AX];
  const long sa = AX == 0 ? 1 : nx;
  const long npencil = AX == 0 ? ny : nx;
  // FULL y sweeps (two x-adjacent pencils per workgroup): consecutive
  // workgroups of one XCD (dispatch round-robins workgroups over the 8 XCDs)
  // take adjacent pencil pairs, so that the row segments they read share that
  // XCD's L2 lines
  const unsigned nb = gridDim.x;
  const unsigned bx = (FULL && AX == 1 && nb % 8 == 0) ? (blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8 : blockIdx.x;
  const long gp = (long)bx * CP + lp;
  const bool pvalid = gp < npencil;
  const long base = pvalid ? (AX == 0 ? gp * nx : gp) : 0;
  const int c0 = blockIdx.y * a.seg_len;
  const int c1 = min(c0 + a.seg_len, n);
  const bool ring = FULL || a.ring != 0;
  // cells finalised in the chunk loop (ghosts at the pencil ends, none on a
  // ring; FULL: cells 2 .. n-4, the others after it)
  const int e0 = FULL ? 2 : c0 == 0 && !ring ? -4 : c0;
  const int e1 = FULL ? n - 3 : c1 == n && !ring ? n + 4 : c1;
  const int kmin = FULL ? 0 : ring ? e0 - 2 : max(e0 - 2, -1);
  const int kmax = FULL ? n - 1 : ring ? e1 + 2 : min(e1 + 2, n + 1);
  // ring index of a cell or interface in [-8, n + 8) (n >= 8)
  auto wrap = [n](int c) -> int { return c < 0 ? c + n : (c >= n ? c - n : c); };
  const double w0 = 3.0 / 2.0, w1 = -3.0 / 10.0, w2 = 1.0 / 30.0;
  const double ih = 1.0 / a.h;

  // Global loads of one ring cell (interior: state and cotangent; ghost: the
  // state of its source cell), issued one chunk before the cell enters the
  // ring so that their latency hides behind a chunk of interface work.
  struct Fetch {
    double u[4], m[4];
  };
  auto fetch = [&](int c, Fetch& f) {
    const bool in = ring || (c >= 0 && c < n);
    const cfdadj::GhostSrc gs = cfdadj::ghost_source(c < 0 ? a.bc_lo : a.bc_hi, c, n);
    const long i = base + (long)(ring ? wrap(c) : in ? c : gs.fixed ? 0 : gs.src) * sa;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      f.u[v] = a.U[v * a.N + i];
      f.m[v] = in ? a.MU[v * a.N + i] : 0.0;
    }
  };
  // primitive state, its rebuilt conservative state and coef * cotangent into the ring
  auto commit = [&](int c, const Fetch& f) {
    const int s = slot_of(c);
    double p[4], m[4] = {0.0, 0.0, 0.0, 0.0};
    if (ring || (c >= 0 && c < n)) {
      cfd::cons_to_prim<2>(f.u, p);
#pragma unroll
      for (int v = 0; v < 4; v++) m[v] = a.coef * f.m[v];
    } else {
      const cfdadj::GhostSrc gs = cfdadj::ghost_source(c < 0 ? a.bc_lo : a.bc_hi, c, n);
      if (gs.fixed) {
#pragma unroll
        for (int v = 0; v < 4; v++) p[v] = a.bc_prim[(c < 0 ? 0 : 4) + v];
      } else {
        cfd::cons_to_prim<2>(f.u, p);
        p[1] = gs.flip == 1 ? -p[1] : p[1];  // selects, not an index into p
        p[2] = gs.flip == 2 ? -p[2] : p[2];
      }
    }
    double cc[4];
    cfd::prim_to_cons<2>(p, cc);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      wc[idx(v, s, lp)] = p[v];
      wc[idx(4 + v, s, lp)] = cc[v];
      mc[idx(v, s, lp)] = m[v];
    }
  };
  auto M = [&](int c, int v) -> double { return mc[idx(v, slot_of(c), lp)]; };
  // the six cells of interface k: ring offsets computed once
  struct Cells {
    const double* wc;
    int off[6];
    __device__ Cells(const double* w, int k, int lp) : wc(w) {
      int s = (k - 3 + 8 * RC) % RC;
      for (int j = 0; j < 6; j++) {
        off[j] = AX == 0 ? lp * RC + s : s * CP + lp;
        s = s + 1 == RC ? 0 : s + 1;
      }
    }
    __device__ double at(int j, int v) const { return wc[off[j] + v * RC * CP]; }
    __device__ double operator()(int j, int v) const { return at(j, v); }
    __device__ double cons(int j, int v) const { return at(j, 4 + v); }
  };

  // the base terms of the output at cell index fi (init's combination, or OUT)
  auto load_base = [&](long fi, double fx[3][4]) {
    if (a.use_init) {
#pragma unroll
      for (int t = 0; t < 3; t++) {
        const double* X = a.init.X[t];
#pragma unroll
        for (int v = 0; v < 4; v++) fx[t][v] = X ? X[v * a.N + fi] : 0.0;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; v++) fx[0][v] = a.OUT[v * a.N + fi];
    }
  };
  // Adds to lwc the staged cotangents of cell c from its interfaces k = c+3,
  // c+2, .., c-2 (in this order, so that every cell's sum has a fixed order):
  // the staging of the chunk starting at k0 for k >= k0, the previous chunk's
  // for k < k0.
  auto gather = [&](int c, int k0, double* lwc) {
    const int par = ((k0 - kmin) / CA) & 1;
    const double* const st_cur = st + par * 24 * CA * CP;
    const double* const st_prev = st + (par ^ 1) * 24 * CA * CP;
#pragma unroll
    for (int j = 0; j < 6; j++) {
      const int kk = c + 3 - j;
      if (kk >= kmin && kk <= kmax) {
        const double* src = kk >= k0 ? st_cur + sidx(j * 4, kk - k0, lp) : st_prev + sidx(j * 4, kk - k0 + CA, lp);
#pragma unroll
        for (int v = 0; v < 4; v++) lwc[v] += src[v * CA * CP];
      }
    }
  };
  // Finalises cell c (ring index; interior: output index fi with base terms
  // fx) from the summed interface cotangents lwc: node flux, the x ghosts of
  // its row (y sweep), consToPrim's transpose; a ghost cell's cotangent goes
  // to GH.
  auto finish = [&](int c, bool interior, long fi, const double fx[3][4], double* lwc) {
    const int s = slot_of(c);
    const int cr = ring ? wrap(c) : c;
    if (ring || (c >= -1 && c <= n)) {  // node fluxes exist on cells -1 .. n
      double p[4], lf[4];
      for (int v = 0; v < 4; v++) {
        p[v] = wc[idx(v, s, lp)];
        lf[v] = w1 * (M(c - 1, v) - M(c + 1, v)) * ih;
      }
      cfdadj::node_flux_vjp<AX>(p, lf, lwc);
    }
    if (AX == 1 && cr >= 0 && cr < n && (gp < 4 || gp >= nx - 4)) {
      // x ghosts of row cr whose source is this column (sources lie within
      // four cells of the row ends)
      const unsigned col = gp < 4 ? (unsigned)gp : (unsigned)(gp - nx + 8);
      for (int q = 0; q < 8; q++)
        if (((a.gx_src >> (4 * q)) & 15u) == col) {
          const int flip = (int)((a.gx_flip >> (2 * q)) & 3u);
          for (int v = 0; v < 4; v++) {
            const double x = a.GHX[((long)cr * 8 + q) * 4 + v];
            lwc[v] += (flip != 0 && v == flip) ? -x : x;
          }
        }
    }
    if (interior) {
      // the conservative state the ring rebuilt from the cell's primitive
      // state (the cell's own state up to rounding, which a derivative
      // does not see)
      double fu[4], g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int v = 0; v < 4; v++) fu[v] = wc[idx(4 + v, s, lp)];
      cfdadj::c2p_vjp(fu, lwc, g);
      for (int v = 0; v < 4; v++) {
        double b;
        if (a.use_init) {
          b = 0.0;
#pragma unroll
          for (int t = 0; t < 3; t++)
            if (a.init.X[t]) b += a.init.b[t] * fx[t][v];
        } else {
          b = fx[0][v];
        }
        a.OUT[v * a.N + fi] = b + g[v];
      }
    } else {
      const int gs = c < 0 ? c + 4 : c - n + 4;  // 0..7
      for (int v = 0; v < 4; v++) a.GH[(gp * 8 + gs) * 4 + v] = lwc[v];
    }
  };

  if (pvalid && la < 5) {
    Fetch f;
    fetch(kmin - 3 + la, f);
    commit(kmin - 3 + la, f);
  }
  Fetch nf;
  if (pvalid && kmin + 2 + la <= kmax + 2) fetch(kmin + 2 + la, nf);
  for (int k0 = kmin; k0 - 3 < e1; k0 += CA) {
    if (pvalid && k0 + 2 + la <= kmax + 2) commit(k0 + 2 + la, nf);
    // this chunk's finalised cell: its base terms, loaded now
    const int cf = k0 - 3 + la;
    const bool fin = pvalid && cf >= e0 && cf < e1;
    const bool fin_in = fin && (ring || (cf >= 0 && cf < n));
    const long fi = base + (long)cf * sa;
    double fx[3][4];
    if (fin_in) load_base(fi, fx);
    // the next chunk's new ring cell
    if (pvalid && k0 + CA - 3 < e1 && k0 + CA + 2 + la <= kmax + 2) fetch(k0 + CA + 2 + la, nf);
    const int k = k0 + la;
    const bool act = pvalid && k <= kmax;
    // REC: this chunk's interface record, in flight across the barrier
    double rq[8];
    unsigned rbits = 0u;
    if (REC && act) {
      const int kr = ring ? wrap(k) : k;  // a ring interface's record: its copy inside the pencil
      const long ri = cfd::rec_index<AX>(gp, kr + 1, n);
#pragma unroll
      for (int v = 0; v < 8; v++) rq[v] = a.rec[v * a.rec_plane + ri];
      rbits = reinterpret_cast<const unsigned*>(a.rec + 8 * a.rec_plane)[ri];
    }
    __syncthreads();
    double lw[6][4];
    for (int j = 0; j < 6; j++)
      for (int v = 0; v < 4; v++) lw[j][v] = 0.0;
    if (act) {
      double lG[4];
      for (int v = 0; v < 4; v++) lG[v] = (w0 * (M(k - 1, v) - M(k, v)) + w2 * (M(k - 2, v) - M(k + 1, v))) * ih;
      const Cells cells(wc, k, lp);
      cfdadj::IfaceFwd f;
      if (REC) {
        cfdadj::Roe2<AX> es;
        cfdadj::iface_from_record<AX>(cells, rq, rbits, f, es);
        cfdadj::iface_backward<AX, IP>(cells, f, lG, lw, es);
      } else {
        iface_forward_exact<AX, IP, SCH>(cells, f, a.ct, a.rc);
        cfdadj::iface_backward<AX, IP>(cells, f, lG, lw);
      }
    }
    const int par = ((k0 - kmin) / CA) & 1;  // this chunk's staging buffer
    if (act) {
      double* const st_cur = st + par * 24 * CA * CP;
#pragma unroll
      for (int j = 0; j < 6; j++)
#pragma unroll
        for (int v = 0; v < 4; v++) st_cur[sidx(j * 4 + v, la, lp)] = lw[j][v];
    }
    __syncthreads();
    if (fin) {
      double lwc[4] = {0.0, 0.0, 0.0, 0.0};
      gather(cf, k0, lwc);
      finish(cf, fin_in, fi, fx, lwc);
    } else if (FULL && k0 == 0 && pvalid && la < 5) {
      double lwc[4] = {0.0, 0.0, 0.0, 0.0};
      gather(cf, k0, lwc);
#pragma unroll
      for (int v = 0; v < 4; v++) carry[(la * 4 + v) * CP + lp] = lwc[v];
    }
    __syncthreads();
  }
  if (FULL) {
    // cell n+2 (node flux of cell n+1), then cells n-3 .. n+1 (n-3 .. n-1, 0,
    // 1): the last chunk's staged interfaces plus the carried first-chunk sums
    if (pvalid && la == 0) {
      Fetch f;
      fetch(n + 2, f);
      commit(n + 2, f);
    }
    __syncthreads();
    if (pvalid && la < 5) {
      const int c = n - 3 + la, cr = wrap(c);
      const int k0 = ((n - 1) / CA) * CA;  // the last chunk
      double lwc[4];
#pragma unroll
      for (int v = 0; v < 4; v++) lwc[v] = carry[(la * 4 + v) * CP + lp];
      gather(c, k0, lwc);
      const long fi = base + (long)cr * sa;
      double fx[3][4];
      load_base(fi, fx);
      finish(c, true, fi, fx, lwc);
    }
  }
}

// Ghost cotangents of the y sweep onto their source cells (the x sweep's are
// folded into the y sweep's cell phase).  One thread per (pencil, ghost slot):
// the first ghost (in slot order) of each source cell sums every ghost that
// copies that cell, in slot order, and applies the transpose of consToPrim
// once -- one writer per cell, deterministic sums.
__global__ __launch_bounds__(256) void k_ghost_fold_y(VjpArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long gp = t >> 3;
  const int g = (int)(t & 7);
  const int nx = a.n[0], n = a.n[1];
  if (gp >= nx) return;
  auto src_of = [&](int q) {
    return cfdadj::ghost_source(q < 4 ? a.bc_lo : a.bc_hi, q < 4 ? q - 4 : n + q - 4, n);
  };
  const cfdadj::GhostSrc gs = src_of(g);
  if (gs.fixed) return;
  for (int q = 0; q < g; q++) {
    const cfdadj::GhostSrc o = src_of(q);
    if (!o.fixed && o.src == gs.src) return;
  }
  // the source cell's state and output first: their loads overlap the ghost reads
  const long i = gp + (long)gs.src * nx;
  double u[4], o4[4], gr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int v = 0; v < 4; v++) {
    u[v] = a.U[v * a.N + i];
    o4[v] = a.OUT[v * a.N + i];
  }
  double l[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = g; q < 8; q++) {
    const cfdadj::GhostSrc o = src_of(q);
    if (o.fixed || o.src != gs.src) continue;
    for (int v = 0; v < 4; v++) {
      const double x = a.GH[(gp * 8 + q) * 4 + v];
      l[v] += v == o.flip ? -x : x;
    }
  }
  cfdadj::c2p_vjp(u, l, gr);
#pragma unroll
  for (int v = 0; v < 4; v++) a.OUT[v * a.N + i] = o4[v] + gr[v];
}

__global__ void k_aos_to_soa4(const double* __restrict__ in, double* __restrict__ out, long N) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    for (int v = 0; v < 4; v++) out[v * N + i] = in[i * 4 + v];
}
__global__ void k_soa_to_aos4(const double* __restrict__ in, double* __restrict__ out, long N) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N)
    for (int v = 0; v < 4; v++) out[i * 4 + v] = in[v * N + i];
}
__global__ void k_put_dt(double* scal, double dt) { scal[0] = dt; }

}  // namespace

struct AdjointStore {
  size_t state_bytes = 0;
  double *LA = nullptr, *LB = nullptr, *LC = nullptr;  // cotangent buffers
  double* GH[2] = {nullptr, nullptr};                  // ghost cotangents per direction
  double* bcprim = nullptr;                            // Dirichlet ghost states [dir][side][4]
  double* stage = nullptr;                             // segment stage states (3K+1)
  int stage_cap = 0;                                   // states the stage buffer holds
  std::vector<double*> ckpt;                           // checkpointed states
  int nsteps = 0, every = 0;
  bool last_in_store = false;  // the stage store holds the last segment (written by the forward)
  double dt = 0.0;
  int vjp_seg[2] = {0, 0};
  bool vjp_full[2] = {false, false};  // ring direction marched one pencil per workgroup lane group
  // interface records of the segment's stage states (recording sweeps), one
  // block per stage state q = 3i (S_i), 3i+1 (u1_i), 3i+2 (u2_i): direction 0
  // then 1, each 8 planes of rec_plane[ax] doubles and the branch words
  double* rec = nullptr;
  int rec_cap = 0;            // stage states the record buffer holds
  long rec_plane[2] = {0, 0};
  size_t rec_off[2] = {0, 0};  // doubles from the block start
  size_t rec_block = 0;        // doubles per stage state
  double* rec_at(int q, int ax) const { return rec + (size_t)q * rec_block + rec_off[ax]; }
};

#define AHIP(expr)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return cfd_fail(CFD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void cfd_adjoint_release(cfd_handle* h) {
  AdjointStore* s = h->adj;
  if (!s) return;
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(s->LA);
  (void)hipFree(s->LB);
  (void)hipFree(s->LC);
  (void)hipFree(s->GH[0]);
  (void)hipFree(s->bcprim);
  (void)hipFree(s->GH[1]);
  (void)hipFree(s->stage);
  (void)hipFree(s->rec);
  for (double* p : s->ckpt) (void)hipFree(p);
  delete s;
  h->adj = nullptr;
}

static int adj_supported(cfd_handle* h) {
  const bool teno_family = h->sch == cfd::SCH_TENO || h->sch == cfd::SCH_TENOS || h->sch == cfd::SCH_TENOA ||
                           h->sch == cfd::SCH_TENOSA || (h->sch == cfd::SCH_WHF && (int)h->ct != 15);
  if (h->fma)
    return cfd_fail(CFD_E_UNSUPP, "the adjoint differentiates the exact FP64 arithmetic (not CFD_PRECISION_FP64_FMA)");
  if (h->D != 2 || h->fp32 || !h->cw || h->flx != 0 || (h->sch != cfd::SCH_JS && !teno_family))
    return cfd_fail(CFD_E_UNSUPP,
                    "the adjoint covers 2D FP64 characteristic-wise WCNS5 (WENO-JS) or TENO-family schemes "
                    "(TENO5, TENO-S, TENO-A, the whf_* variants) with HLLC flux");
  if (h->cfg.gravity[0] != 0.0 || h->cfg.gravity[1] != 0.0 || h->cfg.gravity[2] != 0.0)
    return cfd_fail(CFD_E_UNSUPP, "the adjoint has no source term");
  for (int f = 0; f < 4; f++)
    if (h->cfg.bc[f] == CFD_BC_DMR_TOP || h->cfg.bc[f] == CFD_BC_DMR_WALL)
      return cfd_fail(CFD_E_UNSUPP, "the adjoint has no time- or position-dependent boundaries");
  return 0;
}

// Segment length of the VJP sweep along a direction of n cells.  A segment of
// L cells runs interfaces over L + 5 cells (L + 9 at a pencil end, which also
// finalises the ghosts; L + 5 everywhere on a periodic ring), in chunks of CA; lengths m*CA - 5 keep every chunk of
// an interior segment full.  Cost = lane slots / the fraction of the last
// round of resident workgroups that is busy (the chip holds 256 * occupancy
// workgroups of 256 lanes at once).
static int vjp_plan_seg(int n, long pencil_blocks, int ca, int cp, bool ring) {
  const long resident = 256L * CFD_ADJ_OCC;
  double best = 0.0;
  int best_seg = n;
  for (int m = 1; m <= 4096; m++) {
    int seg = m * ca - 5;
    if (seg < 4) continue;
    const bool last = seg >= n;
    if (last) seg = n;
    long chunks = 0;
    for (int c0 = 0; c0 < n; c0 += seg) {
      const int len = std::min(seg, n - c0);
      const int extra = 5 + (c0 == 0 && !ring ? 2 : 0) + (c0 + len == n && !ring ? 2 : 0);
      chunks += (len + extra + ca - 1) / ca;
    }
    const long blocks = pencil_blocks * ((n + seg - 1) / seg);
    const long rounds = (blocks + resident - 1) / resident;
    const double busy = (double)blocks / (double)(rounds * resident);
    const double cost = (double)pencil_blocks * chunks * ca * cp / busy;
    if (best == 0.0 || cost < best) {
      best = cost;
      best_seg = seg;
    }
    if (last) break;
  }
  return best_seg;
}

static int adj_ensure(cfd_handle* h) {
  if (int rc = adj_supported(h)) return rc;
  if (h->cfg.n[0] < 8 || h->cfg.n[1] < 8) return cfd_fail(CFD_E_UNSUPP, "the adjoint needs at least 8 cells per direction");
  if (h->adj) return 0;
  AdjointStore* s = new AdjointStore();
  h->adj = s;
  s->state_bytes = sizeof(double) * h->N * 4;
  const long np = std::max(h->cfg.n[0], h->cfg.n[1]);
  AHIP(hipMalloc(&s->LA, s->state_bytes));
  AHIP(hipMalloc(&s->LB, s->state_bytes));
  AHIP(hipMalloc(&s->LC, s->state_bytes));
  for (int ax = 0; ax < 2; ax++) AHIP(hipMalloc(&s->GH[ax], sizeof(double) * np * 8 * 4));
  {
    double bp[16];
    for (int f = 0; f < 4; f++)
      for (int v = 0; v < 4; v++) bp[4 * f + v] = h->cfg.bc_prim[f][v];
    AHIP(hipMalloc(&s->bcprim, sizeof(bp)));
    AHIP(hipMemcpy(s->bcprim, bp, sizeof(bp), hipMemcpyHostToDevice));
  }
  for (int ax = 0; ax < 2; ax++) {
    const long pencils = ax == 0 ? h->cfg.n[1] : h->cfg.n[0];
    const int ca = ax == 0 ? VjpTile<0>::CA : VjpTile<1>::CA;
    const int cp = ax == 0 ? VjpTile<0>::CP : VjpTile<1>::CP;
    const bool ring = h->cfg.bc[2 * ax] == CFD_BC_PERIODIC && h->cfg.bc[2 * ax + 1] == CFD_BC_PERIODIC;
    const int n = h->cfg.n[ax];
    s->vjp_seg[ax] = vjp_plan_seg(n, (pencils + cp - 1) / cp, ca, cp, ring);
    // Whole ring pencils (no interface evaluated twice) where that takes no
    // more rounds x chunks of resident workgroups than the segments do.
    const char* env = getenv("CFD_ADJ_FULL");  // 0: never, 1: wherever allowed
    if (ring && n >= 8 && (!env || env[0] != '0')) {
      const long resident = 256L * CFD_ADJ_OCC;
      const long seg = s->vjp_seg[ax];
      long seg_chunks = 0;  // the longest workgroup's chunks
      for (int c0 = 0; c0 < n; c0 += (int)seg) seg_chunks = std::max(seg_chunks, (std::min<long>(seg, n - c0) + 5 + ca - 1) / ca);
      const long seg_blocks = (pencils + cp - 1) / cp * ((n + seg - 1) / seg);
      const long seg_cost = (seg_blocks + resident - 1) / resident * seg_chunks;
      const long full_blocks = (pencils + kFullCP - 1) / kFullCP;
      const long full_cost = (full_blocks + resident - 1) / resident * ((n + kFullCA - 1) / kFullCA);
      s->vjp_full[ax] = (env && env[0] == '1') || full_cost < seg_cost;
    }
  }
  size_t off = 0;
  for (int ax = 0; ax < 2; ax++) {
    const long P = cfd::rec_plane_of(h->cfg.n, ax);
    s->rec_plane[ax] = P;
    s->rec_off[ax] = off;
    off += (size_t)(8 * P + (P + 1) / 2 + 63) / 64 * 64;  // 8 planes + the words, 512-byte aligned
  }
  s->rec_block = off;
  return 0;
}

// the record pointers of stage state q (null: no records)
struct RecPair {
  void* p[2];
};
static RecPair rec_of(const AdjointStore* s, int q) {
  if (!s->rec) return {{nullptr, nullptr}};
  return {{s->rec_at(q, 0), s->rec_at(q, 1)}};
}

// The VJP kernel of a scheme: WENO-JS's reverse pass, or the TENO family's
// (the flag's stencil, the flag from the scheme's own cut-off).
template <int AX, int CA, int CP, bool REC, bool FULL>
static void launch_vjp_rec(int sch, dim3 grid, hipStream_t s, const VjpArgs& a) {
  using cfdadj::JS;
  using cfdadj::Teno;
  const dim3 blk(CA * CP);
  switch (sch) {
  case cfd::SCH_JS: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, JS, cfd::SCH_JS, REC, FULL>), grid, blk, 0, s, a); break;
  case cfd::SCH_TENO: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, Teno, cfd::SCH_TENO, REC, FULL>), grid, blk, 0, s, a); break;
  case cfd::SCH_TENOS: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, Teno, cfd::SCH_TENOS, REC, FULL>), grid, blk, 0, s, a); break;
  case cfd::SCH_TENOA: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, Teno, cfd::SCH_TENOA, REC, FULL>), grid, blk, 0, s, a); break;
  case cfd::SCH_TENOSA: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, Teno, cfd::SCH_TENOSA, REC, FULL>), grid, blk, 0, s, a); break;
  default: hipLaunchKernelGGL((k_rhs_vjp<AX, CA, CP, Teno, cfd::SCH_WHF, REC, FULL>), grid, blk, 0, s, a); break;
  }
}
template <int AX>
static void launch_vjp(int sch, bool full, long pencils, unsigned nseg, hipStream_t s, const VjpArgs& a) {
  if (full) {
    const dim3 grid((unsigned)((pencils + kFullCP - 1) / kFullCP), 1);
    if (a.rec)
      launch_vjp_rec<AX, kFullCA, kFullCP, true, true>(sch, grid, s, a);
    else
      launch_vjp_rec<AX, kFullCA, kFullCP, false, true>(sch, grid, s, a);
    return;
  }
  constexpr int CA = VjpTile<AX>::CA, CP = VjpTile<AX>::CP;
  const dim3 grid((unsigned)((pencils + CP - 1) / CP), nseg);
  if (a.rec)
    launch_vjp_rec<AX, CA, CP, true, false>(sch, grid, s, a);
  else
    launch_vjp_rec<AX, CA, CP, false, false>(sch, grid, s, a);
}

// OUT = base + coef * (d rhs / d U)^T MU at the state U, both directions;
// base = init's combination (null init: the current OUT).
static int adj_vjp(cfd_handle* h, const double* U, const double* MU, double coef, double* OUT,
                   const VjpInit* init, const RecPair& rec) {
  AdjointStore* s = h->adj;
  for (int ax = 0; ax < 2; ax++) {
    VjpArgs a;
    memset(&a, 0, sizeof(a));
    a.rec = static_cast<const double*>(rec.p[ax]);
    a.rec_plane = s->rec_plane[ax];
    a.U = U;
    a.MU = MU;
    a.OUT = OUT;
    a.GH = s->GH[ax];
    a.GHX = s->GH[0];
    a.gx_src = 0u;
    a.gx_flip = 0u;
    const bool xring = h->cfg.bc[0] == CFD_BC_PERIODIC && h->cfg.bc[1] == CFD_BC_PERIODIC;
    for (int q = 0; q < 8; q++) {
      const int nx = h->cfg.n[0];
      const cfdadj::GhostSrc o =
          cfdadj::ghost_source(q < 4 ? h->cfg.bc[0] : h->cfg.bc[1], q < 4 ? q - 4 : nx + q - 4, nx);
      // 15: no source (a fixed state, or no x ghosts at all on an x ring)
      const unsigned code = o.fixed || xring ? 15u : (unsigned)(o.src < 4 ? o.src : o.src - nx + 8);
      a.gx_src |= code << (4 * q);
      a.gx_flip |= (unsigned)(o.flip > 0 ? o.flip : 0) << (2 * q);
    }
    a.coef = coef;
    a.ct = h->ct;
    a.rc = &h->aux[ax].rc;
    a.use_init = ax == 0 && init != nullptr;
    if (a.use_init) a.init = *init;
    a.N = h->N;
    a.n[0] = h->cfg.n[0];
    a.n[1] = h->cfg.n[1];
    a.seg_len = s->vjp_seg[ax];
    a.h = h->cfg.h[ax];
    a.bc_lo = h->cfg.bc[2 * ax];
    a.bc_hi = h->cfg.bc[2 * ax + 1];
    a.ring = a.bc_lo == CFD_BC_PERIODIC && a.bc_hi == CFD_BC_PERIODIC;
    a.bc_prim = s->bcprim + 8 * ax;
    const int n = h->cfg.n[ax];
    const long pencils = ax == 0 ? h->cfg.n[1] : h->cfg.n[0];
    const bool full = s->vjp_full[ax];
    if (full) a.seg_len = n;
    const unsigned nseg = (unsigned)((n + a.seg_len - 1) / a.seg_len);
    if (ax == 0) {
      launch_vjp<0>(h->sch, full, pencils, nseg, h->stream, a);
    } else {
      launch_vjp<1>(h->sch, full, pencils, nseg, h->stream, a);
      if (!a.ring)
        hipLaunchKernelGGL(k_ghost_fold_y, dim3((unsigned)((pencils * 8 + 255) / 256)), dim3(256), 0, h->stream, a);
    }
    AHIP(hipGetLastError());
  }
  return 0;
}

// One RK3 step backwards: s->LA holds l3 on entry and l0 on exit; q0 is the
// stage-state index of u0 in the record buffer (-1: no records).
static int adj_step_back(cfd_handle* h, const double* u0, const double* u1, const double* u2, double dt, int q0) {
  AdjointStore* s = h->adj;
  const RecPair none = {{nullptr, nullptr}};
  int rc;
  const VjpInit i2 = {{s->LA, nullptr, nullptr}, {2.0 / 3.0, 0.0, 0.0}};
  if ((rc = adj_vjp(h, u2, s->LA, -2.0 / 3.0 * dt, s->LB, &i2, q0 < 0 ? none : rec_of(s, q0 + 2)))) return rc;
  const VjpInit i1 = {{s->LB, nullptr, nullptr}, {0.25, 0.0, 0.0}};
  if ((rc = adj_vjp(h, u1, s->LB, -0.25 * dt, s->LC, &i1, q0 < 0 ? none : rec_of(s, q0 + 1)))) return rc;
  const VjpInit i0 = {{s->LA, s->LB, s->LC}, {1.0 / 3.0, 0.75, 1.0}};
  if ((rc = adj_vjp(h, u0, s->LC, -dt, s->LA, &i0, q0 < 0 ? none : rec_of(s, q0)))) return rc;
  return 0;
}

extern "C" {

int cfd_rhs_vjp(cfd_handle* h, const double* cot, double* grad) {
  if (!h || !cot || !grad) return cfd_fail(CFD_E_ARG, "null argument");
  if (int rc = adj_ensure(h)) return rc;
  AdjointStore* s = h->adj;
  const unsigned nb = (unsigned)((h->N + 255) / 256);
  AHIP(hipMemcpyAsync(s->LC, cot, s->state_bytes, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_aos_to_soa4, dim3(nb), dim3(256), 0, h->stream, s->LC, s->LA, h->N);
  const VjpInit zero = {{nullptr, nullptr, nullptr}, {0.0, 0.0, 0.0}};
  if (int rc = adj_vjp(h, h->U0, s->LA, 1.0, s->LB, &zero, {{nullptr, nullptr}})) return rc;
  hipLaunchKernelGGL(k_soa_to_aos4, dim3(nb), dim3(256), 0, h->stream, s->LB, s->LC, h->N);
  AHIP(hipGetLastError());
  AHIP(hipMemcpyAsync(grad, s->LC, s->state_bytes, hipMemcpyDeviceToHost, h->stream));
  AHIP(hipStreamSynchronize(h->stream));
  return CFD_OK;
}

int cfd_adjoint_forward(cfd_handle* h, double dt, int nsteps, int checkpoint_every) {
  if (!h) return cfd_fail(CFD_E_ARG, "null handle");
  if (nsteps < 1 || checkpoint_every < 0 || !(dt > 0.0))
    return cfd_fail(CFD_E_ARG, "need nsteps >= 1, checkpoint_every >= 0 and dt > 0");
  if (int rc = adj_ensure(h)) return rc;
  AdjointStore* s = h->adj;
  s->nsteps = 0;  // valid again only once this forward completes
  s->last_in_store = false;
  const bool keep_all = checkpoint_every == 0;
  const int nck = keep_all ? 0 : (nsteps + checkpoint_every - 1) / checkpoint_every;
  while ((int)s->ckpt.size() > nck) {
    AHIP(hipFree(s->ckpt.back()));
    s->ckpt.pop_back();
  }
  while ((int)s->ckpt.size() < nck) {
    double* p = nullptr;
    AHIP(hipMalloc(&p, s->state_bytes));
    s->ckpt.push_back(p);
  }
  // stage store: S_i at 3i, u1_i at 3i+1, u2_i at 3i+2 for the steps of one
  // segment (all steps when every stage state is kept), plus the end state
  const int seg = keep_all ? nsteps : std::min(checkpoint_every, nsteps);
  const int need = 3 * seg + 1;
  if (s->stage_cap < need) {
    if (s->stage) AHIP(hipFree(s->stage));
    s->stage = nullptr;
    s->stage_cap = 0;
    AHIP(hipMalloc(&s->stage, s->state_bytes * need));
    s->stage_cap = need;
  }
  // interface records of one segment's stage states (not when every stage
  // state is kept: the VJPs then re-evaluate their interfaces)
  if (!keep_all && s->rec_cap < 3 * seg) {
    if (s->rec) AHIP(hipFree(s->rec));
    s->rec = nullptr;
    s->rec_cap = 0;
    AHIP(hipMalloc(&s->rec, sizeof(double) * s->rec_block * 3 * seg));
    s->rec_cap = 3 * seg;
  }
  s->every = checkpoint_every;
  s->dt = dt;
  hipLaunchKernelGGL(k_put_dt, dim3(1), dim3(1), 0, h->stream, h->scal, dt);
  int rc;
  if (keep_all) {
    const size_t nd = (size_t)h->N * 4;
    double* st = s->stage;
    AHIP(hipMemcpyAsync(st, h->U0, s->state_bytes, hipMemcpyDeviceToDevice, h->stream));
    for (int i = 0; i < nsteps; i++) {
      double *S0 = st + (size_t)(3 * i) * nd, *S1 = S0 + nd, *S2 = S1 + nd, *S3 = S2 + nd;
      if ((rc = rk_stage(h, 1, S0, S0, S1, nullptr))) return rc;
      if ((rc = rk_stage(h, 2, S1, S0, S2, nullptr))) return rc;
      if ((rc = rk_stage(h, 3, S2, S0, S3, nullptr))) return rc;
    }
    AHIP(hipMemcpyAsync(h->U0, st + (size_t)(3 * nsteps) * nd, s->state_bytes, hipMemcpyDeviceToDevice,
                        h->stream));
  } else {
    // the last segment runs through the stage store, so that the backward
    // starts on it without recomputing it
    const int last = (nck - 1) * checkpoint_every;
    for (int st = 0; st < last; st++) {
      if (st % checkpoint_every == 0)
        AHIP(hipMemcpyAsync(s->ckpt[st / checkpoint_every], h->U0, s->state_bytes, hipMemcpyDeviceToDevice,
                            h->stream));
      if ((rc = rk_stage(h, 1, h->U0, h->U0, h->Ua, nullptr))) return rc;
      if ((rc = rk_stage(h, 2, h->Ua, h->U0, h->Ub, nullptr))) return rc;
      if ((rc = rk_stage(h, 3, h->Ub, h->U0, h->U0, nullptr))) return rc;
    }
    const size_t nd = (size_t)h->N * 4;
    double* st = s->stage;
    AHIP(hipMemcpyAsync(s->ckpt[nck - 1], h->U0, s->state_bytes, hipMemcpyDeviceToDevice, h->stream));
    AHIP(hipMemcpyAsync(st, h->U0, s->state_bytes, hipMemcpyDeviceToDevice, h->stream));
    for (int i = 0; i < nsteps - last; i++) {
      double *S0 = st + (size_t)(3 * i) * nd, *S1 = S0 + nd, *S2 = S1 + nd, *S3 = S2 + nd;
      if ((rc = rk_stage(h, 1, S0, S0, S1, nullptr, rec_of(s, 3 * i).p))) return rc;
      if ((rc = rk_stage(h, 2, S1, S0, S2, nullptr, rec_of(s, 3 * i + 1).p))) return rc;
      if ((rc = rk_stage(h, 3, S2, S0, S3, nullptr, rec_of(s, 3 * i + 2).p))) return rc;
    }
    AHIP(hipMemcpyAsync(h->U0, st + (size_t)(3 * (nsteps - last)) * nd, s->state_bytes, hipMemcpyDeviceToDevice,
                        h->stream));
  }
  AHIP(hipGetLastError());
  AHIP(hipStreamSynchronize(h->stream));
  s->nsteps = nsteps;
  s->last_in_store = !keep_all;
  return CFD_OK;
}

int cfd_adjoint_backward(cfd_handle* h, const double* cot_final, double* grad_initial) {
  if (!h || !cot_final || !grad_initial) return cfd_fail(CFD_E_ARG, "null argument");
  if (!h->adj || h->adj->nsteps < 1) return cfd_fail(CFD_E_ARG, "cfd_adjoint_forward has not run");
  AdjointStore* s = h->adj;
  const unsigned nb = (unsigned)((h->N + 255) / 256);
  const size_t nd = (size_t)h->N * 4;
  AHIP(hipMemcpyAsync(s->LB, cot_final, s->state_bytes, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_aos_to_soa4, dim3(nb), dim3(256), 0, h->stream, s->LB, s->LA, h->N);
  auto S = [&](int i) { return s->stage + (size_t)(3 * i) * nd; };
  auto U1 = [&](int i) { return s->stage + (size_t)(3 * i + 1) * nd; };
  auto U2 = [&](int i) { return s->stage + (size_t)(3 * i + 2) * nd; };
  int rc;
  if (s->every == 0) {
    for (int i = s->nsteps - 1; i >= 0; i--)
      if ((rc = adj_step_back(h, S(i), U1(i), U2(i), s->dt, -1))) return rc;
  } else {
    hipLaunchKernelGGL(k_put_dt, dim3(1), dim3(1), 0, h->stream, h->scal, s->dt);
    const int nck = (int)s->ckpt.size();
    for (int m = nck - 1; m >= 0; m--) {
      const int first = m * s->every;
      const int len = std::min(s->every, s->nsteps - first);
      if (!(m == nck - 1 && s->last_in_store)) {
        s->last_in_store = false;  // the store now holds another segment
        AHIP(hipMemcpyAsync(S(0), s->ckpt[m], s->state_bytes, hipMemcpyDeviceToDevice, h->stream));
        // every stage records its interfaces; the last step's third stage
        // runs for its record only (its state S(len) is not used)
        for (int i = 0; i < len; i++) {
          if ((rc = rk_stage(h, 1, S(i), S(i), U1(i), nullptr, rec_of(s, 3 * i).p))) return rc;
          if ((rc = rk_stage(h, 2, U1(i), S(i), U2(i), nullptr, rec_of(s, 3 * i + 1).p))) return rc;
          if ((rc = rk_stage(h, 3, U2(i), S(i), S(i + 1), nullptr, rec_of(s, 3 * i + 2).p))) return rc;
        }
      }
      for (int i = len - 1; i >= 0; i--)
        if ((rc = adj_step_back(h, S(i), U1(i), U2(i), s->dt, 3 * i))) return rc;
    }
  }
  hipLaunchKernelGGL(k_soa_to_aos4, dim3(nb), dim3(256), 0, h->stream, s->LA, s->LB, h->N);
  AHIP(hipGetLastError());
  AHIP(hipMemcpyAsync(grad_initial, s->LB, s->state_bytes, hipMemcpyDeviceToHost, h->stream));
  AHIP(hipStreamSynchronize(h->stream));
  return CFD_OK;
}

}  // extern "C"
