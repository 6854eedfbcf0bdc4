// Host build of the adjoint point kernels (cfd_inwsl_amd/csrc/adjoint_point.hpp)
// with C entry points, for CPU tests against finite differences and the
// dual-number oracle (test infrastructure only).
#include "adjoint_point.hpp"

#include <vector>

using namespace cfdadj;

namespace {
// the interface pass with WENO-JS (teno = 0) or TENO5 with cut-off ct
template <int AX>
void iface(int teno, double ct, const double w[6][4], double* F, double lw[6][4], const double* lF) {
  if (teno) {
    Teno t;
    t.ct = ct;
    interface<AX, Teno>(w, F, lw, lF, t);
  } else {
    interface<AX>(w, F, lw, lF);
  }
}
}  // namespace

extern "C" {

void adj_interface_fwd_s(int ax, const double* w, double* F, int teno, double ct) {
  double ww[6][4];
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) ww[j][v] = w[4 * j + v];
  if (ax == 0) iface<0>(teno, ct, ww, F, nullptr, nullptr);
  else iface<1>(teno, ct, ww, F, nullptr, nullptr);
}

void adj_interface_vjp_s(int ax, const double* w, const double* lF, double* lw, int teno, double ct) {
  double ww[6][4], lww[6][4];
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) {
      ww[j][v] = w[4 * j + v];
      lww[j][v] = 0.0;
    }
  double F[4];
  if (ax == 0) iface<0>(teno, ct, ww, F, lww, lF);
  else iface<1>(teno, ct, ww, F, lww, lF);
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) lw[4 * j + v] = lww[j][v];
}

void adj_interface_fwd(int ax, const double* w, double* F) {
  double ww[6][4];
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) ww[j][v] = w[4 * j + v];
  if (ax == 0) interface<0>(ww, F);
  else interface<1>(ww, F);
}

void adj_interface_vjp(int ax, const double* w, const double* lF, double* lw) {
  double ww[6][4], lww[6][4];
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) {
      ww[j][v] = w[4 * j + v];
      lww[j][v] = 0.0;
    }
  double F[4];
  if (ax == 0) interface<0>(ww, F, lww, lF);
  else interface<1>(ww, F, lww, lF);
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) lw[4 * j + v] = lww[j][v];
}

void adj_hllc_fwd(int ax, const double* L, const double* R, double* F) {
  if (ax == 0) hllc<0>(L, R, F);
  else hllc<1>(L, R, F);
}

void adj_hllc_vjp(int ax, const double* L, const double* R, const double* lF, double* lL, double* lR) {
  double F[4];
  for (int i = 0; i < 4; i++) lL[i] = lR[i] = 0.0;
  if (ax == 0) hllc<0>(L, R, F, lL, lR, lF);
  else hllc<1>(L, R, F, lL, lR, lF);
}

double adj_js_fwd(const double* q) {
  JS j;
  j.fwd(q);
  return j.out;
}

void adj_js_vjp(const double* q, double lam, double* g) {
  JS j;
  j.fwd(q);
  for (int i = 0; i < 5; i++) g[i] = 0.0;
  j.bwd(q, lam, g);
}

void adj_node_flux_vjp(int ax, const double* p, const double* lF, double* gp) {
  for (int i = 0; i < 4; i++) gp[i] = 0.0;
  if (ax == 0) node_flux_vjp<0>(p, lF, gp);
  else node_flux_vjp<1>(p, lF, gp);
}

void adj_node_flux(int ax, const double* p, double* F) {
  if (ax == 0) node_flux<0>(p, F);
  else node_flux<1>(p, F);
}

void adj_c2p_vjp(const double* c, const double* lp, double* gc) {
  for (int i = 0; i < 4; i++) gc[i] = 0.0;
  c2p_vjp(c, lp, gc);
}

void adj_interface_vjp_lean(int ax, const double* w, const double* lF, double* lw) {
  struct Acc {
    const double* w;
    double operator()(int j, int v) const { return w[4 * j + v]; }
    double cons(int j, int v) const {
      double c[4];
      p2c(&w[4 * j], c);
      return c[v];
    }
  } acc{w};
  double lww[6][4] = {};
  if (ax == 0) interface_vjp_lean<0>(acc, lF, lww);
  else interface_vjp_lean<1>(acc, lF, lww);
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) lw[4 * j + v] = lww[j][v];
}

// Host statement of the 2D RHS vector-Jacobian product for characteristic-wise
// WENO-JS (or TENO5, teno != 0) + HLLC + MND6 (the map of oracle::rhs_eval):
// g += (d rhs/d U)^T mu.
// Arrays AoS [cell][4], x fastest; bc per face (xlo, xhi, ylo, yhi);
// bc_prim the Dirichlet ghost states (primitive).  Per direction and pencil:
// rhs_i += (w1 (f_{i+1} - f_{i-1}) + w0 (G_{i+1} - G_i) + w2 (G_{i+2} - G_{i-1})) / h
// with G_k the interface flux between cells k-1 and k (cells k-3 .. k+2),
// so lambda_G_k = (w0 (mu_{k-1} - mu_k) + w2 (mu_{k-2} - mu_{k+1})) / h and
// lambda_f_j = w1 (mu_{j-1} - mu_{j+1}) / h, mu = 0 outside the pencil.
void adj_rhs_vjp_s(const int* n, const double* hgrid, const int* bc, const double* bc_prim, const double* U,
                   const double* mu, double* g, int teno, double ct) {
  const double w0 = 3.0 / 2.0, w1 = -3.0 / 10.0, w2 = 1.0 / 30.0;
  const int nx = n[0], ny = n[1];
  for (int ax = 0; ax < 2; ax++) {
    const int nn = n[ax], np = ax == 0 ? ny : nx;
    const double hh = hgrid[ax];
    std::vector<double> w((nn + 8) * 4), lw((nn + 8) * 4), m((nn + 8) * 4);
    for (int ip = 0; ip < np; ip++) {
      auto cell = [&](int i) -> long { return ax == 0 ? (long)ip * nx + i : (long)i * nx + ip; };
      for (int e = -4; e < nn + 4; e++) {
        double* we = &w[(e + 4) * 4];
        double* me = &m[(e + 4) * 4];
        for (int v = 0; v < 4; v++) lw[(e + 4) * 4 + v] = 0.0;
        if (e >= 0 && e < nn) {
          c2p(&U[cell(e) * 4], we);
          for (int v = 0; v < 4; v++) me[v] = mu[cell(e) * 4 + v];
          continue;
        }
        for (int v = 0; v < 4; v++) me[v] = 0.0;
        const int face = 2 * ax + (e < 0 ? 0 : 1);
        GhostSrc gs = ghost_source(bc[face], e, nn);
        if (gs.fixed) {
          for (int v = 0; v < 4; v++) we[v] = bc_prim[face * 4 + v];
        } else {
          c2p(&U[cell(gs.src) * 4], we);
          if (gs.flip >= 0) we[gs.flip] = -we[gs.flip];
        }
      }
      auto M = [&](int e, int v) { return m[(e + 4) * 4 + v]; };
      for (int k = -1; k <= nn + 1; k++) {
        double lG[4];
        for (int v = 0; v < 4; v++) lG[v] = (w0 * (M(k - 1, v) - M(k, v)) + w2 * (M(k - 2, v) - M(k + 1, v))) / hh;
        double ww[6][4], lww[6][4] = {};
        for (int j = 0; j < 6; j++)
          for (int v = 0; v < 4; v++) ww[j][v] = w[(k - 3 + j + 4) * 4 + v];
        double F[4];
        if (ax == 0) iface<0>(teno, ct, ww, F, lww, lG);
        else iface<1>(teno, ct, ww, F, lww, lG);
        for (int j = 0; j < 6; j++)
          for (int v = 0; v < 4; v++) lw[(k - 3 + j + 4) * 4 + v] += lww[j][v];
      }
      for (int j = -1; j <= nn; j++) {
        double lf[4];
        for (int v = 0; v < 4; v++) lf[v] = w1 * (M(j - 1, v) - M(j + 1, v)) / hh;
        if (ax == 0) node_flux_vjp<0>(&w[(j + 4) * 4], lf, &lw[(j + 4) * 4]);
        else node_flux_vjp<1>(&w[(j + 4) * 4], lf, &lw[(j + 4) * 4]);
      }
      // ghosts back onto their source cells
      for (int e = -4; e < nn + 4; e++) {
        if (e >= 0 && e < nn) continue;
        GhostSrc gs = ghost_source(bc[2 * ax + (e < 0 ? 0 : 1)], e, nn);
        if (gs.fixed) continue;
        for (int v = 0; v < 4; v++) {
          double l = lw[(e + 4) * 4 + v];
          lw[(gs.src + 4) * 4 + v] += v == gs.flip ? -l : l;
        }
      }
      for (int i = 0; i < nn; i++) c2p_vjp(&U[cell(i) * 4], &lw[(i + 4) * 4], &g[cell(i) * 4]);
    }
  }
}

void adj_rhs_vjp(const int* n, const double* hgrid, const int* bc, const double* bc_prim, const double* U,
                 const double* mu, double* g) {
  adj_rhs_vjp_s(n, hgrid, bc, bc_prim, U, mu, g, 0, 0.0);
}

}  // extern "C"
