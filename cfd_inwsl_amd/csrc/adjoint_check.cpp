// Host build of the adjoint point kernels (adjoint_point.hpp) with C entry
// points, for unit tests against finite differences on the CPU.
#include "adjoint_point.hpp"

using namespace cfdadj;

extern "C" {

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

}  // extern "C"
