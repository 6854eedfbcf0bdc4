// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY: loaded by the
// test suite, the smoke check and the benchmark's CPU-baseline leg through
// ctypes; never by the product library).
#include "oracle.hpp"
#include "dual.hpp"

#include <omp.h>

using namespace oracle;

extern "C" {

struct oracle_grid {
  int dim;
  int n[3];
  double h[3];
  int scheme;    // oracle::Scheme
  int flux;      // oracle::Flux
  int charwise;  // 1 characteristic-wise, 0 component-wise
  int bc[6];     // oracle::Bc per face (xlo, xhi, ylo, yhi, zlo, zhi)
  double bc_prim[6][5];
  double x0;          // cell centres at x0 + (i + 0.5) h[0]
  double t;           // time at the start (boundaries that depend on time)
  double gravity[3];  // acceleration of the gravity source (0: none)
  int variant;        // 0 zone_FD's path, 1 CongFD1.0's Euler path
  double gamma_num, gamma_den;  // the text of the GAMMA macro: num (den = 0) or num/den
};

static Grid to_grid(const oracle_grid* c) {
  Grid g;
  g.dim = c->dim;
  for (int i = 0; i < 3; i++) { g.n[i] = c->n[i]; g.h[i] = c->h[i]; }
  g.scheme = c->scheme;
  g.flux = c->flux;
  g.charwise = c->charwise;
  for (int i = 0; i < 6; i++) {
    g.bc[i] = c->bc[i];
    for (int v = 0; v < 5; v++) g.bc_prim[i][v] = c->bc_prim[i][v];
  }
  g.x0 = c->x0;
  g.t = c->t;
  for (int i = 0; i < 3; i++) g.gravity[i] = c->gravity[i];
  g.variant = c->variant;
  if (c->gamma_num != 0.0) set_gas(c->gamma_num, c->gamma_den);
  else set_gas(1.4, 0.0);
  return g;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

// The GAMMA macro's text for the point entry points below (the grid entry
// points take it from their grid): num alone (den = 0) or num/den.
void oracle_set_gas(double num, double den) { set_gas(num, den); }

// rhs (AoS [cell][nv], x fastest) of the state cons; prim receives consToPrim.
int oracle_rhs(const oracle_grid* c, const double* cons, double* prim, double* rhs, int nthreads) {
  Grid g = to_grid(c);
  switch (g.dim) {
  case 1: rhs_eval<double, 1>(g, cons, prim, rhs, nthreads); return 0;
  case 2: rhs_eval<double, 2>(g, cons, prim, rhs, nthreads); return 0;
  case 3: rhs_eval<double, 3>(g, cons, prim, rhs, nthreads); return 0;
  }
  return -1;
}

// nsteps SSP-RK3 steps with a fixed dt (cfl <= 0) or the CFL dt of each step.
// t_out: the time reached, advanced as BlockSolver::RK3_SSP + solve do it
// (from the grid's start time).
int oracle_rk3(const oracle_grid* c, double* cons, double dt, double cfl, int nsteps, int nthreads, double* t_out) {
  Grid g = to_grid(c);
  for (int s = 0; s < nsteps; s++) {
    double d = dt;
    switch (g.dim) {
    case 1: if (cfl > 0) d = cfl_dt<double, 1>(g, cons, cfl, nthreads); g.t = rk3_step<double, 1>(g, cons, d, nthreads); break;
    case 2: if (cfl > 0) d = cfl_dt<double, 2>(g, cons, cfl, nthreads); g.t = rk3_step<double, 2>(g, cons, d, nthreads); break;
    case 3: if (cfl > 0) d = cfl_dt<double, 3>(g, cons, cfl, nthreads); g.t = rk3_step<double, 3>(g, cons, d, nthreads); break;
    default: return -1;
    }
  }
  if (t_out) *t_out = g.t;
  return 0;
}

double oracle_cfl_dt(const oracle_grid* c, const double* cons, double cfl, int nthreads) {
  Grid g = to_grid(c);
  switch (g.dim) {
  case 1: return cfl_dt<double, 1>(g, cons, cfl, nthreads);
  case 2: return cfl_dt<double, 2>(g, cons, cfl, nthreads);
  case 3: return cfl_dt<double, 3>(g, cons, cfl, nthreads);
  }
  return -1;
}

// ---- point kernels, for unit tests -------------------------------------
double oracle_inter5(int scheme, const double* q) { return inter5<double>(scheme, q); }

void oracle_cons_to_prim(int dim, const double* c, double* p) {
  if (dim == 1) cons_to_prim<double, 1>(c, p);
  else if (dim == 2) cons_to_prim<double, 2>(c, p);
  else cons_to_prim<double, 3>(c, p);
}

// Left/right eigenvector matrices (row-major NVxNV) of the Roe state of (pl, pr).
void oracle_eigen(int dim, const double* pl, const double* pr, int ax, double* L, double* R) {
  auto run = [&](auto tag) {
    constexpr int D = decltype(tag)::value;
    Eigen<double, D> e;
    e.build(pl, pr, ax);
    for (int i = 0; i < D + 2; i++)
      for (int j = 0; j < D + 2; j++) { L[i * (D + 2) + j] = e.L[i][j]; R[i * (D + 2) + j] = e.R[i][j]; }
  };
  if (dim == 1) run(std::integral_constant<int, 1>());
  else if (dim == 2) run(std::integral_constant<int, 2>());
  else run(std::integral_constant<int, 3>());
}

// primToChar of prim (through the conservative rebuild) and charToPrim back.
void oracle_prim_char_roundtrip(int dim, const double* pl, const double* pr, int ax, const double* prim,
                                double* chr, double* prim_back) {
  auto run = [&](auto tag) {
    constexpr int D = decltype(tag)::value;
    Eigen<double, D> e;
    e.build(pl, pr, ax);
    double c[5], cb[5];
    prim_to_cons<double, D>(prim, c);
    e.to_char(c, chr);
    e.to_cons(chr, cb);
    eig_cons_to_prim<double, D>(cb, prim_back);
  };
  if (dim == 1) run(std::integral_constant<int, 1>());
  else if (dim == 2) run(std::integral_constant<int, 2>());
  else run(std::integral_constant<int, 3>());
}

void oracle_riemann(int dim, int flux, const double* L, const double* R, int ax, double* res) {
  if (dim == 1) riemann<double, 1>(flux, L, R, ax, res);
  else if (dim == 2) riemann<double, 2>(flux, L, R, ax, res);
  else riemann<double, 3>(flux, L, R, ax, res);
}

void oracle_node_flux(int dim, const double* p, int ax, double* F) {
  if (dim == 1) node_flux<double, 1>(p, ax, F);
  else if (dim == 2) node_flux<double, 2>(p, ax, F);
  else node_flux<double, 3>(p, ax, F);
}

// Interface reconstruction from a 6-cell window of primitive states
// (window[j*nv + v], j = cells i-3..i+2).
void oracle_recon(int dim, int scheme, int charwise, const double* window, int ax, double* outL, double* outR) {
  auto run = [&](auto tag) {
    constexpr int D = decltype(tag)::value;
    constexpr int NV = D + 2;
    double cc[6][NV];
    const double* wp[6];
    const double* wc[6];
    for (int j = 0; j < 6; j++) {
      prim_to_cons<double, D>(&window[j * NV], cc[j]);
      wp[j] = &window[j * NV];
      wc[j] = cc[j];
    }
    recon_interface<double, D>(scheme, charwise != 0, wp, wc, ax, outL, outR);
  };
  if (dim == 1) run(std::integral_constant<int, 1>());
  else if (dim == 2) run(std::integral_constant<int, 2>());
  else run(std::integral_constant<int, 3>());
}

// ---- CongFD1.0 AccuracyTest path -----------------------------------------
// getError (CongFD1.0/AccuracyTest.cpp:14-83) through SpaceDis with
// ACCURACYTEST + MND6: flux at half node i-1/2 is the left-biased
// interpolant reconLprim (CongFD1.0/src/src/sp_recon.cpp:431-439,
// sp_flux.cpp:45-55) and difMND6 (CongFD1.0/src/src/sp_difference.cpp:96-136)
// divides each of its three terms by h.  The caller supplies the cell and
// ghost values (10 ghosts per side, ghost 0 adjacent) and h; the result is the
// derivative estimate at every cell.
void oracle_accuracy_rhs(int scheme, int n, const double* cells, const double* ghostL, const double* ghostR,
                         double h, double* rhs) {
  auto at = [&](int i) -> double {
    if (i < 0) return ghostL[-(i + 1)];
    if (i >= n) return ghostR[i - n];
    return cells[i];
  };
  // flux index i in [-1, n+1]: stored at i+1
  std::vector<double> F(n + 3);
  for (int i = -1; i < n + 2; i++) {
    double q[5] = {at(i - 3), at(i - 2), at(i - 1), at(i), at(i + 1)};
    F[i + 1] = inter5<double>(scheme, q);
  }
  const double w0 = 3.0 / 2.0, w1 = -3.0 / 10.0, w2 = 1.0 / 30.0;
  for (int i = 0; i < n; i++) {
    rhs[i] = 0.0;
    rhs[i] += w1 * (at(i + 1) - at(i - 1)) / h + w0 * (F[i + 2] - F[i + 1]) / h + w2 * (F[i + 3] - F[i]) / h;
  }
}

// ---- forward-mode dual-number build of the RHS and RK3 -----------------------
// The oracle instantiated on oracle::Dual: values identical to oracle_rhs /
// oracle_rk3, plus the directional derivative along the tangent dcons.
static void dual_pack(long n, const double* x, const double* dx, std::vector<Dual>& out) {
  out.resize(n);
  for (long i = 0; i < n; i++) out[i] = Dual(x[i], dx[i]);
}

int oracle_dual_rhs(const oracle_grid* c, const double* cons, const double* dcons, double* rhs, double* drhs,
                    int nthreads) {
  Grid g = to_grid(c);
  const long N = (long)g.n[0] * g.n[1] * g.n[2] * (g.dim + 2);
  std::vector<Dual> u, prim(N), r(N);
  dual_pack(N, cons, dcons, u);
  switch (g.dim) {
  case 1: rhs_eval<Dual, 1>(g, u.data(), prim.data(), r.data(), nthreads); break;
  case 2: rhs_eval<Dual, 2>(g, u.data(), prim.data(), r.data(), nthreads); break;
  case 3: rhs_eval<Dual, 3>(g, u.data(), prim.data(), r.data(), nthreads); break;
  default: return -1;
  }
  for (long i = 0; i < N; i++) {
    rhs[i] = r[i].v;
    drhs[i] = r[i].d;
  }
  return 0;
}

// nsteps SSP-RK3 steps with the fixed dt, state and tangent advanced in place.
int oracle_dual_rk3(const oracle_grid* c, double* cons, double* dcons, double dt, int nsteps, int nthreads) {
  Grid g = to_grid(c);
  const long N = (long)g.n[0] * g.n[1] * g.n[2] * (g.dim + 2);
  std::vector<Dual> u;
  dual_pack(N, cons, dcons, u);
  for (int s = 0; s < nsteps; s++) {
    switch (g.dim) {
    case 1: g.t = rk3_step<Dual, 1>(g, u.data(), dt, nthreads); break;
    case 2: g.t = rk3_step<Dual, 2>(g, u.data(), dt, nthreads); break;
    case 3: g.t = rk3_step<Dual, 3>(g, u.data(), dt, nthreads); break;
    default: return -1;
    }
  }
  for (long i = 0; i < N; i++) {
    cons[i] = u[i].v;
    dcons[i] = u[i].d;
  }
  return 0;
}

}  // extern "C"
