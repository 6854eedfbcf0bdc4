// Device point kernels of the CongFD right-hand side: 5-point half-node
// interpolation, Roe-averaged characteristic projection, HLLC / Roe interface
// fluxes, node flux and the conservative<->primitive maps.
//
// Every expression keeps the reference's operation order (no FMA contraction:
// the library is compiled with -ffp-contract=off) so that FP64 results equal
// the reference CPU arithmetic bit for bit.  Citations are relative to the
// reference repository (wanghongfei23/CFD_inWSL).
//
// Variables: prim = (rho, u[, v[, w]], p); cons = (rho, rho u[, ...], rho E);
// NV = D + 2.  The face normal of a sweep along axis AX is the unit vector e_AX.
#pragma once
#include <hip/hip_runtime.h>

namespace cfd {

// zone_FD/src/include/macro.hpp:28
#define CFD_GAMMA 1.4

enum SchemeId { SCH_JS = 0, SCH_Z = 1, SCH_TENO = 2, SCH_TENOS = 3 };
enum FluxId { FLX_HLLC = 0, FLX_ROE = 1 };

__device__ __forceinline__ double sq(double a) { return a * a; }
// std::min / std::max argument semantics
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// ---- 5-point interpolation: value at q2 + h/2 ------------------------------
// Smoothness indicators (zone_FD/src/include/interScheme.hpp:20-27).
__device__ __forceinline__ void beta3(double q0, double q1, double q2, double q3, double q4, double& b0, double& b1,
                                      double& b2) {
  b0 = sq(q0 - 2.0 * q1 + q2) + 0.25 * sq(q0 - 4.0 * q1 + 3.0 * q2);
  b1 = sq(q1 - 2.0 * q2 + q3) + 0.25 * sq(q1 - q3);
  b2 = sq(q2 - 2.0 * q3 + q4) + 0.25 * sq(3.0 * q2 - 4.0 * q3 + q4);
}

// TENO flag -> fixed stencil (interScheme.hpp:417-451)
__device__ __forceinline__ double teno_pick(unsigned flag, double q0, double q1, double q2, double q3, double q4) {
  switch (flag) {
  case 0: return 0.0234375 * q0 - 0.15625 * q1 + 0.703125 * q2 + 0.46875 * q3 - 0.0390625 * q4;
  case 1: return -0.0625 * q1 + 0.5625 * q2 + 0.5625 * q3 - 0.0625 * q4;
  case 2:
  case 3: return 0.375 * q2 + 0.75 * q3 - 0.125 * q4;
  case 4: return 0.0625 * q0 - 0.3125 * q1 + 0.9375 * q2 + 0.3125 * q3;
  case 5: return -0.125 * q1 + 0.75 * q2 + 0.375 * q3;
  case 6: return 0.375 * q0 - 1.25 * q1 + 1.875 * q2;
  default: return q2;
  }
}

template <int SCH>
__device__ __forceinline__ double inter5(double q0, double q1, double q2, double q3, double q4, double ct) {
  double b0, b1, b2;
  beta3(q0, q1, q2, q3, q4, b0, b1, b2);
  if (SCH == SCH_JS || SCH == SCH_Z) {
    // weno5_JSchen (interScheme.hpp:17-43) / weno5_Z (:85-114)
    double u0 = 0.375 * q0 - 1.25 * q1 + 1.875 * q2;
    double u1 = -0.125 * q1 + 0.75 * q2 + 0.375 * q3;
    double u2 = 0.375 * q2 + 0.75 * q3 - 0.125 * q4;
    double a0, a1, a2;
    if (SCH == SCH_JS) {
      a0 = 0.0625 / sq(1e-6 + b0);
      a1 = 0.625 / sq(1e-6 + b1);
      a2 = 0.3125 / sq(1e-6 + b2);
    } else {
      double tau = fabs(b2 - b0);
      a0 = 0.0625 * (1.0 + sq(tau / (b0 + 1e-40)));
      a1 = 0.625 * (1.0 + sq(tau / (b1 + 1e-40)));
      a2 = 0.3125 * (1.0 + sq(tau / (b2 + 1e-40)));
    }
    double s = a0 + a1 + a2;
    double r = a0 * u0 + a1 * u1 + a2 * u2;
    return r / s;
  } else if (SCH == SCH_TENO) {
    // Teno5_Z (interScheme.hpp:384-452); ct = 1e-5 (CT4: 1e-4, CT7: 1e-7)
    double tau = fabs(b2 - b0);
    double t0 = 1.0 + tau / (b0 + 1e-40);
    double t1 = 1.0 + tau / (b1 + 1e-40);
    double t2 = 1.0 + tau / (b2 + 1e-40);
    t0 = t0 * t0;
    t1 = t1 * t1;
    t2 = t2 * t2;
    double g0 = t0 * t0 * t0, g1 = t1 * t1 * t1, g2 = t2 * t2 * t2;
    double CT = ct * (g0 + g1 + g2);
    unsigned flag = (g0 < CT ? 1u : 0u) + (g1 < CT ? 2u : 0u) + (g2 < CT ? 4u : 0u);
    return teno_pick(flag, q0, q1, q2, q3, q4);
  } else {
    // Teno5_CongZ, TENO-S (interScheme.hpp:494-585)
    double bm = b0;
    if (b1 < bm) bm = b1;
    if (b2 < bm) bm = b2;
    const double CT = 0.15704178024750198;
    const double CT_1 = 1 - CT;
    double tau = fabs(b2 - b0);
    double rr = CT * tau - CT_1 * bm;
    double ll = tau * bm;
    unsigned flag = (ll < rr * b0 ? 1u : 0u) + (ll < rr * b1 ? 2u : 0u) + (ll < rr * b2 ? 4u : 0u);
    return teno_pick(flag, q0, q1, q2, q3, q4);
  }
}

// ---- conservative <-> primitive ----------------------------------------------
// Equation::consToPrimEuler1D/2D (zone_FD/src/src/equation.cpp:37-99); the 3D
// form appends w*w to the 2D sum.
template <int D> __device__ __forceinline__ void cons_to_prim(const double* c, double* p) {
  double r = c[0];
  double u = c[1] / r;
  if (D == 1) {
    double E = c[2] / r;
    double e = -u * u / 2 + E;
    double RT = (CFD_GAMMA - 1) * e;
    p[0] = r; p[1] = u; p[2] = r * RT;
  } else if (D == 2) {
    double v = c[2] / r;
    double E = c[3] / r;
    double q2 = (u * u + v * v) / 2;
    double e = -q2 + E;
    double RT = (CFD_GAMMA - 1) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = r * RT;
  } else {
    double v = c[2] / r;
    double w = c[3] / r;
    double E = c[4] / r;
    double q2 = (u * u + v * v + w * w) / 2;
    double e = -q2 + E;
    double RT = (CFD_GAMMA - 1) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = w; p[4] = r * RT;
  }
}

// Conservative state rebuilt inside primToChar (eigenSystem.cpp:91-93, 254-256).
template <int D> __device__ __forceinline__ void prim_to_cons(const double* p, double* c) {
  double r = p[0];
  if (D == 1) {
    double ek = (p[1] * p[1]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = p[2] / (CFD_GAMMA - 1) + r * ek;
  } else if (D == 2) {
    double ek = (p[1] * p[1] + p[2] * p[2]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = p[3] / (CFD_GAMMA - 1) + r * ek;
  } else {
    double ek = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = r * p[3];
    c[4] = p[4] / (CFD_GAMMA - 1) + r * ek;
  }
}

// Back-projected conservative state -> primitive (charToPrim tail,
// eigenSystem.cpp:163-167; 1D consToPrim :309-322).
template <int D> __device__ __forceinline__ void eig_cons_to_prim(const double* c, double* p) {
  double rt = c[0];
  if (D == 1) {
    double ut = c[1] / rt;
    double ekt = (ut * ut) / 2;
    p[0] = rt; p[1] = ut; p[2] = (CFD_GAMMA - 1) * (c[2] - rt * ekt);
  } else if (D == 2) {
    double ut = c[1] / rt, vt = c[2] / rt;
    double rekt = (c[1] * c[1] + c[2] * c[2]) / rt / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = (CFD_GAMMA - 1) * (c[3] - rekt);
  } else {
    double ut = c[1] / rt, vt = c[2] / rt, wt = c[3] / rt;
    double rekt = (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) / rt / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = wt; p[4] = (CFD_GAMMA - 1) * (c[4] - rekt);
  }
}

// ---- Roe-averaged characteristic system (2D / 3D) --------------------------
// eigensystemEuler2D(priml, primr, norm) (eigenSystem.cpp:13-86).  The 3D
// system adds a second shear wave last; for x/y normals the first shear wave
// is the 2D one and the w terms are inserted so that w = 0 reproduces 2D.
// Tangents: x: t1=(0,-1,0) t2=(0,0,1); y: t1=(1,0,0) t2=(0,0,1);
//           z: t1=(1,0,0) t2=(0,1,0).
template <int D, int AX> struct RoeSys {
  static constexpr int NV = D + 2;
  static constexpr double nx = AX == 0 ? 1.0 : 0.0;
  static constexpr double ny = AX == 1 ? 1.0 : 0.0;
  static constexpr double nz = AX == 2 ? 1.0 : 0.0;
  double u, v, w, ek, h, c, Vn;

  __device__ __forceinline__ void build(const double* pl, const double* pr) {
    double rl = pl[0], rr = pr[0];
    double ul = pl[1], ur = pr[1], vl = pl[2], vr = pr[2];
    double ppl = pl[NV - 1], ppr = pr[NV - 1];
    double HL, HR;
    double wl = 0.0, wr = 0.0;
    if (D == 2) {
      HL = (ul * ul + vl * vl) / 2 + ppl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
      HR = (ur * ur + vr * vr) / 2 + ppr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
    } else {
      wl = pl[3];
      wr = pr[3];
      HL = (ul * ul + vl * vl + wl * wl) / 2 + ppl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
      HR = (ur * ur + vr * vr + wr * wr) / 2 + ppr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
    }
    double c1 = sqrt(rl), c2 = sqrt(rr);
    double divisor = 1.0 / (c1 + c2);
    u = (c1 * ul + c2 * ur) * divisor;
    v = (c1 * vl + c2 * vr) * divisor;
    double ht = (c1 * HL + c2 * HR) * divisor;
    if (D == 2) {
      w = 0.0;
      Vn = nx * u + ny * v;
      ek = (u * u + v * v) / 2;
    } else {
      w = (c1 * wl + c2 * wr) * divisor;
      Vn = nx * u + ny * v + nz * w;
      ek = (u * u + v * v + w * w) / 2;
    }
    h = ht - ek;
    c = sqrt((CFD_GAMMA - 1) * h);
  }

  // Row k of the left eigenvector matrix (columns rho, rho u, rho v[, rho w], rho E).
  __device__ __forceinline__ void lrow(int k, double* l) const {
    if (k == 0) {
      if (D == 3 && AX == 2) {
        l[0] = -u; l[1] = 1.0; l[2] = 0.0;
      } else {
        l[0] = (-ny * u + nx * v); l[1] = ny; l[2] = -nx;
      }
      if (D == 3) l[3] = 0.0;
      l[NV - 1] = 0.0;
    } else if (k == 1) {
      l[0] = h - ek; l[1] = u; l[2] = v;
      if (D == 3) l[3] = w;
      l[NV - 1] = -1.0;
    } else if (k == 2) {
      l[0] = (Vn / c + ek / h) / 2;
      l[1] = (-nx / c - u / h) / 2;
      l[2] = (-ny / c - v / h) / 2;
      if (D == 3) l[3] = (-nz / c - w / h) / 2;
      l[NV - 1] = 1.0 / (2.0 * h);
    } else if (k == 3) {
      l[0] = (-Vn / c + ek / h) / 2;
      l[1] = (nx / c - u / h) / 2;
      l[2] = (ny / c - v / h) / 2;
      if (D == 3) l[3] = (nz / c - w / h) / 2;
      l[NV - 1] = 1.0 / (2.0 * h);
    } else {  // k == 4, 3D only: shear 2 along t2
      if (AX == 2) { l[0] = -v; l[1] = 0.0; l[2] = 1.0; l[3] = 0.0; }
      else { l[0] = -w; l[1] = 0.0; l[2] = 0.0; l[3] = 1.0; }
      l[4] = 0.0;
    }
  }

  // Column k of the right eigenvector matrix (rows rho, rho u, ..., rho E).
  __device__ __forceinline__ void rcol(int k, double* r) const {
    if (k == 0) {
      r[0] = 0.0;
      if (D == 3 && AX == 2) { r[1] = 1.0; r[2] = 0.0; r[3] = 0.0; r[4] = u; }
      else {
        r[1] = ny; r[2] = -nx;
        if (D == 3) r[3] = 0.0;
        r[NV - 1] = ny * u - nx * v;
      }
    } else if (k == 1) {
      r[0] = 1.0 / h; r[1] = u / h; r[2] = v / h;
      if (D == 3) r[3] = w / h;
      r[NV - 1] = ek / h;
    } else if (k == 2) {
      r[0] = 1.0; r[1] = u - nx * c; r[2] = v - ny * c;
      if (D == 3) r[3] = w - nz * c;
      r[NV - 1] = h + ek - Vn * c;
    } else if (k == 3) {
      r[0] = 1.0; r[1] = u + nx * c; r[2] = v + ny * c;
      if (D == 3) r[3] = w + nz * c;
      r[NV - 1] = h + ek + Vn * c;
    } else {
      r[0] = 0.0;
      if (AX == 2) { r[1] = 0.0; r[2] = 1.0; r[3] = 0.0; r[4] = v; }
      else { r[1] = 0.0; r[2] = 0.0; r[3] = 1.0; r[4] = w; }
    }
  }
};

// Projection of a conservative state on one left eigenvector row, in the
// summation order of primToChar (eigenSystem.cpp:97-107; 1D :260-264).
template <int D> __device__ __forceinline__ double project(const double* c, const double* l) {
  if (D == 1) return c[0] * l[0] + c[1] * l[1] + c[2] * l[2];
  if (D == 2) return c[1] * l[1] + c[2] * l[2] + c[3] * l[3] + c[0] * l[0];
  return c[1] * l[1] + c[2] * l[2] + c[3] * l[3] + c[4] * l[4] + c[0] * l[0];
}

// ---- 1D Roe system (eigensystemEuler1D, eigenSystem.cpp:217-250) ----------
struct RoeSys1 {
  double L[9], R[9];
  __device__ __forceinline__ void build(const double* pl, const double* pr) {
    double rl = pl[0], ul = pl[1], ppl = pl[2];
    double rr = pr[0], ur = pr[1], ppr = pr[2];
    double HL = (ul * ul) / 2 + ppl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
    double HR = (ur * ur) / 2 + ppr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
    double c1 = sqrt(rl), c2 = sqrt(rr);
    double divisor = 1.0 / (sqrt(rl) + sqrt(rr));
    double u = (c1 * ul + c2 * ur) * divisor;
    double ht = (c1 * HL + c2 * HR) * divisor;
    double ek = (u * u) / 2;
    double h = ht - ek;
    double c = sqrt((CFD_GAMMA - 1) * h);
    const double g1 = CFD_GAMMA - 1;
    double f = 0.5 * g1 / (c * c);
    L[0] = (ht + c * (u - c) / g1) * f;
    L[1] = (-u - c / g1) * f;
    L[2] = 1.0 * f;
    L[3] = (-2.0 * ht + 4.0 * c * c / g1) * f;
    L[4] = (2.0 * u) * f;
    L[5] = -2.0 * f;
    L[6] = (ht - c * (u + c) / g1) * f;
    L[7] = (-u + c / g1) * f;
    L[8] = 1.0 * f;
    R[0] = 1.0; R[1] = 1.0; R[2] = 1.0;
    R[3] = u - c; R[4] = u; R[5] = u + c;
    R[6] = ht - u * c; R[7] = 0.5 * u * u; R[8] = ht + u * c;
  }
};

// variable_positive_limiter (zone_FD/src/include/positive_preserving.hpp:14-46)
// with q = rho (which 0) and q = gamma p (which 1).
__device__ __forceinline__ double lim_q(const double* W, int which) {
  if (which == 0) return W[0];
  double u = W[1] / W[0];
  double ekin = 0.5 * u * u;
  double p = (CFD_GAMMA - 1) * (W[2] - W[0] * ekin);
  return CFD_GAMMA * p;
}
__device__ __forceinline__ void positive_limit(const double* W, double* Wi, int which) {
  double qi = lim_q(Wi, which);
  if (qi >= 1e-10) return;
  double qw = lim_q(W, which);
  if (qw < 1e-10) {
    Wi[0] = W[0]; Wi[1] = W[1]; Wi[2] = W[2];
    return;
  }
  double theta = (1e-10 - qw) / (qi - qw);
  for (int i = 0; i < 3; i++) Wi[i] = (1 - theta) * W[i] + theta * Wi[i];
}

// ---- interface Riemann fluxes ----------------------------------------------
// HLLCFlux1D (zone_FD/src/include/fluxSchemes.hpp:74-125)
__device__ __forceinline__ void hllc1d(const double* L, const double* Rr, double* res) {
  double rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  double HL = pl / rl * CFD_GAMMA / (CFD_GAMMA - 1) + (ul * ul) / 2;
  double HR = pr / rr * CFD_GAMMA / (CFD_GAMMA - 1) + (ur * ur) / 2;
  double cl = sqrt(CFD_GAMMA * (pl / rl));
  double cr = sqrt(CFD_GAMMA * (pr / rr));
  double SL = smin(ur - cr, ul - cl);
  double SR = smax(ul + cl, ur + cr);
  double Sstar = ((pr - pl) + (rl * ul * (SL - ul) - rr * ur * (SR - ur))) / (rl * (SL - ul) - rr * (SR - ur));
  if (SL >= 0) {
    res[0] = rl * ul; res[1] = rl * ul * ul + pl; res[2] = rl * HL * ul;
  } else if (Sstar >= 0) {
    double pStar = pl + rl * (SL - ul) * (Sstar - ul);
    double U[3] = {rl, rl * ul, pl / (CFD_GAMMA - 1) + rl * ul * ul / 2};
    double F[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
    double Dd[3] = {0.0, 1.0, Sstar};
    for (int i = 0; i < 3; i++) res[i] = (Sstar * (SL * U[i] - F[i]) + SL * pStar * Dd[i]) / (SL - Sstar);
  } else if (SR >= 0) {
    double pStar = pr + rr * (SR - ur) * (Sstar - ur);
    double U[3] = {rr, rr * ur, pr / (CFD_GAMMA - 1) + rr * ur * ur / 2};
    double F[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
    double Dd[3] = {0.0, 1.0, Sstar};
    for (int i = 0; i < 3; i++) res[i] = (Sstar * (SR * U[i] - F[i]) + SR * pStar * Dd[i]) / (SR - Sstar);
  } else {
    res[0] = rr * ur; res[1] = rr * ur * ur + pr; res[2] = rr * HR * ur;
  }
}

// roeFlux1D2 (fluxSchemes.hpp:13-71)
__device__ __forceinline__ void roe1d(const double* L, const double* Rr, double* res) {
  double rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  double HL = pl / rl * CFD_GAMMA / (CFD_GAMMA - 1) + (ul * ul) / 2;
  double HR = pr / rr * CFD_GAMMA / (CFD_GAMMA - 1) + (ur * ur) / 2;
  double rBar = sqrt(rl * rr);
  double srl = sqrt(rl), srr = sqrt(rr);
  double uBar = (ul * srl + ur * srr) / (srl + srr);
  double HBar = (HL * srl + HR * srr) / (srl + srr);
  double cBar = sqrt((CFD_GAMMA - 1) * (HBar - uBar * uBar / 2));
  double cBar2 = (CFD_GAMMA - 1) * (HBar - uBar * uBar / 2);
  double dr = rr - rl, du = ur - ul, dp = pr - pl;
  double lam[3] = {fabs(uBar), fabs(uBar + cBar), fabs(uBar - cBar)};
  double eps = 0.1 * (fabs(uBar) + cBar);
  for (int i = 0; i < 3; i++)
    if (lam[i] < eps) lam[i] = (lam[i] * lam[i] + eps * eps) / (2.0 * eps);
  double a0 = lam[0] * (dr - dp / cBar2);
  double a1 = lam[1] / (2.0 * cBar2) * (dp + rBar * cBar * du);
  double a2 = lam[2] / (2.0 * cBar2) * (dp - rBar * cBar * du);
  double a3 = a0 + a1 + a2;
  double a4 = cBar * (a1 - a2);
  double FL[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
  double FR[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
  for (int i = 0; i < 3; i++) res[i] = (FL[i] + FR[i]) / 2.0;
  res[0] -= 0.5 * a3;
  res[1] -= 0.5 * (a3 * uBar + a4);
  res[2] -= 0.5 * (HBar * a3 + uBar * a4 - cBar2 * a0 / (CFD_GAMMA - 1));
}

// HLLCFlux2D2 (fluxSchemes.hpp:202-327), pressure-based wave speeds, with the
// averaged Sstar == 0 branch; 3D writes the star tangential momentum as
// Sstar*n_k + u_k*(sum of the other normal components).
template <int D, int AX> __device__ __forceinline__ void hllc_md(const double* Lp, const double* Rp, double* res) {
  constexpr int NV = D + 2, E = NV - 1;
  const double n[3] = {AX == 0 ? 1.0 : 0.0, AX == 1 ? 1.0 : 0.0, AX == 2 ? 1.0 : 0.0};
  double rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  double HL, HR;
  if (D == 2) {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2]) / 2 + pl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2]) / 2 + pr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
  } else {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2] + Lp[3] * Lp[3]) / 2 + pl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2] + Rp[3] * Rp[3]) / 2 + pr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
  }
  double VnL = Lp[1 + AX], VnR = Rp[1 + AX];
  double cl = sqrt(pl / rl * CFD_GAMMA);
  double cr = sqrt(pr / rr * CFD_GAMMA);
  double rBar = (rl + rr) / 2, cBar = (cl + cr) / 2;
  double ps = smax(0.0, (pl + pr) / 2 - rBar * cBar * (VnR - VnL) / 2);
  double ql = (ps <= pl) ? 1.0 : sqrt(1.0 + (CFD_GAMMA + 1.0) / (2.0 * CFD_GAMMA) * (ps / pl - 1.0));
  double qr = (ps <= pr) ? 1.0 : sqrt(1.0 + (CFD_GAMMA + 1.0) / (2.0 * CFD_GAMMA) * (ps / pr - 1.0));
  double SL = VnL - cl * ql, SR = VnR + cr * qr;
  double Sstar = ((pr - pl) + (rl * VnL * (SL - VnL) - rr * VnR * (SR - VnR))) / (rl * (SL - VnL) - rr * (SR - VnR));

  auto side = [&](const double* P, double r, double p, double H, double Vn, double* F) {
    F[0] = r * Vn;
    for (int k = 0; k < D; k++) F[1 + k] = r * P[1 + k] * Vn + p * n[k];
    F[E] = r * H * Vn;
  };
  auto star = [&](const double* P, double r, double p, double H, double Vn, double S, double* out) {
    double F[NV];
    side(P, r, p, H, Vn, F);
    double Usf = (S - Vn) / (S - Sstar);
    double UE = r * H - p;
    out[0] = F[0] + S * (Usf * r - r);
    for (int k = 0; k < D; k++) {
      double other = 0.0;
      for (int m = 0; m < D; m++)
        if (m != k) other += n[m];
      double Us = Usf * (Sstar * n[k] + P[1 + k] * other) * r;
      out[1 + k] = F[1 + k] + S * (Us - r * P[1 + k]);
    }
    double UsE = Usf * (UE + (Sstar - Vn) * (r * Sstar + p / (S - Vn)));
    out[E] = F[E] + S * (UsE - UE);
  };
  if (SL > 0) {
    side(Lp, rl, pl, HL, VnL, res);
  } else if (Sstar > 0) {
    star(Lp, rl, pl, HL, VnL, SL, res);
  } else if (Sstar == 0) {
    double a[NV], b[NV];
    star(Lp, rl, pl, HL, VnL, SL, a);
    star(Rp, rr, pr, HR, VnR, SR, b);
    for (int i = 0; i < NV; i++) res[i] = a[i] / 2 + b[i] / 2;
  } else if (SR > 0) {
    star(Rp, rr, pr, HR, VnR, SR, res);
  } else {
    side(Rp, rr, pr, HR, VnR, res);
  }
}

// roeFlux2DSym (fluxSchemes.hpp:420-553) with the Harten entropy fix; the 3D
// form sums the shear energy over the tangential components in (u, v, w) order.
template <int D, int AX> __device__ __forceinline__ void roe_md(const double* Lp, const double* Rp, double* res) {
  constexpr int NV = D + 2, E = NV - 1;
  const double n[3] = {AX == 0 ? 1.0 : 0.0, AX == 1 ? 1.0 : 0.0, AX == 2 ? 1.0 : 0.0};
  double rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  double HL, HR;
  if (D == 2) {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2]) / 2 + pl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2]) / 2 + pr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
  } else {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2] + Lp[3] * Lp[3]) / 2 + pl / rl * CFD_GAMMA / (CFD_GAMMA - 1);
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2] + Rp[3] * Rp[3]) / 2 + pr / rr * CFD_GAMMA / (CFD_GAMMA - 1);
  }
  double VnL = Lp[1 + AX], VnR = Rp[1 + AX];
  double c1 = sqrt(rl), c2 = sqrt(rr);
  double divisor = 1.0 / (c1 + c2);
  double rhoAvg = sqrt(rl * rr);
  double va[3];
  for (int k = 0; k < D; k++) va[k] = (c1 * Lp[1 + k] + c2 * Rp[1 + k]) * divisor;
  double HAvg = (c1 * HL + c2 * HR) * divisor;
  double q2 = (D == 2) ? (va[0] * va[0] + va[1] * va[1]) / 2 : (va[0] * va[0] + va[1] * va[1] + va[2] * va[2]) / 2;
  double cAvg = sqrt((CFD_GAMMA - 1) * (HAvg - q2));
  double VnAvg = va[AX];
  double lam0 = fabs(VnAvg - cAvg), lam1 = fabs(VnAvg), lam2 = fabs(VnAvg + cAvg);
  double eps = 0.05 * (fabs(VnAvg) + cAvg);
  if (lam0 < eps) lam0 = (lam0 * lam0 + eps * eps) / (2.0 * eps);
  if (lam1 < eps) lam1 = (lam1 * lam1 + eps * eps) / (2.0 * eps);
  if (lam2 < eps) lam2 = (lam2 * lam2 + eps * eps) / (2.0 * eps);
  double dP = pr - pl, dVn = VnR - VnL, dRho = rr - rl;
  double cf1 = (dP - rhoAvg * cAvg * dVn) / (2.0 * cAvg * cAvg);
  double cf = (dP + rhoAvg * cAvg * dVn) / (2.0 * cAvg * cAvg);
  double cf2 = dRho - dP / (cAvg * cAvg);
  double shear = 0.0;
  bool first = true;
  for (int k = 0; k < D; k++) {
    if (k == AX) continue;
    double t = va[k] * (Rp[1 + k] - Lp[1 + k]);
    shear = first ? t : shear + t;
    first = false;
  }
  // component 0
  {
    double fd = lam0 * (cf1 * 1);
    fd += lam2 * (cf * 1);
    fd += lam1 * (cf2 * 1.0 + rhoAvg * 0.0);
    res[0] = (rl * VnL + rr * VnR - fd) / 2;
  }
  for (int k = 0; k < D; k++) {
    double fd = lam0 * (cf1 * (va[k] - cAvg * n[k]));
    fd += lam2 * (cf * (va[k] + cAvg * n[k]));
    fd += lam1 * (cf2 * va[k] + rhoAvg * (k == AX ? 0.0 : (Rp[1 + k] - Lp[1 + k])));
    double fl = rl * Lp[1 + k] * VnL + pl * n[k];
    double fr = rr * Rp[1 + k] * VnR + pr * n[k];
    res[1 + k] = (fl + fr - fd) / 2;
  }
  {
    double fd = lam0 * (cf1 * (HAvg - cAvg * VnAvg));
    fd += lam2 * (cf * (HAvg + cAvg * VnAvg));
    fd += lam1 * (cf2 * q2 + rhoAvg * shear);
    res[E] = (rl * HL * VnL + rr * HR * VnR - fd) / 2;
  }
}

// Physical flux at a cell (fluxSolveEuler1D / 2D, fluxSchemes.hpp:572-599).
template <int D, int AX> __device__ __forceinline__ void node_flux(const double* p, double* F) {
  if (D == 1) {
    double r = p[0], u = p[1], pp = p[2];
    F[0] = r * u;
    F[1] = r * u * u + pp;
    F[2] = (CFD_GAMMA / (CFD_GAMMA - 1) * pp + r * (u * u) / 2) * u;
    return;
  }
  constexpr int NV = D + 2, E = NV - 1;
  const double n[3] = {AX == 0 ? 1.0 : 0.0, AX == 1 ? 1.0 : 0.0, AX == 2 ? 1.0 : 0.0};
  double r = p[0], pp = p[E];
  double Vn, k2;
  if (D == 2) {
    Vn = (p[1] * n[0] + p[2] * n[1]);
    k2 = (p[1] * p[1] + p[2] * p[2]);
  } else {
    Vn = (p[1] * n[0] + p[2] * n[1] + p[3] * n[2]);
    k2 = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  }
  F[0] = r * Vn;
  for (int k = 0; k < D; k++) F[1 + k] = r * p[1 + k] * Vn + n[k] * pp;
  F[E] = (k2 / 2 * r + CFD_GAMMA / (CFD_GAMMA - 1) * pp) * Vn;
}

}  // namespace cfd
