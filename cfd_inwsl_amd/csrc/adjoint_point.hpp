// Hand-written reverse-mode (vector-Jacobian) counterparts of the 2D
// characteristic-wise WENO-JS / TENO5 + HLLC interface flux, the node flux and the
// conservative->primitive map.  Each *_vjp function recomputes its forward
// intermediates and walks them backwards; branch decisions (HLLC wave
// pattern, ps clipping, shock factors, positivity fallback) are those of the
// forward evaluation, i.e. the derivative of the active branch.
//
// Forward formulas: weno5_JSchen (zone_FD/src/include/interScheme.hpp:17-43),
// Teno5_Z (interScheme.hpp:384-452),
// eigensystemEuler2D (zone_FD/src/src/eigenSystem.cpp:13-168),
// Recon5Order2DEulerEig::reconI (zone_FD/src/include/reconstructor5order.hpp:247-356),
// HLLCFlux2D2 (zone_FD/src/include/fluxSchemes.hpp:202-327),
// fluxSolveEuler2D (fluxSchemes.hpp:583-599), Equation::consToPrimEuler2D
// (zone_FD/src/src/equation.cpp:67-99).
//
// The adjoint needs derivatives, not the reference's rounding, so it is free
// to reassociate; it is compiled for host (unit tests against finite
// differences) and device.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define CFD_HD __host__ __device__
#define CFD_UNROLL _Pragma("unroll")
#else
#define CFD_HD
#define CFD_UNROLL
#endif
// Reverse-pass code needs derivatives, not the forward's rounding: on the
// device multiply-adds in it may contract to FMAs (the forward re-evaluation
// that decides branches is compiled without contraction).
#if defined(__clang__)
#define CFD_CONTRACT _Pragma("clang fp contract(fast)")
#else
#define CFD_CONTRACT
#endif

namespace cfdadj {

constexpr double G = 1.4;
constexpr double G1 = G - 1.0;          // gamma - 1
constexpr double GG1 = G / (G - 1.0);   // gamma / (gamma - 1)
constexpr double IG1 = 1.0 / G1;        // 1 / (gamma - 1)

// Reciprocal for the reverse pass, which needs derivatives, not the forward's
// rounding: on the device the hardware estimate refined by two Newton steps
// (relative error at the level of one rounding), on the host a division.
CFD_HD inline double rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
#else
  return 1.0 / x;
#endif
}

// Square root for the reverse pass (positive arguments: densities, sound
// speeds, shock factors): on the device the hardware reciprocal square root
// refined by one Newton step and one residual correction (within an ulp), on
// the host the library root.
CFD_HD inline double rt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  g = fma(d, h, g);
  return x == 0.0 ? x : g;
#else
  return sqrt(x);
#endif
}

// ---- WENO-JS half-node value and its gradient ---------------------------------
struct JS {
  static constexpr bool kLinear = false;
  double b[3], a[3], u[3], e[3], S, iS, out;
  CFD_HD void fwd(const double* q) {
    CFD_CONTRACT
    double A0 = q[0] - 2.0 * q[1] + q[2], B0 = q[0] - 4.0 * q[1] + 3.0 * q[2];
    double A1 = q[1] - 2.0 * q[2] + q[3], B1 = q[1] - q[3];
    double A2 = q[2] - 2.0 * q[3] + q[4], B2 = 3.0 * q[2] - 4.0 * q[3] + q[4];
    b[0] = A0 * A0 + 0.25 * B0 * B0;
    b[1] = A1 * A1 + 0.25 * B1 * B1;
    b[2] = A2 * A2 + 0.25 * B2 * B2;
    u[0] = 0.375 * q[0] - 1.25 * q[1] + 1.875 * q[2];
    u[1] = -0.125 * q[1] + 0.75 * q[2] + 0.375 * q[3];
    u[2] = 0.375 * q[2] + 0.75 * q[3] - 0.125 * q[4];
    const double d[3] = {0.0625, 0.625, 0.3125};
    // e_k = 1 / (1e-6 + b_k) from one reciprocal of the product (each factor
    // lies in [1e-6, ~1e100] for finite data)
    const double x0 = 1e-6 + b[0], x1 = 1e-6 + b[1], x2 = 1e-6 + b[2];
    const double x01 = x0 * x1;
    const double r = rcp(x01 * x2);
    const double r01 = x2 * r;  // 1 / (x0 x1)
    e[0] = x1 * r01;
    e[1] = x0 * r01;
    e[2] = x01 * r;
    for (int k = 0; k < 3; k++) a[k] = d[k] * e[k] * e[k];
    S = a[0] + a[1] + a[2];
    iS = rcp(S);
    out = (a[0] * u[0] + a[1] * u[1] + a[2] * u[2]) * iS;
  }
  // adds lam * d(out)/dq to g
  CFD_HD void bwd(const double* q, double lam, double* g) const {
    CFD_CONTRACT
    const double iS = lam * this->iS;
    double lu[3], lb[3];
    for (int k = 0; k < 3; k++) {
      lu[k] = a[k] * iS;
      lb[k] = (u[k] - out) * iS * (-2.0 * a[k] * e[k]);
    }
    // u
    g[0] += 0.375 * lu[0];
    g[1] += -1.25 * lu[0] - 0.125 * lu[1];
    g[2] += 1.875 * lu[0] + 0.75 * lu[1] + 0.375 * lu[2];
    g[3] += 0.375 * lu[1] + 0.75 * lu[2];
    g[4] += -0.125 * lu[2];
    // betas
    double A0 = q[0] - 2.0 * q[1] + q[2], B0 = q[0] - 4.0 * q[1] + 3.0 * q[2];
    double A1 = q[1] - 2.0 * q[2] + q[3], B1 = q[1] - q[3];
    double A2 = q[2] - 2.0 * q[3] + q[4], B2 = 3.0 * q[2] - 4.0 * q[3] + q[4];
    double s0A = 2.0 * A0 * lb[0], s0B = 0.5 * B0 * lb[0];
    double s1A = 2.0 * A1 * lb[1], s1B = 0.5 * B1 * lb[1];
    double s2A = 2.0 * A2 * lb[2], s2B = 0.5 * B2 * lb[2];
    g[0] += s0A + s0B;
    g[1] += -2.0 * s0A - 4.0 * s0B + s1A + s1B;
    g[2] += s0A + 3.0 * s0B - 2.0 * s1A + s2A + 3.0 * s2B;
    g[3] += s1A - s1B - 2.0 * s2A - 4.0 * s2B;
    g[4] += s2A + s2B;
  }
};

// ---- TENO5 (Teno5_Z) half-node value and its gradient --------------------------
// The value is the fixed stencil its cut-off flag selects, so the gradient
// is that stencil's coefficients; the flag is a branch of the forward
// (decided by the forward's arithmetic, interScheme.hpp:384-452).
struct Teno {
  static constexpr bool kLinear = true;
  double ct = 1e-5;
  unsigned flag = 0;
  double out = 0.0;
  // flag in the oracle's arithmetic (no contraction): host checks
  CFD_HD void fwd(const double* q) {
    const double b0 = (q[0] - 2.0 * q[1] + q[2]) * (q[0] - 2.0 * q[1] + q[2]) +
                      0.25 * ((q[0] - 4.0 * q[1] + 3.0 * q[2]) * (q[0] - 4.0 * q[1] + 3.0 * q[2]));
    const double b1 = (q[1] - 2.0 * q[2] + q[3]) * (q[1] - 2.0 * q[2] + q[3]) + 0.25 * ((q[1] - q[3]) * (q[1] - q[3]));
    const double b2 = (q[2] - 2.0 * q[3] + q[4]) * (q[2] - 2.0 * q[3] + q[4]) +
                      0.25 * ((3.0 * q[2] - 4.0 * q[3] + q[4]) * (3.0 * q[2] - 4.0 * q[3] + q[4]));
    const double tau = fabs(b2 - b0);
    const double bb[3] = {b0, b1, b2};
    double g[3];
    for (int i = 0; i < 3; i++) {
      double t = 1.0 + tau / (bb[i] + 1e-40);
      t = t * t;
      g[i] = t * t * t;
    }
    const double CT = ct * (g[0] + g[1] + g[2]);
    flag = (g[0] < CT ? 1u : 0u) + (g[1] < CT ? 2u : 0u) + (g[2] < CT ? 4u : 0u);
    double c[5];
    coef(flag, c);
    out = c[0] * q[0] + c[1] * q[1] + c[2] * q[2] + c[3] * q[3] + c[4] * q[4];
  }
  // stencil coefficients of a flag (interScheme.hpp:421-449)
  CFD_HD static void coef(unsigned f, double* c) {
    const double t[8][5] = {{0.0234375, -0.15625, 0.703125, 0.46875, -0.0390625},
                            {0.0, -0.0625, 0.5625, 0.5625, -0.0625},
                            {0.0, 0.0, 0.375, 0.75, -0.125},
                            {0.0, 0.0, 0.375, 0.75, -0.125},
                            {0.0625, -0.3125, 0.9375, 0.3125, 0.0},
                            {0.0, -0.125, 0.75, 0.375, 0.0},
                            {0.375, -1.25, 1.875, 0.0, 0.0},
                            {0.0, 0.0, 1.0, 0.0, 0.0}};
    for (int j = 0; j < 5; j++) c[j] = t[f & 7u][j];
  }
  // adds lam * d(out)/dq to g
  CFD_HD void bwd(const double*, double lam, double* g) const {
    double c[5];
    coef(flag, c);
    for (int j = 0; j < 5; j++) g[j] += lam * c[j];
  }
};

// ---- primitive <-> conservative ---------------------------------------------------
CFD_HD inline void p2c(const double* p, double* c) {
  CFD_CONTRACT
  c[0] = p[0];
  c[1] = p[0] * p[1];
  c[2] = p[0] * p[2];
  c[3] = p[3] * IG1 + 0.5 * p[0] * (p[1] * p[1] + p[2] * p[2]);
}
// g_p += (dc/dp)^T lc
CFD_HD inline void p2c_vjp(const double* p, const double* lc, double* gp) {
  CFD_CONTRACT
  double ek = 0.5 * (p[1] * p[1] + p[2] * p[2]);
  gp[0] += lc[0] + lc[1] * p[1] + lc[2] * p[2] + lc[3] * ek;
  gp[1] += lc[1] * p[0] + lc[3] * p[0] * p[1];
  gp[2] += lc[2] * p[0] + lc[3] * p[0] * p[2];
  gp[3] += lc[3] * IG1;
}
CFD_HD inline void c2p(const double* c, double* p) {
  CFD_CONTRACT
  double r = c[0], ir = rcp(r), u = c[1] * ir, v = c[2] * ir;
  p[0] = r;
  p[1] = u;
  p[2] = v;
  p[3] = G1 * (c[3] - 0.5 * r * (u * u + v * v));
}
// g_c += (dp/dc)^T lp
CFD_HD inline void c2p_vjp(const double* c, const double* lp, double* gc) {
  CFD_CONTRACT
  double r = c[0], ir = rcp(r), u = c[1] * ir, v = c[2] * ir;
  // p1 = c1/r, p2 = c2/r, p3 = G1 (c3 - (c1^2 + c2^2)/(2r))
  gc[0] += lp[0] - lp[1] * u * ir - lp[2] * v * ir + lp[3] * G1 * 0.5 * (u * u + v * v);
  gc[1] += lp[1] * ir - lp[3] * G1 * u;
  gc[2] += lp[2] * ir - lp[3] * G1 * v;
  gc[3] += lp[3] * G1;
}

// ---- node flux (fluxSolveEuler2D) -------------------------------------------------
template <int AX> CFD_HD inline void node_flux(const double* p, double* F) {
  double r = p[0], u = p[1], v = p[2], pp = p[3];
  double Vn = AX == 0 ? u : v;
  F[0] = r * Vn;
  F[1] = r * u * Vn + (AX == 0 ? pp : 0.0);
  F[2] = r * v * Vn + (AX == 1 ? pp : 0.0);
  F[3] = (0.5 * (u * u + v * v) * r + GG1 * pp) * Vn;
}
template <int AX> CFD_HD inline void node_flux_vjp(const double* p, const double* lF, double* gp) {
  CFD_CONTRACT
  double r = p[0], u = p[1], v = p[2], pp = p[3];
  double Vn = AX == 0 ? u : v;
  double E = 0.5 * (u * u + v * v) * r + GG1 * pp;
  double lVn = lF[0] * r + lF[1] * r * u + lF[2] * r * v + lF[3] * E;
  gp[0] += lF[0] * Vn + lF[1] * u * Vn + lF[2] * v * Vn + lF[3] * 0.5 * (u * u + v * v) * Vn;
  gp[1] += lF[1] * r * Vn + lF[3] * r * u * Vn + (AX == 0 ? lVn : 0.0);
  gp[2] += lF[2] * r * Vn + lF[3] * r * v * Vn + (AX == 1 ? lVn : 0.0);
  gp[3] += (AX == 0 ? lF[1] : lF[2]) + lF[3] * GG1 * Vn;
}

// ---- HLLC (HLLCFlux2D2) forward and reverse --------------------------------------
// One side's flux and star-region flux F + S (U* - U).
template <int AX> struct HSide {
  // inputs
  double r, u, v, p;
  // intermediates
  double H, Vn;
};

// Branch bits of one HLLC evaluation: bit 0 pressure clipped to 0, bit 1 left
// shock factor active, bit 2 right shock factor active, bits 3-5 the wave
// region (0 SL > 0, 1 S* > 0, 2 S* == 0, 3 SR > 0, 4 otherwise).
enum { HB_CLIP = 1, HB_SHL = 2, HB_SHR = 4 };
CFD_HD inline int hb_region(int b) { return (b >> 3) & 7; }

// force >= 0: take the branches it names instead of deciding them from this
// function's own arithmetic (the reverse pass follows the branches of the
// forward evaluation, whose rounding can differ where a predicate is tied).
template <int AX> CFD_HD inline void hllc(const double* L, const double* R, double* F, double* lL = nullptr,
                                          double* lR = nullptr, const double* lF = nullptr, int force = -1) {
  CFD_CONTRACT
  const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
  const double rl = L[0], ul = L[1], vl = L[2], pl = L[3];
  const double rr = R[0], ur = R[1], vr = R[2], pr = R[3];
  const double irl = rcp(rl), irr = rcp(rr), ipl = rcp(pl), ipr = rcp(pr);
  const double HL = 0.5 * (ul * ul + vl * vl) + GG1 * pl * irl;
  const double HR = 0.5 * (ur * ur + vr * vr) + GG1 * pr * irr;
  const double VnL = AX == 0 ? ul : vl, VnR = AX == 0 ? ur : vr;
  const double cl = rt(G * pl * irl), cr = rt(G * pr * irr);
  const double rBar = 0.5 * (rl + rr), cBar = 0.5 * (cl + cr);
  const double psr = 0.5 * (pl + pr) - 0.5 * rBar * cBar * (VnR - VnL);
  const bool clip = force >= 0 ? (force & HB_CLIP) != 0 : !(0.0 < psr);
  const double ps = clip ? 0.0 : psr;
  const double kap = (G + 1.0) / (2.0 * G);
  const bool shl = force >= 0 ? (force & HB_SHL) != 0 : !(ps <= pl);
  const bool shr = force >= 0 ? (force & HB_SHR) != 0 : !(ps <= pr);
  const double ql = shl ? rt(1.0 + kap * (ps * ipl - 1.0)) : 1.0;
  const double qr = shr ? rt(1.0 + kap * (ps * ipr - 1.0)) : 1.0;
  const double SL = VnL - cl * ql, SR = VnR + cr * qr;
  const double aL = SL - VnL, aR = SR - VnR;
  const double num = (pr - pl) + (rl * VnL * aL - rr * VnR * aR);
  const double den = rl * aL - rr * aR;
  const double iden = rcp(den);
  const double Ss = num * iden;

  // forward side / star fluxes; reverse accumulators for (r, u, v, p, S, Ss)
  auto side = [&](double r, double u, double v, double p, double H, double Vn, double* out) {
    out[0] = r * Vn;
    out[1] = r * u * Vn + p * nx;
    out[2] = r * v * Vn + p * ny;
    out[3] = r * H * Vn;
  };
  // reverse of side(): lo = cotangent of out; adds to (lr, lu, lv, lp)
  auto side_b = [&](double r, double u, double v, double p, double H, double Vn, const double* lo, double w,
                    double& lr, double& lu, double& lv, double& lp) {
    double l0 = w * lo[0], l1 = w * lo[1], l2 = w * lo[2], l3 = w * lo[3];
    double lVn = l0 * r + l1 * r * u + l2 * r * v + l3 * r * H;
    double lH = l3 * r * Vn;
    lr += l0 * Vn + l1 * u * Vn + l2 * v * Vn + l3 * H * Vn;
    lu += l1 * r * Vn;
    lv += l2 * r * Vn;
    lp += l1 * nx + l2 * ny;
    // H = (u^2+v^2)/2 + GG1 p / r
    lu += lH * u;
    lv += lH * v;
    const double ir = rcp(r);
    lp += lH * GG1 * ir;
    lr += -lH * GG1 * p * ir * ir;
    if (AX == 0) lu += lVn;
    else lv += lVn;
  };
  auto star = [&](double r, double u, double v, double p, double H, double Vn, double S, double* out) {
    double Fs[4];
    side(r, u, v, p, H, Vn, Fs);
    double Usf = (S - Vn) * rcp(S - Ss);
    double U3 = r * H - p;
    double Us0 = Usf * r;
    double Us1 = Usf * (Ss * nx + u * ny) * r;
    double Us2 = Usf * (Ss * ny + v * nx) * r;
    double Us3 = Usf * (U3 + (Ss - Vn) * (r * Ss + p * rcp(S - Vn)));
    out[0] = Fs[0] + S * (Us0 - r);
    out[1] = Fs[1] + S * (Us1 - r * u);
    out[2] = Fs[2] + S * (Us2 - r * v);
    out[3] = Fs[3] + S * (Us3 - U3);
  };
  // reverse of star(): adds to state cotangents and to lS, lSs
  auto star_b = [&](double r, double u, double v, double p, double H, double Vn, double S, const double* lo, double w,
                    double& lr, double& lu, double& lv, double& lp, double& lS, double& lSs) {
    double l0 = w * lo[0], l1 = w * lo[1], l2 = w * lo[2], l3 = w * lo[3];
    double D = S - Ss, A = S - Vn;
    const double iA = rcp(A), iD = rcp(D);
    double Usf = A * iD;
    double U3 = r * H - p;
    double Pt = r * Ss + p * iA;
    double Us0 = Usf * r, Us1 = Usf * (Ss * nx + u * ny) * r, Us2 = Usf * (Ss * ny + v * nx) * r;
    double Us3 = Usf * (U3 + (Ss - Vn) * Pt);
    // out = Fs + S (Us - U)
    double lo4[4] = {l0, l1, l2, l3};
    side_b(r, u, v, p, H, Vn, lo4, 1.0, lr, lu, lv, lp);
    lS += l0 * (Us0 - r) + l1 * (Us1 - r * u) + l2 * (Us2 - r * v) + l3 * (Us3 - U3);
    double lUs0 = l0 * S, lUs1 = l1 * S, lUs2 = l2 * S, lUs3 = l3 * S;
    // -S U terms
    lr += -l0 * S - l1 * S * u - l2 * S * v;
    lu += -l1 * S * r;
    lv += -l2 * S * r;
    double lU3 = -l3 * S;
    // Us0 = Usf r
    double lUsf = lUs0 * r;
    lr += lUs0 * Usf;
    // Us1 = Usf (Ss nx + u ny) r
    double t1 = Ss * nx + u * ny;
    lUsf += lUs1 * t1 * r;
    lr += lUs1 * Usf * t1;
    lSs += lUs1 * Usf * r * nx;
    lu += lUs1 * Usf * r * ny;
    // Us2 = Usf (Ss ny + v nx) r
    double t2 = Ss * ny + v * nx;
    lUsf += lUs2 * t2 * r;
    lr += lUs2 * Usf * t2;
    lSs += lUs2 * Usf * r * ny;
    lv += lUs2 * Usf * r * nx;
    // Us3 = Usf (U3 + (Ss - Vn) Pt)
    double inner = U3 + (Ss - Vn) * Pt;
    lUsf += lUs3 * inner;
    double lin = lUs3 * Usf;
    lU3 += lin;
    lSs += lin * Pt;
    double lVn = -lin * Pt;
    double lPt = lin * (Ss - Vn);
    // Pt = r Ss + p / A
    lr += lPt * Ss;
    lSs += lPt * r;
    lp += lPt * iA;
    double lA = -lPt * p * iA * iA;
    // Usf = A / D
    lA += lUsf * iD;
    double lD = -lUsf * A * iD * iD;
    // A = S - Vn ; D = S - Ss
    lS += lA + lD;
    lVn += -lA;
    lSs += -lD;
    // U3 = r H - p
    lr += lU3 * H;
    double lH = lU3 * r;
    lp += -lU3;
    // H and Vn
    lu += lH * u;
    lv += lH * v;
    const double ir = rcp(r);
    lp += lH * GG1 * ir;
    lr += -lH * GG1 * p * ir * ir;
    if (AX == 0) lu += lVn;
    else lv += lVn;
  };

  int br;
  if (force >= 0) br = hb_region(force);
  else if (SL > 0) br = 0;
  else if (Ss > 0) br = 1;
  else if (Ss == 0) br = 2;
  else if (SR > 0) br = 3;
  else br = 4;
  if (!lF) {
    if (br == 0) side(rl, ul, vl, pl, HL, VnL, F);
    else if (br == 1) star(rl, ul, vl, pl, HL, VnL, SL, F);
    else if (br == 2) {
      double a[4], b[4];
      star(rl, ul, vl, pl, HL, VnL, SL, a);
      star(rr, ur, vr, pr, HR, VnR, SR, b);
      for (int i = 0; i < 4; i++) F[i] = 0.5 * (a[i] + b[i]);
    } else if (br == 3) star(rr, ur, vr, pr, HR, VnR, SR, F);
    else side(rr, ur, vr, pr, HR, VnR, F);
    return;
  }
  // ---- reverse ----
  double lrl = 0, lul = 0, lvl = 0, lpl = 0, lrr = 0, lur = 0, lvr = 0, lpr = 0;
  double lSL = 0, lSR = 0, lSs = 0;
  if (br == 0) side_b(rl, ul, vl, pl, HL, VnL, lF, 1.0, lrl, lul, lvl, lpl);
  else if (br == 4) side_b(rr, ur, vr, pr, HR, VnR, lF, 1.0, lrr, lur, lvr, lpr);
  else {
    const double wl = br == 2 ? 0.5 : (br == 1 ? 1.0 : 0.0);
    const double wr = br == 2 ? 0.5 : (br == 3 ? 1.0 : 0.0);
    if (wl != 0.0) star_b(rl, ul, vl, pl, HL, VnL, SL, lF, wl, lrl, lul, lvl, lpl, lSL, lSs);
    if (wr != 0.0) star_b(rr, ur, vr, pr, HR, VnR, SR, lF, wr, lrr, lur, lvr, lpr, lSR, lSs);
  }
  // Ss = num / den
  double lnum = lSs * iden, lden = -lSs * Ss * iden;
  // num = (pr - pl) + rl VnL aL - rr VnR aR ; den = rl aL - rr aR
  lpr += lnum;
  lpl += -lnum;
  double laL = lnum * rl * VnL + lden * rl;
  double laR = -lnum * rr * VnR - lden * rr;
  lrl += lnum * VnL * aL + lden * aL;
  lrr += -lnum * VnR * aR - lden * aR;
  double lVnL = lnum * rl * aL, lVnR = -lnum * rr * aR;
  // aL = SL - VnL ; aR = SR - VnR
  lSL += laL;
  lVnL += -laL;
  lSR += laR;
  lVnR += -laR;
  // SL = VnL - cl ql ; SR = VnR + cr qr
  lVnL += lSL;
  double lcl = -lSL * ql, lql = -lSL * cl;
  lVnR += lSR;
  double lcr = lSR * qr, lqr = lSR * cr;
  // ql = sqrt(1 + kap (ps/pl - 1)) when shocked
  double lps = 0;
  if (shl) {
    double g = lql * 0.5 * rcp(ql) * kap;
    lps += g * ipl;
    lpl += -g * ps * ipl * ipl;
  }
  if (shr) {
    double g = lqr * 0.5 * rcp(qr) * kap;
    lps += g * ipr;
    lpr += -g * ps * ipr * ipr;
  }
  // ps = max(0, psr); psr = (pl + pr)/2 - rBar cBar (VnR - VnL)/2
  if (!clip) {
    lpl += 0.5 * lps;
    lpr += 0.5 * lps;
    double dv = VnR - VnL;
    double lrBar = -0.5 * lps * cBar * dv, lcBar = -0.5 * lps * rBar * dv;
    lVnR += -0.5 * lps * rBar * cBar;
    lVnL += 0.5 * lps * rBar * cBar;
    lrl += 0.5 * lrBar;
    lrr += 0.5 * lrBar;
    lcl += 0.5 * lcBar;
    lcr += 0.5 * lcBar;
  }
  // cl = sqrt(G pl / rl)
  {
    double g = lcl * 0.5 * rcp(cl) * G;
    lpl += g * irl;
    lrl += -g * pl * irl * irl;
    double h = lcr * 0.5 * rcp(cr) * G;
    lpr += h * irr;
    lrr += -h * pr * irr * irr;
  }
  if (AX == 0) {
    lul += lVnL;
    lur += lVnR;
  } else {
    lvl += lVnL;
    lvr += lVnR;
  }
  lL[0] += lrl;
  lL[1] += lul;
  lL[2] += lvl;
  lL[3] += lpl;
  lR[0] += lrr;
  lR[1] += lur;
  lR[2] += lvr;
  lR[3] += lpr;
}

// ---- Roe-averaged characteristic system of eigensystemEuler2D ------------------
template <int AX> struct Roe2 {
  double c1, c2, D, HL, HR;
  double u, v, ht, Vn, ek, h, c, iD, ic, ih;
  CFD_HD void fwd(const double* pl, const double* pr) {
    CFD_CONTRACT
    HL = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2]) + GG1 * pl[3] * rcp(pl[0]);
    HR = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2]) + GG1 * pr[3] * rcp(pr[0]);
    c1 = rt(pl[0]);
    c2 = rt(pr[0]);
    D = c1 + c2;
    iD = rcp(D);
    u = (c1 * pl[1] + c2 * pr[1]) * iD;
    v = (c1 * pl[2] + c2 * pr[2]) * iD;
    ht = (c1 * HL + c2 * HR) * iD;
    Vn = AX == 0 ? u : v;
    ek = 0.5 * (u * u + v * v);
    h = ht - ek;
    c = rt(G1 * h);
    ic = rcp(c);
    ih = rcp(h);
  }
  // left eigenvector matrix (rows = waves) and right (R[i][k])
  CFD_HD void mats(double L[4][4], double R[4][4]) const {
    CFD_CONTRACT
    const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
    L[0][0] = -ny * u + nx * v; L[0][1] = ny; L[0][2] = -nx; L[0][3] = 0.0;
    L[1][0] = h - ek; L[1][1] = u; L[1][2] = v; L[1][3] = -1.0;
    L[2][0] = 0.5 * (Vn * ic + ek * ih); L[2][1] = 0.5 * (-nx * ic - u * ih); L[2][2] = 0.5 * (-ny * ic - v * ih); L[2][3] = 0.5 * ih;
    L[3][0] = 0.5 * (-Vn * ic + ek * ih); L[3][1] = 0.5 * (nx * ic - u * ih); L[3][2] = 0.5 * (ny * ic - v * ih); L[3][3] = 0.5 * ih;
    R[0][0] = 0.0; R[0][1] = ih; R[0][2] = 1.0; R[0][3] = 1.0;
    R[1][0] = ny; R[1][1] = u * ih; R[1][2] = u - nx * c; R[1][3] = u + nx * c;
    R[2][0] = -nx; R[2][1] = v * ih; R[2][2] = v - ny * c; R[2][3] = v + ny * c;
    R[3][0] = ny * u - nx * v; R[3][1] = ek * ih; R[3][2] = h + ek - Vn * c; R[3][3] = h + ek + Vn * c;
  }
  // row k of L, column k of R (the entries of mats)
  CFD_HD void lrow(int k, double* l) const {
    CFD_CONTRACT
    const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
    if (k == 0) {
      l[0] = -ny * u + nx * v; l[1] = ny; l[2] = -nx; l[3] = 0.0;
    } else if (k == 1) {
      l[0] = h - ek; l[1] = u; l[2] = v; l[3] = -1.0;
    } else if (k == 2) {
      l[0] = 0.5 * (Vn * ic + ek * ih); l[1] = 0.5 * (-nx * ic - u * ih); l[2] = 0.5 * (-ny * ic - v * ih); l[3] = 0.5 * ih;
    } else {
      l[0] = 0.5 * (-Vn * ic + ek * ih); l[1] = 0.5 * (nx * ic - u * ih); l[2] = 0.5 * (ny * ic - v * ih); l[3] = 0.5 * ih;
    }
  }
  CFD_HD void rcol(int k, double* r) const {
    CFD_CONTRACT
    const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
    if (k == 0) {
      r[0] = 0.0; r[1] = ny; r[2] = -nx; r[3] = ny * u - nx * v;
    } else if (k == 1) {
      r[0] = ih; r[1] = u * ih; r[2] = v * ih; r[3] = ek * ih;
    } else if (k == 2) {
      r[0] = 1.0; r[1] = u - nx * c; r[2] = v - ny * c; r[3] = h + ek - Vn * c;
    } else {
      r[0] = 1.0; r[1] = u + nx * c; r[2] = v + ny * c; r[3] = h + ek + Vn * c;
    }
  }
  // Reverse pass in pieces, for register-lean callers: cotangents of the
  // Roe-averaged quantities are accumulated entry by entry of L and R, then
  // chained back to the two cell states.
  struct Acc {
    double lu = 0, lv = 0, lh = 0, lc = 0, lek = 0, lVn = 0;
  };
  // g = cotangent of row k of L
  CFD_HD void acc_Lrow(int k, const double* g, Acc& a) const {
    CFD_CONTRACT
    const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
    const double ic2 = ic * ic, ih2 = ih * ih;
    if (k == 0) {
      a.lu += -ny * g[0];
      a.lv += nx * g[0];
    } else if (k == 1) {
      a.lh += g[0];
      a.lek += -g[0];
      a.lu += g[1];
      a.lv += g[2];
    } else {
      const double sg = k == 2 ? 1.0 : -1.0;  // rows Vn - c and Vn + c differ in the 1/c terms
      a.lVn += sg * g[0] * 0.5 * ic;
      a.lc += -sg * g[0] * 0.5 * Vn * ic2;
      a.lek += g[0] * 0.5 * ih;
      a.lh += -g[0] * 0.5 * ek * ih2;
      a.lc += sg * g[1] * 0.5 * nx * ic2;
      a.lu += -g[1] * 0.5 * ih;
      a.lh += g[1] * 0.5 * u * ih2;
      a.lc += sg * g[2] * 0.5 * ny * ic2;
      a.lv += -g[2] * 0.5 * ih;
      a.lh += g[2] * 0.5 * v * ih2;
      a.lh += -g[3] * 0.5 * ih2;
    }
  }
  // gR(i, k) = cotangent of R[i][k]
  template <class GR> CFD_HD void acc_R(const GR& gR, Acc& a) const {
    CFD_CONTRACT
    const double nx = AX == 0 ? 1.0 : 0.0, ny = AX == 1 ? 1.0 : 0.0;
    const double ih2 = ih * ih;
    const double r30 = gR(3, 0), r01 = gR(0, 1), r11 = gR(1, 1), r21 = gR(2, 1), r31 = gR(3, 1);
    const double r12 = gR(1, 2), r13 = gR(1, 3), r22 = gR(2, 2), r23 = gR(2, 3), r32 = gR(3, 2), r33 = gR(3, 3);
    a.lu += ny * r30;
    a.lv += -nx * r30;
    a.lh += -(r01 + r11 * u + r21 * v + r31 * ek) * ih2;
    a.lu += r11 * ih;
    a.lv += r21 * ih;
    a.lek += r31 * ih;
    a.lu += r12 + r13;
    a.lv += r22 + r23;
    a.lc += -nx * r12 + nx * r13 - ny * r22 + ny * r23;
    a.lh += r32 + r33;
    a.lek += r32 + r33;
    a.lVn += -c * r32 + c * r33;
    a.lc += -Vn * r32 + Vn * r33;
  }
  CFD_HD void finish(const Acc& a0, const double* pl, const double* pr, double* lpl, double* lpr) const {
    CFD_CONTRACT
    double lu = a0.lu, lv = a0.lv, lh = a0.lh, lek = a0.lek;
    // c = sqrt(G1 h); h = ht - ek; ek = (u^2+v^2)/2; Vn = u | v
    lh += a0.lc * 0.5 * G1 * ic;
    const double lht = lh;
    lek += -lh;
    lu += lek * u;
    lv += lek * v;
    if (AX == 0) lu += a0.lVn;
    else lv += a0.lVn;
    // u = (c1 ul + c2 ur)/D etc.
    const double lc1 = (lu * (pl[1] - u) + lv * (pl[2] - v) + lht * (HL - ht)) * iD;
    const double lc2 = (lu * (pr[1] - u) + lv * (pr[2] - v) + lht * (HR - ht)) * iD;
    lpl[1] += lu * c1 * iD;
    lpr[1] += lu * c2 * iD;
    lpl[2] += lv * c1 * iD;
    lpr[2] += lv * c2 * iD;
    const double lHL = lht * c1 * iD, lHR = lht * c2 * iD;
    lpl[0] += lc1 * 0.5 * rcp(c1);
    lpr[0] += lc2 * 0.5 * rcp(c2);
    // HL = (ul^2+vl^2)/2 + GG1 pl/rl
    lpl[1] += lHL * pl[1];
    lpl[2] += lHL * pl[2];
    const double ipl0 = rcp(pl[0]), ipr0 = rcp(pr[0]);
    lpl[3] += lHL * GG1 * ipl0;
    lpl[0] += -lHL * GG1 * pl[3] * ipl0 * ipl0;
    lpr[1] += lHR * pr[1];
    lpr[2] += lHR * pr[2];
    lpr[3] += lHR * GG1 * ipr0;
    lpr[0] += -lHR * GG1 * pr[3] * ipr0 * ipr0;
  }
  // cotangents of L and R -> cotangents of pl, pr (added)
  CFD_HD void bwd(const double* pl, const double* pr, const double gL[4][4], const double gR[4][4], double* lpl,
                  double* lpr) const {
    Acc a;
    for (int k = 0; k < 4; k++) acc_Lrow(k, gL[k], a);
    acc_R([&](int i, int k) { return gR[i][k]; }, a);
    finish(a, pl, pr, lpl, lpr);
  }
};

// ---- interface flux: cells w[0..5] (primitive), interface between w[2], w[3] ------
// Forward (for checks) and reverse: lw[j][v] += (dF/dw[j][v])^T lF.
template <int AX, class IP = JS>
CFD_HD inline void interface(const double w[6][4], double* F, double lw[6][4] = nullptr, const double* lF = nullptr,
                             const IP& proto = IP()) {
  Roe2<AX> es;
  es.fwd(w[2], w[3]);
  double L[4][4], R[4][4];
  es.mats(L, R);
  double c[6][4];
  for (int j = 0; j < 6; j++) p2c(w[j], c[j]);
  double ch[4][6];
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 6; j++) ch[k][j] = L[k][0] * c[j][0] + L[k][1] * c[j][1] + L[k][2] * c[j][2] + L[k][3] * c[j][3];
  IP jl[4], jr[4];
  double QL[4], QR[4];
  for (int k = 0; k < 4; k++) {
    double ql[5] = {ch[k][0], ch[k][1], ch[k][2], ch[k][3], ch[k][4]};
    double qr[5] = {ch[k][5], ch[k][4], ch[k][3], ch[k][2], ch[k][1]};
    jl[k] = proto;
    jr[k] = proto;
    jl[k].fwd(ql);
    jr[k].fwd(qr);
    QL[k] = jl[k].out;
    QR[k] = jr[k].out;
  }
  double cL[4], cR[4];
  for (int i = 0; i < 4; i++) {
    cL[i] = R[i][0] * QL[0] + R[i][1] * QL[1] + R[i][2] * QL[2] + R[i][3] * QL[3];
    cR[i] = R[i][0] * QR[0] + R[i][1] * QR[1] + R[i][2] * QR[2] + R[i][3] * QR[3];
  }
  double oL[4], oR[4];
  c2p(cL, oL);
  c2p(cR, oR);
  const bool fbL = oL[0] < 0 || oL[3] < 0, fbR = oR[0] < 0 || oR[3] < 0;
  if (fbL)
    for (int v = 0; v < 4; v++) oL[v] = w[2][v];
  if (fbR) {
    oR[0] = w[3][0];
    oR[1] = w[2][1];
    oR[2] = w[2][2];
    oR[3] = w[3][3];
  }
  if (!lF) {
    hllc<AX>(oL, oR, F);
    return;
  }
  double loL[4] = {0, 0, 0, 0}, loR[4] = {0, 0, 0, 0};
  double Fd[4];
  hllc<AX>(oL, oR, Fd, loL, loR, lF);
  double lcL[4] = {0, 0, 0, 0}, lcR[4] = {0, 0, 0, 0};
  if (fbL) {
    for (int v = 0; v < 4; v++) lw[2][v] += loL[v];
  } else {
    c2p_vjp(cL, loL, lcL);
  }
  if (fbR) {
    lw[3][0] += loR[0];
    lw[2][1] += loR[1];
    lw[2][2] += loR[2];
    lw[3][3] += loR[3];
  } else {
    c2p_vjp(cR, loR, lcR);
  }
  double gR[4][4], gL[4][4];
  double lQL[4], lQR[4];
  for (int k = 0; k < 4; k++) {
    lQL[k] = R[0][k] * lcL[0] + R[1][k] * lcL[1] + R[2][k] * lcL[2] + R[3][k] * lcL[3];
    lQR[k] = R[0][k] * lcR[0] + R[1][k] * lcR[1] + R[2][k] * lcR[2] + R[3][k] * lcR[3];
    for (int i = 0; i < 4; i++) gR[i][k] = QL[k] * lcL[i] + QR[k] * lcR[i];
  }
  double lch[4][6];
  for (int k = 0; k < 4; k++) {
    for (int j = 0; j < 6; j++) lch[k][j] = 0.0;
    double ql[5] = {ch[k][0], ch[k][1], ch[k][2], ch[k][3], ch[k][4]};
    double qr[5] = {ch[k][5], ch[k][4], ch[k][3], ch[k][2], ch[k][1]};
    double gl[5] = {0, 0, 0, 0, 0}, gr[5] = {0, 0, 0, 0, 0};
    jl[k].bwd(ql, lQL[k], gl);
    jr[k].bwd(qr, lQR[k], gr);
    for (int j = 0; j < 5; j++) lch[k][j] += gl[j];
    for (int j = 0; j < 5; j++) lch[k][5 - j] += gr[j];
  }
  double lc[6][4];
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) lc[j][v] = lch[0][j] * L[0][v] + lch[1][j] * L[1][v] + lch[2][j] * L[2][v] + lch[3][j] * L[3][v];
  for (int k = 0; k < 4; k++)
    for (int v = 0; v < 4; v++) {
      double s = 0;
      for (int j = 0; j < 6; j++) s += lch[k][j] * c[j][v];
      gL[k][v] = s;
    }
  es.bwd(w[2], w[3], gL, gR, lw[2], lw[3]);
  for (int j = 0; j < 6; j++) p2c_vjp(w[j], lc[j], lw[j]);
}

// ---- ghost cells ------------------------------------------------------------------
// Source of ghost cell g (g < 0 or g >= n) of a pencil of n cells under the
// boundary code bc, as the forward sweep fills it: the cell it copies, the
// primitive component it negates (-1: none) and whether it is a fixed state
// (Dirichlet: no cotangent flows back).  Codes as in congfd.h.
struct GhostSrc {
  int src, flip;
  bool fixed;
};
CFD_HD inline GhostSrc ghost_source(int bc, int g, int n) {
  const bool lo = g < 0;
  GhostSrc s{lo ? -1 - g : 2 * n - 1 - g, -1, false};
  switch (bc) {
  case 1: s.src = ((g % n) + n) % n; break;       // periodic
  case 7: s.src = lo ? 0 : n - 1; break;           // outflow
  case 3: s.fixed = true; break;                   // Dirichlet
  case 2:
  case 10: s.flip = 1; break;                      // symmetry, u negated
  case 11: s.flip = 2; break;                      // symmetry, v negated
  default: break;                                  // mirror copy
  }
  return s;
}

// ---- interface flux as a forward record + reverse pass ----------------------------
// What the reverse pass of one interface needs from its forward evaluation.
struct IfaceFwd {
  double QL[4], QR[4];      // characteristic half-node values
  double cL[4], cR[4];      // conservative half-node states
  double oL[4], oR[4];      // primitive states given to the flux
  bool fbL, fbR;            // positivity fallback taken
  int hb;                   // HLLC branch bits, or -1: decided by the reverse pass
  unsigned tflag;           // TENO5 flags: bits 6k..6k+2 the left value of wave k, 6k+3.. the right
};

// HLLC branch bits in HLLCFlux2D2's own arithmetic (fluxSchemes.hpp:202-327),
// expression for expression, so that with the forward's half-node states the
// bits are the branches the forward took.  Requires -ffp-contract=off.
template <int AX> CFD_HD inline int hllc_branch_bits(const double* Lp, const double* Rp) {
  const double rl = Lp[0], rr = Rp[0], pl = Lp[3], pr = Rp[3];
  const double VnL = Lp[1 + AX], VnR = Rp[1 + AX];
  const double cl = sqrt(pl / rl * 1.4), cr = sqrt(pr / rr * 1.4);
  const double rBar = (rl + rr) / 2, cBar = (cl + cr) / 2;
  const double psr = (pl + pr) / 2 - rBar * cBar * (VnR - VnL) / 2;
  const bool clip = !(0.0 < psr);
  const double ps = clip ? 0.0 : psr;
  const bool shl = !(ps <= pl), shr = !(ps <= pr);
  const double kap = (1.4 + 1.0) / (2.0 * 1.4);
  const double ql = shl ? sqrt(1.0 + kap * (ps / pl - 1.0)) : 1.0;
  const double qr = shr ? sqrt(1.0 + kap * (ps / pr - 1.0)) : 1.0;
  const double SL = VnL - cl * ql, SR = VnR + cr * qr;
  const double Ss = ((pr - pl) + (rl * VnL * (SL - VnL) - rr * VnR * (SR - VnR))) / (rl * (SL - VnL) - rr * (SR - VnR));
  int reg;
  if (SL > 0) reg = 0;
  else if (Ss > 0) reg = 1;
  else if (Ss == 0) reg = 2;
  else if (SR > 0) reg = 3;
  else reg = 4;
  return (clip ? HB_CLIP : 0) | (shl ? HB_SHL : 0) | (shr ? HB_SHR : 0) | (reg << 3);
}

// Forward record in this file's own arithmetic (host checks).  The six
// cells are read through w(j, v) (primitive) and w.cons(j, v) (the
// conservative state rebuilt from it); the interface lies between cells 2 and 3.
template <int AX, class IP = JS, class W> CFD_HD inline void iface_forward(const W& w, IfaceFwd& f,
                                                                           const IP& proto = IP()) {
  double pl[4], pr[4];
  for (int v = 0; v < 4; v++) {
    pl[v] = w(2, v);
    pr[v] = w(3, v);
  }
  Roe2<AX> es;
  es.fwd(pl, pr);
  double L[4][4], R[4][4];
  es.mats(L, R);
  f.tflag = 0u;
  for (int k = 0; k < 4; k++) {
    double ch[6];
    for (int j = 0; j < 6; j++)
      ch[j] = L[k][0] * w.cons(j, 0) + L[k][1] * w.cons(j, 1) + L[k][2] * w.cons(j, 2) + L[k][3] * w.cons(j, 3);
    double ql[5] = {ch[0], ch[1], ch[2], ch[3], ch[4]};
    double qr[5] = {ch[5], ch[4], ch[3], ch[2], ch[1]};
    IP a = proto, b = proto;
    a.fwd(ql);
    b.fwd(qr);
    f.QL[k] = a.out;
    f.QR[k] = b.out;
    if constexpr (IP::kLinear) f.tflag |= (a.flag | (b.flag << 3)) << (6 * k);
  }
  for (int i = 0; i < 4; i++) {
    f.cL[i] = R[i][0] * f.QL[0] + R[i][1] * f.QL[1] + R[i][2] * f.QL[2] + R[i][3] * f.QL[3];
    f.cR[i] = R[i][0] * f.QR[0] + R[i][1] * f.QR[1] + R[i][2] * f.QR[2] + R[i][3] * f.QR[3];
  }
  c2p(f.cL, f.oL);
  c2p(f.cR, f.oR);
  f.fbL = f.oL[0] < 0 || f.oL[3] < 0;
  f.fbR = f.oR[0] < 0 || f.oR[3] < 0;
  if (f.fbL)
    for (int v = 0; v < 4; v++) f.oL[v] = pl[v];
  if (f.fbR) {
    f.oR[0] = pr[0];
    f.oR[1] = pl[1];
    f.oR[2] = pl[2];
    f.oR[3] = pr[3];
  }
  f.hb = -1;
}

// Forward record from what the solver's recording sweep kept of the
// interface (sweep_kernel_rec): the characteristic half-node values Q[0..3]
// (left) and Q[4..7] (right) and the branch word (TENO flags in bits 0..23,
// positivity fallback left / right in 24 / 25, HLLC branch bits from 26).
// The branches are the solver's own; the states rebuilt from Q only feed
// derivatives, so they are evaluated in this file's contracted arithmetic.
// es: the Roe-averaged eigensystem of the interface's two cells, built here
// and handed on to the reverse pass.
template <int AX, class W>
CFD_HD inline void iface_from_record(const W& w, const double* Q, unsigned bits, IfaceFwd& f, Roe2<AX>& es) {
  CFD_CONTRACT
  double pl[4], pr[4];
  for (int v = 0; v < 4; v++) {
    pl[v] = w(2, v);
    pr[v] = w(3, v);
  }
  es.fwd(pl, pr);
  for (int i = 0; i < 4; i++) {
    f.QL[i] = Q[i];
    f.QR[i] = Q[4 + i];
    f.cL[i] = 0.0;
    f.cR[i] = 0.0;
  }
  for (int k = 0; k < 4; k++) {
    double r[4];
    es.rcol(k, r);
    for (int i = 0; i < 4; i++) {
      f.cL[i] += f.QL[k] * r[i];
      f.cR[i] += f.QR[k] * r[i];
    }
  }
  c2p(f.cL, f.oL);
  c2p(f.cR, f.oR);
  f.tflag = bits & 0xffffffu;
  f.fbL = (bits >> 24) & 1u;
  f.fbR = (bits >> 25) & 1u;
  f.hb = (int)(bits >> 26);
  if (f.fbL)
    for (int v = 0; v < 4; v++) f.oL[v] = pl[v];
  if (f.fbR) {
    f.oR[0] = pr[0];
    f.oR[1] = pl[1];
    f.oR[2] = pl[2];
    f.oR[3] = pr[3];
  }
}

// Reverse pass of the interface flux given its forward record: lw[j][v] +=
// (dF / dw[j][v])^T lF.  Register-lean for device use: the per-wave WENO
// intermediates and projections are recomputed rather than held, the
// eigenvector cotangents are folded into six scalars as they appear, and the
// projection cotangents go straight through p2c's transpose wave by wave.
template <int AX, class IP = JS, class W>
CFD_HD inline void iface_backward(const W& w, const IfaceFwd& f, const double* lF, double lw[6][4],
                                  const Roe2<AX>& es) {
  CFD_CONTRACT
  double loL[4] = {0, 0, 0, 0}, loR[4] = {0, 0, 0, 0};
  {
    double Fd[4];
    hllc<AX>(f.oL, f.oR, Fd, loL, loR, lF, f.hb);
  }
  double lcL[4] = {0, 0, 0, 0}, lcR[4] = {0, 0, 0, 0};
  if (f.fbL) {
    for (int v = 0; v < 4; v++) lw[2][v] += loL[v];
  } else {
    c2p_vjp(f.cL, loL, lcL);
  }
  if (f.fbR) {
    lw[3][0] += loR[0];
    lw[2][1] += loR[1];
    lw[2][2] += loR[2];
    lw[3][3] += loR[3];
  } else {
    c2p_vjp(f.cR, loR, lcR);
  }
  double pl[4], pr[4];
  for (int v = 0; v < 4; v++) {
    pl[v] = w(2, v);
    pr[v] = w(3, v);
  }
  typename Roe2<AX>::Acc acc;
  es.acc_R([&](int i, int k) { return f.QL[k] * lcL[i] + f.QR[k] * lcR[i]; }, acc);
  double lQL[4], lQR[4];
  CFD_UNROLL
  for (int k = 0; k < 4; k++) {
    double r[4];
    es.rcol(k, r);
    lQL[k] = r[0] * lcL[0] + r[1] * lcL[1] + r[2] * lcL[2] + r[3] * lcL[3];
    lQR[k] = r[0] * lcR[0] + r[1] * lcR[1] + r[2] * lcR[2] + r[3] * lcR[3];
  }
  // conservative cotangents of the six cells, summed over the waves before
  // the transpose of prim -> cons (linear in them) is applied once per cell
  double lc[6][4];
  CFD_UNROLL
  for (int j = 0; j < 6; j++)
    for (int v = 0; v < 4; v++) lc[j][v] = 0.0;
  CFD_UNROLL
  for (int k = 0; k < 4; k++) {
    double Lk[4];
    es.lrow(k, Lk);
    double gl[5] = {0, 0, 0, 0, 0}, gr[5] = {0, 0, 0, 0, 0};
    // row 0 (the shear wave) is (v, 0, -1, 0) for an x normal and (-u, 1, 0, 0)
    // for a y normal: its zero entries are left out of the products below
    auto proj = [&](int j) -> double {
      if (k == 0)
        return AX == 0 ? Lk[0] * w.cons(j, 0) - w.cons(j, 2) : Lk[0] * w.cons(j, 0) + w.cons(j, 1);
      return Lk[0] * w.cons(j, 0) + Lk[1] * w.cons(j, 1) + Lk[2] * w.cons(j, 2) + Lk[3] * w.cons(j, 3);
    };
    if constexpr (IP::kLinear) {  // the recorded flags' stencils
      IP a, b;
      a.flag = (f.tflag >> (6 * k)) & 7u;
      b.flag = (f.tflag >> (6 * k + 3)) & 7u;
      a.bwd(nullptr, lQL[k], gl);
      b.bwd(nullptr, lQR[k], gr);
    } else {
      double ch[6];
      CFD_UNROLL
      for (int j = 0; j < 6; j++) ch[j] = proj(j);
      double ql[5] = {ch[0], ch[1], ch[2], ch[3], ch[4]};
      double qr[5] = {ch[5], ch[4], ch[3], ch[2], ch[1]};
      {
        IP a;
        a.fwd(ql);
        a.bwd(ql, lQL[k], gl);
      }
      {
        IP b;
        b.fwd(qr);
        b.bwd(qr, lQR[k], gr);
      }
    }
    double lch[6];
    lch[0] = gl[0];
    for (int j = 1; j < 5; j++) lch[j] = gl[j] + gr[5 - j];
    lch[5] = gr[0];
    double gLk[4] = {0, 0, 0, 0};
    CFD_UNROLL
    for (int j = 0; j < 6; j++) {
      for (int v = 0; v < 4; v++) gLk[v] += lch[j] * w.cons(j, v);  // row 0: only gLk[0] is used
      if (k == 0) {
        lc[j][0] += lch[j] * Lk[0];
        if (AX == 0)
          lc[j][2] -= lch[j];
        else
          lc[j][1] += lch[j];
      } else {
        for (int v = 0; v < 4; v++) lc[j][v] += lch[j] * Lk[v];
      }
    }
    es.acc_Lrow(k, gLk, acc);
  }
  CFD_UNROLL
  for (int j = 0; j < 6; j++) {
    double p[4];
    for (int v = 0; v < 4; v++) p[v] = w(j, v);
    p2c_vjp(p, lc[j], lw[j]);
  }
  es.finish(acc, pl, pr, lw[2], lw[3]);
}
template <int AX, class IP = JS, class W>
CFD_HD inline void iface_backward(const W& w, const IfaceFwd& f, const double* lF, double lw[6][4]) {
  double pl[4], pr[4];
  for (int v = 0; v < 4; v++) {
    pl[v] = w(2, v);
    pr[v] = w(3, v);
  }
  Roe2<AX> es;
  es.fwd(pl, pr);
  iface_backward<AX, IP>(w, f, lF, lw, es);
}

// Same derivative as interface<AX>(w, F, lw, lF), from the two passes above.
template <int AX, class IP = JS, class W>
CFD_HD inline void interface_vjp_lean(const W& w, const double* lF, double lw[6][4], const IP& proto = IP()) {
  IfaceFwd f;
  iface_forward<AX>(w, f, proto);
  iface_backward<AX, IP>(w, f, lF, lw);
}

}  // namespace cfdadj
