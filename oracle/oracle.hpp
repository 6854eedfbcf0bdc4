// CPU oracle for the CongFD right-hand side and SSP-RK3 time advance.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference solver's arithmetic (wanghongfei23/CFD_inWSL, snapshot zone_FD,
// plus the CongFD1.0 accuracy-test path), kept in the reference's evaluation
// order so that a GPU result can be compared with it bit for bit.  Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it; the product library never links or calls it.
//
// Citations are relative to the reference repository root.
//
// Extensions that the reference does not have (3D, true periodic halo,
// component-wise Euler) are marked "EXTENSION" and are pinned by invariants
// (3D with w=0 equals 2D bit for bit, conservation, free-stream preservation,
// mirror symmetry) rather than by reference vectors.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

// The ratio of specific heats.  The reference defines it as the macro GAMMA
// (zone_FD/src/include/macro.hpp:28, CongFD1.0/src/include/macro.hpp:28-29):
// `1.4` in the default build, the unparenthesised text `5.0/3.0` in the build
// of its Rayleigh-Taylor runs.  The solver routines either copy it into a
// variable (real gamma = GAMMA: the eigensystems, consToPrim, the HLLC wave
// speeds) or expand the macro in place; only one expanded expression of the
// path rounds differently from its variable form: the enthalpy term
// p/r*GAMMA/(GAMMA-1) of the interface fluxes, which reads
// p/r*5.0/3.0/(5.0/3.0-1) = (((p/r)*5.0)/3.0)/(5.0/3.0-1) in that build.
// Every other expansion on the path -- (GAMMA-1)*x, GAMMA*p/r, GAMMA*(p/r),
// GAMMA/(GAMMA-1)*p, x/(GAMMA-1) -- groups as the variable form does.
// The oracle keeps the macro's text as (num, den): den == 0 for a single
// literal.  Set by each oracle entry point before any parallel region.
struct Gas {
  double num = 1.4, den = 0.0;
  double g = 1.4;  // the value of the macro
};
inline Gas g_gas;
inline double gam() { return g_gas.g; }
inline void set_gas(double num, double den) {
  g_gas.num = num;
  g_gas.den = den;
  g_gas.g = den == 0.0 ? num : num / den;
}
// p/r*GAMMA/(GAMMA-1) as the flux routines write it (fluxSchemes.hpp:29,150,224,445;
// CongFD1.0/src/include/fluxScheme.hpp:411-412, src/fluxScheme.cpp:82-83)
template <class T> inline T enth_macro(const T& p, const T& r) {
  if (g_gas.den == 0.0) return p / r * g_gas.g / (g_gas.g - 1);
  return p / r * g_gas.num / g_gas.den / (g_gas.num / g_gas.den - 1);
}

// Interpolation schemes (zone_FD/src/include/interScheme.hpp).
enum Scheme {
  S_JS = 0, S_Z = 1, S_TENO5 = 2, S_TENO5_CT4 = 3, S_TENO5_CT7 = 4, S_TENOS = 5, S_LINEAR = 6,
  S_TENOS_CT4 = 7, S_TENOS_CT7 = 8, S_TENO5_CT10 = 9, S_TENOA = 10,
  S_TENOSA_F002 = 11, S_TENOSA_F203 = 12, S_TENOSA_F202 = 13,
  S_WHF_N = 14, S_WHF_A = 15, S_WHF_AS = 16, S_WHF_S = 17, S_WHF_LAD = 18,
  S_AI_S = 19, S_AI_A = 20, S_AI_AS = 21, S_AI_LADS = 22, S_AI_AZS = 23, S_AI_F002 = 24, S_WHF_COMPARE = 25,
  S_CONGAA = 26, S_MUSCL5 = 27,
  // CongFD1.0's own forms of two research schemes (its InterMethod enum,
  // CongFD1.0/src/include/macro.hpp:45-87, differs from zone_FD's):
  // whf_TCNS_AS_myF203_NoS with its own constants and whf_TCNS_A
  S_C1_F203 = 28, S_C1_TENOA = 29
};
enum Flux { F_HLLC = 0, F_ROE = 1 };

template <class T> inline T sq(const T& a) { return a * a; }
// std::min / std::max semantics: min(a,b) = (b < a) ? b : a, max(a,b) = (a < b) ? b : a
template <class T> inline T smin(const T& a, const T& b) { return (b < a) ? b : a; }
template <class T> inline T smax(const T& a, const T& b) { return (a < b) ? b : a; }

// Smoothness indicators shared by every 5-point scheme
// (zone_FD/src/include/interScheme.hpp:20-27, 384-392).
template <class T> inline void betas(const T q[5], T b[3]) {
  using std::abs;
  b[0] = sq(q[0] - 2.0 * q[1] + q[2]) + 0.25 * sq(q[0] - 4.0 * q[1] + 3.0 * q[2]);
  b[1] = sq(q[1] - 2.0 * q[2] + q[3]) + 0.25 * sq(q[1] - q[3]);
  b[2] = sq(q[2] - 2.0 * q[3] + q[4]) + 0.25 * sq(3.0 * q[2] - 4.0 * q[3] + q[4]);
}

// Candidate stencils u0,u1,u2 (interScheme.hpp:29-32).
template <class T> inline T cand0(const T q[5]) { return 0.375 * q[0] - 1.25 * q[1] + 1.875 * q[2]; }
template <class T> inline T cand1(const T q[5]) { return -0.125 * q[1] + 0.75 * q[2] + 0.375 * q[3]; }
template <class T> inline T cand2(const T q[5]) { return 0.375 * q[2] + 0.75 * q[3] - 0.125 * q[4]; }

// TENO flag -> fixed stencil (interScheme.hpp:417-451).
template <class T> inline T teno_select(unsigned flag, const T q[5]) {
  switch (flag) {
  case 0: return 0.0234375 * q[0] - 0.15625 * q[1] + 0.703125 * q[2] + 0.46875 * q[3] - 0.0390625 * q[4];
  case 1: return -0.0625 * q[1] + 0.5625 * q[2] + 0.5625 * q[3] - 0.0625 * q[4];
  case 2:
  case 3: return cand2(q);
  case 4: return 0.0625 * q[0] - 0.3125 * q[1] + 0.9375 * q[2] + 0.3125 * q[3];
  case 5: return cand1(q);
  case 6: return cand0(q);
  default: return q[2];
  }
}

// weno5_JSchen (interScheme.hpp:17-43)
template <class T> inline T weno5_js(const T q[5]) {
  T b[3];
  betas(q, b);
  const double d[3] = {0.0625, 0.625, 0.3125};
  T u[3] = {cand0(q), cand1(q), cand2(q)};
  T a[3];
  for (int i = 0; i < 3; i++) a[i] = d[i] / sq(1e-6 + b[i]);
  T s = a[0] + a[1] + a[2];
  T r = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
  return r / s;
}

// weno5_Z (interScheme.hpp:85-114)
template <class T> inline T weno5_z(const T q[5]) {
  using std::abs;
  T b[3];
  betas(q, b);
  const double d[3] = {0.0625, 0.625, 0.3125};
  T u[3] = {cand0(q), cand1(q), cand2(q)};
  T tau = abs(b[2] - b[0]);
  T a[3];
  for (int i = 0; i < 3; i++) a[i] = d[i] * (1.0 + sq(tau / (b[i] + 1e-40)));
  T s = a[0] + a[1] + a[2];
  T r = a[0] * u[0] + a[1] * u[1] + a[2] * u[2];
  return r / s;
}

// Teno5_Z (interScheme.hpp:384-452); CT4 / CT7 variants (:116, :184) differ
// only in the cut-off constant.
template <class T> inline T teno5(const T q[5], double ct) {
  using std::abs;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  T g[3];
  for (int i = 0; i < 3; i++) {
    T t = 1.0 + tau / (b[i] + 1e-40);
    t = t * t;
    g[i] = t * t * t;
  }
  T CT = ct * (g[0] + g[1] + g[2]);
  unsigned flag = 0;
  if (g[0] < CT) flag += 1;
  if (g[1] < CT) flag += 2;
  if (g[2] < CT) flag += 4;
  return teno_select(flag, q);
}

// Teno5_CongZ, "TENO-S" (interScheme.hpp:494-585); the CT4 / CT7 variants
// Teno5_CongZCT4 / CT7 (:586, :661) differ only in the constant CT.
template <class T> inline T tenos(const T q[5], const double CT = 0.15704178024750198) {
  using std::abs;
  T b[3];
  betas(q, b);
  int m = 0;  // std::min_element: first minimal index
  if (b[1] < b[m]) m = 1;
  if (b[2] < b[m]) m = 2;
  const double CT_1 = 1 - CT;
  T tau = abs(b[2] - b[0]);
  T rr = CT * tau - CT_1 * b[m];
  T ll = tau * b[m];
  unsigned flag = 0;
  if (ll < rr * b[0]) flag += 1;
  if (ll < rr * b[1]) flag += 2;
  if (ll < rr * b[2]) flag += 4;
  return teno_select(flag, q);
}

// new_whf_TCNS_A, "TENO-A" (interScheme.hpp:3863-3983; InterMethod
// new_WHFTCNSA, named "TENO-A" in info.cpp): Teno5_Z's cut-off with the
// adaptive constant CT = 10^-beta_A, beta_A = floor(10 - 5 (1 - g)),
// g = s^4 (5 - 4 s), s = min(1, 4.1666667 eta_min), eta the smoothness ratios
// of adjacent differences.
// CongFD1.0's whf_TCNS_A (CongFD1.0/src/include/interScheme.hpp:1480-1594) is
// the same function with s = min(1, 4.166667 eta_min).
template <class T> inline T tenoa(const T q[5], double inv_cr = 4.1666667) {
  using std::abs;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  const double epsA = 2.7551 * 1e-7;
  const T d0 = q[0] - q[1], d1 = q[1] - q[2], d2 = q[2] - q[3], d3 = q[3] - q[4];
  const T e0 = (abs(2.0 * d0 * d1) + epsA) / (d0 * d0 + d1 * d1 + epsA);
  const T e1 = (abs(2.0 * d1 * d2) + epsA) / (d1 * d1 + d2 * d2 + epsA);
  const T e2 = (abs(2.0 * d2 * d3) + epsA) / (d2 * d2 + d3 * d3 + epsA);
  T emin = e0;
  if (e1 < emin) emin = e1;
  if (e2 < emin) emin = e2;
  const T x = inv_cr * emin;
  const T mn = x < 1.0 ? x : T(1.0);
  const T gm = mn * mn * mn * mn * (5.0 - 4 * mn);
  const int betaA = (int)std::floor((double)(10.0 - 5.0 * (1.0 - gm)));
  T g[3];
  for (int i = 0; i < 3; i++) {
    T t = 1.0 + tau / (b[i] + 1e-40);
    t = t * t;
    g[i] = t * t * t;
  }
  const double cts[6] = {1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10};
  const double CTA = cts[betaA < 5 ? 0 : (betaA > 10 ? 5 : betaA - 5)];
  T rr = CTA * (g[0] + g[1] + g[2]);
  unsigned flag = 0;
  if (g[0] < rr) flag += 1;
  if (g[1] < rr) flag += 2;
  if (g[2] < rr) flag += 4;
  return teno_select(flag, q);
}

// TENO-S with C_T adapted to eta_min (the smoothness ratios of TENO-A):
//   law 1, whf_zyc_TcnsN_myASF002_1 (interScheme.hpp:3363-3460):
//     s = std::min(1.0, 4.1666667 eta_min), C_T = 0.157/(11.351092 s s - 5.740650 s + 1.506631)
//   law 2, whf_TCNS_AS_myF203_NoS (:3984-4091): m = 1 - std::min(0.24, eta_min)/0.24,
//     C_T = -0.452662 m^3 + 0.590811 m^2 - 0.012189 m + 0.022010 (mm = m m, mmm = m mm)
//   law 3, whf_TCNS_AS_myF202_NoS (:4093-4200): the same m, C_T = -0.088182 m m + 0.259395 m - 0.000616
//   law 4, CongFD1.0's whf_TCNS_AS_myF203_NoS (CongFD1.0/src/include/interScheme.hpp:3170-3279):
//     law 2 with the constants -0.45, 0.59, -0.012, 0.022
// then TENO-S's cut-off with that C_T.  The zyc form writes the differences
// with the opposite sign (q_{i} - q_{i-1}); |2 a b| and a^2 + b^2 are the
// same doubles either way.
template <class T> inline T tenosa(const T q[5], int law) {
  using std::abs;
  T b[3];
  betas(q, b);
  int mb = 0;  // std::min_element: first minimal index
  if (b[1] < b[mb]) mb = 1;
  if (b[2] < b[mb]) mb = 2;
  const double epsA = 2.7551 * 1e-7;
  const T d0 = q[0] - q[1], d1 = q[1] - q[2], d2 = q[2] - q[3], d3 = q[3] - q[4];
  const T e0 = (abs(2.0 * d1 * d0) + epsA) / (d1 * d1 + d0 * d0 + epsA);
  const T e1 = (abs(2.0 * d2 * d1) + epsA) / (d2 * d2 + d1 * d1 + epsA);
  const T e2 = (abs(2.0 * d3 * d2) + epsA) / (d3 * d3 + d2 * d2 + epsA);
  T emin = e0;  // std::min({..}): first minimal
  if (e1 < emin) emin = e1;
  if (e2 < emin) emin = e2;
  T CT;
  if (law == 1) {
    const T x = 4.1666667 * emin;
    const T s = x < 1.0 ? x : T(1.0);
    CT = 0.157 / (11.351092 * s * s - 5.740650 * s + 1.506631);
  } else {
    const T mn = emin < 0.24 ? emin : T(0.24);
    const T m = 1 - mn / 0.24;
    if (law == 2) {
      const T mm = m * m, mmm = m * mm;
      CT = -0.452662 * mmm + 0.590811 * mm - 0.012189 * m + 0.022010;
    } else if (law == 4) {
      const T mm = m * m, mmm = m * mm;
      CT = -0.45 * mmm + 0.59 * mm - 0.012 * m + 0.022;
    } else {
      CT = -0.088182 * m * m + 0.259395 * m - 0.000616;
    }
  }
  const T CT_1 = 1 - CT;
  T tau = abs(b[2] - b[0]);
  T rr = CT * tau - CT_1 * b[mb];
  T ll = tau * b[mb];
  unsigned flag = 0;
  if (ll < rr * b[0]) flag += 1;
  if (ll < rr * b[1]) flag += 2;
  if (ll < rr * b[2]) flag += 4;
  return teno_select(flag, q);
}

inline double value_of(double x) { return x; }

// ---- the whf_TcnsN family (InterMethod whfTCNSN .. whfTCNSNLAD = 17..21) ----
// The interpolated value depends on the data only through the flag, so the
// libm pow calls below decide bits only where a ratio ties its threshold.

// whf_TcnsN (interScheme.hpp:1458-1517): chi_k = gamma_k / sum gamma,
// gamma_k = pow(1 + tau / (beta_k + 1e-40), 6); cut where chi_k < 1e-10.
template <class T> inline T whf_n(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  T g[3];
  for (int i = 0; i < 3; i++) g[i] = pow(1.0 + tau / (b[i] + 1e-40), 6);
  T gs = g[0] + g[1] + g[2];
  unsigned flag = 0;
  if (g[0] / gs < 1e-10) flag += 1;
  if (g[1] / gs < 1e-10) flag += 2;
  if (g[2] / gs < 1e-10) flag += 4;
  return teno_select(flag, q);
}

// whf_TcnsN_A (interScheme.hpp:3253-3362): Teno5_Z's cut-off with
// CT = 10^-beta_A, beta_A = floor(10 - 5 (1 - g)), g = (1-m)^4 (1+4m),
// m = 1 - min(1, 4.1666667 eta_min); its eta pair the differences
// (d0, d2), (d1, d3), (d2, d1).
template <class T> inline T whf_a(const T q[5]) {
  using std::abs;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  const double epsA = 2.7551 * 1e-7;
  const T d0 = q[0] - q[1], d1 = q[1] - q[2], d2 = q[2] - q[3], d3 = q[3] - q[4];
  const T e0 = (abs(2.0 * d0 * d2) + epsA) / (d0 * d0 + d2 * d2 + epsA);
  const T e1 = (abs(2.0 * d1 * d3) + epsA) / (d1 * d1 + d3 * d3 + epsA);
  const T e2 = (abs(2.0 * d2 * d1) + epsA) / (d2 * d2 + d1 * d1 + epsA);
  T emin = e0;
  if (e1 < emin) emin = e1;
  if (e2 < emin) emin = e2;
  const T x = 4.1666667 * emin;
  const T m = 1.0 - (x < 1.0 ? x : T(1.0));
  const T gm = (1 - m) * (1 - m) * (1 - m) * (1 - m) * (1 + 4 * m);
  const int betaA = (int)std::floor((double)(10.0 - 5.0 * (1 - gm)));
  const double cts[6] = {1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10};
  const double CTA = cts[betaA < 5 ? 0 : (betaA > 10 ? 5 : betaA - 5)];
  T g[3];
  for (int i = 0; i < 3; i++) {
    T t = 1.0 + tau / (b[i] + 1e-40);
    t = t * t;
    g[i] = t * t * t;
  }
  T rr = CTA * (g[0] + g[1] + g[2]);
  unsigned flag = 0;
  if (g[0] < rr) flag += 1;
  if (g[1] < rr) flag += 2;
  if (g[2] < rr) flag += 4;
  return teno_select(flag, q);
}

// The cut-off of whf_TcnsN_S / _AS: stencil k is cut where
// beta_k tau < (CT_p tau - (1 - CT_p) beta_k) beta_min.
template <class T> inline T whf_s_cut(const T q[5], const T b[3], double CTp) {
  using std::abs;
  T tau = abs(b[2] - b[0]);
  T bm = b[0];  // std::min({..})
  if (b[1] < bm) bm = b[1];
  if (b[2] < bm) bm = b[2];
  unsigned flag = 0;
  for (int k = 0; k < 3; k++)
    if (b[k] * tau < (CTp * tau - 1.0 * (1.0 - CTp) * b[k]) * bm) flag += 1u << k;
  return teno_select(flag, q);
}

// CT_p = pow(1.5 CT / (1 - CT), 1/6) (interScheme.hpp:1538, :1637)
inline double whf_ct_p(double CT) { return std::pow(1.5 * CT / (1.0 - CT), 1.0 / 6); }

// whf_TcnsN_S (interScheme.hpp:1519-1579): CT = 1e-5.
template <class T> inline T whf_s(const T q[5]) {
  T b[3];
  betas(q, b);
  return whf_s_cut(q, b, whf_ct_p(1e-5));
}

// whf_TcnsN_AS (interScheme.hpp:1581-1679): CT = pow(10, -beta_A) with
// beta_A = floor(10 - 5 (1 - pow(1 - m, 4) (1 + 4 m))), m = 1 - min(1,
// eta_min / 0.24); its eta are taken of the values q themselves (not of their
// differences), eps_A = (0.9 0.24) / (1 - 0.9 0.24) 1e-3 1e-3.
template <class T> inline T whf_as(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  const double epsA = (0.9 * 0.24) / (1.0 - 0.9 * 0.24) * 1e-3 * 1e-3;
  auto eta = [&](T a, T bb, T c) {
    T e1 = (abs(2 * bb * a) + epsA) / (bb * bb + a * a + epsA);
    T e2 = (abs(2 * c * bb) + epsA) / (c * c + bb * bb + epsA);
    return e2 < e1 ? e2 : e1;
  };
  T emin = eta(q[0], q[1], q[2]);
  T ek = eta(q[1], q[2], q[3]);
  if (ek < emin) emin = ek;
  ek = eta(q[2], q[3], q[4]);
  if (ek < emin) emin = ek;
  const T x = emin / 0.24;
  const T m = 1.0 - (x < 1.0 ? x : T(1.0));
  const T gm = pow(1 - m, 4) * (1 + 4 * m);
  const int betaA = (int)std::floor((double)(10.0 - 5.0 * (1 - gm)));
  return whf_s_cut(q, b, whf_ct_p(std::pow(10, -betaA)));
}

// whf_TcnsN_LAD (interScheme.hpp:1681-1760): CT = 10^-m from a table,
// m = clamp(4 + floor(theta 6), 4, 10), theta = 1 / (1 + max_k chi_k / 10),
// chi_k = tau / (beta_k + 1e-40); cut where gamma_k / sum gamma < CT,
// gamma_k = pow(1 + tau / (beta_k + 1e-40), 6).
template <class T> inline T whf_lad(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  T chi[3];
  for (int i = 0; i < 3; i++) chi[i] = tau / (b[i] + 1e-40);
  T cm = chi[0];  // std::max({..})
  if (cm < chi[1]) cm = chi[1];
  if (cm < chi[2]) cm = chi[2];
  const T theta = 1.0 / (1.0 + cm / 10.0);
  int m = (int)(4 + std::floor((double)(theta * (10 - 4))));
  m = m < 4 ? 4 : (m > 10 ? 10 : m);
  const double table[7] = {1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10};
  const double CT = table[m - 4];
  T g[3];
  for (int i = 0; i < 3; i++) g[i] = pow(1.0 + tau / (b[i] + 1e-40), 6);
  T gs = g[0] + g[1] + g[2];
  unsigned flag = 0;
  if (g[0] / gs < CT) flag += 1;
  if (g[1] / gs < CT) flag += 2;
  if (g[2] / gs < CT) flag += 4;
  return teno_select(flag, q);
}

// ---- the whf_ai_TcnsN family (InterMethod whfAITCNSNS .. whfAITCNSNmyASF002_ai1
// = 24..32) and whf_TcnsN_COMPARE (34) ----------------------------------------
// eps_A = (0.9 C_r) / (1 - 0.9 C_r) xi^2 with C_r = 0.24, xi = 1e-3
constexpr double kEpsA2 = (0.9 * 0.24) / (1.0 - 0.9 * 0.24) * 1e-3 * 1e-3;

// min of the smoothness ratios of adjacent differences (d0,d1), (d1,d2),
// (d2,d3); the reference writes them q_i - q_{i-1}, the same doubles
template <class T> inline T eta_adjacent(const T q[5], double epsA) {
  using std::abs;
  const T d0 = q[0] - q[1], d1 = q[1] - q[2], d2 = q[2] - q[3], d3 = q[3] - q[4];
  const T e0 = (abs(2.0 * d1 * d0) + epsA) / (d1 * d1 + d0 * d0 + epsA);
  const T e1 = (abs(2.0 * d2 * d1) + epsA) / (d2 * d2 + d1 * d1 + epsA);
  const T e2 = (abs(2.0 * d3 * d2) + epsA) / (d3 * d3 + d2 * d2 + epsA);
  T e = e0;
  if (e1 < e) e = e1;
  if (e2 < e) e = e2;
  return e;
}

// m = 1 - min(1, eta / 0.24), beta_A = 10 - 5 (1 - pow(1 - m, 4) (1 + 4 m)),
// C_TA = pow(10, -floor(beta_A)) (interScheme.hpp:2300-2305, :2375-2380)
template <class T> inline double ai_cta(const T& eta) {
  using std::pow;
  const T x = eta / 0.24;
  const T m = 1.0 - (x < 1.0 ? x : T(1.0));
  const T g = pow(1.0 - m, 4) * (1.0 + 4.0 * m);
  const T bA = 10.0 - 5.0 * (1.0 - g);
  return std::pow(10.0, -std::floor(value_of(bA)));
}

// TENO-S's cut-off with a given C_T (the delta_k tests of the whf_ai
// schemes: beta_min tau < (C_T tau - (1 - C_T) beta_min) beta_k cuts k)
template <class T, class C> inline T tenos_cut(const T q[5], const T b[3], const C& CT) {
  using std::abs;
  int mb = 0;
  if (b[1] < b[mb]) mb = 1;
  if (b[2] < b[mb]) mb = 2;
  const C CT_1 = 1 - CT;
  T tau = abs(b[2] - b[0]);
  T rr = CT * tau - CT_1 * b[mb];
  T ll = tau * b[mb];
  unsigned flag = 0;
  if (ll < rr * b[0]) flag += 1;
  if (ll < rr * b[1]) flag += 2;
  if (ll < rr * b[2]) flag += 4;
  return teno_select(flag, q);
}

// whf_ai_TcnsN_S_1 (interScheme.hpp:1901-1995): C_T' = pow(1.5e-5 / (1 - 1e-5), 1/6)
template <class T> inline T ai_s(const T q[5]) {
  T b[3];
  betas(q, b);
  return tenos_cut(q, b, whf_ct_p(1e-5));
}

// whf_ai_TcnsN_A_2 (interScheme.hpp:2216-2328): chi_k = gamma_k / sum gamma,
// gamma_k = pow(1 + tau / (beta_k + 1e-40), 6); cut where chi_k < C_TA
template <class T> inline T ai_a(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  T g[3];
  for (int i = 0; i < 3; i++) g[i] = pow(1.0 + tau / (b[i] + 1e-40), 6.0);
  T gs = g[0] + g[1] + g[2];
  const double CTA = ai_cta(eta_adjacent(q, kEpsA2));
  unsigned flag = 0;
  if (g[0] / gs < CTA) flag += 1;
  if (g[1] / gs < CTA) flag += 2;
  if (g[2] / gs < CTA) flag += 4;
  return teno_select(flag, q);
}

// whf_ai_TcnsN_AS_1 / _AS_2 (interScheme.hpp:2329-2511, the same arithmetic):
// TENO-S with C_T' = pow(1.5 C_TA / (1 - C_TA), 1/6)
template <class T> inline T ai_as(const T q[5]) {
  T b[3];
  betas(q, b);
  return tenos_cut(q, b, whf_ct_p(ai_cta(eta_adjacent(q, kEpsA2))));
}

// whf_ai_TcnsN_LADS_yuanbao (interScheme.hpp:2512-2647): C_T = pow(10, -m),
// m = 5 + floor(5 theta), theta = 1 / (1 + max_k chi_k / 10)
template <class T> inline T ai_lads(const T q[5]) {
  using std::abs;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  T chi[3];
  for (int i = 0; i < 3; i++) chi[i] = tau / (b[i] + 1e-40);
  T cm = chi[0];
  if (cm < chi[1]) cm = chi[1];
  if (cm < chi[2]) cm = chi[2];
  const T theta = 1.0 / (1.0 + cm / 10.0);
  const int m = 5 + (int)std::floor(value_of(theta * (10 - 5)));
  return tenos_cut(q, b, whf_ct_p(std::pow(10.0, -m)));
}

// whf_ai_TcnsN_AZS_yuanbao (interScheme.hpp:2862-2935): C_T = min(1e-5,
// |beta0 + beta2 - 2 beta1|), C_T' = pow(1.5 C_T / (1 - C_T), 1/6)
template <class T> inline T ai_azs(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  const T tA = abs(b[0] + b[2] - 2 * b[1]);
  const T C = tA < 1e-5 ? tA : T(1e-5);
  const T CTp = pow(1.5 * C / (1 - C), 1.0 / 6.0);
  return tenos_cut(q, b, CTp);
}

// whf_ai_TcnsN_myASF002_1 / _2 / _ai1 (interScheme.hpp:2936-3252, the same
// values): C_T' = 0.157 / (11.351092 s^2 - 5.740650 s + 1.506631), s = min(1, eta_min / 0.24)
template <class T> inline T ai_f002(const T q[5]) {
  T b[3];
  betas(q, b);
  const T x = eta_adjacent(q, kEpsA2) / 0.24;
  const T s = x < 1.0 ? x : T(1.0);
  const T CT = 0.157 / (11.351092 * s * s - 5.740650 * s + 1.506631);
  return tenos_cut(q, b, CT);
}

// whf_TcnsN_COMPARE (interScheme.hpp:3472-3726): returns its "my" branch,
// whf_TcnsN_A's cut-off with eps_A = (0.9 C_r)/(1 - 0.9 C_r) xi^2 and
// m = 1 - min(1, eta_min / 0.24); the other branch it evaluates is unused.
template <class T> inline T whf_compare(const T q[5]) {
  using std::abs;
  T b[3];
  betas(q, b);
  T tau = abs(b[2] - b[0]);
  const double epsA = kEpsA2;
  const T d0 = q[0] - q[1], d1 = q[1] - q[2], d2 = q[2] - q[3], d3 = q[3] - q[4];
  const T e0 = (abs(2.0 * d0 * d2) + epsA) / (d0 * d0 + d2 * d2 + epsA);
  const T e1 = (abs(2.0 * d1 * d3) + epsA) / (d1 * d1 + d3 * d3 + epsA);
  const T e2 = (abs(2.0 * d2 * d1) + epsA) / (d2 * d2 + d1 * d1 + epsA);
  T emin = e0;
  if (e1 < emin) emin = e1;
  if (e2 < emin) emin = e2;
  const T x = emin / 0.24;
  const T m = 1.0 - (x < 1.0 ? x : T(1.0));
  const T gm = (1 - m) * (1 - m) * (1 - m) * (1 - m) * (1 + 4 * m);
  const int betaA = (int)std::floor(value_of(10.0 - 5.0 * (1 - gm)));
  const double cts[6] = {1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10};
  const double CTA = cts[betaA < 5 ? 0 : (betaA > 10 ? 5 : betaA - 5)];
  T g[3];
  for (int i = 0; i < 3; i++) {
    T t = 1.0 + tau / (b[i] + 1e-40);
    t = t * t;
    g[i] = t * t * t;
  }
  T rr = CTA * (g[0] + g[1] + g[2]);
  unsigned flag = 0;
  if (g[0] < rr) flag += 1;
  if (g[1] < rr) flag += 2;
  if (g[2] < rr) flag += 4;
  return teno_select(flag, q);
}

// Teno5_CongAA (interScheme.hpp:304-382; the 1D Euler path's scheme for
// WCNS5CONGZCT7, solverType.cpp:248-250): TENO-S's cut-off with
// C_T = pow(10, -round(10 - 5 (1 - pow(1 - m, 4) (1 + 4 m)))), m = 1 - min(1,
// eta_min / 0.24), eta of the values q themselves (as whf_TcnsN_AS).
template <class T> inline T congaa(const T q[5]) {
  using std::abs;
  using std::pow;
  T b[3];
  betas(q, b);
  const double epsA = kEpsA2;
  auto eta = [&](T a, T bb, T c) {
    T e1 = (abs(2 * bb * a) + epsA) / (bb * bb + a * a + epsA);
    T e2 = (abs(2 * c * bb) + epsA) / (c * c + bb * bb + epsA);
    return e2 < e1 ? e2 : e1;
  };
  T emin = eta(q[0], q[1], q[2]);
  T ek = eta(q[1], q[2], q[3]);
  if (ek < emin) emin = ek;
  ek = eta(q[2], q[3], q[4]);
  if (ek < emin) emin = ek;
  const T x = emin / 0.24;
  const T m = 1.0 - (x < 1.0 ? x : T(1.0));
  const T g = pow(1 - m, 4) * (1 + 4 * m);
  const double CT = std::pow(10, -std::round(value_of(10.0 - 5.0 * (1 - g))));
  return tenos_cut(q, b, CT);
}

// musclIn5 (interScheme.hpp:1386-1406): q2 + delta / 2 with the minmod slope
// of q1, q2, q3 (beta = 1).
template <class T> inline T muscl5(const T q[5]) {
  const T dm = q[2] - q[1], dp = q[3] - q[2];
  auto mx = [](T a, T b) { return a < b ? b : a; };  // std::max
  auto mn = [](T a, T b) { return b < a ? b : a; };  // std::min
  T delta;
  if (dp > 0)
    delta = mx(T(0.0), mx(mn(1.0 * dm, dp), mn(dm, 1.0 * dp)));
  else
    delta = mn(T(0.0), mn(mx(1.0 * dm, dp), mx(dm, 1.0 * dp)));
  return q[2] + delta * 0.5;
}

// Linear5 (interScheme.hpp:1450-1453)
template <class T> inline T linear5(const T q[5]) { return teno_select(0u, q); }

template <class T> inline T inter5(int s, const T q[5]) {
  switch (s) {
  case S_JS: return weno5_js(q);
  case S_Z: return weno5_z(q);
  case S_TENO5: return teno5(q, 1e-5);
  case S_TENO5_CT4: return teno5(q, 1e-4);
  case S_TENO5_CT7: return teno5(q, 1e-7);
  case S_TENOS: return tenos(q);
  case S_TENOS_CT4: return tenos(q, 0.23050581003334941);
  case S_TENOS_CT7: return tenos(q, 0.072892337360747711);
  case S_TENO5_CT10: return teno5(q, 1e-10);
  case S_TENOA: return tenoa(q);
  case S_TENOSA_F002: return tenosa(q, 1);
  case S_TENOSA_F203: return tenosa(q, 2);
  case S_TENOSA_F202: return tenosa(q, 3);
  case S_WHF_N: return whf_n(q);
  case S_WHF_A: return whf_a(q);
  case S_WHF_AS: return whf_as(q);
  case S_WHF_S: return whf_s(q);
  case S_WHF_LAD: return whf_lad(q);
  case S_AI_S: return ai_s(q);
  case S_AI_A: return ai_a(q);
  case S_AI_AS: return ai_as(q);
  case S_AI_LADS: return ai_lads(q);
  case S_AI_AZS: return ai_azs(q);
  case S_AI_F002: return ai_f002(q);
  case S_WHF_COMPARE: return whf_compare(q);
  case S_CONGAA: return congaa(q);
  case S_MUSCL5: return muscl5(q);
  case S_C1_F203: return tenosa(q, 4);
  case S_C1_TENOA: return tenoa(q, 4.166667);
  default: return linear5(q);
  }
}

// ---------------------------------------------------------------------------
// Euler point kernels.  Variable order: prim = (rho, u[, v[, w]], p),
// cons = (rho, rho u[, rho v[, rho w]], rho E).  NV = D + 2.
// The face normal is the unit vector of axis `ax`.
// ---------------------------------------------------------------------------

// Equation::consToPrimEuler1D / Euler2D (zone_FD/src/src/equation.cpp:37-99);
// EXTENSION: the 3D form adds w*w after the 2D sum.
template <class T, int D> inline void cons_to_prim(const T* c, T* p) {
  T r = c[0];
  T u = c[1] / r;
  if (D == 1) {
    T E = c[2] / r;
    T e = -u * u / 2 + E;
    T RT = (gam() - 1) * e;
    p[0] = r; p[1] = u; p[2] = r * RT;
  } else if (D == 2) {
    T v = c[2] / r;
    T E = c[3] / r;
    T q2 = (u * u + v * v) / 2;
    T e = -q2 + E;
    T RT = (gam() - 1) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = r * RT;
  } else {
    T v = c[2] / r;
    T w = c[3] / r;
    T E = c[4] / r;
    T q2 = (u * u + v * v + w * w) / 2;
    T e = -q2 + E;
    T RT = (gam() - 1) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = w; p[4] = r * RT;
  }
}

// Conservative state rebuilt from a primitive state exactly as
// eigensystemEuler{1D,2D}::primToChar does before projecting
// (zone_FD/src/src/eigenSystem.cpp:91-93, 254-256).
template <class T, int D> inline void prim_to_cons(const T* p, T* c) {
  T r = p[0];
  if (D == 1) {
    T ek = (p[1] * p[1]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = p[2] / (gam() - 1) + r * ek;
  } else if (D == 2) {
    T ek = (p[1] * p[1] + p[2] * p[2]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = p[3] / (gam() - 1) + r * ek;
  } else {
    T ek = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = r * p[3];
    c[4] = p[4] / (gam() - 1) + r * ek;
  }
}

// Roe-averaged characteristic system at one interface.
// 1D: eigensystemEuler1D(priml, primr) (eigenSystem.cpp:217-250).
// 2D: eigensystemEuler2D(priml, primr, norm) (eigenSystem.cpp:13-86).
// 3D (EXTENSION): waves (shear1, entropy, Vn-c, Vn+c, shear2); for the x and y
// normals shear1 is the 2D shear wave and every sum keeps the 2D order with
// the w term inserted so that w = 0 reproduces 2D bit for bit.
template <class T, int D> struct Eigen {
  static constexpr int NV = D + 2;
  T L[NV][NV];  // rows: waves; columns: cons components
  T R[NV][NV];  // rows: cons components; columns: waves

  void build(const T* pl, const T* pr, int ax) {
    using std::sqrt;
    if (D == 1) {
      T rl = pl[0], ul = pl[1], ppl = pl[2];
      T rr = pr[0], ur = pr[1], ppr = pr[2];
      T HL = (ul * ul) / 2 + ppl / rl * gam() / (gam() - 1);
      T HR = (ur * ur) / 2 + ppr / rr * gam() / (gam() - 1);
      T c1 = sqrt(rl), c2 = sqrt(rr);
      T divisor = 1.0 / (sqrt(rl) + sqrt(rr));
      T u = (c1 * ul + c2 * ur) * divisor;
      T ht = (c1 * HL + c2 * HR) * divisor;
      T ek = (u * u) / 2;
      T h = ht - ek;
      T c = sqrt((gam() - 1) * h);
      const double g1 = gam() - 1;
      T l[9] = {ht + c * (u - c) / g1, -u - c / g1, T(1.0),
                -2.0 * ht + 4.0 * c * c / g1, 2.0 * u, T(-2.0),
                ht - c * (u + c) / g1, -u + c / g1, T(1.0)};
      T f = 0.5 * g1 / (c * c);
      for (int i = 0; i < 9; i++) L[i / 3][i % 3] = l[i] * f;
      T r9[9] = {T(1.0), T(1.0), T(1.0), u - c, u, u + c, ht - u * c, 0.5 * u * u, ht + u * c};
      for (int i = 0; i < 9; i++) R[i / 3][i % 3] = r9[i];
      return;
    }
    // 2D / 3D Roe average (eigenSystem.cpp:21-51)
    T rl = pl[0], rr = pr[0];
    T ul = pl[1], ur = pr[1], vl = pl[2], vr = pr[2];
    T wl = (D == 3) ? pl[3] : T(0.0), wr = (D == 3) ? pr[3] : T(0.0);
    T ppl = pl[NV - 1], ppr = pr[NV - 1];
    T HL, HR;
    if (D == 2) {
      HL = (ul * ul + vl * vl) / 2 + ppl / rl * gam() / (gam() - 1);
      HR = (ur * ur + vr * vr) / 2 + ppr / rr * gam() / (gam() - 1);
    } else {
      HL = (ul * ul + vl * vl + wl * wl) / 2 + ppl / rl * gam() / (gam() - 1);
      HR = (ur * ur + vr * vr + wr * wr) / 2 + ppr / rr * gam() / (gam() - 1);
    }
    T c1 = sqrt(rl), c2 = sqrt(rr);
    T divisor = 1.0 / (c1 + c2);
    T u = (c1 * ul + c2 * ur) * divisor;
    T v = (c1 * vl + c2 * vr) * divisor;
    T w = (D == 3) ? (c1 * wl + c2 * wr) * divisor : T(0.0);
    T ht = (c1 * HL + c2 * HR) * divisor;
    const double nx = ax == 0 ? 1.0 : 0.0, ny = ax == 1 ? 1.0 : 0.0, nz = ax == 2 ? 1.0 : 0.0;
    T Vn, ek;
    if (D == 2) {
      Vn = nx * u + ny * v;
      ek = (u * u + v * v) / 2;
    } else {
      Vn = nx * u + ny * v + nz * w;
      ek = (u * u + v * v + w * w) / 2;
    }
    T h = ht - ek;
    T c = sqrt((gam() - 1) * h);
    if (D == 2) {
      // leftEig / rightEig (eigenSystem.cpp:53-85)
      T l[16] = {(-ny * u + nx * v), T(ny), T(-nx), T(0.0),
                 (h - ek), u, v, T(-1.0),
                 (Vn / c + ek / h) / 2, (-nx / c - u / h) / 2, (-ny / c - v / h) / 2, 1.0 / (2.0 * h),
                 (-Vn / c + ek / h) / 2, (nx / c - u / h) / 2, (ny / c - v / h) / 2, 1.0 / (2.0 * h)};
      T r[16] = {T(0.0), 1.0 / h, T(1.0), T(1.0),
                 T(ny), u / h, u - nx * c, u + nx * c,
                 T(-nx), v / h, v - ny * c, v + ny * c,
                 ny * u - nx * v, ek / h, h + ek - Vn * c, h + ek + Vn * c};
      for (int i = 0; i < 16; i++) { L[i / 4][i % 4] = l[i]; R[i / 4][i % 4] = r[i]; }
      return;
    }
    // 3D: tangents t1 (2D shear direction for x/y normals), t2.
    // x: t1=(0,-1,0) t2=(0,0,1); y: t1=(1,0,0) t2=(0,0,1); z: t1=(1,0,0) t2=(0,1,0)
    double t1[3], t2[3];
    if (ax == 0) { t1[0] = 0; t1[1] = -1; t1[2] = 0; t2[0] = 0; t2[1] = 0; t2[2] = 1; }
    else if (ax == 1) { t1[0] = 1; t1[1] = 0; t1[2] = 0; t2[0] = 0; t2[1] = 0; t2[2] = 1; }
    else { t1[0] = 1; t1[1] = 0; t1[2] = 0; t2[0] = 0; t2[1] = 1; t2[2] = 0; }
    // shear-1 row/column: for x/y normals the literal 2D expressions
    T l0rho, r0E;
    if (ax < 2) { l0rho = (-ny * u + nx * v); r0E = ny * u - nx * v; }
    else { l0rho = -u; r0E = u; }
    T t2u = (ax == 2) ? v : w;  // t2 . velocity
    // rows of L: (rho, rho u, rho v, rho w, rho E)
    L[0][0] = l0rho; L[0][1] = T(ax < 2 ? ny : t1[0]); L[0][2] = T(ax < 2 ? -nx : t1[1]); L[0][3] = T(0.0); L[0][4] = T(0.0);
    L[1][0] = h - ek; L[1][1] = u; L[1][2] = v; L[1][3] = w; L[1][4] = T(-1.0);
    L[2][0] = (Vn / c + ek / h) / 2; L[2][1] = (-nx / c - u / h) / 2; L[2][2] = (-ny / c - v / h) / 2;
    L[2][3] = (-nz / c - w / h) / 2; L[2][4] = 1.0 / (2.0 * h);
    L[3][0] = (-Vn / c + ek / h) / 2; L[3][1] = (nx / c - u / h) / 2; L[3][2] = (ny / c - v / h) / 2;
    L[3][3] = (nz / c - w / h) / 2; L[3][4] = 1.0 / (2.0 * h);
    L[4][0] = -t2u; L[4][1] = T(t2[0]); L[4][2] = T(t2[1]); L[4][3] = T(t2[2]); L[4][4] = T(0.0);
    // columns of R (stored R[component][wave])
    R[0][0] = T(0.0); R[0][1] = 1.0 / h; R[0][2] = T(1.0); R[0][3] = T(1.0); R[0][4] = T(0.0);
    R[1][0] = T(ax < 2 ? ny : t1[0]); R[1][1] = u / h; R[1][2] = u - nx * c; R[1][3] = u + nx * c; R[1][4] = T(t2[0]);
    R[2][0] = T(ax < 2 ? -nx : t1[1]); R[2][1] = v / h; R[2][2] = v - ny * c; R[2][3] = v + ny * c; R[2][4] = T(t2[1]);
    R[3][0] = T(t1[2]); R[3][1] = w / h; R[3][2] = w - nz * c; R[3][3] = w + nz * c; R[3][4] = T(t2[2]);
    R[4][0] = r0E; R[4][1] = ek / h; R[4][2] = h + ek - Vn * c; R[4][3] = h + ek + Vn * c; R[4][4] = t2u;
  }

  // primToChar applied to a conservative state built by prim_to_cons
  // (eigenSystem.cpp:97-107, 260-264).
  void to_char(const T* c, T* ch) const {
    for (int k = 0; k < NV; k++) {
      if (D == 1) ch[k] = c[0] * L[k][0] + c[1] * L[k][1] + c[2] * L[k][2];
      else if (D == 2) ch[k] = c[1] * L[k][1] + c[2] * L[k][2] + c[3] * L[k][3] + c[0] * L[k][0];
      else ch[k] = c[1] * L[k][1] + c[2] * L[k][2] + c[3] * L[k][3] + c[4] * L[k][4] + c[0] * L[k][0];
    }
  }
  // charToCons (eigenSystem.cpp:139-149 summation order ch3,ch4,ch2,ch1; 1D :300-302)
  void to_cons(const T* ch, T* c) const {
    for (int i = 0; i < NV; i++) {
      if (D == 1) c[i] = ch[0] * R[i][0] + ch[1] * R[i][1] + ch[2] * R[i][2];
      else if (D == 2) c[i] = ch[2] * R[i][2] + ch[3] * R[i][3] + ch[1] * R[i][1] + ch[0] * R[i][0];
      else c[i] = ch[2] * R[i][2] + ch[3] * R[i][3] + ch[1] * R[i][1] + ch[0] * R[i][0] + ch[4] * R[i][4];
    }
  }
};

// eigensystemEuler2D::charToPrim tail (eigenSystem.cpp:163-167);
// 1D eigensystemEuler1D::consToPrim (:309-322).
template <class T, int D> inline void eig_cons_to_prim(const T* c, T* p) {
  T rt = c[0];
  if (D == 1) {
    T ut = c[1] / rt;
    T ekt = (ut * ut) / 2;
    p[0] = rt; p[1] = ut; p[2] = (gam() - 1) * (c[2] - rt * ekt);
  } else if (D == 2) {
    T ut = c[1] / rt, vt = c[2] / rt;
    T rekt = (c[1] * c[1] + c[2] * c[2]) / rt / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = (gam() - 1) * (c[3] - rekt);
  } else {
    T ut = c[1] / rt, vt = c[2] / rt, wt = c[3] / rt;
    T rekt = (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) / rt / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = wt; p[4] = (gam() - 1) * (c[4] - rekt);
  }
}

// variable_positive_limiter (zone_FD/src/include/positive_preserving.hpp:14-46)
// with the density (q=rho) and rho*c^2 (q=gamma p) functionals used by
// Recon5Order1DEulerEig::reconI (reconstructor5order.hpp:204-232).
template <class T> inline T lim_rho(const T* W) { return W[0]; }
template <class T> inline T lim_rhoc2(const T* W) {
  T u = W[1] / W[0];
  T ekin = 0.5 * u * u;
  T p = (gam() - 1) * (W[2] - W[0] * ekin);
  return gam() * p;
}
template <class T> inline void positive_limit(const T* W, T* Wi, int which) {
  T qi = which == 0 ? lim_rho(Wi) : lim_rhoc2(Wi);
  if (qi >= 1e-10) return;
  T qw = which == 0 ? lim_rho(W) : lim_rhoc2(W);
  if (qw < 1e-10) {
    for (int i = 0; i < 3; i++) Wi[i] = W[i];
    return;
  }
  T theta = (1e-10 - qw) / (qi - qw);
  for (int i = 0; i < 3; i++) Wi[i] = (1 - theta) * W[i] + theta * Wi[i];
}

// Half-node reconstruction at one interface from the 6-cell window
// w[0..5] = cells i-3 .. i+2 (interface between w[2] and w[3]).
// Component-wise: Recon5OrderFaceCenter::reconI (reconstructor5order.hpp:128-151).
// Characteristic: Recon5Order{1D,2D}EulerEig::reconI (:154-244, :247-356).
// variant 1, CongFD1.0's Euler path (SpaceDis::recon1DFaceCenter /
// recon2DFaceCenter, CongFD1.0/src/src/sp_recon.cpp:162-267, with the checks
// of calFluxEuler1D / 2D, sp_flux.cpp:123-252): the same characteristic
// reconstruction with no 1D limiter, and a state with negative density or
// pressure replaced by its own cell (cell i-1 for the left state, cell i for
// the right one) in 1D and 2D alike.
template <class T, int D>
inline void recon_interface(int scheme, bool charwise, const T* const* wp, const T* const* wc, int ax,
                            T* outL, T* outR, int variant = 0) {
  constexpr int NV = D + 2;
  if (!charwise) {
    for (int i = 0; i < NV; i++) {
      T qL[5] = {wp[0][i], wp[1][i], wp[2][i], wp[3][i], wp[4][i]};
      T qR[5] = {wp[5][i], wp[4][i], wp[3][i], wp[2][i], wp[1][i]};
      outL[i] = inter5(scheme, qL);
      outR[i] = inter5(scheme, qR);
    }
    return;
  }
  const T* pL = wp[2];
  const T* pR = wp[3];
  Eigen<T, D> eig;
  eig.build(pL, pR, ax);
  T ch[6][NV];
  for (int j = 0; j < 6; j++) eig.to_char(wc[j], ch[j]);
  T QL[NV], QR[NV];
  for (int k = 0; k < NV; k++) {
    T qL[5] = {ch[0][k], ch[1][k], ch[2][k], ch[3][k], ch[4][k]};
    T qR[5] = {ch[5][k], ch[4][k], ch[3][k], ch[2][k], ch[1][k]};
    QL[k] = inter5(scheme, qL);
    QR[k] = inter5(scheme, qR);
  }
  T cL[NV], cR[NV];
  eig.to_cons(QL, cL);
  eig.to_cons(QR, cR);
  if (variant == 1) {
    eig_cons_to_prim<T, D>(cL, outL);
    eig_cons_to_prim<T, D>(cR, outR);
    if (outL[0] < 0 || outL[NV - 1] < 0)
      for (int i = 0; i < NV; i++) outL[i] = pL[i];
    if (outR[0] < 0 || outR[NV - 1] < 0)
      for (int i = 0; i < NV; i++) outR[i] = pR[i];
    return;
  }
  if (D == 1) {
    T consL[3], consR[3];
    prim_to_cons<T, 1>(pL, consL);
    prim_to_cons<T, 1>(pR, consR);
    positive_limit(consL, cL, 0);
    positive_limit(consR, cR, 0);
    positive_limit(consL, cL, 1);
    positive_limit(consR, cR, 1);
    eig_cons_to_prim<T, 1>(cL, outL);
    eig_cons_to_prim<T, 1>(cR, outR);
    return;
  }
  eig_cons_to_prim<T, D>(cL, outL);
  eig_cons_to_prim<T, D>(cR, outR);
  // simple positivity fallback (reconstructor5order.hpp:301-302); the right
  // state keeps the reference's use of the left cell's velocities.
  if (outL[0] < 0 || outL[NV - 1] < 0)
    for (int i = 0; i < NV; i++) outL[i] = pL[i];
  if (outR[0] < 0 || outR[NV - 1] < 0) {
    outR[0] = pR[0];
    for (int i = 1; i < NV - 1; i++) outR[i] = pL[i];
    outR[NV - 1] = pR[NV - 1];
  }
}

// HLLCFlux1D (zone_FD/src/include/fluxSchemes.hpp:74-125)
template <class T> inline void hllc1d(const T* L, const T* Rr, T* res) {
  using std::sqrt;
  T rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  T HL = enth_macro(pl, rl) + (ul * ul) / 2;
  T HR = enth_macro(pr, rr) + (ur * ur) / 2;
  T cl = sqrt(gam() * (pl / rl));
  T cr = sqrt(gam() * (pr / rr));
  T SL = smin(ur - cr, ul - cl);
  T SR = smax(ul + cl, ur + cr);
  T Sstar = ((pr - pl) + (rl * ul * (SL - ul) - rr * ur * (SR - ur))) / (rl * (SL - ul) - rr * (SR - ur));
  if (SL >= 0) {
    res[0] = rl * ul; res[1] = rl * ul * ul + pl; res[2] = rl * HL * ul;
  } else if (Sstar >= 0) {
    T pStar = pl + rl * (SL - ul) * (Sstar - ul);
    T U[3] = {rl, rl * ul, pl / (gam() - 1) + rl * ul * ul / 2};
    T F[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
    T Dd[3] = {T(0.0), T(1.0), Sstar};
    for (int i = 0; i < 3; i++) res[i] = (Sstar * (SL * U[i] - F[i]) + SL * pStar * Dd[i]) / (SL - Sstar);
  } else if (SR >= 0) {
    T pStar = pr + rr * (SR - ur) * (Sstar - ur);
    T U[3] = {rr, rr * ur, pr / (gam() - 1) + rr * ur * ur / 2};
    T F[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
    T Dd[3] = {T(0.0), T(1.0), Sstar};
    for (int i = 0; i < 3; i++) res[i] = (Sstar * (SR * U[i] - F[i]) + SR * pStar * Dd[i]) / (SR - Sstar);
  } else {
    res[0] = rr * ur; res[1] = rr * ur * ur + pr; res[2] = rr * HR * ur;
  }
}

// roeFlux1D2 (fluxSchemes.hpp:13-71)
template <class T> inline void roe1d(const T* L, const T* Rr, T* res) {
  using std::abs;
  using std::sqrt;
  T rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  T HL = enth_macro(pl, rl) + (ul * ul) / 2;
  T HR = enth_macro(pr, rr) + (ur * ur) / 2;
  T rBar = sqrt(rl * rr);
  T uBar = (ul * sqrt(rl) + ur * sqrt(rr)) / (sqrt(rl) + sqrt(rr));
  T HBar = (HL * sqrt(rl) + HR * sqrt(rr)) / (sqrt(rl) + sqrt(rr));
  T cBar = sqrt((gam() - 1) * (HBar - uBar * uBar / 2));
  T cBar2 = (gam() - 1) * (HBar - uBar * uBar / 2);
  T dr = rr - rl, du = ur - ul, dp = pr - pl;
  T lam[3] = {abs(uBar), abs(uBar + cBar), abs(uBar - cBar)};
  T eps = 0.1 * (abs(uBar) + cBar);
  for (int i = 0; i < 3; i++)
    if (lam[i] < eps) lam[i] = (lam[i] * lam[i] + eps * eps) / (2.0 * eps);
  T a0 = lam[0] * (dr - dp / cBar2);
  T a1 = lam[1] / (2.0 * cBar2) * (dp + rBar * cBar * du);
  T a2 = lam[2] / (2.0 * cBar2) * (dp - rBar * cBar * du);
  T a3 = a0 + a1 + a2;
  T a4 = cBar * (a1 - a2);
  T FL[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
  T FR[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
  for (int i = 0; i < 3; i++) res[i] = (FL[i] + FR[i]) / 2.0;
  res[0] -= 0.5 * a3;
  res[1] -= 0.5 * (a3 * uBar + a4);
  res[2] -= 0.5 * (HBar * a3 + uBar * a4 - cBar2 * a0 / (gam() - 1));
}

// HLLCFlux2D2 (fluxSchemes.hpp:202-327); EXTENSION to 3D with the tangential
// momentum of the star state written as Sstar*n_k + u_k*(sum of the other n's).
template <class T, int D> inline void hllc_md(const T* Lp, const T* Rp, int ax, T* res) {
  using std::sqrt;
  constexpr int NV = D + 2;
  constexpr int E = NV - 1;
  double n[3] = {ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
  T rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  T vel_l[3], vel_r[3];
  for (int k = 0; k < D; k++) { vel_l[k] = Lp[1 + k]; vel_r[k] = Rp[1 + k]; }
  T HL, HR;
  if (D == 2) {
    HL = (vel_l[0] * vel_l[0] + vel_l[1] * vel_l[1]) / 2 + enth_macro(pl, rl);
    HR = (vel_r[0] * vel_r[0] + vel_r[1] * vel_r[1]) / 2 + enth_macro(pr, rr);
  } else {
    HL = (vel_l[0] * vel_l[0] + vel_l[1] * vel_l[1] + vel_l[2] * vel_l[2]) / 2 + enth_macro(pl, rl);
    HR = (vel_r[0] * vel_r[0] + vel_r[1] * vel_r[1] + vel_r[2] * vel_r[2]) / 2 + enth_macro(pr, rr);
  }
  T VnL = vel_l[ax], VnR = vel_r[ax];
  T cl = sqrt(pl / rl * gam());
  T cr = sqrt(pr / rr * gam());
  T rBar = (rl + rr) / 2, cBar = (cl + cr) / 2;
  T ps = smax(T(0.0), (pl + pr) / 2 - rBar * cBar * (VnR - VnL) / 2);
  T ql = (ps <= pl) ? T(1.0) : sqrt(1.0 + (gam() + 1.0) / (2.0 * gam()) * (ps / pl - 1.0));
  T qr = (ps <= pr) ? T(1.0) : sqrt(1.0 + (gam() + 1.0) / (2.0 * gam()) * (ps / pr - 1.0));
  T SL = VnL - cl * ql, SR = VnR + cr * qr;
  T Sstar = ((pr - pl) + (rl * VnL * (SL - VnL) - rr * VnR * (SR - VnR))) / (rl * (SL - VnL) - rr * (SR - VnR));
  auto side_flux = [&](T r, const T* vel, T p, T H, T Vn, T* F) {
    F[0] = r * Vn;
    for (int k = 0; k < D; k++) F[1 + k] = r * vel[k] * Vn + p * n[k];
    F[E] = r * H * Vn;
  };
  // star-region flux F + S*(U* - U) of one side
  auto star_flux = [&](T r, const T* vel, T p, T H, T Vn, T S, T* out) {
    T U[NV], F[NV], Us[NV];
    U[0] = r;
    for (int k = 0; k < D; k++) U[1 + k] = r * vel[k];
    U[E] = r * H - p;
    side_flux(r, vel, p, H, Vn, F);
    T Usf = (S - Vn) / (S - Sstar);
    Us[0] = Usf * r;
    for (int k = 0; k < D; k++) {
      double other = 0.0;  // sum of the other normal components (the 2D n[1]/n[0] swap)
      for (int m = 0; m < D; m++)
        if (m != k) other += n[m];
      Us[1 + k] = Usf * (Sstar * n[k] + vel[k] * other) * r;
    }
    Us[E] = Usf * (U[E] + (Sstar - Vn) * (r * Sstar + p / (S - Vn)));
    for (int i = 0; i < NV; i++) out[i] = F[i] + S * (Us[i] - U[i]);
  };
  if (SL > 0) {
    side_flux(rl, vel_l, pl, HL, VnL, res);
  } else if (Sstar > 0) {
    star_flux(rl, vel_l, pl, HL, VnL, SL, res);
  } else if (Sstar == 0) {
    T a[NV], b[NV];
    star_flux(rl, vel_l, pl, HL, VnL, SL, a);
    star_flux(rr, vel_r, pr, HR, VnR, SR, b);
    for (int i = 0; i < NV; i++) res[i] = a[i] / 2;
    for (int i = 0; i < NV; i++) res[i] += b[i] / 2;
  } else if (SR > 0) {
    star_flux(rr, vel_r, pr, HR, VnR, SR, res);
  } else {
    side_flux(rr, vel_r, pr, HR, VnR, res);
  }
}

// roeFlux2DSym (fluxSchemes.hpp:420-553); EXTENSION to 3D: every tangential
// velocity contributes to the shear part in the order (u, v, w).
template <class T, int D> inline void roe_md(const T* Lp, const T* Rp, int ax, T* res) {
  using std::abs;
  using std::sqrt;
  constexpr int NV = D + 2;
  constexpr int E = NV - 1;
  double n[3] = {ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
  T rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  T vl[3], vr[3];
  for (int k = 0; k < D; k++) { vl[k] = Lp[1 + k]; vr[k] = Rp[1 + k]; }
  T HL, HR;
  if (D == 2) {
    HL = (vl[0] * vl[0] + vl[1] * vl[1]) / 2 + enth_macro(pl, rl);
    HR = (vr[0] * vr[0] + vr[1] * vr[1]) / 2 + enth_macro(pr, rr);
  } else {
    HL = (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]) / 2 + enth_macro(pl, rl);
    HR = (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]) / 2 + enth_macro(pr, rr);
  }
  T VnL = vl[ax], VnR = vr[ax];
  T FcL[NV], FcR[NV];
  FcL[0] = rl * VnL;
  FcR[0] = rr * VnR;
  for (int k = 0; k < D; k++) {
    FcL[1 + k] = rl * vl[k] * VnL + pl * n[k];
    FcR[1 + k] = rr * vr[k] * VnR + pr * n[k];
  }
  FcL[E] = rl * HL * VnL;
  FcR[E] = rr * HR * VnR;
  T c1 = sqrt(rl), c2 = sqrt(rr);
  T divisor = 1.0 / (sqrt(rl) + sqrt(rr));
  T rhoAvg = sqrt(rl * rr);
  T va[3];
  for (int k = 0; k < D; k++) va[k] = (c1 * vl[k] + c2 * vr[k]) * divisor;
  T HAvg = (c1 * HL + c2 * HR) * divisor;
  T q2;
  if (D == 2) q2 = (va[0] * va[0] + va[1] * va[1]) / 2;
  else q2 = (va[0] * va[0] + va[1] * va[1] + va[2] * va[2]) / 2;
  T cAvg = sqrt((gam() - 1) * (HAvg - q2));
  T VnAvg = va[ax];
  T lam[3] = {abs(VnAvg - cAvg), abs(VnAvg), abs(VnAvg + cAvg)};
  T eps = 0.05 * (abs(VnAvg) + cAvg);
  for (int i = 0; i < 3; i++)
    if (lam[i] < eps) lam[i] = (lam[i] * lam[i] + eps * eps) / (2.0 * eps);
  T dP = pr - pl, dVn = VnR - VnL, dRho = rr - rl;
  T dv[3];
  for (int k = 0; k < D; k++) dv[k] = vr[k] - vl[k];
  T FD[NV];
  T cf1 = (dP - rhoAvg * cAvg * dVn) / (2.0 * cAvg * cAvg);
  FD[0] = lam[0] * (cf1 * 1);
  for (int k = 0; k < D; k++) FD[1 + k] = lam[0] * (cf1 * (va[k] - cAvg * n[k]));
  FD[E] = lam[0] * (cf1 * (HAvg - cAvg * VnAvg));
  T cf = (dP + rhoAvg * cAvg * dVn) / (2.0 * cAvg * cAvg);
  FD[0] += lam[2] * (cf * 1);
  for (int k = 0; k < D; k++) FD[1 + k] += lam[2] * (cf * (va[k] + cAvg * n[k]));
  FD[E] += lam[2] * (cf * (HAvg + cAvg * VnAvg));
  T cf2 = dRho - dP / (cAvg * cAvg);
  FD[0] += lam[1] * (cf2 * 1.0 + rhoAvg * 0.0);
  for (int k = 0; k < D; k++) FD[1 + k] += lam[1] * (cf2 * va[k] + rhoAvg * (k == ax ? T(0.0) : dv[k]));
  // shear energy: sum over tangential components in (u, v, w) order
  T shear = T(0.0);
  bool first = true;
  for (int k = 0; k < D; k++) {
    if (k == ax) continue;
    if (first) { shear = va[k] * dv[k]; first = false; }
    else shear = shear + va[k] * dv[k];
  }
  FD[E] += lam[1] * (cf2 * q2 + rhoAvg * shear);
  for (int i = 0; i < NV; i++) res[i] = (FcL[i] + FcR[i] - FD[i]) / 2;
}

// Physical flux at a cell: fluxSolveEuler1D / 2D (fluxSchemes.hpp:572-599).
template <class T, int D> inline void node_flux(const T* p, int ax, T* F) {
  if (D == 1) {
    T r = p[0], u = p[1], pp = p[2];
    F[0] = r * u;
    F[1] = r * u * u + pp;
    F[2] = (gam() / (gam() - 1) * pp + r * (u * u) / 2) * u;
    return;
  }
  constexpr int NV = D + 2;
  constexpr int E = NV - 1;
  double n[3] = {ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
  T r = p[0], pp = p[E];
  T Vn, k2;
  if (D == 2) {
    Vn = (p[1] * n[0] + p[2] * n[1]);
    k2 = (p[1] * p[1] + p[2] * p[2]);
  } else {
    Vn = (p[1] * n[0] + p[2] * n[1] + p[3] * n[2]);
    k2 = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  }
  F[0] = r * Vn;
  for (int k = 0; k < D; k++) F[1 + k] = r * p[1 + k] * Vn + n[k] * pp;
  F[E] = (k2 / 2 * r + gam() / (gam() - 1) * pp) * Vn;
}

template <class T, int D> inline void riemann(int flux, const T* L, const T* R, int ax, T* res) {
  if (D == 1) {
    if (flux == F_HLLC) hllc1d(L, R, res);
    else roe1d(L, R, res);
  } else {
    if (flux == F_HLLC) hllc_md<T, D>(L, R, ax, res);
    else roe_md<T, D>(L, R, ax, res);
  }
}

// ---------------------------------------------------------------------------
// Boundary ("halo") fill per pencil end: OneDBnd::update
// (zone_FD/src/src/oneDBnd.cpp:31-114) with the wiring of
// Initializer::initBnds (initializer.cpp:743-952).  Ghost k = 0 is adjacent to
// the domain.  Codes follow the reference BndType (macro.hpp:33-34) plus
// EXTENSION codes for a true periodic Euler halo and a z symmetry plane.
// ---------------------------------------------------------------------------
enum Bc {
  BC_PERIODIC = 1,       // true periodic (LINEARCONV1D wiring, initializer.cpp:769-773)
  BC_SYMMETRY1D = 2,     // mirror, flip u
  BC_DIRICHLET = 3,      // fixed primitive state
  BC_OUTFLOW = 7,        // SUPERSONICOUTLET: copy the edge cell
  BC_SYMMETRYX = 10,     // mirror, flip u
  BC_SYMMETRYY = 11,     // mirror, flip v
  BC_SYMMETRYZ = 101,    // EXTENSION: mirror, flip w
  BC_MIRROR_COPY = 102,  // the reference's 2D Euler PERIODIC1D wiring: mirror without flip
  BC_DMR_TOP = 8,        // DoubleMachUp: exact oblique-shock state from x and t (oneDBnd.cpp:87-108)
  BC_DMR_WALL = 103,     // initDoubleMachBnds y-low wiring (initializer.cpp:916-934): Dirichlet
                         // for cell centres x < 1/6, SYMMETRYY beyond
};

// Where a ghost layer sits, for the position- and time-dependent codes: the
// pencil's cell-centre x, the cell-centre spacing across the face and the
// current (stage) time.
struct GhostCtx {
  double x = 0.0, dh = 0.0, t = 0.0;
};

template <class T, int D>
inline void fill_ghost(int bc, const double* dir_prim, const T* prim_pencil, int n, int stride_v, bool left,
                       T* ghost /* 4 x NV */, const GhostCtx& ctx = GhostCtx()) {
  constexpr int NV = D + 2;
  if (bc == BC_DMR_WALL) bc = ctx.x < 1.0 / 6.0 ? BC_DIRICHLET : BC_SYMMETRYY;
  for (int k = 0; k < 4; k++) {
    if (bc == BC_DMR_TOP) {
      // oneDBnd.cpp:90-105: ghost k at y = -dh/2 + k dh from the face
      const double y = -ctx.dh / 2 + k * ctx.dh;
      const double gt = 1.0 / 6.0 + std::sqrt(3.0) / 3.0 * (1.0 + 20 * ctx.t);
      const bool post = y > std::sqrt(3.0) * (ctx.x - gt);
      // pre-shock state: rho 1.4 at rest, p 1
      for (int v = 0; v < NV; v++)
        ghost[k * NV + v] = T(post ? dir_prim[v] : v == 0 ? 1.4 : v == NV - 1 ? 1.0 : 0.0);
      continue;
    }
    int src;
    int flip = -1;
    switch (bc) {
    case BC_PERIODIC: src = left ? ((n - 1 - k) % n + n) % n : k % n; break;
    case BC_OUTFLOW: src = left ? 0 : n - 1; break;
    case BC_SYMMETRY1D:
    case BC_SYMMETRYX: src = left ? k : n - 1 - k; flip = 1; break;
    case BC_SYMMETRYY: src = left ? k : n - 1 - k; flip = 2; break;
    case BC_SYMMETRYZ: src = left ? k : n - 1 - k; flip = 3; break;
    case BC_MIRROR_COPY: src = left ? k : n - 1 - k; break;
    default: src = -1; break;
    }
    for (int v = 0; v < NV; v++) {
      T val;
      if (bc == BC_DIRICHLET) val = T(dir_prim[v]);
      else {
        val = prim_pencil[(long)src * stride_v + v];
        if (v == flip) val = -val;
      }
      ghost[k * NV + v] = val;
    }
  }
}

struct Grid {
  int dim;
  int n[3];
  double h[3];
  int scheme;
  int flux;
  int charwise;
  int bc[6];
  double bc_prim[6][5];
  // EXTENSION fields used by the position/time-dependent boundaries and the
  // source term: cell centres are x0 + (i + 0.5) h[0] in every direction
  // (Initializer::initUniformBlock, initializer.cpp:672-677), t is the
  // current time, gravity the acceleration of SourceTerm::calGravitySource.
  double x0 = 0.0;
  double t = 0.0;
  double gravity[3] = {0.0, 0.0, 0.0};
  // 0: zone_FD's path; 1: CongFD1.0's Euler path (recon_interface's variant,
  // and SpaceDis::difMND6's three divisions by h, CongFD1.0/src/src/sp_difference.cpp:130-132)
  int variant = 0;
};

// One direction sweep: SpDistributor::rhsSolve body for direction `ax`
// (zone_FD/src/src/sp_distributor.cpp:9-78): per pencil, reconstruct n+3
// half nodes from 4+n+4 cells, Riemann flux at each, node flux at cells
// -1..n, then MidNodeAndNodeDif6 (zone_FD/src/src/differ.cpp:54-103):
// rhs_i += (w1 (f_{i+1}-f_{i-1}) + w0 (F_{i+1/2}-F_{i-1/2}) + w2 (F_{i+3/2}-F_{i-3/2})) / h.
template <class T, int D>
void sweep(const Grid& g, int ax, const T* prim, T* rhs, int nthreads) {
  constexpr int NV = D + 2;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long sx = 1, sy = nx, sz = (long)nx * ny;
  const int n = g.n[ax];
  const long sa = ax == 0 ? sx : ax == 1 ? sy : sz;
  int m1, m2;
  long s1, s2;
  if (ax == 0) { m1 = ny; m2 = nz; s1 = sy; s2 = sz; }
  else if (ax == 1) { m1 = nx; m2 = nz; s1 = sx; s2 = sz; }
  else { m1 = nx; m2 = ny; s1 = sx; s2 = sy; }
  const double hh = g.h[ax];
  const double w0 = 3.0 / 2.0, w1 = -3.0 / 10.0, w2 = 1.0 / 30.0;
  const long npencil = (long)m1 * m2;
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<T> pp((n + 8) * NV), cc((n + 8) * NV), fh((n + 3) * NV), fn((n + 2) * NV);
#pragma omp for schedule(static)
    for (long ip = 0; ip < npencil; ip++) {
      long i1 = ip % m1, i2 = ip / m1;
      long base = i1 * s1 + i2 * s2;
      // gather pencil with ghosts: pp[(j+4)*NV + v] for cell j in [-4, n+4)
      for (int j = 0; j < n; j++)
        for (int v = 0; v < NV; v++) pp[(j + 4) * NV + v] = prim[(base + j * sa) * NV + v];
      T gl[4 * NV], gr[4 * NV];
      GhostCtx ctx;
      ctx.x = g.x0 + (i1 + 0.5) * g.h[0];  // the pencil's x (y sweeps of the 2D cases)
      ctx.dh = (g.x0 + ((n - 1) + 0.5) * g.h[0]) - (g.x0 + ((n - 2) + 0.5) * g.h[0]);
      ctx.t = g.t;
      fill_ghost<T, D>(g.bc[2 * ax], g.bc_prim[2 * ax], &pp[4 * NV], n, NV, true, gl, ctx);
      fill_ghost<T, D>(g.bc[2 * ax + 1], g.bc_prim[2 * ax + 1], &pp[4 * NV], n, NV, false, gr, ctx);
      for (int k = 0; k < 4; k++)
        for (int v = 0; v < NV; v++) {
          pp[(3 - k) * NV + v] = gl[k * NV + v];
          pp[(n + 4 + k) * NV + v] = gr[k * NV + v];
        }
      for (int j = 0; j < n + 8; j++) prim_to_cons<T, D>(&pp[j * NV], &cc[j * NV]);
      // interfaces: half node m (0..n+2) lies between cells m-2 and m-1
      for (int m = 0; m < n + 3; m++) {
        const T* wp[6];
        const T* wc[6];
        for (int j = 0; j < 6; j++) {
          int cell = m - 4 + j;  // cells m-4 .. m+1
          wp[j] = &pp[(cell + 4) * NV];
          wc[j] = &cc[(cell + 4) * NV];
        }
        T L[NV], R[NV];
        recon_interface<T, D>(g.scheme, g.charwise != 0, wp, wc, ax, L, R, g.variant);
        riemann<T, D>(g.flux, L, R, ax, &fh[m * NV]);
      }
      for (int m = 0; m < n + 2; m++) node_flux<T, D>(&pp[(m - 1 + 4) * NV], ax, &fn[m * NV]);
      for (int i = 0; i < n; i++) {
        T* r = &rhs[(base + i * sa) * NV];
        if (g.variant == 1) {
          for (int v = 0; v < NV; v++)
            r[v] += w1 * (fn[(i + 2) * NV + v] - fn[i * NV + v]) / hh +
                    w0 * (fh[(i + 2) * NV + v] - fh[(i + 1) * NV + v]) / hh +
                    w2 * (fh[(i + 3) * NV + v] - fh[i * NV + v]) / hh;
          continue;
        }
        for (int v = 0; v < NV; v++) {
          r[v] += (w1 * (fn[(i + 2) * NV + v] - fn[i * NV + v]) + w0 * (fh[(i + 2) * NV + v] - fh[(i + 1) * NV + v]) +
                   w2 * (fh[(i + 3) * NV + v] - fh[i * NV + v])) /
                  hh;
        }
      }
    }
  }
}

// Full RHS: setZeros, consToPrim, ghost fill, sweeps x, y, z (rhs = +sum dF/dx).
template <class T, int D> void rhs_eval(const Grid& g, const T* cons, T* prim, T* rhs, int nthreads) {
  constexpr int NV = D + 2;
  const long N = (long)g.n[0] * g.n[1] * g.n[2];
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long i = 0; i < N; i++) {
    cons_to_prim<T, D>(&cons[i * NV], &prim[i * NV]);
    for (int v = 0; v < NV; v++) rhs[i * NV + v] = T(0.0);
  }
  for (int ax = 0; ax < D; ax++) sweep<T, D>(g, ax, prim, rhs, nthreads);
  // SourceTerm::calGravitySource (zone_FD/src/src/sourceTerm.cpp:3-20), written
  // for an acceleration vector: rhs(rho u_d) -= rho g_d, rhs(rho E) -= rho (u . g)
  // (the reference's g = (0, 1) gives its rhs(i,2) -= r, rhs(i,3) -= r v).
  if (g.gravity[0] != 0.0 || g.gravity[1] != 0.0 || g.gravity[2] != 0.0) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long i = 0; i < N; i++) {
      const T* p = &prim[i * NV];
      T ug = p[1] * g.gravity[0];
      for (int d = 1; d < D; d++) ug = ug + p[1 + d] * g.gravity[d];
      for (int d = 0; d < D; d++) rhs[i * NV + 1 + d] -= p[0] * g.gravity[d];
      rhs[i * NV + NV - 1] -= p[0] * ug;
    }
  }
}

// BlockSolver::RK3_SSP (zone_FD/src/src/blockSolver.cpp:43-87) with its time
// bookkeeping (stages at t, t + dt, t + dt - dt/2) and BlockSolver::solve's
// final t += dt (:120-124); g0.t is the time at the start of the step.
template <class T, int D> double rk3_step(const Grid& g0, T* cons, double dt, int nthreads) {
  constexpr int NV = D + 2;
  const long N = (long)g0.n[0] * g0.n[1] * g0.n[2] * NV;
  std::vector<T> u0(cons, cons + N), prim(N), rhs(N);
  Grid g = g0;
  rhs_eval<T, D>(g, cons, prim.data(), rhs.data(), nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long i = 0; i < N; i++) cons[i] = u0[i] - dt * rhs[i];
  g.t += dt;
  rhs_eval<T, D>(g, cons, prim.data(), rhs.data(), nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long i = 0; i < N; i++) cons[i] = 0.75 * u0[i] - 0.25 * dt * rhs[i] + 0.25 * cons[i];
  g.t -= dt / 2;
  rhs_eval<T, D>(g, cons, prim.data(), rhs.data(), nthreads);
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (long i = 0; i < N; i++) cons[i] = 1.0 / 3.0 * u0[i] - 2.0 / 3.0 * dt * rhs[i] + 2.0 / 3.0 * cons[i];
  g.t -= dt / 2;
  return g.t + dt;
}

// BlockSolver::getTimeIntervalExplicit (zone_FD/src/src/blockSolver.cpp:318-348):
// dt = CFL / sum_d max_i (sqrt(gamma p / rho) + |u_d|) / h_d.
template <class T, int D> double cfl_dt(const Grid& g, const T* cons, double cfl, int nthreads) {
  constexpr int NV = D + 2;
  const long N = (long)g.n[0] * g.n[1] * g.n[2];
  double lam = 0;
  for (int ax = 0; ax < D; ax++) {
    double mx = 0;
#pragma omp parallel for num_threads(nthreads) reduction(max : mx)
    for (long i = 0; i < N; i++) {
      T p[NV];
      cons_to_prim<T, D>(&cons[i * NV], p);
      double li = (std::sqrt(gam() * double(p[NV - 1]) / double(p[0])) + std::abs(double(p[ax + 1]))) / g.h[ax];
      mx = std::max(mx, li);
    }
    lam += mx;
  }
  return cfl / lam;
}

}  // namespace oracle
