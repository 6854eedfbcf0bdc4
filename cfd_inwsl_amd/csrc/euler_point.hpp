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

#include <cmath>

namespace cfd {

// zone_FD/src/include/macro.hpp:28
#define CFD_GAMMA 1.4

enum SchemeId { SCH_JS = 0, SCH_Z = 1, SCH_TENO = 2, SCH_TENOS = 3, SCH_TENOA = 4, SCH_TENOSA = 5, SCH_WHF = 6 };
enum FluxId { FLX_HLLC = 0, FLX_ROE = 1 };

template <class T> __device__ __forceinline__ T sq(T a) { return a * a; }
// The reference's 1e-40 regulariser of WENO-Z / TENO is below the FP32
// normal range; the FP32 variant uses 1e-36 (a normal float) instead.
template <class T> __device__ __forceinline__ T eps_tiny() { return T(1e-40); }
template <> __device__ __forceinline__ float eps_tiny<float>() { return 1e-36f; }
// std::min / std::max argument semantics
template <class T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }
template <class T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }

// FP64 division and square root in the instruction sequences the compiler
// emits for IEEE a / b and sqrt(x) on gfx950 (reciprocal or reciprocal square
// root estimate, Newton-Raphson refinement, FMA correction), without their
// range-scaling and special-operand steps (v_div_scale, the scaling of
// v_div_fmas, v_div_fixup; the 2^-767 rescale and class test of sqrt).  Those
// steps are identities for the solver's operands -- nonzero normal divisors,
// dividends above 2^-969 in magnitude or zero, normal quotients, square-root
// arguments above 2^-767 or zero -- so the results are the correctly rounded
// ones, bit for bit (except that a zero dividend gives +0 where IEEE may give
// -0, which no later operation of the path distinguishes).  8 instead of 11
// and 13 instead of 19 instructions; FP32 keeps the compiler's sequences.
//
// The translation units of the opt-in FMA-contracted mode (CFD_FMA_MODE 1)
// take one Newton step fewer in each: the residual correction that ends both
// sequences squares the remaining error once more, so the quotient and root
// stay within an ulp (almost always the correctly rounded one) for two
// instructions less.
#ifndef CFD_FMA_MODE
#define CFD_FMA_MODE 0
#endif
template <class T> __device__ __forceinline__ T fdv(T a, T b) { return a / b; }
template <> __device__ __forceinline__ double fdv<double>(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  if (!CFD_FMA_MODE) {
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
  }
  const double q = a * r;
  const double t = __builtin_fma(-b, q, a);
  return __builtin_fma(t, r, q);
}
template <class T> __device__ __forceinline__ T fsq(T x) { return sqrt(x); }
// FP32 (the storage variant, compared with the FP64 path by its L1 / Linf
// error, not bit for bit): the hardware reciprocal (1 ulp) with one residual
// correction, and the hardware square root (1 ulp), in place of the
// compiler's correctly rounded sequences -- about a quarter of the FP32
// sweeps' instructions.  Operands are nonzero normal divisors and
// non-negative normal (or zero) square-root arguments.
template <> __device__ __forceinline__ float fdv<float>(float a, float b) {
  const float r = __builtin_amdgcn_rcpf(b);
  const float q = a * r;
  const float t = __builtin_fmaf(-b, q, a);
  return __builtin_fmaf(t, r, q);
}
template <> __device__ __forceinline__ float fsq<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double fsq<double>(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  if (!CFD_FMA_MODE) {
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
  }
  return x == 0.0 ? x : g;
}

// a / (2 b) as (a / b) / 2: 2 b is exact and a correctly rounded quotient
// scales exactly by 1/2 (operands far from the exponent limits), so this is
// the reference's a / (2 * b) bit for bit and shares the refined reciprocal
// of b with the other divisions by b (Roe: 2 c^2 and c^2, eigenvectors:
// 2 h and h).  FP32 keeps the plain form.
template <class T> __device__ __forceinline__ T half_quot(T a, T b) { return fdv(a, T(2.0) * b); }
template <> __device__ __forceinline__ double half_quot<double>(double a, double b) { return fdv(a, b) * 0.5; }

// Constants of the research TENO variants, from the host's libm pow as the
// reference computes them: ctp[k-5] = pow(1.5 c / (1 - c), 1/6) with
// c = pow(10, -k), cta[k-5] = pow(10, -k) (k = 5..10), ctp_s the first formula
// with c = 1e-5 (interScheme.hpp:1538, :1637, :2305-2306, :2598-2599).
struct ResearchConsts {
  double ctp[6], cta[6], ctp_s;
};

inline void fill_research_consts(ResearchConsts& x) {
  for (int k = 5; k <= 10; k++) {
    const double c = std::pow(10.0, -k);
    x.cta[k - 5] = c;
    x.ctp[k - 5] = std::pow(1.5 * c / (1.0 - c), 1.0 / 6);
  }
  x.ctp_s = std::pow(1.5 * 1e-5 / (1.0 - 1e-5), 1.0 / 6);
}

// ---- helpers of the research TENO variants -----------------------------------
__device__ __forceinline__ double tpow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float tpow(float x, float y) { return powf(x, y); }
// eps_A = (0.9 C_r) / (1 - 0.9 C_r) xi^2, C_r = 0.24, xi = 1e-3 (interScheme.hpp:1620)
constexpr double kEpsA2 = (0.9 * 0.24) / (1.0 - 0.9 * 0.24) * 1e-3 * 1e-3;
// smoothness ratio (|2 a b| + eps) / (a^2 + b^2 + eps) of two differences
template <class T> __device__ __forceinline__ T eta2(T a, T b, T eps) {
  return fdv(fabs(T(2.0) * a * b) + eps, a * a + b * b + eps);
}
// min of the ratios of adjacent differences (q0-q1, q1-q2), (q1-q2, q2-q3), (q2-q3, q3-q4)
template <class T> __device__ __forceinline__ T eta_adjacent(T q0, T q1, T q2, T q3, T q4, T eps) {
  const T d0 = q0 - q1, d1 = q1 - q2, d2 = q2 - q3, d3 = q3 - q4;
  return smin(smin(eta2(d0, d1, eps), eta2(d1, d2, eps)), eta2(d2, d3, eps));
}
// Integer powers the reference takes with libm pow: the products below differ
// from it by an ulp or two, which moves a cut-off decision only where its
// ratio ties the threshold to that precision.
template <class T> __device__ __forceinline__ T pow4(T x) {
  const T x2 = x * x;
  return x2 * x2;
}
template <class T> __device__ __forceinline__ T pow6(T x) {
  const T x2 = x * x;
  return x2 * x2 * x2;
}
// min of the ratios of adjacent values (q0,q1) .. (q3,q4) -- whf_TcnsN_AS and
// Teno5_CongAA take eta of the values themselves (interScheme.hpp:318-330, :1613-1627)
template <class T> __device__ __forceinline__ T eta_values(T q0, T q1, T q2, T q3, T q4) {
  return smin(smin(eta2(q0, q1, T(kEpsA2)), eta2(q1, q2, T(kEpsA2))),
              smin(eta2(q2, q3, T(kEpsA2)), eta2(q3, q4, T(kEpsA2))));
}
// k = floor(10 - 5 (1 - pow(1 - m, 4) (1 + 4 m))), m = 1 - min(1, eta / 0.24),
// clamped to 5..10 (interScheme.hpp:1633-1634, :2302-2304)
template <class T> __device__ __forceinline__ int adaptive_k(T eta) {
  const T m = T(1.0) - smin(T(1.0), fdv(eta, T(0.24)));
  const T g = pow4(T(1.0) - m) * (T(1.0) + T(4.0) * m);
  const int k = (int)floor(T(10.0) - T(5.0) * (T(1.0) - g));
  return k < 5 ? 5 : (k > 10 ? 10 : k);
}
// theta = 1 / (1 + max_k chi_k / 10), chi_k = tau / (beta_k + 1e-40)
// (interScheme.hpp:1697-1701, :2541-2545)
template <class T> __device__ __forceinline__ T lad_theta(T tau, T b0, T b1, T b2) {
  const T c0 = fdv(tau, b0 + eps_tiny<T>()), c1 = fdv(tau, b1 + eps_tiny<T>()), c2 = fdv(tau, b2 + eps_tiny<T>());
  return fdv(T(1.0), T(1.0) + fdv(smax(smax(c0, c1), c2), T(10.0)));
}
template <class T> __device__ __forceinline__ T ten_pow_neg(int k) {
  switch (k) {
  case 4: return T(1e-4);
  case 5: return T(1e-5);
  case 6: return T(1e-6);
  case 7: return T(1e-7);
  case 8: return T(1e-8);
  case 9: return T(1e-9);
  default: return T(1e-10);
  }
}

// ---- 5-point interpolation: value at q2 + h/2 ------------------------------
// Smoothness indicators (zone_FD/src/include/interScheme.hpp:20-27).
template <class T>
__device__ __forceinline__ void beta3(T q0, T q1, T q2, T q3, T q4, T& b0, T& b1, T& b2) {
  b0 = sq(q0 - T(2.0) * q1 + q2) + T(0.25) * sq(q0 - T(4.0) * q1 + T(3.0) * q2);
  b1 = sq(q1 - T(2.0) * q2 + q3) + T(0.25) * sq(q1 - q3);
  b2 = sq(q2 - T(2.0) * q3 + q4) + T(0.25) * sq(T(3.0) * q2 - T(4.0) * q3 + q4);
}

// TENO flag -> fixed stencil (interScheme.hpp:417-451).  Flag 0 (every
// stencil kept: smooth data, the common case) is tested first, so that a
// wave whose lanes all keep every stencil skips the other cases' blocks.
template <class T>
__device__ __forceinline__ T teno_pick(unsigned flag, T q0, T q1, T q2, T q3, T q4) {
  if (flag == 0u)
    return T(0.0234375) * q0 - T(0.15625) * q1 + T(0.703125) * q2 + T(0.46875) * q3 - T(0.0390625) * q4;
  switch (flag) {
  case 1: return -T(0.0625) * q1 + T(0.5625) * q2 + T(0.5625) * q3 - T(0.0625) * q4;
  case 2:
  case 3: return T(0.375) * q2 + T(0.75) * q3 - T(0.125) * q4;
  case 4: return T(0.0625) * q0 - T(0.3125) * q1 + T(0.9375) * q2 + T(0.3125) * q3;
  case 5: return -T(0.125) * q1 + T(0.75) * q2 + T(0.375) * q3;
  case 6: return T(0.375) * q0 - T(1.25) * q1 + T(1.875) * q2;
  default: return q2;
  }
}

// Every stencil of the gamma-ratio cut-off (Teno5_Z and the schemes that
// share its test) is kept whenever tau <= min_k (beta_k + 1e-40): then
// tau / (beta_k + 1e-40) <= 1, so 1 <= t_k = 1 + tau / (beta_k + 1e-40) <= 2
// and 1 <= gamma_k = t_k^6 <= 64 -- in exact arithmetic and in every
// correctly rounded evaluation alike, rounding being monotone and exact at
// 1, 2, 4, 64 -- hence ct * sum(gamma) <= 192 ct < 1 <= gamma_k for every
// cut-off constant ct <= 1e-4 of the family, and no test gamma_k < ct * sum
// succeeds: the flag is 0, the value the full evaluation returns.  Smooth
// data (tau = O(h^5), beta = O(h^2)) takes this path without any division.
// min_k (beta_k + 1e-40) is (min_k beta_k) + 1e-40 (rounding is monotone);
// the hardware minimum (NaN-ignoring) gives the same decision as std::min
// here: a NaN in beta_0 or beta_2 makes tau NaN and the test false either
// way, and a NaN beta_1 is skipped by both.
template <class T> __device__ __forceinline__ bool teno_keeps_all(T tau, T b0, T b1, T b2) {
  return tau <= fmin(fmin(b0, b1), b2) + eps_tiny<T>();
}

// Teno5_Z's cut-off flags (interScheme.hpp:384-452; ct = 1e-5, CT4 1e-4,
// CT7 1e-7): bit k set where stencil k is cut.
template <class T>
__device__ __forceinline__ unsigned teno5_flag(T q0, T q1, T q2, T q3, T q4, T ct) {
  T b0, b1, b2;
  beta3(q0, q1, q2, q3, q4, b0, b1, b2);
  T tau = fabs(b2 - b0);
  if (teno_keeps_all(tau, b0, b1, b2)) return 0u;
  T t0, t1, t2;
  if (sizeof(T) == sizeof(double)) {
    t0 = T(1.0) + fdv(tau, b0 + eps_tiny<T>());
    t1 = T(1.0) + fdv(tau, b1 + eps_tiny<T>());
    t2 = T(1.0) + fdv(tau, b2 + eps_tiny<T>());
  } else {
    // FP32: (1 + tau / beta)^6 overflows for tau / beta > ~2.6e6.  The
    // cut-off test g_k < ct * sum(g) is invariant to a common factor, so
    // the terms are scaled by min_k beta_k / tau when tau exceeds it.
    T s0 = b0 + eps_tiny<T>(), s1 = b1 + eps_tiny<T>(), s2 = b2 + eps_tiny<T>();
    T sm = smin(smin(s0, s1), s2);
    T base = T(1.0);
    T num = tau;
    if (tau > sm) {
      base = fdv(sm, tau);
      num = sm;
    }
    t0 = base + fdv(num, s0);
    t1 = base + fdv(num, s1);
    t2 = base + fdv(num, s2);
  }
  t0 = t0 * t0;
  t1 = t1 * t1;
  t2 = t2 * t2;
  T g0 = t0 * t0 * t0, g1 = t1 * t1 * t1, g2 = t2 * t2 * t2;
  T CT = ct * (g0 + g1 + g2);
  unsigned flag = (g0 < CT ? 1u : 0u) + (g1 < CT ? 2u : 0u) + (g2 < CT ? 4u : 0u);
  return flag;
}

// Cut-off flag of the TENO-family schemes (bit k set where stencil k is
// cut); the interpolated value is the flag's stencil (teno_pick).
template <int SCH, class T>
__device__ __forceinline__ unsigned cut_flag(T q0, T q1, T q2, T q3, T q4, T ct, const ResearchConsts* rc) {
  T b0, b1, b2;
  beta3(q0, q1, q2, q3, q4, b0, b1, b2);
  if (SCH == SCH_TENO) {
    return teno5_flag(q0, q1, q2, q3, q4, ct);
  } else if (SCH == SCH_TENOA) {
    // new_whf_TCNS_A, TENO-A (interScheme.hpp:3863-3983): Teno5_Z's cut-off with
    // the adaptive constant 10^-beta_A, beta_A = floor(10 - 5 (1 - g)),
    // g = s^4 (5 - 4 s), s = min(1, 4.1666667 eta_min)
    T tau = fabs(b2 - b0);
    // the same cut-off test with a constant 10^-beta_A <= 1e-5: all kept
    // whatever beta_A is, so the smoothness ratios are not needed
    if (teno_keeps_all(tau, b0, b1, b2)) return 0u;
    const T epsA = T(2.7551 * 1e-7);
    const T d0 = q0 - q1, d1 = q1 - q2, d2 = q2 - q3, d3 = q3 - q4;
    const T e0 = fdv(fabs(T(2.0) * d0 * d1) + epsA, d0 * d0 + d1 * d1 + epsA);
    const T e1 = fdv(fabs(T(2.0) * d1 * d2) + epsA, d1 * d1 + d2 * d2 + epsA);
    const T e2 = fdv(fabs(T(2.0) * d2 * d3) + epsA, d2 * d2 + d3 * d3 + epsA);
    const T emin = smin(smin(e0, e1), e2);
    const T mn = smin(T(1.0), T(4.1666667) * emin);
    const T gm = mn * mn * mn * mn * (T(5.0) - T(4) * mn);
    const int betaA = (int)floor(T(10.0) - T(5.0) * (T(1.0) - gm));
    T t0, t1, t2;
    if (sizeof(T) == sizeof(double)) {
      t0 = T(1.0) + fdv(tau, b0 + eps_tiny<T>());
      t1 = T(1.0) + fdv(tau, b1 + eps_tiny<T>());
      t2 = T(1.0) + fdv(tau, b2 + eps_tiny<T>());
    } else {
      // FP32: the scaled terms of the TENO branch (the test is scale invariant)
      T s0 = b0 + eps_tiny<T>(), s1 = b1 + eps_tiny<T>(), s2 = b2 + eps_tiny<T>();
      T sm = smin(smin(s0, s1), s2);
      T base = T(1.0);
      T num = tau;
      if (tau > sm) {
        base = fdv(sm, tau);
        num = sm;
      }
      t0 = base + fdv(num, s0);
      t1 = base + fdv(num, s1);
      t2 = base + fdv(num, s2);
    }
    t0 = t0 * t0;
    t1 = t1 * t1;
    t2 = t2 * t2;
    T g0 = t0 * t0 * t0, g1 = t1 * t1 * t1, g2 = t2 * t2 * t2;
    const T cta = betaA <= 5 ? T(1e-5) : betaA == 6 ? T(1e-6) : betaA == 7 ? T(1e-7)
                : betaA == 8 ? T(1e-8) : betaA == 9 ? T(1e-9) : T(1e-10);
    T rr = cta * (g0 + g1 + g2);
    unsigned flag = (g0 < rr ? 1u : 0u) + (g1 < rr ? 2u : 0u) + (g2 < rr ? 4u : 0u);
    return flag;
  } else if (SCH == SCH_TENOSA) {
    // TENO-S with its constant C_T adapted to the smoothness ratios eta of
    // adjacent differences (those of TENO-A); ct selects the law:
    //   1: whf_zyc_TcnsN_myASF002_1 (interScheme.hpp:3363-3460),
    //      s = min(1, 4.1666667 eta_min), C_T = 0.157 / (11.351092 s^2 - 5.740650 s + 1.506631)
    //   2: whf_TCNS_AS_myF203_NoS (:3984-4091), m = 1 - min(0.24, eta_min) / 0.24,
    //      C_T = -0.452662 m^3 + 0.590811 m^2 - 0.012189 m + 0.022010
    //   3: whf_TCNS_AS_myF202_NoS (:4093-4200), the same m, C_T = -0.088182 m^2 + 0.259395 m - 0.000616
    //   5: whf_ai_TcnsN_AS_1 / _AS_2 (:2329-2511), C_T = ctp[k - 5], k = adaptive_k(eta_min), eps_A = kEpsA2
    //   6: whf_ai_TcnsN_LADS_yuanbao (:2512-2647), C_T = ctp[m - 5], m = 5 + floor(5 theta)
    //   7: whf_ai_TcnsN_AZS_yuanbao (:2862-2935), C_T = pow(1.5 c / (1 - c), 1/6),
    //      c = min(1e-5, |beta0 + beta2 - 2 beta1|)
    //   8: whf_ai_TcnsN_myASF002_1 / _2 / _ai1 (:2936-3252), law 1 with s = min(1, eta_min / 0.24), eps_A = kEpsA2
    //   9: Teno5_CongAA (:304-382), C_T = cta[k - 5] with k = round(10 - 5 (1 - (1-m)^4 (1+4m))),
    //      eta of the values q themselves, eps_A = kEpsA2
    const T bm = smin(smin(b0, b1), b2);
    T tau = fabs(b2 - b0);
    const int law = (int)ct;
    T CT;
    if (law <= 3) {
      const T emin = eta_adjacent(q0, q1, q2, q3, q4, T(2.7551 * 1e-7));
      if (law == 1) {
        const T x = smin(T(1.0), T(4.1666667) * emin);
        CT = fdv(T(0.157), T(11.351092) * x * x - T(5.740650) * x + T(1.506631));
      } else {
        const T m = T(1.0) - fdv(smin(T(0.24), emin), T(0.24));
        if (law == 2) {
          const T mm = m * m, mmm = m * mm;
          CT = T(-0.452662) * mmm + T(0.590811) * mm - T(0.012189) * m + T(0.022010);
        } else {
          CT = T(-0.088182) * m * m + T(0.259395) * m - T(0.000616);
        }
      }
    } else if (law == 5) {
      CT = T(rc->ctp[adaptive_k(eta_adjacent(q0, q1, q2, q3, q4, T(kEpsA2))) - 5]);
    } else if (law == 6) {
      int m = 5 + (int)floor(lad_theta(tau, b0, b1, b2) * T(5.0));
      m = m < 5 ? 5 : (m > 10 ? 10 : m);
      CT = T(rc->ctp[m - 5]);
    } else if (law == 7) {
      const T c = smin(T(1e-5), fabs(b0 + b2 - T(2) * b1));
      CT = tpow(fdv(T(1.5) * c, T(1) - c), T(1.0 / 6.0));
    } else if (law == 8) {
      const T x = smin(T(1.0), fdv(eta_adjacent(q0, q1, q2, q3, q4, T(kEpsA2)), T(0.24)));
      CT = fdv(T(0.157), T(11.351092) * x * x - T(5.740650) * x + T(1.506631));
    } else {
      const T e = eta_values(q0, q1, q2, q3, q4);
      const T m = T(1.0) - smin(T(1.0), fdv(e, T(0.24)));
      const T g = pow4(T(1.0) - m) * (T(1.0) + T(4.0) * m);
      int k = (int)round(T(10.0) - T(5.0) * (T(1.0) - g));
      CT = T(rc->cta[(k < 5 ? 5 : (k > 10 ? 10 : k)) - 5]);
    }
    const T CT_1 = 1 - CT;
    T rr = CT * tau - CT_1 * bm;
    T ll = tau * bm;
    unsigned flag = (ll < rr * b0 ? 1u : 0u) + (ll < rr * b1 ? 2u : 0u) + (ll < rr * b2 ? 4u : 0u);
    return flag;
  } else if (SCH == SCH_WHF) {
    // the whf_TcnsN schemes whose cut-off is not TENO-S's; ct = the
    // InterMethod value.  The value depends on the data through the flag
    // only, so the libm pow of the reference (here pow4 / pow6) decides bits
    // only where a ratio ties its threshold.
    const int var = (int)ct;
    T tau = fabs(b2 - b0);
    unsigned flag;
    if (var == 17 || var == 21 || var == 25) {
      // whf_TcnsN (:1458-1517), whf_TcnsN_LAD (:1681-1760), whf_ai_TcnsN_A_2
      // (:2216-2328): chi_k = gamma_k / sum gamma < C_T cuts k,
      // gamma_k = pow(1 + tau / (beta_k + 1e-40), 6)
      const T g0 = pow6(T(1.0) + fdv(tau, b0 + eps_tiny<T>()));
      const T g1 = pow6(T(1.0) + fdv(tau, b1 + eps_tiny<T>()));
      const T g2 = pow6(T(1.0) + fdv(tau, b2 + eps_tiny<T>()));
      const T gs = g0 + g1 + g2;
      T CT;
      if (var == 17) {
        CT = T(1e-10);
      } else if (var == 21) {  // C_T = 10^-m, m = clamp(4 + floor(6 theta), 4, 10)
        int m = 4 + (int)floor(lad_theta(tau, b0, b1, b2) * T(6.0));
        CT = ten_pow_neg<T>(m < 4 ? 4 : (m > 10 ? 10 : m));
      } else {  // C_TA = pow(10, -k)
        CT = T(rc->cta[adaptive_k(eta_adjacent(q0, q1, q2, q3, q4, T(kEpsA2))) - 5]);
      }
      flag = (fdv(g0, gs) < CT ? 1u : 0u) + (fdv(g1, gs) < CT ? 2u : 0u) + (fdv(g2, gs) < CT ? 4u : 0u);
    } else if (var == 18 || var == 34) {
      // whf_TcnsN_A (:3253-3362), whf_TcnsN_COMPARE (:3472-3726, its returned
      // branch): Teno5_Z's cut-off with C_T = 10^-beta_A, beta_A = floor(10 -
      // 5 (1 - (1-m)^4 (1+4m))); eta of the difference pairs (d0,d2), (d1,d3), (d2,d1)
      const T eps = var == 18 ? T(2.7551 * 1e-7) : T(kEpsA2);
      const T d0 = q0 - q1, d1 = q1 - q2, d2 = q2 - q3, d3 = q3 - q4;
      const T emin = smin(smin(eta2(d0, d2, eps), eta2(d1, d3, eps)), eta2(d2, d1, eps));
      const T x = var == 18 ? T(4.1666667) * emin : fdv(emin, T(0.24));
      const T m = T(1.0) - smin(T(1.0), x);
      const T gm = (1 - m) * (1 - m) * (1 - m) * (1 - m) * (1 + 4 * m);
      const int k = (int)floor(T(10.0) - T(5.0) * (1 - gm));
      const T CT = ten_pow_neg<T>(k < 5 ? 5 : (k > 10 ? 10 : k));
      T t0 = T(1.0) + fdv(tau, b0 + eps_tiny<T>());
      T t1 = T(1.0) + fdv(tau, b1 + eps_tiny<T>());
      T t2 = T(1.0) + fdv(tau, b2 + eps_tiny<T>());
      t0 = t0 * t0;
      t1 = t1 * t1;
      t2 = t2 * t2;
      const T g0 = t0 * t0 * t0, g1 = t1 * t1 * t1, g2 = t2 * t2 * t2;
      const T rr = CT * (g0 + g1 + g2);
      flag = (g0 < rr ? 1u : 0u) + (g1 < rr ? 2u : 0u) + (g2 < rr ? 4u : 0u);
    } else {
      // whf_TcnsN_S (:1519-1579, C_T' = ctp_s) and whf_TcnsN_AS (:1581-1679,
      // C_T' = ctp[k - 5], its eta taken of the values q themselves):
      // beta_k tau < (C_T' tau - (1 - C_T') beta_k) beta_min cuts k
      T CTp;
      if (var == 20) {
        CTp = T(rc->ctp_s);
      } else {
        CTp = T(rc->ctp[adaptive_k(eta_values(q0, q1, q2, q3, q4)) - 5]);
      }
      const T bm = smin(smin(b0, b1), b2);
      const T c1 = T(1.0) - CTp;
      flag = (b0 * tau < (CTp * tau - c1 * b0) * bm ? 1u : 0u) + (b1 * tau < (CTp * tau - c1 * b1) * bm ? 2u : 0u) +
             (b2 * tau < (CTp * tau - c1 * b2) * bm ? 4u : 0u);
    }
    return flag;
  } else {
    // Teno5_CongZ, TENO-S (interScheme.hpp:494-585); ct = its constant CT
    // (0.15704178024750198; Teno5_CongZCT4 / CT7, :586 / :661, differ in it only)
    T bm = b0;
    if (b1 < bm) bm = b1;
    if (b2 < bm) bm = b2;
    const T CT = ct;
    const T CT_1 = 1 - CT;
    T tau = fabs(b2 - b0);
    T rr = CT * tau - CT_1 * bm;
    T ll = tau * bm;
    unsigned flag = (ll < rr * b0 ? 1u : 0u) + (ll < rr * b1 ? 2u : 0u) + (ll < rr * b2 ? 4u : 0u);
    return flag;
  }
}

template <int SCH, class T>
__device__ __forceinline__ T inter5(T q0, T q1, T q2, T q3, T q4, T ct, const ResearchConsts* rc = nullptr) {
  T b0, b1, b2;
  beta3(q0, q1, q2, q3, q4, b0, b1, b2);
  if (SCH == SCH_JS || SCH == SCH_Z) {
    // weno5_JSchen (interScheme.hpp:17-43) / weno5_Z (:85-114)
    T u0 = T(0.375) * q0 - T(1.25) * q1 + T(1.875) * q2;
    T u1 = -T(0.125) * q1 + T(0.75) * q2 + T(0.375) * q3;
    T u2 = T(0.375) * q2 + T(0.75) * q3 - T(0.125) * q4;
    T a0, a1, a2;
    if (SCH == SCH_JS) {
      a0 = fdv(T(0.0625), sq(T(1e-6) + b0));
      a1 = fdv(T(0.625), sq(T(1e-6) + b1));
      a2 = fdv(T(0.3125), sq(T(1e-6) + b2));
    } else if (sizeof(T) == sizeof(double)) {
      T tau = fabs(b2 - b0);
      a0 = T(0.0625) * (T(1.0) + sq(fdv(tau, b0 + eps_tiny<T>())));
      a1 = T(0.625) * (T(1.0) + sq(fdv(tau, b1 + eps_tiny<T>())));
      a2 = T(0.3125) * (T(1.0) + sq(fdv(tau, b2 + eps_tiny<T>())));
    } else {
      // FP32: (tau / beta)^2 overflows the float range once beta is ~19
      // orders below tau, and inf / inf would follow.  The weights are
      // invariant to a common factor, so they are scaled by (min_k beta_k /
      // tau)^2 whenever tau exceeds the smallest beta: every term stays in
      // [0, 2] and the normalised weights are the same real numbers.
      T tau = fabs(b2 - b0);
      T s0 = b0 + eps_tiny<T>(), s1 = b1 + eps_tiny<T>(), s2 = b2 + eps_tiny<T>();
      T sm = smin(smin(s0, s1), s2);
      T base = T(1.0);
      T num = tau;
      if (tau > sm) {
        base = sq(fdv(sm, tau));
        num = sm;
      }
      a0 = T(0.0625) * (base + sq(fdv(num, s0)));
      a1 = T(0.625) * (base + sq(fdv(num, s1)));
      a2 = T(0.3125) * (base + sq(fdv(num, s2)));
    }
    T s = a0 + a1 + a2;
    T r = a0 * u0 + a1 * u1 + a2 * u2;
    return fdv(r, s);
  }
  if (SCH == SCH_WHF && (int)ct == 15) {
    // musclIn5 (interScheme.hpp:1386-1406): q2 + delta / 2, delta the minmod slope of q1, q2, q3
    const T dm = q2 - q1, dp = q3 - q2;
    const T delta = dp > T(0) ? smax(T(0.0), smax(smin(dm, dp), smin(dm, dp)))
                              : smin(T(0.0), smin(smax(dm, dp), smax(dm, dp)));
    return q2 + delta * T(0.5);
  }
  return teno_pick(cut_flag<SCH>(q0, q1, q2, q3, q4, ct, rc), q0, q1, q2, q3, q4);
}

// ---- conservative <-> primitive ----------------------------------------------
// Equation::consToPrimEuler1D/2D (zone_FD/src/src/equation.cpp:37-99); the 3D
// form appends w*w to the 2D sum.
template <int D, class T> __device__ __forceinline__ void cons_to_prim(const T* c, T* p) {
  T r = c[0];
  T u = fdv(c[1], r);
  if (D == 1) {
    T E = fdv(c[2], r);
    T e = -u * u / 2 + E;
    T RT = (T(CFD_GAMMA) - T(1)) * e;
    p[0] = r; p[1] = u; p[2] = r * RT;
  } else if (D == 2) {
    T v = fdv(c[2], r);
    T E = fdv(c[3], r);
    T q2 = (u * u + v * v) / 2;
    T e = -q2 + E;
    T RT = (T(CFD_GAMMA) - T(1)) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = r * RT;
  } else {
    T v = fdv(c[2], r);
    T w = fdv(c[3], r);
    T E = fdv(c[4], r);
    T q2 = (u * u + v * v + w * w) / 2;
    T e = -q2 + E;
    T RT = (T(CFD_GAMMA) - T(1)) * e;
    p[0] = r; p[1] = u; p[2] = v; p[3] = w; p[4] = r * RT;
  }
}

// Conservative state rebuilt inside primToChar (eigenSystem.cpp:91-93, 254-256).
template <int D, class T> __device__ __forceinline__ void prim_to_cons(const T* p, T* c) {
  T r = p[0];
  if (D == 1) {
    T ek = (p[1] * p[1]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = fdv(p[2], T(CFD_GAMMA) - T(1)) + r * ek;
  } else if (D == 2) {
    T ek = (p[1] * p[1] + p[2] * p[2]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = fdv(p[3], T(CFD_GAMMA) - T(1)) + r * ek;
  } else {
    T ek = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) / 2;
    c[0] = r; c[1] = r * p[1]; c[2] = r * p[2]; c[3] = r * p[3];
    c[4] = fdv(p[4], T(CFD_GAMMA) - T(1)) + r * ek;
  }
}

// Back-projected conservative state -> primitive (charToPrim tail,
// eigenSystem.cpp:163-167; 1D consToPrim :309-322).
template <int D, class T> __device__ __forceinline__ void eig_cons_to_prim(const T* c, T* p) {
  T rt = c[0];
  if (D == 1) {
    T ut = fdv(c[1], rt);
    T ekt = (ut * ut) / 2;
    p[0] = rt; p[1] = ut; p[2] = (T(CFD_GAMMA) - T(1)) * (c[2] - rt * ekt);
  } else if (D == 2) {
    T ut = fdv(c[1], rt), vt = fdv(c[2], rt);
    T rekt = fdv(c[1] * c[1] + c[2] * c[2], rt) / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = (T(CFD_GAMMA) - T(1)) * (c[3] - rekt);
  } else {
    T ut = fdv(c[1], rt), vt = fdv(c[2], rt), wt = fdv(c[3], rt);
    T rekt = fdv(c[1] * c[1] + c[2] * c[2] + c[3] * c[3], rt) / 2;
    p[0] = rt; p[1] = ut; p[2] = vt; p[3] = wt; p[4] = (T(CFD_GAMMA) - T(1)) * (c[4] - rekt);
  }
}

template <int D, class T> __device__ __forceinline__ T project(const T* c, const T* l);

// ---- Roe-averaged characteristic system (2D / 3D) --------------------------
// eigensystemEuler2D(priml, primr, norm) (eigenSystem.cpp:13-86).  The 3D
// system adds a second shear wave last; for x/y normals the first shear wave
// is the 2D one and the w terms are inserted so that w = 0 reproduces 2D.
// Tangents: x: t1=(0,-1,0) t2=(0,0,1); y: t1=(1,0,0) t2=(0,0,1);
//           z: t1=(1,0,0) t2=(0,1,0).
template <int D, int AX, class T> struct RoeSys {
  static constexpr int NV = D + 2;
  static constexpr T nx = AX == 0 ? T(1.0) : T(0.0);
  static constexpr T ny = AX == 1 ? T(1.0) : T(0.0);
  static constexpr T nz = AX == 2 ? T(1.0) : T(0.0);
  T u, v, w, ek, h, c, Vn;

  __device__ __forceinline__ void build(const T* pl, const T* pr) {
    T rl = pl[0], rr = pr[0];
    T ul = pl[1], ur = pr[1], vl = pl[2], vr = pr[2];
    T ppl = pl[NV - 1], ppr = pr[NV - 1];
    T HL, HR;
    T wl = T(0.0), wr = T(0.0);
    if (D == 2) {
      HL = (ul * ul + vl * vl) / 2 + fdv(fdv(ppl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
      HR = (ur * ur + vr * vr) / 2 + fdv(fdv(ppr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    } else {
      wl = pl[3];
      wr = pr[3];
      HL = (ul * ul + vl * vl + wl * wl) / 2 + fdv(fdv(ppl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
      HR = (ur * ur + vr * vr + wr * wr) / 2 + fdv(fdv(ppr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    }
    T c1 = fsq(rl), c2 = fsq(rr);
    T divisor = fdv(T(1.0), c1 + c2);
    u = (c1 * ul + c2 * ur) * divisor;
    v = (c1 * vl + c2 * vr) * divisor;
    T ht = (c1 * HL + c2 * HR) * divisor;
    // Vn = nx u + ny v [+ nz w] of eigenSystem.cpp with the constant normal:
    // the zero terms add zeros to the one nonzero product, so Vn is the
    // normal component itself (the sign of an exact zero aside)
    if (D == 2) {
      w = T(0.0);
      Vn = AX == 0 ? u : v;
      ek = (u * u + v * v) / 2;
    } else {
      w = (c1 * wl + c2 * wr) * divisor;
      Vn = AX == 0 ? u : (AX == 1 ? v : w);
      ek = (u * u + v * v + w * w) / 2;
    }
    h = ht - ek;
    c = fsq((T(CFD_GAMMA) - T(1)) * h);
  }

  // Row k of the left eigenvector matrix (columns rho, rho u, rho v[, rho w], rho E).
  __device__ __forceinline__ void lrow(int k, T* l) const {
    if (k == 0) {
      if (D == 3 && AX == 2) {
        l[0] = -u; l[1] = T(1.0); l[2] = T(0.0);
      } else {
        l[0] = (-ny * u + nx * v); l[1] = ny; l[2] = -nx;
      }
      if (D == 3) l[3] = T(0.0);
      l[NV - 1] = T(0.0);
    } else if (k == 1) {
      l[0] = h - ek; l[1] = u; l[2] = v;
      if (D == 3) l[3] = w;
      l[NV - 1] = -T(1.0);
    } else if (k == 2) {
      // (n_d / c - u_d / h) / 2 with n_d = -1 on the normal axis; on the
      // others 0 / c is +0 and (+0 - u_d / h) / 2 = -(u_d / h) / 2
      l[0] = (fdv(Vn, c) + fdv(ek, h)) / 2;
      l[1] = AX == 0 ? (fdv(T(-1.0), c) - fdv(u, h)) / 2 : -fdv(u, h) / 2;
      l[2] = AX == 1 ? (fdv(T(-1.0), c) - fdv(v, h)) / 2 : -fdv(v, h) / 2;
      if (D == 3) l[3] = AX == 2 ? (fdv(T(-1.0), c) - fdv(w, h)) / 2 : -fdv(w, h) / 2;
      l[NV - 1] = half_quot(T(1.0), h);
    } else if (k == 3) {
      l[0] = (fdv(-Vn, c) + fdv(ek, h)) / 2;
      l[1] = AX == 0 ? (fdv(T(1.0), c) - fdv(u, h)) / 2 : -fdv(u, h) / 2;
      l[2] = AX == 1 ? (fdv(T(1.0), c) - fdv(v, h)) / 2 : -fdv(v, h) / 2;
      if (D == 3) l[3] = AX == 2 ? (fdv(T(1.0), c) - fdv(w, h)) / 2 : -fdv(w, h) / 2;
      l[NV - 1] = half_quot(T(1.0), h);
    } else {  // k == 4, 3D only: shear 2 along t2
      if (AX == 2) { l[0] = -v; l[1] = T(0.0); l[2] = T(1.0); l[3] = T(0.0); }
      else { l[0] = -w; l[1] = T(0.0); l[2] = T(0.0); l[3] = T(1.0); }
      l[4] = T(0.0);
    }
  }

  // Column k of the right eigenvector matrix (rows rho, rho u, ..., rho E).
  __device__ __forceinline__ void rcol(int k, T* r) const {
    if (k == 0) {
      r[0] = T(0.0);
      if (D == 3 && AX == 2) { r[1] = T(1.0); r[2] = T(0.0); r[3] = T(0.0); r[4] = u; }
      else {
        r[1] = ny; r[2] = -nx;
        if (D == 3) r[3] = T(0.0);
        r[NV - 1] = ny * u - nx * v;
      }
    } else if (k == 1) {
      r[0] = fdv(T(1.0), h); r[1] = fdv(u, h); r[2] = fdv(v, h);
      if (D == 3) r[3] = fdv(w, h);
      r[NV - 1] = fdv(ek, h);
    } else if (k == 2) {
      r[0] = T(1.0); r[1] = u - nx * c; r[2] = v - ny * c;
      if (D == 3) r[3] = w - nz * c;
      r[NV - 1] = h + ek - Vn * c;
    } else if (k == 3) {
      r[0] = T(1.0); r[1] = u + nx * c; r[2] = v + ny * c;
      if (D == 3) r[3] = w + nz * c;
      r[NV - 1] = h + ek + Vn * c;
    } else {
      r[0] = T(0.0);
      if (AX == 2) { r[1] = T(0.0); r[2] = T(1.0); r[3] = T(0.0); r[4] = v; }
      else { r[1] = T(0.0); r[2] = T(0.0); r[3] = T(1.0); r[4] = w; }
    }
  }

  // Left row k applied to a conservative state: primToChar's sum in
  // project<D>'s order, with the constant zeros and units of the shear rows
  // folded (a product c_i * 0 adds a zero to a nonzero partial sum, c_i * -1
  // is -c_i exactly), so the value is the full sum's, bit for bit, up to the
  // sign of an exact-zero result.  l holds lrow(k) for the other rows.
  //   k = 0, x normal: (v, 0, -1, 0[, 0]);  y normal, 3D z normal: (-u, 1, 0, 0[, 0])
  //   k = 4 (3D), z normal: (-v, 0, 1, 0, 0);  x / y normal: (-w, 0, 0, 1, 0)
  __device__ __forceinline__ T proj(int k, const T* l, const T* cc) const {
    if (k == 0) return AX == 0 ? cc[0] * v - cc[2] : cc[1] - cc[0] * u;
    if (D == 3 && k == 4) return AX == 2 ? cc[2] - cc[0] * v : cc[3] - cc[0] * w;
    return project<D>(cc, l);
  }

  // charToCons: o_i = sum of Q_k r_k[i] over k = 2, 3, 1, 0[, 4] (the
  // reference's summation order), the right vectors' constant zeros and
  // units folded as in proj:
  //   r_2 = (1, u - n_u c, v - n_v c[, w - n_w c], h + ek - Vn c), r_3 the same with + c
  //   r_1 = (1, u, v[, w], ek) / h
  //   r_0: x normal (0, 0, -1[, 0], -v); y normal, 3D z normal (0, 1, 0[, 0], u)
  //   r_4 (3D): z normal (0, 0, 1, 0, v); x / y normal (0, 0, 0, 1, w)
  __device__ __forceinline__ void back(const T* QL, const T* QR, T* oL, T* oR) const {
    constexpr int E = NV - 1;
    auto both = [&](int i, int k, T r) {  // o_i = Q_k r (first term)
      oL[i] = QL[k] * r;
      oR[i] = QR[k] * r;
    };
    auto add = [&](int i, int k, T r) {  // o_i += Q_k r
      oL[i] = oL[i] + QL[k] * r;
      oR[i] = oR[i] + QR[k] * r;
    };
    oL[0] = QL[2];
    oR[0] = QR[2];
    both(1, 2, AX == 0 ? u - c : u);
    both(2, 2, AX == 1 ? v - c : v);
    if (D == 3) both(3, 2, AX == 2 ? w - c : w);
    both(E, 2, h + ek - Vn * c);
    oL[0] = oL[0] + QL[3];
    oR[0] = oR[0] + QR[3];
    add(1, 3, AX == 0 ? u + c : u);
    add(2, 3, AX == 1 ? v + c : v);
    if (D == 3) add(3, 3, AX == 2 ? w + c : w);
    add(E, 3, h + ek + Vn * c);
    add(0, 1, fdv(T(1.0), h));
    add(1, 1, fdv(u, h));
    add(2, 1, fdv(v, h));
    if (D == 3) add(3, 1, fdv(w, h));
    add(E, 1, fdv(ek, h));
    if (AX == 0) {
      oL[2] = oL[2] - QL[0];
      oR[2] = oR[2] - QR[0];
      add(E, 0, -v);
    } else {
      oL[1] = oL[1] + QL[0];
      oR[1] = oR[1] + QR[0];
      add(E, 0, u);
    }
    if (D == 3) {
      oL[AX == 2 ? 2 : 3] = oL[AX == 2 ? 2 : 3] + QL[4];
      oR[AX == 2 ? 2 : 3] = oR[AX == 2 ? 2 : 3] + QR[4];
      add(4, 4, AX == 2 ? v : w);
    }
  }
};

// Projection of a conservative state on one left eigenvector row, in the
// summation order of primToChar (eigenSystem.cpp:97-107; 1D :260-264).
template <int D, class T> __device__ __forceinline__ T project(const T* c, const T* l) {
  if (D == 1) return c[0] * l[0] + c[1] * l[1] + c[2] * l[2];
  if (D == 2) return c[1] * l[1] + c[2] * l[2] + c[3] * l[3] + c[0] * l[0];
  return c[1] * l[1] + c[2] * l[2] + c[3] * l[3] + c[4] * l[4] + c[0] * l[0];
}

// ---- 1D Roe system (eigensystemEuler1D, eigenSystem.cpp:217-250) ----------
template <class T> struct RoeSys1 {
  T L[9], R[9];
  __device__ __forceinline__ void build(const T* pl, const T* pr) {
    T rl = pl[0], ul = pl[1], ppl = pl[2];
    T rr = pr[0], ur = pr[1], ppr = pr[2];
    T HL = (ul * ul) / 2 + fdv(fdv(ppl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    T HR = (ur * ur) / 2 + fdv(fdv(ppr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    T c1 = fsq(rl), c2 = fsq(rr);
    T divisor = fdv(T(1.0), fsq(rl) + fsq(rr));
    T u = (c1 * ul + c2 * ur) * divisor;
    T ht = (c1 * HL + c2 * HR) * divisor;
    T ek = (u * u) / 2;
    T h = ht - ek;
    T c = fsq((T(CFD_GAMMA) - T(1)) * h);
    const T g1 = T(CFD_GAMMA) - T(1);
    T f = fdv(T(0.5) * g1, c * c);
    L[0] = (ht + fdv(c * (u - c), g1)) * f;
    L[1] = (-u - fdv(c, g1)) * f;
    L[2] = T(1.0) * f;
    L[3] = (-T(2.0) * ht + fdv(T(4.0) * c * c, g1)) * f;
    L[4] = (T(2.0) * u) * f;
    L[5] = -T(2.0) * f;
    L[6] = (ht - fdv(c * (u + c), g1)) * f;
    L[7] = (-u + fdv(c, g1)) * f;
    L[8] = T(1.0) * f;
    R[0] = T(1.0); R[1] = T(1.0); R[2] = T(1.0);
    R[3] = u - c; R[4] = u; R[5] = u + c;
    R[6] = ht - u * c; R[7] = T(0.5) * u * u; R[8] = ht + u * c;
  }
};

// variable_positive_limiter (zone_FD/src/include/positive_preserving.hpp:14-46)
// with q = rho (which 0) and q = gamma p (which 1).
template <class T> __device__ __forceinline__ T lim_q(const T* W, int which) {
  if (which == 0) return W[0];
  T u = fdv(W[1], W[0]);
  T ekin = T(0.5) * u * u;
  T p = (T(CFD_GAMMA) - T(1)) * (W[2] - W[0] * ekin);
  return T(CFD_GAMMA) * p;
}
template <class T> __device__ __forceinline__ void positive_limit(const T* W, T* Wi, int which) {
  T qi = lim_q(Wi, which);
  if (qi >= T(1e-10)) return;
  T qw = lim_q(W, which);
  if (qw < T(1e-10)) {
    Wi[0] = W[0]; Wi[1] = W[1]; Wi[2] = W[2];
    return;
  }
  T theta = fdv(T(1e-10) - qw, qi - qw);
  for (int i = 0; i < 3; i++) Wi[i] = (1 - theta) * W[i] + theta * Wi[i];
}

// ---- interface Riemann fluxes ----------------------------------------------
// HLLCFlux1D (zone_FD/src/include/fluxSchemes.hpp:74-125)
template <class T> __device__ __forceinline__ void hllc1d(const T* L, const T* Rr, T* res) {
  T rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  T HL = fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1)) + (ul * ul) / 2;
  T HR = fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1)) + (ur * ur) / 2;
  T cl = fsq(T(CFD_GAMMA) * fdv(pl, rl));
  T cr = fsq(T(CFD_GAMMA) * fdv(pr, rr));
  T SL = smin(ur - cr, ul - cl);
  T SR = smax(ul + cl, ur + cr);
  T Sstar = fdv((pr - pl) + (rl * ul * (SL - ul) - rr * ur * (SR - ur)), rl * (SL - ul) - rr * (SR - ur));
  if (SL >= 0) {
    res[0] = rl * ul; res[1] = rl * ul * ul + pl; res[2] = rl * HL * ul;
  } else if (Sstar >= 0) {
    T pStar = pl + rl * (SL - ul) * (Sstar - ul);
    T U[3] = {rl, rl * ul, fdv(pl, T(CFD_GAMMA) - T(1)) + rl * ul * ul / 2};
    T F[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
    T Dd[3] = {T(0.0), T(1.0), Sstar};
    for (int i = 0; i < 3; i++) res[i] = fdv(Sstar * (SL * U[i] - F[i]) + SL * pStar * Dd[i], SL - Sstar);
  } else if (SR >= 0) {
    T pStar = pr + rr * (SR - ur) * (Sstar - ur);
    T U[3] = {rr, rr * ur, fdv(pr, T(CFD_GAMMA) - T(1)) + rr * ur * ur / 2};
    T F[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
    T Dd[3] = {T(0.0), T(1.0), Sstar};
    for (int i = 0; i < 3; i++) res[i] = fdv(Sstar * (SR * U[i] - F[i]) + SR * pStar * Dd[i], SR - Sstar);
  } else {
    res[0] = rr * ur; res[1] = rr * ur * ur + pr; res[2] = rr * HR * ur;
  }
}

// roeFlux1D2 (fluxSchemes.hpp:13-71)
template <class T> __device__ __forceinline__ void roe1d(const T* L, const T* Rr, T* res) {
  T rl = L[0], ul = L[1], pl = L[2], rr = Rr[0], ur = Rr[1], pr = Rr[2];
  T HL = fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1)) + (ul * ul) / 2;
  T HR = fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1)) + (ur * ur) / 2;
  T rBar = fsq(rl * rr);
  T srl = fsq(rl), srr = fsq(rr);
  T uBar = fdv(ul * srl + ur * srr, srl + srr);
  T HBar = fdv(HL * srl + HR * srr, srl + srr);
  T cBar = fsq((T(CFD_GAMMA) - T(1)) * (HBar - uBar * uBar / 2));
  T cBar2 = (T(CFD_GAMMA) - T(1)) * (HBar - uBar * uBar / 2);
  T dr = rr - rl, du = ur - ul, dp = pr - pl;
  T lam[3] = {fabs(uBar), fabs(uBar + cBar), fabs(uBar - cBar)};
  T eps = T(0.1) * (fabs(uBar) + cBar);
  for (int i = 0; i < 3; i++)
    if (lam[i] < eps) lam[i] = fdv(lam[i] * lam[i] + eps * eps, T(2.0) * eps);
  T a0 = lam[0] * (dr - fdv(dp, cBar2));
  T a1 = half_quot(lam[1], cBar2) * (dp + rBar * cBar * du);
  T a2 = half_quot(lam[2], cBar2) * (dp - rBar * cBar * du);
  T a3 = a0 + a1 + a2;
  T a4 = cBar * (a1 - a2);
  T FL[3] = {rl * ul, rl * ul * ul + pl, rl * HL * ul};
  T FR[3] = {rr * ur, rr * ur * ur + pr, rr * HR * ur};
  for (int i = 0; i < 3; i++) res[i] = (FL[i] + FR[i]) / T(2.0);
  res[0] -= T(0.5) * a3;
  res[1] -= T(0.5) * (a3 * uBar + a4);
  res[2] -= T(0.5) * (HBar * a3 + uBar * a4 - fdv(cBar2 * a0, T(CFD_GAMMA) - T(1)));
}

// HLLCFlux2D2 (fluxSchemes.hpp:202-327), pressure-based wave speeds, with the
// averaged Sstar == 0 branch; 3D writes the star tangential momentum as
// Sstar*n_k + u_k*(sum of the other normal components).  A non-null hb
// receives the branches taken: 1 pressure clipped to 0, 2 left shock factor,
// 4 right shock factor, region << 3 (0 left state, 1 left star, 2 the
// averaged Sstar == 0 star, 3 right star, 4 right state).
template <int D, int AX, class T>
__device__ __forceinline__ void hllc_md(const T* Lp, const T* Rp, T* res, int* hb = nullptr) {
  constexpr int NV = D + 2, E = NV - 1;
  T rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  T HL, HR;
  if (D == 2) {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2]) / 2 + fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2]) / 2 + fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
  } else {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2] + Lp[3] * Lp[3]) / 2 + fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2] + Rp[3] * Rp[3]) / 2 + fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
  }
  T VnL = Lp[1 + AX], VnR = Rp[1 + AX];
  T cl = fsq(fdv(pl, rl) * T(CFD_GAMMA));
  T cr = fsq(fdv(pr, rr) * T(CFD_GAMMA));
  T rBar = (rl + rr) / 2, cBar = (cl + cr) / 2;
  const T psr = (pl + pr) / 2 - rBar * cBar * (VnR - VnL) / 2;
  T ps = smax(T(0.0), psr);
  T ql = (ps <= pl) ? T(1.0) : fsq(T(1.0) + (T(CFD_GAMMA) + T(1.0)) / (T(2.0) * T(CFD_GAMMA)) * (fdv(ps, pl) - T(1.0)));
  T qr = (ps <= pr) ? T(1.0) : fsq(T(1.0) + (T(CFD_GAMMA) + T(1.0)) / (T(2.0) * T(CFD_GAMMA)) * (fdv(ps, pr) - T(1.0)));
  T SL = VnL - cl * ql, SR = VnR + cr * qr;
  T Sstar = fdv((pr - pl) + (rl * VnL * (SL - VnL) - rr * VnR * (SR - VnR)), rl * (SL - VnL) - rr * (SR - VnR));
  if (hb) {
    const int reg = SL > 0 ? 0 : Sstar > 0 ? 1 : Sstar == 0 ? 2 : SR > 0 ? 3 : 4;
    *hb = (T(0.0) < psr ? 0 : 1) | (ps <= pl ? 0 : 2) | (ps <= pr ? 0 : 4) | (reg << 3);
  }

  // products with the normal's zero components left out (zeros added to
  // nonzero terms: the same values up to the sign of an exact zero)
  auto side = [&](const T* P, T r, T p, T H, T Vn, T* F) {
    F[0] = r * Vn;
    for (int k = 0; k < D; k++) F[1 + k] = k == AX ? r * P[1 + k] * Vn + p : r * P[1 + k] * Vn;
    F[E] = r * H * Vn;
  };
  auto star = [&](const T* P, T r, T p, T H, T Vn, T S, T* out) {
    T F[NV];
    side(P, r, p, H, Vn, F);
    T Usf = fdv(S - Vn, S - Sstar);
    T UE = r * H - p;
    out[0] = F[0] + S * (Usf * r - r);
    for (int k = 0; k < D; k++) {
      // Sstar n_k + u_k (sum of the other normal components): Sstar on the
      // normal axis, u_k on the others
      T Us = Usf * (k == AX ? Sstar : P[1 + k]) * r;
      out[1 + k] = F[1 + k] + S * (Us - r * P[1 + k]);
    }
    T UsE = Usf * (UE + (Sstar - Vn) * (r * Sstar + fdv(p, S - Vn)));
    out[E] = F[E] + S * (UsE - UE);
  };
  if (SL > 0) {
    side(Lp, rl, pl, HL, VnL, res);
  } else if (Sstar > 0) {
    star(Lp, rl, pl, HL, VnL, SL, res);
  } else if (Sstar == 0) {
    T a[NV], b[NV];
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
template <int D, int AX, class T> __device__ __forceinline__ void roe_md(const T* Lp, const T* Rp, T* res) {
  constexpr int NV = D + 2, E = NV - 1;
  T rl = Lp[0], rr = Rp[0], pl = Lp[E], pr = Rp[E];
  T HL, HR;
  if (D == 2) {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2]) / 2 + fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2]) / 2 + fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
  } else {
    HL = (Lp[1] * Lp[1] + Lp[2] * Lp[2] + Lp[3] * Lp[3]) / 2 + fdv(fdv(pl, rl) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
    HR = (Rp[1] * Rp[1] + Rp[2] * Rp[2] + Rp[3] * Rp[3]) / 2 + fdv(fdv(pr, rr) * T(CFD_GAMMA), T(CFD_GAMMA) - T(1));
  }
  T VnL = Lp[1 + AX], VnR = Rp[1 + AX];
  T c1 = fsq(rl), c2 = fsq(rr);
  T divisor = fdv(T(1.0), c1 + c2);
  T rhoAvg = fsq(rl * rr);
  T va[3];
  for (int k = 0; k < D; k++) va[k] = (c1 * Lp[1 + k] + c2 * Rp[1 + k]) * divisor;
  T HAvg = (c1 * HL + c2 * HR) * divisor;
  T q2 = (D == 2) ? (va[0] * va[0] + va[1] * va[1]) / 2 : (va[0] * va[0] + va[1] * va[1] + va[2] * va[2]) / 2;
  T cAvg = fsq((T(CFD_GAMMA) - T(1)) * (HAvg - q2));
  T VnAvg = va[AX];
  T lam0 = fabs(VnAvg - cAvg), lam1 = fabs(VnAvg), lam2 = fabs(VnAvg + cAvg);
  T eps = T(0.05) * (fabs(VnAvg) + cAvg);
  if (lam0 < eps) lam0 = fdv(lam0 * lam0 + eps * eps, T(2.0) * eps);
  if (lam1 < eps) lam1 = fdv(lam1 * lam1 + eps * eps, T(2.0) * eps);
  if (lam2 < eps) lam2 = fdv(lam2 * lam2 + eps * eps, T(2.0) * eps);
  T dP = pr - pl, dVn = VnR - VnL, dRho = rr - rl;
  T cf1 = half_quot(dP - rhoAvg * cAvg * dVn, cAvg * cAvg);
  T cf = half_quot(dP + rhoAvg * cAvg * dVn, cAvg * cAvg);
  T cf2 = dRho - fdv(dP, cAvg * cAvg);
  T shear = T(0.0);
  bool first = true;
  for (int k = 0; k < D; k++) {
    if (k == AX) continue;
    T t = va[k] * (Rp[1 + k] - Lp[1 + k]);
    shear = first ? t : shear + t;
    first = false;
  }
  // The products with the constant normal's zero components (c n_k, p n_k,
  // rhoAvg * 0 for the normal momentum) only add zeros to nonzero terms and
  // are left out: the same values, up to the sign of an exact-zero result.
  // component 0
  {
    T fd = lam0 * (cf1 * 1);
    fd += lam2 * (cf * 1);
    fd += lam1 * cf2;
    res[0] = (rl * VnL + rr * VnR - fd) / 2;
  }
  for (int k = 0; k < D; k++) {
    T fd = lam0 * (cf1 * (k == AX ? va[k] - cAvg : va[k]));
    fd += lam2 * (cf * (k == AX ? va[k] + cAvg : va[k]));
    fd += lam1 * (k == AX ? cf2 * va[k] : cf2 * va[k] + rhoAvg * (Rp[1 + k] - Lp[1 + k]));
    T fl = k == AX ? rl * Lp[1 + k] * VnL + pl : rl * Lp[1 + k] * VnL;
    T fr = k == AX ? rr * Rp[1 + k] * VnR + pr : rr * Rp[1 + k] * VnR;
    res[1 + k] = (fl + fr - fd) / 2;
  }
  {
    T fd = lam0 * (cf1 * (HAvg - cAvg * VnAvg));
    fd += lam2 * (cf * (HAvg + cAvg * VnAvg));
    fd += lam1 * (cf2 * q2 + rhoAvg * shear);
    res[E] = (rl * HL * VnL + rr * HR * VnR - fd) / 2;
  }
}

// Physical flux at a cell (fluxSolveEuler1D / 2D, fluxSchemes.hpp:572-599).
template <int D, int AX, class T> __device__ __forceinline__ void node_flux(const T* p, T* F) {
  if (D == 1) {
    T r = p[0], u = p[1], pp = p[2];
    F[0] = r * u;
    F[1] = r * u * u + pp;
    F[2] = (T(CFD_GAMMA) / (T(CFD_GAMMA) - T(1)) * pp + r * (u * u) / 2) * u;
    return;
  }
  constexpr int NV = D + 2, E = NV - 1;
  T r = p[0], pp = p[E];
  // Vn = sum p_k n_k and n_k pp with the constant normal's zero products
  // left out (the same values up to the sign of an exact zero)
  T k2;
  const T Vn = p[1 + AX];
  if (D == 2) {
    k2 = (p[1] * p[1] + p[2] * p[2]);
  } else {
    k2 = (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
  }
  F[0] = r * Vn;
  for (int k = 0; k < D; k++) F[1 + k] = k == AX ? r * p[1 + k] * Vn + pp : r * p[1 + k] * Vn;
  F[E] = (k2 / 2 * r + T(CFD_GAMMA) / (T(CFD_GAMMA) - T(1)) * pp) * Vn;
}

}  // namespace cfd
