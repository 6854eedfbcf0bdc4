// Forward-mode dual numbers for the oracle (TEST INFRASTRUCTURE ONLY).
//
// oracle.hpp is templated on its scalar type; instantiated with Dual it
// carries one directional derivative beside every value.  The value part
// performs exactly the double operations of the plain oracle (same order,
// same rounding), and every branch (std::min/max picks, HLLC wave pattern,
// positivity fallback, shock factors) is decided on the value, so the
// derivative part is the exact directional derivative of the active branch:
// the quantity a hand-written reverse-mode adjoint of the same map must
// reproduce through <adjoint, direction> = <cotangent, tangent>.
#pragma once
#include <cmath>

namespace oracle {

struct Dual {
  double v, d;
  Dual() : v(0.0), d(0.0) {}
  Dual(double x) : v(x), d(0.0) {}  // NOLINT: implicit, constants mix with duals
  Dual(double x, double dx) : v(x), d(dx) {}
  explicit operator double() const { return v; }
  Dual& operator+=(const Dual& b) { d += b.d; v += b.v; return *this; }
  Dual& operator-=(const Dual& b) { d -= b.d; v -= b.v; return *this; }
  Dual& operator*=(const Dual& b) { *this = Dual(v * b.v, d * b.v + v * b.d); return *this; }
  Dual& operator/=(const Dual& b) { *this = Dual(v / b.v, (d * b.v - v * b.d) / (b.v * b.v)); return *this; }
};

inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

inline Dual sqrt(const Dual& a) {
  const double s = std::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
inline Dual abs(const Dual& a) { return a.v < 0 ? -a : a; }
inline Dual fabs(const Dual& a) { return abs(a); }
inline double value_of(const Dual& a) { return a.v; }
// a^n for a constant exponent (libm pow on the value)
inline Dual pow(const Dual& a, double n) {
  const double p = std::pow(a.v, n);
  return Dual(p, n * std::pow(a.v, n - 1.0) * a.d);
}

}  // namespace oracle
