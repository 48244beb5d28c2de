// Forward-mode dual numbers for the oracle's tangent (test infrastructure only).
//
// Dual{v, d}: value v and directional derivative d.  The elementary derivatives are
// the textbook ones; d/dx erf(x) = 2/sqrt(pi) exp(-x^2), the same closed form the
// reference's Tapenade adjoint was hand-patched to use (split_fluxes_b.F90:82-83).
#pragma once
#include <cmath>

namespace orc {

struct Dual {
  double v = 0.0, d = 0.0;
  Dual() = default;
  Dual(double v_) : v(v_), d(0.0) {}
  Dual(double v_, double d_) : v(v_), d(d_) {}
};

inline double val(const Dual& a) { return a.v; }

inline Dual operator+(const Dual& a, const Dual& b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(const Dual& a, const Dual& b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(const Dual& a, const Dual& b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(const Dual& a, const Dual& b) {
  double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline Dual operator-(const Dual& a) { return {-a.v, -a.d}; }
inline Dual operator+(const Dual& a, double b) { return {a.v + b, a.d}; }
inline Dual operator+(double a, const Dual& b) { return {a + b.v, b.d}; }
inline Dual operator-(const Dual& a, double b) { return {a.v - b, a.d}; }
inline Dual operator-(double a, const Dual& b) { return {a - b.v, -b.d}; }
inline Dual operator*(const Dual& a, double b) { return {a.v * b, a.d * b}; }
inline Dual operator*(double a, const Dual& b) { return {a * b.v, a * b.d}; }
inline Dual operator/(const Dual& a, double b) { return {a.v / b, a.d / b}; }
inline Dual operator/(double a, const Dual& b) {
  double q = a / b.v;
  return {q, -q * b.d / b.v};
}

inline Dual exp(const Dual& a) {
  double e = std::exp(a.v);
  return {e, e * a.d};
}
inline Dual log(const Dual& a) { return {std::log(a.v), a.d / a.v}; }
inline Dual log10(const Dual& a) { return {std::log10(a.v), a.d / (a.v * std::log(10.0))}; }
inline Dual sqrt(const Dual& a) {
  double s = std::sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
inline Dual erf(const Dual& a) {
  const double two_by_sqrt_pi = 2.0 / std::sqrt(4.0 * std::atan(1.0));
  return {std::erf(a.v), two_by_sqrt_pi * std::exp(-a.v * a.v) * a.d};
}
inline Dual pow(const Dual& a, double p) {
  if (p == 0.0) return {std::pow(a.v, p), 0.0};
  return {std::pow(a.v, p), p * std::pow(a.v, p - 1.0) * a.d};
}

}  // namespace orc
