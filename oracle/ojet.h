// First-order forward-mode scalar for the oracle: a value with its gradient over up to 12
// inputs (the coordinates of a contact stencil's, hinge's or strand's vertices). Evaluating the
// oracle's templated gradient functions (distance gradients, dihedral and strand gradients) on
// it gives their exact Jacobians inside each branch region (the branch is chosen on the value,
// as the function chooses it), i.e. the second derivatives of the distances and angles.
// TEST INFRASTRUCTURE ONLY: used by the oracle's exact-Jacobian assembly, never by the HIP path.
#pragma once
// Included by omath.h ahead of the templates, so that their qualified std:: calls see the
// overloads at the end of this file.
#include <cmath>
#include <limits>

namespace oracle {

struct Jet {
  static constexpr int N = 12;
  double v = 0.0;
  double g[N] = {};
  Jet() = default;
  Jet(double x) : v(x) {}  // a constant (the templates write T(0), T(1e-12), ...)
  static Jet var(double x, int i) {
    Jet j(x);
    j.g[i] = 1.0;
    return j;
  }
};

// f(a) with f' = d1 at a.v
inline Jet chain1(const Jet &a, double f, double d1) {
  Jet r(f);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = d1 * a.g[i];
  return r;
}

inline Jet operator+(const Jet &a, const Jet &b) {
  Jet r(a.v + b.v);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = a.g[i] + b.g[i];
  return r;
}
inline Jet operator-(const Jet &a) {
  Jet r(-a.v);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = -a.g[i];
  return r;
}
inline Jet operator-(const Jet &a, const Jet &b) {
  Jet r(a.v - b.v);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = a.g[i] - b.g[i];
  return r;
}
inline Jet operator*(const Jet &a, const Jet &b) {
  Jet r(a.v * b.v);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = a.g[i] * b.v + b.g[i] * a.v;
  return r;
}
inline Jet operator/(const Jet &a, const Jet &b) {
  const double q = a.v / b.v;
  Jet r(q);
  for (int i = 0; i < Jet::N; ++i) r.g[i] = (a.g[i] - q * b.g[i]) / b.v;
  return r;
}
inline Jet &operator+=(Jet &a, const Jet &b) { return a = a + b; }
inline Jet &operator-=(Jet &a, const Jet &b) { return a = a - b; }
inline Jet &operator*=(Jet &a, const Jet &b) { return a = a * b; }
inline Jet &operator/=(Jet &a, const Jet &b) { return a = a / b; }
inline bool operator<(const Jet &a, const Jet &b) { return a.v < b.v; }
inline bool operator<=(const Jet &a, const Jet &b) { return a.v <= b.v; }
inline bool operator>(const Jet &a, const Jet &b) { return a.v > b.v; }
inline bool operator>=(const Jet &a, const Jet &b) { return a.v >= b.v; }
inline bool operator==(const Jet &a, const Jet &b) { return a.v == b.v; }
inline bool operator!=(const Jet &a, const Jet &b) { return a.v != b.v; }

inline Jet jet_sqrt(const Jet &a) {
  const double s = std::sqrt(a.v);
  return chain1(a, s, s > 0.0 ? 0.5 / s : 0.0);
}
// atan2(y, x): d/dy = x / r2, d/dx = -y / r2
inline Jet jet_atan2(const Jet &y, const Jet &x) {
  const double r2 = x.v * x.v + y.v * y.v;
  const double fy = x.v / r2, fx = -y.v / r2;
  Jet r(std::atan2(y.v, x.v));
  for (int i = 0; i < Jet::N; ++i) r.g[i] = fy * y.g[i] + fx * x.g[i];
  return r;
}

}  // namespace oracle

// The templates call std::sqrt / std::atan2 / std::fmax / std::fmin / std::fabs / std::isfinite
// and std::numeric_limits<T>::infinity on their scalar.
namespace std {
inline oracle::Jet sqrt(const oracle::Jet &a) { return oracle::jet_sqrt(a); }
inline oracle::Jet atan2(const oracle::Jet &y, const oracle::Jet &x) { return oracle::jet_atan2(y, x); }
inline oracle::Jet fmax(const oracle::Jet &a, const oracle::Jet &b) { return a.v >= b.v ? a : b; }
inline oracle::Jet fmin(const oracle::Jet &a, const oracle::Jet &b) { return a.v <= b.v ? a : b; }
inline oracle::Jet fabs(const oracle::Jet &a) { return a.v < 0.0 ? -a : a; }
inline bool isfinite(const oracle::Jet &a) { return std::isfinite(a.v); }
template <> class numeric_limits<oracle::Jet> {
 public:
  static oracle::Jet infinity() { return oracle::Jet(std::numeric_limits<double>::infinity()); }
};
}  // namespace std
