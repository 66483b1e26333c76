// Oracle math kernel: a scalar Vec3 / Vec2 / Mat2 with the evaluation order of the
// vector library the reference engine is written against (glam 0.30, scalar path):
//   dot   = (x*x' + y*y') + z*z'
//   cross = (y z' - y' z, z x' - z' x, x y' - x' y)
//   length = sqrt(dot(v, v)); normalize = v * (1 / length)  (multiply by the reciprocal)
// TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP path, never the
// product path. Templated on the scalar so the same text gives the FP32 mirror of the
// reference (T = float) and the FP64 twin (T = double).
#pragma once
#include <cmath>
#include <cstdint>

#include "ojet.h"

namespace oracle {

template <class T> struct V3 {
  T x{0}, y{0}, z{0};
  V3() = default;
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline V3<T> &operator+=(V3<T> &a, V3<T> b) { a = a + b; return a; }
template <class T> inline V3<T> &operator-=(V3<T> &a, V3<T> b) { a = a - b; return a; }
template <class T> inline bool operator==(V3<T> a, V3<T> b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T length_sq(V3<T> a) { return dot(a, a); }
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
template <class T> inline V3<T> splat(T s) { return {s, s, s}; }
template <class T> inline bool finite3(V3<T> a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
template <class T> inline V3<T> normalize_or(V3<T> a, V3<T> fb) {
  T rcp = T(1) / length(a);
  if (std::isfinite(rcp) && rcp > T(0)) return a * rcp;
  return fb;
}
template <class T> inline V3<T> normalize_or_zero(V3<T> a) { return normalize_or(a, V3<T>{}); }
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / length(a)); }
template <class T> inline T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class T> struct V2 {
  T x{0}, y{0};
  V2() = default;
  V2(T a, T b) : x(a), y(b) {}
};
template <class T> inline T dot2(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline V2<T> normalize2_or(V2<T> a, V2<T> fb) {
  T rcp = T(1) / std::sqrt(dot2(a, a));
  if (std::isfinite(rcp) && rcp > T(0)) return {a.x * rcp, a.y * rcp};
  return fb;
}

// Column-major 2x2 (x_axis = first column, y_axis = second column).
template <class T> struct M2 {
  V2<T> xa, ya;
  M2() = default;
  M2(V2<T> c0, V2<T> c1) : xa(c0), ya(c1) {}
  T det() const { return xa.x * ya.y - xa.y * ya.x; }
  M2 inverse() const {
    T inv_det = T(1) / det();
    return M2(V2<T>(ya.y * inv_det, -xa.y * inv_det), V2<T>(-ya.x * inv_det, xa.x * inv_det));
  }
  M2 transpose() const { return M2(V2<T>(xa.x, ya.x), V2<T>(xa.y, ya.y)); }
  // element (row r, col c)
  T at(int r, int c) const { return c == 0 ? (r == 0 ? xa.x : xa.y) : (r == 0 ? ya.x : ya.y); }
};

// 3x2 matrix stored as two 3D columns.
template <class T> struct M32 {
  V3<T> c0, c1;
  M32() = default;
  M32(V3<T> a, V3<T> b) : c0(a), c1(b) {}
  M32 mul2(const M2<T> &m) const {
    return M32(c0 * m.xa.x + c1 * m.xa.y, c0 * m.ya.x + c1 * m.ya.y);
  }
  // row-major vec: v[2r+c] = M[r][c]
  void vec6(T out[6]) const {
    out[0] = c0.x; out[1] = c1.x; out[2] = c0.y; out[3] = c1.y; out[4] = c0.z; out[5] = c1.z;
  }
  M2<T> ata() const { return M2<T>(V2<T>(dot(c0, c0), dot(c1, c0)), V2<T>(dot(c0, c1), dot(c1, c1))); }
};

}  // namespace oracle
