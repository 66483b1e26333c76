// Device math of the implicit step: a 3-vector with the reference's evaluation order
// (dot = (x x' + y y') + z z', normalize = v * (1/|v|)), contact primitives (closest point
// on a triangle with barycentrics, the edge-edge witness, distances with analytic
// gradients, the edge-triangle pierce test, log / cubic barriers, lagged friction,
// Rayleigh contact stiffness, additive CCD) and the element models (isotropic shell on
// principal stretches, hinge bending, rod stretch, strand bend).
//
// Every function states the reference function it follows (paths under
// core-engine/crates/engine-core/src/). The library is compiled with -ffp-contract=off so
// distances, activation tests and Morton keys round exactly as the FP32 reference does.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nguo {

template <class T> struct v3 {
  T x, y, z;
};
using f3 = v3<float>;
using d3 = v3<double>;

template <class T> __host__ __device__ inline v3<T> mk(T x, T y, T z) { return v3<T>{x, y, z}; }
template <class T> __host__ __device__ inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __host__ __device__ inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __host__ __device__ inline v3<T> operator-(v3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __host__ __device__ inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __host__ __device__ inline v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> __host__ __device__ inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> __host__ __device__ inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __host__ __device__ inline v3<T> cross(v3<T> a, v3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> __host__ __device__ inline T len(v3<T> a) { return sqrt(dot(a, a)); }
template <class T> __host__ __device__ inline T comp(v3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class T> __host__ __device__ inline void setc(v3<T> &a, int i, T v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}
__host__ __device__ inline float mn(float a, float b) { return fminf(a, b); }
__host__ __device__ inline double mn(double a, double b) { return fmin(a, b); }
__host__ __device__ inline float mx(float a, float b) { return fmaxf(a, b); }
__host__ __device__ inline double mx(double a, double b) { return fmax(a, b); }
template <class T> __host__ __device__ inline v3<T> vmin(v3<T> a, v3<T> b) { return {mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z)}; }
template <class T> __host__ __device__ inline v3<T> vmax(v3<T> a, v3<T> b) { return {mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z)}; }
template <class T> __host__ __device__ inline bool isfin(T v) { return v == v && v - v == v - v; }
template <class T> __host__ __device__ inline v3<T> normalize_or(v3<T> a, v3<T> fb) {
  T rcp = T(1) / len(a);
  if (isfin(rcp) && rcp > T(0)) return a * rcp;
  return fb;
}
template <class T> __host__ __device__ inline v3<T> normalize_or_zero(v3<T> a) { return normalize_or(a, v3<T>{0, 0, 0}); }
template <class T> __host__ __device__ inline T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class T> __host__ __device__ inline d3 to_d(v3<T> a) { return {(double)a.x, (double)a.y, (double)a.z}; }

__device__ inline f3 ld3(const float *p, uint32_t i) { return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]}; }
__device__ inline void st3(float *p, uint32_t i, f3 v) { p[3 * (size_t)i] = v.x; p[3 * (size_t)i + 1] = v.y; p[3 * (size_t)i + 2] = v.z; }

// ------------------------------------------------------------------ contact primitives
// contact.rs:84-131 closest_point_on_triangle_bary.
template <class T> __device__ inline v3<T> closest_point_bary(v3<T> p, v3<T> a, v3<T> b, v3<T> c, T w[3]) {
  v3<T> ab = b - a, ac = c - a, ap = p - a;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) { w[0] = 1; w[1] = 0; w[2] = 0; return a; }
  v3<T> bp = p - b;
  T d3v = dot(ab, bp), d4 = dot(ac, bp);
  if (d3v >= T(0) && d4 <= d3v) { w[0] = 0; w[1] = 1; w[2] = 0; return b; }
  T vc = d1 * d4 - d3v * d2;
  if (vc <= T(0) && d1 >= T(0) && d3v <= T(0)) {
    T den = d1 - d3v;
    T v = den != T(0) ? d1 / den : T(0);
    w[0] = T(1) - v; w[1] = v; w[2] = 0;
    return a + ab * v;
  }
  v3<T> cp = p - c;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) { w[0] = 0; w[1] = 0; w[2] = 1; return c; }
  T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    T den = d2 - d6;
    T ww = den != T(0) ? d2 / den : T(0);
    w[0] = T(1) - ww; w[1] = 0; w[2] = ww;
    return a + ac * ww;
  }
  T va = d3v * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3v) >= T(0) && (d5 - d6) >= T(0)) {
    T den = (d4 - d3v) + (d5 - d6);
    T ww = den != T(0) ? (d4 - d3v) / den : T(0);
    w[0] = 0; w[1] = T(1) - ww; w[2] = ww;
    return b + (c - b) * ww;
  }
  T den = va + vb + vc;
  if (den == T(0)) { w[0] = 1; w[1] = 0; w[2] = 0; return a; }
  T denom = T(1) / den;
  T v = vb * denom, ww = vc * denom;
  w[0] = T(1) - v - ww; w[1] = v; w[2] = ww;
  return a + ab * v + ac * ww;
}

template <class T> __device__ inline T point_triangle_distance(v3<T> p, v3<T> a, v3<T> b, v3<T> c) {
  T w[3];
  return len(p - closest_point_bary(p, a, b, c, w));
}

// contact.rs:836-858 point_tri_distance_grad (FP32 on the step's path; FP64 in the energy
// evaluation).
template <class T> __device__ inline T point_tri_distance_grad(v3<T> p, v3<T> a, v3<T> b, v3<T> c, v3<T> g[4]) {
  T w[3];
  v3<T> q = closest_point_bary(p, a, b, c, w);
  v3<T> e = p - q;
  T d = len(e);
  if (d < T(1.0e-12)) { for (int i = 0; i < 4; ++i) g[i] = v3<T>{0, 0, 0}; return d; }
  v3<T> n = e / d;
  g[0] = n; g[1] = (-w[0]) * n; g[2] = (-w[1]) * n; g[3] = (-w[2]) * n;
  return d;
}

// contact.rs:205-215 point_edge_coeff.
template <class T> __device__ inline void point_edge_coeff(v3<T> p, v3<T> e0, v3<T> e1, T out[2]) {
  v3<T> r = e1 - e0;
  T d = dot(r, r);
  if (d > T(0)) { T t = dot(r, p - e0) / d; out[0] = T(1) - t; out[1] = t; }
  else { out[0] = T(0.5); out[1] = T(0.5); }
}

// contact.rs:238-310 edge_edge_witness.
template <class T> __device__ inline void edge_edge_witness(v3<T> ea0, v3<T> ea1, v3<T> eb0, v3<T> eb1, T out[4]) {
  v3<T> cog = T(0.25) * (ea0 + ea1 + eb0 + eb1);
  v3<T> a0 = ea0 - cog, a1 = ea1 - cog, b0v = eb0 - cog, b1v = eb1 - cog;
  auto dist2 = [&](const T c[4]) {
    v3<T> x0 = c[0] * a0 + c[1] * a1;
    v3<T> x1 = c[2] * b0v + c[3] * b1v;
    v3<T> d = x1 - x0;
    return dot(d, d);
  };
  T pe[2];
  point_edge_coeff(ea0, eb0, eb1, pe);
  T cls[4] = {T(1), T(0), pe[0], pe[1]};
  v3<T> r0 = ea1 - ea0, r1 = eb1 - eb0;
  T a00 = dot(r0, r0), a01 = -dot(r0, r1), a11 = dot(r1, r1);
  v3<T> rhs = eb0 - ea0;
  T b0 = dot(r0, rhs), b1 = -dot(r1, rhs);
  T det = a00 * a11 - a01 * a01;
  if (det != T(0)) {
    T sp = (a11 * b0 - a01 * b1) / det;
    T tp = (a00 * b1 - a01 * b0) / det;
    T direct[4] = {T(1) - sp, sp, T(1) - tp, tp};
    if (dist2(direct) < dist2(cls))
      for (int i = 0; i < 4; ++i) cls[i] = direct[i];
  }
  bool inrange = true;
  for (int i = 0; i < 4; ++i) inrange = inrange && (cls[i] >= T(0) && cls[i] <= T(1));
  if (inrange) { for (int i = 0; i < 4; ++i) out[i] = cls[i]; return; }
  auto snap = [](T c[2]) {
    if (c[0] < T(0)) { c[0] = T(0); c[1] = T(1); }
    else if (c[0] > T(1)) { c[0] = T(1); c[1] = T(0); }
  };
  T c1[2], c2[2], c3[2], c4[2];
  point_edge_coeff(ea0, eb0, eb1, c1); snap(c1);
  point_edge_coeff(ea1, eb0, eb1, c2); snap(c2);
  point_edge_coeff(eb0, ea0, ea1, c3); snap(c3);
  point_edge_coeff(eb1, ea0, ea1, c4); snap(c4);
  T types[4][4] = {{T(1), T(0), c1[0], c1[1]}, {T(0), T(1), c2[0], c2[1]}, {c3[0], c3[1], T(1), T(0)}, {c4[0], c4[1], T(0), T(1)}};
  int best = 0;
  T best_d = T(INFINITY);
  for (int k = 0; k < 4; ++k) {
    T d = dist2(types[k]);
    if (d < best_d) { best_d = d; best = k; }
  }
  for (int i = 0; i < 4; ++i) out[i] = types[best][i];
}

// contact.rs:489-512 segment_segment_distance_grad.
template <class T> __device__ inline T segment_segment_distance_grad(v3<T> p1, v3<T> q1, v3<T> p2, v3<T> q2, v3<T> g[4]) {
  T w[4];
  edge_edge_witness(p1, q1, p2, q2, w);
  v3<T> ca = w[0] * p1 + w[1] * q1;
  v3<T> cb = w[2] * p2 + w[3] * q2;
  v3<T> e = ca - cb;
  T d = len(e);
  if (d < T(1.0e-12)) { for (int i = 0; i < 4; ++i) g[i] = v3<T>{0, 0, 0}; return d; }
  v3<T> n = e / d;
  g[0] = w[0] * n; g[1] = w[1] * n; g[2] = (-w[2]) * n; g[3] = (-w[3]) * n;
  return d;
}

// contact.rs:439-484 segment_segment_distance (clamped pair, for CCD).
template <class T> __device__ inline T segment_segment_distance(v3<T> p1, v3<T> q1, v3<T> p2, v3<T> q2) {
  v3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const T EPS = T(1.0e-12);
  T s, t;
  if (a <= EPS && e <= EPS) return len(p1 - p2);
  if (a <= EPS) { s = 0; t = clampv(f / e, T(0), T(1)); }
  else {
    T c = dot(d1, r);
    if (e <= EPS) { t = 0; s = clampv(-c / a, T(0), T(1)); }
    else {
      T b = dot(d1, d2);
      T denom = a * e - b * b;
      T s0 = denom > EPS ? clampv((b * f - c * e) / denom, T(0), T(1)) : T(0);
      T t0 = (b * s0 + f) / e;
      if (t0 < T(0)) { t = 0; s = clampv(-c / a, T(0), T(1)); }
      else if (t0 > T(1)) { t = 1; s = clampv((b - c) / a, T(0), T(1)); }
      else { s = s0; t = t0; }
    }
  }
  v3<T> c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  return len(c1 - c2);
}

// contact.rs:325-352 edge_triangle_intersect.
__device__ inline bool edge_triangle_intersect(f3 e0, f3 e1, f3 v0, f3 v1, f3 v2) {
  f3 d1 = v1 - v0, d2 = v2 - v0, a0 = e0 - v0, a1 = e1 - v0;
  f3 n = cross(d1, d2);
  float s1 = dot(a0, n), s2 = dot(a1, n);
  if (!(s1 * s2 < 0.0f)) return false;
  float t = s1 / (s1 - s2);
  f3 r = a0 * (1.0f - t) + a1 * t;
  float a = dot(d1, d1), b = dot(d1, d2), c = dot(d2, d2);
  float det = a * c - b * b;
  if (det == 0.0f) return false;
  float p = dot(d1, r), q = dot(d2, r);
  float u = (c * p - b * q) / det;
  float v = (a * q - b * p) / det;
  return u >= 0.0f && v >= 0.0f && u + v <= 1.0f;
}

// contact.rs:1014-1025 log barrier / 580-591 cubic barrier, evaluated in FP64 from the FP32
// gap; out = (energy, dE/dd, d2E/dd2), false when inactive (d >= dhat).
__device__ inline bool barrier_of(int kind, double d, double dhat, double kappa, double out[3]) {
  if (kind == 0) {
    if (d >= dhat) return false;
    d = fmax(d, 1.0e-6);
    double t = d - dhat, ld = log(d / dhat);
    out[0] = -kappa * t * t * ld;
    out[1] = -kappa * (2.0 * t * ld + t * t / d);
    out[2] = -kappa * (2.0 * ld + 4.0 * t / d - (t * t) / (d * d));
    return true;
  }
  double y = d - dhat;
  if (y >= 0.0) return false;
  out[0] = kappa * (-2.0 * (y * y * y) / (3.0 * dhat));
  out[1] = kappa * (-2.0 * y * y / dhat);
  out[2] = kappa * (4.0 * (1.0 - d / dhat));
  return true;
}

// contact.rs:378-431 accd_toi in FP64. kind 0: point (x0[0]) vs triangle, 1: edge vs edge.
// *capped (may be null) is set when the iteration limit ended the advancement.
__device__ inline double accd_toi(const d3 x0[4], const d3 dx[4], int kind, double offset, double max_t, double reduction, uint32_t max_iter,
                                  bool *capped = nullptr) {
  double u_max = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) u_max = fmax(u_max, len(dx[i] - dx[j]));
  if (u_max < 1.0e-12) return max_t;
  double inv_u_max = 1.0 / u_max;
  auto sqd = [&](const d3 *p) {
    double d = kind == 0 ? point_triangle_distance(p[0], p[1], p[2], p[3]) : segment_segment_distance(p[0], p[1], p[2], p[3]);
    return d * d;
  };
  double d0 = sqrt(fmax(sqd(x0), 0.0));
  double eps = reduction * (d0 - offset);
  if (eps <= 0.0) return max_t;
  double target = eps + offset, eps_sqr = eps * eps, toi = 0;
  d3 moved[4];
  uint32_t it = 0;
  for (; it < max_iter; ++it) {
    for (int k = 0; k < 4; ++k) moved[k] = x0[k] + dx[k] * toi;
    double d2 = fmax(sqd(moved), 0.0);
    double dmt = (d2 - target * target) / (sqrt(d2) + target);
    if ((max_t - toi) * u_max < dmt - eps) return max_t;
    if (toi > 0.0 && dmt * dmt < eps_sqr) break;
    double tn = toi + dmt * inv_u_max;
    if (tn != toi) toi = tn; else break;
    if (toi > max_t) return max_t;
  }
  if (capped && it == max_iter) *capped = true;
  return fmax(toi, 0.0);
}

// contact.rs:922-929 surface_normal_at.
__device__ inline f3 surface_normal_at(const float bary[3], f3 face, f3 pn0, f3 pn1, f3 pn2) {
  if (bary[0] > 1.0e-4f && bary[1] > 1.0e-4f && bary[2] > 1.0e-4f) return face;
  return normalize_or_zero(pn0 * bary[0] + pn1 * bary[1] + pn2 * bary[2]);
}

// ------------------------------------------------------------------ elements
// shell_model.rs:59-78 sym_eig2 for [[p, q], [q, r]]: eigenvalues descending, eigenvector
// columns (v0, v1).
__device__ inline void sym_eig2(float p, float q, float r, float lam[2], float v0[2], float v1[2]) {
  float tr = p + r;
  float diff = sqrtf(fmaxf((p - r) * (p - r) + 4.0f * q * q, 0.0f));
  lam[0] = 0.5f * (tr + diff);
  lam[1] = 0.5f * (tr - diff);
  if (fabsf(q) > 1.0e-20f) {
    float ax = q, ay = lam[0] - p;
    float rcp = 1.0f / sqrtf(ax * ax + ay * ay);
    if (isfin(rcp) && rcp > 0.0f) { v0[0] = ax * rcp; v0[1] = ay * rcp; }
    else { v0[0] = 1.0f; v0[1] = 0.0f; }
  } else if (p >= r) { v0[0] = 1.0f; v0[1] = 0.0f; }
  else { v0[0] = 0.0f; v0[1] = 1.0f; }
  v1[0] = -v0[1];
  v1[1] = v0[0];
}

// Column-major 2x2 inverse rest shape: m = [m00 m01; m10 m11] stored (m00, m10, m01, m11).
struct Mat2 {
  float a, b, c, d;  // a = m00, b = m10, c = m01, d = m11
};

// shell_model.rs:84-102 svd3x2 of F = [f0 f1] (columns).
__device__ inline void svd3x2(f3 f0, f3 f1, f3 &u0, f3 &u1, float sigma[2], float v0[2], float v1[2]) {
  float lam[2];
  sym_eig2(dot(f0, f0), dot(f0, f1), dot(f1, f1), lam, v0, v1);
  sigma[0] = sqrtf(fmaxf(lam[0], 0.0f));
  sigma[1] = sqrtf(fmaxf(lam[1], 0.0f));
  float eps = 1.0e-9f * fmaxf(sigma[0], 1.0f);
  f3 X{1, 0, 0}, Y{0, 1, 0};
  f3 fv0 = f0 * v0[0] + f1 * v0[1];
  u0 = normalize_or(sigma[0] > eps ? fv0 / sigma[0] : X, X);
  f3 u1r;
  if (sigma[1] > eps) u1r = (f0 * v1[0] + f1 * v1[1]) / sigma[1];
  else {
    f3 t = fabsf(u0.x) < 0.9f ? X : Y;
    u1r = t - dot(t, u0) * u0;
  }
  u1 = normalize_or(u1r - dot(u1r, u0) * u0, Y);
}

// shell_model.rs: StVK (model 0) / ARAP (model 1) derivative tables on the stretches.
__device__ inline void shell_derivs(const float s[2], int model, float mu, float lambda, float deda[2], float d2[4]) {
  float j = s[0] * s[1];
  float a = s[0], b = s[1];
  float dev0, dev1, h0, h1;
  if (model == 0) {
    float da = a - 1.0f, db = b - 1.0f;
    dev0 = mu * (da * da * da);
    dev1 = mu * (db * db * db);
    h0 = 3.0f * mu * (a - 1.0f) * (a - 1.0f);
    h1 = 3.0f * mu * (b - 1.0f) * (b - 1.0f);
  } else {
    dev0 = mu * (a - 1.0f);
    dev1 = mu * (b - 1.0f);
    h0 = mu;
    h1 = mu;
  }
  deda[0] = dev0 + lambda * s[1] * (j - 1.0f);
  deda[1] = dev1 + lambda * s[0] * (j - 1.0f);
  float crossv = lambda * (2.0f * a * b - 1.0f);
  d2[0] = h0 + lambda * b * b;
  d2[1] = crossv;
  d2[2] = crossv;
  d2[3] = h1 + lambda * a * a;
}

// shell_model.rs:367-436: gradient (3 vertices) and 9x9 PSD Hessian of one shell element.
// project = false keeps negative mode eigenvalues: the exact second derivative.
__device__ inline void shell_element(const f3 p[3], Mat2 dm, float area, int model, float mu, float lambda, float eps, f3 g[3], float *h81,
                                     bool project = true) {
  // F = [p1-p0, p2-p0] * Dm^-1
  f3 e1 = p[1] - p[0], e2 = p[2] - p[0];
  f3 f0 = e1 * dm.a + e2 * dm.b;
  f3 f1 = e1 * dm.c + e2 * dm.d;
  f3 u0, u1;
  float s[2], v0[2], v1[2];
  svd3x2(f0, f1, u0, u1, s, v0, v1);
  float deda[2], d2[4];
  shell_derivs(s, model, mu, lambda, deda, d2);
  // gradient: P = U diag(deda) V^T, H = P Dm^-T, forces on p1, p2; p0 takes minus their sum
  {
    f3 c0 = u0 * deda[0], c1 = u1 * deda[1];
    // (C V^T): V^T = [[v0x, v0y],[v1x, v1y]] as a column-major 2x2 (xa = (v0[0], v1[0]), ya = (v0[1], v1[1]))
    f3 pk0 = c0 * v0[0] + c1 * v1[0];
    f3 pk1 = c0 * v0[1] + c1 * v1[1];
    // times Dm^-T: Dm^-T columns xa = (a, c), ya = (b, d)
    f3 h0 = pk0 * dm.a + pk1 * dm.c;
    f3 h1 = pk0 * dm.b + pk1 * dm.d;
    f3 g1 = h0 * area, g2 = h1 * area;
    g[0] = -(g1 + g2);
    g[1] = g1;
    g[2] = g2;
  }
  // six eigenpairs (shell_model.rs:195-240)
  float ev[2], ev0[2], ev1[2];
  sym_eig2(d2[0], d2[1], d2[3], ev, ev0, ev1);
  const float R = 0.70710678118654752440f;
  float a = s[0], b = s[1];
  float denom = a - b;
  float fallback = 0.5f * (d2[0] + d2[3]) - 0.5f * (d2[1] + d2[2]);
  float lamk[6] = {(deda[0] + deda[1]) / (a + b), fabsf(denom) > eps ? (deda[0] - deda[1]) / denom : fallback,
                   deda[0] / a, deda[1] / b, ev[0], ev[1]};
  // mode matrices Q_k (3x2, rows x y z) given by their columns
  float q[6][6] = {
      // c0.x c0.y c0.z c1.x c1.y c1.z
      {0, -R, 0, R, 0, 0},   {0, R, 0, R, 0, 0},   {0, 0, 1, 0, 0, 0},
      {0, 0, 0, 0, 0, 1},    {ev0[0], 0, 0, 0, ev0[1], 0}, {ev1[0], 0, 0, 0, ev1[1], 0}};
  f3 u2 = cross(u0, u1);
  float a6[36];
  for (int i = 0; i < 36; ++i) a6[i] = 0.0f;
  for (int k = 0; k < 6; ++k) {
    float l = project ? fmaxf(lamk[k], 0.0f) : lamk[k];
    if (l == 0.0f) continue;
    f3 qc0 = u0 * q[k][0] + u1 * q[k][1] + u2 * q[k][2];
    f3 qc1 = u0 * q[k][3] + u1 * q[k][4] + u2 * q[k][5];
    // (U Q) V^T with V^T column-major xa = (v0[0], v1[0]), ya = (v0[1], v1[1])
    f3 m0 = qc0 * v0[0] + qc1 * v1[0];
    f3 m1 = qc0 * v0[1] + qc1 * v1[1];
    float qv[6] = {m0.x, m1.x, m0.y, m1.y, m0.z, m1.z};
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) a6[r * 6 + c] += l * qv[r] * qv[c];
  }
  // B = dF/dx (6x9): row 2i+j; columns 3(k+1)+i = Dm^-1[k][j], column i = -sum
  float dmk[2][2] = {{dm.a, dm.c}, {dm.b, dm.d}};  // dmk[k][j] = Dm^-1[k][j]
  float bm[54];
  for (int i = 0; i < 54; ++i) bm[i] = 0.0f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) {
      int row = 2 * i + j;
      float sum = 0.0f;
      for (int k = 0; k < 2; ++k) {
        float dv = dmk[k][j];
        bm[row * 9 + 3 * (k + 1) + i] = dv;
        sum += dv;
      }
      bm[row * 9 + i] = -sum;
    }
  float ab[54];
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 9; ++c) {
      float acc = 0.0f;
      for (int k = 0; k < 6; ++k) acc += a6[r * 6 + k] * bm[k * 9 + c];
      ab[r * 9 + c] = acc;
    }
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < 9; ++c) {
      float acc = 0.0f;
      for (int k = 0; k < 6; ++k) acc += bm[k * 9 + r] * ab[k * 9 + c];
      h81[r * 9 + c] = area * acc;
    }
}

// bending.rs:34-48 dihedral_angle (FP32 for the step; FP64 for the energy evaluation).
template <class T> __device__ inline T dihedral_angle(const v3<T> p[4]) {
  v3<T> n1 = cross(p[1] - p[0], p[2] - p[0]);
  v3<T> n2 = cross(p[1] - p[2], p[3] - p[2]);
  v3<T> e = normalize_or_zero(p[2] - p[1]);
  v3<T> n1n = normalize_or_zero(n1), n2n = normalize_or_zero(n2);
  T sn = dot(cross(n1n, n2n), e);
  T cs = clampv(dot(n1n, n2n), T(-1), T(1));
  if constexpr (sizeof(T) == 4) return atan2f(sn, cs);
  else return atan2(sn, cs);
}

// bending.rs:78-98 dihedral_grad_analytic (central differences on a degenerate hinge).
__device__ inline void dihedral_grad(const f3 p[4], f3 g[4]) {
  f3 m1 = cross(p[1] - p[0], p[2] - p[0]);
  f3 m2 = cross(p[1] - p[2], p[3] - p[2]);
  f3 e = p[2] - p[1];
  float a1 = len(m1), a2 = len(m2), l2 = dot(e, e);
  if (a1 < 1.0e-12f || a2 < 1.0e-12f || l2 < 1.0e-24f) {
    float scale = fmaxf(len(p[0]) + len(p[1]) + len(p[2]) + len(p[3]), 1.0f);
    float h = 1.0e-4f * scale, inv = 1.0f / (2.0f * h);
    for (int vi = 0; vi < 4; ++vi)
      for (int c = 0; c < 3; ++c) {
        f3 pp[4], pm[4];
        for (int k = 0; k < 4; ++k) { pp[k] = p[k]; pm[k] = p[k]; }
        setc(pp[vi], c, comp(pp[vi], c) + h);
        setc(pm[vi], c, comp(pm[vi], c) - h);
        setc(g[vi], c, (dihedral_angle(pp) - dihedral_angle(pm)) * inv);
      }
    return;
  }
  float l = sqrtf(l2);
  f3 n1 = m1 / a1, n2 = m2 / a2;
  f3 g0 = n1 * (-l / a1);
  f3 g3 = n2 * (-l / a2);
  float ta = dot(p[0] - p[1], e) / l2;
  float tb = dot(p[3] - p[1], e) / l2;
  g[0] = g0;
  g[1] = g0 * (ta - 1.0f) + g3 * (tb - 1.0f);
  g[2] = g0 * (-ta) + g3 * (-tb);
  g[3] = g3;
}

// The analytic dihedral gradient above in FP64 for the exact adjoint's Jacobian; a degenerate
// hinge (zero-area wing or zero-length edge) gets a zero gradient.
__device__ inline void dihedral_grad64(const d3 p[4], d3 g[4]) {
  const d3 m1 = cross(p[1] - p[0], p[2] - p[0]), m2 = cross(p[1] - p[2], p[3] - p[2]), e = p[2] - p[1];
  const double a1 = len(m1), a2 = len(m2), l2 = dot(e, e);
  if (a1 < 1.0e-12 || a2 < 1.0e-12 || l2 < 1.0e-24) {
    for (int k = 0; k < 4; ++k) g[k] = d3{0, 0, 0};
    return;
  }
  const double l = sqrt(l2);
  const d3 g0 = (m1 / a1) * (-l / a1), g3 = (m2 / a2) * (-l / a2);
  const double ta = dot(p[0] - p[1], e) / l2, tb = dot(p[3] - p[1], e) / l2;
  g[0] = g0;
  g[1] = g0 * (ta - 1.0) + g3 * (tb - 1.0);
  g[2] = g0 * (-ta) + g3 * (-tb);
  g[3] = g3;
}

// The strand's gradient (unit stiffness, strand_force_hessian's grad) in FP64.
__device__ inline void strand_grad64(d3 x0, d3 x1, d3 x2, double rest_angle, d3 grad[3]) {
  const d3 e0 = x0 - x1, e1 = x2 - x1, cr = cross(e0, e1);
  const double theta = atan2(len(cr), dot(e0, e1));
  const double cross_sq = dot(cr, cr), e0_sq = dot(e0, e0), e1_sq = dot(e1, e1);
  d3 gt0{0, 0, 0}, gt1{0, 0, 0};
  if (!(cross_sq <= 1.0e-12 * e0_sq * e1_sq)) {
    const d3 n = cr / sqrt(cross_sq);
    gt0 = cross(e0, n) / e0_sq;
    gt1 = -(cross(e1, n)) / e1_sq;
  }
  const double w = theta - rest_angle;
  grad[0] = w * gt0;
  grad[1] = w * (-gt0 - gt1);
  grad[2] = w * gt1;
}

// Central-difference Jacobian of an FP64 gradient function over NV points: J[(3 a + r) * 12 + j]
// = d grad[a][r] / d coordinate j (j < 3 NV), step 1e-6 * scale.
template <int NV, class Grad> __device__ inline void fd_jacobian64(const d3 p[NV], double scale, Grad grad, double *J) {
  const double eps = 1.0e-6 * scale, inv = 1.0 / (2.0 * eps);
  for (int j = 0; j < 3 * NV; ++j) {
    d3 pp[NV], pm[NV], gp[NV], gm[NV];
    for (int k = 0; k < NV; ++k) { pp[k] = p[k]; pm[k] = p[k]; }
    setc(pp[j / 3], j % 3, comp(pp[j / 3], j % 3) + eps);
    setc(pm[j / 3], j % 3, comp(pm[j / 3], j % 3) - eps);
    grad(pp, gp);
    grad(pm, gm);
    for (int a = 0; a < NV; ++a)
      for (int r = 0; r < 3; ++r) J[(3 * a + r) * 12 + j] = (comp(gp[a], r) - comp(gm[a], r)) * inv;
  }
}

// implicit.rs:3287-3308 rod_stretch_block (project = false: compressed segments keep their
// negative transverse stiffness, the exact second derivative).
__device__ inline void rod_stretch_block(f3 d, float rest, float k, float h[9], f3 &ga, bool project = true) {
  float l = len(d);
  if (l < 1.0e-9f) {
    for (int i = 0; i < 9; ++i) h[i] = 0.0f;
    h[0] = k; h[4] = k; h[8] = k;
    ga = f3{0, 0, 0};
    return;
  }
  f3 nrm = d / l;
  float r = (l - rest) / l;
  float c_axial = 1.0f - r;
  float c_trans = project ? fmaxf(r, 0.0f) : r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      float iden = i == j ? 1.0f : 0.0f;
      h[3 * i + j] = k * (c_axial * comp(nrm, i) * comp(nrm, j) + c_trans * iden);
    }
  ga = nrm * (k * (l - rest));
}

// implicit.rs:3240-3261 spring_block: an edge spring's 3x3 block and the gradient on a for
// d = x_a - x_b; the tangential coefficient 1 - rest/l is clamped at 0 (project = false: the
// exact second derivative).
__device__ inline void spring_block(f3 d, float rest, float k, float h[9], f3 &ga, bool project = true) {
  float l = len(d);
  if (l < 1.0e-9f) {
    for (int i = 0; i < 9; ++i) h[i] = 0.0f;
    h[0] = k; h[4] = k; h[8] = k;
    ga = f3{0, 0, 0};
    return;
  }
  f3 nrm = d / l;
  float coeff_t = 1.0f - rest / l;
  if (project) coeff_t = fmaxf(coeff_t, 0.0f);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      float nnt = comp(nrm, i) * comp(nrm, j);
      float iden = i == j ? 1.0f : 0.0f;
      h[3 * i + j] = k * (nnt + coeff_t * (iden - nnt));
    }
  ga = nrm * (k * (l - rest));
}

// rod_bend.rs:81-175 strand_force_hessian (unit stiffness).
__device__ inline void strand_force_hessian(f3 x0, f3 x1, f3 x2, float rest_angle, f3 grad[3], float *hess) {
  f3 e0 = x0 - x1, e1 = x2 - x1;
  float theta = atan2f(len(cross(e0, e1)), dot(e0, e1));
  f3 gt0{0, 0, 0}, gt1{0, 0, 0};
  {
    f3 cr = cross(e0, e1);
    float cross_sq = dot(cr, cr), e0_sq = dot(e0, e0), e1_sq = dot(e1, e1);
    if (!(cross_sq <= 1.0e-12f * e0_sq * e1_sq)) {
      f3 n = cr / sqrtf(cross_sq);
      gt0 = cross(e0, n) / e0_sq;
      gt1 = -(cross(e1, n)) / e1_sq;
    }
  }
  float w = theta - rest_angle;
  grad[0] = w * gt0;
  grad[1] = w * (-gt0 - gt1);
  grad[2] = w * gt1;
  for (int i = 0; i < 81; ++i) hess[i] = 0.0f;
  float l0 = len(e0), l1 = len(e1);
  f3 bcr = cross(e0, e1);
  float lb = len(bcr);
  if (lb <= 1.0e-9f * l0 * l1 || l0 <= 1.0e-9f || l1 <= 1.0e-9f) return;
  f3 tb = bcr / lb;
  f3 e0p = cross(e0, tb), e1p = cross(e1, tb);
  float cos_t = dot(e0, e1) / (l0 * l1);
  float sin_t = lb / (l0 * l1);
  float gamma = (l1 * l1) / (l0 * l0);
  float gm1 = gamma - 1.0f, gp1 = gamma + 1.0f;
  float r = sqrtf(4.0f * w * w * (gm1 / gp1) * (gm1 / gp1) + 1.0f);
  float rm = sqrtf(fmaxf(2.0f * (2.0f * w * w + 1.0f - r), 0.0f));
  float rp = sqrtf(fmaxf(2.0f * (2.0f * w * w + 1.0f + r), 0.0f));
  float lh = 1.0f / (l0 * l0) + 1.0f / (l1 * l1);
  float cc = 4.0f * w / gp1;
  auto accum = [&](float lam, f3 qa, f3 qb) {
    if (lam <= 0.0f) return;
    float n2 = dot(qa, qa) + dot(qb, qb);
    if (n2 <= 1.0e-20f) return;
    float p9[9] = {qa.x, qa.y, qa.z, -qa.x - qb.x, -qa.y - qb.y, -qa.z - qb.z, qb.x, qb.y, qb.z};
    float s = lam / n2;
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) hess[a * 9 + b] += s * p9[a] * p9[b];
  };
  const float SR[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
  const float SBR[4] = {-1.0f, 1.0f, -1.0f, 1.0f};
  for (int m = 0; m < 4; ++m) {
    float big_r = m < 2 ? rm : rp;
    float sr = SR[m], sbr = SBR[m];
    float a = w * (gm1 * gm1 / gp1 + sr * gp1 * r + sbr * gm1 * big_r);
    float b = -0.5f * gamma * big_r * big_r + 2.0f * w * w - sbr * 0.5f * (gm1 + sr * gp1 * r) * big_r;
    float d = 1.0f + sr * r + sbr * big_r;
    float lam = 0.25f * lh * d;
    accum(lam, a * e0 + b * e0p, cc * e1 + d * e1p);
  }
  float beta = (l1 / l0 - l0 / l1) * cos_t;
  float alpha0 = 0.5f * (-beta + sqrtf(beta * beta + 4.0f));
  float ginv = w / sin_t;
  float lam4 = ginv * (cos_t / (l1 * l1) - alpha0 / (l0 * l1));
  float lam5 = ginv * (cos_t / (l1 * l1) + 1.0f / (alpha0 * l0 * l1));
  accum(lam4, alpha0 * tb, tb);
  accum(lam5, tb, -alpha0 * tb);
}

}  // namespace nguo
