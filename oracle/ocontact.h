// Oracle contact primitives: distances with analytic gradients, barriers, friction,
// Rayleigh contact stiffness, additive CCD, the edge-triangle pierce test and
// angle-weighted pseudonormals. Each function restates the reference function named in
// its comment (paths under core-engine/crates/engine-core/src/).
// TEST INFRASTRUCTURE ONLY (checker for the HIP path).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <tuple>
#include <vector>

#include "omath.h"

namespace oracle {

// contact.rs:84-131 closest_point_on_triangle_bary (Ericson 5.1.5, 7 Voronoi regions).
template <class T>
inline V3<T> closest_point_bary(V3<T> p, V3<T> a, V3<T> b, V3<T> c, T w[3]) {
  V3<T> ab = b - a, ac = c - a, ap = p - a;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) { w[0] = 1; w[1] = 0; w[2] = 0; return a; }
  V3<T> bp = p - b;
  T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) { w[0] = 0; w[1] = 1; w[2] = 0; return b; }
  T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    T den = d1 - d3;
    T v = den != T(0) ? d1 / den : T(0);
    w[0] = T(1) - v; w[1] = v; w[2] = 0;
    return a + ab * v;
  }
  V3<T> cp = p - c;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) { w[0] = 0; w[1] = 0; w[2] = 1; return c; }
  T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    T den = d2 - d6;
    T ww = den != T(0) ? d2 / den : T(0);
    w[0] = T(1) - ww; w[1] = 0; w[2] = ww;
    return a + ac * ww;
  }
  T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    T den = (d4 - d3) + (d5 - d6);
    T ww = den != T(0) ? (d4 - d3) / den : T(0);
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

// contact.rs:315-317 point_triangle_distance.
template <class T> inline T point_triangle_distance(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
  T w[3];
  return length(p - closest_point_bary(p, a, b, c, w));
}

// contact.rs:836-858 point_tri_distance_grad (analytic: dd/dp = n, dd/dv = -w_v n).
template <class T>
inline T point_tri_distance_grad(V3<T> p, V3<T> a, V3<T> b, V3<T> c, V3<T> g[4]) {
  T w[3];
  V3<T> q = closest_point_bary(p, a, b, c, w);
  V3<T> e = p - q;
  T d = length(e);
  if (d < T(1.0e-12)) { for (int i = 0; i < 4; ++i) g[i] = V3<T>{}; return d; }
  V3<T> n = e / d;
  g[0] = n; g[1] = (-w[0]) * n; g[2] = (-w[1]) * n; g[3] = (-w[2]) * n;
  return d;
}

// contact.rs:889-902 signed_point_tri_distance_grad.
template <class T>
inline T signed_point_tri_distance_grad(V3<T> p, V3<T> a, V3<T> b, V3<T> c, bool inside, V3<T> g[4]) {
  T d = point_tri_distance_grad(p, a, b, c, g);
  if (inside) { for (int i = 0; i < 4; ++i) g[i] = -g[i]; return -d; }
  return d;
}

// contact.rs:205-215 point_edge_coeff (unclamped; degenerate edge splits evenly).
template <class T> inline void point_edge_coeff(V3<T> p, V3<T> e0, V3<T> e1, T out[2]) {
  V3<T> r = e1 - e0;
  T d = dot(r, r);
  if (d > T(0)) { T t = dot(r, p - e0) / d; out[0] = T(1) - t; out[1] = t; }
  else { out[0] = T(0.5); out[1] = T(0.5); }
}

// contact.rs:238-310 edge_edge_witness (the unclassified edge-edge witness selection).
template <class T> inline void edge_edge_witness(V3<T> ea0, V3<T> ea1, V3<T> eb0, V3<T> eb1, T out[4]) {
  V3<T> cog = T(0.25) * (ea0 + ea1 + eb0 + eb1);
  V3<T> a0 = ea0 - cog, a1 = ea1 - cog, b0v = eb0 - cog, b1v = eb1 - cog;
  auto dist2 = [&](const T c[4]) {
    V3<T> x0 = c[0] * a0 + c[1] * a1;
    V3<T> x1 = c[2] * b0v + c[3] * b1v;
    return length_sq(x1 - x0);
  };
  T pe[2];
  point_edge_coeff(ea0, eb0, eb1, pe);
  T cls[4] = {T(1), T(0), pe[0], pe[1]};
  V3<T> r0 = ea1 - ea0, r1 = eb1 - eb0;
  T a00 = dot(r0, r0), a01 = -dot(r0, r1), a11 = dot(r1, r1);
  V3<T> rhs = eb0 - ea0;
  T b0 = dot(r0, rhs), b1 = -dot(r1, rhs);
  T det = a00 * a11 - a01 * a01;
  if (det != T(0)) {
    T sp = (a11 * b0 - a01 * b1) / det;
    T tp = (a00 * b1 - a01 * b0) / det;
    T direct[4] = {T(1) - sp, sp, T(1) - tp, tp};
    if (dist2(direct) < dist2(cls)) for (int i = 0; i < 4; ++i) cls[i] = direct[i];
  }
  bool inrange = true;
  for (int i = 0; i < 4; ++i) inrange = inrange && (cls[i] >= T(0) && cls[i] <= T(1));
  if (inrange) { for (int i = 0; i < 4; ++i) out[i] = cls[i]; return; }
  auto snap = [](T c[2]) {
    if (c[0] < T(0)) { c[0] = 0; c[1] = 1; }
    else if (c[0] > T(1)) { c[0] = 1; c[1] = 0; }
  };
  T c1[2], c2[2], c3[2], c4[2];
  point_edge_coeff(ea0, eb0, eb1, c1); snap(c1);
  point_edge_coeff(ea1, eb0, eb1, c2); snap(c2);
  point_edge_coeff(eb0, ea0, ea1, c3); snap(c3);
  point_edge_coeff(eb1, ea0, ea1, c4); snap(c4);
  T types[4][4] = {{T(1), T(0), c1[0], c1[1]},
                   {T(0), T(1), c2[0], c2[1]},
                   {c3[0], c3[1], T(1), T(0)},
                   {c4[0], c4[1], T(0), T(1)}};
  int best = 0;
  T best_d = std::numeric_limits<T>::infinity();
  for (int k = 0; k < 4; ++k) {
    T d = dist2(types[k]);
    if (d < best_d) { best_d = d; best = k; }
  }
  for (int i = 0; i < 4; ++i) out[i] = types[best][i];
}

// contact.rs:489-512 segment_segment_distance_grad (one witness for d and gradient).
template <class T>
inline T segment_segment_distance_grad(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2, V3<T> g[4]) {
  T w[4];
  edge_edge_witness(p1, q1, p2, q2, w);
  V3<T> ca = w[0] * p1 + w[1] * q1;
  V3<T> cb = w[2] * p2 + w[3] * q2;
  V3<T> e = ca - cb;
  T d = length(e);
  if (d < T(1.0e-12)) { for (int i = 0; i < 4; ++i) g[i] = V3<T>{}; return d; }
  V3<T> n = e / d;
  g[0] = w[0] * n; g[1] = w[1] * n; g[2] = (-w[2]) * n; g[3] = (-w[3]) * n;
  return d;
}

// contact.rs:439-484 segment_segment_distance (Ericson 5.1.9 clamped pair; used by CCD).
template <class T> inline T segment_segment_distance(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2) {
  V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const T EPS = T(1.0e-12);
  T s, t;
  if (a <= EPS && e <= EPS) return length(p1 - p2);
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
  V3<T> c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  return length(c1 - c2);
}

// contact.rs:325-352 edge_triangle_intersect (strict opposite sides + inside test).
template <class T> inline bool edge_triangle_intersect(V3<T> e0, V3<T> e1, V3<T> v0, V3<T> v1, V3<T> v2) {
  V3<T> d1 = v1 - v0, d2 = v2 - v0, a0 = e0 - v0, a1 = e1 - v0;
  V3<T> n = cross(d1, d2);
  T s1 = dot(a0, n), s2 = dot(a1, n);
  if (!(s1 * s2 < T(0))) return false;
  T t = s1 / (s1 - s2);
  V3<T> r = a0 * (T(1) - t) + a1 * t;
  T a = dot(d1, d1), b = dot(d1, d2), c = dot(d2, d2);
  T det = a * c - b * b;
  if (det == T(0)) return false;
  T p = dot(d1, r), q = dot(d2, r);
  T u = (c * p - b * q) / det;
  T v = (a * q - b * p) / det;
  return u >= T(0) && v >= T(0) && u + v <= T(1);
}

// contact.rs:1014-1025 barrier (IPC log barrier, kappa folded in). Returns false if inactive.
template <class T> inline bool log_barrier(T d, T dhat, T kappa, T out[3]) {
  if (d >= dhat) return false;
  d = std::fmax(d, T(1.0e-6));
  T t = d - dhat;
  T ld = std::log(d / dhat);
  out[0] = -kappa * t * t * ld;
  out[1] = -kappa * (T(2) * t * ld + t * t / d);
  out[2] = -kappa * (T(2) * ld + T(4) * t / d - (t * t) / (d * d));
  return true;
}

// contact.rs:580-591 cubic_barrier (offset form), unscaled.
template <class T> inline bool cubic_barrier(T g, T ghat, T offset, T out[3]) {
  g = g - offset;
  T y = g - ghat;
  if (y >= T(0)) return false;
  out[0] = -T(2) * (y * y * y) / (T(3) * ghat);
  out[1] = -T(2) * y * y / ghat;
  out[2] = T(4) * (T(1) - g / ghat);
  return true;
}

enum BarrierKind { BARRIER_LOG = 0, BARRIER_CUBIC = 1 };

// contact.rs:723-730 barrier_of.
template <class T> inline bool barrier_of(int kind, T d, T dhat, T kappa, T out[3]) {
  if (kind == BARRIER_LOG) return log_barrier(d, dhat, kappa, out);
  if (!cubic_barrier(d, dhat, T(0), out)) return false;
  out[0] = kappa * out[0]; out[1] = kappa * out[1]; out[2] = kappa * out[2];
  return true;
}

// contact.rs:750-832 rayleigh_stiffness: (local . w) . w with local = elastic + m/g^2 I on
// the diagonal blocks, w = normalize(weights[i] * e); zero mass replaced by the stencil max.
template <class T>
inline T rayleigh_stiffness(const T (*elastic)[9], const T *mass, const T *weights, int n, V3<T> e, T offset) {
  T g = length(e) - offset;
  if (g <= T(0) || n == 0) return T(0);
  T sqr_x = g * g;
  T m_max = 0;
  for (int i = 0; i < n; ++i) m_max = std::fmax(m_max, mass[i]);
  T w[12];
  for (int i = 0; i < n; ++i) {
    V3<T> v = weights[i] * e;
    w[3 * i] = v.x; w[3 * i + 1] = v.y; w[3 * i + 2] = v.z;
  }
  T ss = 0;
  for (int i = 0; i < 3 * n; ++i) ss += w[i] * w[i];
  T norm = std::sqrt(ss);
  if (norm <= T(0)) return T(0);
  for (int i = 0; i < 3 * n; ++i) w[i] /= norm;
  T acc = 0;
  for (int ii = 0; ii < n; ++ii) {
    T m_ii = mass[ii] > T(0) ? mass[ii] : m_max;
    for (int jj = 0; jj < n; ++jj) {
      const T *blk = elastic[ii * n + jj];
      for (int r = 0; r < 3; ++r) {
        T row = 0;
        for (int c = 0; c < 3; ++c) {
          T v = blk[3 * r + c];
          if (ii == jj && r == c) v += m_ii / sqr_x;
          row += v * w[3 * jj + c];
        }
        acc += w[3 * ii + r] * row;
      }
    }
  }
  return acc;
}

// contact.rs:613-643 friction (lagged Hessian lambda * P). Returns false when zero.
template <class T>
inline bool friction(V3<T> force_contact, V3<T> dx, V3<T> normal, T mu, T min_dx, V3<T> &grad, T hess[9]) {
  if (mu <= T(0)) return false;
  T contact = -dot(normal, force_contact);
  V3<T> pdx = dx - dot(normal, dx) * normal;
  T eps = std::fmax(min_dx, T(1.0e-12));
  T denom = std::fmax(length(pdx), eps);
  T lambda = mu * contact / denom;
  if (lambda <= T(0)) return false;
  grad = pdx * lambda;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      T p = (r == c ? T(1) : T(0)) - normal[r] * normal[c];
      hess[3 * r + c] = lambda * p;
    }
  return true;
}

// contact.rs:658-706 stencil_friction: proximity weights w_a = g_a . n, slip = sum w_a dx_a.
// Outputs the per-vertex gradient and the pieces of the 12x12 Hessian (w_a w_b * fh).
template <class T>
inline bool stencil_friction(const V3<T> g[4], const V3<T> dx[4], T barrier_grad, T mu, T min_dx,
                             V3<T> fgrad[4], T w[4], T fh[9]) {
  if (mu <= T(0)) return false;
  T nlen = length(g[0]);
  if (nlen < T(1.0e-12)) return false;
  V3<T> n = g[0] / nlen;
  V3<T> slip{};
  for (int a = 0; a < 4; ++a) { w[a] = dot(g[a], n); slip += dx[a] * w[a]; }
  V3<T> f_c = n * barrier_grad;
  V3<T> fg{};
  if (!friction(f_c, slip, n, mu, min_dx, fg, fh)) return false;
  if (fg == V3<T>{}) return false;
  for (int a = 0; a < 4; ++a) fgrad[a] = fg * w[a];
  return true;
}

// contact.rs:378-431 accd_toi (conservative advancement, u_max over all point pairs).
template <class T, class F>
inline T accd_toi(const V3<T> x0[4], const V3<T> dx[4], T offset, T max_t, T reduction, uint32_t max_iter, F square_dist) {
  T u_max = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) u_max = std::fmax(u_max, length(dx[i] - dx[j]));
  if (u_max < T(1.0e-12)) return max_t;
  T inv_u_max = T(1) / u_max;
  T d0 = std::sqrt(std::fmax(square_dist(x0), T(0)));
  T eps = reduction * (d0 - offset);
  if (eps <= T(0)) return max_t;
  T target = eps + offset;
  T eps_sqr = eps * eps;
  T toi = 0;
  V3<T> moved[4];
  for (uint32_t it = 0; it < max_iter; ++it) {
    for (int k = 0; k < 4; ++k) moved[k] = x0[k] + dx[k] * toi;
    T d2 = std::fmax(square_dist(moved), T(0));
    T dmt = (d2 - target * target) / (std::sqrt(d2) + target);
    if ((max_t - toi) * u_max < dmt - eps) return max_t;
    if (toi > T(0) && dmt * dmt < eps_sqr) break;
    T toi_next = toi + dmt * inv_u_max;
    if (toi_next != toi) toi = toi_next; else break;
    if (toi > max_t) return max_t;
  }
  return std::fmax(toi, T(0));
}

// contact.rs:922-929 surface_normal_at.
template <class T> inline V3<T> surface_normal_at(const T bary[3], V3<T> face, const V3<T> pn[3]) {
  if (bary[0] > T(1.0e-4) && bary[1] > T(1.0e-4) && bary[2] > T(1.0e-4)) return face;
  return normalize_or_zero(pn[0] * bary[0] + pn[1] * bary[1] + pn[2] * bary[2]);
}

// contact.rs:943-981 pseudonormals: angle-weighted, vertices welded on exact position bits
// (-0.0 folded to 0.0), zero-area faces skipped.
template <class T>
inline std::vector<V3<T>> pseudonormals(const std::vector<V3<T>> &pos, const std::vector<std::array<uint32_t, 3>> &tris) {
  using Key = std::tuple<T, T, T>;
  auto fold = [](T c) { return c == T(0) ? T(0) : c; };
  std::map<Key, uint32_t> weld;
  std::vector<uint32_t> of(pos.size());
  for (size_t i = 0; i < pos.size(); ++i) {
    Key k{fold(pos[i].x), fold(pos[i].y), fold(pos[i].z)};
    auto it = weld.find(k);
    if (it == weld.end()) { uint32_t nx = (uint32_t)weld.size(); weld.emplace(k, nx); of[i] = nx; }
    else of[i] = it->second;
  }
  std::vector<V3<T>> acc(weld.size());
  for (auto &t : tris) {
    V3<T> p[3] = {pos[t[0]], pos[t[1]], pos[t[2]]};
    V3<T> face = normalize_or_zero(cross(p[1] - p[0], p[2] - p[0]));
    if (face == V3<T>{}) continue;
    for (int k = 0; k < 3; ++k) {
      V3<T> u = normalize_or_zero(p[(k + 1) % 3] - p[k]);
      V3<T> v = normalize_or_zero(p[(k + 2) % 3] - p[k]);
      acc[of[t[k]]] += face * std::acos(clampv(dot(u, v), T(-1), T(1)));
    }
  }
  std::vector<V3<T>> out(pos.size());
  for (size_t i = 0; i < pos.size(); ++i) out[i] = normalize_or_zero(acc[of[i]]);
  return out;
}

}  // namespace oracle
