// Oracle element models on the configs' path: isotropic shells (StVK / ARAP on the two
// principal stretches of a 3x2 deformation gradient, six-eigenpair PSD filter),
// discrete-shell hinge bending (analytic dtheta/dx, Gauss-Newton Hessian), rod stretch and
// Wu-Kim strand bending, and the Lame conversion.
// TEST INFRASTRUCTURE ONLY (checker for the HIP path).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <vector>

#include "omath.h"

namespace oracle {

enum ShellModel { SHELL_STVK = 0, SHELL_ARAP = 1 };

// material.rs:31-44 lame_from_young_poisson.
template <class T> inline void lame_from_young_poisson(T young, T poisson, T &mu, T &lambda) {
  T nu = std::isfinite(poisson) ? clampv(poisson, T(1.0e-6), T(0.5) - T(1.0e-4)) : T(0.35);
  mu = young / (T(2) * (T(1) + nu));
  lambda = young * nu / ((T(1) + nu) * (T(1) - T(2) * nu));
}

// shell_model.rs:59-78 sym_eig2: eigenvalues descending, eigenvectors as columns.
template <class T> inline void sym_eig2(const M2<T> &a, T lam[2], M2<T> &vecs) {
  T p = a.xa.x, q = a.ya.x, r = a.ya.y;
  T tr = p + r;
  T diff = std::sqrt(std::fmax((p - r) * (p - r) + T(4) * q * q, T(0)));
  T l0 = T(0.5) * (tr + diff), l1 = T(0.5) * (tr - diff);
  V2<T> v0;
  if (std::fabs(q) > T(1.0e-20)) v0 = normalize2_or(V2<T>(q, l0 - p), V2<T>(1, 0));
  else if (p >= r) v0 = V2<T>(1, 0);
  else v0 = V2<T>(0, 1);
  V2<T> v1(-v0.y, v0.x);
  lam[0] = l0; lam[1] = l1;
  vecs = M2<T>(v0, v1);
}

// shell_model.rs:84-102 svd3x2 (thin SVD, sigma descending, non-negative).
template <class T> inline void svd3x2(const M32<T> &f, M32<T> &u, T sigma[2], M2<T> &v) {
  T lam[2];
  sym_eig2(f.ata(), lam, v);
  sigma[0] = std::sqrt(std::fmax(lam[0], T(0)));
  sigma[1] = std::sqrt(std::fmax(lam[1], T(0)));
  T eps = T(1.0e-9) * std::fmax(sigma[0], T(1));
  auto fv = [&](V2<T> vv) { return f.c0 * vv.x + f.c1 * vv.y; };
  V3<T> X(1, 0, 0), Y(0, 1, 0);
  V3<T> u0 = normalize_or(sigma[0] > eps ? fv(v.xa) / sigma[0] : X, X);
  V3<T> u1_raw;
  if (sigma[1] > eps) u1_raw = fv(v.ya) / sigma[1];
  else {
    V3<T> t = std::fabs(u0.x) < T(0.9) ? X : Y;
    u1_raw = t - dot(t, u0) * u0;
  }
  V3<T> u1 = normalize_or(u1_raw - dot(u1_raw, u0) * u0, Y);
  u = M32<T>(u0, u1);
}

// shell_model.rs:131-152 shell_dpsi_dsigma (StVK: mu (s-1)^3; ARAP: mu (s-1); + detsqr).
template <class T> inline void shell_dpsi(const T s[2], int model, T mu, T lambda, T out[2]) {
  T j = s[0] * s[1];
  auto dev = [&](T x) {
    if (model == SHELL_STVK) { T d = x - T(1); return mu * (d * d * d); }
    return mu * (x - T(1));
  };
  out[0] = dev(s[0]) + lambda * s[1] * (j - T(1));
  out[1] = dev(s[1]) + lambda * s[0] * (j - T(1));
}

// shell_model.rs:156-176 shell_d2psi_dsigma2 (row-major 2x2).
template <class T> inline void shell_d2psi(const T s[2], int model, T mu, T lambda, T out[4]) {
  T a = s[0], b = s[1];
  auto dev = [&](T x) {
    if (model == SHELL_STVK) return T(3) * mu * (x - T(1)) * (x - T(1));
    return mu;
  };
  T crossv = lambda * (T(2) * a * b - T(1));
  out[0] = dev(a) + lambda * b * b;
  out[1] = crossv;
  out[2] = crossv;
  out[3] = dev(b) + lambda * a * a;
}

// shell_model.rs:112-127 shell_density (energy per unit rest area).
template <class T> inline T shell_density(const T s[2], int model, T mu, T lambda) {
  T j = s[0] * s[1];
  T dev;
  if (model == SHELL_STVK) {
    T a = s[0] - T(1), b = s[1] - T(1);
    dev = mu * (a * a * a * a + b * b * b * b) / T(4);
  } else {
    dev = T(0.5) * mu * ((s[0] - T(1)) * (s[0] - T(1)) + (s[1] - T(1)) * (s[1] - T(1)));
  }
  return dev + T(0.5) * lambda * (j - T(1)) * (j - T(1));
}

// shell_model.rs:195-240 shell_eigenpairs_from_derivs: six mode matrices (3x2 in the SVD
// frame) and their unclamped eigenvalues.
template <class T>
inline void shell_eigenpairs(const T s[2], const T deda[2], const T d2[4], T eps, M32<T> q[6], T lam[6]) {
  T ev[2];
  M2<T> evec;
  sym_eig2(M2<T>(V2<T>(d2[0], d2[2]), V2<T>(d2[1], d2[3])), ev, evec);
  const T R = T(0.70710678118654752440);
  auto m = [](T r00, T r01, T r10, T r11, T r20, T r21) {
    return M32<T>(V3<T>(r00, r10, r20), V3<T>(r01, r11, r21));
  };
  q[0] = m(0, R, -R, 0, 0, 0);
  q[1] = m(0, R, R, 0, 0, 0);
  q[2] = m(0, 0, 0, 0, 1, 0);
  q[3] = m(0, 0, 0, 0, 0, 1);
  q[4] = m(evec.xa.x, 0, 0, evec.xa.y, 0, 0);
  q[5] = m(evec.ya.x, 0, 0, evec.ya.y, 0, 0);
  T a = s[0], b = s[1];
  T denom = a - b;
  T fallback = T(0.5) * (d2[0] + d2[3]) - T(0.5) * (d2[1] + d2[2]);
  lam[0] = (deda[0] + deda[1]) / (a + b);
  lam[1] = std::fabs(denom) > eps ? (deda[0] - deda[1]) / denom : fallback;
  lam[2] = deda[0] / a;
  lam[3] = deda[1] / b;
  lam[4] = ev[0];
  lam[5] = ev[1];
}

// shell_model.rs:262-288 shell_d2psi_df2_from_derivs: sum max(lam_k,0) q_k q_k^T.
template <class T>
inline void shell_d2psi_df2(const M32<T> &u, const T s[2], const M2<T> &v, const T deda[2], const T d2[4], T eps, T out[36],
                            bool project = true) {
  M32<T> q[6];
  T lam[6];
  shell_eigenpairs(s, deda, d2, eps, q, lam);
  V3<T> u2 = cross(u.c0, u.c1);
  for (int i = 0; i < 36; ++i) out[i] = 0;
  M2<T> vt = v.transpose();
  for (int k = 0; k < 6; ++k) {
    T l = project ? std::fmax(lam[k], T(0)) : lam[k];
    if (l == T(0)) continue;
    auto col = [&](V3<T> c) { return u.c0 * c.x + u.c1 * c.y + u2 * c.z; };
    M32<T> uq(col(q[k].c0), col(q[k].c1));
    T qv[6];
    uq.mul2(vt).vec6(qv);
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) out[r * 6 + c] += l * qv[r] * qv[c];
  }
}

template <class T> struct ShellElem {
  uint32_t v[3];
  M2<T> dm_inv;
  T area;
  uint32_t face = 0;  // index among the faces handed to build_shells
};

// shell_model.rs:309-349 build_shells (intrinsic in-plane frame; skips degenerate faces).
template <class T>
inline std::vector<ShellElem<T>> build_shells(const std::vector<std::array<uint32_t, 3>> &tris, const std::vector<V3<T>> &rest, T min_area) {
  std::vector<ShellElem<T>> out;
  for (uint32_t fi = 0; fi < (uint32_t)tris.size(); ++fi) {
    const auto &t = tris[fi];
    V3<T> p0 = rest[t[0]], p1 = rest[t[1]], p2 = rest[t[2]];
    V3<T> e1 = p1 - p0, e2 = p2 - p0;
    V3<T> n = cross(e1, e2);
    T area = T(0.5) * length(n);
    if (area <= min_area) continue;
    V3<T> ax = normalize(e1);
    V3<T> ay = normalize(e2 - dot(e2, ax) * ax);
    M2<T> dm(V2<T>(dot(e1, ax), dot(e1, ay)), V2<T>(dot(e2, ax), dot(e2, ay)));
    if (std::fabs(dm.det()) < T(1e-20)) continue;
    ShellElem<T> el;
    el.v[0] = t[0]; el.v[1] = t[1]; el.v[2] = t[2];
    el.dm_inv = dm.inverse();
    el.area = area;
    el.face = fi;
    out.push_back(el);
  }
  return out;
}

template <class T> inline M32<T> shell_F(const V3<T> p[3], const M2<T> &dm_inv) {
  return M32<T>(p[1] - p[0], p[2] - p[0]).mul2(dm_inv);
}

template <class T> inline T shell_energy(const V3<T> p[3], const ShellElem<T> &el, int model, T mu, T lambda) {
  M32<T> f = shell_F(p, el.dm_inv), u;
  T s[2];
  M2<T> v;
  svd3x2(f, u, s, v);
  return el.area * shell_density(s, model, mu, lambda);
}

// shell_model.rs:367-380 shell_gradient_analytic.
template <class T> inline void shell_gradient(const V3<T> p[3], const ShellElem<T> &el, int model, T mu, T lambda, V3<T> g[3]) {
  M32<T> f = shell_F(p, el.dm_inv), u;
  T s[2];
  M2<T> v;
  svd3x2(f, u, s, v);
  T d[2];
  shell_dpsi(s, model, mu, lambda, d);
  M32<T> pk1 = M32<T>(u.c0 * d[0], u.c1 * d[1]).mul2(v.transpose());
  M32<T> h = pk1.mul2(el.dm_inv.transpose());
  V3<T> g1 = h.c0 * el.area, g2 = h.c1 * el.area;
  g[0] = -(g1 + g2); g[1] = g1; g[2] = g2;
}

// shell_model.rs:384-400 shell_dfdx (6x9 constant map).
template <class T> inline void shell_dfdx(const M2<T> &dm_inv, T b[54]) {
  for (int i = 0; i < 54; ++i) b[i] = 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 2; ++j) {
      int row = 2 * i + j;
      T sum = 0;
      for (int k = 0; k < 2; ++k) {
        T d = dm_inv.at(k, j);
        b[row * 9 + 3 * (k + 1) + i] = d;
        sum += d;
      }
      b[row * 9 + i] = -sum;
    }
}

// shell_model.rs:403-436 shell_hessian_analytic: area * B^T A6 B.
// project = false: the exact second derivative (negative mode eigenvalues kept).
template <class T>
inline void shell_hessian(const V3<T> p[3], const ShellElem<T> &el, int model, T mu, T lambda, T eps, T h[81], bool project = true) {
  M32<T> f = shell_F(p, el.dm_inv), u;
  T s[2];
  M2<T> v;
  svd3x2(f, u, s, v);
  T deda[2], d2[4], a6[36], b[54], ab[54];
  shell_dpsi(s, model, mu, lambda, deda);
  shell_d2psi(s, model, mu, lambda, d2);
  shell_d2psi_df2(u, s, v, deda, d2, eps, a6, project);
  shell_dfdx(el.dm_inv, b);
  for (int r = 0; r < 6; ++r)
    for (int c = 0; c < 9; ++c) {
      T acc = 0;
      for (int k = 0; k < 6; ++k) acc += a6[r * 6 + k] * b[k * 9 + c];
      ab[r * 9 + c] = acc;
    }
  for (int r = 0; r < 9; ++r)
    for (int c = 0; c < 9; ++c) {
      T acc = 0;
      for (int k = 0; k < 6; ++k) acc += b[k * 9 + r] * ab[k * 9 + c];
      h[r * 9 + c] = el.area * acc;
    }
}

// ---------------------------------------------------------------- hinge bending
template <class T> struct Hinge {
  uint32_t v[4];  // opp_a, shared_0, shared_1, opp_b
  T rest, w;
};

// bending.rs:34-48 dihedral_angle (signed; flat flap reads 0).
template <class T> inline T dihedral_angle(const V3<T> p[4]) {
  V3<T> n1 = cross(p[1] - p[0], p[2] - p[0]);
  V3<T> n2 = cross(p[1] - p[2], p[3] - p[2]);
  V3<T> e = normalize_or_zero(p[2] - p[1]);
  V3<T> n1n = normalize_or_zero(n1), n2n = normalize_or_zero(n2);
  T sn = dot(cross(n1n, n2n), e);
  T cs = clampv(dot(n1n, n2n), T(-1), T(1));
  return std::atan2(sn, cs);
}

// bending.rs:51-67 dihedral_grad (central differences; degenerate fallback only).
template <class T> inline void dihedral_grad_fd(const V3<T> p[4], V3<T> g[4]) {
  T scale = std::fmax(length(p[0]) + length(p[1]) + length(p[2]) + length(p[3]), T(1));
  T h = T(1.0e-4) * scale;
  T inv = T(1) / (T(2) * h);
  for (int vi = 0; vi < 4; ++vi)
    for (int c = 0; c < 3; ++c) {
      V3<T> pp[4], pm[4];
      for (int k = 0; k < 4; ++k) { pp[k] = p[k]; pm[k] = p[k]; }
      pp[vi][c] += h;
      pm[vi][c] -= h;
      g[vi][c] = (dihedral_angle(pp) - dihedral_angle(pm)) * inv;
    }
}

// bending.rs:78-98 dihedral_grad_analytic.
template <class T> inline void dihedral_grad(const V3<T> p[4], V3<T> g[4]) {
  V3<T> m1 = cross(p[1] - p[0], p[2] - p[0]);
  V3<T> m2 = cross(p[1] - p[2], p[3] - p[2]);
  V3<T> e = p[2] - p[1];
  T a1 = length(m1), a2 = length(m2), l2 = length_sq(e);
  if (a1 < T(1.0e-12) || a2 < T(1.0e-12) || l2 < T(1.0e-24)) { dihedral_grad_fd(p, g); return; }
  T l = std::sqrt(l2);
  V3<T> n1 = m1 / a1, n2 = m2 / a2;
  V3<T> g0 = n1 * (-l / a1);
  V3<T> g3 = n2 * (-l / a2);
  T ta = dot(p[0] - p[1], e) / l2;
  T tb = dot(p[3] - p[1], e) / l2;
  g[0] = g0;
  g[1] = g0 * (ta - T(1)) + g3 * (tb - T(1));
  g[2] = g0 * (-ta) + g3 * (-tb);
  g[3] = g3;
}

// bending.rs:128-158 build_hinges: interior edges (exactly two incident faces), sorted by
// the (min,max) edge key; opposite vertices in face order.
template <class T>
inline std::vector<Hinge<T>> build_hinges(const std::vector<std::array<uint32_t, 3>> &tris, const std::vector<V3<T>> &pos) {
  std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> m;
  for (auto &t : tris) {
    uint32_t e[3][3] = {{t[1], t[2], t[0]}, {t[0], t[2], t[1]}, {t[0], t[1], t[2]}};
    for (auto &x : e) {
      auto key = x[0] < x[1] ? std::make_pair(x[0], x[1]) : std::make_pair(x[1], x[0]);
      m[key].push_back(x[2]);
    }
  }
  std::vector<Hinge<T>> out;
  for (auto &kv : m) {
    if (kv.second.size() != 2) continue;
    Hinge<T> h;
    h.v[0] = kv.second[0]; h.v[1] = kv.first.first; h.v[2] = kv.first.second; h.v[3] = kv.second[1];
    V3<T> p[4] = {pos[h.v[0]], pos[h.v[1]], pos[h.v[2]], pos[h.v[3]]};
    h.rest = dihedral_angle(p);
    h.w = T(1);
    out.push_back(h);
  }
  return out;
}

// ---------------------------------------------------------------- rods
// implicit.rs:3287-3308 rod_stretch_block: 3x3 block (row-major) and gradient on a.
template <class T> inline void rod_stretch_block(V3<T> d, T rest, T k, T h[9], V3<T> &ga, bool project = true) {
  T l = length(d);
  if (l < T(1.0e-9)) {
    for (int i = 0; i < 9; ++i) h[i] = 0;
    h[0] = k; h[4] = k; h[8] = k;
    ga = V3<T>{};
    return;
  }
  V3<T> nrm = d / l;
  T r = (l - rest) / l;
  T c_axial = T(1) - r;
  T c_trans = project ? std::fmax(r, T(0)) : r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T iden = i == j ? T(1) : T(0);
      h[3 * i + j] = k * (c_axial * nrm[i] * nrm[j] + c_trans * iden);
    }
  ga = nrm * (k * (l - rest));
}

// implicit.rs:3240-3261 spring_block: the edge spring's 3x3 block (row-major) and gradient on
// a for d = x_a - x_b; the tangential coefficient 1 - rest/l is clamped at 0 (SPD projection)
// unless project is false (the exact second derivative).
template <class T> inline void spring_block(V3<T> d, T rest, T k, T h[9], V3<T> &ga, bool project = true) {
  T l = length(d);
  if (l < T(1.0e-9)) {
    for (int i = 0; i < 9; ++i) h[i] = 0;
    h[0] = k; h[4] = k; h[8] = k;
    ga = V3<T>{};
    return;
  }
  V3<T> nrm = d / l;
  T coeff_t = T(1) - rest / l;
  if (project) coeff_t = std::fmax(coeff_t, T(0));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T nnt = nrm[i] * nrm[j];
      T iden = i == j ? T(1) : T(0);
      h[3 * i + j] = k * (nnt + coeff_t * (iden - nnt));
    }
  ga = nrm * (k * (l - rest));
}

// rod_bend.rs:43-57 gradient_theta.
template <class T> inline void strand_gradient_theta(V3<T> e0, V3<T> e1, V3<T> out[2]) {
  V3<T> cr = cross(e0, e1);
  T cross_sq = length_sq(cr);
  T e0_sq = dot(e0, e0), e1_sq = dot(e1, e1);
  if (cross_sq <= T(1.0e-12) * e0_sq * e1_sq) { out[0] = V3<T>{}; out[1] = V3<T>{}; return; }
  V3<T> n = cr / std::sqrt(cross_sq);
  out[0] = cross(e0, n) / e0_sq;
  out[1] = -(cross(e1, n)) / e1_sq;
}

// rod_bend.rs:81-175 strand_force_hessian (unit stiffness; Wu-Kim eigen-projected 9x9).
template <class T> inline void strand_force_hessian(V3<T> x0, V3<T> x1, V3<T> x2, T rest_angle, V3<T> grad[3], T hess[81]) {
  V3<T> e0 = x0 - x1, e1 = x2 - x1;
  T theta = std::atan2(length(cross(e0, e1)), dot(e0, e1));
  V3<T> gt[2];
  strand_gradient_theta(e0, e1, gt);
  T w = theta - rest_angle;
  grad[0] = w * gt[0];
  grad[1] = w * (-gt[0] - gt[1]);
  grad[2] = w * gt[1];
  for (int i = 0; i < 81; ++i) hess[i] = 0;
  T l0 = length(e0), l1 = length(e1);
  V3<T> bcr = cross(e0, e1);
  T lb = length(bcr);
  if (lb <= T(1.0e-9) * l0 * l1 || l0 <= T(1.0e-9) || l1 <= T(1.0e-9)) return;
  V3<T> tb = bcr / lb;
  V3<T> e0p = cross(e0, tb), e1p = cross(e1, tb);
  T cos_t = dot(e0, e1) / (l0 * l1);
  T sin_t = lb / (l0 * l1);
  T gamma = (l1 * l1) / (l0 * l0);
  T gm1 = gamma - T(1), gp1 = gamma + T(1);
  T r = std::sqrt(T(4) * w * w * (gm1 / gp1) * (gm1 / gp1) + T(1));
  T rm = std::sqrt(std::fmax(T(2) * (T(2) * w * w + T(1) - r), T(0)));
  T rp = std::sqrt(std::fmax(T(2) * (T(2) * w * w + T(1) + r), T(0)));
  T lh = T(1) / (l0 * l0) + T(1) / (l1 * l1);
  T cc = T(4) * w / gp1;
  auto accum = [&](T lam, V3<T> qa, V3<T> qb) {
    if (lam <= T(0)) return;
    T n2 = dot(qa, qa) + dot(qb, qb);
    if (n2 <= T(1.0e-20)) return;
    T p9[9];
    for (int c = 0; c < 3; ++c) { p9[c] = qa[c]; p9[3 + c] = -qa[c] - qb[c]; p9[6 + c] = qb[c]; }
    T s = lam / n2;
    for (int a = 0; a < 9; ++a)
      for (int b = 0; b < 9; ++b) hess[a * 9 + b] += s * p9[a] * p9[b];
  };
  const T SR[4] = {T(-1), T(-1), T(1), T(1)};
  const T SBR[4] = {T(-1), T(1), T(-1), T(1)};
  for (int m = 0; m < 4; ++m) {
    T big_r = m < 2 ? rm : rp;
    T sr = SR[m], sbr = SBR[m];
    T a = w * (gm1 * gm1 / gp1 + sr * gp1 * r + sbr * gm1 * big_r);
    T b = -T(0.5) * gamma * big_r * big_r + T(2) * w * w - sbr * T(0.5) * (gm1 + sr * gp1 * r) * big_r;
    T d = T(1) + sr * r + sbr * big_r;
    T lam = T(0.25) * lh * d;
    accum(lam, a * e0 + b * e0p, cc * e1 + d * e1p);
  }
  T beta = (l1 / l0 - l0 / l1) * cos_t;
  T alpha0 = T(0.5) * (-beta + std::sqrt(beta * beta + T(4)));
  T ginv = w / sin_t;
  T lam4 = ginv * (cos_t / (l1 * l1) - alpha0 / (l0 * l1));
  T lam5 = ginv * (cos_t / (l1 * l1) + T(1) / (alpha0 * l0 * l1));
  accum(lam4, alpha0 * tb, tb);
  accum(lam5, tb, -alpha0 * tb);
}

template <class T> inline T strand_angle(V3<T> x0, V3<T> x1, V3<T> x2) {
  V3<T> e0 = x0 - x1, e1 = x2 - x1;
  return std::atan2(length(cross(e0, e1)), dot(e0, e1));
}

}  // namespace oracle
