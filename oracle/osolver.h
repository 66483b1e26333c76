// Oracle implicit solver: a restatement of engine-core's ImplicitSolver for the terms the
// workloads use (inertia, rods, rod bends, isotropic shells, hinge bending, analytic
// colliders, vertex-triangle and edge-edge self/body contact with Rayleigh stiffness,
// barrier and lagged friction), the Dirichlet-filtered scalar CSR, the f64 block-Jacobi CG
// and the ACCD line search.
// TEST INFRASTRUCTURE ONLY (checker for the HIP path).
//
// Candidate sets: the reference enumerates candidates with a uniform hash grid
// (broadphase.rs) whose output is a superset of "the point lies in the triangle's AABB
// padded by the search radius" (resp. "the two edge AABBs overlap once one is padded").
// This oracle enumerates exactly that AABB relation. For the barrier stencils the active
// filter (d < dhat) makes the result independent of the superset; for CCD the minimum over
// candidates is the same except where a grid-only extra pair would itself clamp the step.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

#include "../include/nguo_hip.h"
#include "ocontact.h"
#include "oelements.h"

namespace oracle {

struct Csr {
  uint32_t n = 0;
  std::vector<uint32_t> row_ptr, col_idx;
  std::vector<float> values;
  std::vector<double> values64;  // FP64 copies of the values (when the assembly keeps them)
};

template <class T> struct Stencil {
  uint32_t idx[4];
  T d;
  V3<T> g[4];
  T ghat;
  uint8_t ee = 0;      // 0: vertex-triangle (idx = v, t0, t1, t2), 1: edge-edge (ea0, ea1, eb0, eb1)
  uint8_t inside = 0;  // vertex-triangle against a collider facet with the vertex inside (signed gap)
};

// `count` points as Jet coordinates: point a's coordinate c is input 3 a + c.
inline void jet_points(const V3<double> *p, int count, V3<Jet> *out) {
  for (int a = 0; a < count; ++a)
    for (int c = 0; c < 3; ++c) out[a][c] = Jet::var(p[a][c], 3 * a + c);
}

struct CgStats {
  uint32_t iters = 0;
  double residual = 0;
};

// linsolve.rs:38-61 invert3x3 (f64).
inline bool invert3x3(const double m[3][3], double out[3][3]) {
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  if (std::fabs(det) < 1e-20) return false;
  double id = 1.0 / det;
  out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
  out[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  out[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  out[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  out[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
  return true;
}

// linsolve.rs:272-411 cg_cpu_precond(Block) + cg_run: f64 PCG, residual checked every
// iteration; FP64 right-hand side and result.
inline std::vector<double> cg_block_jacobi64(const Csr &a, const std::vector<double> &bf, double tol, uint32_t max_iter, uint32_t *iters_out) {
  const uint32_t n = a.n;
  const bool v64 = !a.values64.empty();
  auto val = [&](uint32_t k) { return v64 ? a.values64[k] : (double)a.values[k]; };
  bool block3 = n % 3 == 0;
  std::vector<double> inv_diag(n, 1.0);
  for (uint32_t r = 0; r < n; ++r)
    for (uint32_t k = a.row_ptr[r]; k < a.row_ptr[r + 1]; ++k)
      if (a.col_idx[k] == r) { double d = val(k); inv_diag[r] = d != 0.0 ? 1.0 / d : 1.0; }
  std::vector<std::array<double, 9>> binv;
  if (block3) {
    binv.resize(n / 3);
    for (uint32_t i = 0; i < n / 3; ++i) {
      double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, o[3][3];
      for (int r = 0; r < 3; ++r) {
        uint32_t row = 3 * i + r;
        for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) {
          uint32_t c = a.col_idx[k];
          if (c >= 3 * i && c < 3 * i + 3) m[r][c - 3 * i] = val(k);
        }
      }
      if (invert3x3(m, o)) {
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) binv[i][3 * r + c] = o[r][c];
      } else {
        binv[i] = {inv_diag[3 * i], 0, 0, 0, inv_diag[3 * i + 1], 0, 0, 0, inv_diag[3 * i + 2]};
      }
    }
  }
  auto precond = [&](const std::vector<double> &r, std::vector<double> &z) {
    if (block3) {
      for (uint32_t i = 0; i < n / 3; ++i) {
        const auto &bi = binv[i];
        double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        z[3 * i] = bi[0] * r0 + bi[1] * r1 + bi[2] * r2;
        z[3 * i + 1] = bi[3] * r0 + bi[4] * r1 + bi[5] * r2;
        z[3 * i + 2] = bi[6] * r0 + bi[7] * r1 + bi[8] * r2;
      }
    } else {
      for (uint32_t i = 0; i < n; ++i) z[i] = r[i] * inv_diag[i];
    }
  };
  auto spmv = [&](const std::vector<double> &x, std::vector<double> &out) {
    for (uint32_t row = 0; row < n; ++row) {
      double s = 0;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) s += val(k) * x[a.col_idx[k]];
      out[row] = s;
    }
  };
  auto dotv = [](const std::vector<double> &u, const std::vector<double> &v) {
    double s = 0;
    for (size_t i = 0; i < u.size(); ++i) s += u[i] * v[i];
    return s;
  };
  double bnorm = std::fmax(std::sqrt(dotv(bf, bf)), 1e-30);
  std::vector<double> x(n, 0.0), r = bf, z(n), p, ap(n);
  precond(r, z);
  p = z;
  double rz = dotv(r, z);
  uint32_t iters = 0;
  double residual = std::sqrt(dotv(r, r)) / bnorm;
  while (iters < max_iter) {
    spmv(p, ap);
    double pap = dotv(p, ap);
    if (pap <= 0.0 || std::isnan(pap)) break;
    double alpha = rz / pap;
    for (uint32_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
    iters++;
    residual = std::sqrt(dotv(r, r)) / bnorm;
    if (residual <= tol) break;
    precond(r, z);
    double rz_new = dotv(r, z);
    double beta = rz_new / rz;
    for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz_new;
  }
  if (iters_out) *iters_out = iters;
  return x;
}

// The same with an f32 right-hand side and result (the reference's solve seam).
inline std::vector<float> cg_block_jacobi(const Csr &a, const std::vector<float> &b, float tol, uint32_t max_iter, CgStats *st) {
  const uint32_t n = a.n;
  std::vector<double> bf(n);
  for (uint32_t i = 0; i < n; ++i) bf[i] = b[i];
  uint32_t iters = 0;
  std::vector<double> x = cg_block_jacobi64(a, bf, (double)tol, max_iter, &iters);
  std::vector<float> out(n);
  for (uint32_t i = 0; i < n; ++i) out[i] = (float)x[i];
  if (st) {
    // linsolve.rs:445-455 relative_residual with the f32 solution and an f32 SpMV.
    double num = 0, den = 0;
    for (uint32_t row = 0; row < n; ++row) {
      float s = 0.0f;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) s += a.values[k] * out[a.col_idx[k]];
      double d = (double)b[row] - (double)s;
      num += d * d;
      den += (double)b[row] * (double)b[row];
    }
    st->iters = iters;
    st->residual = std::sqrt(num) / std::fmax(std::sqrt(den), 1e-30);
  }
  return out;
}

// linsolve.rs:80-99 cholesky (lower, row-major); false if a pivot is <= 1e-14.
inline bool cholesky(const std::vector<double> &a, int m, std::vector<double> &l) {
  l.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = a[(size_t)i * m + j];
      for (int k = 0; k < j; ++k) s -= l[(size_t)i * m + k] * l[(size_t)j * m + k];
      if (i == j) {
        if (s <= 1e-14) return false;
        l[(size_t)i * m + i] = std::sqrt(s);
      } else {
        l[(size_t)i * m + j] = s / l[(size_t)j * m + j];
      }
    }
  return true;
}

// linsolve.rs:102-118 chol_solve.
inline void chol_solve(const std::vector<double> &l, int m, const double *b, double *out) {
  std::vector<double> y(m);
  for (int i = 0; i < m; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= l[(size_t)i * m + k] * y[k];
    y[i] = s / l[(size_t)i * m + i];
  }
  for (int i = m - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < m; ++k) s -= l[(size_t)k * m + i] * out[k];
    out[i] = s / l[(size_t)i * m + i];
  }
}

// linsolve.rs:126-258 SchwarzTopology: greedy BFS aggregation of the block graph into
// disjoint domains of `domain` vertices; per-domain dense block with the diagonal floor
// 1e-6 tr/m + 1e-8, f64 Cholesky (diagonal fallback), block solve.
struct Schwarz {
  std::vector<std::vector<uint32_t>> domains;
  std::vector<std::vector<double>> factors;
  std::vector<int> dims;
  static std::vector<std::vector<uint32_t>> build_domains(const Csr &a, uint32_t domain) {
    uint32_t nv = a.n / 3;
    std::vector<std::vector<uint32_t>> adj(nv);
    for (uint32_t row = 0; row < a.n; ++row) {
      uint32_t i = row / 3;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) {
        uint32_t j = a.col_idx[k] / 3;
        if (j != i) adj[i].push_back(j);
      }
    }
    for (auto &v : adj) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); }
    std::vector<uint8_t> assigned(nv, 0);
    std::vector<std::vector<uint32_t>> doms;
    for (uint32_t seed = 0; seed < nv; ++seed) {
      if (assigned[seed]) continue;
      std::vector<uint32_t> dom{seed};
      assigned[seed] = 1;
      size_t fi = 0;
      while (dom.size() < domain && fi < dom.size()) {
        uint32_t v = dom[fi++];
        for (uint32_t nb : adj[v])
          if (!assigned[nb] && dom.size() < domain) { assigned[nb] = 1; dom.push_back(nb); }
      }
      doms.push_back(dom);
    }
    return doms;
  }
  void factorize(const Csr &a) {
    factors.clear();
    dims.clear();
    for (auto &dom : domains) {
      int m = 3 * (int)dom.size();
      std::vector<uint32_t> dofs;
      for (uint32_t v : dom) { dofs.push_back(3 * v); dofs.push_back(3 * v + 1); dofs.push_back(3 * v + 2); }
      std::vector<double> ad((size_t)m * m, 0.0);
      for (int p = 0; p < m; ++p)
        for (uint32_t k = a.row_ptr[dofs[p]]; k < a.row_ptr[dofs[p] + 1]; ++k)
          for (int q = 0; q < m; ++q)
            if (a.col_idx[k] == dofs[q]) ad[(size_t)p * m + q] = (double)a.values[k];
      double tr = 0;
      for (int k = 0; k < m; ++k) tr += ad[(size_t)k * m + k];
      double floor = 1e-6 * (tr / m) + 1e-8;
      for (int k = 0; k < m; ++k) ad[(size_t)k * m + k] += floor;
      std::vector<double> l;
      if (!cholesky(ad, m, l)) {
        l.assign((size_t)m * m, 0.0);
        for (int p = 0; p < m; ++p) {
          double v = ad[(size_t)p * m + p];
          l[(size_t)p * m + p] = v > 1e-14 ? std::sqrt(v) : 1.0;
        }
      }
      factors.push_back(l);
      dims.push_back(m);
    }
  }
  void apply(const std::vector<double> &r, std::vector<double> &z) const {
    for (size_t d = 0; d < domains.size(); ++d) {
      const auto &dom = domains[d];
      int m = dims[d];
      std::vector<double> rd(m), xd(m);
      for (size_t p = 0; p < dom.size(); ++p)
        for (int c = 0; c < 3; ++c) rd[3 * p + c] = r[3 * dom[p] + c];
      chol_solve(factors[d], m, rd.data(), xd.data());
      for (size_t p = 0; p < dom.size(); ++p)
        for (int c = 0; c < 3; ++c) z[3 * dom[p] + c] = xd[3 * p + c];
    }
  }
};

// linsolve.rs:522-657 MultilevelSchwarz (levels = 2): z = S0 r + w C1^T S1 (C1 r). The
// coarse node of a fine vertex is its level-0 domain; C1 sums a domain's rows (restriction
// by summation, prolongation by injection); A1 = C1 A C1^T is the Galerkin coarse operator,
// assembled by summing every fine entry into (3 agg(row/3) + row%3, 3 agg(col/3) + col%3) in
// CSR order; S1 is a one-level Schwarz over A1 with the same domain size. Weight 0 leaves
// the one-level preconditioner S0.
struct Schwarz2 {
  Schwarz l0, l1;
  std::vector<uint32_t> agg0;
  uint32_t n_coarse = 0;
  double weight = 0.0;
  Csr coarse;
  void build(const Csr &a, uint32_t domain, double w) {
    weight = w;
    l0.domains = Schwarz::build_domains(a, domain);
    l0.factorize(a);
    if (weight == 0.0) return;
    uint32_t nv = a.n / 3;
    agg0.assign(nv, 0);
    for (size_t d = 0; d < l0.domains.size(); ++d)
      for (uint32_t v : l0.domains[d]) agg0[v] = (uint32_t)d;
    n_coarse = (uint32_t)l0.domains.size();
    // coarse CSR pattern: sorted unique coarse columns per coarse row
    const uint32_t nc = 3 * n_coarse;
    std::vector<std::vector<uint32_t>> cols(nc);
    for (uint32_t row = 0; row < a.n; ++row) {
      uint32_t cr = 3 * agg0[row / 3] + row % 3;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) {
        uint32_t col = a.col_idx[k];
        cols[cr].push_back(3 * agg0[col / 3] + col % 3);
      }
    }
    coarse.n = nc;
    coarse.row_ptr.assign(1, 0);
    coarse.col_idx.clear();
    for (uint32_t r = 0; r < nc; ++r) {
      auto &c = cols[r];
      std::sort(c.begin(), c.end());
      c.erase(std::unique(c.begin(), c.end()), c.end());
      coarse.col_idx.insert(coarse.col_idx.end(), c.begin(), c.end());
      coarse.row_ptr.push_back((uint32_t)coarse.col_idx.size());
    }
    // f32 values summed in fine CSR order (linsolve.rs:613-625 coarse_matrix)
    coarse.values.assign(coarse.col_idx.size(), 0.0f);
    for (uint32_t row = 0; row < a.n; ++row) {
      uint32_t cr = 3 * agg0[row / 3] + row % 3;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) {
        uint32_t col = a.col_idx[k];
        uint32_t cc = 3 * agg0[col / 3] + col % 3;
        auto it = std::lower_bound(coarse.col_idx.begin() + coarse.row_ptr[cr], coarse.col_idx.begin() + coarse.row_ptr[cr + 1], cc);
        coarse.values[it - coarse.col_idx.begin()] += a.values[k];
      }
    }
    l1.domains = Schwarz::build_domains(coarse, domain);
    l1.factorize(coarse);
  }
  void apply(const std::vector<double> &r, std::vector<double> &z) const {
    l0.apply(r, z);
    if (weight == 0.0) return;
    const uint32_t nv = (uint32_t)agg0.size();
    std::vector<double> rc(3 * (size_t)n_coarse, 0.0), ec(3 * (size_t)n_coarse, 0.0);
    for (uint32_t i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) rc[3 * agg0[i] + c] += r[3 * i + c];
    l1.apply(rc, ec);
    for (uint32_t i = 0; i < nv; ++i)
      for (int c = 0; c < 3; ++c) z[3 * i + c] += weight * ec[3 * agg0[i] + c];
  }
};

// linsolve.rs:272-411 with Precond::Schwarz / SchwarzMl (and the cg_run loop of
// cg_block_jacobi): weight 0 = one level, otherwise the two-level form above.
inline std::vector<float> cg_schwarz(const Csr &a, const std::vector<float> &b, float tol, uint32_t max_iter, uint32_t domain, double weight,
                                     CgStats *st) {
  const uint32_t n = a.n;
  Schwarz2 sz;
  sz.build(a, domain, weight);
  auto spmv = [&](const std::vector<double> &x, std::vector<double> &out) {
    for (uint32_t row = 0; row < n; ++row) {
      double s = 0;
      for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k) s += (double)a.values[k] * x[a.col_idx[k]];
      out[row] = s;
    }
  };
  auto dotv = [](const std::vector<double> &u, const std::vector<double> &v) {
    double s = 0;
    for (size_t i = 0; i < u.size(); ++i) s += u[i] * v[i];
    return s;
  };
  std::vector<double> bf(n);
  for (uint32_t i = 0; i < n; ++i) bf[i] = b[i];
  double bnorm = std::fmax(std::sqrt(dotv(bf, bf)), 1e-30);
  std::vector<double> x(n, 0.0), r = bf, z(n), p, ap(n);
  sz.apply(r, z);
  p = z;
  double rz = dotv(r, z);
  uint32_t iters = 0;
  double residual = std::sqrt(dotv(r, r)) / bnorm;
  while (iters < max_iter) {
    spmv(p, ap);
    double pap = dotv(p, ap);
    if (pap <= 0.0 || std::isnan(pap)) break;
    double alpha = rz / pap;
    for (uint32_t i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
    iters++;
    residual = std::sqrt(dotv(r, r)) / bnorm;
    if (residual <= (double)tol) break;
    sz.apply(r, z);
    double rz_new = dotv(r, z);
    double beta = rz_new / rz;
    for (uint32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rz = rz_new;
  }
  std::vector<float> out(n);
  for (uint32_t i = 0; i < n; ++i) out[i] = (float)x[i];
  if (st) { st->iters = iters; st->residual = residual; }
  return out;
}

template <class T> class Solver {
 public:
  using V = V3<T>;
  using Tri = std::array<uint32_t, 3>;
  using Edge = std::array<uint32_t, 2>;

  // state
  std::vector<V> pos, vel, prev;
  std::vector<float> inv_mass;  // the reference keeps inv_mass in f32 in both tiers
  double time = 0;
  nguo_params prm{};
  // topology
  std::vector<Tri> tris;
  std::vector<Hinge<T>> hinges;
  std::vector<ShellElem<T>> shells;
  // Per shell face rest-shape scale s (empty = 1): the rest shape of the face is scaled by s,
  // i.e. Dm^-1 / s and area s^2 (the parameters of the drape adjoint).
  std::vector<T> face_scale;
  std::vector<V> last_xt;  // inertial target of the last step
  // implicit.rs:430-444 kinematic_dx: the static collider's displacement set since the last
  // step (the friction slip override), consumed and cleared by step(); empty = never moved
  std::vector<V> kinematic_dx;
  V body_dx(uint32_t i) const { return kinematic_dx.empty() ? V{} : kinematic_dx[i]; }
  bool project_elastic = true;  // PSD-project shell / rod-segment Hessians (false: exact)
  bool with_rod_hinges = true;  // false: leave the strand bending terms out of the assembly

  // What one Newton iteration holds fixed when it assembles at a point x_ref: the stencil set,
  // each stencil's Rayleigh stiffness (implicit.rs:755-795, recomputed per assembly but a
  // constant of it), the friction normal force, contact normal and proximity weights
  // (contact.rs:613-706, lagged), and for the analytic colliders the same per (vertex,
  // collider). The step energy at x is the energy with these taken at x itself; its gradient
  // at x is the assembled gradient. hinge_g / fric_lam / col_lam additionally fix the search
  // directions of the Gauss-Newton terms: the assembled Hessian is the Jacobian of the
  // gradient with those directions held (the frozen-direction gradient model below).
  struct Frozen {
    std::vector<Stencil<T>> st;
    std::vector<T> kappa;
    std::vector<int> kind;                 // barrier kind per stencil
    std::vector<uint8_t> fric;             // friction term present at x_ref
    std::vector<T> fric_c;                 // mu * normal force magnitude (-mu b'(d))
    std::vector<V> fric_n;
    std::vector<std::array<T, 4>> fric_w;
    std::vector<T> fric_lam;               // lambda = mu c / max(eps, |P slip|) at x_ref
    std::vector<std::array<V, 4>> hinge_g;  // dtheta/dx at x_ref
    std::vector<uint8_t> col_fric;         // per (vertex, collider), vertex-major
    std::vector<V> col_n;
    std::vector<T> col_c, col_lam;
  };
  const Frozen *gn = nullptr;  // non-null: assemble the frozen-direction gradient model
  // With gn: the frozen-parameter model instead. Only the stencil set, Rayleigh stiffnesses,
  // barrier kinds and the friction quantities (normal, weights, lambda) stay those of x_ref;
  // distances, distance gradients, hinge and strand angles and collider normals are those of
  // x, and the assembled Hessian is the exact Jacobian of the assembled gradient: the barrier's
  // b'(d) d(grad d)/dx, b'' unclamped, the hinge's (theta - rest) d(grad theta)/dx, the exact
  // strand Jacobian and a sphere collider's b'(d) (I - n n^T) / r join the Gauss-Newton terms.
  // The adjoint of an equilibrium of this gradient is its exact derivative.
  bool exact_jac = false;
  bool skip_stencil_jac = false;  // with exact_jac: leave the barrier's b'(d) d(grad d)/dx out
  bool csr64 = false;  // the assembled CSR keeps FP64 values (Csr::values64) as well
  ShellElem<T> scaled(const ShellElem<T> &el) const {
    if (face_scale.empty()) return el;
    const T sc = face_scale[el.face];
    ShellElem<T> e = el;
    e.dm_inv = M2<T>(V2<T>(el.dm_inv.xa.x / sc, el.dm_inv.xa.y / sc), V2<T>(el.dm_inv.ya.x / sc, el.dm_inv.ya.y / sc));
    e.area = el.area * sc * sc;
    return e;
  }
  struct Rod { uint32_t a, b; T rest, k; };
  struct Spring { uint32_t a, b; T rest; };
  std::vector<Spring> springs;  // every unique triangle edge when stretch_k > 0
  struct RodHinge { uint32_t j, i, k; T rest, stiff; };
  std::vector<Rod> rods;
  std::vector<RodHinge> rod_hinges;
  std::vector<nguo_collider> colliders;
  std::vector<Edge> edges;
  std::vector<std::vector<uint32_t>> exclusion;
  std::vector<uint8_t> static_vert;  // empty = no static mesh collider
  std::vector<Tri> collider_tris;
  std::vector<V> collider_pn;
  T char_len = 1;
  // diagnostics of the last step
  double last_max_reresid = 0;
  uint32_t last_max_cg_iters = 0, last_newton = 0, last_num_stencils = 0;
  float last_toi = 1;
  bool last_penetration_free = true, last_all_finite = true;
  std::vector<std::vector<V>> newton_trace;  // x after every Newton iteration of the last step
  std::vector<T> toi_trace;

  explicit Solver(const nguo_scene_desc &s) {
    uint32_t n = s.n_verts;
    pos.resize(n); vel.resize(n); inv_mass.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      pos[i] = V(T(s.positions[3 * i]), T(s.positions[3 * i + 1]), T(s.positions[3 * i + 2]));
      if (s.velocities) vel[i] = V(T(s.velocities[3 * i]), T(s.velocities[3 * i + 1]), T(s.velocities[3 * i + 2]));
      inv_mass[i] = s.inv_mass[i];
    }
    prev = pos;
    prm = s.params;
    tris.resize(s.n_tris);
    for (uint32_t f = 0; f < s.n_tris; ++f) tris[f] = {s.tris[3 * f], s.tris[3 * f + 1], s.tris[3 * f + 2]};
    // implicit.rs:530-601 ImplicitSolver::new: hinges over ALL triangles, 1-ring exclusion,
    // unique edges, characteristic length.
    hinges = build_hinges<T>(tris, pos);
    exclusion.assign(n, {});
    for (uint32_t i = 0; i < n; ++i) exclusion[i].push_back(i);
    for (auto &t : tris)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          if (a != b) exclusion[t[a]].push_back(t[b]);
    for (auto &e : exclusion) { std::sort(e.begin(), e.end()); e.erase(std::unique(e.begin(), e.end()), e.end()); }
    {
      std::vector<Edge> es;
      for (auto &t : tris)
        for (auto ab : {std::make_pair(0, 1), std::make_pair(1, 2), std::make_pair(2, 0)}) {
          uint32_t i = t[ab.first], j = t[ab.second];
          es.push_back({std::min(i, j), std::max(i, j)});
        }
      std::sort(es.begin(), es.end());
      es.erase(std::unique(es.begin(), es.end()), es.end());
      edges = es;
      // state.rs:39-65 SimState::from_mesh: a stretch constraint per unique edge (sorted, as
      // geometry.rs:37-45 derive_unique_edges), rest length from the scene positions
      if (s.params.stretch_k > 0.0f)
        for (auto &e : es) springs.push_back({e[0], e[1], length(pos[e[0]] - pos[e[1]])});
    }
    // implicit.rs:1492-1499 set_shell_faces (min_area 1e-12).
    std::vector<Tri> sf(s.n_shell_faces);
    for (uint32_t f = 0; f < s.n_shell_faces; ++f) sf[f] = {s.shell_faces[3 * f], s.shell_faces[3 * f + 1], s.shell_faces[3 * f + 2]};
    shells = build_shells<T>(sf, pos, T(1.0e-12));
    // implicit.rs:873-893 set_rods + rebuild_collision_topology (rod segments appended to
    // the collision edges with offset 0 and the scene-wide gap).
    {
      std::vector<Edge> seen = edges;  // (min, max) keys, sorted
      for (uint32_t r = 0; r < s.n_rods; ++r) {
        Rod rd{s.rod_verts[2 * r], s.rod_verts[2 * r + 1], T(s.rod_rest[r]), T(s.rod_k[r])};
        rods.push_back(rd);
        Edge key{std::min(rd.a, rd.b), std::max(rd.a, rd.b)};
        auto it = std::lower_bound(seen.begin(), seen.end(), key);
        if (it != seen.end() && *it == key) continue;
        seen.insert(it, key);
        edges.push_back({rd.a, rd.b});
      }
    }
    for (uint32_t h = 0; h < s.n_rod_hinges; ++h)
      rod_hinges.push_back({s.rod_hinge_verts[3 * h], s.rod_hinge_verts[3 * h + 1], s.rod_hinge_verts[3 * h + 2],
                            T(s.rod_hinge_rest[h]), T(s.rod_hinge_stiff[h])});
    for (uint32_t c = 0; c < s.n_colliders; ++c) colliders.push_back(s.colliders[c]);
    if (edges.empty()) char_len = 1;
    else {
      T sum = 0;
      for (auto &e : edges) sum += length(pos[e[0]] - pos[e[1]]);
      char_len = std::fmax(sum / T(edges.size()), T(1e-6));
    }
    if (s.static_collider) set_static_collider_verts(s.static_collider);
  }

  // implicit.rs:1258-1322 set_static_collider_verts (winding check, pseudonormals).
  void set_static_collider_verts(const uint8_t *mask) {
    const T INF = std::numeric_limits<T>::infinity();
    V lo = splat(INF), hi = splat(-INF);
    std::vector<Tri> marked;
    for (auto &t : tris) {
      if (!(mask[t[0]] && mask[t[1]] && mask[t[2]])) continue;
      for (int k = 0; k < 3; ++k) { lo = vmin(lo, pos[t[k]]); hi = vmax(hi, pos[t[k]]); }
      marked.push_back(t);
    }
    if (marked.empty()) return;
    V e = hi - lo;
    double box_vol = (double)(e.x * e.y * e.z);
    V mid = T(0.5) * (lo + hi);
    double vol = 0;
    for (auto &t : marked) {
      V a = pos[t[0]] - mid, b = pos[t[1]] - mid, c = pos[t[2]] - mid;
      vol += (double)dot(cross(a, b), c) / 6.0;
    }
    if (vol < -0.05 * box_vol) return;
    static_vert.assign(mask, mask + pos.size());
    collider_tris = marked;
    collider_pn = pseudonormals<T>(pos, collider_tris);
  }

  // implicit.rs:1368-1398 set_kinematic_collider_positions: teleport the static collider's
  // vertices to x (only its rows are read), set their velocities to the move over dt,
  // accumulate the move into kinematic_dx and re-bake the pseudonormals. false (and nothing
  // changed) without a collider or with a non-finite position on a collider row.
  bool set_kinematic_collider_positions(const std::vector<V> &x) {
    if (static_vert.empty() || x.size() != pos.size()) return false;
    for (size_t i = 0; i < x.size(); ++i)
      if (static_vert[i] && !finite3(x[i])) return false;
    if (kinematic_dx.empty()) kinematic_dx.assign(x.size(), V{});
    const T inv_h = T(1) / T(prm.dt);
    for (size_t i = 0; i < x.size(); ++i) {
      if (!static_vert[i]) continue;
      const V d = x[i] - pos[i];
      kinematic_dx[i] += d;
      pos[i] = x[i];
      vel[i] = d * inv_h;
    }
    collider_pn = pseudonormals<T>(pos, collider_tris);
    return true;
  }

  bool movable(uint32_t v) const { return inv_mass[v] > 0.0f; }
  bool tri_movable(const Tri &t) const { return movable(t[0]) || movable(t[1]) || movable(t[2]); }
  bool excluded(uint32_t v, const Tri &t) const {
    for (uint32_t e : exclusion[v])
      if (e == t[0] || e == t[1] || e == t[2]) return true;
    return false;
  }
  bool is_collider_tri(const uint32_t idx[4]) const {
    return !static_vert.empty() && static_vert[idx[1]] && static_vert[idx[2]] && static_vert[idx[3]];
  }
  T dhat() const { return T(prm.contact_dhat); }

  // ---------------------------------------------------------------- candidate sets
  // Grid cell of a point: exact integer coordinates (no hashing, so no two cells share a list).
  struct Cell {
    int64_t a, b, c;
    bool operator==(const Cell &o) const { return a == o.a && b == o.b && c == o.c; }
  };
  struct CellHash {
    size_t operator()(const Cell &k) const {
      return (size_t)((uint64_t)(k.a * 73856093) ^ (uint64_t)(k.b * 19349663) ^ (uint64_t)(k.c * 83492791));
    }
  };

  // All (v, t) with t over `tl` whose vertex v lies in a grid cell that AABB(t) padded by
  // `pad` touches (cell = max(pad, mean triangle extent, 1e-6), keys floor(coordinate / cell)),
  // not excluded, and not both immovable (broadphase.rs:269-329), each pair once, in
  // ascending (t, v) order. The test is on cells, as the reference's, not on the padded box:
  // an unsigned gap below `pad` lies inside the box anyway, but a signed (negative) gap to a
  // collider facet does not, and such a facet is a candidate exactly when its cells are.
  template <class F> void for_each_vt(const std::vector<V> &x, const std::vector<Tri> &tl, T pad, F visit) const {
    if (x.empty() || tl.empty()) return;
    // grid over vertices with cell = max(pad, mean tri extent)
    T ext_sum = 0;
    for (auto &t : tl) {
      V a = x[t[0]], b = x[t[1]], c = x[t[2]];
      V d = vmax(vmax(a, b), c) - vmin(vmin(a, b), c);
      ext_sum += std::fmax(std::fmax(d.x, d.y), d.z);
    }
    T cell = std::fmax(std::fmax(pad, ext_sum / T(tl.size())), T(1e-6));
    T inv = T(1) / cell;
    auto key = [&](T v) { return (int64_t)std::floor(v * inv); };
    std::unordered_map<Cell, std::vector<uint32_t>, CellHash> grid;
    grid.reserve(x.size() * 2);
    for (uint32_t i = 0; i < x.size(); ++i) grid[Cell{key(x[i].x), key(x[i].y), key(x[i].z)}].push_back(i);
    std::vector<uint32_t> cand;
    for (uint32_t ti = 0; ti < tl.size(); ++ti) {
      const Tri &t = tl[ti];
      V a = x[t[0]], b = x[t[1]], c = x[t[2]];
      V lo = vmin(vmin(a, b), c) - splat(pad), hi = vmax(vmax(a, b), c) + splat(pad);
      bool tm = tri_movable(t);
      cand.clear();
      for (int64_t cx = key(lo.x); cx <= key(hi.x); ++cx)
        for (int64_t cy = key(lo.y); cy <= key(hi.y); ++cy)
          for (int64_t cz = key(lo.z); cz <= key(hi.z); ++cz) {
            auto it = grid.find(Cell{cx, cy, cz});
            if (it == grid.end()) continue;
            for (uint32_t v : it->second) {
              if (!tm && !movable(v)) continue;
              if (excluded(v, t)) continue;
              cand.push_back(v);
            }
          }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (uint32_t v : cand) visit(v, ti);
    }
  }

  // All (ei < ej) with AABB(ei) overlapping AABB(ej) padded by `pad`, no shared vertex, at
  // least one edge movable (broadphase.rs:394-522), sorted ascending, each once.
  std::vector<std::pair<uint32_t, uint32_t>> ee_pairs(const std::vector<V> &x, T pad) const {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    if (edges.size() < 2) return out;
    T len_sum = 0;
    for (auto &e : edges) len_sum += length(x[e[0]] - x[e[1]]);
    T cell = std::fmax(std::fmax(pad, len_sum / T(edges.size())), T(1e-6));
    T inv = T(1) / cell;
    auto key = [&](T v) { return (int64_t)std::floor(v * inv); };
    auto hk = [](int64_t a, int64_t b, int64_t c) {
      return (uint64_t)(a * 73856093) ^ (uint64_t)(b * 19349663) ^ (uint64_t)(c * 83492791);
    };
    auto emov = [&](uint32_t e) { return movable(edges[e][0]) || movable(edges[e][1]); };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    for (uint32_t ei = 0; ei < edges.size(); ++ei) {
      if (!emov(ei)) continue;
      V p = x[edges[ei][0]], q = x[edges[ei][1]];
      V lo = vmin(p, q), hi = vmax(p, q);
      for (int64_t cx = key(lo.x); cx <= key(hi.x); ++cx)
        for (int64_t cy = key(lo.y); cy <= key(hi.y); ++cy)
          for (int64_t cz = key(lo.z); cz <= key(hi.z); ++cz) grid[hk(cx, cy, cz)].push_back(ei);
    }
    for (uint32_t ej = 0; ej < edges.size(); ++ej) {
      const Edge &eb = edges[ej];
      V p = x[eb[0]], q = x[eb[1]];
      V blo = vmin(p, q) - splat(pad), bhi = vmax(p, q) + splat(pad);
      for (int64_t cx = key(blo.x); cx <= key(bhi.x); ++cx)
        for (int64_t cy = key(blo.y); cy <= key(bhi.y); ++cy)
          for (int64_t cz = key(blo.z); cz <= key(bhi.z); ++cz) {
            auto it = grid.find(hk(cx, cy, cz));
            if (it == grid.end()) continue;
            for (uint32_t ei : it->second) {
              if (ei == ej || (ei > ej && emov(ej))) continue;
              const Edge &ea = edges[ei];
              if (ea[0] == eb[0] || ea[0] == eb[1] || ea[1] == eb[0] || ea[1] == eb[1]) continue;
              V ap = x[ea[0]], aq = x[ea[1]];
              V alo = vmin(ap, aq), ahi = vmax(ap, aq);
              if (alo.x > bhi.x || alo.y > bhi.y || alo.z > bhi.z || ahi.x < blo.x || ahi.y < blo.y || ahi.z < blo.z) continue;
              out.push_back({std::min(ei, ej), std::max(ei, ej)});
            }
          }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }

  // implicit.rs:1431-1476 collider_sides: side from the closest collider facet (ties: lowest
  // facet index), +1 beyond dhat.
  std::vector<T> collider_sides(const std::vector<V> &x) const {
    if (collider_tris.empty()) return {};
    T dh = dhat();
    std::vector<T> best(x.size(), std::numeric_limits<T>::infinity());
    std::vector<T> side(x.size(), T(1));
    for_each_vt(x, collider_tris, dh, [&](uint32_t v, uint32_t ti) {
      if (static_vert[v]) return;
      const Tri &t = collider_tris[ti];
      V a = x[t[0]], b = x[t[1]], c = x[t[2]];
      T l[3];
      V q = closest_point_bary(x[v], a, b, c, l);
      V e = x[v] - q;
      T d2 = length_sq(e);
      if (d2 >= best[v]) return;
      best[v] = d2;
      V pn[3] = {collider_pn[t[0]], collider_pn[t[1]], collider_pn[t[2]]};
      V n = surface_normal_at(l, normalize_or_zero(cross(b - a, c - a)), pn);
      side[v] = dot(e, n) < T(0) ? T(-1) : T(1);
    });
    for (size_t i = 0; i < x.size(); ++i)
      if (best[i] >= dh * dh) side[i] = T(1);
    return side;
  }

  // implicit.rs:678-753 self_collision_stencils (sorted V-T pairs, then sorted E-E pairs).
  std::vector<Stencil<T>> stencils(const std::vector<V> &x) const {
    std::vector<Stencil<T>> out;
    T dh = dhat();
    std::vector<T> side = collider_sides(x);
    std::vector<std::pair<uint32_t, uint32_t>> vt;
    for_each_vt(x, tris, dh, [&](uint32_t v, uint32_t ti) { vt.push_back({v, ti}); });
    std::sort(vt.begin(), vt.end());
    for (auto &pr : vt) {
      const Tri &t = tris[pr.second];
      Stencil<T> s;
      s.idx[0] = pr.first; s.idx[1] = t[0]; s.idx[2] = t[1]; s.idx[3] = t[2];
      T d;
      if (is_collider_tri(s.idx)) {
        s.inside = side[s.idx[0]] < T(0) ? 1 : 0;
        d = signed_point_tri_distance_grad(x[s.idx[0]], x[s.idx[1]], x[s.idx[2]], x[s.idx[3]], side[s.idx[0]] < T(0), s.g);
      } else
        d = point_tri_distance_grad(x[s.idx[0]], x[s.idx[1]], x[s.idx[2]], x[s.idx[3]], s.g);
      if (d >= dh) continue;
      s.d = d; s.ghat = dh;
      out.push_back(s);
    }
    T ee_search = dh;  // max_edge_gap (= dhat, no per-edge gaps) + 2 * max_edge_offset (= 0)
    for (auto &pr : ee_pairs(x, ee_search)) {
      const Edge &ea = edges[pr.first], &eb = edges[pr.second];
      Stencil<T> s;
      s.idx[0] = ea[0]; s.idx[1] = ea[1]; s.idx[2] = eb[0]; s.idx[3] = eb[1];
      s.ee = 1;
      T d = segment_segment_distance_grad(x[s.idx[0]], x[s.idx[1]], x[s.idx[2]], x[s.idx[3]], s.g);
      T pair_ghat = T(0.5) * (dh + dh);
      if (d - T(0) >= pair_ghat) continue;
      s.d = d - T(0); s.ghat = pair_ghat;
      out.push_back(s);
    }
    return out;
  }

  // ---------------------------------------------------------------- assembly
  struct Blocks {
    std::unordered_map<uint64_t, uint32_t> at;
    std::vector<std::array<T, 9>> val;
    static uint64_t key(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }
    void push(uint32_t r, uint32_t c, T v) {  // scalar dof indices
      uint64_t k = key(r / 3, c / 3);
      auto it = at.find(k);
      uint32_t slot;
      if (it == at.end()) { slot = (uint32_t)val.size(); at.emplace(k, slot); val.push_back({}); }
      else slot = it->second;
      val[slot][3 * (r % 3) + (c % 3)] += v;
    }
    const std::array<T, 9> *get(uint32_t a, uint32_t b) const {
      auto it = at.find(key(a, b));
      return it == at.end() ? nullptr : &val[it->second];
    }
  };

  Blocks *keep_blocks = nullptr;  // non-null: the next assembly leaves its summed blocks here

  // implicit.rs:755-795 self_contact_rayleigh.
  T self_contact_rayleigh(const Blocks &blk, const Stencil<T> &s) const {
    T nlen = length(s.g[0]);
    if (nlen < T(1.0e-12)) return T(prm.contact_kappa);
    V nrm = s.g[0] / nlen;
    T weights[4], mass[4];
    T elastic[16][9];
    for (int a = 0; a < 4; ++a) {
      weights[a] = dot(s.g[a], nrm);
      float im = inv_mass[s.idx[a]];
      mass[a] = im > 0.0f ? T(1) / T(im) : T(0);
    }
    for (int ia = 0; ia < 4; ++ia)
      for (int ib = 0; ib < 4; ++ib) {
        const auto *b = blk.get(s.idx[ia], s.idx[ib]);
        for (int k = 0; k < 9; ++k) elastic[ia * 4 + ib][k] = b ? (*b)[k] : T(0);
      }
    T k = rayleigh_stiffness<T>(elastic, mass, weights, 4, s.d * nrm, T(0));
    return k > T(0) ? k : T(prm.contact_kappa);
  }

  // implicit.rs:2280-3155 assemble -> (Dirichlet-filtered CSR, rhs). `stencils_out` (may be
  // null) receives the active stencils used; `kappa_out` their stiffnesses; `kappa_in` (may
  // be null) supplies them instead of the Rayleigh stiffness of this assembly's blocks.
  Csr assemble(const std::vector<V> &x, const std::vector<V> &xt, T inv_h2, std::vector<float> &rhs_out,
               std::vector<Stencil<T>> *stencils_out = nullptr, std::vector<T> *kappa_out = nullptr,
               std::vector<T> *grad_out = nullptr, const std::vector<T> *kappa_in = nullptr) const {
    const uint32_t n = (uint32_t)x.size();
    std::vector<T> grad(3 * n, T(0));
    Blocks blk;
    auto push = [&](uint32_t r, uint32_t c, T v) { blk.push(r, c, v); };
    std::vector<Stencil<T>> sc = gn ? gn->st : (prm.self_collision ? stencils(x) : std::vector<Stencil<T>>{});
    if (gn)
      for (auto &s0 : sc) {
        if (exact_jac) stencil_eval(x, s0);  // gap and direction of x
        else s0.d = stencil_gap(x, s0);      // the gap moves, its direction stays
      }
    // inertia
    for (uint32_t i = 0; i < n; ++i) {
      float im = inv_mass[i];
      if (im <= 0.0f) continue;
      T m_over_h2 = (T(1) / T(im)) * inv_h2;
      V diff = x[i] - xt[i];
      for (int c = 0; c < 3; ++c) {
        push(3 * i + c, 3 * i + c, m_over_h2);
        grad[3 * i + c] += m_over_h2 * diff[c];
      }
    }
    // rod segments (implicit.rs:2407-2424)
    for (auto &rd : rods) {
      T hk[9];
      V ga;
      rod_stretch_block(x[rd.a] - x[rd.b], rd.rest, rd.k, hk, ga, project_elastic);
      for (int c = 0; c < 3; ++c) { grad[3 * rd.a + c] += ga[c]; grad[3 * rd.b + c] -= ga[c]; }
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) {
          T v = hk[3 * r + cc];
          push(3 * rd.a + r, 3 * rd.a + cc, v);
          push(3 * rd.b + r, 3 * rd.b + cc, v);
          push(3 * rd.a + r, 3 * rd.b + cc, -v);
          push(3 * rd.b + r, 3 * rd.a + cc, -v);
        }
    }
    // rod bending (implicit.rs:2426-2452)
    for (auto &h : rod_hinges) {
      if (h.stiff <= T(0) || !with_rod_hinges) continue;
      uint32_t v[3] = {h.j, h.i, h.k};
      V g[3];
      T hh[81];
      strand_force_hessian(x[v[0]], x[v[1]], x[v[2]], h.rest, g, hh);
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) grad[3 * v[a] + c] += h.stiff * g[a][c];
      if (exact_jac) {
        const V pv[3] = {x[v[0]], x[v[1]], x[v[2]]};
        double J[9][9];
        strand_jacobian(pv, h.rest, J);
        for (int a = 0; a < 9; ++a)
          for (int b = 0; b < 9; ++b) hh[a * 9 + b] = T(J[a][b]);
      }
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
              T val = h.stiff * hh[(3 * a + r) * 9 + 3 * b + cc];
              if (val != T(0)) push(3 * v[a] + r, 3 * v[b] + cc, val);
            }
    }
    // edge springs (implicit.rs:2487-2507)
    for (auto &sp : springs) {
      T hk[9];
      V ga;
      spring_block(x[sp.a] - x[sp.b], sp.rest, T(prm.stretch_k), hk, ga, project_elastic);
      for (int c = 0; c < 3; ++c) { grad[3 * sp.a + c] += ga[c]; grad[3 * sp.b + c] -= ga[c]; }
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) {
          T v = hk[3 * r + cc];
          push(3 * sp.a + r, 3 * sp.a + cc, v);
          push(3 * sp.b + r, 3 * sp.b + cc, v);
          push(3 * sp.a + r, 3 * sp.b + cc, -v);
          push(3 * sp.b + r, 3 * sp.a + cc, -v);
        }
    }
    // isotropic shells (implicit.rs:2608-2652)
    if (!shells.empty() && (prm.shell_mu > 0.0f || prm.shell_lambda > 0.0f)) {
      T smu = T(prm.shell_mu), slam = T(prm.shell_lambda);
      for (auto &el0 : shells) {
        const ShellElem<T> el = scaled(el0);
        V p[3] = {x[el.v[0]], x[el.v[1]], x[el.v[2]]};
        V eg[3];
        T eh[81];
        shell_gradient(p, el, prm.shell_model, smu, slam, eg);
        shell_hessian(p, el, prm.shell_model, smu, slam, T(prm.eiganalysis_eps), eh, project_elastic);
        for (int a = 0; a < 3; ++a) {
          uint32_t va = el.v[a];
          for (int c = 0; c < 3; ++c) grad[3 * va + c] += eg[a][c];
          for (int b = 0; b < 3; ++b) {
            uint32_t vb = el.v[b];
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc) {
                T v = eh[(3 * a + r) * 9 + (3 * b + cc)];
                if (v != T(0)) push(3 * va + r, 3 * vb + cc, v);
              }
          }
        }
      }
    }
    // hinge bending (implicit.rs:2697-2740)
    T kb = T(prm.bend_stiffness);
    if (kb > T(0)) {
      for (size_t hi = 0; hi < hinges.size(); ++hi) {
        const auto &hg = hinges[hi];
        V p[4] = {x[hg.v[0]], x[hg.v[1]], x[hg.v[2]], x[hg.v[3]]};
        T theta = dihedral_angle(p);
        V g[4];
        if (gn && !exact_jac) for (int a = 0; a < 4; ++a) g[a] = gn->hinge_g[hi][a];
        else dihedral_grad(p, g);
        T kwt = kb * hg.w;
        if (kwt <= T(0)) continue;
        T gval = theta - hg.rest;
        double dth[12], Jg[12][12];
        if (exact_jac) hinge_jacobian(p, dth, Jg);
        for (int a = 0; a < 4; ++a) {
          uint32_t va = hg.v[a];
          for (int c = 0; c < 3; ++c) grad[3 * va + c] += kwt * gval * g[a][c];
          for (int b = 0; b < 4; ++b) {
            uint32_t vb = hg.v[b];
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc) {
                T v = exact_jac ? kwt * (g[a][r] * T(dth[3 * b + cc]) + gval * T(Jg[3 * a + r][3 * b + cc])) : kwt * g[a][r] * g[b][cc];
                if (v != T(0)) push(3 * va + r, 3 * vb + cc, v);
              }
          }
        }
      }
    }
    // analytic colliders (implicit.rs:2954-3021; vertex_contact_full contact.rs:1082-1125)
    if (!colliders.empty()) {
      T dh = dhat(), kappa = T(prm.contact_kappa);
      for (uint32_t i = 0; i < n; ++i) {
        if (inv_mass[i] <= 0.0f) continue;
        T mu_i = T(prm.friction);
        V gsum{};
        T hsum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t ci = 0; ci < colliders.size(); ++ci) {
          const auto &c = colliders[ci];
          const size_t fi = i * colliders.size() + ci;
          T d;
          V nn;
          collider_distance(c, x[i], d, nn);
          if (gn && !exact_jac) nn = gn->col_n[fi];
          T k = kappa;  // contact_dynamic = 0 -> dynamic_stiffness returns the base kappa
          T bo[3];
          if (!barrier_of(prm.barrier_kind, d, dh, k, bo)) continue;
          V f_c = nn * bo[1];
          gsum += f_c;
          T hp = exact_jac ? bo[2] : std::fmax(bo[2], T(0));
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) hsum[3 * r + cc] += hp * nn[r] * nn[cc];
          if (exact_jac && c.kind != NGUO_COLLIDER_HALFSPACE) {
            // d n / d x of a sphere's outward normal: (I - n n^T) / |x - c|
            const T l = length(x[i] - V(T(c.a[0]), T(c.a[1]), T(c.a[2])));
            if (l >= T(1.0e-9))
              for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) hsum[3 * r + cc] += bo[1] * ((r == cc ? T(1) : T(0)) - nn[r] * nn[cc]) / l;
          }
          if (gn) {
            if (gn->col_fric[fi]) {
              const V u = x[i] - pos[i], n0 = gn->col_n[fi];
              gsum += (u - dot(n0, u) * n0) * gn->col_lam[fi];
              for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) hsum[3 * r + cc] += gn->col_lam[fi] * ((r == cc ? T(1) : T(0)) - n0[r] * n0[cc]);
            }
          } else if (mu_i > T(0)) {
            V fg;
            T fh[9];
            if (friction(f_c, x[i] - pos[i], nn, mu_i, T(prm.friction_eps), fg, fh)) {
              gsum += fg;
              for (int k2 = 0; k2 < 9; ++k2) hsum[k2] += fh[k2];
            }
          }
        }
        for (int c = 0; c < 3; ++c) grad[3 * i + c] += gsum[c];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc)
            if (hsum[3 * r + cc] != T(0)) push(3 * i + r, 3 * i + cc, hsum[3 * r + cc]);
      }
    }
    // self / body contact (implicit.rs:3026-3103)
    if (!sc.empty()) {
      Blocks snap = blk;
      size_t si = 0;
      for (auto &s : sc) {
        T k = gn ? gn->kappa[si] : (kappa_in ? (*kappa_in)[si] : self_contact_rayleigh(snap, s));
        const size_t fi = si;
        ++si;
        if (kappa_out) kappa_out->push_back(k);
        int kind = gn ? gn->kind[fi] : (s.d < T(0) ? BARRIER_CUBIC : prm.barrier_kind);
        T bo[3];
        if (!barrier_of(kind, s.d, s.ghat, k, bo)) continue;
        T bg = bo[1], hp = exact_jac ? bo[2] : std::fmax(bo[2], T(0));
        for (int a = 0; a < 4; ++a) {
          for (int c = 0; c < 3; ++c) grad[3 * s.idx[a] + c] += bg * s.g[a][c];
          for (int b = 0; b < 4; ++b)
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc) {
                T vh = hp * s.g[a][r] * s.g[b][cc];
                if (vh != T(0)) push(3 * s.idx[a] + r, 3 * s.idx[b] + cc, vh);
              }
        }
        if (exact_jac && !skip_stencil_jac && bg != T(0)) {
          double J[12][12];
          stencil_jacobian(x, s, J);
          for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
              for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) {
                  T vh = bg * T(J[3 * a + r][3 * b + cc]);
                  if (vh != T(0)) push(3 * s.idx[a] + r, 3 * s.idx[b] + cc, vh);
                }
        }
        T mu_s = T(prm.friction);
        if (gn) {
          if (gn->fric[fi]) {
            const V n0 = gn->fric_n[fi];
            const auto &w = gn->fric_w[fi];
            V slip{};
            for (int a = 0; a < 4; ++a) slip += (x[s.idx[a]] - pos[s.idx[a]] + body_dx(s.idx[a])) * w[a];
            const T lam = gn->fric_lam[fi];
            const V fg = (slip - dot(n0, slip) * n0) * lam;
            for (int a = 0; a < 4; ++a) {
              for (int c = 0; c < 3; ++c) grad[3 * s.idx[a] + c] += fg[c] * w[a];
              for (int b = 0; b < 4; ++b) {
                T wab = w[a] * w[b];
                if (wab == T(0)) continue;
                for (int r = 0; r < 3; ++r)
                  for (int cc = 0; cc < 3; ++cc) {
                    T vh = wab * lam * ((r == cc ? T(1) : T(0)) - n0[r] * n0[cc]);
                    if (vh != T(0)) push(3 * s.idx[a] + r, 3 * s.idx[b] + cc, vh);
                  }
              }
            }
          }
        } else if (mu_s > T(0)) {
          // implicit.rs:3068-3080: the slip of a kinematic body vertex includes its move
          V dxs[4];
          for (int a = 0; a < 4; ++a) dxs[a] = x[s.idx[a]] - pos[s.idx[a]] + body_dx(s.idx[a]);
          V fg[4];
          T w[4], fh[9];
          if (stencil_friction(s.g, dxs, bg, mu_s, T(prm.friction_eps), fg, w, fh)) {
            for (int a = 0; a < 4; ++a) {
              for (int c = 0; c < 3; ++c) grad[3 * s.idx[a] + c] += fg[a][c];
              for (int b = 0; b < 4; ++b) {
                T wab = w[a] * w[b];
                if (wab == T(0)) continue;
                for (int r = 0; r < 3; ++r)
                  for (int cc = 0; cc < 3; ++cc) {
                    T vh = wab * fh[3 * r + cc];
                    if (vh != T(0)) push(3 * s.idx[a] + r, 3 * s.idx[b] + cc, vh);
                  }
              }
            }
          }
        }
      }
    }
    if (stencils_out) *stencils_out = sc;
    if (grad_out) *grad_out = grad;
    if (keep_blocks) *keep_blocks = blk;
    // rhs = -grad, pinned rows zero; Dirichlet finish (csr.rs:385-416)
    rhs_out.assign(3 * n, 0.0f);
    for (uint32_t i = 0; i < 3 * n; ++i) rhs_out[i] = (float)(-grad[i]);
    for (uint32_t i = 0; i < n; ++i)
      if (inv_mass[i] <= 0.0f) rhs_out[3 * i] = rhs_out[3 * i + 1] = rhs_out[3 * i + 2] = 0.0f;
    // per block row: sorted block columns
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> rows(n);
    for (auto &kv : blk.at) rows[(uint32_t)(kv.first >> 32)].push_back({(uint32_t)(kv.first & 0xffffffffu), kv.second});
    Csr a;
    a.n = 3 * n;
    a.row_ptr.push_back(0);
    for (uint32_t i = 0; i < n; ++i) {
      std::sort(rows[i].begin(), rows[i].end());
      bool pin_i = inv_mass[i] <= 0.0f;
      for (int r = 0; r < 3; ++r) {
        if (pin_i) {
          a.col_idx.push_back(3 * i + r);
          a.values.push_back(1.0f);
          if (csr64) a.values64.push_back(1.0);
        } else {
          for (auto &cb : rows[i]) {
            if (inv_mass[cb.first] <= 0.0f) continue;
            for (int c = 0; c < 3; ++c) {
              a.col_idx.push_back(3 * cb.first + c);
              a.values.push_back((float)blk.val[cb.second][3 * r + c]);
              if (csr64) a.values64.push_back((double)blk.val[cb.second][3 * r + c]);
            }
          }
        }
        a.row_ptr.push_back((uint32_t)a.col_idx.size());
      }
    }
    return a;
  }

  static void collider_distance(const nguo_collider &c, V x, T &d, V &nrm) {
    if (c.kind == NGUO_COLLIDER_HALFSPACE) {
      V p(T(c.a[0]), T(c.a[1]), T(c.a[2])), nn(T(c.b[0]), T(c.b[1]), T(c.b[2]));
      d = dot(x - p, nn);
      nrm = nn;
    } else {
      V ctr(T(c.a[0]), T(c.a[1]), T(c.a[2]));
      V v = x - ctr;
      T l = length(v);
      if (l < T(1.0e-9)) { d = -T(c.radius); nrm = V(0, 1, 0); }
      else { d = l - T(c.radius); nrm = v / l; }
    }
  }

  // The gap and distance gradient of stencil s at positions x.
  void stencil_eval(const std::vector<V> &x, Stencil<T> &s) const {
    const V a = x[s.idx[0]], b = x[s.idx[1]], c = x[s.idx[2]], d = x[s.idx[3]];
    if (s.ee) s.d = segment_segment_distance_grad(a, b, c, d, s.g);
    else if (s.inside) s.d = signed_point_tri_distance_grad(a, b, c, d, true, s.g);
    else s.d = point_tri_distance_grad(a, b, c, d, s.g);
  }
  // J[3 a + r][3 b + c] = d g[a][r] / d x[idx[b]][c], g the distance gradient of stencil s at x
  // (forward mode through the same function, so exact inside its branch region).
  void stencil_jacobian(const std::vector<V> &x, const Stencil<T> &s, double J[12][12]) const {
    V3<double> p[4];
    for (int a = 0; a < 4; ++a) p[a] = V3<double>((double)x[s.idx[a]].x, (double)x[s.idx[a]].y, (double)x[s.idx[a]].z);
    V3<Jet> q[4], g[4];
    jet_points(p, 4, q);
    if (s.ee) segment_segment_distance_grad(q[0], q[1], q[2], q[3], g);
    else if (s.inside) signed_point_tri_distance_grad(q[0], q[1], q[2], q[3], true, g);
    else point_tri_distance_grad(q[0], q[1], q[2], q[3], g);
    for (int a = 0; a < 4; ++a)
      for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 12; ++j) J[3 * a + r][j] = g[a][r].g[j];
  }
  // Hinge: dth[j] = d theta / dx_j and Jg[3 a + r][j] = d g[a][r] / dx_j, with theta the dihedral
  // angle and g the analytic gradient the assembly uses.
  static void hinge_jacobian(const V p[4], double dth[12], double Jg[12][12]) {
    V3<double> pd[4];
    for (int a = 0; a < 4; ++a) pd[a] = V3<double>((double)p[a].x, (double)p[a].y, (double)p[a].z);
    V3<Jet> q[4], g[4];
    jet_points(pd, 4, q);
    const Jet th = dihedral_angle(q);
    dihedral_grad(q, g);
    for (int j = 0; j < 12; ++j) dth[j] = th.g[j];
    for (int a = 0; a < 4; ++a)
      for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 12; ++j) Jg[3 * a + r][j] = g[a][r].g[j];
  }
  // The strand's gradient (unit stiffness) as strand_force_hessian returns it, and its exact
  // Jacobian J[3 a + r][3 b + c] over the three vertices.
  template <class S> static void strand_grad(V3<S> x0, V3<S> x1, V3<S> x2, S rest, V3<S> grad[3]) {
    const V3<S> e0 = x0 - x1, e1 = x2 - x1;
    const S theta = std::atan2(length(cross(e0, e1)), dot(e0, e1));
    V3<S> gt[2];
    strand_gradient_theta(e0, e1, gt);
    const S w = theta - rest;
    grad[0] = w * gt[0];
    grad[1] = w * (-gt[0] - gt[1]);
    grad[2] = w * gt[1];
  }
  static void strand_jacobian(const V p[3], T rest, double J[9][9]) {
    V3<double> pd[3];
    for (int a = 0; a < 3; ++a) pd[a] = V3<double>((double)p[a].x, (double)p[a].y, (double)p[a].z);
    V3<Jet> q[3], g[3];
    jet_points(pd, 3, q);
    strand_grad(q[0], q[1], q[2], Jet((double)rest), g);
    for (int a = 0; a < 3; ++a)
      for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 9; ++j) J[3 * a + r][j] = g[a][r].g[j];
  }

  // The gap of stencil s at positions x (its distance function, signed for a collider facet
  // with the vertex inside), without the gradient.
  T stencil_gap(const std::vector<V> &x, const Stencil<T> &s) const {
    V g[4];
    const V a = x[s.idx[0]], b = x[s.idx[1]], c = x[s.idx[2]], d = x[s.idx[3]];
    if (s.ee) return segment_segment_distance_grad(a, b, c, d, g);
    if (s.inside) return signed_point_tri_distance_grad(a, b, c, d, true, g);
    return point_tri_distance_grad(a, b, c, d, g);
  }

  // The quantities an assembly at x_ref holds fixed (see Frozen).
  Frozen freeze(const std::vector<V> &x, const std::vector<V> &xt, T inv_h2) const {
    Frozen f;
    std::vector<float> rhs;
    (void)assemble(x, xt, inv_h2, rhs, &f.st, &f.kappa);
    const T eps = std::fmax(T(prm.friction_eps), T(1.0e-12)), mu = T(prm.friction);
    for (size_t i = 0; i < f.st.size(); ++i) {
      const Stencil<T> &s = f.st[i];
      const int kind = s.d < T(0) ? BARRIER_CUBIC : prm.barrier_kind;
      f.kind.push_back(kind);
      T bo[3] = {0, 0, 0};
      const bool act = barrier_of(kind, s.d, s.ghat, f.kappa[i], bo);
      V dxs[4], fg[4];
      T w[4] = {0, 0, 0, 0}, fh[9];
      for (int a = 0; a < 4; ++a) dxs[a] = x[s.idx[a]] - pos[s.idx[a]] + body_dx(s.idx[a]);
      const bool fr = act && mu > T(0) && stencil_friction(s.g, dxs, bo[1], mu, eps, fg, w, fh);
      V n{};
      T lam = 0;
      if (fr) {
        n = s.g[0] / length(s.g[0]);
        V slip{};
        for (int a = 0; a < 4; ++a) slip += dxs[a] * w[a];
        lam = mu * (-bo[1]) / std::fmax(length(slip - dot(n, slip) * n), eps);
      }
      f.fric.push_back(fr ? 1 : 0);
      f.fric_c.push_back(fr ? -mu * bo[1] : T(0));
      f.fric_n.push_back(n);
      f.fric_w.push_back({w[0], w[1], w[2], w[3]});
      f.fric_lam.push_back(lam);
    }
    for (auto &hg : hinges) {
      V p[4] = {x[hg.v[0]], x[hg.v[1]], x[hg.v[2]], x[hg.v[3]]}, g[4];
      dihedral_grad(p, g);
      f.hinge_g.push_back({g[0], g[1], g[2], g[3]});
    }
    const size_t nc = colliders.size();
    f.col_fric.assign(x.size() * nc, 0);
    f.col_n.assign(x.size() * nc, V{});
    f.col_c.assign(x.size() * nc, T(0));
    f.col_lam.assign(x.size() * nc, T(0));
    for (uint32_t i = 0; i < x.size(); ++i) {
      if (inv_mass[i] <= 0.0f) continue;
      for (size_t ci = 0; ci < nc; ++ci) {
        T d, bo[3];
        V nn, fg;
        T fh[9];
        collider_distance(colliders[ci], x[i], d, nn);
        const size_t fi = i * nc + ci;
        f.col_n[fi] = nn;
        if (!barrier_of(prm.barrier_kind, d, dhat(), T(prm.contact_kappa), bo)) continue;
        const V u = x[i] - pos[i];
        if (mu > T(0) && friction(nn * bo[1], u, nn, mu, eps, fg, fh)) {
          f.col_fric[fi] = 1;
          f.col_c[fi] = -mu * bo[1];
          f.col_lam[fi] = mu * (-bo[1]) / std::fmax(length(u - dot(nn, u) * nn), eps);
        }
      }
    }
    return f;
  }

  // Terms of the step energy (the incremental potential each Newton iteration decreases).
  enum EnergyTerm { E_INERTIA, E_SHELL, E_HINGE, E_ROD, E_ROD_BEND, E_BARRIER, E_FRICTION, E_COLLIDER, E_SPRING, E_TERMS };

  // The smoothed Coulomb potential whose gradient is contact.rs:613-643 friction with the
  // normal force, normal and weights held: c f0(|P u|), f0(s) = s^2 / (2 eps) up to eps, then
  // s - eps / 2 (c = mu times the normal force).
  static T friction_potential(T c, V n, V u, T eps) {
    const T s = length(u - dot(n, u) * n);
    return c * (s <= eps ? s * s / (T(2) * eps) : s - T(0.5) * eps);
  }

  // Step energy at x towards the inertial target xt (friction against the step's start pos),
  // with the contact quantities of `f`; out[E_TERMS] receives the terms. Every term is the one
  // whose gradient the assembly adds: inertia 1/2 m/h^2 |x - xt|^2 (implicit.rs:2392-2405),
  // rod stretch and edge springs 1/2 k (l - rest)^2 (implicit.rs:3287-3308, the rod tests'
  // energy; implicit.rs:1-10 the mass-spring potential), strand
  // bending 1/2 k (theta - rest)^2 (rod_bend.rs:1-20), shells area psi(sigma)
  // (shell_model.rs:112-127), hinges 1/2 k_b w (theta - rest)^2 (bending.rs:105-123), the
  // barriers (contact.rs:580-591, 1014-1025) and the friction potentials.
  T energy(const std::vector<V> &x, const std::vector<V> &xt, T inv_h2, const Frozen &f, T out[E_TERMS]) const {
    for (int k = 0; k < E_TERMS; ++k) out[k] = T(0);
    const uint32_t n = (uint32_t)x.size();
    for (uint32_t i = 0; i < n; ++i) {
      float im = inv_mass[i];
      if (im <= 0.0f) continue;
      out[E_INERTIA] += T(0.5) * (T(1) / T(im)) * inv_h2 * length_sq(x[i] - xt[i]);
    }
    for (auto &rd : rods) {
      const T e = length(x[rd.a] - x[rd.b]) - rd.rest;
      out[E_ROD] += T(0.5) * rd.k * e * e;
    }
    for (auto &sp : springs) {
      const T e = length(x[sp.a] - x[sp.b]) - sp.rest;
      out[E_SPRING] += T(0.5) * T(prm.stretch_k) * e * e;
    }
    for (auto &h : rod_hinges) {
      if (h.stiff <= T(0) || !with_rod_hinges) continue;
      const T w = strand_angle(x[h.j], x[h.i], x[h.k]) - h.rest;
      out[E_ROD_BEND] += T(0.5) * h.stiff * w * w;
    }
    if (!shells.empty() && (prm.shell_mu > 0.0f || prm.shell_lambda > 0.0f))
      for (auto &el0 : shells) {
        const ShellElem<T> el = scaled(el0);
        V p[3] = {x[el.v[0]], x[el.v[1]], x[el.v[2]]};
        out[E_SHELL] += shell_energy(p, el, prm.shell_model, T(prm.shell_mu), T(prm.shell_lambda));
      }
    const T kb = T(prm.bend_stiffness);
    if (kb > T(0))
      for (auto &hg : hinges) {
        const T kwt = kb * hg.w;
        if (kwt <= T(0)) continue;
        V p[4] = {x[hg.v[0]], x[hg.v[1]], x[hg.v[2]], x[hg.v[3]]};
        const T e = dihedral_angle(p) - hg.rest;
        out[E_HINGE] += T(0.5) * kwt * e * e;
      }
    const T eps = std::fmax(T(prm.friction_eps), T(1.0e-12));
    const size_t nc = colliders.size();
    for (uint32_t i = 0; i < n && nc; ++i) {
      if (inv_mass[i] <= 0.0f) continue;
      for (size_t ci = 0; ci < nc; ++ci) {
        T d, bo[3];
        V nn;
        collider_distance(colliders[ci], x[i], d, nn);
        if (barrier_of(prm.barrier_kind, d, dhat(), T(prm.contact_kappa), bo)) out[E_COLLIDER] += bo[0];
        const size_t fi = i * nc + ci;
        if (f.col_fric[fi]) out[E_COLLIDER] += friction_potential(f.col_c[fi], f.col_n[fi], x[i] - pos[i], eps);
      }
    }
    for (size_t si = 0; si < f.st.size(); ++si) {
      const Stencil<T> &s = f.st[si];
      T bo[3];
      if (barrier_of(f.kind[si], stencil_gap(x, s), s.ghat, f.kappa[si], bo)) out[E_BARRIER] += bo[0];
      if (f.fric[si]) {
        V slip{};
        for (int a = 0; a < 4; ++a) slip += (x[s.idx[a]] - pos[s.idx[a]] + body_dx(s.idx[a])) * f.fric_w[si][a];
        out[E_FRICTION] += friction_potential(f.fric_c[si], f.fric_n[si], slip, eps);
      }
    }
    T tot = 0;
    for (int k = 0; k < E_TERMS; ++k) tot += out[k];
    return tot;
  }

  // ---------------------------------------------------------------- CCD
  // implicit.rs:1690-1947 ccd_max_step (analytic colliders + widening V-T / E-E ACCD).
  T ccd_max_step(const std::vector<V> &x, const std::vector<V> &dx) const {
    bool has_self = prm.self_collision != 0;
    if (colliders.empty() && !has_self) return T(1);
    const T GAP = T(0.2), T_MIN = T(0);
    T t = T(1);
    const uint32_t n = (uint32_t)x.size();
    for (uint32_t i = 0; i < n; ++i) {
      if (inv_mass[i] <= 0.0f) continue;
      for (auto &c : colliders) {
        T d0;
        V nrm;
        collider_distance(c, x[i], d0, nrm);
        if (d0 <= T(0)) continue;
        if (c.kind == NGUO_COLLIDER_HALFSPACE) {
          T rate = dot(dx[i], nrm);
          if (rate < T(0)) t = std::fmin(t, (GAP - T(1)) * d0 / rate);
        } else {
          T l = length(dx[i]);
          if (l > T(1.0e-12)) t = std::fmin(t, (T(1) - GAP) * d0 / l);
        }
      }
    }
    if (has_self) {
      const T MAX_T = T(1.25), REDUCTION = T(0.01);
      uint32_t iters = std::max(prm.ccd_max_iter, 1u);
      T max_disp = 0;
      for (auto &d : dx) max_disp = std::fmax(max_disp, length(d));
      T fixed = dhat();
      T allowed = std::fmax(std::fmin(T(0.5) * char_len, max_disp), T(1.0e-9));
      std::vector<T> side = collider_sides(x);
      while (true) {
        bool full = allowed >= max_disp;
        T search = fixed + T(2) * allowed;
        T t_self = T(1);
        for_each_vt(x, tris, search, [&](uint32_t v, uint32_t ti) {
          const Tri &tr = tris[ti];
          uint32_t idx[4] = {v, tr[0], tr[1], tr[2]};
          if (is_collider_tri(idx) && side[v] < T(0)) return;
          V x0[4], dd[4];
          for (int k = 0; k < 4; ++k) { x0[k] = x[idx[k]]; dd[k] = dx[idx[k]]; }
          T toi = accd_toi(x0, dd, T(0), MAX_T, REDUCTION, iters, [](const V *p) {
            T d = point_triangle_distance(p[0], p[1], p[2], p[3]);
            return d * d;
          });
          t_self = std::fmin(t_self, toi / MAX_T);
        });
        for (auto &pr : ee_pairs(x, search)) {
          const Edge &ea = edges[pr.first], &eb = edges[pr.second];
          uint32_t idx[4] = {ea[0], ea[1], eb[0], eb[1]};
          V x0[4], dd[4];
          for (int k = 0; k < 4; ++k) { x0[k] = x[idx[k]]; dd[k] = dx[idx[k]]; }
          T toi = accd_toi(x0, dd, T(0), MAX_T, REDUCTION, iters, [](const V *p) {
            T d = segment_segment_distance(p[0], p[1], p[2], p[3]);
            return d * d;
          });
          t_self = std::fmin(t_self, toi / MAX_T);
        }
        if (full || t_self * max_disp <= allowed) { t = std::fmin(t, t_self); break; }
        allowed = std::fmin(allowed * T(4), max_disp);
      }
    }
    return clampv(t, T_MIN, T(1));
  }

  // implicit.rs:618-650 positions_self_intersect (edge pierces triangle, non-inert pairs).
  bool self_intersects(const std::vector<V> &x) const {
    if (edges.empty()) return false;
    bool hit = false;
    // every (edge, tri) pair whose AABBs overlap: query each triangle's AABB with the
    // edge endpoints' box; reuse the V-T grid by testing edges against triangle boxes.
    T len_sum = 0;
    for (auto &e : edges) len_sum += length(x[e[0]] - x[e[1]]);
    T cell = std::fmax(len_sum / T(edges.size()), T(1e-6));
    T inv = T(1) / cell;
    auto key = [&](T v) { return (int64_t)std::floor(v * inv); };
    auto hk = [](int64_t a, int64_t b, int64_t c) {
      return (uint64_t)(a * 73856093) ^ (uint64_t)(b * 19349663) ^ (uint64_t)(c * 83492791);
    };
    std::unordered_map<uint64_t, std::vector<uint32_t>> grid;
    for (uint32_t ei = 0; ei < edges.size(); ++ei) {
      V p = x[edges[ei][0]], q = x[edges[ei][1]];
      V lo = vmin(p, q), hi = vmax(p, q);
      for (int64_t cx = key(lo.x); cx <= key(hi.x); ++cx)
        for (int64_t cy = key(lo.y); cy <= key(hi.y); ++cy)
          for (int64_t cz = key(lo.z); cz <= key(hi.z); ++cz) grid[hk(cx, cy, cz)].push_back(ei);
    }
    std::vector<uint32_t> cand;
    for (uint32_t ti = 0; ti < tris.size() && !hit; ++ti) {
      const Tri &t = tris[ti];
      V a = x[t[0]], b = x[t[1]], c = x[t[2]];
      V lo = vmin(vmin(a, b), c), hi = vmax(vmax(a, b), c);
      bool tm = tri_movable(t);
      cand.clear();
      for (int64_t cx = key(lo.x); cx <= key(hi.x); ++cx)
        for (int64_t cy = key(lo.y); cy <= key(hi.y); ++cy)
          for (int64_t cz = key(lo.z); cz <= key(hi.z); ++cz) {
            auto it = grid.find(hk(cx, cy, cz));
            if (it != grid.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
          }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (uint32_t ei : cand) {
        const Edge &e = edges[ei];
        if (!tm && !movable(e[0]) && !movable(e[1])) continue;
        if (e[0] == t[0] || e[0] == t[1] || e[0] == t[2] || e[1] == t[0] || e[1] == t[1] || e[1] == t[2]) continue;
        if (edge_triangle_intersect(x[e[0]], x[e[1]], a, b, c)) { hit = true; break; }
      }
    }
    return hit;
  }

  // ---------------------------------------------------------------- one Newton iteration
  // implicit.rs:1979-2030 loop body. Returns the CCD step fraction; x is updated in place.
  T newton_iteration(std::vector<V> &x, const std::vector<V> &xt, T inv_h2, CgStats &cst, T *max_step_out) {
    std::vector<float> rhs;
    Csr a = assemble(x, xt, inv_h2, rhs);
    std::vector<float> dxf = cg_block_jacobi(a, rhs, prm.cg_tol, prm.cg_max_iter, &cst);
    const uint32_t n = (uint32_t)x.size();
    std::vector<V> dx(n);
    for (uint32_t i = 0; i < n; ++i) dx[i] = V(T(dxf[3 * i]), T(dxf[3 * i + 1]), T(dxf[3 * i + 2]));
    T max_norm = 0;
    for (auto &d : dx) max_norm = std::fmax(max_norm, length(d));
    if (max_norm > T(0)) {
      T rescale = std::fmin(T(prm.max_dx) / max_norm, T(1));
      if (rescale < T(1))
        for (auto &d : dx) d = d * rescale;
    }
    T t = ccd_max_step(x, dx);
    T max_step = 0;
    for (uint32_t i = 0; i < n; ++i) {
      V d = dx[i] * t;
      x[i] += d;
      max_step = std::fmax(max_step, length(d));
    }
    if (max_step_out) *max_step_out = max_step;
    return t;
  }

  // Adjoint of the last step's equilibrium (the drape adjoint's reference). At the converged
  // end state x* the step gradient vanishes, grad E_step(x*; s) = 0, so for a loss L(x*):
  //   dL/ds_f = -lambda^T d(grad E_step)/ds_f,  H lambda = dL/dx,
  // with H the Newton Hessian the step assembles at x* (inertial target and friction reference
  // of that step). Only the shells depend on s; their derivative is taken by central
  // differences of the element gradient in s (relative step 1e-6). dLdx is per vertex (3 n);
  // pinned rows are ignored. exact: unprojected shell / rod-segment blocks. xstar (3 n, may be
  // null) evaluates at that point instead of the step's own end state. Returns d L / d s per
  // scene shell face (nfaces).
  std::vector<double> adjoint(const std::vector<double> &dLdx, uint32_t nfaces, double cg_tol, CgStats *st, bool exact = false,
                              const double *xstar = nullptr) {
    const uint32_t n = (uint32_t)pos.size();
    const T h = T(prm.dt), inv_h2 = T(1) / (h * h);
    std::vector<V> xs = pos;
    if (xstar)
      for (uint32_t i = 0; i < n; ++i) xs[i] = V(T(xstar[3 * i]), T(xstar[3 * i + 1]), T(xstar[3 * i + 2]));
    std::vector<V> keep = pos;
    pos = prev;  // friction reference of the last step
    std::vector<float> rhs;
    std::vector<T> kap;
    Csr a;
    if (gn) {
      // the caller's held quantities and switches (the frozen-parameter model)
      a = assemble(xs, last_xt, inv_h2, rhs);
    } else {
      a = assemble(xs, last_xt, inv_h2, rhs, nullptr, &kap);
    }
    if (exact && !gn) {
      // contact stiffnesses stay those of the step (Rayleigh against the projected blocks)
      project_elastic = false;
      a = assemble(xs, last_xt, inv_h2, rhs, nullptr, nullptr, nullptr, &kap);
      project_elastic = true;
    }
    pos = keep;
    std::vector<double> b(3 * n, 0.0);
    for (uint32_t i = 0; i < n; ++i)
      if (inv_mass[i] > 0.0f)
        for (int c = 0; c < 3; ++c) b[3 * i + c] = dLdx[3 * i + c];
    uint32_t iters = 0;
    std::vector<double> lam = cg_block_jacobi64(a, b, cg_tol, 100000, &iters);
    if (st) {
      // FP64 relative residual of the solve
      double num = 0, den = 0;
      for (uint32_t row = 0; row < a.n; ++row) {
        double sacc = 0;
        for (uint32_t k = a.row_ptr[row]; k < a.row_ptr[row + 1]; ++k)
          sacc += (a.values64.empty() ? (double)a.values[k] : a.values64[k]) * lam[a.col_idx[k]];
        num += (b[row] - sacc) * (b[row] - sacc);
        den += b[row] * b[row];
      }
      st->iters = iters;
      st->residual = std::sqrt(num) / std::fmax(std::sqrt(den), 1e-300);
    }
    std::vector<double> out(nfaces, 0.0);
    if (!(prm.shell_mu > 0.0f || prm.shell_lambda > 0.0f)) return out;
    const T smu = T(prm.shell_mu), slam = T(prm.shell_lambda);
    for (auto &el0 : shells) {
      const T s0 = face_scale.empty() ? T(1) : face_scale[el0.face];
      const T ds = T(1e-6) * s0;
      V p[3] = {xs[el0.v[0]], xs[el0.v[1]], xs[el0.v[2]]};
      V gp[3], gm[3];
      ShellElem<T> ep = el0, em = el0;
      const T sp = s0 + ds, sm = s0 - ds;
      ep.dm_inv = M2<T>(V2<T>(el0.dm_inv.xa.x / sp, el0.dm_inv.xa.y / sp), V2<T>(el0.dm_inv.ya.x / sp, el0.dm_inv.ya.y / sp));
      ep.area = el0.area * sp * sp;
      em.dm_inv = M2<T>(V2<T>(el0.dm_inv.xa.x / sm, el0.dm_inv.xa.y / sm), V2<T>(el0.dm_inv.ya.x / sm, el0.dm_inv.ya.y / sm));
      em.area = el0.area * sm * sm;
      shell_gradient(p, ep, prm.shell_model, smu, slam, gp);
      shell_gradient(p, em, prm.shell_model, smu, slam, gm);
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) {
        const uint32_t v = el0.v[k];
        if (inv_mass[v] <= 0.0f) continue;
        for (int c = 0; c < 3; ++c) acc += lam[3 * v + c] * (double)((gp[k][c] - gm[k][c]) / (sp - sm));
      }
      out[el0.face] -= acc;
    }
    return out;
  }

  // The held quantities (Frozen) of an assembly at the last step's end state (or at xat, 3 n,
  // when given), with that step's inertial target and friction reference (its start).
  Frozen freeze_end(const double *xat = nullptr) {
    const T h = T(prm.dt);
    std::vector<V> keep = pos, x = pos;
    if (xat)
      for (size_t i = 0; i < x.size(); ++i) x[i] = V(T(xat[3 * i]), T(xat[3 * i + 1]), T(xat[3 * i + 2]));
    pos = prev;
    Frozen f = freeze(x, last_xt, T(1) / (h * h));
    pos = keep;
    return f;
  }

  // Newton on the assembled gradient under the caller's switches (meant for the
  // frozen-parameter model: gn with exact_jac and csr64), from x, with the last step's inertial
  // target and friction reference: the FP64 Jacobian solved by the FP64 block-Jacobi CG to
  // 1e-13, the step halved until the largest free gradient entry decreases. Stops once the
  // largest update is below tol or after max_iter updates; returns the updates taken and the
  // final largest free gradient entry in *res.
  uint32_t equilibrium(std::vector<V> &x, double tol, uint32_t max_iter, double *res) {
    const uint32_t n = (uint32_t)pos.size();
    const T h = T(prm.dt), inv_h2 = T(1) / (h * h);
    std::vector<V> keep = pos;
    pos = prev;
    auto gmax = [&](const std::vector<T> &g) {
      double m = 0;
      for (uint32_t i = 0; i < n; ++i)
        if (inv_mass[i] > 0.0f)
          for (int c = 0; c < 3; ++c) m = std::isfinite((double)g[3 * i + c]) ? std::fmax(m, std::fabs((double)g[3 * i + c])) : INFINITY;
      return m;
    };
    std::vector<float> rhs;
    std::vector<T> gr;
    Csr a = assemble(x, last_xt, inv_h2, rhs, nullptr, nullptr, &gr);
    double g0 = gmax(gr);
    uint32_t it = 0;
    while (it < max_iter) {
      std::vector<double> b(3 * n, 0.0);
      for (uint32_t i = 0; i < n; ++i)
        if (inv_mass[i] > 0.0f)
          for (int c = 0; c < 3; ++c) b[3 * i + c] = -(double)gr[3 * i + c];
      uint32_t cgi = 0;
      const std::vector<double> dx = cg_block_jacobi64(a, b, 1e-13, 100000, &cgi);
      double dmax = 0;
      for (double v : dx) dmax = std::fmax(dmax, std::fabs(v));
      double step = 1.0, gnew = g0;
      std::vector<V> xn;
      std::vector<T> grn;
      Csr an;
      for (int ls = 0; ls < 40; ++ls, step *= 0.5) {
        xn = x;
        for (uint32_t i = 0; i < n; ++i)
          if (inv_mass[i] > 0.0f)
            for (int c = 0; c < 3; ++c) xn[i][c] += T(step * dx[3 * i + c]);
        an = assemble(xn, last_xt, inv_h2, rhs, nullptr, nullptr, &grn);
        gnew = gmax(grn);
        if (gnew < g0 || step * dmax < tol) break;
      }
      x = xn;
      a = std::move(an);
      gr = std::move(grn);
      g0 = gnew;
      ++it;
      if (step * dmax < tol) break;
    }
    pos = keep;
    if (res) *res = g0;
    return it;
  }

  // implicit.rs:1951-2191 step.
  void step() {
    const uint32_t n = (uint32_t)pos.size();
    T h = T(prm.dt);
    T inv_h2 = T(1) / (h * h);
    V g(T(prm.gravity[0]), T(prm.gravity[1]), T(prm.gravity[2]));
    std::vector<V> x0 = pos, xt(n);
    for (uint32_t i = 0; i < n; ++i) xt[i] = inv_mass[i] > 0.0f ? x0[i] + vel[i] * h + g * (h * h) : x0[i];
    last_xt = xt;
    std::vector<V> x = x0;
    last_max_reresid = 0;
    last_max_cg_iters = 0;
    newton_trace.clear();
    toi_trace.clear();
    uint32_t cap = std::max(prm.newton_iters, 1u), taken = 0;
    bool final_step = false;
    for (uint32_t it = 0; it < cap; ++it) {
      taken++;
      CgStats cst;
      T max_step = 0;
      T t = newton_iteration(x, xt, inv_h2, cst, &max_step);
      last_max_reresid = std::fmax(last_max_reresid, cst.residual);
      last_max_cg_iters = std::max(last_max_cg_iters, cst.iters);
      last_toi = (float)t;
      newton_trace.push_back(x);
      toi_trace.push_back(t);
      if (max_step < T(1.0e-7)) break;
      if (final_step) break;
      if (prm.target_toi > 0.0f && t >= T(prm.target_toi) && taken > prm.min_newton_steps) final_step = true;
    }
    T floor = T(prm.floor_y);
    for (uint32_t i = 0; i < n; ++i)
      if (inv_mass[i] > 0.0f && std::isfinite(floor) && x[i].y < floor) x[i].y = floor;
    bool fin = true;
    for (auto &p : x) fin = fin && finite3(p);
    T inv_h = T(1) / h;
    for (uint32_t i = 0; i < n; ++i) {
      if (fin) {
        vel[i] = inv_mass[i] > 0.0f ? (x[i] - x0[i]) * inv_h : V{};
        prev[i] = x0[i];
        pos[i] = x[i];
      } else {
        vel[i] = V{};
        prev[i] = x0[i];
        pos[i] = x0[i];
      }
    }
    time += (double)h;
    // implicit.rs:2065-2070: the collider's move described this step only
    for (auto &d : kinematic_dx) d = V{};
    last_newton = taken;
    last_all_finite = fin;
    // implicit.rs:2235-2266 penetration_free: analytic colliders, floor, edge-triangle pierce.
    bool pf = fin;
    T tol = -T(1e-3) * std::fmax(dhat(), T(1e-6));
    for (uint32_t i = 0; i < n && pf; ++i) {
      if (inv_mass[i] <= 0.0f) continue;
      for (auto &c : colliders) {
        T d;
        V nn;
        collider_distance(c, pos[i], d, nn);
        if (d < tol) pf = false;
      }
      if (std::isfinite(floor) && pos[i].y < floor - T(1e-3)) pf = false;
    }
    if (pf && prm.self_collision) pf = !self_intersects(pos);
    last_penetration_free = pf;
  }
};

}  // namespace oracle
