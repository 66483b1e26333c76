// C entry points of the oracle library (liboracle.so), loaded by the tests through ctypes.
// TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference that the HIP path is
// checked against. Nothing in the product library links or calls this.
#include <cstring>
#include <memory>

#include "obvh.h"
#include "osolver.h"

using namespace oracle;

namespace {
struct Handle {
  int fp64;
  std::unique_ptr<Solver<float>> s32;
  std::unique_ptr<Solver<double>> s64;
  std::unique_ptr<Solver<float>::Frozen> fz32;
  std::unique_ptr<Solver<double>::Frozen> fz64;
  Csr csr;
  std::vector<float> rhs;
};
template <class T> std::vector<V3<T>> load(const float *x, uint32_t n) {
  std::vector<V3<T>> v(n);
  for (uint32_t i = 0; i < n; ++i) v[i] = V3<T>(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
  return v;
}
template <class T> void store(const std::vector<V3<T>> &v, float *out) {
  for (size_t i = 0; i < v.size(); ++i) { out[3 * i] = (float)v[i].x; out[3 * i + 1] = (float)v[i].y; out[3 * i + 2] = (float)v[i].z; }
}
template <class T> void store_d(const std::vector<V3<T>> &v, double *out) {
  for (size_t i = 0; i < v.size(); ++i) { out[3 * i] = (double)v[i].x; out[3 * i + 1] = (double)v[i].y; out[3 * i + 2] = (double)v[i].z; }
}
template <class T> V3<T> v3(const double *p) { return V3<T>(T(p[0]), T(p[1]), T(p[2])); }
}  // namespace

#define DISPATCH(h, CALL) ((h)->fp64 ? ((h)->s64->CALL) : ((h)->s32->CALL))

extern "C" {

void *oracle_solver_create(const nguo_scene_desc *s, int fp64) {
  auto *h = new Handle();
  h->fp64 = fp64;
  if (fp64) h->s64.reset(new Solver<double>(*s));
  else h->s32.reset(new Solver<float>(*s));
  return h;
}
void oracle_solver_destroy(void *p) { delete (Handle *)p; }

uint32_t oracle_solver_nverts(void *p) {
  auto *h = (Handle *)p;
  return (uint32_t)(h->fp64 ? h->s64->pos.size() : h->s32->pos.size());
}

void oracle_solver_step(void *p, uint32_t nsteps) {
  auto *h = (Handle *)p;
  for (uint32_t k = 0; k < nsteps; ++k) {
    if (h->fp64) h->s64->step(); else h->s32->step();
  }
}

void oracle_solver_positions(void *p, float *out) {
  auto *h = (Handle *)p;
  if (h->fp64) store(h->s64->pos, out); else store(h->s32->pos, out);
}
void oracle_solver_positions_d(void *p, double *out) {
  auto *h = (Handle *)p;
  if (h->fp64) store_d(h->s64->pos, out); else store_d(h->s32->pos, out);
}
void oracle_solver_velocities(void *p, float *out) {
  auto *h = (Handle *)p;
  if (h->fp64) store(h->s64->vel, out); else store(h->s32->vel, out);
}
void oracle_solver_set_state(void *p, const float *x, const float *v) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  if (h->fp64) { h->s64->pos = load<double>(x, n); if (v) h->s64->vel = load<double>(v, n); }
  else { h->s32->pos = load<float>(x, n); if (v) h->s32->vel = load<float>(v, n); }
}

// Per shell face rest scale (nfaces = the scene's shell face count).
void oracle_solver_set_rest_scale(void *p, const double *scale, uint32_t nfaces) {
  auto *h = (Handle *)p;
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    s.face_scale.assign(nfaces, T(1));
    for (uint32_t f = 0; f < nfaces; ++f) s.face_scale[f] = T(scale[f]);
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
}

// Adjoint of the last step's equilibrium: dLds[f] for every shell face given dL/dx (3 n).
// exact: unprojected shell / rod-segment blocks; xstar (may be null): evaluation point.
uint32_t oracle_solver_adjoint(void *p, const double *dLdx, uint32_t nfaces, double cg_tol, int exact, const double *xstar, double *dLds,
                               double *residual) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  std::vector<double> g(dLdx, dLdx + 3 * (size_t)n);
  CgStats st;
  std::vector<double> out = h->fp64 ? h->s64->adjoint(g, nfaces, cg_tol, &st, exact != 0, xstar)
                                    : h->s32->adjoint(g, nfaces, cg_tol, &st, exact != 0, xstar);
  for (uint32_t f = 0; f < nfaces; ++f) dLds[f] = out[f];
  if (residual) *residual = st.residual;
  return st.iters;
}

// out: time, max_reresid, max_cg_iters, newton_iters, last_toi, penetration_free, all_finite, n_edges
void oracle_solver_metrics(void *p, double *out) {
  auto *h = (Handle *)p;
  auto fill = [&](auto &s) {
    out[0] = s.time; out[1] = s.last_max_reresid; out[2] = s.last_max_cg_iters; out[3] = s.last_newton;
    out[4] = s.last_toi; out[5] = s.last_penetration_free; out[6] = s.last_all_finite; out[7] = (double)s.edges.size();
  };
  if (h->fp64) fill(*h->s64); else fill(*h->s32);
}

// Positions after Newton iteration `k` of the last step (and its CCD fraction).
int oracle_solver_newton_trace(void *p, uint32_t k, float *out, double *toi) {
  auto *h = (Handle *)p;
  auto get = [&](auto &s) -> int {
    if (k >= s.newton_trace.size()) return -1;
    store(s.newton_trace[k], out);
    if (toi) *toi = (double)s.toi_trace[k];
    return (int)s.newton_trace.size();
  };
  return h->fp64 ? get(*h->s64) : get(*h->s32);
}

// Active contact stencils at x (sorted V-T then sorted E-E). Returns the count; writes up to
// cap entries: idx [4*cap], d [cap], g [12*cap], kind [cap] (0 = V-T, 1 = E-E).
int64_t oracle_solver_stencils(void *p, const float *x, uint32_t cap, uint32_t *idx, float *d, float *g) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> int64_t {
    using T = decltype(tag);
    auto st = s.stencils(load<T>(x, n));
    for (size_t k = 0; k < st.size() && k < cap; ++k) {
      for (int a = 0; a < 4; ++a) idx[4 * k + a] = st[k].idx[a];
      d[k] = (float)st[k].d;
      if (g)
        for (int a = 0; a < 4; ++a)
          for (int c = 0; c < 3; ++c) g[12 * k + 3 * a + c] = (float)st[k].g[a][c];
    }
    return (int64_t)st.size();
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

// implicit.rs:1368-1398 set_kinematic_collider_positions (x: all vertices, f32); 1 = applied.
int oracle_solver_set_kinematic(void *p, const float *x) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  return h->fp64 ? (h->s64->set_kinematic_collider_positions(load<double>(x, n)) ? 1 : 0)
                 : (h->s32->set_kinematic_collider_positions(load<float>(x, n)) ? 1 : 0);
}

// Collider side (+1 / -1) per vertex at x; returns 0 when the scene has no mesh collider.
int oracle_solver_collider_sides(void *p, const float *x, float *side) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> int {
    using T = decltype(tag);
    auto sd = s.collider_sides(load<T>(x, n));
    if (sd.empty()) return 0;
    for (uint32_t i = 0; i < n; ++i) side[i] = (float)sd[i];
    return 1;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

// Assemble the Newton system at (x, x_tilde); kept in the handle. Returns nnz.
int64_t oracle_solver_assemble(void *p, const float *x, const float *xt, double inv_h2, float *kappa_out, uint32_t kcap) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    std::vector<T> kap;
    h->csr = s.assemble(load<T>(x, n), load<T>(xt, n), T(inv_h2), h->rhs, nullptr, &kap, nullptr);
    if (kappa_out)
      for (size_t k = 0; k < kap.size() && k < kcap; ++k) kappa_out[k] = (float)kap[k];
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
  return (int64_t)h->csr.values.size();
}
void oracle_solver_csr(void *p, uint32_t *row_ptr, uint32_t *col_idx, float *values, float *rhs) {
  auto *h = (Handle *)p;
  std::memcpy(row_ptr, h->csr.row_ptr.data(), h->csr.row_ptr.size() * 4);
  std::memcpy(col_idx, h->csr.col_idx.data(), h->csr.col_idx.size() * 4);
  std::memcpy(values, h->csr.values.data(), h->csr.values.size() * 4);
  std::memcpy(rhs, h->rhs.data(), h->rhs.size() * 4);
}

// Elastic+contact gradient at x (f64 out), for energy/gradient checks.
void oracle_solver_gradient(void *p, const float *x, const float *xt, double inv_h2, double *grad) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    std::vector<float> rhs;
    std::vector<T> gr;
    s.assemble(load<T>(x, n), load<T>(xt, n), T(inv_h2), rhs, nullptr, nullptr, &gr);
    for (size_t i = 0; i < gr.size(); ++i) grad[i] = (double)gr[i];
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
}

// As oracle_solver_gradient with FP64 positions and inertial target (derivative checks).
void oracle_solver_gradient_d(void *p, const double *x, const double *xt, double inv_h2, double *grad) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    std::vector<V3<T>> xv(n), tv(n);
    for (uint32_t i = 0; i < n; ++i) {
      xv[i] = V3<T>(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
      tv[i] = V3<T>(T(xt[3 * i]), T(xt[3 * i + 1]), T(xt[3 * i + 2]));
    }
    std::vector<float> rhs;
    std::vector<T> gr;
    s.assemble(xv, tv, T(inv_h2), rhs, nullptr, nullptr, &gr);
    for (size_t i = 0; i < gr.size(); ++i) grad[i] = (double)gr[i];
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
}

// ---------------------------------------------------------------- step energy
}  // extern "C"
namespace {
template <class T> std::vector<V3<T>> load_d(const double *x, uint32_t n) {
  std::vector<V3<T>> v(n);
  for (uint32_t i = 0; i < n; ++i) v[i] = V3<T>(T(x[3 * i]), T(x[3 * i + 1]), T(x[3 * i + 2]));
  return v;
}
template <class T> typename Solver<T>::Frozen *frozen_of(Handle *h);
template <> Solver<float>::Frozen *frozen_of<float>(Handle *h) { return h->fz32.get(); }
template <> Solver<double>::Frozen *frozen_of<double>(Handle *h) { return h->fz64.get(); }
}  // namespace
extern "C" {

// Holds the contact quantities of an assembly at x (Solver::Frozen) for the calls below.
void oracle_solver_freeze(void *p, const double *x, const double *xt, double inv_h2) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  if (h->fp64) h->fz64.reset(new Solver<double>::Frozen(h->s64->freeze(load_d<double>(x, n), load_d<double>(xt, n), inv_h2)));
  else h->fz32.reset(new Solver<float>::Frozen(h->s32->freeze(load_d<float>(x, n), load_d<float>(xt, n), (float)inv_h2)));
}

// Step energy at x (terms[8]: inertia, shells, hinges, rods, rod bends, barriers, friction,
// analytic colliders). use_frozen: with the held quantities; else with those of x itself.
double oracle_solver_energy(void *p, const double *x, const double *xt, double inv_h2, int use_frozen, double *terms) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> double {
    using T = decltype(tag);
    auto xv = load_d<T>(x, n), tv = load_d<T>(xt, n);
    T out[Solver<T>::E_TERMS];
    T e;
    if (use_frozen) {
      auto *f = frozen_of<T>(h);
      if (!f) return std::nan("");
      e = s.energy(xv, tv, T(inv_h2), *f, out);
    } else {
      e = s.energy(xv, tv, T(inv_h2), s.freeze(xv, tv, T(inv_h2)), out);
    }
    if (terms)
      for (int k = 0; k < Solver<T>::E_TERMS; ++k) terms[k] = (double)out[k];
    return (double)e;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

}  // extern "C"
namespace {
// Assembly switches of the derivative checks: 1 = the frozen-direction gradient model of the
// held quantities, 2 = unprojected shell and rod-segment blocks, 4 = no strand bending terms,
// 8 (with 1) = the frozen-parameter model instead (directions of x, exact Jacobian, FP64 CSR),
// 16 (with 9) = that Jacobian without the barrier's b'(d) d(grad d)/dx blocks.
template <class S> struct Switches {
  S &s;
  Switches(S &solver, const typename S::Frozen *f, int flags) : s(solver) {
    s.gn = (flags & 1) ? f : nullptr;
    s.project_elastic = !(flags & 2);
    s.with_rod_hinges = !(flags & 4);
    s.exact_jac = (flags & 9) == 9;
    s.csr64 = (flags & 9) == 9;
    s.skip_stencil_jac = (flags & 25) == 25;
  }
  ~Switches() {
    s.gn = nullptr; s.project_elastic = true; s.with_rod_hinges = true; s.keep_blocks = nullptr;
    s.exact_jac = false; s.csr64 = false; s.skip_stencil_jac = false;
  }
};
}  // namespace
extern "C" {

// The held contact quantities (oracle_solver_freeze) per stencil, for building the contact
// operator outside: idx [4], gap, distance gradient [12], Rayleigh stiffness, barrier kind,
// ghat, friction flag, lambda, normal [3], weights [4]. Returns the stencil count (arrays of
// at least cap entries; nothing written beyond cap).
uint32_t oracle_solver_frozen_stencils(void *p, uint32_t cap, uint32_t *idx, double *d, double *g12, double *kappa, int32_t *kind,
                                       double *ghat, int32_t *fric, double *lam, double *n3, double *w4) {
  auto *h = (Handle *)p;
  auto go = [&](auto *f) -> uint32_t {
    if (!f) return 0;
    const uint32_t ns = (uint32_t)f->st.size();
    for (uint32_t i = 0; i < ns && i < cap; ++i) {
      const auto &s = f->st[i];
      for (int a = 0; a < 4; ++a) {
        idx[4 * i + a] = s.idx[a];
        for (int c = 0; c < 3; ++c) g12[12 * i + 3 * a + c] = (double)s.g[a][c];
        w4[4 * i + a] = (double)f->fric_w[i][a];
      }
      d[i] = (double)s.d; kappa[i] = (double)f->kappa[i]; kind[i] = f->kind[i]; ghat[i] = (double)s.ghat;
      fric[i] = f->fric[i]; lam[i] = (double)f->fric_lam[i];
      for (int c = 0; c < 3; ++c) n3[3 * i + c] = (double)f->fric_n[i][c];
    }
    return ns;
  };
  return h->fp64 ? go(h->fz64.get()) : go(h->fz32.get());
}

// Assembled gradient at x under the switches.
int oracle_solver_model_gradient(void *p, const double *x, const double *xt, double inv_h2, int flags, double *grad) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> int {
    using T = decltype(tag);
    auto *f = frozen_of<T>(h);
    if ((flags & 1) && !f) return -1;
    Switches<Solver<T>> sw(s, f, flags);
    std::vector<float> rhs;
    std::vector<T> gr;
    s.assemble(load_d<T>(x, n), load_d<T>(xt, n), T(inv_h2), rhs, nullptr, nullptr, &gr);
    for (size_t i = 0; i < gr.size(); ++i) grad[i] = (double)gr[i];
    return 0;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

// hu = H u with H the assembled Hessian at x under the switches (summed blocks, before the
// Dirichlet filter and the f32 rounding of the CSR); rows and columns of pinned vertices are
// left out (their entries of hu are 0).
int oracle_solver_hess_vec(void *p, const double *x, const double *xt, double inv_h2, int flags, const double *u, double *hu) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> int {
    using T = decltype(tag);
    auto *f = frozen_of<T>(h);
    if ((flags & 1) && !f) return -1;
    Switches<Solver<T>> sw(s, f, flags);
    typename Solver<T>::Blocks blk;
    s.keep_blocks = &blk;
    std::vector<float> rhs;
    s.assemble(load_d<T>(x, n), load_d<T>(xt, n), T(inv_h2), rhs);
    for (uint32_t i = 0; i < 3 * n; ++i) hu[i] = 0.0;
    for (auto &kv : blk.at) {
      const uint32_t r = (uint32_t)(kv.first >> 32), c = (uint32_t)(kv.first & 0xffffffffu);
      if (s.inv_mass[r] <= 0.0f || s.inv_mass[c] <= 0.0f) continue;
      const auto &v = blk.val[kv.second];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) hu[3 * r + a] += (double)v[3 * a + b] * u[3 * c + b];
    }
    return 0;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

// Holds the quantities of an assembly at the last step's end state (x null) or at x [3 n],
// with that step's inertial target and its start as the friction reference, for the
// frozen-parameter calls below.
void oracle_solver_freeze_end(void *p, const double *x) {
  auto *h = (Handle *)p;
  if (h->fp64) h->fz64.reset(new Solver<double>::Frozen(h->s64->freeze_end(x)));
  else h->fz32.reset(new Solver<float>::Frozen(h->s32->freeze_end(x)));
}

// Equilibrium of the frozen-parameter gradient (held quantities of oracle_solver_freeze_end,
// the current rest scales, the last step's inertial target and friction reference) by Newton
// from x_in (null: the state's positions). x_out [3 n]; res: the final largest free gradient
// entry. Returns the Newton updates taken, -1 without held quantities.
int oracle_solver_equilibrium(void *p, const double *x_in, double tol, uint32_t max_iter, double *x_out, double *res) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> int {
    using T = decltype(tag);
    auto *f = frozen_of<T>(h);
    if (!f) return -1;
    Switches<Solver<T>> sw(s, f, 1 | 2 | 8);
    std::vector<V3<T>> x = x_in ? load_d<T>(x_in, n) : s.pos;
    int it = (int)s.equilibrium(x, tol, max_iter, res);
    for (uint32_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) x_out[3 * i + c] = (double)x[i][c];
    return it;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

// The drape adjoint of the frozen-parameter equilibrium at xeval [3 n] (the exact Jacobian
// of the held-quantity gradient, FP64 CG to cg_tol): dLds [nfaces]. Returns the CG
// iterations, or -1 without held quantities.
int oracle_solver_adjoint_frozen(void *p, const double *dLdx, uint32_t nfaces, double cg_tol, const double *xeval, double *dLds,
                                 double *residual) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  std::vector<double> g(dLdx, dLdx + 3 * (size_t)n);
  auto go = [&](auto &s, auto tag) -> int {
    using T = decltype(tag);
    auto *f = frozen_of<T>(h);
    if (!f) return -1;
    Switches<Solver<T>> sw(s, f, 1 | 2 | 8);
    CgStats st;
    std::vector<double> out = s.adjoint(g, nfaces, cg_tol, &st, true, xeval);
    for (uint32_t k = 0; k < nfaces; ++k) dLds[k] = out[k];
    if (residual) *residual = st.residual;
    return (int)st.iters;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

double oracle_solver_ccd(void *p, const float *x, const float *dx) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  if (h->fp64) return h->s64->ccd_max_step(load<double>(x, n), load<double>(dx, n));
  return (double)h->s32->ccd_max_step(load<float>(x, n), load<float>(dx, n));
}

int oracle_solver_self_intersects(void *p, const float *x) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  if (h->fp64) return h->s64->self_intersects(load<double>(x, n)) ? 1 : 0;
  return h->s32->self_intersects(load<float>(x, n)) ? 1 : 0;
}

// One Newton iteration from x (in/out) towards x_tilde. Returns the CCD fraction.
double oracle_solver_newton_iteration(void *p, float *x, const float *xt, double *residual, uint32_t *iters) {
  auto *h = (Handle *)p;
  uint32_t n = oracle_solver_nverts(p);
  auto go = [&](auto &s, auto tag) -> double {
    using T = decltype(tag);
    auto xv = load<T>(x, n);
    T hh = T(s.prm.dt);
    CgStats cst;
    T t = s.newton_iteration(xv, load<T>(xt, n), T(1) / (hh * hh), cst, nullptr);
    store(xv, x);
    if (residual) *residual = cst.residual;
    if (iters) *iters = cst.iters;
    return (double)t;
  };
  return h->fp64 ? go(*h->s64, 0.0) : go(*h->s32, 0.0f);
}

void oracle_solver_x_tilde(void *p, float *out) {
  auto *h = (Handle *)p;
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    T hh = T(s.prm.dt);
    V3<T> g(T(s.prm.gravity[0]), T(s.prm.gravity[1]), T(s.prm.gravity[2]));
    std::vector<V3<T>> xt(s.pos.size());
    for (size_t i = 0; i < s.pos.size(); ++i) xt[i] = s.inv_mass[i] > 0.0f ? s.pos[i] + s.vel[i] * hh + g * (hh * hh) : s.pos[i];
    store(xt, out);
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
}

void oracle_solver_x_tilde_d(void *p, double *out) {
  auto *h = (Handle *)p;
  auto go = [&](auto &s, auto tag) {
    using T = decltype(tag);
    T hh = T(s.prm.dt);
    V3<T> g(T(s.prm.gravity[0]), T(s.prm.gravity[1]), T(s.prm.gravity[2]));
    std::vector<V3<T>> xt(s.pos.size());
    for (size_t i = 0; i < s.pos.size(); ++i) xt[i] = s.inv_mass[i] > 0.0f ? s.pos[i] + s.vel[i] * hh + g * (hh * hh) : s.pos[i];
    store_d(xt, out);
  };
  if (h->fp64) go(*h->s64, 0.0); else go(*h->s32, 0.0f);
}

// ---------------------------------------------------------------- linear solve
void oracle_cg(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *b, float tol,
               uint32_t max_iter, float *x_out, uint32_t *iters, double *residual) {
  Csr a;
  a.n = n;
  a.row_ptr.assign(row_ptr, row_ptr + n + 1);
  a.col_idx.assign(col_idx, col_idx + row_ptr[n]);
  a.values.assign(values, values + row_ptr[n]);
  std::vector<float> bv(b, b + n);
  CgStats st;
  auto x = cg_block_jacobi(a, bv, tol, max_iter, &st);
  std::memcpy(x_out, x.data(), n * 4);
  if (iters) *iters = st.iters;
  if (residual) *residual = st.residual;
}

// CG with the one-level Schwarz preconditioner (linsolve.rs:463-520 SchwarzCpu semantics).
void oracle_cg_schwarz(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *b, float tol,
                       uint32_t max_iter, uint32_t domain, double weight, float *x_out, uint32_t *iters, double *residual) {
  Csr a;
  a.n = n;
  a.row_ptr.assign(row_ptr, row_ptr + n + 1);
  a.col_idx.assign(col_idx, col_idx + row_ptr[n]);
  a.values.assign(values, values + row_ptr[n]);
  std::vector<float> bv(b, b + n);
  CgStats st;
  auto x = cg_schwarz(a, bv, tol, max_iter, domain, weight, &st);
  std::memcpy(x_out, x.data(), n * 4);
  if (iters) *iters = st.iters;
  if (residual) *residual = st.residual;
}

// Schwarz domains (vertex index lists, concatenated) and z = M^-1 r (f64); weight 0 = one
// level, otherwise the two-level form with that coarse weight.
uint32_t oracle_schwarz(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, uint32_t domain, double weight,
                        uint32_t *dom_ptr, uint32_t *dom_vert, const double *r, double *z) {
  Csr a;
  a.n = n;
  a.row_ptr.assign(row_ptr, row_ptr + n + 1);
  a.col_idx.assign(col_idx, col_idx + row_ptr[n]);
  a.values.assign(values, values + row_ptr[n]);
  Schwarz2 sz;
  sz.build(a, domain, weight);
  if (dom_ptr) {
    uint32_t o = 0;
    dom_ptr[0] = 0;
    for (size_t d = 0; d < sz.l0.domains.size(); ++d) {
      for (uint32_t v : sz.l0.domains[d]) dom_vert[o++] = v;
      dom_ptr[d + 1] = o;
    }
  }
  if (r && z) {
    std::vector<double> rv(r, r + n), zv(n, 0.0);
    sz.apply(rv, zv);
    std::memcpy(z, zv.data(), n * 8);
  }
  return (uint32_t)sz.l0.domains.size();
}

// f32 CSR SpMV (csr.rs:98-110 spmv_cpu).
void oracle_spmv(uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx, const float *values, const float *x, float *y) {
  for (uint32_t r = 0; r < n; ++r) {
    float s = 0.0f;
    for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += values[k] * x[col_idx[k]];
    y[r] = s;
  }
}

// ---------------------------------------------------------------- element-level
#define PREC_SWITCH(fp64, ...) \
  if (fp64) { using T = double; __VA_ARGS__ } else { using T = float; __VA_ARGS__ }

double oracle_point_tri_distance_grad(int fp64, const double *p12, double *g12) {
  PREC_SWITCH(fp64, {
    V3<T> g[4];
    T d = point_tri_distance_grad(v3<T>(p12), v3<T>(p12 + 3), v3<T>(p12 + 6), v3<T>(p12 + 9), g);
    for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) g12[3 * a + c] = (double)g[a][c];
    return (double)d;
  })
}
double oracle_segment_distance_grad(int fp64, const double *p12, double *g12) {
  PREC_SWITCH(fp64, {
    V3<T> g[4];
    T d = segment_segment_distance_grad(v3<T>(p12), v3<T>(p12 + 3), v3<T>(p12 + 6), v3<T>(p12 + 9), g);
    for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) g12[3 * a + c] = (double)g[a][c];
    return (double)d;
  })
}
double oracle_segment_distance(int fp64, const double *p12) {
  PREC_SWITCH(fp64, { return (double)segment_segment_distance(v3<T>(p12), v3<T>(p12 + 3), v3<T>(p12 + 6), v3<T>(p12 + 9)); })
}
double oracle_point_tri_distance(int fp64, const double *p12) {
  PREC_SWITCH(fp64, { return (double)point_triangle_distance(v3<T>(p12), v3<T>(p12 + 3), v3<T>(p12 + 6), v3<T>(p12 + 9)); })
}
int oracle_edge_triangle_intersect(int fp64, const double *p15) {
  PREC_SWITCH(fp64, {
    return edge_triangle_intersect(v3<T>(p15), v3<T>(p15 + 3), v3<T>(p15 + 6), v3<T>(p15 + 9), v3<T>(p15 + 12)) ? 1 : 0;
  })
}
// kind 0 = vertex-triangle (p = vertex, a, b, c), 1 = edge-edge.
double oracle_accd_toi(int fp64, int kind, const double *x12, const double *dx12, double offset, double max_t, double reduction,
                       uint32_t max_iter) {
  PREC_SWITCH(fp64, {
    V3<T> x0[4], dd[4];
    for (int k = 0; k < 4; ++k) { x0[k] = v3<T>(x12 + 3 * k); dd[k] = v3<T>(dx12 + 3 * k); }
    if (kind == 0)
      return (double)accd_toi(x0, dd, T(offset), T(max_t), T(reduction), max_iter, [](const V3<T> *p) {
        T d = point_triangle_distance(p[0], p[1], p[2], p[3]);
        return d * d;
      });
    return (double)accd_toi(x0, dd, T(offset), T(max_t), T(reduction), max_iter, [](const V3<T> *p) {
      T d = segment_segment_distance(p[0], p[1], p[2], p[3]);
      return d * d;
    });
  })
}
int oracle_barrier(int fp64, int kind, double d, double dhat, double kappa, double *out3) {
  PREC_SWITCH(fp64, {
    T o[3];
    if (!barrier_of(kind, T(d), T(dhat), T(kappa), o)) return 0;
    for (int i = 0; i < 3; ++i) out3[i] = (double)o[i];
    return 1;
  })
}
double oracle_rayleigh_stiffness(int fp64, const double *elastic, const double *mass, const double *weights, int n, const double *e,
                                 double offset) {
  PREC_SWITCH(fp64, {
    T el[16][9], m[4], w[4];
    for (int i = 0; i < n * n; ++i) for (int k = 0; k < 9; ++k) el[i][k] = T(elastic[9 * i + k]);
    for (int i = 0; i < n; ++i) { m[i] = T(mass[i]); w[i] = T(weights[i]); }
    return (double)rayleigh_stiffness<T>(el, m, w, n, v3<T>(e), T(offset));
  })
}
// Friction for one contact: returns 1 if active; grad3, hess9 out.
int oracle_friction(int fp64, const double *force, const double *dx, const double *normal, double mu, double min_dx, double *grad3,
                    double *hess9) {
  PREC_SWITCH(fp64, {
    V3<T> g;
    T hh[9];
    if (!friction(v3<T>(force), v3<T>(dx), v3<T>(normal), T(mu), T(min_dx), g, hh)) return 0;
    for (int c = 0; c < 3; ++c) grad3[c] = (double)g[c];
    for (int k = 0; k < 9; ++k) hess9[k] = (double)hh[k];
    return 1;
  })
}
void oracle_lame(double young, double poisson, double *mu, double *lambda) {
  float m, l;
  lame_from_young_poisson<float>((float)young, (float)poisson, m, l);
  *mu = m;
  *lambda = l;
}
// Shell element on rest positions rest9 and current p9: energy, gradient (9), Hessian (81).
double oracle_shell_element(int fp64, int model, double mu, double lambda, double eps, const double *rest9, const double *p9, double *grad9,
                            double *hess81) {
  PREC_SWITCH(fp64, {
    std::vector<V3<T>> rest = {v3<T>(rest9), v3<T>(rest9 + 3), v3<T>(rest9 + 6)};
    std::vector<std::array<uint32_t, 3>> tri = {{0u, 1u, 2u}};
    auto els = build_shells<T>(tri, rest, T(1e-12));
    if (els.empty()) return 0.0;
    V3<T> p[3] = {v3<T>(p9), v3<T>(p9 + 3), v3<T>(p9 + 6)};
    V3<T> g[3];
    T h[81];
    shell_gradient(p, els[0], model, T(mu), T(lambda), g);
    shell_hessian(p, els[0], model, T(mu), T(lambda), T(eps), h);
    for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) grad9[3 * a + c] = (double)g[a][c];
    for (int k = 0; k < 81; ++k) hess81[k] = (double)h[k];
    return (double)shell_energy(p, els[0], model, T(mu), T(lambda));
  })
}
// Hinge: angle, dtheta/dx (12).
double oracle_dihedral(int fp64, const double *p12, double *g12) {
  PREC_SWITCH(fp64, {
    V3<T> p[4] = {v3<T>(p12), v3<T>(p12 + 3), v3<T>(p12 + 6), v3<T>(p12 + 9)};
    V3<T> g[4];
    dihedral_grad(p, g);
    for (int a = 0; a < 4; ++a) for (int c = 0; c < 3; ++c) g12[3 * a + c] = (double)g[a][c];
    return (double)dihedral_angle(p);
  })
}
// Strand bend at unit stiffness: returns angle; grad (9), Hessian (81).
double oracle_strand(int fp64, const double *p9, double rest, double *grad9, double *hess81) {
  PREC_SWITCH(fp64, {
    V3<T> g[3];
    T h[81];
    strand_force_hessian(v3<T>(p9), v3<T>(p9 + 3), v3<T>(p9 + 6), T(rest), g, h);
    for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) grad9[3 * a + c] = (double)g[a][c];
    for (int k = 0; k < 81; ++k) hess81[k] = (double)h[k];
    return (double)strand_angle(v3<T>(p9), v3<T>(p9 + 3), v3<T>(p9 + 6));
  })
}
void oracle_rod_stretch(int fp64, const double *d3, double rest, double k, double *h9, double *ga3) {
  PREC_SWITCH(fp64, {
    T h[9];
    V3<T> ga;
    rod_stretch_block(v3<T>(d3), T(rest), T(k), h, ga);
    for (int i = 0; i < 9; ++i) h9[i] = (double)h[i];
    for (int c = 0; c < 3; ++c) ga3[c] = (double)ga[c];
  })
}
void oracle_spring(int fp64, const double *d3, double rest, double k, int project, double *h9, double *ga3) {
  PREC_SWITCH(fp64, {
    T h[9];
    V3<T> ga;
    spring_block(v3<T>(d3), T(rest), T(k), h, ga, project != 0);
    for (int i = 0; i < 9; ++i) h9[i] = (double)h[i];
    for (int c = 0; c < 3; ++c) ga3[c] = (double)ga[c];
  })
}

// ---------------------------------------------------------------- LBVH stages
void oracle_morton_codes(uint32_t nv, const float *pos, uint32_t nt, const uint32_t *tris, uint32_t *codes) {
  std::vector<V3<float>> p(nv);
  for (uint32_t i = 0; i < nv; ++i) p[i] = V3<float>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
  std::vector<std::array<uint32_t, 3>> t(nt);
  for (uint32_t f = 0; f < nt; ++f) t[f] = {tris[3 * f], tris[3 * f + 1], tris[3 * f + 2]};
  auto c = morton_codes(p, t);
  std::memcpy(codes, c.data(), nt * 4);
}
void oracle_sort_pairs(uint32_t n, uint32_t *keys, uint32_t *vals) {
  std::vector<uint32_t> k(keys, keys + n), v(vals, vals + n);
  sort_pairs(k, v);
  std::memcpy(keys, k.data(), n * 4);
  std::memcpy(vals, v.data(), n * 4);
}
void oracle_karras(uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right) {
  std::vector<uint32_t> c(codes, codes + n), l, r;
  karras_tree(c, l, r);
  if (n > 1) { std::memcpy(left, l.data(), (n - 1) * 4); std::memcpy(right, r.data(), (n - 1) * 4); }
}
void oracle_propagate_aabb(uint32_t n, const uint32_t *left, const uint32_t *right, const float *leaf6, float *node6) {
  std::vector<uint32_t> l(left, left + (n > 0 ? n - 1 : 0)), r(right, right + (n > 0 ? n - 1 : 0));
  std::vector<std::array<float, 6>> lf(n);
  for (uint32_t i = 0; i < n; ++i) for (int k = 0; k < 6; ++k) lf[i][k] = leaf6[6 * i + k];
  auto nd = propagate_aabb(l, r, lf);
  for (size_t i = 0; i < nd.size(); ++i) for (int k = 0; k < 6; ++k) node6[6 * i + k] = nd[i][k];
}

}  // extern "C"
