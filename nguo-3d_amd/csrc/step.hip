// The resident implicit step: engine-core/src/implicit.rs:1951-2191 ImplicitSolver::step
// (and the wgpu GpuImplicitSolver it generalises, engine-gpu/src/gpu_implicit.rs:94-659)
// on one MI355X. Per Newton iteration, all on the device:
//   1. element kernels (shells, hinges, rods, rod bends, edge springs) write local blocks;
//   2. a fixed-order gather sums them into the SELL-64 3x3 BSR over the free vertices
//      (push order of implicit.rs:2392-2740: inertia, rods, rod bends, edge springs, shells,
//      hinges, analytic colliders) - no float atomics, bitwise deterministic;
//   3. LBVH queries produce the active vertex-triangle and edge-edge stencils, sorted by
//      (v, t) / (ei, ej) exactly as implicit.rs:678-753 orders them;
//   4. per stencil: Rayleigh stiffness from the assembled elastic blocks, barrier (FP64),
//      lagged friction; the contact Hessian stays matrix-free (rank one per stencil plus
//      the friction projector) and the block-Jacobi diagonal absorbs it;
//   5. FP32 PCG with FP64 refinement to the requested relative residual;
//   6. the ACCD line search (FP64 time of impact) over LBVH candidates with the reference's
//      widening search radius (implicit.rs:1834-1943).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "device_common.h"
#include "devmath.h"
#include "lbvh.h"
#include "pcg.h"

namespace nguo {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t CCD_DEFER_CAP = 1u << 18;  // CCD pairs one pass can defer (more run inline)
constexpr uint32_t C_DENSE = 0u, C_ROD_POS = 1u, C_ROD_NEG = 2u, C_HINGE = 3u;
__host__ __device__ inline uint32_t code(uint32_t type, uint32_t payload) { return (type << 30) | payload; }

// ---------------------------------------------------------------- element sets
struct ElemDev {
  uint32_t ns, nh, nr, nrh, nsp;
  const uint32_t *sh_v;  const float *sh_dm; const float *sh_area;
  const float *sh_scale;  // rest-shape scale per shell (Dm^-1 / s, area s^2)
  const uint32_t *hg_v;  const float *hg_rest;
  const uint32_t *rd_v;  const float *rd_rest; const float *rd_k;
  const uint32_t *rh_v;  const float *rh_rest; const float *rh_stiff;
  const uint32_t *sp_v;  const float *sp_rest;  // edge springs (stiffness Params::ksp)
  const double *sp_rest64;  // their rest lengths in FP64 (energy evaluation)
  float *out_dense;  // 81 floats per rod bend, then per shell (9 blocks of 9)
  float *out_dg;     // 9 floats per rod bend, then per shell
  float *out_rod;    // 12 per rod, then per edge spring: hk (9) + ga (3)
  float *out_hinge;  // 14 per hinge: kwt, kwt*(theta-rest), dtheta/dx (12)
};

struct Params {
  int shell_model;
  float shell_mu, shell_lambda, eig_eps, kb;
  float ksp;  // edge-spring stiffness
  float inv_h2;
  float dhat, kappa, mu, feps;
  int barrier_kind;
  int project;  // PSD-project element Hessians (Newton); 0 = exact second derivatives (adjoint)
  uint32_t ncoll;
  nguo_collider coll[4];
};

__global__ __launch_bounds__(BLOCK) void k_shells(ElemDev E, Params P, const float *__restrict__ x) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.ns) return;
  const uint32_t *v = E.sh_v + 3 * (size_t)e;
  f3 p[3] = {ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2])};
  const float *dm = E.sh_dm + 4 * (size_t)e;
  const float sc = E.sh_scale[e];
  Mat2 m{dm[0] / sc, dm[1] / sc, dm[2] / sc, dm[3] / sc};
  f3 g[3];
  float h[81];
  shell_element(p, m, E.sh_area[e] * sc * sc, P.shell_model, P.shell_mu, P.shell_lambda, P.eig_eps, g, h, P.project != 0);
  float *ob = E.out_dense + 81 * ((size_t)E.nrh + e);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) ob[(a * 3 + b) * 9 + 3 * r + c] = h[(3 * a + r) * 9 + 3 * b + c];
  float *og = E.out_dg + 9 * ((size_t)E.nrh + e);
  for (int a = 0; a < 3; ++a) { og[3 * a] = g[a].x; og[3 * a + 1] = g[a].y; og[3 * a + 2] = g[a].z; }
}

// Drape adjoint, per active shell e with scale s: the derivative of its gradient in s.
// With Dm'^-1 = Dm^-1 / s and area' = s^2 area, F' = F / s leaves U and V unchanged and
// scales the stretches, so d g / d s = (1/s) area' U diag(c) V^T Dm'^-T (the gradient's
// assembly) with c_i = dPsi/dsigma_i - sum_j d2Psi/dsigma_i dsigma_j sigma_j. Writes
// out[face] = -lambda . dg/ds over the element's vertices (lambda on free dofs by vdof).
__global__ __launch_bounds__(BLOCK) void k_shell_dgds(ElemDev E, Params P, const float *__restrict__ x, const uint32_t *__restrict__ vdof,
                                                      const float *__restrict__ lam, const uint32_t *__restrict__ face, float *out) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.ns) return;
  const uint32_t *v = E.sh_v + 3 * (size_t)e;
  f3 p[3] = {ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2])};
  const float *dm = E.sh_dm + 4 * (size_t)e;
  const float sc = E.sh_scale[e];
  Mat2 m{dm[0] / sc, dm[1] / sc, dm[2] / sc, dm[3] / sc};
  const float area = E.sh_area[e] * sc * sc;
  f3 e1 = p[1] - p[0], e2 = p[2] - p[0];
  f3 f0 = e1 * m.a + e2 * m.b;
  f3 f1 = e1 * m.c + e2 * m.d;
  f3 u0, u1;
  float sg[2], v0[2], v1[2];
  svd3x2(f0, f1, u0, u1, sg, v0, v1);
  float deda[2], d2[4];
  shell_derivs(sg, P.shell_model, P.shell_mu, P.shell_lambda, deda, d2);
  const float c0s = deda[0] - (d2[0] * sg[0] + d2[1] * sg[1]);
  const float c1s = deda[1] - (d2[2] * sg[0] + d2[3] * sg[1]);
  f3 c0 = u0 * c0s, c1 = u1 * c1s;
  f3 pk0 = c0 * v0[0] + c1 * v1[0];
  f3 pk1 = c0 * v0[1] + c1 * v1[1];
  f3 h0 = pk0 * m.a + pk1 * m.c;
  f3 h1 = pk0 * m.b + pk1 * m.d;
  const float k = area / sc;
  f3 g[3] = {(h0 + h1) * (-k), h0 * k, h1 * k};
  float acc = 0.0f;
  for (int a = 0; a < 3; ++a) {
    uint32_t d = vdof[v[a]];
    if (d == 0xffffffffu) continue;
    f3 l = ld3(lam, d);
    acc += dot(l, g[a]);
  }
  out[face[e]] = -acc;
}

// A stencil's gap and distance gradient in FP64 (the FP64 scheme's own closest points),
// signed like its FP32 gap d32 (negative for a vertex inside the static body).
__device__ inline double stencil_geom64(bool vt, const uint32_t id[4], const float *__restrict__ x, float d32, d3 g[4]) {
  const d3 p0 = to_d(ld3(x, id[0])), p1 = to_d(ld3(x, id[1])), p2 = to_d(ld3(x, id[2])), p3 = to_d(ld3(x, id[3]));
  double d = vt ? point_tri_distance_grad(p0, p1, p2, p3, g) : segment_segment_distance_grad(p0, p1, p2, p3, g);
  if (d32 < 0.0f) {
    d = -d;
    for (int a = 0; a < 4; ++a) g[a] = -g[a];
  }
  return d;
}

// ---- exact adjoint (NGUO_ADJ_EXACT): the Jacobian terms the Newton Hessian leaves out, as
// dense FP64 blocks over up to four vertices (dof[4 k + a]: free block row, NONE: no row or
// column; blk[144 k + (3 a + r) * 12 + 3 b + c]). Per contact stencil kappa b'(d) times the
// Jacobian of the distance gradient (kappa, barrier kind and stencil set those of the step's
// end state, as the step's assembly there holds them); per hinge kb (theta - rest) times the
// Jacobian of the dihedral gradient; per strand the exact Jacobian of its gradient minus the
// Wu-Kim block the Newton Hessian holds; per shell, rod segment and edge spring the
// unprojected block minus the projected one; per vertex the spheres' b'(d) (I - n n^T) /
// |x - c|. The Jacobians of the FP64 gradients are central differences (devmath
// fd_jacobian64).
struct AdjBlocks {
  uint32_t *dof;
  double *blk;
};

__device__ inline double span64(const d3 *p, int nv) {
  double m = 0.0;
  for (int a = 0; a < nv; ++a)
    for (int b = a + 1; b < nv; ++b) m = fmax(m, len(p[a] - p[b]));
  return m > 0.0 ? m : 1.0;
}

__global__ __launch_bounds__(BLOCK) void k_adj_contacts(uint32_t ns, uint32_t nvt, const uint32_t *__restrict__ sidx, const float *__restrict__ sd,
                                                       const float *__restrict__ kappa, float ghat_vt, float ghat_ee, Params P,
                                                       const uint32_t *__restrict__ vdof, const float *__restrict__ x, AdjBlocks B) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= ns) return;
  uint32_t id[4];
  d3 p[4];
  for (int a = 0; a < 4; ++a) {
    id[a] = sidx[4 * (size_t)s + a];
    p[a] = to_d(ld3(x, id[a]));
    B.dof[4 * (size_t)s + a] = vdof[id[a]];
  }
  double *J = B.blk + 144 * (size_t)s;
  const bool vt = s < nvt, neg = sd[s] < 0.0f;
  auto grad = [&](const d3 *q, d3 *g) {
    if (vt) (void)point_tri_distance_grad(q[0], q[1], q[2], q[3], g);
    else (void)segment_segment_distance_grad(q[0], q[1], q[2], q[3], g);
    if (neg)
      for (int a = 0; a < 4; ++a) g[a] = -g[a];
  };
  d3 g[4];
  const double d = stencil_geom64(vt, id, x, sd[s], g);
  const int kind = neg ? 1 : P.barrier_kind;
  double bo[3] = {0, 0, 0};
  const bool act = barrier_of(kind, d, (double)(vt ? ghat_vt : ghat_ee), (double)kappa[s], bo);
  if (!act || bo[1] == 0.0) {
    for (int k = 0; k < 144; ++k) J[k] = 0.0;
    return;
  }
  fd_jacobian64<4>(p, span64(p, 4), grad, J);
  for (int k = 0; k < 144; ++k) J[k] *= bo[1];
}

__global__ __launch_bounds__(BLOCK) void k_adj_hinges(ElemDev E, Params P, const uint32_t *__restrict__ vdof, const float *__restrict__ x, AdjBlocks B) {
  const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nh) return;
  d3 p[4];
  for (int a = 0; a < 4; ++a) {
    const uint32_t v = E.hg_v[4 * (size_t)e + a];
    p[a] = to_d(ld3(x, v));
    B.dof[4 * (size_t)e + a] = vdof[v];
  }
  double *J = B.blk + 144 * (size_t)e;
  const double c = (double)P.kb * (dihedral_angle(p) - (double)E.hg_rest[e]);
  if (!(P.kb > 0.0f) || c == 0.0) {
    for (int k = 0; k < 144; ++k) J[k] = 0.0;
    return;
  }
  fd_jacobian64<4>(p, span64(p, 4), [](const d3 *q, d3 *g) { dihedral_grad64(q, g); }, J);
  for (int k = 0; k < 144; ++k) J[k] *= c;
}

__global__ __launch_bounds__(BLOCK) void k_adj_strands(ElemDev E, const uint32_t *__restrict__ vdof, const float *__restrict__ x, AdjBlocks B) {
  const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nrh) return;
  d3 p[3];
  for (int a = 0; a < 3; ++a) {
    const uint32_t v = E.rh_v[3 * (size_t)e + a];
    p[a] = to_d(ld3(x, v));
    B.dof[4 * (size_t)e + a] = vdof[v];
  }
  B.dof[4 * (size_t)e + 3] = NONE;
  double *J = B.blk + 144 * (size_t)e;
  for (int k = 0; k < 144; ++k) J[k] = 0.0;
  const double rest = (double)E.rh_rest[e], st = (double)E.rh_stiff[e];
  if (!(st > 0.0)) return;
  fd_jacobian64<3>(p, span64(p, 3), [rest](const d3 *q, d3 *g) { strand_grad64(q[0], q[1], q[2], rest, g); }, J);
  const float *ob = E.out_dense + 81 * (size_t)e;  // the Newton Hessian's block (stiffness included)
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double &v = J[(3 * a + r) * 12 + 3 * b + c];
          v = st * v - (double)ob[(a * 3 + b) * 9 + 3 * r + c];
        }
}

__global__ __launch_bounds__(BLOCK) void k_adj_shells(ElemDev E, Params P, const uint32_t *__restrict__ vdof, const float *__restrict__ x, AdjBlocks B) {
  const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.ns) return;
  const uint32_t *v = E.sh_v + 3 * (size_t)e;
  for (int a = 0; a < 3; ++a) B.dof[4 * (size_t)e + a] = vdof[v[a]];
  B.dof[4 * (size_t)e + 3] = NONE;
  f3 p[3] = {ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2])};
  const float *dm = E.sh_dm + 4 * (size_t)e;
  const float sc = E.sh_scale[e];
  Mat2 m{dm[0] / sc, dm[1] / sc, dm[2] / sc, dm[3] / sc};
  f3 g[3];
  float h[81];
  shell_element(p, m, E.sh_area[e] * sc * sc, P.shell_model, P.shell_mu, P.shell_lambda, P.eig_eps, g, h, false);
  const float *ob = E.out_dense + 81 * ((size_t)E.nrh + e);  // the projected block of the assembly
  double *J = B.blk + 144 * (size_t)e;
  for (int k = 0; k < 144; ++k) J[k] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[(3 * a + r) * 12 + 3 * b + c] = (double)h[(3 * a + r) * 9 + 3 * b + c] - (double)ob[(a * 3 + b) * 9 + 3 * r + c];
}

__global__ __launch_bounds__(BLOCK) void k_adj_rods(ElemDev E, float ksp, const uint32_t *__restrict__ vdof, const float *__restrict__ x, AdjBlocks B) {
  const uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nr + E.nsp) return;
  float h[9];
  f3 ga;
  uint32_t va, vb;
  if (e < E.nr) {
    va = E.rd_v[2 * (size_t)e];
    vb = E.rd_v[2 * (size_t)e + 1];
    rod_stretch_block(ld3(x, va) - ld3(x, vb), E.rd_rest[e], E.rd_k[e], h, ga, false);
  } else {
    const uint32_t sp = e - E.nr;
    va = E.sp_v[2 * (size_t)sp];
    vb = E.sp_v[2 * (size_t)sp + 1];
    spring_block(ld3(x, va) - ld3(x, vb), E.sp_rest[sp], ksp, h, ga, false);
  }
  B.dof[4 * (size_t)e] = vdof[va];
  B.dof[4 * (size_t)e + 1] = vdof[vb];
  B.dof[4 * (size_t)e + 2] = NONE;
  B.dof[4 * (size_t)e + 3] = NONE;
  const float *o = E.out_rod + 12 * (size_t)e;  // the projected block of the assembly
  double *J = B.blk + 144 * (size_t)e;
  for (int k = 0; k < 144; ++k) J[k] = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const double dv = (double)h[3 * r + c] - (double)o[3 * r + c];
      J[r * 12 + c] = dv;
      J[(3 + r) * 12 + 3 + c] = dv;
      J[r * 12 + 3 + c] = -dv;
      J[(3 + r) * 12 + c] = -dv;
    }
}

__global__ __launch_bounds__(BLOCK) void k_adj_spheres(uint32_t V, const float *__restrict__ im, const uint32_t *__restrict__ vdof, const float *__restrict__ x,
                                                      Params P, AdjBlocks B) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V) return;
  B.dof[4 * (size_t)i] = im[i] > 0.0f ? vdof[i] : NONE;
  for (int a = 1; a < 4; ++a) B.dof[4 * (size_t)i + a] = NONE;
  double *J = B.blk + 144 * (size_t)i;
  for (int k = 0; k < 144; ++k) J[k] = 0.0;
  if (!(im[i] > 0.0f)) return;
  const d3 xi = to_d(ld3(x, i));
  for (uint32_t k = 0; k < P.ncoll; ++k) {
    const nguo_collider &c = P.coll[k];
    if (c.kind == NGUO_COLLIDER_HALFSPACE) continue;
    const d3 v = xi - d3{c.a[0], c.a[1], c.a[2]};
    const double l = len(v);
    if (l < 1.0e-9) continue;
    const d3 n = v / l;
    double bo[3];
    if (!barrier_of(P.barrier_kind, l - (double)c.radius, (double)P.dhat, (double)P.kappa, bo)) continue;
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) J[r * 12 + cc] += bo[1] * ((r == cc ? 1.0 : 0.0) - comp(n, r) * comp(n, cc)) / l;
  }
}

// y += H2 x over the blocks (FP64; blocks sharing a row add atomically).
__global__ __launch_bounds__(BLOCK) void k_adj_apply(uint32_t nb, AdjBlocks B, const double *__restrict__ x, double *y) {
  const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= nb) return;
  uint32_t dof[4];
  for (int a = 0; a < 4; ++a) dof[a] = B.dof[4 * (size_t)k + a];
  const double *J = B.blk + 144 * (size_t)k;
  for (int a = 0; a < 4; ++a) {
    if (dof[a] == NONE) continue;
    for (int r = 0; r < 3; ++r) {
      double acc = 0.0;
      for (int b = 0; b < 4; ++b) {
        if (dof[b] == NONE) continue;
        for (int c = 0; c < 3; ++c) acc += J[(3 * a + r) * 12 + 3 * b + c] * x[3 * (size_t)dof[b] + c];
      }
      if (acc != 0.0) atomicAdd(&y[3 * (size_t)dof[a] + r], acc);
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_hinges(ElemDev E, Params P, const float *__restrict__ x) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nh) return;
  const uint32_t *v = E.hg_v + 4 * (size_t)e;
  f3 p[4] = {ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2]), ld3(x, v[3])};
  float theta = dihedral_angle(p);
  f3 g[4];
  dihedral_grad(p, g);
  float kwt = P.kb * 1.0f;  // hinge weight w = 1 (bending.rs:151-155)
  float *o = E.out_hinge + 14 * (size_t)e;
  o[0] = kwt;
  o[1] = kwt * (theta - E.hg_rest[e]);
  for (int a = 0; a < 4; ++a) { o[2 + 3 * a] = g[a].x; o[3 + 3 * a] = g[a].y; o[4 + 3 * a] = g[a].z; }
}

// Rod segments (implicit.rs:2407-2424), then the edge springs (implicit.rs:2487-2507), one
// thread per two-vertex element.
__global__ __launch_bounds__(BLOCK) void k_rods(ElemDev E, int project, float ksp, const float *__restrict__ x) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nr + E.nsp) return;
  float h[9];
  f3 ga;
  if (e < E.nr) {
    uint32_t a = E.rd_v[2 * (size_t)e], b = E.rd_v[2 * (size_t)e + 1];
    rod_stretch_block(ld3(x, a) - ld3(x, b), E.rd_rest[e], E.rd_k[e], h, ga, project != 0);
  } else {
    const uint32_t sp = e - E.nr;
    uint32_t a = E.sp_v[2 * (size_t)sp], b = E.sp_v[2 * (size_t)sp + 1];
    spring_block(ld3(x, a) - ld3(x, b), E.sp_rest[sp], ksp, h, ga, project != 0);
  }
  float *o = E.out_rod + 12 * (size_t)e;
  for (int i = 0; i < 9; ++i) o[i] = h[i];
  o[9] = ga.x; o[10] = ga.y; o[11] = ga.z;
}

__global__ __launch_bounds__(BLOCK) void k_rod_hinges(ElemDev E, const float *__restrict__ x) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= E.nrh) return;
  const uint32_t *v = E.rh_v + 3 * (size_t)e;
  f3 g[3];
  float h[81];
  strand_force_hessian(ld3(x, v[0]), ld3(x, v[1]), ld3(x, v[2]), E.rh_rest[e], g, h);
  float st = E.rh_stiff[e];
  float *ob = E.out_dense + 81 * (size_t)e;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) ob[(a * 3 + b) * 9 + 3 * r + c] = st * h[(3 * a + r) * 9 + 3 * b + c];
  float *og = E.out_dg + 9 * (size_t)e;
  for (int a = 0; a < 3; ++a) { og[3 * a] = st * g[a].x; og[3 * a + 1] = st * g[a].y; og[3 * a + 2] = st * g[a].z; }
}

// contact.rs:34-49 Collider::distance.
__device__ inline void collider_distance(const nguo_collider &c, f3 x, float &d, f3 &n) {
  if (c.kind == NGUO_COLLIDER_HALFSPACE) {
    f3 p{c.a[0], c.a[1], c.a[2]}, nn{c.b[0], c.b[1], c.b[2]};
    d = dot(x - p, nn);
    n = nn;
  } else {
    f3 ctr{c.a[0], c.a[1], c.a[2]};
    f3 v = x - ctr;
    float l = len(v);
    if (l < 1.0e-9f) { d = -c.radius; n = f3{0, 1, 0}; }
    else { d = l - c.radius; n = v / l; }
  }
}

// contact.rs:613-643 friction: returns lambda (0 when inactive), gradient in g.
__device__ inline float friction(f3 f_c, f3 dx, f3 n, float mu, float min_dx, f3 &g) {
  g = f3{0, 0, 0};
  if (mu <= 0.0f) return 0.0f;
  float contact = -dot(n, f_c);
  f3 pdx = dx - dot(n, dx) * n;
  float eps = fmaxf(min_dx, 1.0e-12f);
  float denom = fmaxf(len(pdx), eps);
  float lambda = mu * contact / denom;
  if (lambda <= 0.0f) return 0.0f;
  g = pdx * lambda;
  return lambda;
}

// contact.rs:1082-1125 vertex_contact_full for the analytic colliders (kappa fixed,
// contact_dynamic = 0): gradient and 3x3 Hessian added to g / h.
__device__ inline void collider_terms(const Params &P, f3 x, f3 xs, f3 &g, float h[9]) {
  for (uint32_t k = 0; k < P.ncoll; ++k) {
    float d;
    f3 n;
    collider_distance(P.coll[k], x, d, n);
    double bo[3];
    if (!barrier_of(P.barrier_kind, (double)d, (double)P.dhat, (double)P.kappa, bo)) continue;
    f3 f_c = n * (float)bo[1];
    g = g + f_c;
    float hp = fmaxf((float)bo[2], 0.0f);
    float nn[3] = {n.x, n.y, n.z};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) h[3 * r + c] += hp * nn[r] * nn[c];
    if (P.mu > 0.0f) {
      f3 fg;
      float lam = friction(f_c, x - xs, n, P.mu, P.feps, fg);
      if (lam > 0.0f) {
        g = g + fg;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) h[3 * r + c] += lam * ((r == c ? 1.0f : 0.0f) - nn[r] * nn[c]);
      }
    }
  }
}

// Fixed-order gather of the element blocks into the BSR slots (+ inertia and analytic
// colliders on diagonal slots of the free system).
__global__ __launch_bounds__(BLOCK) void k_gather_hess(uint32_t nslots, const uint32_t *__restrict__ slot_row, const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ hc_ptr, const uint32_t *__restrict__ hc, ElemDev E,
                                                      const uint32_t *__restrict__ row_vert, const float *__restrict__ im, int active,
                                                      Params P, const float *__restrict__ x, const float *__restrict__ xs, float *__restrict__ val) {
  uint32_t sl = blockIdx.x * BLOCK + threadIdx.x;
  if (sl >= nslots) return;
  float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t row = slot_row[sl];
  if (row != NONE) {
    bool diag = col[sl] == row;
    uint32_t vtx = row_vert[row];
    if (active && diag) {
      float m_over_h2 = (1.0f / im[vtx]) * P.inv_h2;
      acc[0] = m_over_h2; acc[4] = m_over_h2; acc[8] = m_over_h2;
    }
    for (uint32_t k = hc_ptr[sl]; k < hc_ptr[sl + 1]; ++k) {
      uint32_t c = hc[k], type = c >> 30, pl = c & 0x3fffffffu;
      if (type == C_DENSE) {
        const float *b = E.out_dense + 9 * (size_t)pl;
        for (int e = 0; e < 9; ++e) acc[e] += b[e];
      } else if (type == C_ROD_POS) {
        const float *b = E.out_rod + 12 * (size_t)pl;
        for (int e = 0; e < 9; ++e) acc[e] += b[e];
      } else if (type == C_ROD_NEG) {
        const float *b = E.out_rod + 12 * (size_t)pl;
        for (int e = 0; e < 9; ++e) acc[e] += -b[e];
      } else {
        uint32_t h = pl >> 4, a = (pl >> 2) & 3u, bb = pl & 3u;
        const float *o = E.out_hinge + 14 * (size_t)h;
        const float *ga = o + 2 + 3 * a, *gb = o + 2 + 3 * bb;
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) acc[3 * r + cc] += (o[0] * ga[r]) * gb[cc];
      }
    }
    if (active && diag && P.ncoll) {
      f3 g{0, 0, 0};
      collider_terms(P, ld3(x, vtx), ld3(xs, vtx), g, acc);
    }
  }
  for (int e = 0; e < 9; ++e) val[sell_val_index(sl, e)] = acc[e];
}

// Elastic (+ inertia, colliders) gradient per free vertex, in push order.
__global__ __launch_bounds__(BLOCK) void k_gather_grad(uint32_t nf, const uint32_t *__restrict__ free_list, const uint32_t *__restrict__ gc_ptr,
                                                      const uint32_t *__restrict__ gc, ElemDev E, const float *__restrict__ im, Params P,
                                                      const float *__restrict__ x, const float *__restrict__ xt, const float *__restrict__ xs,
                                                      float *__restrict__ grad) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nf) return;
  uint32_t v = free_list[i];
  float m_over_h2 = (1.0f / im[v]) * P.inv_h2;
  f3 xv = ld3(x, v);
  f3 diff = xv - ld3(xt, v);
  f3 g{m_over_h2 * diff.x, m_over_h2 * diff.y, m_over_h2 * diff.z};
  for (uint32_t k = gc_ptr[i]; k < gc_ptr[i + 1]; ++k) {
    uint32_t c = gc[k], type = c >> 30, pl = c & 0x3fffffffu;
    if (type == C_DENSE) {
      const float *b = E.out_dg + 3 * (size_t)pl;
      g.x += b[0]; g.y += b[1]; g.z += b[2];
    } else if (type == C_ROD_POS) {
      const float *b = E.out_rod + 12 * (size_t)pl + 9;
      g.x += b[0]; g.y += b[1]; g.z += b[2];
    } else if (type == C_ROD_NEG) {
      const float *b = E.out_rod + 12 * (size_t)pl + 9;
      g.x -= b[0]; g.y -= b[1]; g.z -= b[2];
    } else {
      uint32_t h = pl >> 2, a = pl & 3u;
      const float *o = E.out_hinge + 14 * (size_t)h;
      g.x += o[1] * o[2 + 3 * a]; g.y += o[1] * o[3 + 3 * a]; g.z += o[1] * o[4 + 3 * a];
    }
  }
  if (P.ncoll) {
    float hdummy[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    f3 gc2{0, 0, 0};
    collider_terms(P, xv, ld3(xs, v), gc2, hdummy);
    g = g + gc2;
  }
  st3(grad, i, g);
}

// ---------------------------------------------------------------- contact detection
struct Topo {
  uint32_t V;
  const float *im;
  const uint32_t *tris, *edges;
  const uint32_t *ex_ptr, *ex;
  const uint8_t *static_v;  // null = no mesh collider
  const float *side;        // null = no mesh collider
  const float *pn;          // [3V] pseudonormals
  const uint32_t *vdof, *sdof;
  const uint32_t *xdof;     // vertex -> row of the mixed-element table; null = no such element
};

__device__ inline bool movable(const Topo &T, uint32_t v) { return T.im[v] > 0.0f; }
__device__ inline bool excluded(const Topo &T, uint32_t v, uint32_t t0, uint32_t t1, uint32_t t2) {
  for (uint32_t k = T.ex_ptr[v]; k < T.ex_ptr[v + 1]; ++k) {
    uint32_t e = T.ex[k];
    if (e == t0 || e == t1 || e == t2) return true;
  }
  return false;
}
// The query vertex's exclusion list (its 1-ring), up to EXC_REG entries held in registers for
// the leaf tests; a longer list has its tail read from memory.
constexpr uint32_t EXC_REG = 12;
struct ExcList {
  uint32_t r[EXC_REG];
  uint32_t b, e;
  __device__ inline void load(const Topo &T, uint32_t v) {
    b = T.ex_ptr[v];
    e = T.ex_ptr[v + 1];
#pragma unroll
    for (uint32_t u = 0; u < EXC_REG; ++u) r[u] = b + u < e ? T.ex[b + u] : 0xffffffffu;
  }
  __device__ inline bool has(const Topo &T, uint32_t t0, uint32_t t1, uint32_t t2) const {
    bool hit = false;
#pragma unroll
    for (uint32_t u = 0; u < EXC_REG; ++u) hit |= r[u] == t0 || r[u] == t1 || r[u] == t2;
    for (uint32_t k = b + EXC_REG; k < e && !hit; ++k) {
      const uint32_t x = T.ex[k];
      hit = x == t0 || x == t1 || x == t2;
    }
    return hit;
  }
};

__device__ inline bool inside_padded(f3 p, f3 lo, f3 hi, float pad) {
  f3 l = lo - f3{pad, pad, pad}, h = hi + f3{pad, pad, pad};
  return !(p.x < l.x || p.y < l.y || p.z < l.z || p.x > h.x || p.y > h.y || p.z > h.z);
}

// implicit.rs:1431-1476 collider_sides: the side of each movable non-collider vertex from
// its closest collider facet within dhat (ties: lowest facet index); +1 otherwise. A vertex
// found inside sets below_any[0] (may be null).
__global__ __launch_bounds__(BLOCK) void k_sides(Topo T, BvhView coll, const float *__restrict__ x, float dhat, float *side, uint32_t *err,
                                                uint32_t *below_any) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
  if (v >= T.V) return;
  float s = 1.0f;
  if (movable(T, v) && !T.static_v[v]) {
    f3 p = ld3(x, v);
    float best = INFINITY;
    uint32_t best_t = NONE;
    float best_side = 1.0f;
    bvh_traverse(coll, err, tstack, [&](f3 lo, f3 hi) { return inside_padded(p, lo, hi, dhat); },
                 [&](uint32_t t) {
                   uint32_t t0 = T.tris[3 * (size_t)t], t1 = T.tris[3 * (size_t)t + 1], t2 = T.tris[3 * (size_t)t + 2];
                   f3 a = ld3(x, t0), b = ld3(x, t1), c = ld3(x, t2);
                   if (!inside_padded(p, vmin(vmin(a, b), c), vmax(vmax(a, b), c), dhat)) return;
                   if (excluded(T, v, t0, t1, t2)) return;
                   float l[3];
                   f3 q = closest_point_bary(p, a, b, c, l);
                   f3 e = p - q;
                   float d2 = dot(e, e);
                   if (d2 > best || (d2 == best && t > best_t)) return;
                   best = d2;
                   best_t = t;
                   f3 n = surface_normal_at(l, normalize_or_zero(cross(b - a, c - a)), ld3(T.pn, t0), ld3(T.pn, t1), ld3(T.pn, t2));
                   best_side = dot(e, n) < 0.0f ? -1.0f : 1.0f;
                 });
    if (best < dhat * dhat) s = best_side;
  }
  side[v] = s;
  if (below_any && s < 0.0f) atomicOr(below_any, 1u);
}

// The vertex-triangle broad phase's grid cell (broadphase.rs:209-241 grid_cell and
// mean_tri_extent): max(dhat, mean over all triangles of the largest side of the triangle's
// AABB, 1e-6), the mean summed in FP32 in triangle order as the reference sums it. Only a
// vertex inside a collider within dhat needs it: its signed gap to a collider facet is
// negative whatever the distance, so the reference's grid cells alone decide which facets it
// meets. cell[0] is the flag k_sides raised; without it the kernel returns at once. cell[1]
// receives the cell's FP32 bits. One block: the extents of 4096 triangles at a time go to LDS
// in parallel, then thread 0 adds them in order.
constexpr uint32_t CELL_BLOCK = 256, CELL_CHUNK = 4096;
__global__ __launch_bounds__(CELL_BLOCK) void k_vt_cell(const uint32_t *__restrict__ tris, uint32_t F, const float *__restrict__ x, float dhat,
                                                       uint32_t *cell) {
  if (cell[0] == 0) return;
  __shared__ float ext[CELL_CHUNK];
  float sum = 0.0f;
  for (uint32_t base = 0; base < F; base += CELL_CHUNK) {
    const uint32_t n = min(CELL_CHUNK, F - base);
    for (uint32_t k = threadIdx.x; k < n; k += CELL_BLOCK) {
      const size_t t = base + k;
      f3 a = ld3(x, tris[3 * t]), b = ld3(x, tris[3 * t + 1]), c = ld3(x, tris[3 * t + 2]);
      f3 d = vmax(vmax(a, b), c) - vmin(vmin(a, b), c);
      ext[k] = fmaxf(fmaxf(d.x, d.y), d.z);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (uint32_t k = 0; k < n; ++k) sum += ext[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) cell[1] = __float_as_uint(fmaxf(fmaxf(dhat, F ? sum / (float)F : 0.0f), 1e-6f));
}

__device__ inline int cell_key(float v, float inv) { return (int)floorf(v * inv); }

struct StencilOut {
  uint32_t cap;
  uint32_t *count;
  uint64_t *key;
  uint32_t *order;  // payload for the sort: slot index
  float *d;         // [cap]
  float *g;         // [12 cap]
};

__device__ inline void emit(const StencilOut &O, uint64_t key, float d, const f3 g[4]) {
  uint32_t slot = atomicAdd(O.count, 1u);
  if (slot >= O.cap) return;
  O.key[slot] = key;
  O.order[slot] = slot;
  O.d[slot] = d;
  float *o = O.g + 12 * (size_t)slot;
  for (int a = 0; a < 4; ++a) { o[3 * a] = g[a].x; o[3 * a + 1] = g[a].y; o[3 * a + 2] = g[a].z; }
}

// implicit.rs:683-717 vertex-triangle stencils: every (v, t) with v in AABB(t) + dhat, not
// excluded, not both immovable, with (signed for collider facets) distance < dhat. A vertex
// inside a collider (side -1) meets a collider facet when its grid cell lies in the cells of
// the facet's padded AABB (broadphase.rs:299-325; the cell from k_vt_cell in cell[1]), as the
// reference's hash grid offers it; any other pair's gap below dhat puts v inside the padded
// AABB already, so the box test is the same filter there.
__global__ __launch_bounds__(BLOCK) void k_vt_stencils(Topo T, BvhView dyn, BvhView stat, const float *__restrict__ x, float dhat,
                                                      const uint32_t *__restrict__ cell, StencilOut O, uint32_t *err) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
  if (v >= T.V) return;
  f3 p = ld3(x, v);
  bool mv = movable(T, v);
  const bool below = T.side && cell ? T.side[v] < 0.0f : false;
  // an inside vertex's traversal pad covers every facet whose padded cells can hold it: a
  // facet AABB within dhat + one cell of p, with a second cell of margin for the rounding of
  // the keys
  float inv = 0.0f, tpad = dhat;
  int kx = 0, ky = 0, kz = 0;
  if (below) {
    const float c = __uint_as_float(cell[1]);
    inv = 1.0f / c;
    tpad = dhat + 2.0f * c;
    kx = cell_key(p.x, inv); ky = cell_key(p.y, inv); kz = cell_key(p.z, inv);
  }
  ExcList exl;
  exl.load(T, v);
  // a leaf loads its positions and flags in one memory round trip (the triangle's vertex ids
  // come with the traversal record) and tests the filters (predicates, in any order) from
  // registers
  auto leaf = [&](uint32_t t, uint3 vid) {
    const uint32_t t0 = vid.x, t1 = vid.y, t2 = vid.z;
    f3 a = ld3(x, t0), b = ld3(x, t1), c = ld3(x, t2);
    const float m0 = T.im[t0], m1 = T.im[t1], m2 = T.im[t2];
    const bool st_all = T.static_v ? (T.static_v[t0] && T.static_v[t1] && T.static_v[t2]) : false;
    if (st_all && below) {
      const f3 lo = vmin(vmin(a, b), c) - f3{dhat, dhat, dhat}, hi = vmax(vmax(a, b), c) + f3{dhat, dhat, dhat};
      if (kx < cell_key(lo.x, inv) || kx > cell_key(hi.x, inv) || ky < cell_key(lo.y, inv) || ky > cell_key(hi.y, inv) ||
          kz < cell_key(lo.z, inv) || kz > cell_key(hi.z, inv))
        return;
    } else if (!inside_padded(p, vmin(vmin(a, b), c), vmax(vmax(a, b), c), dhat))
      return;
    bool tm = m0 > 0.0f || m1 > 0.0f || m2 > 0.0f;
    if (!tm && !mv) return;
    if (exl.has(T, t0, t1, t2)) return;
    f3 g[4];
    float d = point_tri_distance_grad(p, a, b, c, g);
    if (st_all && below) {
      d = -d;
      for (int k = 0; k < 4; ++k) g[k] = -g[k];
    }
    if (d >= dhat) return;
    emit(O, ((uint64_t)v << 32) | t, d, g);
  };
  auto ov = [&](f3 lo, f3 hi) { return inside_padded(p, lo, hi, tpad); };
  // parts 0 .. TRAV_PARTS-1: the movable tree's subtrees; the rest: the static tree's
  const uint32_t part = blockIdx.y;
  if (part < TRAV_PARTS) bvh_traverse_part(dyn, part, err, tstack, ov, leaf);
  else if (mv) bvh_traverse_part(stat, part - TRAV_PARTS, err, tstack, ov, leaf);
}

// implicit.rs:719-751 edge-edge stencils. Query edges are padded, tree edges are not (the
// overlap is symmetric). self_query: the queries are the tree's own edges and each pair is met
// once, by the query whose sorted leaf slot is the lower; otherwise (movable edges against
// the static edges' tree) every leaf is a partner. Keys are (lower, higher edge id).
__global__ __launch_bounds__(BLOCK) void k_ee_stencils(Topo T, BvhView tree, const uint32_t *qids, uint32_t nq, int self_query,
                                                      const float *__restrict__ x, float pad, float ghat, StencilOut O, uint32_t *err) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= nq) return;
  uint32_t ej = qids[q];
  uint32_t b0 = T.edges[2 * (size_t)ej], b1 = T.edges[2 * (size_t)ej + 1];
  f3 pb = ld3(x, b0), qb = ld3(x, b1);
  f3 blo = vmin(pb, qb) - f3{pad, pad, pad}, bhi = vmax(pb, qb) + f3{pad, pad, pad};
  auto ov = [&](f3 lo, f3 hi) { return !(lo.x > bhi.x || lo.y > bhi.y || lo.z > bhi.z || hi.x < blo.x || hi.y < blo.y || hi.z < blo.z); };
  // a leaf reuses the query edge's positions from registers (the pair's vertices in
  // edge-index order: the lower edge's two, then the higher's); a movable query meets each
  // other movable edge once: only the leaves after its own sorted slot
  const uint32_t min_slot = self_query ? tree.slot[q] + 1 : 0u;
  bvh_traverse_part(tree, blockIdx.y, err, tstack, ov, [&](uint32_t ei, uint3 vid) {
    if (ei == ej) return;
    const uint32_t a0 = vid.x, a1 = vid.y;
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    f3 pa = ld3(x, a0), qa = ld3(x, a1);
    if (!ov(vmin(pa, qa), vmax(pa, qa))) return;
    const bool a_lo = ei < ej;
    const uint32_t lo_e = min(ei, ej), hi_e = max(ei, ej);
    f3 g[4];
    float d = segment_segment_distance_grad(a_lo ? pa : pb, a_lo ? qa : qb, a_lo ? pb : pa, a_lo ? qb : qa, g);
    if (d - 0.0f >= ghat) return;
    emit(O, ((uint64_t)lo_e << 32) | hi_e, d - 0.0f, g);
  }, min_slot);
}

// Gathers the sorted stencils (V-T block then E-E block) into one array.
__global__ __launch_bounds__(BLOCK) void k_stencil_gather(uint32_t nvt, uint32_t nee, const uint64_t *vt_key, const uint32_t *vt_ord,
                                                         const float *vt_d, const float *vt_g, const uint64_t *ee_key, const uint32_t *ee_ord,
                                                         const float *ee_d, const float *ee_g, const uint32_t *tris, const uint32_t *edges,
                                                         uint32_t *sidx, float *sd, float *sg) {
  uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= nvt + nee) return;
  uint32_t id[4];
  uint32_t src;
  const float *dd, *gg;
  if (s < nvt) {
    uint64_t k = vt_key[s];
    uint32_t v = (uint32_t)(k >> 32), t = (uint32_t)k;
    id[0] = v; id[1] = tris[3 * (size_t)t]; id[2] = tris[3 * (size_t)t + 1]; id[3] = tris[3 * (size_t)t + 2];
    src = vt_ord[s]; dd = vt_d; gg = vt_g;
  } else {
    uint64_t k = ee_key[s - nvt];
    uint32_t ei = (uint32_t)(k >> 32), ej = (uint32_t)k;
    id[0] = edges[2 * (size_t)ei]; id[1] = edges[2 * (size_t)ei + 1]; id[2] = edges[2 * (size_t)ej]; id[3] = edges[2 * (size_t)ej + 1];
    src = ee_ord[s - nvt]; dd = ee_d; gg = ee_g;
  }
  for (int a = 0; a < 4; ++a) sidx[4 * (size_t)s + a] = id[a];
  sd[s] = dd[src];
  for (int k = 0; k < 12; ++k) sg[12 * (size_t)s + k] = gg[12 * (size_t)src + k];
}

// Adds block (r, c) of M to out when the pattern holds it.
__device__ inline void sell_add_block(const SellView &M, uint32_t r, uint32_t c, float out[9]) {
  if (M.nb == 0) return;
  uint32_t base = M.slice_ptr[r >> 6], lane = r & 63u, lo = 0, hi = M.row_len[r];
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (M.col[base + 64u * mid + lane] < c) lo = mid + 1; else hi = mid;
  }
  if (lo < M.row_len[r] && M.col[base + 64u * lo + lane] == c) {
    uint32_t sl = base + 64u * lo + lane;
    for (int e = 0; e < 9; ++e) out[e] += M.val[sell_val_index(sl, e)];
  }
}

// Block (va, vb) of the elastic Hessian gathered so far, pinned rows and columns included (the
// reference's fixed_hess_in, implicit.rs:3175-3228): from the free system for two free
// vertices; otherwise from the static table (elements on pinned vertices only) plus the
// mixed-element table (elements joining pinned and free vertices).
__device__ inline void lookup_block(const Topo &T, const SellView &K, const SellView &S, const SellView &X, uint32_t va, uint32_t vb,
                                    float out[9]) {
  for (int e = 0; e < 9; ++e) out[e] = 0.0f;
  if (T.vdof[va] != NONE && T.vdof[vb] != NONE) {
    sell_add_block(K, T.vdof[va], T.vdof[vb], out);
    return;
  }
  if (T.sdof && T.sdof[va] != NONE && T.sdof[vb] != NONE) sell_add_block(S, T.sdof[va], T.sdof[vb], out);
  if (T.xdof && T.xdof[va] != NONE) sell_add_block(X, T.xdof[va], vb, out);
}

struct ContactOut {
  uint32_t *dof;
  float *g, *w, *coef, *n, *cgrad;  // cgrad: 24 per stencil (barrier part, friction part)
  float *kappa;
};

// implicit.rs:3026-3103 per-stencil barrier + friction, implicit.rs:755-795 Rayleigh stiffness.
__global__ __launch_bounds__(BLOCK) void k_stencil_eval(uint32_t ns, const uint32_t *__restrict__ sidx, const float *__restrict__ sd,
                                                       const float *__restrict__ sg, float ghat_vt, float ghat_ee, uint32_t nvt, Topo T,
                                                       SellView K, SellView S, SellView X, Params P, const float *__restrict__ x, const float *__restrict__ xs,
                                                       const float *__restrict__ kdx,
                                                       ContactOut C) {
  uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= ns) return;
  uint32_t id[4];
  f3 g[4];
  for (int a = 0; a < 4; ++a) {
    id[a] = sidx[4 * (size_t)s + a];
    const float *q = sg + 12 * (size_t)s + 3 * a;
    g[a] = f3{q[0], q[1], q[2]};
  }
  float d = sd[s];
  float ghat = s < nvt ? ghat_vt : ghat_ee;
  // Rayleigh stiffness (contact.rs:750-832) along the weighted normal
  float k = P.kappa;
  float nlen = len(g[0]);
  f3 nrm{0, 0, 0};
  float weights[4] = {0, 0, 0, 0};
  if (!(nlen < 1.0e-12f)) {
    nrm = g[0] / nlen;
    float mass[4];
    for (int a = 0; a < 4; ++a) {
      weights[a] = dot(g[a], nrm);
      float im = T.im[id[a]];
      mass[a] = im > 0.0f ? 1.0f / im : 0.0f;
    }
    f3 e = d * nrm;
    float gap = len(e) - 0.0f;
    float kk = 0.0f;
    if (gap > 0.0f) {
      float sqr_x = gap * gap;
      float m_max = 0.0f;
      for (int a = 0; a < 4; ++a) m_max = fmaxf(m_max, mass[a]);
      float w[12];
      for (int a = 0; a < 4; ++a) { f3 v = weights[a] * e; w[3 * a] = v.x; w[3 * a + 1] = v.y; w[3 * a + 2] = v.z; }
      float ss = 0.0f;
      for (int i = 0; i < 12; ++i) ss += w[i] * w[i];
      float norm = sqrtf(ss);
      if (norm > 0.0f) {
        for (int i = 0; i < 12; ++i) w[i] /= norm;
        float acc = 0.0f;
        for (int ii = 0; ii < 4; ++ii) {
          float m_ii = mass[ii] > 0.0f ? mass[ii] : m_max;
          for (int jj = 0; jj < 4; ++jj) {
            float blk[9];
            lookup_block(T, K, S, X, id[ii], id[jj], blk);
            for (int r = 0; r < 3; ++r) {
              float row = 0.0f;
              for (int c = 0; c < 3; ++c) {
                float v = blk[3 * r + c];
                if (ii == jj && r == c) v += m_ii / sqr_x;
                row += v * w[3 * jj + c];
              }
              acc += w[3 * ii + r] * row;
            }
          }
        }
        kk = acc;
      }
    }
    k = kk > 0.0f ? kk : P.kappa;
  }
  C.kappa[s] = k;
  int kind = d < 0.0f ? 1 : P.barrier_kind;
  double bo[3] = {0, 0, 0};
  bool act = barrier_of(kind, (double)d, (double)ghat, (double)k, bo);
  float bg = act ? (float)bo[1] : 0.0f;
  float hp = act ? fmaxf((float)bo[2], 0.0f) : 0.0f;
  float lam = 0.0f;
  f3 fn{0, 0, 0};
  f3 fg{0, 0, 0};
  float wv[4] = {0, 0, 0, 0};
  if (act && P.mu > 0.0f && !(nlen < 1.0e-12f)) {
    // contact.rs:658-706 stencil_friction
    f3 slip{0, 0, 0};
    // implicit.rs:3068-3080: a kinematic body vertex's slip includes the move it was given
    // before this step (kdx, null when the body did not move)
    for (int a = 0; a < 4; ++a) {
      wv[a] = dot(g[a], nrm);
      f3 dxa = ld3(x, id[a]) - ld3(xs, id[a]);
      if (kdx) dxa = dxa + ld3(kdx, id[a]);
      slip = slip + dxa * wv[a];
    }
    f3 f_c = nrm * bg;
    lam = friction(f_c, slip, nrm, P.mu, P.feps, fg);
    if (lam > 0.0f && fg.x == 0.0f && fg.y == 0.0f && fg.z == 0.0f) lam = 0.0f;
    fn = nrm;
  }
  for (int a = 0; a < 4; ++a) {
    C.dof[4 * (size_t)s + a] = T.vdof[id[a]];
    C.w[4 * (size_t)s + a] = wv[a];
    float *cg = C.cgrad + 24 * (size_t)s + 3 * a;
    cg[0] = bg * g[a].x; cg[1] = bg * g[a].y; cg[2] = bg * g[a].z;
    float *cf = C.cgrad + 24 * (size_t)s + 12 + 3 * a;
    f3 f = lam > 0.0f ? fg * wv[a] : f3{0, 0, 0};
    cf[0] = f.x; cf[1] = f.y; cf[2] = f.z;
    for (int c = 0; c < 3; ++c) C.g[12 * (size_t)s + 3 * a + c] = comp(g[a], c);
  }
  C.coef[2 * (size_t)s] = hp;
  C.coef[2 * (size_t)s + 1] = lam;
  C.n[3 * (size_t)s] = fn.x; C.n[3 * (size_t)s + 1] = fn.y; C.n[3 * (size_t)s + 2] = fn.z;
}

// Incidence keys: (free dof << 32) | (4 s + a); pinned participants get dof = nf (sort last).
// pos[4 s + a] = slot of (s, a) in the vertex-sorted incidence list (~0u for a pinned
// participant, which has no incidence).
__global__ __launch_bounds__(BLOCK) void k_inc_pos(uint32_t nf, uint32_t m, const uint32_t *dof, const uint32_t *inc_ptr, const uint32_t *inc,
                                                  uint32_t *pos) {
  uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= m) return;
  if (dof[k] == 0xffffffffu) pos[k] = 0xffffffffu;
  if (k < inc_ptr[nf]) pos[inc[k]] = k;
}

__global__ __launch_bounds__(BLOCK) void k_inc_keys(uint32_t ns, const uint32_t *dof, uint32_t nf, uint64_t *key, uint32_t *val) {
  uint32_t e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= 4 * ns) return;
  uint32_t d = dof[e];
  key[e] = ((uint64_t)(d == NONE ? nf : d) << 32) | e;
  val[e] = e;
}
__global__ __launch_bounds__(BLOCK) void k_inc_ptr(uint32_t nf, uint32_t m, const uint64_t *key, uint32_t *ptr, uint32_t *inc) {
  uint32_t r = blockIdx.x * BLOCK + threadIdx.x;
  if (r < m) inc[r] = (uint32_t)key[r];
  if (r > nf) return;
  uint64_t target = (uint64_t)r << 32;
  uint32_t lo = 0, hi = m;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (key[mid] < target) lo = mid + 1; else hi = mid;
  }
  ptr[r] = lo;
}

// rhs = -(elastic gradient + contact gradient in stencil order).
__global__ __launch_bounds__(BLOCK) void k_rhs(uint32_t nf, const float *__restrict__ eg, const uint32_t *inc_ptr, const uint32_t *inc,
                                              const float *cgrad, int has_contact, float *rhs) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nf) return;
  f3 g = ld3(eg, i);
  if (has_contact) {
    for (uint32_t k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
      uint32_t e = inc[k], s = e >> 2, a = e & 3u;
      const float *b = cgrad + 24 * (size_t)s + 3 * a;
      const float *f = cgrad + 24 * (size_t)s + 12 + 3 * a;
      g.x += b[0]; g.y += b[1]; g.z += b[2];
      g.x += f[0]; g.y += f[1]; g.z += f[2];
    }
  }
  st3(rhs, i, f3{-g.x, -g.y, -g.z});
}

// ---------------------------------------------------------------- step kernels
__global__ __launch_bounds__(BLOCK) void k_predict(uint32_t V, const float *im, const float *x0, const float *v, float h, f3 grav, float *xt, float *x) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V) return;
  f3 p = ld3(x0, i);
  f3 t = im[i] > 0.0f ? p + ld3(v, i) * h + grav * (h * h) : p;
  st3(xt, i, t);
  st3(x, i, p);
}

// dx on all vertices (0 on pinned) and per-block max |dx|.
__global__ __launch_bounds__(BLOCK) void k_scatter_dx(uint32_t V, const uint32_t *vdof, const float *dxf, float *dx, float *part) {
  __shared__ float lds[4];
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float m = 0.0f;
  if (i < V) {
    uint32_t d = vdof[i];
    f3 v = d == NONE ? f3{0, 0, 0} : ld3(dxf, d);
    st3(dx, i, v);
    m = len(v);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

__global__ __launch_bounds__(BLOCK) void k_max_final(uint32_t n, const float *part, float *out) {
  __shared__ float lds[4];
  float m = 0.0f;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) m = fmaxf(m, part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

__global__ __launch_bounds__(BLOCK) void k_scale(uint32_t n, float *a, float s) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) a[i] = a[i] * s;
}

// Analytic-collider CCD clamp (implicit.rs:1792-1817), min into tmin (double bits).
__global__ __launch_bounds__(BLOCK) void k_ccd_colliders(uint32_t V, const float *im, const float *x, const float *dx, Params P,
                                                        unsigned long long *tmin) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V || !(im[i] > 0.0f)) return;
  const float GAP = 0.2f;
  float t = 1.0f;
  f3 xi = ld3(x, i), di = ld3(dx, i);
  for (uint32_t k = 0; k < P.ncoll; ++k) {
    float d0;
    f3 n;
    collider_distance(P.coll[k], xi, d0, n);
    if (d0 <= 0.0f) continue;
    if (P.coll[k].kind == NGUO_COLLIDER_HALFSPACE) {
      float rate = dot(di, n);
      if (rate < 0.0f) t = fminf(t, (GAP - 1.0f) * d0 / rate);
    } else {
      float l = len(di);
      if (l > 1.0e-12f) t = fminf(t, (1.0f - GAP) * d0 / l);
    }
  }
  if (t < 1.0f) atomicMin(tmin, (unsigned long long)__double_as_longlong((double)t));
}

// Vertex-triangle ACCD over the padded-AABB candidates (implicit.rs:1856-1900), FP64 TOI.
// A pair whose advancement is still running after ACCD_INLINE iterations is handed to
// k_ccd_deferred (one pair per thread, full iteration limit, same arithmetic from the start,
// so the same time of impact): a traversal thread never walks a long series of slow pairs.
constexpr uint32_t ACCD_INLINE = 16;

// Exclusion before the FP64 advancement: at its first iterate accd_toi returns its horizon
// (no time of impact, so the pair cannot lower the step) when 1.25 u_max < 0.98 d0, with
// d0 the pair's distance (eps = 0.01 d0). The gap between the two point sets' boxes is a lower
// bound of d0 and the FP32 u_max an upper bound of the FP64 one once widened; the margins
// cover the FP32 rounding, so a pair skipped here is one the FP64 test skips too.
__device__ inline bool accd_surely_clear(const f3 *pa, int na, const f3 *pb, int nb, const f3 dxs[4]) {
  f3 alo = pa[0], ahi = pa[0], blo = pb[0], bhi = pb[0];
  for (int k = 1; k < na; ++k) { alo = vmin(alo, pa[k]); ahi = vmax(ahi, pa[k]); }
  for (int k = 1; k < nb; ++k) { blo = vmin(blo, pb[k]); bhi = vmax(bhi, pb[k]); }
  const float gx = fmaxf(fmaxf(blo.x - ahi.x, alo.x - bhi.x), 0.0f);
  const float gy = fmaxf(fmaxf(blo.y - ahi.y, alo.y - bhi.y), 0.0f);
  const float gz = fmaxf(fmaxf(blo.z - ahi.z, alo.z - bhi.z), 0.0f);
  const float gap = sqrtf(gx * gx + gy * gy + gz * gz) * (1.0f - 1.0e-5f) - 1.0e-9f;
  float u = 0.0f;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) u = fmaxf(u, len(dxs[i] - dxs[j]));
  return 1.25f * (u * (1.0f + 1.0e-5f) + 1.0e-12f) < 0.98f * gap;
}

// The same argument bounds the search (dq: the query's largest |dx|, every vertex has
// |dx| <= max_disp). u_max, the largest |dx_i - dx_j| over the pair's four vertices, is at
// most 2 max_disp, so a pair whose boxes are more than 1.3 (2 max_disp) apart returns the
// horizon at the first iterate. A tighter bound holds for the pair's distance: every point
// of one primitive moves relative to every point of the other by at most W = dq + max_disp,
// so with boxes more than 1.3 W apart the distance stays above 0.0385 d0 for t <= 1.25; the
// advancement then moves by at least 0.0185 of the step per iteration and returns the
// horizon within 68 iterations, never a time of impact. With an iteration limit of at least
// CCD_TIGHT_ITERS the query's padding is capped at 1.3 W, otherwise at 1.3 (2 max_disp); the
// padding is the smaller of the widening pass's and that cap.
constexpr uint32_t CCD_TIGHT_ITERS = 80;
__device__ inline float ccd_query_pad(float pad, float dq, float max_disp, uint32_t iters) {
  const float w = iters >= CCD_TIGHT_ITERS ? dq + max_disp : 2.0f * max_disp;
  return fminf(pad, 1.3f * w * (1.0f + 1.0e-5f) + 1.0e-9f);
}
// A second bound of the same kind from the displacements' spread: every vertex of a candidate
// pair has each component of its dx within [lo, hi], the union of the query's own range and
// the other side's (the movable primitives' range, k_dx_range, or 0 for a tree whose vertices
// do not move), so u_max, the largest |dx_i - dx_j| over the pair's four vertices, is at most
// U = |hi - lo|, and boxes more than 1.3 U apart are surely clear (accd_surely_clear) and
// return the horizon. When the cloth moves nearly rigidly (falling as a whole), U is far below
// 2 max_disp; the pairs it drops return the horizon either way, so the result is unchanged.
__device__ inline float ccd_spread_pad(float pad, f3 qlo, f3 qhi, f3 olo, f3 ohi) {
  if (!(olo.x <= ohi.x)) return pad;  // no movable vertex recorded: no bound
  const f3 r{fmaxf(qhi.x, ohi.x) - fminf(qlo.x, olo.x), fmaxf(qhi.y, ohi.y) - fminf(qlo.y, olo.y),
             fmaxf(qhi.z, ohi.z) - fminf(qlo.z, olo.z)};
  return fminf(pad, 1.3f * len(r) * (1.0f + 1.0e-5f) + 1.0e-9f);
}
// Order-preserving map of a float to a u32 (atomic min / max on the keys).
__device__ inline uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float fkey_inv(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }
// The movable primitives' displacement range from k_dx_range (keys: min xyz, max xyz).
__device__ inline void dx_range(const uint32_t *keys, f3 &lo, f3 &hi) {
  lo = f3{fkey_inv(keys[0]), fkey_inv(keys[1]), fkey_inv(keys[2])};
  hi = f3{fkey_inv(keys[3]), fkey_inv(keys[4]), fkey_inv(keys[5])};
}
// keys = (+inf x3, -inf x3), or the range {0} when a movable primitive has a fixed vertex
__global__ void k_dx_range_init(uint32_t *keys, int with_zero) {
  const uint32_t t = threadIdx.x;
  if (t < 6) keys[t] = with_zero ? fkey(0.0f) : fkey(t < 3 ? INFINITY : -INFINITY);
}
// Component-wise min / max of dx over the free vertices (every vertex of a movable primitive
// is free or fixed with dx = 0, the latter covered by k_dx_range_init's with_zero).
__global__ __launch_bounds__(BLOCK) void k_dx_range(uint32_t V, const uint32_t *__restrict__ vdof, const float *__restrict__ dx, uint32_t *keys) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  const bool ok = i < V && vdof[i] != NONE;
  const f3 d = ok ? ld3(dx, i) : f3{0, 0, 0};
  float lo[3] = {ok ? d.x : INFINITY, ok ? d.y : INFINITY, ok ? d.z : INFINITY};
  float hi[3] = {ok ? d.x : -INFINITY, ok ? d.y : -INFINITY, ok ? d.z : -INFINITY};
  for (int c = 0; c < 3; ++c) {
    lo[c] = wave_min(lo[c]);
    hi[c] = wave_max(hi[c]);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0)
    for (int c = 0; c < 3; ++c) {
      if (lo[c] <= hi[c]) {
        atomicMin(&keys[c], fkey(lo[c]));
        atomicMax(&keys[3 + c], fkey(hi[c]));
      }
    }
}
struct CcdDefer {
  uint32_t *count;  // pairs appended
  uint32_t *pairs;  // [2 cap]: (a | kind << 31, b): vertex-triangle (v, t) or edge-edge (lo, hi)
  uint32_t cap;
};
__device__ inline bool defer_pair(const CcdDefer &D, uint32_t a, uint32_t b, uint32_t kind) {
  const uint32_t k = atomicAdd(D.count, 1u);
  if (k >= D.cap) return false;
  D.pairs[2 * (size_t)k] = a | (kind << 31);
  D.pairs[2 * (size_t)k + 1] = b;
  return true;
}

// Leaves load their positions, motions and flags in one memory round trip (the triangle's
// vertex ids come with the traversal record) and test the filters from registers: the
// filters are predicates, so their order changes nothing but the latency.
__global__ __launch_bounds__(BLOCK) void k_ccd_vt(Topo T, BvhView dyn, BvhView stat, const float *__restrict__ x, const float *__restrict__ dx,
                                                 float pad0, float max_disp, uint32_t iters, const uint32_t *__restrict__ dxr, CcdDefer D,
                                                 unsigned long long *tmin, uint32_t *err) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t v = blockIdx.x * BLOCK + threadIdx.x;
  double best = 1.0;
  if (v < T.V) {
    f3 p = ld3(x, v);
    const f3 dvv = ld3(dx, v);
    bool mv = movable(T, v);
    const bool below = T.side ? T.side[v] < 0.0f : false;
    ExcList exl;
    exl.load(T, v);
    // parts 0 .. TRAV_PARTS-1: the movable tree's subtrees; the rest: the static tree's, whose
    // vertices do not move
    const uint32_t part = blockIdx.y;
    f3 olo{0, 0, 0}, ohi{0, 0, 0};
    if (part < TRAV_PARTS) dx_range(dxr, olo, ohi);
    const float pad = ccd_spread_pad(ccd_query_pad(pad0, len(dvv), max_disp, iters), dvv, dvv, olo, ohi);
    auto leaf = [&](uint32_t t, uint3 vid) {
      const uint32_t t0 = vid.x, t1 = vid.y, t2 = vid.z;
      f3 a = ld3(x, t0), b = ld3(x, t1), c = ld3(x, t2);
      const f3 da = ld3(dx, t0), db = ld3(dx, t1), dc = ld3(dx, t2);
      const float m0 = T.im[t0], m1 = T.im[t1], m2 = T.im[t2];
      const bool st_all = T.static_v ? (T.static_v[t0] && T.static_v[t1] && T.static_v[t2]) : false;
      if (!inside_padded(p, vmin(vmin(a, b), c), vmax(vmax(a, b), c), pad)) return;
      bool tm = m0 > 0.0f || m1 > 0.0f || m2 > 0.0f;
      if (!tm && !mv) return;
      if (exl.has(T, t0, t1, t2)) return;
      if (st_all && below) return;
      const f3 dv[4] = {dvv, da, db, dc};
      const f3 tri[3] = {a, b, c};
      if (accd_surely_clear(&p, 1, tri, 3, dv)) return;
      d3 x0[4] = {to_d(p), to_d(a), to_d(b), to_d(c)};
      d3 d4[4] = {to_d(dv[0]), to_d(dv[1]), to_d(dv[2]), to_d(dv[3])};
      bool capped = false;
      double toi = accd_toi(x0, d4, 0, 0.0, 1.25, 0.01, min(iters, ACCD_INLINE), &capped);
      if (capped && iters > ACCD_INLINE) {
        if (defer_pair(D, v, t, 0u)) return;
        toi = accd_toi(x0, d4, 0, 0.0, 1.25, 0.01, iters);
      }
      best = fmin(best, toi / 1.25);
    };
    auto ov = [&](f3 lo, f3 hi) { return inside_padded(p, lo, hi, pad); };
    if (part < TRAV_PARTS) bvh_traverse_part(dyn, part, err, tstack, ov, leaf);
    else if (mv) bvh_traverse_part(stat, part - TRAV_PARTS, err, tstack, ov, leaf);
  }
  // one atomic per wavefront: the minimum is order-independent
  const double wb = wave_min(best);
  if ((threadIdx.x & (WAVE - 1)) == 0 && wb < 1.0) atomicMin(tmin, (unsigned long long)__double_as_longlong(wb));
}

// Edge-edge ACCD (implicit.rs:1902-1936).
__global__ __launch_bounds__(BLOCK) void k_ccd_ee(Topo T, BvhView tree, const uint32_t *qids, uint32_t nq, int self_query, const float *__restrict__ x,
                                                 const float *__restrict__ dx, float pad0, float max_disp, uint32_t iters,
                                                 const uint32_t *__restrict__ dxr, CcdDefer D, unsigned long long *tmin, uint32_t *err) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t q = blockIdx.x * BLOCK + threadIdx.x;
  double best = 1.0;
  if (q < nq) {
    uint32_t ej = qids[q];
    uint32_t b0 = T.edges[2 * (size_t)ej], b1 = T.edges[2 * (size_t)ej + 1];
    f3 pb = ld3(x, b0), qb = ld3(x, b1);
    const f3 db0 = ld3(dx, b0), db1 = ld3(dx, b1);
    // the other side: the movable edges (self query) or the static edges (dx = 0)
    f3 olo{0, 0, 0}, ohi{0, 0, 0};
    if (self_query) dx_range(dxr, olo, ohi);
    const float pad = ccd_spread_pad(ccd_query_pad(pad0, fmaxf(len(db0), len(db1)), max_disp, iters), vmin(db0, db1), vmax(db0, db1), olo, ohi);
    f3 blo = vmin(pb, qb) - f3{pad, pad, pad}, bhi = vmax(pb, qb) + f3{pad, pad, pad};
    auto ov = [&](f3 lo, f3 hi) { return !(lo.x > bhi.x || lo.y > bhi.y || lo.z > bhi.z || hi.x < blo.x || hi.y < blo.y || hi.z < blo.z); };
    // a leaf loads its edge's positions and motions in one memory round trip (the vertex ids
    // come with the traversal record), and reuses the query edge's from registers. A query of the tree's own edges meets
    // each other edge once: only the leaves after its own sorted slot (whole subtrees before
    // it are skipped).
    const uint32_t min_slot = self_query ? tree.slot[q] + 1 : 0u;
    bvh_traverse_part(tree, blockIdx.y, err, tstack, ov, [&](uint32_t ei, uint3 vid) {
      if (ei == ej) return;
      const uint32_t a0 = vid.x, a1 = vid.y;
      if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
      f3 pa = ld3(x, a0), qa = ld3(x, a1);
      const f3 da0 = ld3(dx, a0), da1 = ld3(dx, a1);
      if (!ov(vmin(pa, qa), vmax(pa, qa))) return;
      // the pair's vertices in edge-index order: the lower edge's two, then the higher's
      const bool a_lo = ei < ej;
      const uint32_t lo_e = min(ei, ej), hi_e = max(ei, ej);
      const f3 xs[4] = {a_lo ? pa : pb, a_lo ? qa : qb, a_lo ? pb : pa, a_lo ? qb : qa};
      const f3 dv[4] = {a_lo ? da0 : db0, a_lo ? da1 : db1, a_lo ? db0 : da0, a_lo ? db1 : da1};
      if (accd_surely_clear(xs, 2, xs + 2, 2, dv)) return;
      d3 x0[4] = {to_d(xs[0]), to_d(xs[1]), to_d(xs[2]), to_d(xs[3])};
      d3 d4[4] = {to_d(dv[0]), to_d(dv[1]), to_d(dv[2]), to_d(dv[3])};
      bool capped = false;
      double toi = accd_toi(x0, d4, 1, 0.0, 1.25, 0.01, min(iters, ACCD_INLINE), &capped);
      if (capped && iters > ACCD_INLINE) {
        if (defer_pair(D, lo_e, hi_e, 1u)) return;
        toi = accd_toi(x0, d4, 1, 0.0, 1.25, 0.01, iters);
      }
      best = fmin(best, toi / 1.25);
    }, min_slot);
  }
  // one atomic per wavefront: the minimum is order-independent
  const double wb = wave_min(best);
  if ((threadIdx.x & (WAVE - 1)) == 0 && wb < 1.0) atomicMin(tmin, (unsigned long long)__double_as_longlong(wb));
}

// The deferred pairs of one pass, one per thread, with the full iteration limit.
__global__ __launch_bounds__(BLOCK) void k_ccd_deferred(Topo T, CcdDefer D, const float *__restrict__ x, const float *__restrict__ dx, uint32_t iters,
                                                       unsigned long long *tmin) {
  const uint32_t k = blockIdx.x * BLOCK + threadIdx.x;
  if (k >= min(*D.count, D.cap)) return;
  const uint32_t w0 = D.pairs[2 * (size_t)k], b = D.pairs[2 * (size_t)k + 1];
  const uint32_t kind = w0 >> 31, a = w0 & 0x7fffffffu;
  uint32_t i[4];
  if (kind == 0) {
    i[0] = a; i[1] = T.tris[3 * (size_t)b]; i[2] = T.tris[3 * (size_t)b + 1]; i[3] = T.tris[3 * (size_t)b + 2];
  } else {
    i[0] = T.edges[2 * (size_t)a]; i[1] = T.edges[2 * (size_t)a + 1]; i[2] = T.edges[2 * (size_t)b]; i[3] = T.edges[2 * (size_t)b + 1];
  }
  d3 x0[4], d4[4];
  for (int q = 0; q < 4; ++q) { x0[q] = to_d(ld3(x, i[q])); d4[q] = to_d(ld3(dx, i[q])); }
  const double toi = accd_toi(x0, d4, (int)kind, 0.0, 1.25, 0.01, iters) / 1.25;
  if (toi < 1.0) atomicMin(tmin, (unsigned long long)__double_as_longlong(toi));
}

// x += t dx; per-block max |t dx|.
__global__ __launch_bounds__(BLOCK) void k_apply(uint32_t V, float *x, const float *dx, float t, float *part) {
  __shared__ float lds[4];
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float m = 0.0f;
  if (i < V) {
    f3 d = ld3(dx, i) * t;
    st3(x, i, ld3(x, i) + d);
    m = len(d);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

// implicit.rs:2032-2064: floor projection, finiteness guard, velocity update, commit.
__global__ __launch_bounds__(BLOCK) void k_finite(uint32_t V, const float *x, uint32_t *bad) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V) return;
  f3 p = ld3(x, i);
  if (!(isfin(p.x) && isfin(p.y) && isfin(p.z))) atomicOr(bad, 1u);
}
__global__ __launch_bounds__(BLOCK) void k_commit(uint32_t V, const float *im, float *x, float *x0, float *v, float inv_h, float floor_y,
                                                 const uint32_t *bad, float *part) {
  __shared__ float lds[4];
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float m = 0.0f;
  if (i < V) {
    f3 p = ld3(x, i), p0 = ld3(x0, i);
    bool free_v = im[i] > 0.0f;
    if (free_v && isfin(floor_y) && p.y < floor_y) p.y = floor_y;
    if (*bad == 0u) {
      f3 vel = free_v ? (p - p0) * inv_h : f3{0, 0, 0};
      st3(v, i, vel);
      st3(x0, i, p);
      st3(x, i, p);
      if (free_v) m = len(vel);
    } else {
      st3(v, i, f3{0, 0, 0});
      st3(x, i, p0);
    }
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
}

// implicit.rs:2212-2230 count_active_contacts and the analytic-collider part of
// penetration_free (implicit.rs:2235-2266), per free vertex: out[0] counts the vertices within
// dhat of a collider, out[1] flags one more than 1e-3 max(dhat, 1e-6) inside one.
__global__ __launch_bounds__(BLOCK) void k_collider_metrics(uint32_t V, const float *im, const float *x, Params P, uint32_t *out) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= V || !(im[i] > 0.0f)) return;
  const f3 p = ld3(x, i);
  const float tol = -1.0e-3f * fmaxf(P.dhat, 1.0e-6f);
  bool near = false, pen = false;
  for (uint32_t k = 0; k < P.ncoll; ++k) {
    float d;
    f3 n;
    collider_distance(P.coll[k], p, d, n);
    near = near || d < P.dhat;
    pen = pen || d < tol;
  }
  if (near) atomicAdd(out, 1u);
  if (pen) atomicOr(out + 1, 1u);
}

// implicit.rs:2193-2209 max_stretch_ratio: the largest |x_a - x_b| / rest over the mesh's
// unique edges (state.rs:39-65), 1 for an edge with rest <= 1e-9; a NaN ratio drops out as
// f32::max drops it. Ratios are non-negative, so their bit patterns order as the values.
__global__ __launch_bounds__(BLOCK) void k_max_sigma(uint32_t n, const uint32_t *__restrict__ ev, const float *__restrict__ rest,
                                                    const float *__restrict__ x, uint32_t *out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float r = 0.0f;
  if (i < n) {
    const f3 d = ld3(x, ev[2 * (size_t)i]) - ld3(x, ev[2 * (size_t)i + 1]);
    const float l = sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    r = rest[i] > 1.0e-9f ? l / rest[i] : 1.0f;
    if (!(r >= 0.0f)) r = 0.0f;
  }
  r = wave_max(r);
  if ((threadIdx.x & 63) == 0 && r > 0.0f) atomicMax(out, __float_as_uint(r));
}

// implicit.rs:618-650 positions_self_intersect: any non-inert, non-adjacent (edge, triangle)
// with a transverse pierce. Each edge queries both triangle trees (movable edges) or the
// dynamic one (static edges).
__global__ __launch_bounds__(BLOCK) void k_pierce(Topo T, BvhView dyn, BvhView stat, uint32_t E, const float *__restrict__ x, uint32_t *hit,
                                                 uint32_t *err) {
  __shared__ uint32_t tstack[TRAV_LDS * BLOCK];
  uint32_t ei = blockIdx.x * BLOCK + threadIdx.x;
  if (ei >= E) return;
  uint32_t e0 = T.edges[2 * (size_t)ei], e1 = T.edges[2 * (size_t)ei + 1];
  f3 p = ld3(x, e0), q = ld3(x, e1);
  f3 lo = vmin(p, q), hi = vmax(p, q);
  bool em = movable(T, e0) || movable(T, e1);
  auto ov = [&](f3 l, f3 h) { return !(l.x > hi.x || l.y > hi.y || l.z > hi.z || h.x < lo.x || h.y < lo.y || h.z < lo.z); };
  auto leaf = [&](uint32_t t) {
    uint32_t t0 = T.tris[3 * (size_t)t], t1 = T.tris[3 * (size_t)t + 1], t2 = T.tris[3 * (size_t)t + 2];
    if (e0 == t0 || e0 == t1 || e0 == t2 || e1 == t0 || e1 == t1 || e1 == t2) return;
    if (edge_triangle_intersect(p, q, ld3(x, t0), ld3(x, t1), ld3(x, t2))) atomicOr(hit, 1u);
  };
  bvh_traverse(dyn, err, tstack, ov, leaf);
  if (em) bvh_traverse(stat, err, tstack, ov, leaf);
}

}  // namespace nguo

// ================================================================ host side
using namespace nguo;

// ---------------------------------------------------------------- step energy (diagnostic)
// The step energy, the incremental potential each Newton iteration decreases, at the current
// positions: inertia, shells, hinges, rod stretch, strand bending, the barriers and friction
// potentials of the active stencils and of the analytic colliders. Each term is the one
// whose gradient the assembly adds (implicit.rs:2392-3103): inertia 1/2 m/h^2 |x - x~|^2,
// shells area psi(sigma) (shell_model.rs:112-127), hinges 1/2 k_b (theta - theta0)^2
// (bending.rs:105-123), rods and edge springs 1/2 k (l - l0)^2 (implicit.rs:3287-3308, :1-10), strand bends
// 1/2 k (theta - theta0)^2 (rod_bend.rs:1-20), barriers (contact.rs:580-591, 1014-1025) and the
// smoothed Coulomb potential c f0(|P u|) whose gradient is contact.rs:613-706. Everything is
// evaluated in FP64 from the FP32 positions (gaps from the stencils' own witness weights,
// Rayleigh stiffnesses from the assembled blocks, shells from FP64 rest frames), so the value
// is that of the FP64 scheme at these positions. It is a diagnostic, run on request only.
enum : int { EN_INERTIA, EN_SHELL, EN_HINGE, EN_ROD, EN_ROD_BEND, EN_BARRIER, EN_FRICTION, EN_COLLIDER, EN_SPRING, EN_TERMS };

// Sums each term over the workgroup's wavefronts into acc (all lanes active).
__device__ inline void energy_commit(const double e[EN_TERMS], double *acc) {
#pragma unroll
  for (int k = 0; k < EN_TERMS; ++k) {
    const double v = wave_sum(e[k]);
    if ((threadIdx.x & (WAVE - 1)) == 0 && v != 0.0) atomicAdd(acc + k, v);
  }
}

// c f0(|P u|), f0(s) = s^2 / (2 eps) up to eps, then s - eps / 2: the potential whose gradient
// is contact.rs:613-643 friction with its normal force (c = mu times it) and normal held.
__device__ inline double friction_potential(double c, d3 n, d3 u, double eps) {
  const double sl = len(u - n * dot(n, u));
  return c * (sl <= eps ? sl * sl / (2.0 * eps) : sl - 0.5 * eps);
}

// Inertia and analytic colliders, one thread per vertex.
__global__ __launch_bounds__(BLOCK) void k_energy_verts(uint32_t V, const float *__restrict__ im, const float *__restrict__ x,
                                                       const float *__restrict__ xt, const float *__restrict__ xs, double inv_h2, Params P,
                                                       double *acc) {
  double e[EN_TERMS] = {};
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < V && im[i] > 0.0f) {
    const d3 xi = to_d(ld3(x, i)), dv = xi - to_d(ld3(xt, i));
    e[EN_INERTIA] = 0.5 * (1.0 / (double)im[i]) * inv_h2 * dot(dv, dv);
    const double eps = fmax((double)P.feps, 1.0e-12);
    for (uint32_t k = 0; k < P.ncoll; ++k) {
      const nguo_collider &c = P.coll[k];
      double dd;
      d3 n;
      if (c.kind == NGUO_COLLIDER_HALFSPACE) {
        n = d3{c.b[0], c.b[1], c.b[2]};
        dd = dot(xi - d3{c.a[0], c.a[1], c.a[2]}, n);
      } else {
        const d3 v = xi - d3{c.a[0], c.a[1], c.a[2]};
        const double l = len(v);
        if (l < 1.0e-9) { dd = -(double)c.radius; n = d3{0, 1, 0}; }
        else { dd = l - (double)c.radius; n = v / l; }
      }
      double bo[3];
      if (!barrier_of(P.barrier_kind, dd, (double)P.dhat, (double)P.kappa, bo)) continue;
      e[EN_COLLIDER] += bo[0];
      const double cf = -(double)P.mu * bo[1];
      if (P.mu > 0.0f && cf > 0.0) e[EN_COLLIDER] += friction_potential(cf, n, xi - to_d(ld3(xs, i)), eps);
    }
  }
  energy_commit(e, acc);
}

// Shells, hinges, rod segments, strand bends and edge springs of one element set, one thread
// per element.
__global__ __launch_bounds__(BLOCK) void k_energy_elems(ElemDev E, const double *__restrict__ dm64, const double *__restrict__ area64, Params P,
                                                       const float *__restrict__ x, double *acc) {
  double e[EN_TERMS] = {};
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t b1 = E.ns, b2 = b1 + E.nh, b3 = b2 + E.nr, b4 = b3 + E.nrh, b5 = b4 + E.nsp;
  if (i < b1) {
    const uint32_t *v = E.sh_v + 3 * (size_t)i;
    const d3 p0 = to_d(ld3(x, v[0])), e1 = to_d(ld3(x, v[1])) - p0, e2 = to_d(ld3(x, v[2])) - p0;
    const double sc = (double)E.sh_scale[i];
    const double *dm = dm64 + 4 * (size_t)i;
    const d3 f0 = e1 * (dm[0] / sc) + e2 * (dm[1] / sc), f1 = e1 * (dm[2] / sc) + e2 * (dm[3] / sc);
    // eigenvalues of F^T F - I, then sigma - 1 = l / (sigma + 1) and J - 1 without cancellation
    const double c00 = dot(f0, f0) - 1.0, c11 = dot(f1, f1) - 1.0, c01 = dot(f0, f1);
    const double mm = 0.5 * (c00 + c11), hd = 0.5 * (c00 - c11), r = sqrt(hd * hd + c01 * c01);
    const double l1 = mm + r, l2 = mm - r;
    const double s1 = sqrt(fmax(1.0 + l1, 0.0)), s2 = sqrt(fmax(1.0 + l2, 0.0));
    const double a1 = l1 / (s1 + 1.0), a2 = l2 / (s2 + 1.0);
    const double jm = (l1 + l2 + l1 * l2) / (s1 * s2 + 1.0);
    const double mu = P.shell_mu, lam = P.shell_lambda;
    const double dev = P.shell_model == NGUO_SHELL_STVK ? mu * (a1 * a1 * a1 * a1 + a2 * a2 * a2 * a2) / 4.0 : 0.5 * mu * (a1 * a1 + a2 * a2);
    e[EN_SHELL] = area64[i] * sc * sc * (dev + 0.5 * lam * jm * jm);
  } else if (i < b2) {
    const uint32_t h = i - b1;
    const uint32_t *v = E.hg_v + 4 * (size_t)h;
    const d3 p[4] = {to_d(ld3(x, v[0])), to_d(ld3(x, v[1])), to_d(ld3(x, v[2])), to_d(ld3(x, v[3]))};
    const double w = dihedral_angle(p) - (double)E.hg_rest[h];
    e[EN_HINGE] = 0.5 * (double)P.kb * w * w;  // hinge weight 1
  } else if (i < b3) {
    const uint32_t r = i - b2;
    const double l = len(to_d(ld3(x, E.rd_v[2 * (size_t)r])) - to_d(ld3(x, E.rd_v[2 * (size_t)r + 1]))) - (double)E.rd_rest[r];
    e[EN_ROD] = 0.5 * (double)E.rd_k[r] * l * l;
  } else if (i < b4) {
    const uint32_t r = i - b3;
    const float stiff = E.rh_stiff[r];
    if (stiff > 0.0f) {
      const uint32_t *v = E.rh_v + 3 * (size_t)r;
      const d3 c = to_d(ld3(x, v[1])), e0 = to_d(ld3(x, v[0])) - c, e1 = to_d(ld3(x, v[2])) - c;
      const double w = atan2(len(cross(e0, e1)), dot(e0, e1)) - (double)E.rh_rest[r];
      e[EN_ROD_BEND] = 0.5 * (double)stiff * w * w;
    }
  } else if (i < b5) {
    const uint32_t sp = i - b4;
    const double l = len(to_d(ld3(x, E.sp_v[2 * (size_t)sp])) - to_d(ld3(x, E.sp_v[2 * (size_t)sp + 1]))) - E.sp_rest64[sp];
    e[EN_SPRING] = 0.5 * (double)P.ksp * l * l;
  }
  energy_commit(e, acc);
}

// Barrier and friction potentials of the active stencils, one thread per stencil: the
// stiffness and the friction quantities as k_stencil_eval forms them (normal from the first
// participant's distance gradient, proximity weights, the contact_kappa fallback), all from the
// FP64 gap and gradient.
__global__ __launch_bounds__(BLOCK) void k_energy_contacts(uint32_t ns, const uint32_t *__restrict__ sidx, const float *__restrict__ sd,
                                                          float ghat_vt, float ghat_ee, uint32_t nvt, Topo T,
                                                          SellView K, SellView S, SellView X, Params P, const float *__restrict__ x,
                                                          const float *__restrict__ xs, const float *__restrict__ kdx, double *acc) {
  double e[EN_TERMS] = {};
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s < ns) {
    uint32_t id[4];
    for (int a = 0; a < 4; ++a) id[a] = sidx[4 * (size_t)s + a];
    const float d32 = sd[s];
    const double ghat = s < nvt ? ghat_vt : ghat_ee;
    d3 g[4];
    const double d = stencil_geom64(s < nvt, id, x, d32, g);
    double k = P.kappa, w[4] = {0, 0, 0, 0};
    d3 n{0, 0, 0};
    const bool has_n = !(len(g[0]) < 1.0e-12);
    if (has_n) {
      n = g[0] / len(g[0]);
      double mass[4], m_max = 0.0;
      for (int a = 0; a < 4; ++a) {
        w[a] = dot(g[a], n);
        const float imv = T.im[id[a]];
        mass[a] = imv > 0.0f ? 1.0 / (double)imv : 0.0;
        m_max = fmax(m_max, mass[a]);
      }
      const double gap = fabs(d);
      double kk = 0.0;
      if (gap > 0.0) {
        double q[12], ss = 0.0;
        for (int a = 0; a < 4; ++a) {
          const d3 v = n * (w[a] * d);
          q[3 * a] = v.x; q[3 * a + 1] = v.y; q[3 * a + 2] = v.z;
        }
        for (int c = 0; c < 12; ++c) ss += q[c] * q[c];
        const double nq = sqrt(ss);
        if (nq > 0.0) {
          for (int c = 0; c < 12; ++c) q[c] /= nq;
          for (int ii = 0; ii < 4; ++ii) {
            const double m_ii = mass[ii] > 0.0 ? mass[ii] : m_max;
            for (int jj = 0; jj < 4; ++jj) {
              float blk[9];
              lookup_block(T, K, S, X, id[ii], id[jj], blk);
              for (int r = 0; r < 3; ++r) {
                double row = 0.0;
                for (int c = 0; c < 3; ++c) {
                  double v = (double)blk[3 * r + c];
                  if (ii == jj && r == c) v += m_ii / (gap * gap);
                  row += v * q[3 * jj + c];
                }
                kk += q[3 * ii + r] * row;
              }
            }
          }
        }
      }
      k = kk > 0.0 ? kk : (double)P.kappa;
    }
    const int kind = d32 < 0.0f ? 1 : P.barrier_kind;
    double bo[3];
    if (barrier_of(kind, d, ghat, k, bo)) {
      e[EN_BARRIER] = bo[0];
      const double cf = -(double)P.mu * bo[1];
      if (P.mu > 0.0f && has_n && cf > 0.0) {
        d3 slip{0, 0, 0};
        for (int a = 0; a < 4; ++a) {
          d3 dxa = to_d(ld3(x, id[a])) - to_d(ld3(xs, id[a]));
          if (kdx) dxa = dxa + to_d(ld3(kdx, id[a]));
          slip = slip + dxa * w[a];
        }
        e[EN_FRICTION] = friction_potential(cf, n, slip, fmax((double)P.feps, 1.0e-12));
      }
    }
  }
  energy_commit(e, acc);
}

struct ElemHost {
  std::vector<uint32_t> sh_v, hg_v, rd_v, rh_v, sp_v;
  std::vector<uint32_t> sh_face;  // scene shell face index of each shell
  std::vector<float> sh_dm, sh_area, hg_rest, rd_rest, rd_k, rh_rest, rh_stiff, sp_rest;
  std::vector<double> sh_dm64, sh_area64, sp_rest64;  // rest frames and lengths in FP64 (energy evaluation)
  uint32_t ns() const { return (uint32_t)sh_area.size(); }
  uint32_t nh() const { return (uint32_t)hg_rest.size(); }
  uint32_t nr() const { return (uint32_t)rd_rest.size(); }
  uint32_t nrh() const { return (uint32_t)rh_rest.size(); }
  uint32_t nsp() const { return (uint32_t)sp_rest.size(); }
};

struct ElemSet {
  uint32_t ns = 0, nh = 0, nr = 0, nrh = 0, nsp = 0;
  DevBuf<uint32_t> sh_v, hg_v, rd_v, rh_v, sp_v;
  DevBuf<float> sh_dm, sh_area, hg_rest, rd_rest, rd_k, rh_rest, rh_stiff, sp_rest;
  DevBuf<float> sh_scale;
  DevBuf<double> sh_dm64, sh_area64, sp_rest64;
  DevBuf<uint32_t> sh_face;
  std::vector<uint32_t> h_sh_face;
  std::vector<float> h_sh_scale;  // host copy (kept alive for the asynchronous upload)
  DevBuf<float> out_dense, out_dg, out_rod, out_hinge;
  SellMatrix M;
  DevBuf<uint32_t> slot_row, hc_ptr, hc, gc_ptr, gc, row_vert;
  ElemDev dev() const {
    return ElemDev{ns, nh, nr, nrh, nsp, sh_v.p, sh_dm.p, sh_area.p, sh_scale.p, hg_v.p, hg_rest.p, rd_v.p, rd_rest.p, rd_k.p,
                   rh_v.p, rh_rest.p, rh_stiff.p, sp_v.p, sp_rest.p, sp_rest64.p, out_dense.p, out_dg.p, out_rod.p, out_hinge.p};
  }
};

struct SolveGate;

// Wall time of the step's phases, for diagnostics: with NGUO_PHASE_TIMES=1 in the environment
// at step creation, every phase boundary synchronises the stream and each step prints one line
// of milliseconds per phase to stderr. Off (no synchronisation, nothing printed) otherwise.
enum Phase { PH_ASSEMBLE, PH_CONTACTS, PH_DOMAINS, PH_SOLVE, PH_CCD, PH_UPDATE, PH_POST, PH_N };
struct PhaseClock {
  bool on = false;
  double ms[PH_N] = {};
  std::chrono::steady_clock::time_point t;
  void start(hipStream_t s) {
    if (!on) return;
    (void)hipStreamSynchronize(s);
    for (double &m : ms) m = 0.0;
    t = std::chrono::steady_clock::now();
  }
  void mark(int ph, hipStream_t s) {
    if (!on) return;
    (void)hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    ms[ph] += std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
  }
  void report(double time, uint32_t verts) const {
    if (!on) return;
    static const char *names[PH_N] = {"assemble", "contacts", "domains", "solve", "ccd", "update", "post"};
    double tot = 0.0;
    std::string line;
    char buf[64];
    for (int i = 0; i < PH_N; ++i) {
      tot += ms[i];
      std::snprintf(buf, sizeof buf, " %s %.3f", names[i], ms[i]);
      line += buf;
    }
    std::fprintf(stderr, "nguo phase ms V=%u t=%.4f:%s total %.3f\n", verts, time, line.c_str(), tot);
  }
};

struct nguo_step {
  nguo_ctx *ctx = nullptr;
  hipStream_t s = nullptr;
  nguo_params prm{};
  Params P{};
  uint32_t V = 0, NF = 0, F = 0, E = 0;
  std::vector<uint32_t> h_vdof, h_free, h_tris, h_edges;
  std::vector<float> h_im;
  std::vector<float> h_pin_x;  // positions of the pinned vertices (xyz per pinned vertex, in vertex order)
  std::vector<uint8_t> h_static;
  bool has_collider = false;
  float char_len = 1.0f;
  double time = 0.0;
  // device state
  DevBuf<float> x, v, x0, xt, dx, dxf, im, rhs, egrad, side, pn, part;
  DevBuf<uint32_t> vt_cell;  // [0] some vertex is inside a collider, [1] the broad phase's grid cell (FP32 bits)
  DevBuf<uint32_t> vdof, sdof, free_list, tris, edges, ex_ptr, ex, err, flag;
  DevBuf<uint8_t> static_v;
  DevBuf<unsigned long long> tmin;
  // trees and their primitive id lists
  DevBuf<uint32_t> tri_dyn_ids, tri_stat_ids, tri_coll_ids, edge_dyn_ids, edge_stat_ids;
  DevBuf<uint32_t> dxr;       // the movable primitives' dx range per CCD call (k_dx_range keys)
  uint64_t tree_epoch = 1, built_epoch = 0;  // the dynamic trees are rebuilt once per step (epoch)
  bool dyn_has_fixed = false;  // some movable primitive has a fixed vertex (dx = 0 in the range)
  uint32_t n_tri_dyn = 0, n_tri_stat = 0, n_tri_coll = 0, n_edge_dyn = 0, n_edge_stat = 0;
  Lbvh tree_tri_dyn, tree_tri_stat, tree_tri_coll, tree_edge_dyn, tree_edge_stat;
  // elements: active (system), static (all pinned) and mixed (pinned and free: contact blocks)
  ElemSet act, stat, mix;
  DevBuf<uint32_t> xdof;  // vertex -> row of the mixed-element table
  // kinematic collider (implicit.rs:1324-1398): the move given to its vertices since the last
  // step, read by the friction slip and cleared by the step; the collider triangles' ids
  DevBuf<float> kdx;
  std::vector<float> h_kdx;
  bool kdx_on = false;
  std::vector<uint32_t> h_coll_ids;
  // contact buffers
  uint32_t cap = 0;
  DevBuf<uint32_t> vt_count, vt_ord, ee_ord;
  DevBuf<uint64_t> vt_key, ee_key, inc_key;
  DevBuf<float> vt_d, vt_g, ee_d, ee_g;
  DevBuf<uint32_t> sidx, inc_val, inc, inc_ptr, cdof;
  DevBuf<uint32_t> inc_pos;  // incidence slot of each (stencil, participant)
  DevBuf<float> contrib;     // per-incidence contact blocks of the PCG's p (float4 each)
  DevBuf<float> sd, sg, cg, cw, ccoef, cn, cgrad, csums, ckappa;
  uint32_t ns = 0, nvt = 0, nee = 0;
  RadixSort sorter;
  std::unique_ptr<Pcg> pcg;
  std::unique_ptr<Pcg64> pcg64;  // the FP64 PCG (precision NGUO_PCG_FP64)
  int precision = NGUO_PCG_FP32_REFINED;
  std::vector<uint32_t> h_sdof;  // contact stencil data read back for the Schwarz domains
  std::vector<float> h_scoef, h_sg, h_sw;
  std::unique_ptr<Schwarz> sz;
  // host scratch (pinned)
  float *hf = nullptr;
  uint32_t *hu = nullptr;
  unsigned long long *hl = nullptr;
  nguo_step_metrics metrics{};
  DevBuf<double> eacc;  // step energy terms (nguo_step_energy)
  DevBuf<uint32_t> sig_v;  // the mesh's unique edges and their rest lengths (max_sigma)
  DevBuf<float> sig_rest;
  uint32_t n_sig = 0;
  DevBuf<float> trace;  // x after every Newton iteration of the last step (when trace_on)
  bool trace_on = false;
  uint32_t trace_count = 0;  // iterates held in trace
  DevBuf<uint32_t> defer_count, defer_pairs;  // CCD pairs handed to k_ccd_deferred in a pass
  // drape adjoint: start of the last step (its friction reference), adjoint vector and rhs
  uint32_t n_shell_faces = 0;
  DevBuf<float> xprev, lam, adj_b, dgds;
  std::vector<float> h_adj_b;
  DevBuf<uint32_t> adj_dof;  // exact adjoint: the Jacobian blocks the Newton Hessian leaves out
  DevBuf<double> adj_blk, adj_x64, adj_y64;
  bool adj_valid = false;  // the last step can be differentiated (cleared by set_state / set_rest_scale)
  std::vector<float> trace_toi;
  bool own_stream = false;
  PhaseClock clock;
  // batch membership: Newton solves go through the batch's gate (null = solved alone)
  SolveGate *gate = nullptr;
  hipEvent_t gate_ev = nullptr;  // this step's assembly, recorded before it joins a batched solve
  ~nguo_step() {
    if (gate_ev) (void)hipEventDestroy(gate_ev);
    if (own_stream && s) (void)hipStreamSynchronize(s);
    if (hf) (void)hipHostFree(hf);
    if (hu) (void)hipHostFree(hu);
    if (hl) (void)hipHostFree(hl);
    if (own_stream && s) (void)hipStreamDestroy(s);
  }
};

// Batched Newton solves of several step objects advancing on their own threads: each thread
// that reaches its linear solve deposits the system and waits; when every member thread has
// deposited (or left because its Newton loop ended), the last one to arrive solves all of them
// together (Pcg::solve_batch, at most MAX_BATCH per call) on the batch stream and releases
// the others. Stream order: the batch stream waits for each system's assembly event; each
// member stream waits for the batch's completion event.
struct SolveGate {
  std::mutex m;
  std::condition_variable cv;
  uint32_t members = 0, arrived = 0;
  uint64_t gen = 0;
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  std::vector<PcgBatchItem> items;
  std::vector<nguo_step *> owners;
  std::vector<PcgStats> results;           // by owner, valid after release
  std::vector<std::exception_ptr> errors;  // by owner
  ~SolveGate() {
    if (done) (void)hipEventDestroy(done);
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
  void join() {
    std::lock_guard<std::mutex> lk(m);
    ++members;
  }
  void leave() {
    std::unique_lock<std::mutex> lk(m);
    --members;
    if (arrived > 0 && arrived == members) run();
  }
  PcgStats solve(nguo_step *st, const SellMatrix &M, const ContactView &C, const float *b, float tol, uint32_t max_iter, float *x) {
    if (!st->gate_ev) NGUO_HIP(hipEventCreateWithFlags(&st->gate_ev, hipEventDisableTiming));
    NGUO_HIP(hipEventRecord(st->gate_ev, st->s));
    std::unique_lock<std::mutex> lk(m);
    const uint32_t me = (uint32_t)owners.size();
    items.push_back(PcgBatchItem{st->pcg.get(), &M, C, b, tol, max_iter, x, PcgStats{}});
    owners.push_back(st);
    ++arrived;
    const uint64_t my = gen;
    if (arrived == members) run();
    else cv.wait(lk, [&] { return gen != my; });
    if (errors[me]) std::rethrow_exception(errors[me]);
    return results[me];
  }
  // With the lock held and every member arrived: solve the deposited systems, release.
  void run() {
    const size_t n = items.size();
    results.assign(n, PcgStats{});
    errors.assign(n, nullptr);
    try {
      if (!s) NGUO_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      if (!done) NGUO_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
      bool batch = true;
      for (auto &it : items) batch = batch && it.pcg->batchable();
      if (batch) {
        for (auto *st : owners) NGUO_HIP(hipStreamWaitEvent(s, st->gate_ev, 0));
        for (size_t o = 0; o < n; o += MAX_BATCH) Pcg::solve_batch(s, items.data() + o, (uint32_t)std::min<size_t>(MAX_BATCH, n - o));
        NGUO_HIP(hipEventRecord(done, s));
        for (auto *st : owners) NGUO_HIP(hipStreamWaitEvent(st->s, done, 0));
        for (size_t i = 0; i < n; ++i) results[i] = items[i].stats;
      } else {
        for (size_t i = 0; i < n; ++i) {
          PcgBatchItem &it = items[i];
          results[i] = it.pcg->solve(*it.M, it.C, it.b, it.tol, it.max_iter, it.x_out);
        }
      }
    } catch (...) {
      for (auto &e : errors) e = std::current_exception();
    }
    items.clear();
    owners.clear();
    arrived = 0;
    ++gen;
    cv.notify_all();
  }
};

namespace {

std::vector<int> digit_shifts(uint64_t maxval_lo, uint64_t maxval_hi) {
  std::vector<int> sh;
  int lob = 0, hib = 0;
  while (lob < 32 && (maxval_lo >> lob) != 0) lob++;
  while (hib < 32 && (maxval_hi >> hib) != 0) hib++;
  for (int b = 0; b < lob; b += 8) sh.push_back(b);
  for (int b = 0; b < hib; b += 8) sh.push_back(32 + b);
  if (sh.empty()) sh.push_back(0);
  return sh;
}

// shell_model.rs:309-349 build_shells (rest frame, Dm^-1, area) on the host.
// shell_model.rs:309-349 the rest frame of one face: Dm^-1 (column-major) and area, in FP32
// for the step and in FP64 for the energy evaluation.
template <class T> bool shell_rest(const float *p0, const float *p1, const float *p2, T dm_out[4], T &area) {
  using V = v3<T>;
  V a{T(p0[0]), T(p0[1]), T(p0[2])}, b{T(p1[0]), T(p1[1]), T(p1[2])}, c{T(p2[0]), T(p2[1]), T(p2[2])};
  V e1 = b - a, e2 = c - a;
  V n = cross(e1, e2);
  area = T(0.5) * std::sqrt(dot(n, n));
  if (area <= T(1.0e-12)) return false;
  V ax = e1 * (T(1) / std::sqrt(dot(e1, e1)));
  V t = e2 - dot(e2, ax) * ax;
  V ay = t * (T(1) / std::sqrt(dot(t, t)));
  T m00 = dot(e1, ax), m10 = dot(e1, ay), m01 = dot(e2, ax), m11 = dot(e2, ay);
  T det = m00 * m11 - m10 * m01;
  if (std::fabs(det) < T(1e-20)) return false;
  T inv_det = T(1) / det;
  // inverse, column-major (xa = (m11, -m10) / det, ya = (-m01, m00) / det)
  dm_out[0] = m11 * inv_det;
  dm_out[1] = -m10 * inv_det;
  dm_out[2] = -m01 * inv_det;
  dm_out[3] = m00 * inv_det;
  return true;
}

// bending.rs:34-48 dihedral angle on the host (rest angles).
float dihedral_host(const float *p0, const float *p1, const float *p2, const float *p3) {
  f3 a{p0[0], p0[1], p0[2]}, b{p1[0], p1[1], p1[2]}, c{p2[0], p2[1], p2[2]}, d{p3[0], p3[1], p3[2]};
  f3 n1 = cross(b - a, c - a), n2 = cross(b - c, d - c);
  f3 e = normalize_or_zero(c - b);
  f3 n1n = normalize_or_zero(n1), n2n = normalize_or_zero(n2);
  float sn = dot(cross(n1n, n2n), e);
  float cs = clampv(dot(n1n, n2n), -1.0f, 1.0f);
  return std::atan2(sn, cs);
}

// Builds the pattern, slot->row map and the fixed-order contribution lists of one element
// set over `rows` (vertex -> row map, NONE = not in this system). Columns use the same map,
// or, with `cmap`, their own (the mixed-element table: global vertex columns); with `pinned`,
// only blocks with a pinned row or column vertex are kept.
void build_gather(ElemSet &S, const ElemHost &H, uint32_t nrows, const std::vector<uint32_t> &vmap, const std::vector<uint32_t> &row_vert,
                  bool with_grad, hipStream_t s, const std::vector<uint32_t> *cmap = nullptr, const std::vector<uint8_t> *pinned = nullptr) {
  // (row, column) of block (va, vb), or NONE
  auto at = [&](uint32_t va, uint32_t vb) -> std::pair<uint32_t, uint32_t> {
    const uint32_t r = vmap[va], c = cmap ? (*cmap)[vb] : vmap[vb];
    if (r == NONE || c == NONE) return {NONE, NONE};
    if (pinned && !(*pinned)[va] && !(*pinned)[vb]) return {NONE, NONE};
    return {r, c};
  };
  std::vector<std::vector<uint32_t>> cols(nrows);
  if (!cmap)
    for (uint32_t r = 0; r < nrows; ++r) cols[r].push_back(r);
  auto couple = [&](const uint32_t *v, int k) {
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) {
        auto rc = at(v[a], v[b]);
        if (rc.first != NONE) cols[rc.first].push_back(rc.second);
      }
  };
  for (uint32_t e = 0; e < H.nr(); ++e) couple(&H.rd_v[2 * e], 2);
  for (uint32_t e = 0; e < H.nrh(); ++e) couple(&H.rh_v[3 * e], 3);
  for (uint32_t e = 0; e < H.nsp(); ++e) couple(&H.sp_v[2 * e], 2);
  for (uint32_t e = 0; e < H.ns(); ++e) couple(&H.sh_v[3 * e], 3);
  for (uint32_t e = 0; e < H.nh(); ++e) couple(&H.hg_v[4 * e], 4);
  for (auto &c : cols) { std::sort(c.begin(), c.end()); c.erase(std::unique(c.begin(), c.end()), c.end()); }
  S.M.build_pattern(nrows, cols, s);
  // slot -> row
  std::vector<uint32_t> srow(S.M.nslots, NONE);
  for (uint32_t r = 0; r < nrows; ++r)
    for (uint32_t k = 0; k < S.M.h_row_len[r]; ++k) srow[S.M.slot(r, k)] = r;
  // contributions (slot, code) in push order
  std::vector<std::pair<uint32_t, uint32_t>> hlist;
  std::vector<std::pair<uint32_t, uint32_t>> glist;
  const uint32_t nrh = H.nrh();
  // a two-vertex element's blocks [[H, -H], [-H, H]] and gradient (+g, -g), out_rod entry e
  auto block = [&](uint32_t va, uint32_t vb, uint32_t c) {
    auto rc = at(va, vb);
    if (rc.first != NONE) hlist.push_back({S.M.find(rc.first, rc.second), c});
  };
  auto pair = [&](const uint32_t *v, uint32_t e) {
    block(v[0], v[0], code(C_ROD_POS, e));
    block(v[1], v[1], code(C_ROD_POS, e));
    block(v[0], v[1], code(C_ROD_NEG, e));
    block(v[1], v[0], code(C_ROD_NEG, e));
    uint32_t ra = vmap[v[0]], rb = vmap[v[1]];
    if (ra != NONE) glist.push_back({ra, code(C_ROD_POS, e)});
    if (rb != NONE) glist.push_back({rb, code(C_ROD_NEG, e)});
  };
  for (uint32_t e = 0; e < H.nr(); ++e) pair(&H.rd_v[2 * e], e);
  auto dense = [&](const uint32_t *v, uint32_t elem_base) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) block(v[a], v[b], code(C_DENSE, elem_base * 9 + 3 * a + b));
    for (int a = 0; a < 3; ++a)
      if (vmap[v[a]] != NONE) glist.push_back({vmap[v[a]], code(C_DENSE, elem_base * 3 + a)});
  };
  for (uint32_t e = 0; e < nrh; ++e) dense(&H.rh_v[3 * e], e);
  for (uint32_t e = 0; e < H.nsp(); ++e) pair(&H.sp_v[2 * e], H.nr() + e);
  for (uint32_t e = 0; e < H.ns(); ++e) dense(&H.sh_v[3 * e], nrh + e);
  for (uint32_t e = 0; e < H.nh(); ++e) {
    const uint32_t *v = &H.hg_v[4 * e];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) block(v[a], v[b], code(C_HINGE, e * 16 + 4 * a + b));
    for (int a = 0; a < 4; ++a)
      if (vmap[v[a]] != NONE) glist.push_back({vmap[v[a]], code(C_HINGE, e * 4 + a)});
  }
  auto to_csr = [](const std::vector<std::pair<uint32_t, uint32_t>> &lst, uint32_t nkeys, std::vector<uint32_t> &ptr, std::vector<uint32_t> &val) {
    ptr.assign(nkeys + 1, 0);
    for (auto &p : lst) ptr[p.first + 1]++;
    for (uint32_t i = 0; i < nkeys; ++i) ptr[i + 1] += ptr[i];
    val.assign(lst.size(), 0);
    std::vector<uint32_t> cur(ptr.begin(), ptr.end() - 1);
    for (auto &p : lst) val[cur[p.first]++] = p.second;  // stable: keeps push order
  };
  std::vector<uint32_t> hp, hv, gp, gv;
  to_csr(hlist, S.M.nslots, hp, hv);
  S.slot_row.upload(srow, s);
  S.hc_ptr.upload(hp, s);
  S.hc.upload(hv.empty() ? std::vector<uint32_t>{0} : hv, s);
  if (with_grad) {
    to_csr(glist, nrows, gp, gv);
    S.gc_ptr.upload(gp, s);
    S.gc.upload(gv.empty() ? std::vector<uint32_t>{0} : gv, s);
  }
  S.row_vert.upload(row_vert.empty() ? std::vector<uint32_t>{0} : row_vert, s);
}

void upload_elems(ElemSet &S, const ElemHost &H, hipStream_t s) {
  S.ns = H.ns(); S.nh = H.nh(); S.nr = H.nr(); S.nrh = H.nrh(); S.nsp = H.nsp();
  auto up = [&](auto &buf, const auto &vec) {
    using T = typename std::decay_t<decltype(vec)>::value_type;
    if (vec.empty()) buf.reserve(1);
    else buf.upload(vec.data(), vec.size(), s);
    (void)sizeof(T);
  };
  up(S.sh_v, H.sh_v); up(S.sh_dm, H.sh_dm); up(S.sh_area, H.sh_area);
  up(S.sh_dm64, H.sh_dm64); up(S.sh_area64, H.sh_area64);
  S.h_sh_face = H.sh_face;
  S.h_sh_scale.assign(H.ns(), 1.0f);
  up(S.sh_scale, S.h_sh_scale);
  up(S.sh_face, S.h_sh_face);
  up(S.hg_v, H.hg_v); up(S.hg_rest, H.hg_rest);
  up(S.rd_v, H.rd_v); up(S.rd_rest, H.rd_rest); up(S.rd_k, H.rd_k);
  up(S.rh_v, H.rh_v); up(S.rh_rest, H.rh_rest); up(S.rh_stiff, H.rh_stiff);
  up(S.sp_v, H.sp_v); up(S.sp_rest, H.sp_rest); up(S.sp_rest64, H.sp_rest64);
  S.out_dense.reserve(81 * ((size_t)S.nrh + S.ns) + 1);
  S.out_dg.reserve(9 * ((size_t)S.nrh + S.ns) + 1);
  S.out_rod.reserve(12 * ((size_t)S.nr + S.nsp) + 1);
  S.out_hinge.reserve(14 * (size_t)S.nh + 1);
}

void run_elements(ElemSet &S, const Params &P, const float *x, hipStream_t s) {
  ElemDev E = S.dev();
  if (S.nr + S.nsp) hipLaunchKernelGGL(k_rods, dim3(div_up(S.nr + S.nsp, BLOCK)), dim3(BLOCK), 0, s, E, P.project, P.ksp, x);
  if (S.nrh) hipLaunchKernelGGL(k_rod_hinges, dim3(div_up(S.nrh, BLOCK)), dim3(BLOCK), 0, s, E, x);
  if (S.ns) hipLaunchKernelGGL(k_shells, dim3(div_up(S.ns, BLOCK)), dim3(BLOCK), 0, s, E, P, x);
  if (S.nh) hipLaunchKernelGGL(k_hinges, dim3(div_up(S.nh, BLOCK)), dim3(BLOCK), 0, s, E, P, x);
  NGUO_HIP(hipGetLastError());
}

void gather(ElemSet &S, const Params &P, int active, const float *im, const float *x, const float *xs, hipStream_t s) {
  if (S.M.nslots == 0) return;
  hipLaunchKernelGGL(k_gather_hess, dim3(div_up(S.M.nslots, BLOCK)), dim3(BLOCK), 0, s, S.M.nslots, (const uint32_t *)S.slot_row.p,
                     (const uint32_t *)S.M.col.p, (const uint32_t *)S.hc_ptr.p, (const uint32_t *)S.hc.p, S.dev(), (const uint32_t *)S.row_vert.p,
                     im, active, P, x, xs, S.M.val.p);
  NGUO_HIP(hipGetLastError());
}

// The elastic blocks at x: the free system (inertia, elements, analytic colliders) and, when
// elements join pinned and free vertices, the mixed-element table's blocks on pinned rows and
// columns (read by the contact stiffness).
void assemble_elastic(nguo_step *st, const Params &P, hipStream_t s) {
  run_elements(st->act, P, st->x.p, s);
  gather(st->act, P, 1, st->im.p, st->x.p, st->x0.p, s);
  if (st->mix.M.nb) {
    run_elements(st->mix, P, st->x.p, s);
    gather(st->mix, P, 0, st->im.p, st->x.p, st->x0.p, s);
  }
}

Topo topo(nguo_step *st) {
  Topo T{};
  T.V = st->V;
  T.im = st->im.p;
  T.tris = st->tris.p;
  T.edges = st->edges.p;
  T.ex_ptr = st->ex_ptr.p;
  T.ex = st->ex.p;
  T.static_v = st->has_collider ? st->static_v.p : nullptr;
  T.side = st->has_collider ? st->side.p : nullptr;
  T.pn = st->pn.p;
  T.vdof = st->vdof.p;
  T.sdof = st->sdof.p;
  T.xdof = st->mix.M.nb ? st->xdof.p : nullptr;
  return T;
}

ContactView contact_view(nguo_step *st) {
  ContactView C{};
  C.ns = st->ns;
  C.dof = st->cdof.p;
  C.g = st->cg.p;
  C.w = st->cw.p;
  C.coef = st->ccoef.p;
  C.n = st->cn.p;
  C.sums = st->csums.p;
  C.inc_ptr = st->inc_ptr.p;
  C.inc = st->inc.p;
  C.pos = st->inc_pos.p;
  C.contrib = (float4 *)st->contrib.p;
  return C;
}

// The trees over the pinned primitives are built from the state's pinned positions, so entry
// points that evaluate at caller-given positions require x to keep them.
void record_pinned(nguo_step *st, const float *x) {
  st->h_pin_x.clear();
  for (uint32_t i = 0; i < st->V; ++i)
    if (!(st->h_im[i] > 0.0f)) st->h_pin_x.insert(st->h_pin_x.end(), x + 3 * (size_t)i, x + 3 * (size_t)i + 3);
}
void require_pinned(const nguo_step *st, const float *x, const char *who) {
  size_t k = 0;
  for (uint32_t i = 0; i < st->V; ++i) {
    if (st->h_im[i] > 0.0f) continue;
    for (int c = 0; c < 3; ++c, ++k)
      NGUO_REQUIRE(x[3 * (size_t)i + c] == st->h_pin_x[k], NGUO_E_INVALID,
                   std::string(who) + ": x moves pinned vertex " + std::to_string(i) + " (set the state first)");
  }
}

void ensure_contact_cap(nguo_step *st, uint32_t want) {
  if (want <= st->cap) return;
  uint32_t c = std::max<uint32_t>(want, 1u << 14);
  st->vt_key.reserve(c); st->vt_ord.reserve(c); st->vt_d.reserve(c); st->vt_g.reserve(12 * (size_t)c);
  st->ee_key.reserve(c); st->ee_ord.reserve(c); st->ee_d.reserve(c); st->ee_g.reserve(12 * (size_t)c);
  size_t c2 = 2 * (size_t)c;
  st->sidx.reserve(4 * c2); st->sd.reserve(c2); st->sg.reserve(12 * c2);
  st->cdof.reserve(4 * c2); st->cg.reserve(12 * c2); st->cw.reserve(4 * c2); st->ccoef.reserve(2 * c2); st->cn.reserve(3 * c2);
  st->cgrad.reserve(24 * c2); st->csums.reserve(4 * c2); st->ckappa.reserve(c2);
  st->inc_key.reserve(4 * c2); st->inc_val.reserve(4 * c2); st->inc.reserve(4 * c2);
  st->inc_pos.reserve(4 * c2); st->contrib.reserve(16 * c2);
  st->cap = c;
}

uint32_t read_u32(nguo_step *st, const uint32_t *dev) {
  NGUO_HIP(hipMemcpyAsync(st->hu, dev, 4, hipMemcpyDeviceToHost, st->s));
  NGUO_HIP(hipStreamSynchronize(st->s));
  return st->hu[0];
}

// Active stencils at the current x, sorted by (v, t) then (ei, ej), gathered into sidx/sd/sg.
void detect(nguo_step *st) {
  hipStream_t s = st->s;
  st->ns = st->nvt = st->nee = 0;
  if (!st->prm.self_collision) return;
  Topo T = topo(st);
  float dhat = st->prm.contact_dhat;
  // dynamic trees at x: built at a step's first detection, refit at its later Newton
  // iterations (the candidate sets, and so the sorted stencil lists, do not depend on the
  // topology; only the traversal's pruning does)
  if (st->tree_epoch != st->built_epoch) {
    st->tree_tri_dyn.build(0, st->n_tri_dyn, st->tri_dyn_ids.p, st->tris.p, st->x.p, s);
    st->tree_edge_dyn.build(1, st->n_edge_dyn, st->edge_dyn_ids.p, st->edges.p, st->x.p, s);
    st->built_epoch = st->tree_epoch;
  } else {
    st->tree_tri_dyn.refit(0, st->tris.p, st->x.p, s);
    st->tree_edge_dyn.refit(1, st->edges.p, st->x.p, s);
  }
  if (st->has_collider) {
    NGUO_HIP(hipMemsetAsync(st->vt_cell.p, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_sides, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, T, st->tree_tri_coll.view(), (const float *)st->x.p, dhat,
                       st->side.p, st->err.p, st->vt_cell.p);
    hipLaunchKernelGGL(k_vt_cell, dim3(1), dim3(CELL_BLOCK), 0, s, (const uint32_t *)st->tris.p, st->F, (const float *)st->x.p, dhat,
                       st->vt_cell.p);
  }
  for (int attempt = 0; attempt < 4; ++attempt) {
    NGUO_HIP(hipMemsetAsync(st->vt_count.p, 0, 2 * sizeof(uint32_t), s));
    StencilOut ovt{st->cap, st->vt_count.p, st->vt_key.p, st->vt_ord.p, st->vt_d.p, st->vt_g.p};
    StencilOut oee{st->cap, st->vt_count.p + 1, st->ee_key.p, st->ee_ord.p, st->ee_d.p, st->ee_g.p};
    hipLaunchKernelGGL(k_vt_stencils, dim3(div_up(st->V, BLOCK), 2 * TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_tri_dyn.view(), st->tree_tri_stat.view(),
                       (const float *)st->x.p, dhat, st->has_collider ? (const uint32_t *)st->vt_cell.p : nullptr, ovt, st->err.p);
    if (st->n_edge_dyn)
      hipLaunchKernelGGL(k_ee_stencils, dim3(div_up(st->n_edge_dyn, BLOCK), TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_edge_dyn.view(),
                         (const uint32_t *)st->edge_dyn_ids.p, st->n_edge_dyn, 1, (const float *)st->x.p, dhat, dhat, oee, st->err.p);
    // movable edges against the static edges' tree (built once: its positions never change)
    if (st->n_edge_stat && st->n_edge_dyn)
      hipLaunchKernelGGL(k_ee_stencils, dim3(div_up(st->n_edge_dyn, BLOCK), TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_edge_stat.view(),
                         (const uint32_t *)st->edge_dyn_ids.p, st->n_edge_dyn, 0, (const float *)st->x.p, dhat, dhat, oee, st->err.p);
    NGUO_HIP(hipGetLastError());
    NGUO_HIP(hipMemcpyAsync(st->hu, st->vt_count.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NGUO_HIP(hipMemcpyAsync(st->hu + 2, st->err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NGUO_HIP(hipStreamSynchronize(s));
    NGUO_REQUIRE(st->hu[2] == 0, NGUO_E_CAPACITY, "LBVH traversal stack overflow");
    uint32_t nvt = st->hu[0], nee = st->hu[1];
    if (nvt <= st->cap && nee <= st->cap) { st->nvt = nvt; st->nee = nee; break; }
    ensure_contact_cap(st, (uint32_t)(std::max(nvt, nee) * 1.5) + 1024);
    NGUO_REQUIRE(attempt < 3, NGUO_E_CAPACITY, "contact stencil list did not fit after regrowth");
  }
  st->ns = st->nvt + st->nee;
  if (st->ns == 0) return;
  st->sorter.sort_u64(st->nvt, st->vt_key.p, st->vt_ord.p, digit_shifts(st->F, st->V), s);
  st->sorter.sort_u64(st->nee, st->ee_key.p, st->ee_ord.p, digit_shifts(st->E, st->E), s);
  hipLaunchKernelGGL(k_stencil_gather, dim3(div_up(st->ns, BLOCK)), dim3(BLOCK), 0, s, st->nvt, st->nee, (const uint64_t *)st->vt_key.p,
                     (const uint32_t *)st->vt_ord.p, (const float *)st->vt_d.p, (const float *)st->vt_g.p, (const uint64_t *)st->ee_key.p,
                     (const uint32_t *)st->ee_ord.p, (const float *)st->ee_d.p, (const float *)st->ee_g.p, (const uint32_t *)st->tris.p,
                     (const uint32_t *)st->edges.p, st->sidx.p, st->sd.p, st->sg.p);
  NGUO_HIP(hipGetLastError());
}

// Detection + per-stencil evaluation + vertex incidence: the contact view of the system.
void contacts(nguo_step *st) {
  hipStream_t s = st->s;
  detect(st);
  if (st->ns == 0) return;
  Topo T = topo(st);
  float dhat = st->prm.contact_dhat;
  ContactOut CO{st->cdof.p, st->cg.p, st->cw.p, st->ccoef.p, st->cn.p, st->cgrad.p, st->ckappa.p};
  hipLaunchKernelGGL(k_stencil_eval, dim3(div_up(st->ns, BLOCK)), dim3(BLOCK), 0, s, st->ns, (const uint32_t *)st->sidx.p, (const float *)st->sd.p,
                     (const float *)st->sg.p, dhat, 0.5f * (dhat + dhat), st->nvt, T, st->act.M.view(), st->stat.M.view(), st->mix.M.view(), st->P,
                     (const float *)st->x.p, (const float *)st->x0.p, st->kdx_on ? (const float *)st->kdx.p : nullptr, CO);
  uint32_t m = 4 * st->ns;
  hipLaunchKernelGGL(k_inc_keys, dim3(div_up(m, BLOCK)), dim3(BLOCK), 0, s, st->ns, (const uint32_t *)st->cdof.p, st->NF, st->inc_key.p, st->inc_val.p);
  st->sorter.sort_u64(m, st->inc_key.p, st->inc_val.p, digit_shifts(m, st->NF + 1), s);
  hipLaunchKernelGGL(k_inc_ptr, dim3(div_up(std::max(m, st->NF + 1), BLOCK)), dim3(BLOCK), 0, s, st->NF, m, (const uint64_t *)st->inc_key.p,
                     st->inc_ptr.p, st->inc.p);
  hipLaunchKernelGGL(k_inc_pos, dim3(div_up(m, BLOCK)), dim3(BLOCK), 0, s, st->NF, m, (const uint32_t *)st->cdof.p,
                     (const uint32_t *)st->inc_ptr.p, (const uint32_t *)st->inc.p, st->inc_pos.p);
  NGUO_HIP(hipGetLastError());
}

float reduce_max(nguo_step *st, uint32_t nparts) {
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(BLOCK), 0, st->s, nparts, (const float *)st->part.p, st->part.p + st->part.n - 1);
  NGUO_HIP(hipMemcpyAsync(st->hf, st->part.p + st->part.n - 1, sizeof(float), hipMemcpyDeviceToHost, st->s));
  NGUO_HIP(hipStreamSynchronize(st->s));
  return st->hf[0];
}

// implicit.rs:1690-1947 ccd_max_step.
float ccd_max_step(nguo_step *st, float max_disp) {
  hipStream_t s = st->s;
  bool has_self = st->prm.self_collision != 0;
  if (st->P.ncoll == 0 && !has_self) return 1.0f;
  double t = 1.0;
  Topo T = topo(st);
  auto run_min = [&](auto launch) -> double {
    st->hl[0] = (unsigned long long)__builtin_bit_cast(unsigned long long, 1.0);
    NGUO_HIP(hipMemcpyAsync(st->tmin.p, st->hl, sizeof(unsigned long long), hipMemcpyHostToDevice, s));
    launch();
    NGUO_HIP(hipGetLastError());
    NGUO_HIP(hipMemcpyAsync(st->hl, st->tmin.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    NGUO_HIP(hipStreamSynchronize(s));
    return __builtin_bit_cast(double, st->hl[0]);
  };
  if (st->P.ncoll) {
    t = std::min(t, run_min([&] {
      hipLaunchKernelGGL(k_ccd_colliders, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, st->V, (const float *)st->im.p, (const float *)st->x.p,
                         (const float *)st->dx.p, st->P, st->tmin.p);
    }));
  }
  if (has_self && std::isfinite(max_disp)) {
    uint32_t iters = std::max(st->prm.ccd_max_iter, 1u);
    // the movable primitives' displacement range (dx is fixed across the passes)
    hipLaunchKernelGGL(k_dx_range_init, dim3(1), dim3(WAVE), 0, s, st->dxr.p, st->dyn_has_fixed ? 1 : 0);
    hipLaunchKernelGGL(k_dx_range, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, st->V, (const uint32_t *)st->vdof.p, (const float *)st->dx.p,
                       st->dxr.p);
    float fixed = st->prm.contact_dhat;
    float allowed = std::max(std::min(0.5f * st->char_len, max_disp), 1.0e-9f);
    for (int pass = 0; pass < 64; ++pass) {
      bool full = allowed >= max_disp;
      float search = fixed + 2.0f * allowed;
      const CcdDefer D{st->defer_count.p, st->defer_pairs.p, CCD_DEFER_CAP};
      double t_self = run_min([&] {
        NGUO_HIP(hipMemsetAsync(st->defer_count.p, 0, sizeof(uint32_t), s));
        hipLaunchKernelGGL(k_ccd_vt, dim3(div_up(st->V, BLOCK), 2 * TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_tri_dyn.view(), st->tree_tri_stat.view(),
                           (const float *)st->x.p, (const float *)st->dx.p, search, max_disp, iters, (const uint32_t *)st->dxr.p, D, st->tmin.p,
                           st->err.p);
        if (st->n_edge_dyn)
          hipLaunchKernelGGL(k_ccd_ee, dim3(div_up(st->n_edge_dyn, BLOCK), TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_edge_dyn.view(),
                             (const uint32_t *)st->edge_dyn_ids.p, st->n_edge_dyn, 1, (const float *)st->x.p, (const float *)st->dx.p, search, max_disp,
                             iters, (const uint32_t *)st->dxr.p, D, st->tmin.p, st->err.p);
        // movable edges against the static edges' tree: a movable query's padding covers
        // every pair the static side's padding would (its own motion is part of the cap), and
        // a pair beyond the static side's cap returns the horizon, so the minimum is the same
        if (st->n_edge_stat && st->n_edge_dyn)
          hipLaunchKernelGGL(k_ccd_ee, dim3(div_up(st->n_edge_dyn, BLOCK), TRAV_PARTS), dim3(BLOCK), 0, s, T, st->tree_edge_stat.view(),
                             (const uint32_t *)st->edge_dyn_ids.p, st->n_edge_dyn, 0, (const float *)st->x.p, (const float *)st->dx.p, search,
                             max_disp, iters, (const uint32_t *)st->dxr.p, D, st->tmin.p, st->err.p);
        if (iters > ACCD_INLINE)
          hipLaunchKernelGGL(k_ccd_deferred, dim3(div_up(CCD_DEFER_CAP, BLOCK)), dim3(BLOCK), 0, s, T, D, (const float *)st->x.p, (const float *)st->dx.p,
                             iters, st->tmin.p);
      });
      if (full || (float)t_self * max_disp <= allowed) { t = std::min(t, t_self); break; }
      // implicit.rs:1942 widens x4. Widening to exactly the motion t_self permits returns the
      // same t (the reference's own argument: a pass whose radius covers t * max_disp of motion
      // is the full-radius query for any t it returns): the next pass holds every pair of this
      // one, so its t_self' <= t_self and t_self' * max_disp <= allowed certifies it. At most two
      // passes, the second with the smallest radius that certifies.
      allowed = std::min((float)t_self * max_disp, max_disp);
    }
    NGUO_REQUIRE(read_u32(st, st->err.p) == 0, NGUO_E_CAPACITY, "LBVH traversal stack overflow");
  }
  return (float)std::min(std::max(t, 0.0), 1.0);
}

// Trees over the pinned primitives (triangles and edges with every vertex pinned) and the
// collider's triangles: their positions do not change while stepping, so they are built from
// the current positions at creation and when a state is set, not per Newton iteration.
// contact.rs:943-981 pseudonormals of the collider triangles at positions X (3V floats):
// vertices welded on exact bits (-0 folded), face normals summed with their corner angles.
std::vector<float> collider_pseudonormals(const nguo_step *st, const float *X) {
  const uint32_t V = st->V;
  std::vector<float> pn(3 * (size_t)V, 0.0f);
  std::vector<std::pair<std::array<uint32_t, 3>, uint32_t>> keys(V);
  auto bits = [](float c) { uint32_t u; float z = c == 0.0f ? 0.0f : c; std::memcpy(&u, &z, 4); return u; };
  for (uint32_t i = 0; i < V; ++i) keys[i] = {{bits(X[3 * i]), bits(X[3 * i + 1]), bits(X[3 * i + 2])}, i};
  std::sort(keys.begin(), keys.end());
  std::vector<uint32_t> of(V);
  uint32_t nw = 0;
  for (uint32_t i = 0; i < V; ++i) {
    if (i > 0 && keys[i].first != keys[i - 1].first) nw++;
    of[keys[i].second] = nw;
  }
  std::vector<f3> acc(nw + 1, f3{0, 0, 0});
  for (uint32_t f : st->h_coll_ids) {
    const uint32_t *t = &st->h_tris[3 * f];
    f3 p[3];
    for (int k = 0; k < 3; ++k) p[k] = f3{X[3 * t[k]], X[3 * t[k] + 1], X[3 * t[k] + 2]};
    f3 face = normalize_or_zero(cross(p[1] - p[0], p[2] - p[0]));
    if (face.x == 0.0f && face.y == 0.0f && face.z == 0.0f) continue;
    for (int k = 0; k < 3; ++k) {
      f3 u = normalize_or_zero(p[(k + 1) % 3] - p[k]);
      f3 w = normalize_or_zero(p[(k + 2) % 3] - p[k]);
      acc[of[t[k]]] = acc[of[t[k]]] + face * std::acos(clampv(dot(u, w), -1.0f, 1.0f));
    }
  }
  for (uint32_t i = 0; i < V; ++i) {
    f3 n = normalize_or_zero(acc[of[i]]);
    pn[3 * i] = n.x; pn[3 * i + 1] = n.y; pn[3 * i + 2] = n.z;
  }
  return pn;
}

void build_static_trees(nguo_step *st, hipStream_t s) {
  st->tree_tri_stat.build(0, st->n_tri_stat, st->tri_stat_ids.p, st->tris.p, st->x.p, s);
  st->tree_edge_stat.build(1, st->n_edge_stat, st->edge_stat_ids.p, st->edges.p, st->x.p, s);
  if (st->has_collider) st->tree_tri_coll.build(0, st->n_tri_coll, st->tri_coll_ids.p, st->tris.p, st->x.p, s);
}

// True when the step's Newton solves use the FP64 PCG (its Schwarz preconditioner is one-level).
bool uses_fp64(const nguo_step *st) { return st->precision == NGUO_PCG_FP64 && st->sz && st->sz->nd1 == 0; }

// The Newton system (act.M + contacts) x = b with the step's solver: the FP64 PCG, or the FP32
// PCG with FP64 refinement (through the batch's gate when the step is a batch member).
// The Schwarz domains of the system about to be solved: with contact stencils, built around
// their couplings (Schwarz::build_contact, from the stencils' rows and coefficients read back);
// without, the elastic-pattern domains of the step object.
void prepare_schwarz(nguo_step *st) {
  if (!st->sz) return;
  hipStream_t s = st->s;
  if (st->ns == 0) {
    st->sz->restore_elastic(st->act.M, s);
    return;
  }
  const size_t ns = st->ns;
  st->h_sdof.resize(4 * ns);
  st->h_scoef.resize(2 * ns);
  st->h_sg.resize(12 * ns);
  st->h_sw.resize(4 * ns);
  st->cdof.download(st->h_sdof.data(), 4 * ns, s);
  st->ccoef.download(st->h_scoef.data(), 2 * ns, s);
  st->cg.download(st->h_sg.data(), 12 * ns, s);
  st->cw.download(st->h_sw.data(), 4 * ns, s);
  NGUO_HIP(hipStreamSynchronize(s));
  st->sz->build_contact(st->act.M, st->ns, st->h_sdof.data(), st->h_scoef.data(), st->h_sg.data(), st->h_sw.data(), s);
}

PcgStats solve_system(nguo_step *st, const float *b, float tol, float *x) {
  prepare_schwarz(st);
  st->clock.mark(PH_DOMAINS, st->s);
  if (uses_fp64(st)) {
    if (!st->pcg64) st->pcg64.reset(new Pcg64(st->s));
    return st->pcg64->solve(st->act.M, contact_view(st), *st->sz, b, tol, st->prm.cg_max_iter, x);
  }
  if (st->gate) return st->gate->solve(st, st->act.M, contact_view(st), b, tol, st->prm.cg_max_iter, x);
  return st->pcg->solve(st->act.M, contact_view(st), b, tol, st->prm.cg_max_iter, x);
}

void assemble_and_solve(nguo_step *st, nguo_stats &stats) {
  hipStream_t s = st->s;
  assemble_elastic(st, st->P, s);
  if (st->NF)
    hipLaunchKernelGGL(k_gather_grad, dim3(div_up(st->NF, BLOCK)), dim3(BLOCK), 0, s, st->NF, (const uint32_t *)st->free_list.p,
                       (const uint32_t *)st->act.gc_ptr.p, (const uint32_t *)st->act.gc.p, st->act.dev(), (const float *)st->im.p, st->P,
                       (const float *)st->x.p, (const float *)st->xt.p, (const float *)st->x0.p, st->egrad.p);
  st->clock.mark(PH_ASSEMBLE, s);
  contacts(st);
  if (st->NF)
    hipLaunchKernelGGL(k_rhs, dim3(div_up(st->NF, BLOCK)), dim3(BLOCK), 0, s, st->NF, (const float *)st->egrad.p, (const uint32_t *)st->inc_ptr.p,
                       (const uint32_t *)st->inc.p, (const float *)st->cgrad.p, st->ns > 0 ? 1 : 0, st->rhs.p);
  NGUO_HIP(hipGetLastError());
  st->clock.mark(PH_CONTACTS, s);
  if (st->NF == 0) { stats = nguo_stats{}; return; }
  PcgStats ps = solve_system(st, st->rhs.p, st->prm.cg_tol, st->dxf.p);
  st->clock.mark(PH_SOLVE, s);
  stats.iters = ps.iters;
  stats.sweeps = ps.sweeps;
  stats.residual = ps.residual;
}

// A batch member for the duration of one step's Newton loop (leaves on every exit path).
struct GateMembership {
  SolveGate *g;
  explicit GateMembership(SolveGate *gate) : g(gate) { if (g) g->join(); }
  void leave() { if (g) { g->leave(); g = nullptr; } }
  ~GateMembership() { leave(); }
};

void advance_one(nguo_step *st) {
  hipStream_t s = st->s;
  ++st->tree_epoch;
  const uint64_t allocs0 = g_dev_allocs, frees0 = g_dev_frees;
  GateMembership member(uses_fp64(st) ? nullptr : st->gate);  // FP64 solves run on the instance's own stream
  const uint32_t V = st->V;
  const unsigned GV = div_up(V, BLOCK);
  float h = st->prm.dt;
  st->P.inv_h2 = 1.0f / (h * h);
  st->adj_valid = false;
  st->clock.start(s);
  NGUO_HIP(hipMemcpyAsync(st->xprev.p, st->x0.p, 3 * (size_t)V * sizeof(float), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_predict, dim3(GV), dim3(BLOCK), 0, s, V, (const float *)st->im.p, (const float *)st->x0.p, (const float *)st->v.p, h,
                     f3{st->prm.gravity[0], st->prm.gravity[1], st->prm.gravity[2]}, st->xt.p, st->x.p);
  uint32_t cap_it = std::max(st->prm.newton_iters, 1u), taken = 0;
  st->trace_toi.clear();
  st->trace_count = 0;
  bool final_step = false;
  double max_res = 0;
  uint32_t max_it = 0, sum_it = 0;
  float last_t = 1.0f;
  for (uint32_t it = 0; it < cap_it; ++it) {
    taken++;
    nguo_stats ls{};
    assemble_and_solve(st, ls);
    max_res = std::max(max_res, ls.residual);
    max_it = std::max(max_it, ls.iters);
    sum_it += ls.iters;
    hipLaunchKernelGGL(k_scatter_dx, dim3(GV), dim3(BLOCK), 0, s, V, (const uint32_t *)st->vdof.p, (const float *)st->dxf.p, st->dx.p, st->part.p);
    float max_norm = reduce_max(st, GV);
    if (max_norm > 0.0f) {
      float rescale = std::min(st->prm.max_dx / max_norm, 1.0f);
      if (rescale < 1.0f) {
        hipLaunchKernelGGL(k_scale, dim3(div_up(3 * (size_t)V, BLOCK)), dim3(BLOCK), 0, s, 3 * V, st->dx.p, rescale);
        max_norm = max_norm * rescale;
      }
    }
    st->clock.mark(PH_UPDATE, s);
    float t = ccd_max_step(st, max_norm);
    st->clock.mark(PH_CCD, s);
    last_t = t;
    hipLaunchKernelGGL(k_apply, dim3(GV), dim3(BLOCK), 0, s, V, st->x.p, (const float *)st->dx.p, t, st->part.p);
    if (st->trace_on) {  // per-iterate positions for parity checks (off in the production step)
      st->trace.reserve(3 * (size_t)V * cap_it);
      NGUO_HIP(hipMemcpyAsync(st->trace.p + 3 * (size_t)V * it, st->x.p, 3 * (size_t)V * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    st->trace_toi.resize(it + 1);
    st->trace_toi[it] = t;
    if (st->trace_on) st->trace_count = it + 1;
    float max_step = reduce_max(st, GV);
    st->clock.mark(PH_UPDATE, s);
    if (max_step < 1.0e-7f) break;
    if (final_step) break;
    if (st->prm.target_toi > 0.0f && t >= st->prm.target_toi && taken > st->prm.min_newton_steps) final_step = true;
  }
  member.leave();
  NGUO_HIP(hipMemsetAsync(st->flag.p, 0, 2 * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_finite, dim3(GV), dim3(BLOCK), 0, s, V, (const float *)st->x.p, st->flag.p);
  hipLaunchKernelGGL(k_commit, dim3(GV), dim3(BLOCK), 0, s, V, (const float *)st->im.p, st->x.p, st->x0.p, st->v.p, 1.0f / h, st->prm.floor_y,
                     (const uint32_t *)st->flag.p, st->part.p);
  float vmax = reduce_max(st, GV);
  uint32_t bad = read_u32(st, st->flag.p);
  st->time += (double)h;
  if (st->kdx_on) {  // implicit.rs:2065-2070: the collider's move described this step only
    NGUO_HIP(hipMemsetAsync(st->kdx.p, 0, 3 * (size_t)V * sizeof(float), s));
    std::fill(st->h_kdx.begin(), st->h_kdx.end(), 0.0f);
    st->kdx_on = false;
  }
  // penetration_free: finite, analytic colliders respected, no edge pierces a triangle
  uint32_t hit = 0, ncontacts = 0;
  if (st->P.ncoll && !bad) {
    NGUO_HIP(hipMemsetAsync(st->flag.p + 2, 0, 2 * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_collider_metrics, dim3(GV), dim3(BLOCK), 0, s, V, (const float *)st->im.p, (const float *)st->x.p, st->P, st->flag.p + 2);
    NGUO_HIP(hipMemcpyAsync(st->hu + 4, st->flag.p + 2, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    NGUO_HIP(hipStreamSynchronize(s));
    ncontacts = st->hu[4];
    hit |= st->hu[5];
  }
  if (st->prm.self_collision && !bad) {
    if (st->built_epoch == st->tree_epoch) st->tree_tri_dyn.refit(0, st->tris.p, st->x.p, s);  // built this step
    else st->tree_tri_dyn.build(0, st->n_tri_dyn, st->tri_dyn_ids.p, st->tris.p, st->x.p, s);
    Topo T = topo(st);
    hipLaunchKernelGGL(k_pierce, dim3(div_up(st->E, BLOCK)), dim3(BLOCK), 0, s, T, st->tree_tri_dyn.view(), st->tree_tri_stat.view(), st->E,
                       (const float *)st->x.p, st->flag.p + 1, st->err.p);
    NGUO_HIP(hipGetLastError());
    hit |= read_u32(st, st->flag.p + 1);
  }
  nguo_step_metrics &m = st->metrics;
  m.time = st->time;
  m.max_reresid = max_res;
  m.max_cg_iters = max_it;
  m.sum_cg_iters = sum_it;
  m.newton_iters = taken;
  m.num_stencils = st->ns;
  m.max_velocity = vmax;
  m.all_finite = bad ? 0 : 1;
  m.penetration_free = (!bad && !hit) ? 1 : 0;
  m.last_toi = last_t;
  m.num_contacts = ncontacts;
  m.max_sigma = 0.0f;
  if (st->n_sig) {
    NGUO_HIP(hipMemsetAsync(st->flag.p + 3, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_max_sigma, dim3(div_up(st->n_sig, BLOCK)), dim3(BLOCK), 0, s, st->n_sig, (const uint32_t *)st->sig_v.p,
                       (const float *)st->sig_rest.p, (const float *)st->x.p, st->flag.p + 3);
    uint32_t bits = read_u32(st, st->flag.p + 3);
    std::memcpy(&m.max_sigma, &bits, sizeof bits);
  }
  m.device_alloc = (uint32_t)(g_dev_allocs - allocs0);
  m.device_free = (uint32_t)(g_dev_frees - frees0);
  st->adj_valid = !bad;
  st->clock.mark(PH_POST, s);
  st->clock.report(st->time, st->V);
}

}  // namespace

extern "C" {

int nguo_step_create(nguo_ctx *ctx, const nguo_scene_desc *sc, nguo_step **out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && sc && out, NGUO_E_INVALID, "nguo_step_create: null argument");
    NGUO_REQUIRE(sc->n_verts > 0 && sc->positions && sc->inv_mass, NGUO_E_INVALID, "nguo_step_create: empty scene");
    NGUO_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<nguo_step> st(new nguo_step());
    if (const char *e = std::getenv("NGUO_PHASE_TIMES")) st->clock.on = e[0] == '1';
    st->ctx = ctx;
    // each step object owns a stream (default priority), so several can advance concurrently
    NGUO_HIP(hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking));
    st->own_stream = true;
    hipStream_t s = st->s;
    const uint32_t V = sc->n_verts;
    st->V = V;
    st->prm = sc->params;
    NGUO_REQUIRE(sc->n_colliders <= 4, NGUO_E_UNSUPPORTED, "at most 4 analytic colliders");
    NGUO_REQUIRE(sc->params.stretch_k >= 0.0f && std::isfinite(sc->params.stretch_k), NGUO_E_INVALID,
                 "edge-spring stiffness must be finite and >= 0");
    // validation (implicit.rs:826-838 style)
    for (uint32_t i = 0; i < 3 * sc->n_tris; ++i) NGUO_REQUIRE(sc->tris[i] < V, NGUO_E_INVALID, "triangle index out of range");
    for (uint32_t i = 0; i < 3 * sc->n_shell_faces; ++i) NGUO_REQUIRE(sc->shell_faces[i] < V, NGUO_E_INVALID, "shell face index out of range");
    for (uint32_t r = 0; r < sc->n_rods; ++r) {
      uint32_t a = sc->rod_verts[2 * r], b = sc->rod_verts[2 * r + 1];
      NGUO_REQUIRE(a < V && b < V && a != b, NGUO_E_INVALID, "rod out of range or degenerate");
      NGUO_REQUIRE(sc->rod_rest[r] > 0.0f && sc->rod_k[r] > 0.0f, NGUO_E_INVALID, "rod rest length and stiffness must be > 0");
    }
    for (uint32_t h = 0; h < sc->n_rod_hinges; ++h) {
      const uint32_t *v = sc->rod_hinge_verts + 3 * h;
      NGUO_REQUIRE(v[0] < V && v[1] < V && v[2] < V, NGUO_E_INVALID, "rod hinge out of range");
      NGUO_REQUIRE(v[0] != v[1] && v[1] != v[2] && v[0] != v[2], NGUO_E_INVALID, "rod hinge repeats a vertex");
    }
    st->h_im.assign(sc->inv_mass, sc->inv_mass + V);
    st->h_vdof.assign(V, NONE);
    for (uint32_t i = 0; i < V; ++i)
      if (st->h_im[i] > 0.0f) { st->h_vdof[i] = (uint32_t)st->h_free.size(); st->h_free.push_back(i); }
    st->NF = (uint32_t)st->h_free.size();
    record_pinned(st.get(), sc->positions);
    st->F = sc->n_tris;
    st->h_tris.assign(sc->tris, sc->tris + 3 * (size_t)sc->n_tris);
    // Params
    Params &P = st->P;
    P.shell_model = st->prm.shell_model;
    P.shell_mu = st->prm.shell_mu;
    P.shell_lambda = st->prm.shell_lambda;
    P.eig_eps = st->prm.eiganalysis_eps;
    P.kb = st->prm.bend_stiffness;
    P.ksp = st->prm.stretch_k;
    P.dhat = st->prm.contact_dhat;
    P.kappa = st->prm.contact_kappa;
    P.mu = st->prm.friction;
    P.feps = st->prm.friction_eps;
    P.barrier_kind = st->prm.barrier_kind;
    P.project = 1;
    P.ncoll = sc->n_colliders;
    P.inv_h2 = 1.0f / (st->prm.dt * st->prm.dt);
    for (uint32_t c = 0; c < sc->n_colliders; ++c) {
      P.coll[c] = sc->colliders[c];
      if (P.coll[c].kind == NGUO_COLLIDER_HALFSPACE) {  // scene.rs:24-27 normalize_or_zero
        f3 nn{P.coll[c].b[0], P.coll[c].b[1], P.coll[c].b[2]};
        nn = normalize_or_zero(nn);
        P.coll[c].b[0] = nn.x; P.coll[c].b[1] = nn.y; P.coll[c].b[2] = nn.z;
      }
    }
    const float *X = sc->positions;
    auto pinned = [&](uint32_t v) { return !(st->h_im[v] > 0.0f); };
    // elements, split into active (any free vertex: their free-free blocks and free rows build
    // the system) and static (all pinned: blocks gathered once); an element joining pinned and
    // free vertices is also listed in the mixed table, whose blocks on a pinned row or column
    // serve the contact stiffness (the reference's unfiltered blocks, implicit.rs:3175-3228)
    ElemHost ha, hs, hm;
    auto classify = [&](const uint32_t *v, int k) -> int {
      int np = 0;
      for (int a = 0; a < k; ++a) np += pinned(v[a]) ? 1 : 0;
      if (np == 0) return 0;
      if (np == k) return 1;
      return 2;
    };
    auto homes = [&](int c) -> std::vector<ElemHost *> {
      if (c == 0) return {&ha};
      if (c == 1) return {&hs};
      return {&ha, &hm};
    };
    if (!sc->shell_faces || !(st->prm.shell_mu > 0.0f || st->prm.shell_lambda > 0.0f)) {
      // shells contribute nothing (implicit.rs:2608-2611)
    } else {
      for (uint32_t f = 0; f < sc->n_shell_faces; ++f) {
        const uint32_t *t = sc->shell_faces + 3 * f;
        float dm[4], area;
        double dm64[4], area64;
        if (!shell_rest(X + 3 * t[0], X + 3 * t[1], X + 3 * t[2], dm, area)) continue;
        (void)shell_rest(X + 3 * t[0], X + 3 * t[1], X + 3 * t[2], dm64, area64);
        for (ElemHost *H : homes(classify(t, 3))) {
          H->sh_v.insert(H->sh_v.end(), t, t + 3);
          H->sh_dm.insert(H->sh_dm.end(), dm, dm + 4);
          H->sh_area.push_back(area);
          H->sh_dm64.insert(H->sh_dm64.end(), dm64, dm64 + 4);
          H->sh_area64.push_back(area64);
          H->sh_face.push_back(f);
        }
      }
    }
    if (st->prm.bend_stiffness > 0.0f) {
      // bending.rs:128-158 build_hinges over all triangles
      std::vector<std::pair<uint64_t, uint32_t>> ek;  // (edge key, opposite) in face order
      ek.reserve(3 * (size_t)st->F);
      for (uint32_t f = 0; f < st->F; ++f) {
        const uint32_t *t = &st->h_tris[3 * f];
        uint32_t tri[3][3] = {{t[1], t[2], t[0]}, {t[0], t[2], t[1]}, {t[0], t[1], t[2]}};
        for (auto &e : tri) {
          uint32_t a = std::min(e[0], e[1]), b = std::max(e[0], e[1]);
          ek.push_back({((uint64_t)a << 32) | b, e[2]});
        }
      }
      std::stable_sort(ek.begin(), ek.end(), [](const auto &p, const auto &q) { return p.first < q.first; });
      for (size_t i = 0; i < ek.size();) {
        size_t j = i;
        while (j < ek.size() && ek[j].first == ek[i].first) ++j;
        if (j - i == 2) {
          uint32_t v[4] = {ek[i].second, (uint32_t)(ek[i].first >> 32), (uint32_t)ek[i].first, ek[i + 1].second};
          const float rest = dihedral_host(X + 3 * v[0], X + 3 * v[1], X + 3 * v[2], X + 3 * v[3]);
          for (ElemHost *H : homes(classify(v, 4))) {
            H->hg_v.insert(H->hg_v.end(), v, v + 4);
            H->hg_rest.push_back(rest);
          }
        }
        i = j;
      }
    }
    for (uint32_t r = 0; r < sc->n_rods; ++r) {
      const uint32_t *v = sc->rod_verts + 2 * r;
      for (ElemHost *H : homes(classify(v, 2))) {
        H->rd_v.insert(H->rd_v.end(), v, v + 2);
        H->rd_rest.push_back(sc->rod_rest[r]);
        H->rd_k.push_back(sc->rod_k[r]);
      }
    }
    for (uint32_t h = 0; h < sc->n_rod_hinges; ++h) {
      if (!(sc->rod_hinge_stiff[h] > 0.0f)) continue;
      const uint32_t *v = sc->rod_hinge_verts + 3 * h;
      for (ElemHost *H : homes(classify(v, 3))) {
        H->rh_v.insert(H->rh_v.end(), v, v + 3);
        H->rh_rest.push_back(sc->rod_hinge_rest[h]);
        H->rh_stiff.push_back(sc->rod_hinge_stiff[h]);
      }
    }
    // state.rs:39-65 SimState::from_mesh: one stretch constraint per unique triangle edge, in
    // the sorted order of geometry.rs:37-45, rest length |x_a - x_b| of the scene positions.
    // They carry the max_sigma channel always and the edge springs when stretch_k > 0.
    std::vector<std::pair<uint32_t, uint32_t>> ue;
    ue.reserve(3 * (size_t)st->F);
    for (uint32_t f = 0; f < st->F; ++f) {
      const uint32_t *t = &st->h_tris[3 * f];
      for (int k = 0; k < 3; ++k) ue.push_back({std::min(t[k], t[(k + 1) % 3]), std::max(t[k], t[(k + 1) % 3])});
    }
    std::sort(ue.begin(), ue.end());
    ue.erase(std::unique(ue.begin(), ue.end()), ue.end());
    std::vector<uint32_t> sig_v;
    std::vector<float> sig_rest;
    sig_v.reserve(2 * ue.size());
    sig_rest.reserve(ue.size());
    for (auto &e : ue) {
      const uint32_t v[2] = {e.first, e.second};
      const float dx = X[3 * v[0]] - X[3 * v[1]], dy = X[3 * v[0] + 1] - X[3 * v[1] + 1], dz = X[3 * v[0] + 2] - X[3 * v[1] + 2];
      const float rest = std::sqrt(dx * dx + dy * dy + dz * dz);
      sig_v.insert(sig_v.end(), v, v + 2);
      sig_rest.push_back(rest);
      if (!(st->prm.stretch_k > 0.0f)) continue;
      const double ex = (double)X[3 * v[0]] - X[3 * v[1]], ey = (double)X[3 * v[0] + 1] - X[3 * v[1] + 1],
                   ez = (double)X[3 * v[0] + 2] - X[3 * v[1] + 2];
      for (ElemHost *H : homes(classify(v, 2))) {
        H->sp_v.insert(H->sp_v.end(), v, v + 2);
        H->sp_rest.push_back(rest);
        H->sp_rest64.push_back(std::sqrt(ex * ex + ey * ey + ez * ez));
      }
    }
    st->n_sig = (uint32_t)ue.size();
    if (st->n_sig) {
      st->sig_v.upload(sig_v.data(), sig_v.size(), s);
      st->sig_rest.upload(sig_rest.data(), sig_rest.size(), s);
    }
    // static rows: pinned vertices touched by static elements
    std::vector<uint32_t> sdof(V, NONE), srow_vert;
    auto touch = [&](const std::vector<uint32_t> &vv) {
      for (uint32_t v : vv)
        if (sdof[v] == NONE) { sdof[v] = (uint32_t)srow_vert.size(); srow_vert.push_back(v); }
    };
    touch(hs.rd_v); touch(hs.rh_v); touch(hs.sp_v); touch(hs.sh_v); touch(hs.hg_v);
    // device uploads
    st->x.upload(X, 3 * (size_t)V, s);
    st->x0.upload(X, 3 * (size_t)V, s);
    if (sc->velocities) st->v.upload(sc->velocities, 3 * (size_t)V, s);
    else st->v.zero(3 * (size_t)V, s);
    st->xt.reserve(3 * (size_t)V);
    st->dx.reserve(3 * (size_t)V);
    st->xprev.reserve(3 * (size_t)V);
    st->n_shell_faces = sc->shell_faces ? sc->n_shell_faces : 0;
    st->im.upload(st->h_im, s);
    st->vdof.upload(st->h_vdof, s);
    st->sdof.upload(sdof, s);
    st->free_list.upload(st->h_free.empty() ? std::vector<uint32_t>{0} : st->h_free, s);
    st->dxf.reserve(3 * (size_t)std::max(st->NF, 1u));
    st->rhs.reserve(3 * (size_t)std::max(st->NF, 1u));
    st->egrad.reserve(3 * (size_t)std::max(st->NF, 1u));
    st->part.reserve(div_up(V, BLOCK) + 2);
    st->err.zero(1, s);
    st->flag.zero(4, s);
    st->tmin.reserve(1);
    st->defer_count.zero(1, s);
    st->defer_pairs.reserve(2 * (size_t)CCD_DEFER_CAP);
    st->vt_count.zero(2, s);
    st->inc_ptr.zero((size_t)st->NF + 2, s);
    upload_elems(st->act, ha, s);
    upload_elems(st->stat, hs, s);
    upload_elems(st->mix, hm, s);
    build_gather(st->act, ha, st->NF, st->h_vdof, st->h_free, true, s);
    build_gather(st->stat, hs, (uint32_t)srow_vert.size(), sdof, srow_vert, false, s);
    {
      // the mixed-element table: a row per vertex of a mixed element, global vertex columns,
      // blocks with a pinned row or column vertex only
      std::vector<uint32_t> xdof(V, NONE), xrow_vert, ident(V);
      std::vector<uint8_t> pin(V);
      for (uint32_t v = 0; v < V; ++v) { ident[v] = v; pin[v] = pinned(v) ? 1 : 0; }
      for (const auto *vv : {&hm.rd_v, &hm.rh_v, &hm.sp_v, &hm.sh_v, &hm.hg_v})
        for (uint32_t v : *vv)
          if (xdof[v] == NONE) { xdof[v] = (uint32_t)xrow_vert.size(); xrow_vert.push_back(v); }
      build_gather(st->mix, hm, (uint32_t)xrow_vert.size(), xdof, xrow_vert, false, s, &ident, &pin);
      st->xdof.upload(xdof, s);
    }
    // static element blocks once (the body never moves)
    if (st->stat.M.nb) {
      run_elements(st->stat, P, st->x.p, s);
      gather(st->stat, P, 0, st->im.p, st->x.p, st->x0.p, s);
    }
    // collision topology: unique edges (+ rods), 1-ring exclusion, trees
    std::vector<std::pair<uint32_t, uint32_t>> es;
    for (uint32_t f = 0; f < st->F; ++f) {
      const uint32_t *t = &st->h_tris[3 * f];
      for (int k = 0; k < 3; ++k) {
        uint32_t i = t[k], j = t[(k + 1) % 3];
        es.push_back({std::min(i, j), std::max(i, j)});
      }
    }
    std::sort(es.begin(), es.end());
    es.erase(std::unique(es.begin(), es.end()), es.end());
    std::vector<std::pair<uint32_t, uint32_t>> seen = es;
    for (auto &e : es) { st->h_edges.push_back(e.first); st->h_edges.push_back(e.second); }
    for (uint32_t r = 0; r < sc->n_rods; ++r) {
      uint32_t a = sc->rod_verts[2 * r], b = sc->rod_verts[2 * r + 1];
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = std::lower_bound(seen.begin(), seen.end(), key);
      if (it != seen.end() && *it == key) continue;
      seen.insert(it, key);
      st->h_edges.push_back(a);
      st->h_edges.push_back(b);
    }
    st->E = (uint32_t)(st->h_edges.size() / 2);
    {
      double sum = 0;
      float fsum = 0.0f;
      for (uint32_t e = 0; e < st->E; ++e) {
        const float *a = X + 3 * st->h_edges[2 * e], *b = X + 3 * st->h_edges[2 * e + 1];
        f3 d{a[0] - b[0], a[1] - b[1], a[2] - b[2]};
        fsum += std::sqrt(dot(d, d));
      }
      (void)sum;
      st->char_len = st->E ? std::max(fsum / (float)st->E, 1e-6f) : 1.0f;
    }
    std::vector<std::vector<uint32_t>> exl(V);
    for (uint32_t i = 0; i < V; ++i) exl[i].push_back(i);
    for (uint32_t f = 0; f < st->F; ++f) {
      const uint32_t *t = &st->h_tris[3 * f];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          if (a != b) exl[t[a]].push_back(t[b]);
    }
    std::vector<uint32_t> exp(V + 1, 0), exv;
    for (uint32_t i = 0; i < V; ++i) {
      std::sort(exl[i].begin(), exl[i].end());
      exl[i].erase(std::unique(exl[i].begin(), exl[i].end()), exl[i].end());
      exp[i + 1] = exp[i] + (uint32_t)exl[i].size();
      exv.insert(exv.end(), exl[i].begin(), exl[i].end());
    }
    st->tris.upload(st->h_tris.empty() ? std::vector<uint32_t>{0, 0, 0} : st->h_tris, s);
    st->edges.upload(st->h_edges.empty() ? std::vector<uint32_t>{0, 0} : st->h_edges, s);
    st->ex_ptr.upload(exp, s);
    st->ex.upload(exv, s);
    // static mesh collider (implicit.rs:1258-1322): winding check, pseudonormals
    std::vector<uint32_t> coll_ids;
    if (sc->static_collider) {
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (uint32_t f = 0; f < st->F; ++f) {
        const uint32_t *t = &st->h_tris[3 * f];
        if (!(sc->static_collider[t[0]] && sc->static_collider[t[1]] && sc->static_collider[t[2]])) continue;
        coll_ids.push_back(f);
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 3; ++c) { lo[c] = std::fmin(lo[c], X[3 * t[k] + c]); hi[c] = std::fmax(hi[c], X[3 * t[k] + c]); }
      }
      if (!coll_ids.empty()) {
        f3 mid{0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])};
        double box = (double)((hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]));
        double vol = 0;
        for (uint32_t f : coll_ids) {
          const uint32_t *t = &st->h_tris[3 * f];
          f3 a = f3{X[3 * t[0]], X[3 * t[0] + 1], X[3 * t[0] + 2]} - mid;
          f3 b = f3{X[3 * t[1]], X[3 * t[1] + 1], X[3 * t[1] + 2]} - mid;
          f3 c = f3{X[3 * t[2]], X[3 * t[2] + 1], X[3 * t[2] + 2]} - mid;
          vol += (double)dot(cross(a, b), c) / 6.0;
        }
        if (vol < -0.05 * box) coll_ids.clear();  // wound inward: unsigned contact only
      }
    }
    st->has_collider = !coll_ids.empty();
    st->h_coll_ids = coll_ids;
    std::vector<float> pn(3 * (size_t)V, 0.0f);
    if (st->has_collider) {
      st->h_static.assign(sc->static_collider, sc->static_collider + V);
      pn = collider_pseudonormals(st.get(), X);
      st->static_v.upload(st->h_static, s);
    } else {
      st->static_v.zero(V, s);
    }
    st->pn.upload(pn, s);
    st->side.reserve(V);
    st->vt_cell.reserve(2);
    {
      std::vector<float> ones(V, 1.0f);
      st->side.upload(ones, s);
    }
    // tree membership
    std::vector<uint32_t> tdyn, tstat, edyn, estat;
    for (uint32_t f = 0; f < st->F; ++f) {
      const uint32_t *t = &st->h_tris[3 * f];
      (pinned(t[0]) && pinned(t[1]) && pinned(t[2]) ? tstat : tdyn).push_back(f);
    }
    for (uint32_t e = 0; e < st->E; ++e) (pinned(st->h_edges[2 * e]) && pinned(st->h_edges[2 * e + 1]) ? estat : edyn).push_back(e);
    auto upl = [&](DevBuf<uint32_t> &b, const std::vector<uint32_t> &v, uint32_t &n) {
      n = (uint32_t)v.size();
      b.upload(v.empty() ? std::vector<uint32_t>{0} : v, s);
    };
    for (uint32_t f : tdyn)
      for (int k = 0; k < 3; ++k) st->dyn_has_fixed = st->dyn_has_fixed || pinned(st->h_tris[3 * f + k]);
    for (uint32_t e : edyn)
      for (int k = 0; k < 2; ++k) st->dyn_has_fixed = st->dyn_has_fixed || pinned(st->h_edges[2 * e + k]);
    st->dxr.reserve(6);
    upl(st->tri_dyn_ids, tdyn, st->n_tri_dyn);
    upl(st->tri_stat_ids, tstat, st->n_tri_stat);
    upl(st->tri_coll_ids, coll_ids, st->n_tri_coll);
    upl(st->edge_dyn_ids, edyn, st->n_edge_dyn);
    upl(st->edge_stat_ids, estat, st->n_edge_stat);
    build_static_trees(st.get(), s);
    ensure_contact_cap(st.get(), std::max<uint32_t>(4 * V, 1u << 16));
    st->pcg.reset(new Pcg(s));
    NGUO_REQUIRE(st->prm.precond == NGUO_PRECOND_BLOCK || st->prm.precond == NGUO_PRECOND_SCHWARZ, NGUO_E_INVALID, "unknown preconditioner");
    if (st->prm.precond == NGUO_PRECOND_SCHWARZ && st->NF) {
      NGUO_REQUIRE(st->prm.schwarz_domain >= 1 && st->prm.schwarz_domain <= SCHWARZ_MAX_DOMAIN, NGUO_E_UNSUPPORTED,
                   "Schwarz domains hold 1..32 vertices");
      st->sz.reset(new Schwarz());
      NGUO_REQUIRE(st->prm.schwarz_weight >= 0.0f && std::isfinite(st->prm.schwarz_weight), NGUO_E_INVALID,
                   "Schwarz coarse weight must be finite and >= 0");
      // contact-built domain sets vary in count: room for twice a full packing of the free
      // vertices, reserved before build() uploads anything (reserving reallocates)
      const uint32_t nd_cap = 2 * div_up(st->NF, st->prm.schwarz_domain) + 64;
      st->sz->reserve(nd_cap);
      st->sz->build(st->act.M, st->prm.schwarz_domain, s, st->prm.schwarz_weight);
      st->pcg->set_schwarz(st->sz.get());
    }
    // capacities from the mesh, so that the growing contact of a drape does not reallocate:
    // the step's sorts (incidence keys: 4 per stencil) and the solver's per-stencil and
    // per-domain buffers
    st->sorter.reserve_u64(4 * st->cap);
    if (st->NF) st->pcg->reserve(st->NF, 2 * st->cap, st->sz ? 2 * div_up(st->NF, st->prm.schwarz_domain) + 64 : 0);
    NGUO_HIP(hipHostMalloc((void **)&st->hf, 16 * sizeof(float), hipHostMallocDefault));
    NGUO_HIP(hipHostMalloc((void **)&st->hu, 16 * sizeof(uint32_t), hipHostMallocDefault));
    NGUO_HIP(hipHostMalloc((void **)&st->hl, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    NGUO_HIP(hipStreamSynchronize(s));
    *out = st.release();
    return NGUO_OK;
  });
}

void nguo_step_destroy(nguo_step *st) { delete st; }

int nguo_step_advance(nguo_step *st, uint32_t nsteps) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st, NGUO_E_INVALID, "nguo_step_advance: null handle");
    for (uint32_t k = 0; k < nsteps; ++k) advance_one(st);
    NGUO_HIP(hipStreamSynchronize(st->s));
    return NGUO_OK;
  });
}

int nguo_step_get_positions(nguo_step *st, float *out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && out, NGUO_E_INVALID, "nguo_step_get_positions: null argument");
    st->x0.download(out, 3 * (size_t)st->V, st->s);
    NGUO_HIP(hipStreamSynchronize(st->s));
    return NGUO_OK;
  });
}

int nguo_step_get_velocities(nguo_step *st, float *out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && out, NGUO_E_INVALID, "nguo_step_get_velocities: null argument");
    st->v.download(out, 3 * (size_t)st->V, st->s);
    NGUO_HIP(hipStreamSynchronize(st->s));
    return NGUO_OK;
  });
}

int nguo_step_set_state(nguo_step *st, const float *positions, const float *velocities) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && positions, NGUO_E_INVALID, "nguo_step_set_state: null argument");
    st->x0.upload(positions, 3 * (size_t)st->V, st->s);
    st->x.upload(positions, 3 * (size_t)st->V, st->s);
    if (velocities) st->v.upload(velocities, 3 * (size_t)st->V, st->s);
    build_static_trees(st, st->s);  // the state names the pinned vertices' positions too
    record_pinned(st, positions);
    st->adj_valid = false;
    NGUO_HIP(hipStreamSynchronize(st->s));
    return NGUO_OK;
  });
}

int nguo_step_set_kinematic_collider_positions(nguo_step *st, const float *x) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && x, NGUO_E_INVALID, "nguo_step_set_kinematic_collider_positions: null argument");
    NGUO_REQUIRE(st->has_collider, NGUO_E_INVALID, "nguo_step_set_kinematic_collider_positions: the scene has no static mesh collider");
    const uint32_t V = st->V;
    for (uint32_t i = 0; i < V; ++i) {
      if (!st->h_static[i]) continue;
      NGUO_REQUIRE(std::isfinite(x[3 * i]) && std::isfinite(x[3 * i + 1]) && std::isfinite(x[3 * i + 2]), NGUO_E_INVALID,
                   "nguo_step_set_kinematic_collider_positions: non-finite position on collider vertex " + std::to_string(i));
      NGUO_REQUIRE(!(st->h_im[i] > 0.0f), NGUO_E_UNSUPPORTED,
                   "nguo_step_set_kinematic_collider_positions: collider vertex " + std::to_string(i) + " is not pinned");
    }
    hipStream_t s = st->s;
    std::vector<float> pos(3 * (size_t)V), vel(3 * (size_t)V);
    st->x.download(pos.data(), 3 * (size_t)V, s);
    st->v.download(vel.data(), 3 * (size_t)V, s);
    NGUO_HIP(hipStreamSynchronize(s));
    if (st->h_kdx.empty()) st->h_kdx.assign(3 * (size_t)V, 0.0f);
    st->kdx.reserve(3 * (size_t)V);
    const float inv_h = 1.0f / st->prm.dt;
    for (uint32_t i = 0; i < V; ++i) {
      if (!st->h_static[i]) continue;
      for (int c = 0; c < 3; ++c) {
        const size_t k = 3 * (size_t)i + c;
        const float d = x[k] - pos[k];
        st->h_kdx[k] += d;
        pos[k] = x[k];
        vel[k] = d * inv_h;
      }
    }
    st->x.upload(pos, s);
    st->x0.upload(pos, s);
    st->v.upload(vel, s);
    st->kdx.upload(st->h_kdx, s);
    st->kdx_on = true;
    record_pinned(st, pos.data());
    st->pn.upload(collider_pseudonormals(st, pos.data()), s);
    build_static_trees(st, s);
    if (st->stat.M.nb) {  // the body's own element blocks at its new pose
      run_elements(st->stat, st->P, st->x.p, s);
      gather(st->stat, st->P, 0, st->im.p, st->x.p, st->x0.p, s);
    }
    st->adj_valid = false;
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_step_set_precision(nguo_step *st, int precision) {
  if (!st) return fail(NGUO_E_INVALID, "nguo_step_set_precision: null handle");
  if (precision != NGUO_PCG_FP32_REFINED && precision != NGUO_PCG_FP64) return fail(NGUO_E_INVALID, "nguo_step_set_precision: unknown precision");
  if (precision == NGUO_PCG_FP64 && !(st->sz && st->sz->nd1 == 0))
    return fail(NGUO_E_UNSUPPORTED, "nguo_step_set_precision: the FP64 PCG needs the one-level Schwarz preconditioner");
  st->precision = precision;
  return NGUO_OK;
}

int nguo_step_set_trace(nguo_step *st, int enable) {
  if (!st) return fail(NGUO_E_INVALID, "nguo_step_set_trace: null handle");
  st->trace_on = enable != 0;
  return NGUO_OK;
}

int nguo_step_get_newton_trace(nguo_step *st, uint32_t k, float *out, uint32_t *count, float *toi) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st, NGUO_E_INVALID, "nguo_step_get_newton_trace: null handle");
    uint32_t n = (uint32_t)st->trace_toi.size();
    if (count) *count = n;
    if (toi && k < n) *toi = st->trace_toi[k];
    if (!out) return NGUO_OK;
    NGUO_REQUIRE(k < n, NGUO_E_INVALID, "nguo_step_get_newton_trace: iteration index out of range");
    NGUO_REQUIRE(k < st->trace_count, NGUO_E_INVALID, "nguo_step_get_newton_trace: positions were not traced (nguo_step_set_trace)");
    NGUO_HIP(hipMemcpyAsync(out, st->trace.p + 3 * (size_t)st->V * k, 3 * (size_t)st->V * sizeof(float), hipMemcpyDeviceToHost, st->s));
    NGUO_HIP(hipStreamSynchronize(st->s));
    if (toi) *toi = st->trace_toi[k];
    return NGUO_OK;
  });
}

int nguo_step_stencils(nguo_step *st, const float *x, uint32_t cap, uint32_t *idx, float *d, float *g, float *kappa, uint32_t *count) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && x && count, NGUO_E_INVALID, "nguo_step_stencils: null argument");
    require_pinned(st, x, "nguo_step_stencils");
    hipStream_t s = st->s;
    st->x.upload(x, 3 * (size_t)st->V, s);
    assemble_elastic(st, st->P, s);
    contacts(st);
    *count = st->ns;
    uint32_t m = std::min(cap, st->ns);
    if (m) {
      if (idx) st->sidx.download(idx, 4 * (size_t)m, s);
      if (d) st->sd.download(d, m, s);
      if (g) st->sg.download(g, 12 * (size_t)m, s);
      if (kappa) st->ckappa.download(kappa, m, s);
    }
    NGUO_HIP(hipMemcpyAsync(st->x.p, st->x0.p, 3 * (size_t)st->V * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_step_collider_sides(nguo_step *st, const float *x, float *side) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && x && side, NGUO_E_INVALID, "nguo_step_collider_sides: null argument");
    require_pinned(st, x, "nguo_step_collider_sides");
    hipStream_t s = st->s;
    if (!st->has_collider) {
      std::fill(side, side + st->V, 1.0f);
      return NGUO_OK;
    }
    st->x.upload(x, 3 * (size_t)st->V, s);
    hipLaunchKernelGGL(k_sides, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, topo(st), st->tree_tri_coll.view(), (const float *)st->x.p,
                       st->prm.contact_dhat, st->side.p, st->err.p, nullptr);
    NGUO_HIP(hipGetLastError());
    st->side.download(side, st->V, s);
    NGUO_HIP(hipMemcpyAsync(st->x.p, st->x0.p, 3 * (size_t)st->V * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_step_set_rest_scale(nguo_step *st, const float *face_scale, uint32_t nfaces) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && face_scale, NGUO_E_INVALID, "nguo_step_set_rest_scale: null argument");
    NGUO_REQUIRE(nfaces == st->n_shell_faces, NGUO_E_INVALID, "nguo_step_set_rest_scale: nfaces must equal the scene's shell face count");
    for (uint32_t f = 0; f < nfaces; ++f)
      NGUO_REQUIRE(std::isfinite(face_scale[f]) && face_scale[f] > 0.0f, NGUO_E_INVALID, "nguo_step_set_rest_scale: scales must be finite and > 0");
    hipStream_t s = st->s;
    NGUO_HIP(hipStreamSynchronize(s));  // the host copies below may still feed an earlier upload
    for (ElemSet *S : {&st->act, &st->stat, &st->mix}) {
      if (!S->ns) continue;
      for (uint32_t e = 0; e < S->ns; ++e) S->h_sh_scale[e] = face_scale[S->h_sh_face[e]];
      S->sh_scale.upload(S->h_sh_scale, s);
    }
    // the static element blocks are evaluated once; refresh them for the new rest shapes
    if (st->stat.M.nb) {
      run_elements(st->stat, st->P, st->x.p, s);
      gather(st->stat, st->P, 0, st->im.p, st->x.p, st->x0.p, s);
    }
    st->adj_valid = false;
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

// The exact adjoint solve (H0 + H2) lam = b, with H0 the matrix and contact operator
// assembled at x* (the elastic blocks unprojected, or, with elastic_diffs, projected as the
// Newton Hessian has them) and H2 the Jacobian blocks H0 leaves out (k_adj_*; with
// elastic_diffs also the elastic blocks' unprojected-minus-projected differences). lam is the fixed point of G(lam) = H0^-1 (b - H2 lam), each G one
// solve of the step's PCG to `tol`; Anderson mixing over the last EXACT_ADJ_DEPTH updates
// (the affine fixed point's GMRES) converges where the fixed-point map's spectral radius
// (that of H0^-1 H2) alone would be slow. Stops when the relative update |G(lam) - lam| /
// |G(lam)| falls to tol, after EXACT_ADJ_MAX solves, or after EXACT_ADJ_GIVE_UP solves while
// still above 1e-2. b = adj_b on entry (FP32, free dofs),
// the solution in lam; returns the last solve's stats with iters summed over all solves.
}  // extern "C"
namespace {
constexpr uint32_t EXACT_ADJ_DEPTH = 6, EXACT_ADJ_MAX = 60, EXACT_ADJ_GIVE_UP = 15;
constexpr float EXACT_ADJ_ACCEPT = 1e-3f;  // relative update above which the result is refused
template <class Solve>
PcgStats solve_exact_adjoint(nguo_step *st, float tol, bool elastic_diffs, Solve solve_h0, uint32_t *outer, double *outer_res) {
  hipStream_t s = st->s;
  const size_t n = 3 * (size_t)st->NF;
  const ElemDev E = st->act.dev();
  uint32_t n_sph = 0;
  for (uint32_t k = 0; k < st->P.ncoll; ++k) n_sph += st->P.coll[k].kind != NGUO_COLLIDER_HALFSPACE;
  const uint32_t ne_s = elastic_diffs ? E.ns : 0, ne_r = elastic_diffs ? E.nr + E.nsp : 0;
  const uint32_t o_h = st->ns, o_r = o_h + E.nh, o_s = o_r + E.nrh, o_e = o_s + ne_s, o_v = o_e + ne_r;
  const uint32_t nb = o_v + (n_sph ? st->V : 0);
  st->adj_dof.reserve(4 * (size_t)nb);
  st->adj_blk.reserve(144 * (size_t)nb);
  st->adj_x64.reserve(n);
  st->adj_y64.reserve(n);
  auto at = [&](uint32_t off) { return AdjBlocks{st->adj_dof.p + 4 * (size_t)off, st->adj_blk.p + 144 * (size_t)off}; };
  const float dhat = st->P.dhat;
  if (st->ns)
    hipLaunchKernelGGL(k_adj_contacts, dim3(div_up(st->ns, BLOCK)), dim3(BLOCK), 0, s, st->ns, st->nvt, (const uint32_t *)st->sidx.p,
                       (const float *)st->sd.p, (const float *)st->ckappa.p, dhat, 0.5f * (dhat + dhat), st->P, (const uint32_t *)st->vdof.p,
                       (const float *)st->x.p, at(0));
  if (E.nh)
    hipLaunchKernelGGL(k_adj_hinges, dim3(div_up(E.nh, BLOCK)), dim3(BLOCK), 0, s, E, st->P, (const uint32_t *)st->vdof.p, (const float *)st->x.p,
                       at(o_h));
  if (E.nrh)
    hipLaunchKernelGGL(k_adj_strands, dim3(div_up(E.nrh, BLOCK)), dim3(BLOCK), 0, s, E, (const uint32_t *)st->vdof.p, (const float *)st->x.p,
                       at(o_r));
  if (ne_s)
    hipLaunchKernelGGL(k_adj_shells, dim3(div_up(ne_s, BLOCK)), dim3(BLOCK), 0, s, E, st->P, (const uint32_t *)st->vdof.p, (const float *)st->x.p,
                       at(o_s));
  if (ne_r)
    hipLaunchKernelGGL(k_adj_rods, dim3(div_up(ne_r, BLOCK)), dim3(BLOCK), 0, s, E, st->P.ksp, (const uint32_t *)st->vdof.p,
                       (const float *)st->x.p, at(o_e));
  if (n_sph)
    hipLaunchKernelGGL(k_adj_spheres, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, st->V, (const float *)st->im.p,
                       (const uint32_t *)st->vdof.p, (const float *)st->x.p, st->P, at(o_v));
  NGUO_HIP(hipGetLastError());
  std::vector<double> b(n), lam(n, 0.0), g(n), f(n), y(n), g_prev, f_prev;
  for (size_t i = 0; i < n; ++i) b[i] = (double)st->h_adj_b[i];
  std::vector<float> rf(n), lf(n);
  std::vector<std::vector<double>> dG, dF;
  PcgStats ps{};
  uint32_t total = 0;
  auto nrm = [](const std::vector<double> &v) {
    double a = 0.0;
    for (double e : v) a += e * e;
    return std::sqrt(a);
  };
  *outer = 0;
  *outer_res = INFINITY;
  bool first = true;
  while (*outer < EXACT_ADJ_MAX) {
    // G(lam) = H0^-1 (b - H2 lam)
    if (first) {
      for (size_t i = 0; i < n; ++i) rf[i] = (float)b[i];
    } else {
      st->adj_x64.upload(lam.data(), n, s);
      st->adj_y64.zero(n, s);
      hipLaunchKernelGGL(k_adj_apply, dim3(div_up(nb, BLOCK)), dim3(BLOCK), 0, s, nb, at(0), (const double *)st->adj_x64.p, st->adj_y64.p);
      NGUO_HIP(hipGetLastError());
      st->adj_y64.download(y.data(), n, s);
      NGUO_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < n; ++i) rf[i] = (float)(b[i] - y[i]);
    }
    st->adj_b.upload(rf.data(), n, s);
    ps = solve_h0();
    total += ps.iters;
    ++*outer;
    if (ps.breakdown) break;
    st->lam.download(lf.data(), n, s);
    NGUO_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < n; ++i) {
      g[i] = (double)lf[i];
      f[i] = g[i] - lam[i];
    }
    if (nb == 0) {  // no blocks: H2 = 0 and the one solve is the answer
      *outer_res = 0.0;
      break;
    }
    const double gn = nrm(g);
    *outer_res = gn > 0.0 ? nrm(f) / gn : 0.0;
    if (*outer_res <= (double)tol) break;
    // a mixing that has not reached 1e-2 in EXACT_ADJ_GIVE_UP solves is not converging (the
    // exact Jacobian far from H0, typically indefinite): stop rather than spend the budget
    if (*outer >= EXACT_ADJ_GIVE_UP && *outer_res > 1e-2) break;
    if (!first) {
      std::vector<double> df(n), dg(n);
      for (size_t i = 0; i < n; ++i) { df[i] = f[i] - f_prev[i]; dg[i] = g[i] - g_prev[i]; }
      dF.push_back(std::move(df));
      dG.push_back(std::move(dg));
      if (dF.size() > EXACT_ADJ_DEPTH) { dF.erase(dF.begin()); dG.erase(dG.begin()); }
    }
    f_prev = f;
    g_prev = g;
    first = false;
    // lam = g - dG gamma, gamma = argmin |f - dF gamma| (normal equations, m <= depth)
    const size_t m = dF.size();
    std::vector<double> gam(m, 0.0);
    if (m) {
      std::vector<double> A(m * m), r(m);
      double tr = 0.0;
      for (size_t i = 0; i < m; ++i) {
        for (size_t j = 0; j < m; ++j) {
          double a = 0.0;
          for (size_t k = 0; k < n; ++k) a += dF[i][k] * dF[j][k];
          A[i * m + j] = a;
        }
        double a = 0.0;
        for (size_t k = 0; k < n; ++k) a += dF[i][k] * f[k];
        r[i] = a;
        tr += A[i * m + i];
      }
      for (size_t i = 0; i < m; ++i) A[i * m + i] += 1e-14 * tr;
      // Gaussian elimination with partial pivoting on the small system
      for (size_t c = 0; c < m; ++c) {
        size_t pv = c;
        for (size_t i = c + 1; i < m; ++i)
          if (std::fabs(A[i * m + c]) > std::fabs(A[pv * m + c])) pv = i;
        if (pv != c) {
          for (size_t j = 0; j < m; ++j) std::swap(A[c * m + j], A[pv * m + j]);
          std::swap(r[c], r[pv]);
        }
        if (A[c * m + c] == 0.0) continue;
        for (size_t i = c + 1; i < m; ++i) {
          const double q = A[i * m + c] / A[c * m + c];
          for (size_t j = c; j < m; ++j) A[i * m + j] -= q * A[c * m + j];
          r[i] -= q * r[c];
        }
      }
      for (size_t c = m; c-- > 0;) {
        double a = r[c];
        for (size_t j = c + 1; j < m; ++j) a -= A[c * m + j] * gam[j];
        gam[c] = A[c * m + c] != 0.0 ? a / A[c * m + c] : 0.0;
      }
    }
    for (size_t i = 0; i < n; ++i) {
      double v = g[i];
      for (size_t k = 0; k < m; ++k) v -= dG[k][i] * gam[k];
      lam[i] = v;
    }
  }
  // the last G(lam) is in st->lam (its solve's output)
  ps.iters = total;
  return ps;
}
}  // namespace
extern "C" {

int nguo_step_adjoint(nguo_step *st, const float *dLdx, float tol, int hessian, float *dLds, nguo_stats *stats) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && dLdx && dLds, NGUO_E_INVALID, "nguo_step_adjoint: null argument");
    NGUO_REQUIRE(st->adj_valid, NGUO_E_INVALID,
                 "nguo_step_adjoint: no step to differentiate (advance first; set_state and set_rest_scale clear the last step)");
    NGUO_REQUIRE(tol > 0.0f, NGUO_E_INVALID, "nguo_step_adjoint: tol must be > 0");
    NGUO_REQUIRE(hessian == NGUO_ADJ_NEWTON || hessian == NGUO_ADJ_EXACT_ELASTIC || hessian == NGUO_ADJ_EXACT, NGUO_E_INVALID,
                 "nguo_step_adjoint: unknown hessian mode");
    hipStream_t s = st->s;
    const uint32_t nfc = st->n_shell_faces;
    if (nfc) std::memset(dLds, 0, nfc * sizeof(float));
    if (st->NF == 0 || nfc == 0) {
      if (stats) *stats = nguo_stats{};
      return NGUO_OK;
    }
    // H at x* (= x = x0 after the commit) with the last step's inertial target (xt) and
    // friction reference (its start, xprev)
    NGUO_HIP(hipMemcpyAsync(st->x0.p, st->xprev.p, 3 * (size_t)st->V * sizeof(float), hipMemcpyDeviceToDevice, s));
    assemble_elastic(st, st->P, s);
    contacts(st);  // contact stiffnesses from the step's (projected) elastic blocks
    if (hessian != NGUO_ADJ_NEWTON) {
      Params PA = st->P;
      PA.project = 0;
      assemble_elastic(st, PA, s);
    }
    // rhs: dL/dx on the free dofs
    std::vector<float> &hb = st->h_adj_b;
    hb.resize(3 * (size_t)st->NF);
    for (uint32_t k = 0; k < st->NF; ++k)
      for (int c = 0; c < 3; ++c) hb[3 * (size_t)k + c] = dLdx[3 * (size_t)st->h_free[k] + c];
    st->adj_b.upload(hb, s);
    st->lam.reserve(3 * (size_t)st->NF);
    PcgStats ps;
    prepare_schwarz(st);
    auto solve_h0 = [&]() {
      if (uses_fp64(st)) {
        if (!st->pcg64) st->pcg64.reset(new Pcg64(s));
        return st->pcg64->solve(st->act.M, contact_view(st), *st->sz, st->adj_b.p, tol, st->prm.cg_max_iter, st->lam.p);
      }
      return st->pcg->solve(st->act.M, contact_view(st), st->adj_b.p, tol, st->prm.cg_max_iter, st->lam.p);
    };
    uint32_t outer = 0;
    double outer_res = 0.0;
    if (hessian == NGUO_ADJ_EXACT) {
      // H0 with the unprojected elastic blocks is the closer splitting (fewer solves); where it
      // is indefinite (its first solve breaks down) H0 becomes the Newton Hessian and the
      // elastic blocks' differences join H2
      ps = solve_exact_adjoint(st, tol, false, solve_h0, &outer, &outer_res);
      if (ps.breakdown && outer == 1) {
        const uint32_t first_iters = ps.iters;
        assemble_elastic(st, st->P, s);
        prepare_schwarz(st);
        ps = solve_exact_adjoint(st, tol, true, solve_h0, &outer, &outer_res);
        ps.iters += first_iters;
      }
    } else {
      ps = solve_h0();
    }
    NGUO_HIP(hipMemcpyAsync(st->x0.p, st->x.p, 3 * (size_t)st->V * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_REQUIRE(!ps.breakdown, NGUO_E_UNSUPPORTED,
                 "nguo_step_adjoint: the adjoint Hessian is not positive definite at x* (PCG breakdown); use the Newton Hessian");
    NGUO_REQUIRE(hessian != NGUO_ADJ_EXACT || outer_res <= (double)EXACT_ADJ_ACCEPT, NGUO_E_UNSUPPORTED,
                 "nguo_step_adjoint: the exact adjoint's outer iteration did not converge (relative update " + std::to_string(outer_res) +
                     " after " + std::to_string(outer) + " solves); the exact Jacobian is far from the Newton Hessian at x*");
    st->dgds.zero(nfc, s);
    if (st->act.ns)
      hipLaunchKernelGGL(k_shell_dgds, dim3(div_up(st->act.ns, BLOCK)), dim3(BLOCK), 0, s, st->act.dev(), st->P, (const float *)st->x.p,
                         (const uint32_t *)st->vdof.p, (const float *)st->lam.p, (const uint32_t *)st->act.sh_face.p, st->dgds.p);
    NGUO_HIP(hipGetLastError());
    st->dgds.download(dLds, nfc, s);
    NGUO_HIP(hipStreamSynchronize(s));
    if (stats) {
      stats->iters = ps.iters;
      stats->sweeps = hessian == NGUO_ADJ_EXACT ? outer : ps.sweeps;
      stats->residual = hessian == NGUO_ADJ_EXACT ? outer_res : ps.residual;
    }
    return NGUO_OK;
  });
}

int nguo_step_energy(nguo_step *st, const float *x, double *terms, double *total) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && x, NGUO_E_INVALID, "nguo_step_energy: null argument");
    NGUO_REQUIRE(st->adj_valid, NGUO_E_INVALID,
                 "nguo_step_energy: no step to evaluate (advance first; set_state and set_rest_scale clear the last step)");
    require_pinned(st, x, "nguo_step_energy");
    hipStream_t s = st->s;
    const size_t n3 = 3 * (size_t)st->V;
    // the committed end state waits in dx; the step's friction reference (its start) goes to
    // x0 for the contact evaluation, x takes the positions asked for
    st->dx.reserve(n3);
    NGUO_HIP(hipMemcpyAsync(st->dx.p, st->x0.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipMemcpyAsync(st->x0.p, st->xprev.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    st->x.upload(x, n3, s);
    assemble_elastic(st, st->P, s);
    contacts(st);
    st->eacc.zero(EN_TERMS, s);
    const double h = (double)st->prm.dt;
    hipLaunchKernelGGL(k_energy_verts, dim3(div_up(st->V, BLOCK)), dim3(BLOCK), 0, s, st->V, (const float *)st->im.p, (const float *)st->x.p,
                       (const float *)st->xt.p, (const float *)st->x0.p, 1.0 / (h * h), st->P, st->eacc.p);
    for (ElemSet *es : {&st->act, &st->stat}) {
      const uint32_t ne = es->ns + es->nh + es->nr + es->nrh + es->nsp;
      if (ne)
        hipLaunchKernelGGL(k_energy_elems, dim3(div_up(ne, BLOCK)), dim3(BLOCK), 0, s, es->dev(), (const double *)es->sh_dm64.p,
                           (const double *)es->sh_area64.p, st->P, (const float *)st->x.p, st->eacc.p);
    }
    if (st->ns) {
      const float dhat = st->prm.contact_dhat;
      hipLaunchKernelGGL(k_energy_contacts, dim3(div_up(st->ns, BLOCK)), dim3(BLOCK), 0, s, st->ns, (const uint32_t *)st->sidx.p,
                         (const float *)st->sd.p, dhat, 0.5f * (dhat + dhat), st->nvt, topo(st), st->act.M.view(),
                         st->stat.M.view(), st->mix.M.view(), st->P, (const float *)st->x.p, (const float *)st->x0.p,
                         st->kdx_on ? (const float *)st->kdx.p : nullptr, st->eacc.p);
    }
    NGUO_HIP(hipGetLastError());
    double e[EN_TERMS];
    st->eacc.download(e, EN_TERMS, s);
    NGUO_HIP(hipMemcpyAsync(st->x0.p, st->dx.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipMemcpyAsync(st->x.p, st->dx.p, n3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipStreamSynchronize(s));
    double tot = 0.0;
    for (int k = 0; k < EN_TERMS; ++k) {
      tot += e[k];
      if (terms) terms[k] = e[k];
    }
    if (total) *total = tot;
    return NGUO_OK;
  });
}

int nguo_step_bench_spmv(nguo_step *st, uint32_t reps, int iteration_form, double *avg_ms, double *bytes) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && avg_ms && bytes && reps > 0, NGUO_E_INVALID, "nguo_step_bench_spmv: bad argument");
    hipStream_t s = st->s;
    if (st->NF == 0) { *avg_ms = 0; *bytes = 0; return NGUO_OK; }
    // the system of the current state (elastic blocks + contact term) and a probe vector
    assemble_elastic(st, st->P, s);
    contacts(st);
    ContactView C = contact_view(st);
    float *xin = st->egrad.p;
    NGUO_HIP(hipMemcpyAsync(xin, st->x.p, 3 * (size_t)st->NF * sizeof(float), hipMemcpyDeviceToDevice, s));
    *avg_ms = st->pcg->time_spmv(st->act.M.view(), C, xin, reps, iteration_form != 0);
    // the iteration launch also reads each contact stencil's participants, gradients,
    // weights, coefficients, normal and incidence slots (116 B; the blocks it writes are not
    // counted)
    *bytes = st->act.M.spmv_bytes(iteration_form != 0) + (iteration_form ? 116.0 * (double)C.ns : 0.0);
    return NGUO_OK;
  });
}

int nguo_step_bench_precond(nguo_step *st, uint32_t reps, double *avg_ms, double *bytes) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(st && avg_ms && bytes && reps > 0, NGUO_E_INVALID, "nguo_step_bench_precond: bad argument");
    hipStream_t s = st->s;
    if (st->NF == 0) { *avg_ms = 0; *bytes = 0; return NGUO_OK; }
    // the system of the current state (elastic blocks + contact term)
    assemble_elastic(st, st->P, s);
    contacts(st);
    *avg_ms = st->pcg->time_precond(st->act.M, contact_view(st), reps, bytes);
    return NGUO_OK;
  });
}

// ---- batch of independent scenes advanced concurrently (C3, graded sizes)
struct nguo_batch {
  nguo_ctx *ctx = nullptr;
  std::vector<nguo_step *> inst;
  SolveGate gate;
  ~nguo_batch() {
    for (auto *st : inst) delete st;
  }
};

int nguo_batch_create(nguo_ctx *ctx, uint32_t n, const nguo_scene_desc *scenes, nguo_batch **out) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && scenes && out && n > 0, NGUO_E_INVALID, "nguo_batch_create: bad argument");
    std::unique_ptr<nguo_batch> b(new nguo_batch());
    b->ctx = ctx;
    for (uint32_t i = 0; i < n; ++i) {
      nguo_step *st = nullptr;
      int rc = nguo_step_create(ctx, &scenes[i], &st);
      if (rc != NGUO_OK) return fail(rc, "nguo_batch_create: instance " + std::to_string(i) + ": " + std::string(nguo_last_error()));
      b->inst.push_back(st);
    }
    *out = b.release();
    return NGUO_OK;
  });
}

void nguo_batch_destroy(nguo_batch *b) { delete b; }

// Every instance advances on its own stream from its own host thread (assembly, contacts,
// CCD and update interleave on the device). With nguo_batch_set_solve(b, 1) the instances'
// Newton solves meet in the batch's gate, which runs them as one batched PCG (two launches
// per iteration for all of them); by default each solves on its own stream. Each instance runs
// exactly the single-scene step (per-instance Newton loop, tolerance, CCD step) and its solves
// take the same iterations and values either way, so its trajectory is the one
// nguo_step_advance gives.
int nguo_batch_advance(nguo_batch *b, uint32_t nsteps) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(b, NGUO_E_INVALID, "nguo_batch_advance: null handle");
    const size_t n = b->inst.size();
    std::vector<int> code(n, NGUO_OK);
    std::vector<std::string> msg(n);
    std::vector<std::thread> th;
    th.reserve(n);
    for (size_t i = 0; i < n; ++i)
      th.emplace_back([&, i]() {
        try {
          NGUO_HIP(hipSetDevice(b->ctx->device));
          for (uint32_t k = 0; k < nsteps; ++k) advance_one(b->inst[i]);
          NGUO_HIP(hipStreamSynchronize(b->inst[i]->s));
        } catch (const HipError &e) {
          code[i] = e.code;
          msg[i] = e.msg;
        } catch (const std::bad_alloc &) {
          code[i] = NGUO_E_NOMEM;
          msg[i] = "host allocation failed";
        }
      });
    for (auto &t : th) t.join();
    for (size_t i = 0; i < n; ++i)
      if (code[i] != NGUO_OK) return fail(code[i], "nguo_batch_advance: instance " + std::to_string(i) + ": " + msg[i]);
    return NGUO_OK;
  });
}

int nguo_batch_set_solve(nguo_batch *b, int batched) {
  if (!b || (batched != 0 && batched != 1)) return fail(NGUO_E_INVALID, "nguo_batch_set_solve: bad argument");
  for (auto *st : b->inst) st->gate = batched ? &b->gate : nullptr;
  return NGUO_OK;
}

int nguo_batch_instance(nguo_batch *b, uint32_t i, nguo_step **out) {
  if (!b || !out || i >= b->inst.size()) return fail(NGUO_E_INVALID, "nguo_batch_instance: bad argument");
  *out = b->inst[i];
  return NGUO_OK;
}

int nguo_step_get_metrics(nguo_step *st, nguo_step_metrics *out) {
  if (!st || !out) return fail(NGUO_E_INVALID, "nguo_step_get_metrics: null argument");
  *out = st->metrics;
  return NGUO_OK;
}

}  // extern "C"
