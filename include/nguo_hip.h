/*
 * nguo_hip.h — C ABI of the MI355X (gfx950) implicit-IPC inner loop.
 *
 * One shared library, libnguo_hip.so, exports everything declared here. The boundary
 * carries plain pointers and sizes only: the caller owns every host buffer it passes,
 * the library owns every device buffer it creates. Every entry point returns an int
 * status (NGUO_OK == 0, negative on error); nguo_last_error() returns the message of the
 * last failure on the calling thread. A handle is used from one thread at a time and may
 * be moved between threads (the reference's solver objects are Send, not Sync).
 *
 * Reference interfaces each group replaces (paths under core-engine/crates/):
 *   nguo_ls_*    engine-core/src/linsolve.rs:21-28   trait LinearSolver { solve, last_stats }
 *                engine-gpu/src/pcg.rs:737-774       GpuLinearSolver (the wgpu PCG)
 *   nguo_step_*  engine-gpu/src/gpu_implicit.rs:159-659  GpuImplicitSolver::{new,step,positions}
 *                engine-core/src/implicit.rs:1951-2191   ImplicitSolver::step (the CPU tier)
 *   nguo_lbvh_*  engine-gpu/src/morton.rs:61-230, radix.rs:102-135, karras.rs:91-262,
 *                bvh_aabb.rs, bvh_traverse.rs:77-260 (GpuMorton/GpuRadixSort/GpuKarras/...)
 *   nguo_scene_* engine-api/src/scene.rs:89-229 (Scene::lower + build_implicit): the
 *                synthetic workload generators that stand in for the scene JSON.
 */
#ifndef NGUO_HIP_H
#define NGUO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGUO_OK 0
#define NGUO_E_INVALID (-1)     /* bad argument / malformed scene (Result::Err in the reference) */
#define NGUO_E_DEVICE (-2)      /* HIP runtime error or no device */
#define NGUO_E_NOMEM (-3)       /* device allocation failed */
#define NGUO_E_CAPACITY (-4)    /* a fixed-capacity list overflowed (contacts, candidates) */
#define NGUO_E_UNSUPPORTED (-5) /* input the device path does not handle (stated in the message) */

#define NGUO_SHELL_STVK 0
#define NGUO_SHELL_ARAP 1
#define NGUO_BARRIER_LOG 0
#define NGUO_BARRIER_CUBIC 1
#define NGUO_PRECOND_BLOCK 0   /* 3x3 block Jacobi (engine-gpu/src/pcg.rs GpuPrecond::Block) */
#define NGUO_PRECOND_SCHWARZ 1 /* additive Schwarz over vertex domains (GpuPrecond::SchwarzMl) */

#define NGUO_COLLIDER_HALFSPACE 0
#define NGUO_COLLIDER_SPHERE 1

/* engine-core/src/contact.rs:21-31 Collider. HalfSpace: a = point, b = unit normal.
 * Sphere: a = centre, radius. */
typedef struct nguo_collider {
  int32_t kind;
  float a[3];
  float b[3];
  float radius;
} nguo_collider;

/* The subset of engine-core/src/implicit.rs:30-229 ImplicitParams the path reads. */
typedef struct nguo_params {
  float dt;
  float gravity[3];
  uint32_t newton_iters;
  float bend_stiffness;   /* hinge bending k_b (weight w = 1 per hinge) */
  float shell_mu;         /* isotropic shell Lame mu */
  float shell_lambda;     /* isotropic shell Lame lambda */
  int32_t shell_model;    /* NGUO_SHELL_* */
  float eiganalysis_eps;  /* degenerate-eigenvalue guard of the shell eigensystem */
  float contact_dhat;     /* barrier activation distance */
  float contact_kappa;    /* barrier base stiffness (fallback of the Rayleigh stiffness) */
  float friction;         /* Coulomb mu for contact stencils */
  float friction_eps;     /* slip regularisation */
  int32_t barrier_kind;   /* NGUO_BARRIER_* */
  int32_t self_collision; /* vertex-triangle + edge-edge contact (incl. static mesh colliders) */
  float max_dx;           /* uniform search-direction cap, INFINITY = off */
  float target_toi;       /* Newton termination target, 0 = off */
  uint32_t min_newton_steps;
  uint32_t ccd_max_iter;
  float cg_tol;           /* relative residual ||b-Ax||/||b|| of each Newton solve */
  uint32_t cg_max_iter;
  float floor_y;          /* -INFINITY = no floor */
  int32_t precond;        /* NGUO_PRECOND_*: 3x3 block Jacobi or additive Schwarz */
  uint32_t schwarz_domain;/* vertices per Schwarz domain (1..32) */
  float schwarz_weight;   /* coarse weight of the two-level Schwarz form; 0 = one level
                             (engine-core/src/linsolve.rs:522-657, GpuPrecond::SchwarzMl) */
  float stretch_k;        /* edge-spring stiffness k over every unique triangle edge, rest
                             lengths from the scene positions (ImplicitParams::stiffness over
                             SimState::stretch, implicit.rs:2487-2507, spring_block
                             :3240-3261); 0 = no springs */
} nguo_params;

/* A scene in the solver's vertex space (engine-api/src/scene.rs:89-152 `lower` output plus
 * the ImplicitSolver setters implicit.rs:819-1563). All arrays are caller-owned. */
typedef struct nguo_scene_desc {
  uint32_t n_verts;
  const float *positions;        /* [3*n_verts] */
  const float *velocities;       /* [3*n_verts] or NULL (zero) */
  const float *inv_mass;         /* [n_verts], 0 = pinned */
  uint32_t n_tris;
  const uint32_t *tris;          /* [3*n_tris] all triangles: collision + hinge topology */
  const uint8_t *static_collider;/* [n_verts] static mesh collider mask or NULL */
  uint32_t n_shell_faces;
  const uint32_t *shell_faces;   /* [3*n_shell_faces] faces carrying the shell model */
  uint32_t n_rods;
  const uint32_t *rod_verts;     /* [2*n_rods] */
  const float *rod_rest;         /* [n_rods] */
  const float *rod_k;            /* [n_rods] */
  uint32_t n_rod_hinges;
  const uint32_t *rod_hinge_verts; /* [3*n_rod_hinges] (j, i = centre, k) */
  const float *rod_hinge_rest;   /* [n_rod_hinges] rest angle, pi = straight */
  const float *rod_hinge_stiff;  /* [n_rod_hinges] */
  uint32_t n_colliders;
  const nguo_collider *colliders;
  nguo_params params;
} nguo_scene_desc;

/* engine-core/src/linsolve.rs:13-18 LinearStats (+ refinement sweeps). */
typedef struct nguo_stats {
  uint32_t iters;        /* inner PCG iterations, summed over refinement sweeps */
  uint32_t sweeps;       /* FP64 refinement sweeps */
  double residual;       /* final ||b - A x|| / ||b||, evaluated in FP64 */
} nguo_stats;

/* engine-core/src/metrics.rs:12-35 StepMetrics (the fields the device path fills). */
typedef struct nguo_step_metrics {
  double time;
  double max_reresid;
  uint32_t max_cg_iters;
  uint32_t newton_iters;
  uint32_t num_stencils;    /* active contact stencils at the last Newton iteration */
  float max_velocity;
  int32_t penetration_free; /* 1 = no edge pierces a triangle and all finite */
  int32_t all_finite;
  float last_toi;           /* CCD step fraction of the last Newton iteration */
  uint32_t num_contacts;    /* free vertices within dhat of an analytic collider (implicit.rs:2212-2230) */
  float max_sigma;          /* max |x_a - x_b| / rest over the triangle mesh's unique edges, rest from
                               the scene positions (implicit.rs:2193-2209, state.rs:39-65); 0: no triangles */
  uint32_t device_alloc;    /* device allocations made during the step (0 once capacities settle) */
  uint32_t device_free;
  uint32_t sum_cg_iters;    /* inner PCG iterations summed over the step's Newton solves */
} nguo_step_metrics;

typedef struct nguo_ctx nguo_ctx;
typedef struct nguo_ls nguo_ls;
typedef struct nguo_step nguo_step;
typedef struct nguo_scene_buf nguo_scene_buf;

const char *nguo_last_error(void);

/* ---- device context (engine-gpu/src/lib.rs:79-196 GpuContext::new) */
int nguo_ctx_create(int device, nguo_ctx **out);
void nguo_ctx_destroy(nguo_ctx *ctx);
/* Device name and CU count, for reports. */
int nguo_ctx_info(nguo_ctx *ctx, char *name, size_t name_len, int *cu_count);

/* ---- LinearSolver seam (linsolve.rs:21-28; pcg.rs:737-774 GpuLinearSolver::solve).
 * create(): precond NGUO_PRECOND_BLOCK or NGUO_PRECOND_SCHWARZ with schwarz_domain vertices
 * per domain (1..32); schwarz_weight 0 = one level, > 0 = the two-level form with that coarse
 * weight (GpuPrecond::{Block, SchwarzMl{domain, weight}}, pcg.rs:32-39).
 * solve(): A is the scalar CSR the reference assembles (n rows, a multiple of 3 for the
 * block preconditioner), b the right-hand side, tol relative to ||b||. x_out receives the
 * solution; stats (may be NULL) receives what last_stats() would return. */
int nguo_ls_create(nguo_ctx *ctx, int precond, uint32_t schwarz_domain, float schwarz_weight, nguo_ls **out);
void nguo_ls_destroy(nguo_ls *ls);
int nguo_ls_solve(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                  const float *values, const float *b, float tol, uint32_t max_iter,
                  float *x_out, nguo_stats *stats);
int nguo_ls_last_stats(nguo_ls *ls, nguo_stats *out);
/* z = M^-1 r with the solver's preconditioner built on this matrix (parity checks). */
int nguo_ls_precond_apply(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                          const float *values, const float *r, float *z);
/* y = A x through the device 3x3 BSR SpMV (the PCG's operator), for parity checks. */
int nguo_ls_spmv(nguo_ls *ls, uint32_t n, const uint32_t *row_ptr, const uint32_t *col_idx,
                 const float *values, const float *x, float *y);

/* ---- whole implicit step (GpuImplicitSolver / ImplicitSolver::step) */
int nguo_step_create(nguo_ctx *ctx, const nguo_scene_desc *scene, nguo_step **out);
void nguo_step_destroy(nguo_step *st);
int nguo_step_advance(nguo_step *st, uint32_t nsteps);
int nguo_step_get_positions(nguo_step *st, float *out /* [3*n_verts] */);
int nguo_step_get_velocities(nguo_step *st, float *out /* [3*n_verts] */);
/* positions of every vertex (pinned ones included: the trees over the pinned primitives are
 * rebuilt from them); velocities may be null (kept) */
int nguo_step_set_state(nguo_step *st, const float *positions, const float *velocities);
int nguo_step_get_metrics(nguo_step *st, nguo_step_metrics *out);
/* implicit.rs:1368-1398 set_kinematic_collider_positions: moves the static mesh collider to the
 * collider rows of x ([3*n_verts]; the other rows are not read), between steps, once per step.
 * Its positions and velocities (move / dt), pseudonormals, trees and element blocks follow, and
 * the move enters the next step's friction slip at contacts with it (implicit.rs:430-444
 * kinematic_dx, cleared by the step). Like the reference it teleports the body: motion beyond
 * the barrier's reach within one step is not resolved. NGUO_E_INVALID (nothing changed) without
 * a collider or with a non-finite collider position, where the reference returns false. */
int nguo_step_set_kinematic_collider_positions(nguo_step *st, const float *x);
/* Positions after Newton iteration k of the last advanced step (the oracle's per-iteration
 * parity point); *count receives the number of iterations taken, toi their CCD fraction.
 * out == NULL only queries the count (and toi of iteration k). Positions are kept only for steps advanced with
 * tracing on (nguo_step_set_trace(st, 1); off by default: the copy is diagnostics and stays
 * out of the production step); asking for them otherwise is NGUO_E_INVALID. No reference
 * counterpart: the reference logs per-iteration scalars only (implicit.rs:2183-2190). */
int nguo_step_set_trace(nguo_step *st, int enable);
/* The Newton solves' arithmetic. NGUO_PCG_FP32_REFINED (default): FP32 PCG (vectors, matrix,
 * Schwarz inverses) with FP64 iterative refinement on the FP64 residual. NGUO_PCG_FP64: PCG
 * with FP64 vectors, scalars and Schwarz inverses over the same FP32 matrix values and contact
 * data; needs the one-level Schwarz preconditioner (NGUO_E_UNSUPPORTED otherwise). Both stop at
 * the FP64 residual cg_tol. No reference counterpart (the reference's CPU CG is f64, its wgpu
 * PCG f32, engine-core/src/linsolve.rs:350-411, engine-gpu/src/pcg.rs:488-673). */
#define NGUO_PCG_FP32_REFINED 0
#define NGUO_PCG_FP64 1
int nguo_step_set_precision(nguo_step *st, int precision);
int nguo_step_get_newton_trace(nguo_step *st, uint32_t k, float *out, uint32_t *count, float *toi);
/* The active contact stencils at positions x, as implicit.rs:678-753
 * self_collision_stencils returns them (vertex-triangle sorted by (v, t), then edge-edge
 * sorted by (ei, ej)): idx [4 cap], gap d [cap], distance gradient g [12 cap] and the
 * Rayleigh stiffness kappa [cap] (implicit.rs:755-795) against the elastic Hessian at x.
 * *count receives the number of stencils; at most cap are written. State is unchanged. */
int nguo_step_stencils(nguo_step *st, const float *x, uint32_t cap, uint32_t *idx, float *d, float *g, float *kappa, uint32_t *count);
/* implicit.rs:1431-1476 collider_sides at positions x: -1 inside the static body, +1 else.
 * Both this and nguo_step_stencils evaluate at any x that keeps the state's pinned positions
 * (the trees over pinned primitives are built from the state); an x that moves a pinned
 * vertex is NGUO_E_INVALID (set the state first). */
int nguo_step_collider_sides(nguo_step *st, const float *x, float *side);

/* Step energy at positions x (the incremental potential the Newton iteration of the last
 * advanced step decreases: its inertial target and friction reference, contact quantities at
 * x), evaluated in FP64 from the FP32 positions. terms (may be NULL) receives the
 * NGUO_ENERGY_TERMS parts: inertia 1/2 m/h^2 |x - x~|^2 (implicit.rs:2392-2405), shells
 * (shell_model.rs:112-127), hinges (bending.rs:105-123), rod stretch (implicit.rs:3287-3308),
 * strand bends (rod_bend.rs:1-20), stencil barriers (contact.rs:580-591, 1014-1025), stencil
 * friction potentials (contact.rs:613-706 integrated with the normal force held), analytic
 * colliders (barrier + friction, contact.rs:1082-1125), edge springs 1/2 k (l - l0)^2
 * (implicit.rs:2487-2507); total (may be NULL) their sum. x
 * must keep the state's pinned positions. No reference counterpart: the reference logs only
 * the kinetic energy (implicit.rs:2183-2190, state.rs:84-92). State is unchanged. */
#define NGUO_ENERGY_TERMS 9
int nguo_step_energy(nguo_step *st, const float *x, double *terms, double *total);

/* ---- drape adjoint (the differentiation form, C4). No reference counterpart: the reference
 * steps a fixed rest shape (engine-core/src/shell_model.rs:309-349 build_shells) and has no
 * gradient. Parameters are per shell face rest-shape scales s_f: face f's rest shape is
 * scaled by s_f (Dm^-1 / s_f, area s_f^2); default 1. nfaces = scene.n_shell_faces; scales
 * must be finite and > 0. Faces that build_shells skips (degenerate) get gradient 0. */
int nguo_step_set_rest_scale(nguo_step *st, const float *face_scale, uint32_t nfaces);
/* Gradient of a loss L(x*) through the equilibrium x* of the last advanced step: with
 * g(x; s) the step's gradient (inertia to its target, elastic terms, contact and lagged
 * friction), g(x*; s) = 0, so dL/ds_f = -lambda . dg/ds_f with H lambda = dL/dx, H the Newton
 * Hessian the step assembles (PSD-projected elastic blocks, rank-one barrier and friction
 * terms; hessian = NGUO_ADJ_EXACT_ELASTIC leaves the shell and rod-segment blocks unprojected),
 * solved by the same FP32 PCG with FP64 refinement to relative residual `tol`. A PCG
 * breakdown (H not positive definite at x*) returns NGUO_E_UNSUPPORTED.
 * dLdx [3 n_verts] (pinned rows ignored); dLds [n_shell_faces]. State is unchanged; the last
 * step must not have been followed by set_state or set_rest_scale. */
#define NGUO_ADJ_NEWTON 0        /* H as the Newton iteration assembles it */
#define NGUO_ADJ_EXACT_ELASTIC 1 /* shells and rod segments unprojected (their exact second derivative) */
/* The exact Jacobian of the step's gradient with the contact parameters held at x* (the stencil
 * set, Rayleigh stiffnesses, barrier kinds, friction normal / weights / lambda: the
 * frozen-parameter drape, whose derivative this is): the unprojected elastic blocks plus
 * kappa b'(d) d(grad d)/dx per stencil, kb (theta - rest) d(grad theta)/dx per hinge, the exact
 * strand blocks and a sphere's normal curvature. Solved as the fixed point of
 * lam = H0^-1 (dL/dx - H2 lam) (H0 the Newton Hessian, H2 = H - H0, each step one PCG solve to
 * tol) with Anderson mixing: stats->iters sums the PCG iterations, stats->sweeps counts the
 * solves, stats->residual is the final relative update; NGUO_E_UNSUPPORTED above 1e-3 (the
 * mixing also stops after 15 solves still above 1e-2: an indefinite exact Jacobian, as where
 * the step ended short of its root). */
#define NGUO_ADJ_EXACT 2
int nguo_step_adjoint(nguo_step *st, const float *dLdx, float tol, int hessian, float *dLds, nguo_stats *stats);

/* ---- batch of independent scenes (the graded-size form, C3). Each instance is a full step
 * object on its own stream; nguo_batch_advance advances all of them concurrently (one host
 * thread each), every instance taking exactly the steps nguo_step_advance would. Instances are
 * reached through nguo_batch_instance for positions, metrics and the other nguo_step_* calls
 * (not concurrently with nguo_batch_advance). No reference counterpart: the reference steps
 * one scene per ImplicitSolver (engine-core/src/implicit.rs:1951). */
typedef struct nguo_batch nguo_batch;
int nguo_batch_create(nguo_ctx *ctx, uint32_t n, const nguo_scene_desc *scenes, nguo_batch **out);
void nguo_batch_destroy(nguo_batch *batch);
int nguo_batch_advance(nguo_batch *batch, uint32_t nsteps);
int nguo_batch_instance(nguo_batch *batch, uint32_t i, nguo_step **out);
/* batched = 1: the instances' Newton solves run as one batched PCG whose two launches per
 * iteration span all instances (at most 8 per launch); 0 (default): each instance solves on
 * its own stream, concurrently with the others. The results are bitwise the same either way. */
int nguo_batch_set_solve(nguo_batch *batch, int batched);

/* ---- measurement (no reference counterpart: the reference has no GPU timestamps,
 * core-engine/docs/WGPU_PIN.md:142-144). Average duration of the PCG operator y = (A+C) x on
 * the current Newton system, timed with HIP events on the library stream over `reps`
 * back-to-back launches. iteration_form = 0: y = A x plus the contact blocks (evaluated once
 * before the timed launches), *bytes = nnzb*40 + nb*28 + 4; iteration_form = 1: the launch
 * the PCG iteration makes (also reduces the preconditioner's partials, gathers z, evaluates
 * and stores p = z + beta p_old, evaluates the contact stencils' blocks),
 * *bytes = nnzb*40 + nb*52 + 4 + 116 per contact stencil. */
int nguo_step_bench_spmv(nguo_step *st, uint32_t reps, int iteration_form, double *avg_ms, double *bytes);
/* The preconditioner's launch of the PCG iteration on the current Newton system (one-level
 * Schwarz: alpha, the row's contact blocks added to q, r and d updated, z = A_d^-1 r per
 * domain), event-timed the same way after building the preconditioner for that system;
 * *bytes = the inverses' upper triangles (2 m (m + 1) per domain of m dof) + 32 per dof
 * + with contact 8 per vertex and 12 per incidence. 0 ms for a block-Jacobi scene. */
int nguo_step_bench_precond(nguo_step *st, uint32_t reps, double *avg_ms, double *bytes);
/* A full 4-pass LSD radix sort of n random (u32 key, u32 index) pairs on the device (the
 * LBVH's sort), event-timed per sort; *bytes = 80 n (per pass: 4 B histogram read, 16 B
 * scatter read + write per pair). */
int nguo_bench_sort(nguo_ctx *ctx, uint32_t n, uint32_t reps, double *avg_ms, double *bytes);
/* The same operator on a synthetic nx x ny cloth-grid BSR (13 blocks per interior row). */
int nguo_bench_spmv_grid(nguo_ctx *ctx, uint32_t nx, uint32_t ny, uint32_t reps, double *avg_ms, double *bytes, double *nnzb);

/* ---- LBVH stages on a triangle set (bit-exact checks; engine-gpu/src/morton.rs:61-230,
 * radix.rs:112-135, karras.rs:91-262, bvh_aabb.rs). Host buffers in and out. */
int nguo_lbvh_morton(nguo_ctx *ctx, uint32_t nv, const float *pos, uint32_t nt, const uint32_t *tris, uint32_t *codes);
int nguo_lbvh_sort_pairs(nguo_ctx *ctx, uint32_t n, uint32_t *keys, uint32_t *vals);
int nguo_lbvh_karras(nguo_ctx *ctx, uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right);
int nguo_lbvh_aabb(nguo_ctx *ctx, uint32_t n, const uint32_t *left, const uint32_t *right, const float *leaf_box, float *node_box);

/* The sort of nguo_lbvh_sort_pairs on DEVICE buffers, allocating nothing: no hipMalloc,
 * hipFree or allocating library call. No counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). keys_dev / vals_dev (n u32 each, device memory of the
 * context's device) are sorted in place, stably by key, bitwise as nguo_lbvh_sort_pairs
 * sorts them. The intermediate buffers live in one file-scope device array per GPU (9.4 MB,
 * held until the process ends), so n is at most NGUO_NOALLOC_SORT_MAX (NGUO_E_CAPACITY
 * beyond). Every call on a GPU shares that GPU's array: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller. The call queues on
 * the context's stream and returns after the sort has finished. */
#define NGUO_NOALLOC_SORT_MAX (1u << 20)
int nguo_lbvh_sort_pairs_noalloc(nguo_ctx *ctx, uint32_t n, uint32_t *keys_dev, uint32_t *vals_dev);

/* ---- scene generators for the benchmark workloads (host only, no device needed).
 * config: 0 sheet on icosphere, 1 tube garment on a capsule body with seam rods,
 * 2 two tangled concentric tubes, 3 graded tube (param = circumference factor),
 * 100 small sheet on icosphere (tests). */
int nguo_scene_generate(int config, uint64_t seed, double param, nguo_scene_buf **out);
int nguo_scene_view(const nguo_scene_buf *buf, nguo_scene_desc *out);
void nguo_scene_free(nguo_scene_buf *buf);

#ifdef __cplusplus
}
#endif

#endif /* NGUO_HIP_H */
