// FP32 preconditioned CG with FP64 iterative refinement on a SELL-64 3x3-block matrix plus
// an optional matrix-free contact term. Used by the LinearSolver seam (nguo_ls_*) and by
// the resident Newton step.
#pragma once
#include <vector>

#include "device_common.h"
#include "nguo_internal.h"
#include "pcg_device.h"

namespace nguo {

// Pattern + values of a SELL-64 3x3-block matrix on the device, with its host pattern.
struct SellMatrix {
  uint32_t nb = 0, nslices = 0, nslots = 0, nnzb = 0;
  // nonzero: all slices padded to this width (when that costs at most ELL_MAX_PAD more
  // slots), so a slice's first slot is 64 s ell_w without reading slice_ptr
  uint32_t ell_w = 0;
  std::vector<uint32_t> h_slice_ptr, h_row_len, h_col;
  uint64_t pattern_id = 0;  // distinct for every build_pattern call in the process
  DevBuf<uint32_t> slice_ptr, row_len, col, diag_slot;
  DevBuf<float> val;
  // cols[i] = sorted block columns of block row i (must contain i).
  void build_pattern(uint32_t nblocks, const std::vector<std::vector<uint32_t>> &cols, hipStream_t s);
  uint32_t slot(uint32_t row, uint32_t k) const { return h_slice_ptr[row / 64] + 64u * k + (row & 63u); }
  // slot of block (row, col) from the host pattern, or 0xffffffff.
  uint32_t find(uint32_t row, uint32_t colj) const;
  SellView view() const { return SellView{nb, slice_ptr.p, row_len.p, col.p, val.p, ell_w}; }
  // Algorithmic bytes of one y = A x (f32 values, u32 column index per 3x3 block; row
  // length, x and y per block row). The PCG iteration's instance also reads z and writes p
  // (24 more bytes per block row).
  double spmv_bytes(bool iteration_form = false) const {
    return (double)nnzb * 40.0 + (double)nb * (iteration_form ? 52.0 : 28.0) + 4.0;
  }
};

// One-level additive Schwarz over disjoint vertex domains (engine-core/src/linsolve.rs:126-258
// SchwarzTopology): greedy BFS aggregation of the block graph into domains of `domain`
// vertices (at most 32, so a domain block is at most 96 x 96); per solve, one workgroup per
// domain assembles its dense block (elastic blocks + contact terms) in LDS, adds the
// diagonal floor 1e-6 tr/m + 1e-8, factors it in FP64 and writes the explicit FP32 inverse
// A_d^-1 = L^-T L^-1; the apply is one dense 96 x 96 product per domain.
constexpr uint32_t SCHWARZ_MAX_DOMAIN = 32;
constexpr uint32_t SCHWARZ_MAX_DOF = 3 * SCHWARZ_MAX_DOMAIN;
// floats of one domain's FP32 inverse: the 21 upper 16 x 16 tiles of the symmetric block
constexpr uint32_t SCHWARZ_INV_FLOATS = (SCHWARZ_MAX_DOF / 16) * (SCHWARZ_MAX_DOF / 16 + 1) / 2 * 256;

// A domain's tiled inverse: the upper 16 x 16 tiles (P <= Q < 6) of the symmetric 96 x 96
// block, tile (P, Q) at SCH_TILE64 sym_tile64(P, Q), each row-major.
constexpr uint32_t SCH_TILE64 = 256;
__host__ __device__ constexpr uint32_t sym_tile64(uint32_t P, uint32_t Q) { return P * 6u - P * (P - 1u) / 2u + (Q - P); }

struct Schwarz {
  uint32_t nd = 0, domain = SCHWARZ_MAX_DOMAIN;
  std::vector<uint32_t> h_dom_ptr, h_dom_row;
  // the matrix pattern's block neighbours of every row in row order (CSR), for the contact
  // domains' growth; rebuilt when the pattern (its column array) is another
  std::vector<uint32_t> h_nbr_ptr, h_nbr;
  uint64_t nbr_of = 0;  // SellMatrix::pattern_id of the lists (0: none)
  DevBuf<uint32_t> dom_ptr, dom_row, row_dom, row_pos, fails;
  DevBuf<uint32_t> dom_dof;  // [nd * 96]: stride-4 vector index of dof t of domain d, or ~0u past m
  DevBuf<float> ainv;  // nd blocks of SCHWARZ_INV_FLOATS (upper 16 x 16 tiles, row-major each)
  DevBuf<double> ainv64;  // the same tiles in FP64 (the FP64 PCG)
  // [2 nd SCHWARZ_MAX_DOMAIN], written with the inverses: the contact incidence range
  // (inc_ptr[row], inc_ptr[row + 1]) of row p of domain d at 2 (d SCHWARZ_MAX_DOMAIN + p)
  DevBuf<uint32_t> inc_rng;
  // Two-level form (engine-core/src/linsolve.rs:522-657 MultilevelSchwarz): the coarse node of
  // a vertex is its domain; the Galerkin coarse operator A1 = C A C^T sums the blocks between
  // two domains; a one-level Schwarz over A1 (coarse domains: BFS over the coarse graph, same
  // size) is the coarse smoother; z = S0 r + weight C^T S1 C r. weight 0 = one level.
  float weight = 0.0f;
  uint32_t nd1 = 0;
  std::vector<uint32_t> h_cdom_ptr, h_cdom_node;
  DevBuf<uint32_t> cdom_ptr, cdom_node, node_cdom, node_cpos;
  DevBuf<float> ainv1;  // nd1 coarse blocks of SCHWARZ_MAX_DOF^2, row-major with stride SCHWARZ_MAX_DOF
  DevBuf<float> rc;     // [4 nd] restricted residual (sum of a domain's rows), per apply
  void build(const SellMatrix &M, uint32_t domain_size, hipStream_t s, float coarse_weight = 0.0f);
  // Domains for a system with contact stencils (ns; their free-vertex rows, (hp, lambda),
  // distance gradients and proximity weights on the host): the strongest couplings kept inside
  // domains (see the definition). Domain size and coarse weight stay those of build().
  void build_contact(const SellMatrix &M, uint32_t ns, const uint32_t *dof, const float *coef, const float *g, const float *w, hipStream_t s);
  // Back to build()'s elastic-pattern domains (after build_contact).
  void restore_elastic(const SellMatrix &M, hipStream_t s);
  // Sizes the per-domain device arrays for up to nd_cap domains (contact-built domain sets
  // vary in count from system to system). Reallocates without keeping contents: call before
  // build().
  void reserve(uint32_t nd_cap);
  bool contact_built = false;
  std::vector<uint32_t> h_row_dom, h_row_pos, h_dom_dof, h_node_cdom, h_node_cpos;  // upload sources
  std::vector<uint32_t> elastic_ptr, elastic_row;  // build()'s domains

 private:
  void upload(const SellMatrix &M, hipStream_t s);
};

struct PcgStats {
  uint32_t iters = 0, sweeps = 0;
  double residual = 0;
  bool breakdown = false;
};

// Additive Schwarz preconditioner kernel of the PCG iteration (one workgroup per domain).
// init: r = (f32) r64, d = 0, z = M^-1 r, p = z. Otherwise iteration j: alpha = rz_j / pAp
// (npap iteration-launch partials; rz_j in scal[8 + (j & 1)]), q += the row's contact blocks
// contrib over the row's incidence range recorded by the build (in order; contrib null: no
// contact term), r -= alpha q, d += alpha p, z = M^-1 r. Writes the r.z | r.r partials of the
// new z to nxt (2 nd).
void launch_schwarz_pcg(hipStream_t s, const Schwarz &sz, bool init, const double *r64, float *r, float *z, float *dsol, float *p,
                        const float *q, const float4 *contrib, const double *pap_part, uint32_t npap, double *nxt, double *scal, uint32_t *flag,
                        uint32_t j);
// Iteration form of launch_schwarz_pcg over several one-level systems in one launch.
void launch_schwarz_pcg_batch(hipStream_t s, const BatchOf<SchArgs> &B, uint32_t j);
// Two-level form only, right after launch_schwarz_pcg: z += weight C^T S1 (C r) on every
// domain's rows (and p = z for init); replaces the r.z partials in nxt.
void launch_schwarz_coarse(hipStream_t s, const Schwarz &sz, bool init, const float *r, float *z, float *p, double *nxt, const uint32_t *flag);

class Pcg;
// One system of a batched solve (Pcg::solve_batch); stats receives its result.
struct PcgBatchItem {
  Pcg *pcg;
  const SellMatrix *M;
  ContactView C;
  const float *b;
  float tol;
  uint32_t max_iter;
  float *x_out;
  PcgStats stats;
};

class Pcg {
 public:
  explicit Pcg(hipStream_t s) : s_(s) {}
  ~Pcg();
  // Solve (A + C) x = b to ||b - (A+C) x|| <= tol ||b|| (FP64 residual). b is a device
  // f32 vector [3 nb]; the solution lands in x_out (device f32 [3 nb]).
  PcgStats solve(const SellMatrix &M, const ContactView &C, const float *b, float tol, uint32_t max_iter, float *x_out);
  // Solves n <= MAX_BATCH systems, each with its own Pcg (one-level Schwarz), together on
  // stream s: refinement sweeps run in lockstep over the systems that still need one, and
  // each PCG iteration is two launches spanning all of them (a system whose sweep has
  // converged exits those launches at its first instruction). Every system takes exactly
  // the iterations, sweeps and values its own solve() would. The caller orders s after the
  // systems' assembly and its own streams after s.
  static void solve_batch(hipStream_t s, PcgBatchItem *items, uint32_t n);
  bool batchable() const { return sz_ && sz_->nd > 0 && sz_->nd1 == 0; }
  // Sizes the solver's buffers for nb block rows, up to ns_cap contact stencils and nz_cap
  // preconditioner partials (no allocation on later solves within these).
  void reserve(uint32_t nb, uint32_t ns_cap, uint32_t nz_cap);
  // y = (A + C) x in f32 (stride-3 vectors), for parity checks.
  void apply(const SellView &A, const ContactView &C, const float *x, float *y);
  // Average duration (ms) of the PCG's SpMV kernel on (A + C) x over `reps` back-to-back
  // launches, between two HIP events on the solver stream. iteration_form: the iteration
  // launch (also reduces the preconditioner's partials, gathers z, evaluates p = z + beta
  // p_old, stores p, evaluates the contact stencils' blocks and the p.q partials); otherwise
  // y = A x plus the contact blocks (evaluated once, before the timed launches).
  double time_spmv(const SellView &A, const ContactView &C, const float *x, uint32_t reps, bool iteration_form);
  // Average duration (ms) of the preconditioner's iteration launch (the Schwarz apply with the
  // PCG update and the row's contact blocks) over `reps` back-to-back launches on (M, C), after
  // building the preconditioner for it; bytes receives the launch's algorithmic bytes (the
  // inverses' upper triangles, 32 B per dof of vectors and domain rows, the contact ranges and
  // blocks). 0 without a Schwarz preconditioner.
  double time_precond(const SellMatrix &M, const ContactView &C, uint32_t reps, double *bytes);
  // Block-Jacobi inverse of (A + C)'s diagonal blocks (FP64 inversion, scalar fallback).
  void build_precond(const SellMatrix &M, const ContactView &C);
  // Use `sz` (built over the same matrix pattern) instead of 3x3 block Jacobi; null = Jacobi.
  void set_schwarz(Schwarz *sz) { sz_ = sz; }
  Schwarz *schwarz() const { return sz_; }
  // z = M^-1 r with the current preconditioner (after build_precond), for parity checks.
  void apply_precond(const SellMatrix &M, const float *r, float *z);
  uint32_t chunk = 8;  // iterations per host convergence check, at least (grows with the sweep)
  // SpMV slot-group size (slots whose loads are in flight at once), picked by timing the
  // instantiated sizes on the live system at the first solve of a system size
  uint32_t spmv_group() const { return spmv_group_; }

 private:
  void ensure(uint32_t nb, uint32_t ns);
  void launch_minv(const SellMatrix &M, const ContactView &C);
  void launch_precond_init(const SellMatrix &M, double *part);
  hipStream_t s_;
  Schwarz *sz_ = nullptr;
  uint32_t cap_nb_ = 0, cap_ns_ = 0;
  uint32_t spmv_group_ = 16, tuned_nb_ = 0;     // single-system SpMV
  uint32_t batch_group_ = 16, batch_tuned_ = 0;  // batched SpMV (tuned on the total slices)
  DevBuf<float> r_, z_, p_, p2_, q_, d_, minv_;
  DevBuf<double> x64_, r64_, part_, scal_, sums64_;
  DevBuf<uint32_t> flag_;
  uint32_t *h_flag_ = nullptr;  // two slots of 4: state, iterations, cap (coherent pinned memory)
  uint32_t *d_flag_ = nullptr;  // the device's address of h_flag_
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double *h_scal_ = nullptr;
};

// The FP64 PCG's preconditioner build: Schwarz::build's domains, blocks and factorisation
// with the inverses kept in FP64 (sz.ainv64).
void build_schwarz64(hipStream_t s, const SellMatrix &M, const ContactView &C, Schwarz &sz);

// Preconditioned CG in FP64 on the same system: the matrix keeps its FP32 values (read into
// FP64 sums), the contact term is applied matrix-free from its FP32 stencil data in FP64,
// vectors, scalars and the one-level Schwarz inverses are FP64. At the drape's contact states
// the barrier curvature reaches ~1e11 against a mass term of ~0.04, so the system's condition
// number is ~1e12: FP32 vectors and inverses cannot hold both scales (the FP32 iteration
// stalls near 1e-4 and leans on refinement sweeps), while FP64 converges in the iteration
// count of exact arithmetic. Each iteration is four launches: the stencils' blocks of H p,
// q = (A + C) p with the p.q partials, the Schwarz launch (alpha, r -= alpha q, x += alpha p,
// z = A_d^-1 r, the r.z | r.r partials), and p = z + beta p with the convergence state.
class Pcg64 {
 public:
  explicit Pcg64(hipStream_t s) : s_(s) {}
  ~Pcg64();
  // Solve (A + C) x = b to ||b - (A+C) x|| <= tol ||b|| (FP64 residual) with the one-level
  // Schwarz `sz` (built over M's pattern); b, x_out device f32 [3 nb].
  PcgStats solve(const SellMatrix &M, const ContactView &C, Schwarz &sz, const float *b, float tol, uint32_t max_iter, float *x_out);
  uint32_t chunk = 16;  // iterations per host convergence check

 private:
  void ensure(uint32_t nb, uint32_t ns, uint32_t nd);
  hipStream_t s_;
  uint32_t cap_nb_ = 0, cap_ns_ = 0, cap_nd_ = 0;
  DevBuf<double> r_, z_, p_, q_, contrib_, x64_, r64_, part_, nxt_, scal_, sums64_;
  DevBuf<uint32_t> flag_;
  uint32_t *h_flag_ = nullptr;  // two slots of 4 (pinned)
  double *h_scal_ = nullptr;
  hipEvent_t ev_[2] = {nullptr, nullptr};
};

}  // namespace nguo
