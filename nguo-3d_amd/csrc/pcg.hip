// SELL-64 3x3-block SpMV and the FP32 PCG / FP64 iterative-refinement solver.
//
// Reference being replaced: engine-gpu/src/pcg.rs:488-673 Pcg::solve (scalar-CSR SpMV, one
// thread per scalar row, f32 alpha/beta, residual read back every 16 iterations) and the
// CPU oracle engine-core/src/linsolve.rs:350-411 cg_run (f64 block-Jacobi PCG). Here the
// matrix is stored as 3x3 blocks (36 B values + 4 B column per block instead of 9 x 8 B),
// the scalars alpha/beta/rz live on the device in FP64, vectors are FP32, and the true
// residual of each refinement sweep is recomputed in FP64.
#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pcg.h"
#include "pcg_device.h"

namespace nguo {

constexpr uint32_t NONE = 0xffffffffu;

// ---------------------------------------------------------------- pattern
// Padding (in slots) accepted to give every slice the same width.
constexpr double ELL_MAX_PAD = 1.15;

void SellMatrix::build_pattern(uint32_t nblocks, const std::vector<std::vector<uint32_t>> &cols, hipStream_t s) {
  static std::atomic<uint64_t> next_id{1};
  pattern_id = next_id.fetch_add(1);
  nb = nblocks;
  nslices = (nb + 63) / 64;
  h_slice_ptr.assign(nslices + 1, 0);
  h_row_len.assign(nb, 0);
  nnzb = 0;
  std::vector<uint32_t> width(nslices, 0);
  uint64_t sum_w = 0;
  uint32_t max_w = 0;
  for (uint32_t sl = 0; sl < nslices; ++sl) {
    uint32_t mx = 0;
    for (uint32_t r = sl * 64; r < std::min(nb, sl * 64 + 64); ++r) {
      h_row_len[r] = (uint32_t)cols[r].size();
      mx = std::max(mx, h_row_len[r]);
      nnzb += h_row_len[r];
    }
    width[sl] = mx;
    sum_w += mx;
    max_w = std::max(max_w, mx);
  }
  ell_w = (max_w > 0 && (double)max_w * nslices <= ELL_MAX_PAD * (double)sum_w) ? max_w : 0u;
  for (uint32_t sl = 0; sl < nslices; ++sl) h_slice_ptr[sl + 1] = h_slice_ptr[sl] + 64u * (ell_w ? ell_w : width[sl]);
  nslots = h_slice_ptr[nslices];
  h_col.assign(nslots, 0);
  std::vector<uint32_t> dslot(nb, NONE);
  for (uint32_t r = 0; r < nb; ++r) {
    uint32_t sl = r / 64, mx = (h_slice_ptr[sl + 1] - h_slice_ptr[sl]) / 64;
    for (uint32_t k = 0; k < mx; ++k) {
      uint32_t sidx = h_slice_ptr[sl] + 64u * k + (r & 63u);
      h_col[sidx] = k < cols[r].size() ? cols[r][k] : r;
      if (k < cols[r].size() && cols[r][k] == r) dslot[r] = sidx;
    }
  }
  // padding slots of rows past nb in the last slice keep col 0 (value 0, never read)
  slice_ptr.upload(h_slice_ptr, s);
  row_len.upload(h_row_len, s);
  col.upload(h_col, s);
  diag_slot.upload(dslot, s);
  val.zero((size_t)nslots * 9, s);
}

uint32_t SellMatrix::find(uint32_t row, uint32_t colj) const {
  uint32_t len = h_row_len[row];
  // columns are sorted: binary search over k
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) / 2;
    uint32_t c = h_col[slot(row, mid)];
    if (c < colj) lo = mid + 1;
    else hi = mid;
  }
  if (lo < len && h_col[slot(row, lo)] == colj) return slot(row, lo);
  return NONE;
}

// ---------------------------------------------------------------- kernels
__device__ inline void contact_gather64(const ContactView &C, uint32_t i, const double *sums, double &y0, double &y1, double &y2) {
  for (uint32_t k = C.inc_ptr[i]; k < C.inc_ptr[i + 1]; ++k) {
    uint32_t e = C.inc[k], s = e >> 2, a = e & 3u;
    double hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
    const double *sm = sums + 4 * (size_t)s;
    double t = hp * sm[0];
    const float *g = C.g + 12 * (size_t)s + 3 * a;
    y0 += t * g[0]; y1 += t * g[1]; y2 += t * g[2];
    if (lam != 0.0) {
      const float *n = C.n + 3 * (size_t)s;
      double nd = n[0] * sm[1] + n[1] * sm[2] + n[2] * sm[3];
      double pw = lam * (double)C.w[4 * (size_t)s + a];
      y0 += pw * (sm[1] - n[0] * nd); y1 += pw * (sm[2] - n[1] * nd); y2 += pw * (sm[3] - n[2] * nd);
    }
  }
}

template <class TX, class TS>
__global__ __launch_bounds__(BLOCK) void k_contact_sums(ContactView C, const TX *__restrict__ x, TS *__restrict__ sums, const uint32_t *flag) {
  if (flag && flag[0]) return;
  uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= C.ns) return;
  TS s1 = 0, a0 = 0, a1 = 0, a2 = 0;
  for (int a = 0; a < 4; ++a) {
    uint32_t d = C.dof[4 * (size_t)s + a];
    if (d == NONE) continue;
    TS x0 = x[3 * (size_t)d], x1 = x[3 * (size_t)d + 1], x2 = x[3 * (size_t)d + 2];
    const float *g = C.g + 12 * (size_t)s + 3 * a;
    s1 += (TS)g[0] * x0 + (TS)g[1] * x1 + (TS)g[2] * x2;
    TS w = C.w[4 * (size_t)s + a];
    a0 += w * x0; a1 += w * x1; a2 += w * x2;
  }
  sums[4 * (size_t)s] = s1; sums[4 * (size_t)s + 1] = a0; sums[4 * (size_t)s + 2] = a1; sums[4 * (size_t)s + 3] = a2;
}

// ---------------------------------------------------------------- PCG iteration kernels
// PCG vectors are stored with a stride of 4 floats per block row (x, y, z, 0): a gathered
// neighbour is one 16-byte load, and a row's three components are one 16-byte store. The
// fourth component is zero and never enters the arithmetic.
//
// One iteration j is two launches:
//   k_pcg_iter  every workgroup reduces rz_j and rr_j from the preconditioner's partials
//               (same order everywhere; slice workgroup 0 records the state) and forms
//               p_j = z_j + beta_j p_{j-1} wherever it reads p; the SELL slices' workgroups
//               (one wavefront, 64 block rows each) write q = A p_j and store p_j, the
//               stencils' workgroups (64 stencils each) write each contact stencil's blocks
//               of H_s p_j to its incidence slots; partials of p.q: p.(A p) per slice,
//               sum_a p_a . (H_s p)_a per stencil workgroup;
//   preconditioner (k_schwarz_pcg / k_jacobi_pcg) alpha = rz_j / pAp, q += the row's
//               contact blocks in incidence order, r -= alpha q, d += alpha p, z = M^-1 r,
//               partials of r.z and r.r.
// At these sizes a launch costs a few dependent memory round trips, and a workgroup's loads
// queue on its compute unit's memory pipeline, so the SpMV spreads over one wavefront per
// workgroup and issues every load of a group of slots before its FMAs. The plain form (y =
// (A + C) x, no iteration state) is a stencil launch, then the SpMV adding the blocks.
constexpr uint32_t SPMV_GROUP = 16;   // default: slots whose loads are in flight together
constexpr uint32_t PRE_STENCILS = 64; // stencils per stencil workgroup (one wavefront)
enum SpmvMode { SPMV_PLAIN = 0, SPMV_FIRST = 1, SPMV_NEXT = 2 };

__device__ inline float4 pnext(const float4 &zz, const float4 &pp, float beta) {
  return make_float4(zz.x + beta * pp.x, zz.y + beta * pp.y, zz.z + beta * pp.z, 0.0f);
}

// Loads of a group of slots: the block values and the gathered z / pold of the column. Every
// slot of the group is loaded, with no branch (slots past the slice's width repeat its last
// slot; padding slots hold a valid column), so all of the group's loads stay in flight
// together: a load under a branch leaves the compiler unable to count the loads outstanding
// at the next wait, which then drains all of them, one slot's round trip after another.
// Needs width >= 1.
template <int MODE, uint32_t G> struct SlotGroup {
  BlockVals v[G];
  float4 zc[G], pc[G];
  __device__ inline void load(const SellView &A, const float4 *z, const float4 *pold, uint32_t base, uint32_t k0, uint32_t width, uint32_t lane) {
    uint32_t grp[G], cc[G];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) grp[u] = base + 64u * min(k0 + u, width - 1u);
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) cc[u] = A.col[grp[u] + lane];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) v[u] = load_block(A.val, grp[u], lane);
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) {
      pc[u] = pold[cc[u]];
      if (MODE == SPMV_NEXT) zc[u] = z[cc[u]];
    }
  }
  // y += sum over the group's slots of B p (p = pold, or z + beta pold), slots ascending
  __device__ inline void accumulate(float &y0, float &y1, float &y2, float beta, uint32_t k0, uint32_t len) const {
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) {
      if (k0 + u < len) {
        float x0 = pc[u].x, x1 = pc[u].y, x2 = pc[u].z;
        if (MODE == SPMV_NEXT) { x0 = zc[u].x + beta * x0; x1 = zc[u].y + beta * x1; x2 = zc[u].z + beta * x2; }
        y0 += v[u].a.x * x0 + v[u].a.y * x1 + v[u].a.z * x2;
        y1 += v[u].a.w * x0 + v[u].b.x * x1 + v[u].b.y * x2;
        y2 += v[u].b.z * x0 + v[u].b.w * x1 + v[u].c * x2;
      }
    }
  }
};

// Opens iteration j in a workgroup of NT threads: rz_j, rr_j from the preconditioner's
// partials cur[0..nz) and cur[nz..2nz) (summed by the workgroup itself, in the same order as
// every other), beta_j = rz_j / rz_{j-1} (rz by iteration parity in scal[8 + (j & 1)]). The
// iteration stops on convergence (rr_j <= threshold) or at the cap; the writer workgroup
// records the state (a workgroup that reads the new state at its start exits as this test
// would make it). False when the iteration stops.
template <int NT>
__device__ inline bool open_iteration(const double *cur, uint32_t nz, double *scal, uint32_t *flag, uint32_t j, bool writer, double *lds,
                                      float &beta, uint32_t *hflag = nullptr) {
  const double *src[2] = {cur, cur + nz};
  const uint32_t n[2] = {nz, nz};
  double v[2];
  sum_parts_all<NT, 2, (NT < BLOCK ? 16u : 4u)>(src, n, v, lds);
  const double b = v[0] / scal[8 + ((j - 1) & 1u)];
  const bool conv = v[1] <= scal[5], capped = j >= flag[2];
  if (writer && threadIdx.x == 0) {
    flag[1] = j;
    scal[8 + (j & 1u)] = v[0];
    scal[3] = b;
    scal[4] = v[1];
    if (conv) flag[0] = 1;
    else if (capped) flag[0] = 3;
    if (hflag) {
      hflag[0] = conv ? 1u : (capped ? 3u : 0u);
      hflag[1] = j;
      hflag[2] = flag[2];
    }
  }
  beta = (float)b;
  return !(conv || capped);
}

// Contact stencils sb * 64 .. sb * 64 + 63, one per lane of a one-wavefront workgroup: the
// sums g . p and sum_a w_a p_a over the participants in order, p evaluated exactly as the
// SpMV evaluates it (NEXT: z + beta pold; otherwise pold), and each free participant's block
// of H_s p, hp (g.p) g_a + lambda w_a (s_w - n (n . s_w)), stored at its incidence slot.
// Iteration forms also write dpart[sb] = sum over the stencils of sum_a p_a . block_a. The
// stencil's data and its participants' z / pold entries need no scalar, so they are loaded
// before the iteration's reduction and are in flight while it runs.
template <int MODE>
__device__ __forceinline__ void pcg_stencil_body(uint32_t sb, const ContactView &C, const float4 *__restrict__ z, const float4 *__restrict__ pold,
                                                 const double *cur, uint32_t nz, double *scal, uint32_t *flag, uint32_t j, double *dpart) {
  __shared__ double lds[2];
  const uint32_t state = MODE != SPMV_PLAIN ? flag[0] : 0u;
  const uint32_t s = sb * PRE_STENCILS + threadIdx.x;
  const bool act = s < C.ns;
  uint32_t db[4], kp[4];
  float g[12], w[4], n[3], hp = 0.f, lam = 0.f;
  float4 zc[4], pc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) { db[b] = NONE; kp[b] = NONE; w[b] = 0.f; }
  if (act) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      db[b] = C.dof[4 * (size_t)s + b];
      kp[b] = C.pos[4 * (size_t)s + b];
      w[b] = C.w[4 * (size_t)s + b];
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) g[k] = C.g[12 * (size_t)s + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) n[k] = C.n[3 * (size_t)s + k];
    hp = C.coef[2 * (size_t)s];
    lam = C.coef[2 * (size_t)s + 1];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (db[b] != NONE) {
        pc[b] = pold[db[b]];
        if (MODE == SPMV_NEXT) zc[b] = z[db[b]];
      }
  }
  if (state) return;
  float beta = 0.0f;
  if (MODE == SPMV_NEXT && !open_iteration<PRE_STENCILS>(cur, nz, scal, flag, j, false, lds, beta)) return;
  double dp = 0.0;
  if (act) {
    float s1 = 0.f, a0 = 0.f, a1 = 0.f, a2 = 0.f;
    float4 xs[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      xs[b] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (db[b] == NONE) continue;
      const float4 x = MODE == SPMV_NEXT ? pnext(zc[b], pc[b], beta) : pc[b];
      xs[b] = x;
      s1 += g[3 * b] * x.x + g[3 * b + 1] * x.y + g[3 * b + 2] * x.z;
      a0 += w[b] * x.x; a1 += w[b] * x.y; a2 += w[b] * x.z;
    }
    const float t = hp * s1;
    const float nd = n[0] * a0 + n[1] * a1 + n[2] * a2;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      if (kp[a] == NONE) continue;
      float c0 = t * g[3 * a], c1 = t * g[3 * a + 1], c2 = t * g[3 * a + 2];
      if (lam != 0.0f) {
        const float pw = lam * w[a];
        c0 += pw * (a0 - n[0] * nd); c1 += pw * (a1 - n[1] * nd); c2 += pw * (a2 - n[2] * nd);
      }
      C.contrib[kp[a]] = make_float4(c0, c1, c2, 0.0f);
      dp += (double)xs[a].x * c0 + (double)xs[a].y * c1 + (double)xs[a].z * c2;
    }
  }
  if (MODE == SPMV_PLAIN) return;
  dp = wave_sum(dp);
  if (threadIdx.x == 0) dpart[sb] = dp;
}

// A SELL slice (64 block rows, one per lane) of a one-wavefront workgroup: SPMV_GROUP slots at
// a time, column indices, then every value and gather load of the group, then the FMAs in
// slot order. PLAIN: q = (A + C) pold, the row's contact blocks (from the stencil launch)
// added in incidence order. FIRST: q = A p_0 (p_0 = pold). NEXT: q = A p_j, p_j = z + beta
// pold wherever p is read; the row's owner also stores p_j to pnew, and every reader of a
// neighbour's p evaluates the same expression, so they agree bitwise. Iteration forms write
// part[slice] = p.q over the slice.
template <int MODE, uint32_t G>
__device__ __forceinline__ void pcg_spmv_body(uint32_t slice, const SellView &A, const ContactView &C, const float4 *__restrict__ z,
                                              const float4 *__restrict__ pold, float4 *__restrict__ pnew, float4 *__restrict__ q, double *part,
                                              const double *cur, uint32_t nz, double *scal, uint32_t *flag, uint32_t j,
                                              uint32_t *hflag = nullptr) {
  __shared__ double lds[2];
  const uint32_t state = MODE != SPMV_PLAIN ? flag[0] : 0u;
  const uint32_t lane = threadIdx.x, row = slice * 64u + lane;
  const bool active = row < A.nb;
  // lanes past nb (last slice) read row nb - 1 and the slice's padding slots, and add nothing
  const uint32_t rr = active ? row : A.nb - 1u;
  const uint32_t base = A.ell_w ? 64u * A.ell_w * slice : A.slice_ptr[slice];
  const uint32_t width = A.ell_w ? A.ell_w : (A.slice_ptr[slice + 1] - base) / 64u;
  const uint32_t len0 = A.row_len[rr];
  uint32_t ib = 0, ie = 0;
  if (MODE == SPMV_PLAIN && C.ns) {
    ib = C.inc_ptr[rr];
    ie = C.inc_ptr[rr + 1];
  }
  const uint32_t len = active ? len0 : 0u;
  if (!active) ie = ib;
  SlotGroup<MODE, G> g;
  if (width) g.load(A, z, pold, base, 0, width, lane);
  float4 prow = make_float4(0.f, 0.f, 0.f, 0.f), zrow = prow;
  if (MODE != SPMV_PLAIN) {
    prow = pold[rr];
    if (MODE == SPMV_NEXT) zrow = z[rr];
  }
  float4 cv[8];
  if (MODE == SPMV_PLAIN && C.ns) {
    // contrib holds at least one entry whenever there are stencils
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) cv[u] = C.contrib[ie > ib ? min(ib + u, ie - 1u) : 0u];
  }
  // the first slice's workgroup publishes the flag to the host slot (hflag) on every path
  const bool publish = hflag && slice == 0 && lane == 0;
  if (state) {
    if (publish) { hflag[0] = state; hflag[1] = flag[1]; hflag[2] = flag[2]; }
    return;
  }
  if (MODE == SPMV_FIRST && publish) { hflag[0] = 0u; hflag[1] = flag[1]; hflag[2] = flag[2]; }
  float beta = 0.0f;
  if (MODE == SPMV_NEXT && !open_iteration<WAVE>(cur, nz, scal, flag, j, slice == 0, lds, beta, hflag)) return;
  double dp = 0.0;
  if (active) {
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    g.accumulate(y0, y1, y2, beta, 0, len);
    for (uint32_t k0 = G; k0 < width; k0 += G) {
      g.load(A, z, pold, base, k0, width, lane);
      g.accumulate(y0, y1, y2, beta, k0, len);
    }
    if (MODE == SPMV_PLAIN) {
      // past the first 8 blocks, 8 more per memory round trip (clamped to the row's last)
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u)
        if (ib + u < ie) { y0 += cv[u].x; y1 += cv[u].y; y2 += cv[u].z; }
      for (uint32_t k0 = ib + 8; k0 < ie; k0 += 8) {
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u) cv[u] = C.contrib[min(k0 + u, ie - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < 8; ++u)
          if (k0 + u < ie) { y0 += cv[u].x; y1 += cv[u].y; y2 += cv[u].z; }
      }
    }
    q[row] = make_float4(y0, y1, y2, 0.0f);
    if (MODE != SPMV_PLAIN) {
      const float4 pr = MODE == SPMV_NEXT ? pnext(zrow, prow, beta) : prow;
      if (MODE == SPMV_NEXT) pnew[row] = pr;
      dp = (double)pr.x * y0 + (double)pr.y * y1 + (double)pr.z * y2;
    }
  }
  if (MODE == SPMV_PLAIN) return;
  dp = wave_sum(dp);
  if (lane == 0) part[slice] = dp;
}

// Workgroup b of a system's iteration launch: a slice (b < slices) or a stencil group.
template <int MODE, uint32_t G> __device__ __forceinline__ void pcg_iter_body(uint32_t b, const SpmvArgs &a, uint32_t j) {
  const uint32_t ns = (a.A.nb + 63u) / 64u;
  if (b < ns)
    pcg_spmv_body<MODE, G>(b, a.A, a.C, a.z, a.pold, a.pnew, a.q, a.part, a.cur, a.nz, a.scal, a.flag, j, a.hflag);
  else
    pcg_stencil_body<MODE>(b - ns, a.C, a.z, a.pold, a.cur, a.nz, a.scal, a.flag, j, a.part + ns);
}

template <int MODE, uint32_t G> __global__ __launch_bounds__(WAVE) void k_pcg_iter(SpmvArgs a, uint32_t j) {
  pcg_iter_body<MODE, G>(blockIdx.x, a, j);
}

// The same over several systems (one grid, system i's workgroups at B.off[i]).
template <int MODE, uint32_t G> __global__ __launch_bounds__(WAVE) void kb_pcg_iter(BatchOf<SpmvArgs> B, uint32_t j) {
  const uint32_t i = batch_slot(B.off, B.k, blockIdx.x);
  pcg_iter_body<MODE, G>(blockIdx.x - B.off[i], B.a[i], j);
}

template <uint32_t G>
__global__ __launch_bounds__(WAVE) void k_pcg_spmv_plain(SellView A, ContactView C, const float4 *__restrict__ x, float4 *__restrict__ y) {
  pcg_spmv_body<SPMV_PLAIN, G>(blockIdx.x, A, C, nullptr, x, nullptr, y, nullptr, nullptr, 0u, nullptr, nullptr, 0u);
}
__global__ __launch_bounds__(WAVE) void k_pcg_stencils_plain(ContactView C, const float4 *__restrict__ x) {
  pcg_stencil_body<SPMV_PLAIN>(blockIdx.x, C, nullptr, x, nullptr, 0u, nullptr, nullptr, 0u, nullptr);
}

inline unsigned stencil_groups(const ContactView &C) { return div_up(C.ns, PRE_STENCILS); }
inline unsigned iter_grid(const SellView &A, const ContactView &C) { return div_up(A.nb, 64) + stencil_groups(C); }

// Host-side dispatch over the slot-group sizes the SpMV is instantiated for; the group only
// changes how many slots' loads are in flight at once (and so the registers and occupancy),
// not the order of the row's sum, so every size gives bitwise the same q.
template <int MODE> static void launch_iter(uint32_t G, hipStream_t s, const SpmvArgs &a, uint32_t j) {
  const unsigned grid = iter_grid(a.A, a.C);
  if (G == 8) hipLaunchKernelGGL((k_pcg_iter<MODE, 8>), dim3(grid), dim3(WAVE), 0, s, a, j);
  else hipLaunchKernelGGL((k_pcg_iter<MODE, 16>), dim3(grid), dim3(WAVE), 0, s, a, j);
}
template <int MODE> static void launch_iter_batch(uint32_t G, hipStream_t s, const BatchOf<SpmvArgs> &B, uint32_t j) {
  if (!B.k || !B.off[B.k]) return;
  if (G == 8) hipLaunchKernelGGL((kb_pcg_iter<MODE, 8>), dim3(B.off[B.k]), dim3(WAVE), 0, s, B, j);
  else hipLaunchKernelGGL((kb_pcg_iter<MODE, 16>), dim3(B.off[B.k]), dim3(WAVE), 0, s, B, j);
}
static void launch_plain(uint32_t G, hipStream_t s, const SellView &A, const ContactView &C, const float4 *x, float4 *y) {
  if (C.ns) hipLaunchKernelGGL(k_pcg_stencils_plain, dim3(stencil_groups(C)), dim3(WAVE), 0, s, C, x);
  if (G == 8) hipLaunchKernelGGL((k_pcg_spmv_plain<8>), dim3(div_up(A.nb, 64)), dim3(WAVE), 0, s, A, C, x, y);
  else hipLaunchKernelGGL((k_pcg_spmv_plain<16>), dim3(div_up(A.nb, 64)), dim3(WAVE), 0, s, A, C, x, y);
}
constexpr uint32_t SPMV_GROUP_CHOICES[2] = {8, 16};

// Average ms of `reps` launches of f() between two events on s (after two warm-up launches).
template <class F> static double time_launches(hipStream_t s, uint32_t reps, F f) {
  f();
  f();
  hipEvent_t e0, e1;
  NGUO_HIP(hipEventCreate(&e0));
  NGUO_HIP(hipEventCreate(&e1));
  NGUO_HIP(hipEventRecord(e0, s));
  for (uint32_t r = 0; r < reps; ++r) f();
  NGUO_HIP(hipEventRecord(e1, s));
  NGUO_HIP(hipEventSynchronize(e1));
  float ms = 0;
  NGUO_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  NGUO_HIP(hipGetLastError());
  return (double)ms / reps;
}

// The autotuner: the slot-group size whose launches of `launch(G)` are fastest on the live
// system (ties keep the larger group).
template <class L> static uint32_t pick_group(hipStream_t s, L launch) {
  uint32_t best = SPMV_GROUP;
  double best_ms = 1e30;
  for (uint32_t G : SPMV_GROUP_CHOICES) {
    const double ms = time_launches(s, 10, [&] { launch(G); });
    if (ms < best_ms * 0.98 || (ms <= best_ms && G > best)) { best_ms = std::min(best_ms, ms); best = G; }
  }
  return best;
}

// 3x3 block-Jacobi preconditioner step, one block row per thread (vectors of stride 4,
// r64 of stride 3). INIT: r = (f32) r64, d = 0, z = Minv r, p = z. Otherwise: pAp from the
// npap iteration-launch partials (every workgroup, same order), rz_j from scal[8 + (j & 1)],
// stop on breakdown, q += the row's contact blocks (incidence order; contrib null without
// contact), r -= alpha q, d += alpha p, z = Minv r.
// Partials r.z -> nxt[block], r.r -> nxt[gridDim + block].
template <bool INIT>
__global__ __launch_bounds__(BLOCK) void k_jacobi_pcg(uint32_t nb, const float *__restrict__ minv, const double *__restrict__ r64,
                                                     float *__restrict__ r, float *__restrict__ z, float *__restrict__ d,
                                                     float *__restrict__ p, const float *__restrict__ q, const uint32_t *inc_ptr,
                                                     const float4 *contrib, const double *pap_part, uint32_t npap, double *nxt, double *scal,
                                                     uint32_t *flag, uint32_t j) {
  __shared__ double lds[2 * BLOCK / WAVE];
  float alpha = 0.0f;
  if (!INIT) {
    if (flag[0]) return;
    const double *src[1] = {pap_part};
    const uint32_t n[1] = {npap};
    double v[1];
    sum_parts_all<BLOCK, 1>(src, n, v, lds);
    if (!(v[0] > 0.0)) {
      if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 2;
      return;
    }
    const double a = scal[8 + (j & 1u)] / v[0];
    alpha = (float)a;
    if (blockIdx.x == 0 && threadIdx.x == 0) { scal[1] = v[0]; scal[2] = a; }
  }
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < nb) {
    const size_t o = 4 * (size_t)i;
    float r0, r1, r2;
    if (INIT) {
      r0 = (float)r64[3 * (size_t)i]; r1 = (float)r64[3 * (size_t)i + 1]; r2 = (float)r64[3 * (size_t)i + 2];
      d[o] = 0.f; d[o + 1] = 0.f; d[o + 2] = 0.f;
    } else {
      float q0 = q[o], q1 = q[o + 1], q2 = q[o + 2];
      if (contrib)
        for (uint32_t k = inc_ptr[i]; k < inc_ptr[i + 1]; ++k) {
          const float4 c = contrib[k];
          q0 += c.x; q1 += c.y; q2 += c.z;
        }
      r0 = r[o] - alpha * q0; r1 = r[o + 1] - alpha * q1; r2 = r[o + 2] - alpha * q2;
      d[o] += alpha * p[o]; d[o + 1] += alpha * p[o + 1]; d[o + 2] += alpha * p[o + 2];
    }
    r[o] = r0; r[o + 1] = r1; r[o + 2] = r2;
    const float *m = minv + 9 * (size_t)i;
    const float z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
    const float z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
    const float z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
    z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
    if (INIT) { p[o] = z0; p[o + 1] = z1; p[o + 2] = z2; }
    v[0] = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
    v[1] = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
  }
  wg_sum_all<BLOCK, 2>(v, lds);
  if (threadIdx.x == 0) { nxt[blockIdx.x] = v[0]; nxt[gridDim.x + blockIdx.x] = v[1]; }
}

// r = b - (A + C) x in FP64; part = r.r per block.
__global__ __launch_bounds__(BLOCK) void k_residual64(SellView A, ContactView C, const double *sums64, const float *__restrict__ b,
                                                     const double *__restrict__ x, double *__restrict__ r, double *part) {
  __shared__ double lds[4];
  uint32_t row = blockIdx.x * BLOCK + threadIdx.x;
  double rr = 0;
  if (row < A.nb) {
    uint32_t base = A.slice_ptr[row >> 6], len = A.row_len[row], lane = row & 63u;
    double y0 = 0, y1 = 0, y2 = 0;
    for (uint32_t k = 0; k < len; ++k) {
      uint32_t c = A.col[base + 64u * k + lane];
      const BlockVals v = load_block(A.val, base + 64u * k, lane);
      double x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
      y0 += (double)v.a.x * x0 + (double)v.a.y * x1 + (double)v.a.z * x2;
      y1 += (double)v.a.w * x0 + (double)v.b.x * x1 + (double)v.b.y * x2;
      y2 += (double)v.b.z * x0 + (double)v.b.w * x1 + (double)v.c * x2;
    }
    if (C.ns) contact_gather64(C, row, sums64, y0, y1, y2);
    double r0 = (double)b[3 * (size_t)row] - y0, r1 = (double)b[3 * (size_t)row + 1] - y1, r2 = (double)b[3 * (size_t)row + 2] - y2;
    r[3 * (size_t)row] = r0; r[3 * (size_t)row + 1] = r1; r[3 * (size_t)row + 2] = r2;
    rr = r0 * r0 + r1 * r1 + r2 * r2;
  }
  double t = block_sum(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sums `n` partials in a fixed order (single workgroup); valid in thread 0.
__device__ inline double reduce_parts(const double *part, uint32_t n, double *lds) {
  double acc = 0;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) acc += part[i];
  return block_sum(acc, lds);
}

// Opens a sweep from the (r.z | r.r) partials of the preconditioner's INIT launch.
__global__ __launch_bounds__(BLOCK) void k_init_scal(double *scal, uint32_t *flag, const double *part, uint32_t npart, double eta2,
                                                    uint32_t cap) {
  __shared__ double lds[4];
  double rz = reduce_parts(part, npart, lds);
  double rr = reduce_parts(part + npart, npart, lds);
  if (threadIdx.x == 0) {
    scal[8] = rz; scal[4] = rr; scal[6] = rr; scal[5] = eta2 * rr;
    flag[0] = rr == 0.0 ? 1u : (cap == 0 ? 3u : 0u);
    flag[1] = 0;
    flag[2] = cap;
  }
}

__global__ __launch_bounds__(BLOCK) void k_sum_parts(const double *part, uint32_t npart, double *out) {
  __shared__ double lds[4];
  double s = reduce_parts(part, npart, lds);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_f32_to_f64(uint32_t n, const float *a, double *b) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) b[i] = a[i];
}
__global__ void k_fill64(uint32_t n, double *a, double v) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) a[i] = v;
}
// x (stride 3, FP64) += d (stride 4)
__global__ void k_accum(uint32_t n, double *x, const float *d) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) x[i] += (double)d[4 * (i / 3) + i % 3];
}
// stride 3 <-> stride 4 copies of a vector of nb rows (fourth component zero)
__global__ void k_pack4(uint32_t nb, const float *a, float4 *b) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < nb) b[i] = make_float4(a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2], 0.0f);
}
__global__ void k_unpack4(uint32_t nb, const float4 *a, float *b) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < nb) {
    const float4 v = a[i];
    b[3 * (size_t)i] = v.x; b[3 * (size_t)i + 1] = v.y; b[3 * (size_t)i + 2] = v.z;
  }
}
__global__ void k_to_f32(uint32_t n, const double *x, float *out) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = (float)x[i];
}

// Block-Jacobi inverse (engine-core/src/linsolve.rs:293-315): the 3x3 diagonal block of
// (A + C) inverted in FP64; |det| < 1e-20 falls back to 1/diag (1 where the diagonal is 0).
__global__ __launch_bounds__(BLOCK) void k_minv(SellView A, ContactView C, const uint32_t *diag_slot, float *minv) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= A.nb) return;
  double m[9];
  uint32_t ds = diag_slot[i];
  for (int e = 0; e < 9; ++e) m[e] = ds == NONE ? 0.0 : (double)A.val[sell_val_index(ds, e)];
  if (C.ns) {
    for (uint32_t k = C.inc_ptr[i]; k < C.inc_ptr[i + 1]; ++k) {
      uint32_t e = C.inc[k], s = e >> 2, a = e & 3u;
      double hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
      const float *g = C.g + 12 * (size_t)s + 3 * a;
      const float *n = C.n + 3 * (size_t)s;
      double w = C.w[4 * (size_t)s + a];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          m[3 * r + c] += hp * (double)g[r] * (double)g[c] + lam * w * w * ((r == c ? 1.0 : 0.0) - (double)n[r] * (double)n[c]);
    }
  }
  double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
  float *o = minv + 9 * (size_t)i;
  if (fabs(det) < 1e-20) {
    for (int e = 0; e < 9; ++e) o[e] = 0.f;
    for (int r = 0; r < 3; ++r) {
      float dv = (float)m[4 * r];
      o[4 * r] = dv != 0.0f ? (float)(1.0 / (double)dv) : 1.0f;
    }
    return;
  }
  double id = 1.0 / det;
  o[0] = (float)((m[4] * m[8] - m[5] * m[7]) * id);
  o[1] = (float)((m[2] * m[7] - m[1] * m[8]) * id);
  o[2] = (float)((m[1] * m[5] - m[2] * m[4]) * id);
  o[3] = (float)((m[5] * m[6] - m[3] * m[8]) * id);
  o[4] = (float)((m[0] * m[8] - m[2] * m[6]) * id);
  o[5] = (float)((m[2] * m[3] - m[0] * m[5]) * id);
  o[6] = (float)((m[3] * m[7] - m[4] * m[6]) * id);
  o[7] = (float)((m[1] * m[6] - m[0] * m[7]) * id);
  o[8] = (float)((m[0] * m[4] - m[1] * m[3]) * id);
}

// ---------------------------------------------------------------- host driver
// A refinement sweep asks the FP32 iteration for a reduction of its own (recursive) residual
// by eta = clamp(0.25 tol / res, REFINE_ETA_MIN, 0.5) relative to the sweep's start. The floor
// sits at the FP32 unit roundoff: a smaller reduction of the recursive residual would no
// longer improve the FP32 iterate. At a 1e-6 tolerance the first sweep thus aims at 2.5e-7
// and most solves take one sweep (a floor of 1e-5 forced a second, restarted Krylov sweep:
// 7% more iterations on the same drape states).
constexpr double REFINE_ETA_MIN = 1.0e-7;
// Iterations queued per host convergence check: at least Pcg::chunk, growing with the sweep
// to a sixteenth of the iterations done (at most CHUNK_MAX), so the host checks thin out on
// long solves while the launches queued past convergence stay a small share.
constexpr uint32_t CHUNK_MAX = 32;
inline uint32_t chunk_len(uint32_t chunk, uint32_t j) { return std::min(CHUNK_MAX, std::max(chunk, j / 16)); }

Pcg::~Pcg() {
  for (auto &e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (h_flag_) (void)hipHostFree(h_flag_);
  if (h_scal_) (void)hipHostFree(h_scal_);
}

void Pcg::ensure(uint32_t nb, uint32_t ns) {
  {
    // iteration-launch partials (one per slice and per stencil group), then the (r.z | r.r)
    // partials of the preconditioner kernel (per row block or per domain); the FP64 residual
    // reuses the front
    size_t S = div_up(nb, 64) + div_up(ns, PRE_STENCILS), G = div_up(nb, BLOCK), nz = std::max<size_t>(G, sz_ ? sz_->nd : 0);
    part_.reserve(std::max(S, G) + 2 * nz + 8);
  }
  if (!h_flag_) {
    // coherent and mapped: the iteration launch stores a chunk's flag here directly
    NGUO_HIP(hipHostMalloc((void **)&h_flag_, 8 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
    NGUO_HIP(hipHostGetDevicePointer((void **)&d_flag_, h_flag_, 0));
    for (auto &e : ev_) NGUO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    NGUO_HIP(hipHostMalloc((void **)&h_scal_, 8 * sizeof(double), hipHostMallocDefault));
  }
  if (nb > cap_nb_) {
    size_t n3 = 3 * (size_t)nb, n4 = 4 * (size_t)nb;
    // stride-4 vectors: the fourth components stay zero
    r_.zero(n4, s_); z_.zero(n4, s_); p_.zero(n4, s_); p2_.zero(n4, s_); q_.zero(n4, s_); d_.zero(n4, s_);
    minv_.reserve(9 * (size_t)nb);
    x64_.reserve(n3); r64_.reserve(n3);
    cap_nb_ = nb;
  }
  scal_.reserve(16);
  flag_.reserve(4);
  if (ns > cap_ns_) {
    const uint32_t c = std::max(ns, 1u << 14);  // contact stencils: sized like the step's contact lists
    sums64_.reserve(4 * (size_t)c);
    cap_ns_ = c;
  }
}

void Pcg::reserve(uint32_t nb, uint32_t ns_cap, uint32_t nz_cap) {
  ensure(nb, ns_cap);
  const size_t S = div_up(nb, 64) + div_up(ns_cap, PRE_STENCILS), G = div_up(nb, BLOCK);
  part_.reserve(std::max(S, G) + 2 * std::max<size_t>(G, nz_cap) + 8);
}

void Pcg::launch_minv(const SellMatrix &M, const ContactView &C) {
  hipLaunchKernelGGL(k_minv, dim3(div_up(M.nb, BLOCK)), dim3(BLOCK), 0, s_, M.view(), C, M.diag_slot.p, minv_.p);
}

__global__ __launch_bounds__(BLOCK) void k_bj_apply(uint32_t nb, const float *minv, const float *r, float *z) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= nb) return;
  const float *m = minv + 9 * (size_t)i;
  size_t o = 3 * (size_t)i;
  float r0 = r[o], r1 = r[o + 1], r2 = r[o + 2];
  z[o] = m[0] * r0 + m[1] * r1 + m[2] * r2;
  z[o + 1] = m[3] * r0 + m[4] * r1 + m[5] * r2;
  z[o + 2] = m[6] * r0 + m[7] * r1 + m[8] * r2;
}

// Opens a sweep: r = (f32) r64, d = 0, z = M^-1 r, p_ = z; (r.z | r.r) partials -> part.
void Pcg::launch_precond_init(const SellMatrix &M, double *part) {
  if (sz_) {
    launch_schwarz_pcg(s_, *sz_, true, r64_.p, r_.p, z_.p, d_.p, p_.p, q_.p, nullptr, nullptr, 0, part, scal_.p, flag_.p, 0);
    launch_schwarz_coarse(s_, *sz_, true, r_.p, z_.p, p_.p, part, flag_.p);
    return;
  }
  hipLaunchKernelGGL(k_jacobi_pcg<true>, dim3(div_up(M.nb, BLOCK)), dim3(BLOCK), 0, s_, M.nb, (const float *)minv_.p,
                     (const double *)r64_.p, r_.p, z_.p, d_.p, p_.p, (const float *)q_.p, (const uint32_t *)nullptr, (const float4 *)nullptr,
                     (const double *)nullptr, 0u, part, scal_.p, flag_.p, 0u);
}

void Pcg::apply_precond(const SellMatrix &M, const float *r, float *z) {
  const uint32_t nb = M.nb, n3 = 3 * nb;
  ensure(nb, 0);
  if (sz_) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(div_up(n3, BLOCK)), dim3(BLOCK), 0, s_, n3, r, r64_.p);
    launch_precond_init(M, part_.p);
    hipLaunchKernelGGL(k_unpack4, dim3(div_up(nb, BLOCK)), dim3(BLOCK), 0, s_, nb, (const float4 *)z_.p, z);
  } else {
    hipLaunchKernelGGL(k_bj_apply, dim3(div_up(nb, BLOCK)), dim3(BLOCK), 0, s_, nb, (const float *)minv_.p, r, z);
  }
  NGUO_HIP(hipGetLastError());
}

PcgStats Pcg::solve(const SellMatrix &M, const ContactView &C, const float *b, float tol, uint32_t max_iter, float *x_out) {
  PcgStats st;
  const uint32_t nb = M.nb;
  ensure(nb, C.ns);
  const SellView A = M.view();
  const uint32_t n3 = 3 * nb;
  const unsigned G = div_up(nb, BLOCK), G3 = div_up(n3, BLOCK), GS = div_up(C.ns, BLOCK), NS = div_up(nb, 64);
  build_precond(M, C);
  // ||b||^2 through the FP64 residual at x = 0
  hipLaunchKernelGGL(k_fill64, dim3(G3), dim3(BLOCK), 0, s_, n3, x64_.p, 0.0);
  if (C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s_, C, (const double *)x64_.p, sums64_.p, nullptr);
  hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double *)sums64_.p, b, (const double *)x64_.p, r64_.p, part_.p);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s_, (const double *)part_.p, G, scal_.p + 7);
  NGUO_HIP(hipMemcpyAsync(h_scal_ + 7, scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s_));
  NGUO_HIP(hipStreamSynchronize(s_));
  const double bb = h_scal_[7];
  double rr = bb;
  if (!(bb > 0.0)) {
    hipLaunchKernelGGL(k_to_f32, dim3(G3), dim3(BLOCK), 0, s_, n3, (const double *)x64_.p, x_out);
    NGUO_HIP(hipGetLastError());
    st.residual = 0.0;
    return st;
  }
  const uint32_t nz = sz_ ? sz_->nd : G;  // preconditioner partials: per domain or per row block
  const unsigned NI = iter_grid(A, C);     // iteration-launch partials: slices, then stencil groups
  double *part_spmv = part_.p, *part_pc = part_.p + std::max(NI, G);
  const float4 *contrib = C.ns ? (const float4 *)C.contrib : nullptr;
  for (uint32_t sweep = 0; sweep < 12; ++sweep) {
    double res = std::sqrt(rr / bb);
    if (res <= (double)tol || st.iters >= max_iter) break;
    double eta = std::min(0.5, std::max(0.25 * (double)tol / res, REFINE_ETA_MIN));
    launch_precond_init(M, part_pc);
    hipLaunchKernelGGL(k_init_scal, dim3(1), dim3(BLOCK), 0, s_, scal_.p, flag_.p, (const double *)part_pc, nz, eta * eta, max_iter - st.iters);
    // iteration j: p_j in p_ (even j) or p2_ (odd j); FIRST reads p_0 = z_0 from p_
    const SpmvArgs first{A, C, (const float4 *)z_.p, (const float4 *)p_.p, nullptr, (float4 *)q_.p, part_spmv, part_pc, nz, scal_.p, flag_.p};
    SpmvArgs next_odd = first, next_even = first;
    next_odd.pold = (const float4 *)p_.p; next_odd.pnew = (float4 *)p2_.p;
    next_even.pold = (const float4 *)p2_.p; next_even.pnew = (float4 *)p_.p;
    if (tuned_nb_ != nb) {
      // autotune on the live system: the first iteration's launch (idempotent; iteration 0
      // recomputes its q and partials below)
      spmv_group_ = pick_group(s_, [&](uint32_t G) { launch_iter<SPMV_FIRST>(G, s_, first, 0); });
      tuned_nb_ = nb;
    }
    // Chunks of chunk_len iterations are queued one ahead of the one whose state the host
    // reads, so the device never idles on the convergence check; iterations queued after
    // convergence exit at their first instruction.
    uint32_t j = 0, queued = 0, waited = 0;
    auto queue_chunk = [&]() {
      const uint32_t sl = queued & 1u;
      for (uint32_t c = 0, n = chunk_len(chunk, j); c < n; ++c, ++j) {
        float *pj = (j & 1) ? p2_.p : p_.p;
        // the chunk's last iteration launch publishes the flag to host slot sl
        SpmvArgs a = j == 0 ? first : ((j & 1) ? next_odd : next_even);
        a.hflag = c + 1 == n ? d_flag_ + 4 * sl : nullptr;
        if (j == 0) launch_iter<SPMV_FIRST>(spmv_group_, s_, a, j);
        else launch_iter<SPMV_NEXT>(spmv_group_, s_, a, j);
        if (sz_) {
          launch_schwarz_pcg(s_, *sz_, false, nullptr, r_.p, z_.p, d_.p, pj, q_.p, contrib, part_spmv, NI, part_pc, scal_.p, flag_.p, j);
          launch_schwarz_coarse(s_, *sz_, false, r_.p, z_.p, pj, part_pc, flag_.p);
        } else hipLaunchKernelGGL(k_jacobi_pcg<false>, dim3(G), dim3(BLOCK), 0, s_, nb, (const float *)minv_.p, (const double *)nullptr, r_.p,
                                z_.p, d_.p, pj, (const float *)q_.p, C.inc_ptr, contrib, (const double *)part_spmv, NI, part_pc, scal_.p,
                                flag_.p, j);
      }
      NGUO_HIP(hipGetLastError());
      NGUO_HIP(hipEventRecord(ev_[sl], s_));
      ++queued;
    };
    const uint32_t *fl = nullptr;
    queue_chunk();
    const uint32_t sweep_cap = max_iter - st.iters;
    while (true) {
      queue_chunk();
      const uint32_t sl = waited & 1u;
      NGUO_HIP(hipEventSynchronize(ev_[sl]));
      ++waited;
      if (h_flag_[4 * sl]) { fl = h_flag_ + 4 * sl; break; }
      // the cap sets a state within the chunk that reaches it: a flag still clear two chunks
      // later means the host slot is not being written
      // (in 64 bits: a cap near UINT32_MAX, "no cap", must not wrap the bound)
      NGUO_REQUIRE((uint64_t)j <= (uint64_t)sweep_cap + 2ull * CHUNK_MAX + 2ull, NGUO_E_DEVICE,
                   "PCG: the convergence flag did not reach the host");
    }
    st.iters += fl[1];
    st.sweeps += 1;
    if (fl[0] == 2) st.breakdown = true;
    hipLaunchKernelGGL(k_accum, dim3(G3), dim3(BLOCK), 0, s_, n3, x64_.p, (const float *)d_.p);
    if (C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s_, C, (const double *)x64_.p, sums64_.p, nullptr);
    hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double *)sums64_.p, b, (const double *)x64_.p, r64_.p, part_.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s_, (const double *)part_.p, G, scal_.p + 7);
    NGUO_HIP(hipMemcpyAsync(h_scal_ + 7, scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s_));
    NGUO_HIP(hipStreamSynchronize(s_));
    rr = h_scal_[7];
    if (st.breakdown && fl[1] == 0) break;
  }
  hipLaunchKernelGGL(k_to_f32, dim3(G3), dim3(BLOCK), 0, s_, n3, (const double *)x64_.p, x_out);
  NGUO_HIP(hipGetLastError());
  st.residual = std::sqrt(rr / bb);
  return st;
}

void Pcg::solve_batch(hipStream_t s, PcgBatchItem *items, uint32_t n) {
  NGUO_REQUIRE(n <= MAX_BATCH, NGUO_E_INVALID, "batched PCG: too many systems");
  for (uint32_t i = 0; i < n; ++i)
    NGUO_REQUIRE(items[i].pcg->batchable(), NGUO_E_UNSUPPORTED, "batched PCG: every system needs one-level Schwarz");
  // every Pcg works on the batch stream for the duration
  std::vector<hipStream_t> own(n);
  for (uint32_t i = 0; i < n; ++i) { own[i] = items[i].pcg->s_; items[i].pcg->s_ = s; }
  struct Restore {
    PcgBatchItem *it; std::vector<hipStream_t> &own; uint32_t n;
    ~Restore() { for (uint32_t i = 0; i < n; ++i) it[i].pcg->s_ = own[i]; }
  } restore{items, own, n};
  std::vector<double> bb(n), rr(n);
  std::vector<uint8_t> done(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    PcgBatchItem &it = items[i];
    Pcg &P = *it.pcg;
    it.stats = PcgStats{};
    const uint32_t nb = it.M->nb, n3 = 3 * nb;
    P.ensure(nb, it.C.ns);
    const SellView A = it.M->view();
    const unsigned G = div_up(nb, BLOCK), G3 = div_up(n3, BLOCK), GS = div_up(it.C.ns, BLOCK);
    P.build_precond(*it.M, it.C);
    hipLaunchKernelGGL(k_fill64, dim3(G3), dim3(BLOCK), 0, s, n3, P.x64_.p, 0.0);
    if (it.C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s, it.C, (const double *)P.x64_.p, P.sums64_.p, nullptr);
    hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s, A, it.C, (const double *)P.sums64_.p, it.b, (const double *)P.x64_.p, P.r64_.p,
                       P.part_.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s, (const double *)P.part_.p, G, P.scal_.p + 7);
    NGUO_HIP(hipMemcpyAsync(P.h_scal_ + 7, P.scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  NGUO_HIP(hipStreamSynchronize(s));
  for (uint32_t i = 0; i < n; ++i) {
    bb[i] = rr[i] = items[i].pcg->h_scal_[7];
    if (!(bb[i] > 0.0)) done[i] = 1;
  }
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto &e : ev) NGUO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct EvFree {
    hipEvent_t *e;
    ~EvFree() { for (int k = 0; k < 2; ++k) if (e[k]) (void)hipEventDestroy(e[k]); }
  } evfree{ev};
  const uint32_t chunk = items[0].pcg->chunk;
  for (uint32_t sweep = 0; sweep < 12; ++sweep) {
    // systems that need this sweep
    std::vector<uint32_t> act;
    for (uint32_t i = 0; i < n; ++i) {
      if (done[i]) continue;
      const PcgBatchItem &it = items[i];
      const double res = std::sqrt(rr[i] / bb[i]);
      if (res <= (double)it.tol || it.stats.iters >= it.max_iter) { done[i] = 1; continue; }
      act.push_back(i);
    }
    if (act.empty()) break;
    BatchOf<SpmvArgs> sf{}, sn0{}, sn1{};
    BatchOf<SchArgs> c0{}, c1{};
    const uint32_t k = (uint32_t)act.size();
    sf.k = sn0.k = sn1.k = c0.k = c1.k = k;
    for (uint32_t a = 0; a < k; ++a) {
      PcgBatchItem &it = items[act[a]];
      Pcg &P = *it.pcg;
      const uint32_t nb = it.M->nb, nd = P.sz_->nd;
      const SellView A = it.M->view();
      const unsigned NI = iter_grid(A, it.C);
      const double res = std::sqrt(rr[act[a]] / bb[act[a]]);
      const double eta = std::min(0.5, std::max(0.25 * (double)it.tol / res, REFINE_ETA_MIN));
      double *part_spmv = P.part_.p, *part_pc = P.part_.p + std::max<uint32_t>(NI, div_up(nb, BLOCK));
      P.launch_precond_init(*it.M, part_pc);
      hipLaunchKernelGGL(k_init_scal, dim3(1), dim3(BLOCK), 0, s, P.scal_.p, P.flag_.p, (const double *)part_pc, nd, eta * eta,
                         it.max_iter - it.stats.iters);
      const float4 *z4 = (const float4 *)P.z_.p;
      float4 *p4 = (float4 *)P.p_.p, *p24 = (float4 *)P.p2_.p;
      // FIRST (j = 0): p_0 in p_; NEXT with j odd: pold = p_, pnew = p2_; j even: the reverse
      sf.a[a] = SpmvArgs{A, it.C, z4, p4, nullptr, (float4 *)P.q_.p, part_spmv, part_pc, nd, P.scal_.p, P.flag_.p};
      sn1.a[a] = sf.a[a];
      sn1.a[a].pnew = p24;
      sn0.a[a] = sf.a[a];
      sn0.a[a].pold = p24;
      sn0.a[a].pnew = p4;
      c0.a[a] = SchArgs{nd, P.sz_->dom_ptr.p, P.sz_->dom_dof.p, P.sz_->ainv.p, P.r_.p, P.z_.p, P.d_.p, P.p_.p, P.q_.p, P.sz_->inc_rng.p,
                        it.C.ns ? (const float4 *)it.C.contrib : nullptr, part_spmv, NI, part_pc, P.scal_.p, P.flag_.p};
      c1.a[a] = c0.a[a];
      c1.a[a].p = P.p2_.p;
      sf.off[a + 1] = sn0.off[a + 1] = sn1.off[a + 1] = sf.off[a] + NI;
      c0.off[a + 1] = c1.off[a + 1] = c0.off[a] + nd;
    }
    uint32_t &group = items[0].pcg->batch_group_;
    if (items[0].pcg->batch_tuned_ != sf.off[k]) {
      group = pick_group(s, [&](uint32_t G) { launch_iter_batch<SPMV_FIRST>(G, s, sf, 0); });
      items[0].pcg->batch_tuned_ = sf.off[k];
    }
    uint32_t j = 0, queued = 0, waited = 0;
    auto queue_chunk = [&]() {
      for (uint32_t c = 0, n = chunk_len(chunk, j); c < n; ++c, ++j) {
        if (j == 0) launch_iter_batch<SPMV_FIRST>(group, s, sf, j);
        else launch_iter_batch<SPMV_NEXT>(group, s, (j & 1) ? sn1 : sn0, j);
        launch_schwarz_pcg_batch(s, (j & 1) ? c1 : c0, j);
      }
      NGUO_HIP(hipGetLastError());
      const uint32_t sl = queued & 1u;
      for (uint32_t a = 0; a < k; ++a) {
        Pcg &P = *items[act[a]].pcg;
        NGUO_HIP(hipMemcpyAsync(P.h_flag_ + 4 * sl, P.flag_.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      }
      NGUO_HIP(hipEventRecord(ev[sl], s));
      ++queued;
    };
    uint32_t sl_done = 0;
    queue_chunk();
    while (true) {
      queue_chunk();
      const uint32_t sl = waited & 1u;
      NGUO_HIP(hipEventSynchronize(ev[sl]));
      ++waited;
      bool all = true;
      for (uint32_t a = 0; a < k && all; ++a) all = items[act[a]].pcg->h_flag_[4 * sl] != 0;
      if (all) { sl_done = sl; break; }
    }
    for (uint32_t a = 0; a < k; ++a) {
      PcgBatchItem &it = items[act[a]];
      Pcg &P = *it.pcg;
      const uint32_t *fl = P.h_flag_ + 4 * sl_done;
      const uint32_t nb = it.M->nb, n3 = 3 * nb;
      const unsigned G = div_up(nb, BLOCK), G3 = div_up(n3, BLOCK), GS = div_up(it.C.ns, BLOCK);
      it.stats.iters += fl[1];
      it.stats.sweeps += 1;
      if (fl[0] == 2) it.stats.breakdown = true;
      if (it.stats.breakdown && fl[1] == 0) done[act[a]] = 1;
      hipLaunchKernelGGL(k_accum, dim3(G3), dim3(BLOCK), 0, s, n3, P.x64_.p, (const float *)P.d_.p);
      if (it.C.ns)
        hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s, it.C, (const double *)P.x64_.p, P.sums64_.p, nullptr);
      hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s, it.M->view(), it.C, (const double *)P.sums64_.p, it.b, (const double *)P.x64_.p,
                         P.r64_.p, P.part_.p);
      hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s, (const double *)P.part_.p, G, P.scal_.p + 7);
      NGUO_HIP(hipMemcpyAsync(P.h_scal_ + 7, P.scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    NGUO_HIP(hipGetLastError());
    NGUO_HIP(hipStreamSynchronize(s));
    for (uint32_t a = 0; a < k; ++a) rr[act[a]] = items[act[a]].pcg->h_scal_[7];
  }
  for (uint32_t i = 0; i < n; ++i) {
    PcgBatchItem &it = items[i];
    const uint32_t n3 = 3 * it.M->nb;
    hipLaunchKernelGGL(k_to_f32, dim3(div_up(n3, BLOCK)), dim3(BLOCK), 0, s, n3, (const double *)it.pcg->x64_.p, it.x_out);
    it.stats.residual = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : 0.0;
  }
  NGUO_HIP(hipGetLastError());
}

void Pcg::apply(const SellView &A, const ContactView &C, const float *x, float *y) {
  ensure(A.nb, C.ns);
  if (!A.nb) return;
  const unsigned G = div_up(A.nb, BLOCK);
  hipLaunchKernelGGL(k_pack4, dim3(G), dim3(BLOCK), 0, s_, A.nb, x, (float4 *)p2_.p);
  launch_plain(spmv_group_, s_, A, C, (const float4 *)p2_.p, (float4 *)q_.p);
  hipLaunchKernelGGL(k_unpack4, dim3(G), dim3(BLOCK), 0, s_, A.nb, (const float4 *)q_.p, y);
  NGUO_HIP(hipGetLastError());
}

// State under which the iteration launch runs without stopping: running, no cap, no
// convergence threshold, rz_{j-1} = 1 and zero preconditioner partials (beta 0).
__global__ void k_bench_state(uint32_t *flag, double *scal, double *cur, uint32_t ncur) {
  for (uint32_t i = threadIdx.x; i < ncur; i += blockDim.x) cur[i] = 0.0;
  if (threadIdx.x == 0) {
    flag[0] = 0u; flag[1] = 0u; flag[2] = 0xffffffffu;
    scal[5] = -1.0;
    scal[8] = 1.0; scal[9] = 1.0;
  }
}

double Pcg::time_spmv(const SellView &A, const ContactView &C, const float *x, uint32_t reps, bool iteration_form) {
  ensure(A.nb, C.ns);
  if (!A.nb || !reps) return 0.0;
  const unsigned NS = div_up(A.nb, 64), G = div_up(A.nb, BLOCK), NI = iter_grid(A, C);
  const uint32_t nz = sz_ ? sz_->nd : G;
  double *part_spmv = part_.p, *part_pc = part_.p + std::max(NI, G);
  hipLaunchKernelGGL(k_pack4, dim3(G), dim3(BLOCK), 0, s_, A.nb, x, (float4 *)p2_.p);
  hipLaunchKernelGGL(k_pack4, dim3(G), dim3(BLOCK), 0, s_, A.nb, x, (float4 *)z_.p);
  if (iteration_form) hipLaunchKernelGGL(k_bench_state, dim3(1), dim3(BLOCK), 0, s_, flag_.p, scal_.p, part_pc, 2 * nz);
  // plain form: the contact blocks of x once, outside the timed launches
  if (!iteration_form && C.ns) hipLaunchKernelGGL(k_pcg_stencils_plain, dim3(stencil_groups(C)), dim3(WAVE), 0, s_, C, (const float4 *)p2_.p);
  const SpmvArgs it{A, C, (const float4 *)z_.p, (const float4 *)p2_.p, (float4 *)p_.p, (float4 *)q_.p, part_spmv, part_pc, nz, scal_.p, flag_.p};
  auto launch_g = [&](uint32_t Gs) {
    if (iteration_form) {
      launch_iter<SPMV_NEXT>(Gs, s_, it, 1);
    } else if (Gs == 8) {
      hipLaunchKernelGGL((k_pcg_spmv_plain<8>), dim3(NS), dim3(WAVE), 0, s_, A, C, (const float4 *)p2_.p, (float4 *)q_.p);
    } else {
      hipLaunchKernelGGL((k_pcg_spmv_plain<16>), dim3(NS), dim3(WAVE), 0, s_, A, C, (const float4 *)p2_.p, (float4 *)q_.p);
    }
  };
  if (tuned_nb_ != A.nb) {  // a system no solve has tuned yet
    spmv_group_ = pick_group(s_, launch_g);
    tuned_nb_ = A.nb;
  }
  return time_launches(s_, reps, [&] { launch_g(spmv_group_); });
}

// State under which the preconditioner's iteration launch runs without stopping: running,
// rz_j = 1 and pAp partials of 1 (alpha = 1 / npap).
__global__ void k_bench_precond_state(uint32_t *flag, double *scal, double *pap, uint32_t npap) {
  for (uint32_t i = threadIdx.x; i < npap; i += blockDim.x) pap[i] = 1.0;
  if (threadIdx.x == 0) {
    flag[0] = 0u; flag[1] = 0u; flag[2] = 0xffffffffu;
    scal[8] = 1.0; scal[9] = 1.0;
  }
}

double Pcg::time_precond(const SellMatrix &M, const ContactView &C, uint32_t reps, double *bytes) {
  *bytes = 0.0;
  if (!sz_ || !M.nb || !reps) return 0.0;
  ensure(M.nb, C.ns);
  build_precond(M, C);
  if (!sz_->nd) return 0.0;
  const SellView A = M.view();
  const unsigned G = div_up(M.nb, BLOCK), NI = iter_grid(A, C);
  double *part_spmv = part_.p, *part_pc = part_.p + std::max(NI, G);
  hipLaunchKernelGGL(k_bench_precond_state, dim3(1), dim3(BLOCK), 0, s_, flag_.p, scal_.p, part_spmv, NI);
  const float4 *contrib = C.ns ? (const float4 *)C.contrib : nullptr;
  const double ms = time_launches(s_, reps, [&] {
    launch_schwarz_pcg(s_, *sz_, false, nullptr, r_.p, z_.p, d_.p, p_.p, q_.p, contrib, part_spmv, NI, part_pc, scal_.p, flag_.p, 1);
  });
  double b = 0.0;
  for (uint32_t d = 0; d < sz_->nd; ++d) {
    const double m = 3.0 * (sz_->h_dom_ptr[d + 1] - sz_->h_dom_ptr[d]);
    b += 2.0 * m * (m + 1.0);  // upper triangle of the FP32 inverse
  }
  b += 32.0 * 3.0 * M.nb;  // r, d (read + write), q, p (read), z (write), domain rows
  if (C.ns) {
    uint32_t ninc = 0;
    NGUO_HIP(hipMemcpyAsync(&ninc, C.inc_ptr + M.nb, sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
    NGUO_HIP(hipStreamSynchronize(s_));
    b += 8.0 * M.nb + 12.0 * ninc;  // incidence ranges, the blocks' three components
  }
  *bytes = b;
  return ms;
}

// ---------------------------------------------------------------- FP64 PCG
// Vectors r, z, p, q of stride 4 doubles per block row (x, y, z, 0; one 32-byte load per
// gathered neighbour), the solution accumulated in the FP64 iterate of stride 3. scal: 1 pAp,
// 2 alpha, 4 rr, 5 threshold (eta^2 rr0), 8 / 9 rz by iteration parity. flag: 0 state
// (0 running, 1 converged, 2 breakdown, 3 cap), 1 iterations done, 2 cap.

// Stencil s (one thread each): the sums g . p and sum_a w_a p_a over its free participants,
// then each free participant's block of H_s p, hp (g . p) g_a + lambda w_a (s_w - n (n . s_w)),
// stored at its incidence slot.
__global__ __launch_bounds__(BLOCK) void k64_stencils(ContactView C, const double4 *__restrict__ p, double4 *__restrict__ contrib,
                                                     const uint32_t *flag) {
  if (flag[0]) return;
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= C.ns) return;
  uint32_t db[4], kp[4];
  double g[12], w[4], x[4][3];
  for (int b = 0; b < 4; ++b) {
    db[b] = C.dof[4 * (size_t)s + b];
    kp[b] = C.pos[4 * (size_t)s + b];
    w[b] = C.w[4 * (size_t)s + b];
  }
  for (int k = 0; k < 12; ++k) g[k] = C.g[12 * (size_t)s + k];
  const double n0 = C.n[3 * (size_t)s], n1 = C.n[3 * (size_t)s + 1], n2 = C.n[3 * (size_t)s + 2];
  const double hp = C.coef[2 * (size_t)s], lam = C.coef[2 * (size_t)s + 1];
  double s1 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int b = 0; b < 4; ++b) {
    x[b][0] = x[b][1] = x[b][2] = 0.0;
    if (db[b] == NONE) continue;
    const double4 v = p[db[b]];
    x[b][0] = v.x; x[b][1] = v.y; x[b][2] = v.z;
    s1 += g[3 * b] * v.x + g[3 * b + 1] * v.y + g[3 * b + 2] * v.z;
    a0 += w[b] * v.x; a1 += w[b] * v.y; a2 += w[b] * v.z;
  }
  const double t = hp * s1, nd = n0 * a0 + n1 * a1 + n2 * a2;
  for (int a = 0; a < 4; ++a) {
    if (kp[a] == NONE) continue;
    double c0 = t * g[3 * a], c1 = t * g[3 * a + 1], c2 = t * g[3 * a + 2];
    if (lam != 0.0) {
      const double pw = lam * w[a];
      c0 += pw * (a0 - n0 * nd); c1 += pw * (a1 - n1 * nd); c2 += pw * (a2 - n2 * nd);
    }
    contrib[kp[a]] = make_double4(c0, c1, c2, 0.0);
  }
}

// q = (A + C) p, one block row per thread (FP32 block values, FP64 sums, the row's contact
// blocks added in incidence order); part[workgroup] = p . q over its rows.
__global__ __launch_bounds__(BLOCK) void k64_spmv(SellView A, ContactView C, const double4 *__restrict__ p, const double4 *__restrict__ contrib,
                                                 double4 *__restrict__ q, double *part, const uint32_t *flag) {
  __shared__ double lds[4];
  if (flag[0]) return;
  const uint32_t row = blockIdx.x * BLOCK + threadIdx.x;
  double dp = 0.0;
  if (row < A.nb) {
    const uint32_t base = A.ell_w ? 64u * A.ell_w * (row >> 6) : A.slice_ptr[row >> 6], lane = row & 63u, len = A.row_len[row];
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
    for (uint32_t k = 0; k < len; ++k) {
      const uint32_t c = A.col[base + 64u * k + lane];
      const BlockVals v = load_block(A.val, base + 64u * k, lane);
      const double4 x = p[c];
      y0 += (double)v.a.x * x.x + (double)v.a.y * x.y + (double)v.a.z * x.z;
      y1 += (double)v.a.w * x.x + (double)v.b.x * x.y + (double)v.b.y * x.z;
      y2 += (double)v.b.z * x.x + (double)v.b.w * x.y + (double)v.c * x.z;
    }
    if (C.ns)
      for (uint32_t k = C.inc_ptr[row]; k < C.inc_ptr[row + 1]; ++k) {
        const double4 cv = contrib[k];
        y0 += cv.x; y1 += cv.y; y2 += cv.z;
      }
    q[row] = make_double4(y0, y1, y2, 0.0);
    const double4 pr = p[row];
    dp = pr.x * y0 + pr.y * y1 + pr.z * y2;
  }
  const double t = block_sum(dp, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Schwarz launch of the FP64 PCG, one workgroup per domain. INIT: r = r64 (stride 3), z =
// A_d^-1 r, p = z. Otherwise: alpha = rz / pAp (npq SpMV partials, reduced by every workgroup
// in the same order; breakdown stops the solve), r -= alpha q, x += alpha p, z = A_d^-1 r.
// The inverse is read from the 21 upper 16 x 16 FP64 tiles: row t's entry k comes from tile
// (t / 16, k / 16) directly, or transposed from tile (k / 16, t / 16) below the diagonal.
// Partials r.z -> nxt[d], r.r -> nxt[nd + d].
template <bool INIT>
__global__ __launch_bounds__(BLOCK) void k64_schwarz(uint32_t nd, const uint32_t *__restrict__ dom_ptr, const uint32_t *__restrict__ dom_dof,
                                                    const double *__restrict__ ainv, const double *__restrict__ r64, double *__restrict__ r,
                                                    double *__restrict__ z, double *__restrict__ p, const double *__restrict__ q,
                                                    double *__restrict__ x64, const double *pq_part, uint32_t npq, double *nxt, double *scal,
                                                    uint32_t *flag, uint32_t j) {
  __shared__ double rs[SCHWARZ_MAX_DOF];
  __shared__ double lds[2 * BLOCK / WAVE];
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  if (!INIT && flag[0]) return;
  double alpha = 0.0;
  if (!INIT) {
    const double *src[1] = {pq_part};
    const uint32_t n[1] = {npq};
    double v[1];
    sum_parts_all<BLOCK, 1>(src, n, v, lds);
    if (!(v[0] > 0.0)) {
      if (d == 0 && t == 0) flag[0] = 2;
      return;
    }
    alpha = scal[8 + (j & 1u)] / v[0];
    if (d == 0 && t == 0) { scal[1] = v[0]; scal[2] = alpha; }
  }
  const uint32_t m = 3 * (dom_ptr[d + 1] - dom_ptr[d]);
  const uint32_t od = t < m ? dom_dof[(size_t)d * SCHWARZ_MAX_DOF + t] : NONE;
  const size_t i3 = od != NONE ? 3 * (size_t)(od >> 2) + (od & 3u) : 0;
  double rv = 0.0;
  if (t < m) {
    if (INIT) {
      rv = r64[i3];
    } else {
      rv = r[od] - alpha * q[od];
      x64[i3] += alpha * p[od];
    }
    r[od] = rv;
  }
  if (t < SCHWARZ_MAX_DOF) rs[t] = rv;
  __syncthreads();
  double v[2] = {0.0, 0.0};
  if (t < m) {
    const double *a = ainv + (size_t)d * SCHWARZ_INV_FLOATS;
    const uint32_t P = t / 16, i = t % 16;
    double zv = 0.0;
    for (uint32_t k = 0; k < m; ++k) {
      const uint32_t Q = k / 16, jj = k % 16;
      zv += (P <= Q ? a[SCH_TILE64 * sym_tile64(P, Q) + 16 * i + jj] : a[SCH_TILE64 * sym_tile64(Q, P) + 16 * jj + i]) * rs[k];
    }
    z[od] = zv;
    if (INIT) p[od] = zv;
    v[0] = rv * zv;
    v[1] = rv * rv;
  }
  wg_sum_all<BLOCK, 2>(v, lds);
  if (t == 0) { nxt[d] = v[0]; nxt[nd + d] = v[1]; }
}

// Opens the solve: rz0 and rr0 from the Schwarz partials, threshold eta^2 rr0, state.
__global__ __launch_bounds__(BLOCK) void k64_init(const double *nxt, uint32_t nd, double eta2, uint32_t cap, double *scal, uint32_t *flag) {
  __shared__ double lds[2 * BLOCK / WAVE];
  const double *src[2] = {nxt, nxt + nd};
  const uint32_t n[2] = {nd, nd};
  double v[2];
  sum_parts_all<BLOCK, 2>(src, n, v, lds);
  if (threadIdx.x == 0) {
    scal[8] = v[0]; scal[4] = v[1]; scal[5] = eta2 * v[1];
    flag[0] = v[1] == 0.0 ? 1u : (cap == 0 ? 3u : 0u);
    flag[1] = 0;
    flag[2] = cap;
  }
}

// Closes iteration j: rz_{j+1}, rr_{j+1} from the Schwarz partials (every workgroup, same
// order); converged (rr <= threshold) or capped stops the solve, else p = z + beta p.
__global__ __launch_bounds__(BLOCK) void k64_pnext(uint32_t n4, const double *__restrict__ z, double *__restrict__ p, const double *nxt,
                                                  uint32_t nd, double *scal, uint32_t *flag, uint32_t j) {
  __shared__ double lds[2 * BLOCK / WAVE];
  if (flag[0]) return;
  const double *src[2] = {nxt, nxt + nd};
  const uint32_t n[2] = {nd, nd};
  double v[2];
  sum_parts_all<BLOCK, 2>(src, n, v, lds);
  const bool conv = v[1] <= scal[5], capped = j + 1 >= flag[2];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    flag[1] = j + 1;
    scal[4] = v[1];
    scal[8 + ((j + 1) & 1u)] = v[0];
    if (conv) flag[0] = 1;
    else if (capped) flag[0] = 3;
  }
  if (conv || capped) return;
  const double beta = v[0] / scal[8 + (j & 1u)];
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n4) p[i] = z[i] + beta * p[i];
}

Pcg64::~Pcg64() {
  for (auto &e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (h_flag_) (void)hipHostFree(h_flag_);
  if (h_scal_) (void)hipHostFree(h_scal_);
}

void Pcg64::ensure(uint32_t nb, uint32_t ns, uint32_t nd) {
  if (!h_flag_) {
    NGUO_HIP(hipHostMalloc((void **)&h_flag_, 8 * sizeof(uint32_t), hipHostMallocDefault));
    NGUO_HIP(hipHostMalloc((void **)&h_scal_, 8 * sizeof(double), hipHostMallocDefault));
    for (auto &e : ev_) NGUO_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  if (nb > cap_nb_) {
    const size_t n4 = 4 * (size_t)nb, n3 = 3 * (size_t)nb;
    r_.zero(n4, s_); z_.zero(n4, s_); p_.zero(n4, s_); q_.zero(n4, s_);  // fourth components stay zero
    x64_.reserve(n3); r64_.reserve(n3);
    cap_nb_ = nb;
  }
  part_.reserve(std::max<size_t>(div_up(nb, BLOCK), 1) + 8);
  if (nd > cap_nd_) { nxt_.reserve(2 * (size_t)nd + 2); cap_nd_ = nd; }
  if (ns > cap_ns_) {
    const uint32_t c = std::max(ns, 1u << 14);
    contrib_.reserve(16 * (size_t)c);
    sums64_.reserve(4 * (size_t)c);
    cap_ns_ = c;
  }
  scal_.reserve(16);
  flag_.reserve(4);
}

PcgStats Pcg64::solve(const SellMatrix &M, const ContactView &C, Schwarz &sz, const float *b, float tol, uint32_t max_iter, float *x_out) {
  PcgStats st;
  const uint32_t nb = M.nb, n3 = 3 * nb, n4 = 4 * nb, nd = sz.nd;
  NGUO_REQUIRE(nd > 0 && sz.nd1 == 0, NGUO_E_UNSUPPORTED, "FP64 PCG: needs the one-level Schwarz preconditioner");
  ensure(nb, C.ns, nd);
  const SellView A = M.view();
  const unsigned G = div_up(nb, BLOCK), G3 = div_up(n3, BLOCK), G4 = div_up(n4, BLOCK), GS = div_up(C.ns, BLOCK);
  build_schwarz64(s_, M, C, sz);
  // ||b||^2 through the FP64 residual at x = 0
  hipLaunchKernelGGL(k_fill64, dim3(G3), dim3(BLOCK), 0, s_, n3, x64_.p, 0.0);
  if (C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s_, C, (const double *)x64_.p, sums64_.p, nullptr);
  hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double *)sums64_.p, b, (const double *)x64_.p, r64_.p, part_.p);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s_, (const double *)part_.p, G, scal_.p + 7);
  NGUO_HIP(hipMemcpyAsync(h_scal_, scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s_));
  NGUO_HIP(hipStreamSynchronize(s_));
  const double bb = h_scal_[0];
  double rr = bb;
  double4 *r4 = (double4 *)r_.p, *p4 = (double4 *)p_.p, *q4 = (double4 *)q_.p;
  (void)r4;
  if (bb > 0.0) {
    for (uint32_t sweep = 0; sweep < 8; ++sweep) {
      const double res = std::sqrt(rr / bb);
      if (res <= (double)tol || st.iters >= max_iter) break;
      // one sweep reaches the tolerance unless the recursive residual drifts from the true one
      const double eta = std::min(0.5, 0.5 * (double)tol / res);
      hipLaunchKernelGGL(k64_schwarz<true>, dim3(nd), dim3(BLOCK), 0, s_, nd, (const uint32_t *)sz.dom_ptr.p, (const uint32_t *)sz.dom_dof.p,
                         (const double *)sz.ainv64.p, (const double *)r64_.p, r_.p, z_.p, p_.p, (const double *)q_.p, x64_.p,
                         (const double *)part_.p, G, nxt_.p, scal_.p, flag_.p, 0u);
      hipLaunchKernelGGL(k64_init, dim3(1), dim3(BLOCK), 0, s_, (const double *)nxt_.p, nd, eta * eta, max_iter - st.iters, scal_.p, flag_.p);
      uint32_t j = 0, queued = 0, waited = 0, done_it = 0, state = 0;
      auto queue_chunk = [&]() {
        const uint32_t sl = queued & 1u;
        for (uint32_t c = 0; c < chunk; ++c, ++j) {
          if (C.ns) hipLaunchKernelGGL(k64_stencils, dim3(GS), dim3(BLOCK), 0, s_, C, (const double4 *)p4, (double4 *)contrib_.p, (const uint32_t *)flag_.p);
          hipLaunchKernelGGL(k64_spmv, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double4 *)p4, (const double4 *)contrib_.p, q4, part_.p,
                             (const uint32_t *)flag_.p);
          hipLaunchKernelGGL(k64_schwarz<false>, dim3(nd), dim3(BLOCK), 0, s_, nd, (const uint32_t *)sz.dom_ptr.p, (const uint32_t *)sz.dom_dof.p,
                             (const double *)sz.ainv64.p, (const double *)nullptr, r_.p, z_.p, p_.p, (const double *)q_.p, x64_.p,
                             (const double *)part_.p, G, nxt_.p, scal_.p, flag_.p, j);
          hipLaunchKernelGGL(k64_pnext, dim3(G4), dim3(BLOCK), 0, s_, n4, (const double *)z_.p, p_.p, (const double *)nxt_.p, nd, scal_.p, flag_.p, j);
        }
        NGUO_HIP(hipGetLastError());
        NGUO_HIP(hipMemcpyAsync(h_flag_ + 4 * sl, flag_.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
        NGUO_HIP(hipEventRecord(ev_[sl], s_));
        ++queued;
      };
      queue_chunk();
      const uint64_t sweep_cap = (uint64_t)max_iter - st.iters;
      while (true) {
        queue_chunk();
        const uint32_t sl = waited & 1u;
        NGUO_HIP(hipEventSynchronize(ev_[sl]));
        ++waited;
        state = h_flag_[4 * sl];
        done_it = h_flag_[4 * sl + 1];
        if (state) break;
        NGUO_REQUIRE((uint64_t)j <= sweep_cap + 2ull * chunk + 2ull, NGUO_E_DEVICE, "FP64 PCG: the convergence state did not reach the host");
      }
      NGUO_HIP(hipStreamSynchronize(s_));  // the chunk queued past convergence exits at once
      st.iters += done_it;
      st.sweeps += 1;
      if (state == 2) st.breakdown = true;
      if (C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s_, C, (const double *)x64_.p, sums64_.p, nullptr);
      hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double *)sums64_.p, b, (const double *)x64_.p, r64_.p, part_.p);
      hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s_, (const double *)part_.p, G, scal_.p + 7);
      NGUO_HIP(hipMemcpyAsync(h_scal_, scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s_));
      NGUO_HIP(hipStreamSynchronize(s_));
      rr = h_scal_[0];
      if (st.breakdown && done_it == 0) break;
    }
  }
  hipLaunchKernelGGL(k_to_f32, dim3(G3), dim3(BLOCK), 0, s_, n3, (const double *)x64_.p, x_out);
  NGUO_HIP(hipGetLastError());
  st.residual = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
  return st;
}

}  // namespace nguo
