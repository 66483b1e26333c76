// One-level additive Schwarz preconditioner on the device (see pcg.h). Reference:
// engine-core/src/linsolve.rs:126-258 SchwarzTopology::{build, factorize, apply} and the GPU
// apply it fed, engine-gpu/src/schwarz.rs + shaders/schwarz.wgsl (one thread per domain,
// serial triangular solves). Here a workgroup owns a domain: the dense block lives in LDS,
// the Cholesky and the triangular inverse run in FP64 across the workgroup, and the PCG
// applies the explicit inverse with coalesced column reads.
#include <algorithm>

#include "pcg.h"
#include "pcg_device.h"

namespace nguo {

constexpr uint32_t NONE_S = 0xffffffffu;
typedef double double4_t __attribute__((ext_vector_type(4)));
constexpr uint32_t MD = SCHWARZ_MAX_DOF;

void Schwarz::build(const SellMatrix &M, uint32_t domain_size, hipStream_t s, float coarse_weight) {
  domain = std::max(1u, std::min(domain_size, SCHWARZ_MAX_DOMAIN));
  weight = coarse_weight;
  const uint32_t nv = M.nb;
  // linsolve.rs:134-170: adjacency = off-diagonal block columns (sorted), greedy BFS from
  // seeds in index order.
  std::vector<uint8_t> assigned(nv, 0);
  h_dom_ptr.assign(1, 0);
  h_dom_row.clear();
  std::vector<uint32_t> dom;
  if (nbr_of != M.pattern_id || h_nbr_ptr.size() != (size_t)nv + 1) {
    h_nbr_ptr.assign(1, 0);
    h_nbr.clear();
    for (uint32_t v = 0; v < nv; ++v) {
      for (uint32_t k = 0; k < M.h_row_len[v]; ++k) h_nbr.push_back(M.h_col[M.slot(v, k)]);
      h_nbr_ptr.push_back((uint32_t)h_nbr.size());
    }
    nbr_of = M.pattern_id;
  }
  for (uint32_t seed = 0; seed < nv; ++seed) {
    if (assigned[seed]) continue;
    dom.assign(1, seed);
    assigned[seed] = 1;
    size_t fi = 0;
    while (dom.size() < domain && fi < dom.size()) {
      uint32_t v = dom[fi++];
      for (uint32_t k = 0; k < M.h_row_len[v]; ++k) {
        uint32_t nb = M.h_col[M.slot(v, k)];
        if (nb == v) continue;
        if (!assigned[nb] && dom.size() < domain) { assigned[nb] = 1; dom.push_back(nb); }
      }
    }
    h_dom_row.insert(h_dom_row.end(), dom.begin(), dom.end());
    h_dom_ptr.push_back((uint32_t)h_dom_row.size());
  }
  elastic_ptr = h_dom_ptr;
  elastic_row = h_dom_row;
  upload(M, s);
  contact_built = false;
}

void Schwarz::restore_elastic(const SellMatrix &M, hipStream_t s) {
  if (!contact_built) return;
  h_dom_ptr = elastic_ptr;
  h_dom_row = elastic_row;
  upload(M, s);
  contact_built = false;
}

// Domains for a Newton system with contact. A contact stencil couples its free participants
// a, b through hp g_a g_b^T + lambda w_a w_b (I - n n^T), with hp the barrier curvature, which
// near small gaps reaches 1e11 .. 1e16 against a mass term of ~0.04; a coupling between two
// domains is seen only by the Krylov iteration, and the iteration count follows the strongest
// couplings the domains cut. So the strongest couplings decide the domains: vertex pairs are
// merged by union-find in order of coupling strength hp |g_a . g_b| + 2 lambda |w_a w_b| (summed
// over the pair's stencils; ties by pair), as long as the merged cluster has at most `domain`
// vertices; then, from seeds in index order, a domain takes the seed's cluster and grows over
// contact partners (strongest first) and then the matrix pattern, adding whole clusters that
// still fit. Measured on the tube garment's Newton systems (FP64 PCG, one-level, 32-vertex
// domains): 1138 -> 103 and 170 -> 37 iterations.
void Schwarz::build_contact(const SellMatrix &M, uint32_t ns, const uint32_t *dof, const float *coef, const float *g, const float *w, hipStream_t s) {
  const uint32_t nv = M.nb;
  struct Pair {
    uint64_t key;
    double strength;
  };
  std::vector<Pair> pairs;
  pairs.reserve(12 * (size_t)ns);
  for (uint32_t st = 0; st < ns; ++st) {
    const double hp = coef[2 * (size_t)st], lam = coef[2 * (size_t)st + 1];
    for (uint32_t a = 0; a < 4; ++a) {
      const uint32_t da = dof[4 * (size_t)st + a];
      if (da == NONE_S) continue;
      for (uint32_t b = 0; b < 4; ++b) {
        const uint32_t db = dof[4 * (size_t)st + b];
        if (b == a || db == NONE_S || db == da) continue;
        const float *ga = g + 12 * (size_t)st + 3 * a, *gb = g + 12 * (size_t)st + 3 * b;
        const double gg = std::fabs((double)ga[0] * gb[0] + (double)ga[1] * gb[1] + (double)ga[2] * gb[2]);
        const double ww = std::fabs((double)w[4 * (size_t)st + a] * (double)w[4 * (size_t)st + b]);
        pairs.push_back(Pair{((uint64_t)da << 32) | db, hp * gg + 2.0 * lam * ww});
      }
    }
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) { return x.key < y.key; });
  size_t o = 0;
  for (size_t i = 0; i < pairs.size(); ++i) {
    if (o && pairs[o - 1].key == pairs[i].key) pairs[o - 1].strength += pairs[i].strength;
    else pairs[o++] = pairs[i];
  }
  pairs.resize(o);
  // contact partners of each vertex, strongest first
  std::vector<uint32_t> cptr(nv + 1, 0);
  for (const Pair &p : pairs) ++cptr[(p.key >> 32) + 1];
  for (uint32_t v = 0; v < nv; ++v) cptr[v + 1] += cptr[v];
  std::vector<std::pair<double, uint32_t>> cadj(pairs.size());
  {
    std::vector<uint32_t> fill(cptr.begin(), cptr.end() - 1);
    for (const Pair &p : pairs) cadj[fill[p.key >> 32]++] = {p.strength, (uint32_t)(p.key & 0xffffffffu)};
    for (uint32_t v = 0; v < nv; ++v)
      std::sort(cadj.begin() + cptr[v], cadj.begin() + cptr[v + 1], [](const auto &x, const auto &y) {
        return x.first > y.first || (x.first == y.first && x.second < y.second);
      });
  }
  // union-find over the pairs (each unordered pair once), strongest first
  std::vector<Pair> order;
  order.reserve(pairs.size() / 2 + 1);
  for (const Pair &p : pairs)
    if ((p.key >> 32) < (p.key & 0xffffffffu)) order.push_back(p);
  std::sort(order.begin(), order.end(), [](const Pair &x, const Pair &y) { return x.strength > y.strength || (x.strength == y.strength && x.key < y.key); });
  std::vector<uint32_t> parent(nv), size(nv, 1);
  for (uint32_t v = 0; v < nv; ++v) parent[v] = v;
  auto find = [&](uint32_t a) {
    while (parent[a] != a) { parent[a] = parent[parent[a]]; a = parent[a]; }
    return a;
  };
  for (const Pair &p : order) {
    uint32_t ra = find((uint32_t)(p.key >> 32)), rb = find((uint32_t)(p.key & 0xffffffffu));
    if (ra == rb || size[ra] + size[rb] > domain) continue;
    if (size[ra] < size[rb] || (size[ra] == size[rb] && rb < ra)) std::swap(ra, rb);
    parent[rb] = ra;
    size[ra] += size[rb];
  }
  // cluster members in index order
  std::vector<uint32_t> root(nv), mptr(nv + 1, 0), mem(nv);
  for (uint32_t v = 0; v < nv; ++v) { root[v] = find(v); ++mptr[root[v] + 1]; }
  for (uint32_t v = 0; v < nv; ++v) mptr[v + 1] += mptr[v];
  {
    std::vector<uint32_t> fill(mptr.begin(), mptr.end() - 1);
    for (uint32_t v = 0; v < nv; ++v) mem[fill[root[v]]++] = v;
  }
  std::vector<uint8_t> assigned(nv, 0);
  h_dom_ptr.assign(1, 0);
  h_dom_row.clear();
  std::vector<uint32_t> dom;
  auto take = [&](uint32_t v) {
    const uint32_t r = root[v];
    if (assigned[v] || dom.size() + (mptr[r + 1] - mptr[r]) > domain) return;
    for (uint32_t k = mptr[r]; k < mptr[r + 1]; ++k) { assigned[mem[k]] = 1; dom.push_back(mem[k]); }
  };
  for (uint32_t seed = 0; seed < nv; ++seed) {
    if (assigned[seed]) continue;
    dom.clear();
    take(seed);
    size_t fi = 0;
    while (dom.size() < domain && fi < dom.size()) {
      const uint32_t v = dom[fi++];
      for (uint32_t k = cptr[v]; k < cptr[v + 1] && dom.size() < domain; ++k) take(cadj[k].second);
      for (uint32_t k = h_nbr_ptr[v]; k < h_nbr_ptr[v + 1] && dom.size() < domain; ++k) {
        const uint32_t nb = h_nbr[k];
        if (nb != v) take(nb);
      }
    }
    h_dom_row.insert(h_dom_row.end(), dom.begin(), dom.end());
    h_dom_ptr.push_back((uint32_t)h_dom_row.size());
  }
  upload(M, s);
  contact_built = true;
}

void Schwarz::reserve(uint32_t nd_cap) {
  nd_cap = std::max(nd_cap, 1u);
  ainv.reserve((size_t)nd_cap * SCHWARZ_INV_FLOATS);
  dom_dof.reserve((size_t)nd_cap * MD);
  dom_ptr.reserve((size_t)nd_cap + 1);
  inc_rng.reserve(2 * (size_t)nd_cap * SCHWARZ_MAX_DOMAIN);
  if (weight > 0.0f) rc.reserve(4 * (size_t)nd_cap);
}

// Device copies of the domains in h_dom_ptr / h_dom_row (and the coarse level over them).
void Schwarz::upload(const SellMatrix &M, hipStream_t s) {
  const uint32_t nv = M.nb;
  std::vector<uint32_t> dom;
  nd = (uint32_t)h_dom_ptr.size() - 1;
  // host copies live in the object: the uploads below are asynchronous
  std::vector<uint32_t> &rd = h_row_dom, &rp = h_row_pos, &dd = h_dom_dof;
  rd.assign(nv, NONE_S);
  rp.assign(nv, 0);
  for (uint32_t d = 0; d < nd; ++d)
    for (uint32_t k = h_dom_ptr[d]; k < h_dom_ptr[d + 1]; ++k) { rd[h_dom_row[k]] = d; rp[h_dom_row[k]] = k - h_dom_ptr[d]; }
  dd.assign((size_t)std::max(nd, 1u) * MD, NONE_S);
  for (uint32_t d = 0; d < nd; ++d)
    for (uint32_t k = h_dom_ptr[d]; k < h_dom_ptr[d + 1]; ++k)
      for (uint32_t c = 0; c < 3; ++c) dd[(size_t)d * MD + 3 * (k - h_dom_ptr[d]) + c] = 4 * h_dom_row[k] + c;
  dom_dof.upload(dd, s);
  dom_ptr.upload(h_dom_ptr, s);
  if (h_dom_row.empty()) h_dom_row.push_back(0);  // nd = 0: one unread entry
  dom_row.upload(h_dom_row, s);
  if (rd.empty()) { rd.push_back(0); rp.push_back(0); }
  row_dom.upload(rd, s);
  row_pos.upload(rp, s);
  ainv.reserve((size_t)std::max(nd, 1u) * SCHWARZ_INV_FLOATS);
  inc_rng.zero(2 * (size_t)std::max(nd, 1u) * SCHWARZ_MAX_DOMAIN, s);
  fails.zero(1, s);
  nd1 = 0;
  if (weight > 0.0f && nd > 0) {
    // coarse graph: domains G ~ H when a block of a row of G lies in a column of H
    // (linsolve.rs:573-605: the coarse sparsity of every fine entry); BFS as for level 0
    std::vector<std::vector<uint32_t>> adj(nd);
    for (uint32_t i = 0; i < nv; ++i)
      for (uint32_t k = 0; k < M.h_row_len[i]; ++k) {
        const uint32_t H = rd[M.h_col[M.slot(i, k)]];
        if (H != rd[i]) adj[rd[i]].push_back(H);
      }
    for (auto &a : adj) { std::sort(a.begin(), a.end()); a.erase(std::unique(a.begin(), a.end()), a.end()); }
    std::vector<uint8_t> cassigned(nd, 0);
    h_cdom_ptr.assign(1, 0);
    h_cdom_node.clear();
    for (uint32_t seed = 0; seed < nd; ++seed) {
      if (cassigned[seed]) continue;
      dom.assign(1, seed);
      cassigned[seed] = 1;
      size_t fi = 0;
      while (dom.size() < domain && fi < dom.size()) {
        const uint32_t g = dom[fi++];
        for (uint32_t h : adj[g])
          if (!cassigned[h] && dom.size() < domain) { cassigned[h] = 1; dom.push_back(h); }
      }
      h_cdom_node.insert(h_cdom_node.end(), dom.begin(), dom.end());
      h_cdom_ptr.push_back((uint32_t)h_cdom_node.size());
    }
    nd1 = (uint32_t)h_cdom_ptr.size() - 1;
    std::vector<uint32_t> &ncd = h_node_cdom, &ncp = h_node_cpos;
    ncd.assign(nd, 0);
    ncp.assign(nd, 0);
    for (uint32_t D = 0; D < nd1; ++D)
      for (uint32_t k = h_cdom_ptr[D]; k < h_cdom_ptr[D + 1]; ++k) { ncd[h_cdom_node[k]] = D; ncp[h_cdom_node[k]] = k - h_cdom_ptr[D]; }
    cdom_ptr.upload(h_cdom_ptr, s);
    cdom_node.upload(h_cdom_node, s);
    node_cdom.upload(ncd, s);
    node_cpos.upload(ncp, s);
    ainv1.reserve((size_t)nd1 * MD * MD);
    rc.zero(4 * (size_t)nd, s);
  }
}

// The level-0 inverses are stored as the 21 upper 16 x 16 tiles (P <= Q < 6) of the
// symmetric 96 x 96 block, tile (P, Q) at SCH_TILE * sym_tile(P, Q), row-major; every entry
// outside the domain's m x m block is zero. 5376 floats per domain instead of 9216.
constexpr uint32_t SCH_TB = MD / 16;                   // row blocks of 16
constexpr uint32_t SCH_TILE = 256;                     // floats per tile
constexpr uint32_t SCH_NT = SCH_TB * (SCH_TB + 1) / 2;  // tiles
static_assert(MD % 16 == 0, "the tiled inverse needs a whole number of 16-row blocks");
static_assert(SCHWARZ_INV_FLOATS == SCH_NT * SCH_TILE, "tiled inverse size");
__host__ __device__ constexpr uint32_t sym_tile(uint32_t P, uint32_t Q) { return P * SCH_TB - P * (P - 1) / 2 + (Q - P); }
__device__ inline void sym_tile_pq(uint32_t tile, uint32_t &P, uint32_t &Q) {
  uint32_t rem = tile;
  P = 0;
  while (rem >= SCH_TB - P) { rem -= SCH_TB - P; ++P; }
  Q = P + rem;
}

// The build keeps a block's lower triangle packed by rows: entry (i, j), j <= i, at tri(i) + j.
// The factorisation and the inverse read nothing above the diagonal.
__host__ __device__ constexpr uint32_t tri(uint32_t i) { return i * (i + 1) / 2; }

// Lower 16 x 16 tiles of the symmetric 96 x 96 block during the factorisation: tile (I, J),
// J <= I, has index I (I + 1) / 2 + J; wavefront w of the workgroup holds tiles w, w + 4, ...
// (at most SW_K per wavefront) in registers, in the matrix cores' result layout (lane l,
// register v: row 4 v + l / 16, column l % 16). Diagonal tiles are held whole.
constexpr uint32_t SW_K = (SCH_NT + 3) / 4;
__device__ inline void lower_tile_ij(uint32_t tile, uint32_t &I, uint32_t &J) {
  I = 0;
  while (tile > I) { tile -= I + 1; ++I; }
  J = tile;
}
__device__ inline double pick4(const double4_t &a, uint32_t v) { return v == 0 ? a[0] : v == 1 ? a[1] : v == 2 ? a[2] : a[3]; }
// x[k], k < 4 lane-dependent, as two levels of selects (no divergent branches)
__device__ inline double sel4(uint32_t k, double x0, double x1, double x2, double x3) {
  const double lo = (k & 1u) ? x1 : x0, hi = (k & 1u) ? x3 : x2;
  return (k & 2u) ? hi : lo;
}
__device__ inline void put4(double4_t &a, uint32_t v, double x) {
  if (v == 0) a[0] = x;
  else if (v == 1) a[1] = x;
  else if (v == 2) a[2] = x;
  else a[3] = x;
}

// Floor, FP64 inverse (diagonal fallback) of the symmetric m x m block whose lower triangle
// is packed in L (entry (i, j), j <= i, at tri(i) + j), written to out: TILED, as the 21 upper
// tiles above; otherwise row-major with row stride MD (columns m..m4-1 zero). All threads of
// the workgroup call it; L and diag0 are the caller's LDS.
//
// The inverse is the symmetric sweep (Gauss-Jordan elimination on an SPD matrix) over pivot
// blocks of four indices p0 .. p0 + 3: with D = S^-1 of the current 4 x 4 pivot block S,
//   A_ij <- A_ij - (A_iP D) A_Pj   (i, j outside P),  A_iP <- A_iP D,  A_PP <- -D,
// and after every index has been swept the block holds -A^-1. The Schur update of a 16 x 16
// tile is one rank-4 product on the FP64 matrix cores (v_mfma_f64_16x16x4f64: A operand lane
// a = (row a % 16, k a / 16), B operand lane b = (k b / 16, column b % 16)), with the tiles
// kept in registers across all sweeps; the four pivot columns are published to LDS by the
// tiles that hold them (double-buffered, one workgroup barrier per sweep). Every thread
// factors the 4 x 4 pivot block itself (Cholesky, then D = W^T W with W = L^-1). The pivots
// of those Cholesky factorisations are the Schur-complement pivots a_jj - sum_k l_jk^2 of the
// full Cholesky in order, so the reference's failure test (pivot <= 1e-14, linsolve.rs:80-99)
// and its diagonal fallback apply unchanged. Per sweep a wavefront loads the operands of all
// its tiles, issues the rank-4 products back to back, then rewrites the pivot rows and
// columns; lane-dependent choices are selects, not branches. The kernels that call this fit
// 168 VGPRs (three workgroups per compute unit: the 651 domains of a 20k-vertex garment run
// in one round).
template <bool TILED, class TOut = float>
__device__ void factor_invert(double *L, double *diag0, double &floor_v, uint32_t m, TOut *out, uint32_t *fails) {
  const uint32_t t = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(t / WAVE), lane = t % WAVE, r16 = lane % 16, q4 = lane / 16;
  // floor 1e-6 tr/m + 1e-8 (linsolve.rs:232-238)
  if (t == 0) {
    double tr = 0.0;
    for (uint32_t k = 0; k < m; ++k) tr += L[tri(k) + k];
    floor_v = 1e-6 * (tr / (double)m) + 1e-8;
  }
  __syncthreads();
  for (uint32_t k = t; k < m; k += BLOCK) { L[tri(k) + k] += floor_v; diag0[k] = L[tri(k) + k]; }
  __syncthreads();
  // registers <- tiles; indices m .. 16 nt - 1 padded as identity
  const uint32_t nt = (m + 15) / 16;
  double4_t acc[SW_K];
  uint32_t TI[SW_K], TJ[SW_K];
  bool act[SW_K];
#pragma unroll
  for (uint32_t k = 0; k < SW_K; ++k) {
    const uint32_t tile = wave + 4 * k;
    lower_tile_ij(tile < SCH_NT ? tile : 0, TI[k], TJ[k]);
    act[k] = tile < SCH_NT && TI[k] < nt;
    acc[k] = double4_t{0.0, 0.0, 0.0, 0.0};
    if (!act[k]) continue;
#pragma unroll
    for (uint32_t v = 0; v < 4; ++v) {
      const uint32_t i = 16 * TI[k] + 4 * v + q4, j = 16 * TJ[k] + r16;
      acc[k][v] = (i < m && j < m) ? (i >= j ? L[tri(i) + j] : L[tri(j) + i]) : (i == j ? 1.0 : 0.0);
    }
  }
  __syncthreads();
  // L's storage becomes the two pivot-column buffers: row i of buffer b at cbuf[b][4 i .. 4 i + 3]
  double *cbuf = L;
  auto publish = [&](uint32_t p0, double *cb) {
    const uint32_t Pt = p0 / 16, c0 = p0 % 16;
#pragma unroll
    for (uint32_t k = 0; k < SW_K; ++k) {
      if (!act[k]) continue;
      if (TJ[k] == Pt) {  // tiles (I, Pt): column p0 + (r16 - c0) of rows 16 I + 4 v + q4
        if (r16 >= c0 && r16 < c0 + 4)
#pragma unroll
          for (uint32_t v = 0; v < 4; ++v) cb[4 * (16 * TI[k] + 4 * v + q4) + (r16 - c0)] = acc[k][v];
      } else if (TI[k] == Pt) {  // tiles (Pt, J), J < Pt: row p0 + q4 is column p0 + q4 of row 16 J + r16
        cb[4 * (16 * TJ[k] + r16) + q4] = pick4(acc[k], c0 / 4);
      }
    }
  };
  publish(0, cbuf);
  __syncthreads();
  bool bad = false;
  uint32_t buf = 0;
  for (uint32_t p0 = 0; p0 < m; p0 += 4) {
    const double *cb = cbuf + buf * (4 * MD);
    const uint32_t Pt = p0 / 16, c0 = p0 % 16;
    // S = A_PP (row p0 + a at cb[4 (p0 + a) + b]); Cholesky with the pivot test, W = L^-1, D = W^T W.
    // 1 / l_jj = rsqrt(pivot): the hardware estimate and two Newton steps (full FP64 precision).
    const double *S = cb + 4 * p0;
    auto rsqrt64 = [](double x) {
      double y = __builtin_amdgcn_rsq(x);
      y = y * fma(-0.5 * x, y * y, 1.5);
      return y * fma(-0.5 * x, y * y, 1.5);
    };
    const double d0 = S[0];
    if (d0 <= 1e-14) { bad = true; break; }
    const double w00 = rsqrt64(d0), l10 = S[4] * w00, l20 = S[8] * w00, l30 = S[12] * w00;
    const double d1 = fma(-l10, l10, S[5]);
    if (d1 <= 1e-14) { bad = true; break; }
    const double w11 = rsqrt64(d1), l21 = fma(-l20, l10, S[9]) * w11, l31 = fma(-l30, l10, S[13]) * w11;
    const double d2 = fma(-l21, l21, fma(-l20, l20, S[10]));
    if (d2 <= 1e-14) { bad = true; break; }
    const double w22 = rsqrt64(d2), l32 = fma(-l31, l21, fma(-l30, l20, S[14])) * w22;
    const double d3 = fma(-l32, l32, fma(-l31, l31, fma(-l30, l30, S[15])));
    if (d3 <= 1e-14) { bad = true; break; }
    const double w33 = rsqrt64(d3);
    const double w10 = -w11 * (l10 * w00);
    const double w21 = -w22 * (l21 * w11), w20 = -w22 * fma(l21, w10, l20 * w00);
    const double w32 = -w33 * (l32 * w22), w31 = -w33 * fma(l32, w21, l31 * w11), w30 = -w33 * fma(l32, w20, fma(l31, w10, l30 * w00));
    // D = W^T W (symmetric)
    const double D00 = fma(w30, w30, fma(w20, w20, fma(w10, w10, w00 * w00)));
    const double D11 = fma(w31, w31, fma(w21, w21, w11 * w11));
    const double D22 = fma(w32, w32, w22 * w22);
    const double D33 = w33 * w33;
    const double D01 = fma(w31, w30, fma(w21, w20, w11 * w10));
    const double D02 = fma(w32, w30, w22 * w20);
    const double D03 = w33 * w30;
    const double D12 = fma(w32, w31, w22 * w21);
    const double D13 = w33 * w31;
    const double D23 = w33 * w32;
    // column q4 of D for the A operand and the pivot rows, column r16 - c0 for the pivot columns
    const uint32_t qc = (r16 - c0) & 3u;
    const double dk[4] = {sel4(q4, D00, D01, D02, D03), sel4(q4, D01, D11, D12, D13), sel4(q4, D02, D12, D22, D23), sel4(q4, D03, D13, D23, D33)};
    const double dc[4] = {sel4(qc, D00, D01, D02, D03), sel4(qc, D01, D11, D12, D13), sel4(qc, D02, D12, D22, D23), sel4(qc, D03, D13, D23, D33)};
    auto row_times = [&](uint32_t row, const double *dd) {
      const double *rr = cb + 4 * row;
      return fma(rr[3], dd[3], fma(rr[2], dd[2], fma(rr[1], dd[1], rr[0] * dd[0])));
    };
    // operands of every tile first (the LDS reads of all tiles in flight together), then the
    // rank-4 products back to back, then the pivot rows and columns
    double av[SW_K], bv[SW_K];
#pragma unroll
    for (uint32_t k = 0; k < SW_K; ++k) {
      if (!act[k]) continue;
      // A_IJ -= (A_IP D) (A_JP)^T, pivot rows and columns left out
      const uint32_t i = 16 * TI[k] + r16, j = 16 * TJ[k] + r16;
      const double a = -row_times(i, dk), b = cb[4 * j + q4];
      av[k] = i - p0 >= 4u ? a : 0.0;
      bv[k] = j - p0 >= 4u ? b : 0.0;
    }
#pragma unroll
    for (uint32_t k = 0; k < SW_K; ++k)
      if (act[k]) acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k], bv[k], acc[k], 0, 0, 0);
#pragma unroll
    for (uint32_t k = 0; k < SW_K; ++k) {
      if (!act[k]) continue;
      const uint32_t I = TI[k], J = TJ[k];
      if (J == Pt) {
        // pivot columns p0 + qc (lanes r16 = c0 + qc): A_iP D for rows outside P, -D on the pivot block
        double x[4];
#pragma unroll
        for (uint32_t v = 0; v < 4; ++v) {
          const uint32_t row = 16 * I + 4 * v + q4, pr = (row - p0) & 3u;
          const double piv = -sel4(pr, dc[0], dc[1], dc[2], dc[3]);
          x[v] = row - p0 < 4u ? piv : row_times(row, dc);
        }
        if (r16 - c0 < 4u) acc[k] = double4_t{x[0], x[1], x[2], x[3]};
      }
      if (I == Pt) {
        // pivot rows p0 + q4 (register c0 / 4) of tiles (Pt, J), columns outside P
        const uint32_t j = 16 * J + r16;
        const double y = row_times(j, dk);
        if (j - p0 >= 4u) put4(acc[k], c0 / 4, y);
      }
    }
    buf ^= 1u;
    if (p0 + 4 < m) publish(p0 + 4, cbuf + buf * (4 * MD));
    __syncthreads();
  }
  const uint32_t m4 = (m + 3) & ~3u;
  if (bad) {
    // diagonal fallback: factor sqrt(a_pp) (or 1) -> inverse 1 / a_pp
    if (t == 0) atomicAdd(fails, 1u);
    if (TILED) {
      for (uint32_t e = t; e < SCH_NT * SCH_TILE; e += BLOCK) {
        uint32_t P, Q;
        sym_tile_pq(e / SCH_TILE, P, Q);
        const uint32_t p = 16 * P + (e % SCH_TILE) / 16, q = 16 * Q + e % 16;
        const double v = p < m ? diag0[p] : 0.0;
        out[e] = (p == q && p < m) ? (TOut)(v > 1e-14 ? 1.0 / v : 1.0) : TOut(0);
      }
      return;
    }
    for (uint32_t i = t; i < m * m4; i += BLOCK) {
      uint32_t p = i / m4, q = i % m4;
      double v = diag0[p];
      out[p * MD + q] = p == q ? (TOut)(v > 1e-14 ? 1.0 / v : 1.0) : TOut(0);
    }
    return;
  }
  // the registers hold -A^-1: lower tile (I, J) is the transpose of upper tile (J, I)
#pragma unroll
  for (uint32_t k = 0; k < SW_K; ++k) {
    const uint32_t tile = wave + 4 * k;
    if (tile >= SCH_NT) continue;
    const uint32_t I = TI[k], J = TJ[k];
#pragma unroll
    for (uint32_t v = 0; v < 4; ++v) {
      const uint32_t i = 16 * I + 4 * v + q4, j = 16 * J + r16;
      const TOut val = (act[k] && i < m && j < m) ? (TOut)(-acc[k][v]) : TOut(0);
      if (TILED) {
        out[SCH_TILE * sym_tile(J, I) + 16 * r16 + 4 * v + q4] = val;
      } else if (act[k]) {
        if (i < m && j < m4) out[i * MD + j] = val;
        if (I != J && j < m && i < m4) out[j * MD + i] = val;
      }
    }
  }
}

// Dense block of domain d (elastic + contact), FP64 Cholesky with the reference's floor and
// diagonal fallback, explicit inverse A^-1 = L^-T L^-1 written as FP32 (the FP32 PCG) or FP64
// (the FP64 PCG).
template <class TOut>
__global__ __launch_bounds__(BLOCK, 3) void k_schwarz_build(SellView A, ContactView C, const uint32_t *__restrict__ dom_ptr,
                                                        const uint32_t *__restrict__ dom_row, const uint32_t *__restrict__ row_dom,
                                                        const uint32_t *__restrict__ row_pos, TOut *__restrict__ ainv, uint32_t *inc_rng,
                                                        uint32_t *fails) {
  __shared__ double L[MD * (MD + 1) / 2];
  __shared__ double diag0[MD];
  __shared__ double floor_v;
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const uint32_t base = dom_ptr[d], nv = dom_ptr[d + 1] - base, m = 3 * nv;
  for (uint32_t i = t; i < tri(m); i += BLOCK) L[i] = 0.0;
  __syncthreads();
  // elastic blocks of the domain rows: 8 threads per row, one writer per block
  for (uint32_t p = t / 8; p < nv; p += BLOCK / 8) {
    const uint32_t row = dom_row[base + p];
    const uint32_t sbase = A.slice_ptr[row >> 6], lane = row & 63u, len = A.row_len[row];
    for (uint32_t k = t % 8; k < len; k += 8) {
      const uint32_t sl = sbase + 64u * k + lane;
      const uint32_t c = A.col[sl];
      if (row_dom[c] != d) continue;
      const uint32_t q = row_pos[c];
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc)
          if (3 * q + cc <= 3 * p + r) L[tri(3 * p + r) + 3 * q + cc] = (double)A.val[sell_val_index(sl, 3 * r + cc)];
    }
  }
  __syncthreads();
  // contact: hp g_a g_b^T + lambda w_a w_b (I - n n^T) for partners inside the domain. One
  // thread per scalar row 3 p + r (so each entry of L has one writer): the row's incidences
  // are added in list order, partners in stencil order; the row's incidence range is kept
  // for the apply
  if (C.ns) {
    for (uint32_t i = t; i < 3 * nv; i += BLOCK) {
      const uint32_t p = i / 3, r = i % 3;
      const uint32_t row = dom_row[base + p];
      const uint32_t k0 = C.inc_ptr[row], k1 = C.inc_ptr[row + 1];
      if (r == 0) {
        inc_rng[2 * ((size_t)d * SCHWARZ_MAX_DOMAIN + p)] = k0;
        inc_rng[2 * ((size_t)d * SCHWARZ_MAX_DOMAIN + p) + 1] = k1;
      }
      double *Lr = L + tri(3 * p + r);
      for (uint32_t k = k0; k < k1; ++k) {
        const uint32_t e = C.inc[k], s = e >> 2, a = e & 3u;
        const double hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
        const double gar = (double)C.g[12 * (size_t)s + 3 * a + r];
        const float *n = C.n + 3 * (size_t)s;
        const double nr = (double)n[r];
        const double wa = C.w[4 * (size_t)s + a];
        for (uint32_t b = 0; b < 4; ++b) {
          const uint32_t db = C.dof[4 * (size_t)s + b];
          if (db == NONE_S || row_dom[db] != d) continue;
          const uint32_t q = row_pos[db];
          const float *gb = C.g + 12 * (size_t)s + 3 * b;
          const double wb = C.w[4 * (size_t)s + b];
          for (uint32_t cc = 0; cc < 3; ++cc)
            if (3 * q + cc <= 3 * p + r)
              Lr[3 * q + cc] += hp * gar * (double)gb[cc] + lam * wa * wb * ((r == cc ? 1.0 : 0.0) - nr * (double)n[cc]);
        }
      }
    }
    __syncthreads();
  }
  factor_invert<true, TOut>(L, diag0, floor_v, m, ainv + (size_t)d * SCHWARZ_INV_FLOATS, fails);
}

// Coarse block of coarse domain D (linsolve.rs:613-625 coarse_matrix, 230-258 factorize):
// entry (3a + r, 3b + c) sums the r,c entries of every block (i, j), i in domain G_a, j in
// domain G_b, elastic and contact, G_a and G_b members a, b of D. Thread 3a + r owns coarse
// row 3a + r and walks the rows of G_a in domain order (no write conflicts, fixed order);
// then the same floor, FP64 factorisation and FP32 inverse as level 0.
__global__ __launch_bounds__(BLOCK, 3) void k_schwarz_build_coarse(SellView A, ContactView C, const uint32_t *__restrict__ dom_ptr,
                                                               const uint32_t *__restrict__ dom_row, const uint32_t *__restrict__ row_dom,
                                                               const uint32_t *__restrict__ cdom_ptr, const uint32_t *__restrict__ cdom_node,
                                                               const uint32_t *__restrict__ node_cdom, const uint32_t *__restrict__ node_cpos,
                                                               float *__restrict__ ainv1, uint32_t *fails) {
  __shared__ double L[MD * (MD + 1) / 2];
  __shared__ double diag0[MD];
  __shared__ double floor_v;
  const uint32_t D = blockIdx.x, t = threadIdx.x;
  const uint32_t cb = cdom_ptr[D], n1 = cdom_ptr[D + 1] - cb, m = 3 * n1;
  for (uint32_t i = t; i < tri(m); i += BLOCK) L[i] = 0.0;
  __syncthreads();
  if (t < m) {
    const uint32_t a = t / 3, r = t % 3, G = cdom_node[cb + a];
    double *lrow = L + tri(t);  // row t of the packed lower triangle: columns 0 .. t
    for (uint32_t k = dom_ptr[G]; k < dom_ptr[G + 1]; ++k) {
      const uint32_t row = dom_row[k];
      const uint32_t sbase = A.slice_ptr[row >> 6], lane = row & 63u, len = A.row_len[row];
      for (uint32_t kk = 0; kk < len; ++kk) {
        const uint32_t sl = sbase + 64u * kk + lane;
        const uint32_t H = row_dom[A.col[sl]];
        if (node_cdom[H] != D) continue;
        const uint32_t b = node_cpos[H];
        for (int c = 0; c < 3; ++c)
          if (3 * b + c <= t) lrow[3 * b + c] += (double)A.val[sell_val_index(sl, 3 * r + c)];
      }
      if (C.ns) {
        for (uint32_t kk = C.inc_ptr[row]; kk < C.inc_ptr[row + 1]; ++kk) {
          const uint32_t e = C.inc[kk], s = e >> 2, ai = e & 3u;
          const double hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
          const double gar = C.g[12 * (size_t)s + 3 * ai + r], nr = C.n[3 * (size_t)s + r];
          const double wa = C.w[4 * (size_t)s + ai];
          for (uint32_t bi = 0; bi < 4; ++bi) {
            const uint32_t db = C.dof[4 * (size_t)s + bi];
            if (db == NONE_S) continue;
            const uint32_t H = row_dom[db];
            if (node_cdom[H] != D) continue;
            const uint32_t b = node_cpos[H];
            const double wb = C.w[4 * (size_t)s + bi];
            for (int c = 0; c < 3; ++c)
              if (3 * b + c <= t)
                lrow[3 * b + c] += hp * gar * (double)C.g[12 * (size_t)s + 3 * bi + c] +
                                   lam * wa * wb * ((r == (uint32_t)c ? 1.0 : 0.0) - nr * (double)C.n[3 * (size_t)s + c]);
          }
        }
      }
    }
  }
  __syncthreads();
  factor_invert<false>(L, diag0, floor_v, m, ainv1 + (size_t)D * MD * MD, fails);
}

// Coarse correction, one workgroup per domain d (member a of coarse domain D): e =
// (A1_D^-1 rc_D) rows 3a..3a+2, then z += weight e on d's rows (p = z too for init), and the
// r.z partial of the corrected z.
constexpr int COARSE_THREADS = 128;
template <bool INIT>
__global__ __launch_bounds__(COARSE_THREADS) void k_schwarz_coarse(uint32_t nd, const uint32_t *__restrict__ dom_dof,
                                                                  const uint32_t *__restrict__ cdom_ptr, const uint32_t *__restrict__ cdom_node,
                                                                  const uint32_t *__restrict__ node_cdom, const uint32_t *__restrict__ node_cpos,
                                                                  const float *__restrict__ ainv1, const float4 *__restrict__ rc, float weight,
                                                                  const float *__restrict__ r, float *__restrict__ z, float *__restrict__ p,
                                                                  double *nxt, const uint32_t *flag) {
  __shared__ float rcs[MD];
  __shared__ float ec[3];
  __shared__ double lds[COARSE_THREADS / WAVE];
  if (!INIT && flag[0]) return;
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const uint32_t D = node_cdom[d], a = node_cpos[d], cb = cdom_ptr[D], n1 = cdom_ptr[D + 1] - cb, m1 = 3 * n1;
  const uint32_t o = t < MD ? dom_dof[(size_t)d * MD + t] : NONE_S;
  float r0 = 0.f, z0 = 0.f;
  if (o != NONE_S) { r0 = r[o]; z0 = z[o]; }
  if (t < n1) {
    const float4 v = rc[cdom_node[cb + t]];
    rcs[3 * t] = v.x; rcs[3 * t + 1] = v.y; rcs[3 * t + 2] = v.z;
  }
  __syncthreads();
  // rows 3a + c of the coarse inverse: 32 lanes per row, fixed-order lane sums
  if (t < 3 * 32) {
    const uint32_t c = t / 32, l = t % 32;
    const float *row = ainv1 + (size_t)D * MD * MD + (size_t)(3 * a + c) * MD;
    float acc = 0.0f;
    for (uint32_t k = l; k < m1; k += 32) acc += row[k] * rcs[k];
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 32);
    if (l == 0) ec[c] = acc;
  }
  __syncthreads();
  double v = 0.0;
  if (o != NONE_S) {
    const float zn = z0 + weight * ec[t % 3];
    z[o] = zn;
    if (INIT) p[o] = zn;
    v = (double)r0 * zn;
  }
  v = wave_sum(v);
  if ((t & (WAVE - 1)) == 0) lds[t / WAVE] = v;
  __syncthreads();
  if (t == 0) nxt[d] = lds[0] + lds[1];
}

void launch_schwarz_coarse(hipStream_t s, const Schwarz &sz, bool init, const float *r, float *z, float *p, double *nxt, const uint32_t *flag) {
  if (!sz.nd1) return;
  if (init)
    hipLaunchKernelGGL(k_schwarz_coarse<true>, dim3(sz.nd), dim3(COARSE_THREADS), 0, s, sz.nd, (const uint32_t *)sz.dom_dof.p,
                       (const uint32_t *)sz.cdom_ptr.p, (const uint32_t *)sz.cdom_node.p, (const uint32_t *)sz.node_cdom.p,
                       (const uint32_t *)sz.node_cpos.p, (const float *)sz.ainv1.p, (const float4 *)sz.rc.p, sz.weight, r, z, p, nxt, flag);
  else
    hipLaunchKernelGGL(k_schwarz_coarse<false>, dim3(sz.nd), dim3(COARSE_THREADS), 0, s, sz.nd, (const uint32_t *)sz.dom_dof.p,
                       (const uint32_t *)sz.cdom_ptr.p, (const uint32_t *)sz.cdom_node.p, (const uint32_t *)sz.node_cdom.p,
                       (const uint32_t *)sz.node_cpos.p, (const float *)sz.ainv1.p, (const float4 *)sz.rc.p, sz.weight, r, z, p, nxt, flag);
}

void Pcg::build_precond(const SellMatrix &M, const ContactView &C) {
  ensure(M.nb, C.ns);
  if (sz_) {
    if (sz_->nd) hipLaunchKernelGGL(k_schwarz_build<float>, dim3(sz_->nd), dim3(BLOCK), 0, s_, M.view(), C, (const uint32_t *)sz_->dom_ptr.p,
                                    (const uint32_t *)sz_->dom_row.p, (const uint32_t *)sz_->row_dom.p, (const uint32_t *)sz_->row_pos.p,
                                    sz_->ainv.p, sz_->inc_rng.p, sz_->fails.p);
    if (sz_->nd1)
      hipLaunchKernelGGL(k_schwarz_build_coarse, dim3(sz_->nd1), dim3(BLOCK), 0, s_, M.view(), C, (const uint32_t *)sz_->dom_ptr.p,
                         (const uint32_t *)sz_->dom_row.p, (const uint32_t *)sz_->row_dom.p, (const uint32_t *)sz_->cdom_ptr.p,
                         (const uint32_t *)sz_->cdom_node.p, (const uint32_t *)sz_->node_cdom.p, (const uint32_t *)sz_->node_cpos.p,
                         sz_->ainv1.p, sz_->fails.p);
  } else {
    launch_minv(M, C);
  }
  NGUO_HIP(hipGetLastError());
}

// Preconditioner kernel of the PCG iteration, one workgroup per domain. Not INIT: the
// workgroup reduces pAp (npap SpMV partials) itself, takes rz_j from scal[8 + (j & 1)],
// stops on breakdown,
// then applies the PCG update to the domain's rows (r -= alpha q, d += alpha p); INIT:
// r = (f32) r64, d = 0, p = z. Then z = A_d^-1 r from the tiled inverse: thread t < 168 owns
// the 4-column quarter t % 4 of rows 2s + (t / 4) % 2 (s < 8) of tile t / 8, eight float4
// loads issued before anything else (8 adjacent lanes read 128 contiguous bytes); it forms
// the direct products (row 16P + i of the tile against r's block Q) and, off the diagonal,
// the transposed ones (column 16Q + j against r's block P). Lanes combine them by shuffles,
// and row k of z sums its 6 tile partials from LDS in tile order. Each row belongs to exactly
// one domain, so r, d and z have a single writer. Vectors have stride 4 per block row, r64
// stride 3. Partials r.z -> nxt[d], r.r -> nxt[nd + d].
constexpr uint32_t SCH_TPT = 8;                    // threads per tile
constexpr uint32_t SCH_THREADS = SCH_NT * SCH_TPT;  // 168
constexpr uint32_t SCH_ROWS = 16 / (SCH_TPT / 4);   // tile rows per thread
static_assert(SCH_THREADS <= (uint32_t)BLOCK && SCH_TPT == 8, "tile threads must fit the workgroup");

template <bool INIT>
__device__ __forceinline__ void schwarz_pcg_body(uint32_t d, uint32_t nd, const uint32_t *__restrict__ dom_ptr, const uint32_t *__restrict__ dom_dof,
                                                 const float *__restrict__ ainv, const double *__restrict__ r64, float *__restrict__ r,
                                                 float *__restrict__ z, float *__restrict__ dsol, float *__restrict__ p,
                                                 const float *__restrict__ q, const uint32_t *__restrict__ inc_rng,
                                                 const float4 *__restrict__ contrib, const double *pap_part, uint32_t npap, double *nxt,
                                                 double *scal, uint32_t *flag, uint32_t j, float4 *rc) {
  __shared__ float4 rs4[MD / 4];
  __shared__ float zdir[SCH_NT][16];
  __shared__ float4 ztr[SCH_NT][4];
  __shared__ double lds[2 * BLOCK / WAVE];
  float *rs = reinterpret_cast<float *>(rs4);
  const uint32_t t = threadIdx.x;
  const uint32_t state = INIT ? 0u : flag[0];
  // loads that need no scalar: this thread's rows of its tile (every tile entry is written by
  // the build, zero outside the domain's m x m block), the iteration launch's partials, the
  // domain rows and their incidence ranges; then this thread's r / q / p / d entries and the
  // first eight contact blocks of its row. None sits under a per-thread branch (threads past
  // the tiles read the last tile, rows past the domain read entry 0, missing blocks re-read a
  // present one, and none uses what it read), so the compiler can count the loads in flight
  // and never drains them one by one; the block-wide test d < nd (always true) keeps them
  // ahead of the state test rather than sunk below it. Two memory round trips in all.
  // lane l of a tile's 8: rows of parity h = l / 4, column quarter cq = l (h = 0) or 7 - l
  // (h = 1), so the half-row mirror pairs the two row parities of one quarter
  const uint32_t tile = t / SCH_TPT, h = (t / 4) % 2, cq = h ? 3 - t % 4 : t % 4;
  const bool mac = t < SCH_THREADS;
  float4 w[SCH_ROWS];
  uint32_t od0 = NONE_S, ib = 0, ie = 0;
  constexpr uint32_t NCB = 8;  // contact blocks of a row loaded together
  float r0 = 0.f, q0 = 0.f, p0 = 0.f, d0 = 0.f, cv[NCB];
#pragma unroll
  for (uint32_t u = 0; u < NCB; ++u) cv[u] = 0.f;
#pragma unroll
  for (uint32_t i = 0; i < SCH_ROWS; ++i) w[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  const double *src[1] = {pap_part};
  const uint32_t npp[1] = {npap};
  PartSums<BLOCK, 1> ps;
  const uint32_t base = dom_ptr[d], m = 3 * (dom_ptr[d + 1] - base);
  if (d < nd) {
    const uint32_t tl = mac ? tile : SCH_NT - 1;
    const float4 *a4 = reinterpret_cast<const float4 *>(ainv + (size_t)d * SCHWARZ_INV_FLOATS) + (size_t)tl * (SCH_TILE / 4) + 4 * h + cq;
#pragma unroll
    for (uint32_t i = 0; i < SCH_ROWS; ++i) w[i] = a4[8 * i];
    if (!INIT) ps.load(src, npp);
    od0 = dom_dof[(size_t)d * MD + min(t, MD - 1)];
    uint2 rg = make_uint2(0u, 0u);
    if (!INIT) rg = reinterpret_cast<const uint2 *>(inc_rng)[(size_t)d * SCHWARZ_MAX_DOMAIN + min(t / 3, SCHWARZ_MAX_DOMAIN - 1)];
    // every load above is issued before anything waits on the domain rows
    __builtin_amdgcn_sched_barrier(0);
    const size_t oc = (t < MD && od0 != NONE_S) ? od0 : 0u;
    if (INIT) {
      r0 = (float)r64[3 * (oc >> 2) + (oc & 3)];
    } else {
      ib = contrib ? rg.x : 0u;
      ie = contrib ? rg.y : 0u;
      r0 = r[oc]; q0 = q[oc]; p0 = p[oc]; d0 = dsol[oc];
      // contrib holds at least one entry whenever it is given; otherwise r stands in
      const float *cf = contrib ? reinterpret_cast<const float *>(contrib) + (oc & 3) : r;
#pragma unroll
      for (uint32_t u = 0; u < NCB; ++u) cv[u] = cf[4 * (size_t)(ie > ib ? min(ib + u, ie - 1u) : 0u)];
    }
  }
  const uint32_t od = t < MD ? od0 : NONE_S;
  const size_t o = od != NONE_S ? od : 0u;
  if (state) return;
  float alpha = 0.0f;
  if (!INIT) {
    double v[1];
    ps.sum(src, npp, v, lds);
    if (!(v[0] > 0.0)) {
      if (d == 0 && threadIdx.x == 0) flag[0] = 2;
      return;
    }
    const double a = scal[8 + (j & 1u)] / v[0];
    alpha = (float)a;
    if (d == 0 && threadIdx.x == 0) { scal[1] = v[0]; scal[2] = a; }
  }
  float rv = 0.0f;
  if (t < m) {
    if (INIT) {
      rv = r0;
      dsol[o] = 0.0f;
    } else {
      // the row's contact blocks of H p, in incidence order; past the first NCB, NCB more per
      // memory round trip (clamped to the row's last block, so no load waits on a branch)
      float qv = q0;
#pragma unroll
      for (uint32_t u = 0; u < NCB; ++u)
        if (ib + u < ie) qv += cv[u];
      for (uint32_t k0 = ib + NCB; k0 < ie; k0 += NCB) {
        const float *cf = reinterpret_cast<const float *>(contrib) + (o & 3);
#pragma unroll
        for (uint32_t u = 0; u < NCB; ++u) cv[u] = cf[4 * (size_t)min(k0 + u, ie - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < NCB; ++u)
          if (k0 + u < ie) qv += cv[u];
      }
      rv = r0 - alpha * qv;
      dsol[o] = d0 + alpha * p0;
    }
    r[o] = rv;
    rs[t] = rv;
  } else if (t < MD) {
    rs[t] = 0.0f;
  }
  __syncthreads();
  if (rc && t < 3) {
    // restriction for the coarse level: the domain's rows summed (domain order)
    float acc = 0.0f;
    for (uint32_t k = t; k < m; k += 3) acc += rs[k];
    reinterpret_cast<float *>(rc)[4 * (size_t)d + t] = acc;
  }
  if (mac) {
    uint32_t P, Q;
    sym_tile_pq(tile, P, Q);
    const float4 rq = rs4[4 * Q + cq];
    float dz[SCH_ROWS];
    float4 tz = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (uint32_t i = 0; i < SCH_ROWS; ++i) {
      const float4 a = w[i];
      dz[i] = a.x * rq.x + a.y * rq.y + a.z * rq.z + a.w * rq.w;
      const float rp = rs[16 * P + 2 * i + h];
      tz.x += a.x * rp; tz.y += a.y * rp; tz.z += a.z * rp; tz.w += a.w * rp;
    }
    // DPP exchanges: the four quarters of a row inside each quad, the two row parities of a
    // quarter by the half-row mirror; a + b == b + a, so both lanes of each exchange hold the
    // same sum
#pragma unroll
    for (uint32_t i = 0; i < SCH_ROWS; ++i) {
      dz[i] += dpp_f<DPP_XOR1>(dz[i]);
      dz[i] += dpp_f<DPP_XOR2>(dz[i]);
    }
    tz.x += dpp_f<DPP_HALF_MIRROR>(tz.x); tz.y += dpp_f<DPP_HALF_MIRROR>(tz.y);
    tz.z += dpp_f<DPP_HALF_MIRROR>(tz.z); tz.w += dpp_f<DPP_HALF_MIRROR>(tz.w);
    if (cq == 0)
#pragma unroll
      for (uint32_t i = 0; i < SCH_ROWS; ++i) zdir[tile][2 * i + h] = dz[i];
    if (h == 0) ztr[tile][cq] = tz;
  }
  __syncthreads();
  double v[2] = {0.0, 0.0};
  if (t < m) {
    const uint32_t R = t / 16, k = t % 16;
    const float *zt = reinterpret_cast<const float *>(&ztr[0][0]);
    float zv = zdir[sym_tile(R, R)][k];
    for (uint32_t Q = R + 1; Q < SCH_TB; ++Q) zv += zdir[sym_tile(R, Q)][k];
    for (uint32_t P = 0; P < R; ++P) zv += zt[16 * sym_tile(P, R) + k];
    z[o] = zv;
    if (INIT) p[o] = zv;
    v[0] = (double)rv * zv;
    v[1] = (double)rv * rv;
  }
  wg_sum_all<BLOCK, 2>(v, lds);
  if (t == 0) { nxt[d] = v[0]; nxt[nd + d] = v[1]; }
}

template <bool INIT>
__global__ __launch_bounds__(BLOCK) void k_schwarz_pcg(uint32_t nd, const uint32_t *__restrict__ dom_ptr, const uint32_t *__restrict__ dom_dof,
                                                      const float *__restrict__ ainv, const double *__restrict__ r64, float *__restrict__ r,
                                                      float *__restrict__ z, float *__restrict__ dsol, float *__restrict__ p,
                                                      const float *__restrict__ q, const uint32_t *__restrict__ inc_rng,
                                                      const float4 *__restrict__ contrib, const double *pap_part, uint32_t npap, double *nxt,
                                                      double *scal, uint32_t *flag, uint32_t j, float4 *rc) {
  schwarz_pcg_body<INIT>(blockIdx.x, nd, dom_ptr, dom_dof, ainv, r64, r, z, dsol, p, q, inc_rng, contrib, pap_part, npap, nxt, scal, flag, j, rc);
}

// Iteration form (not INIT, one level) over several systems.
__global__ __launch_bounds__(BLOCK) void kb_schwarz_pcg(BatchOf<SchArgs> B, uint32_t j) {
  const uint32_t i = batch_slot(B.off, B.k, blockIdx.x);
  const SchArgs &a = B.a[i];
  schwarz_pcg_body<false>(blockIdx.x - B.off[i], a.nd, a.dom_ptr, a.dom_dof, a.ainv, nullptr, a.r, a.z, a.dsol, a.p, a.q, a.inc_rng, a.contrib,
                          a.pap_part, a.npap, a.nxt, a.scal, a.flag, j, nullptr);
}

void launch_schwarz_pcg_batch(hipStream_t s, const BatchOf<SchArgs> &B, uint32_t j) {
  if (!B.k || !B.off[B.k]) return;
  hipLaunchKernelGGL(kb_schwarz_pcg, dim3(B.off[B.k]), dim3(BLOCK), 0, s, B, j);
}

void launch_schwarz_pcg(hipStream_t s, const Schwarz &sz, bool init, const double *r64, float *r, float *z, float *dsol, float *p,
                        const float *q, const float4 *contrib, const double *pap_part, uint32_t npap, double *nxt, double *scal, uint32_t *flag,
                        uint32_t j) {
  if (!sz.nd) return;
  if (init)
    hipLaunchKernelGGL(k_schwarz_pcg<true>, dim3(sz.nd), dim3(BLOCK), 0, s, sz.nd, (const uint32_t *)sz.dom_ptr.p, (const uint32_t *)sz.dom_dof.p,
                       (const float *)sz.ainv.p, r64, r, z, dsol, p, q, (const uint32_t *)nullptr, (const float4 *)nullptr, pap_part, npap, nxt,
                       scal, flag, j, sz.nd1 ? (float4 *)sz.rc.p : (float4 *)nullptr);
  else
    hipLaunchKernelGGL(k_schwarz_pcg<false>, dim3(sz.nd), dim3(BLOCK), 0, s, sz.nd, (const uint32_t *)sz.dom_ptr.p, (const uint32_t *)sz.dom_dof.p,
                       (const float *)sz.ainv.p, r64, r, z, dsol, p, q, (const uint32_t *)sz.inc_rng.p, contrib, pap_part, npap, nxt, scal, flag, j,
                       sz.nd1 ? (float4 *)sz.rc.p : (float4 *)nullptr);
}

// The FP64 PCG's preconditioner: the same domain blocks, factorisation and floor, inverses
// kept in FP64 (sz.ainv64, the 21 upper tiles per domain).
void build_schwarz64(hipStream_t s, const SellMatrix &M, const ContactView &C, Schwarz &sz) {
  sz.ainv64.reserve((size_t)std::max<uint32_t>(sz.nd, 1) * SCHWARZ_INV_FLOATS);
  if (sz.nd)
    hipLaunchKernelGGL(k_schwarz_build<double>, dim3(sz.nd), dim3(BLOCK), 0, s, M.view(), C, (const uint32_t *)sz.dom_ptr.p,
                       (const uint32_t *)sz.dom_row.p, (const uint32_t *)sz.row_dom.p, (const uint32_t *)sz.row_pos.p, sz.ainv64.p,
                       sz.inc_rng.p, sz.fails.p);
  NGUO_HIP(hipGetLastError());
}

}  // namespace nguo
