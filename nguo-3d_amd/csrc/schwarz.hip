// One-level additive Schwarz preconditioner on the device (see pcg.h). Reference:
// engine-core/src/linsolve.rs:126-258 SchwarzTopology::{build, factorize, apply} and the GPU
// apply it fed, engine-gpu/src/schwarz.rs + shaders/schwarz.wgsl (one thread per domain,
// serial triangular solves). Here a workgroup owns a domain: the dense block lives in LDS,
// the Cholesky and the triangular inverse run in FP64 across the workgroup, and the PCG
// applies the explicit inverse with coalesced column reads.
#include <algorithm>

#include "pcg.h"

namespace nguo {

constexpr uint32_t NONE_S = 0xffffffffu;
constexpr uint32_t MD = SCHWARZ_MAX_DOF;

void Schwarz::build(const SellMatrix &M, uint32_t domain_size, hipStream_t s) {
  domain = std::max(1u, std::min(domain_size, SCHWARZ_MAX_DOMAIN));
  const uint32_t nv = M.nb;
  // linsolve.rs:134-170: adjacency = off-diagonal block columns (sorted), greedy BFS from
  // seeds in index order.
  std::vector<uint8_t> assigned(nv, 0);
  h_dom_ptr.assign(1, 0);
  h_dom_row.clear();
  std::vector<uint32_t> dom;
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
  nd = (uint32_t)h_dom_ptr.size() - 1;
  std::vector<uint32_t> rd(nv, NONE_S), rp(nv, 0);
  for (uint32_t d = 0; d < nd; ++d)
    for (uint32_t k = h_dom_ptr[d]; k < h_dom_ptr[d + 1]; ++k) { rd[h_dom_row[k]] = d; rp[h_dom_row[k]] = k - h_dom_ptr[d]; }
  dom_ptr.upload(h_dom_ptr, s);
  dom_row.upload(h_dom_row.empty() ? std::vector<uint32_t>{0} : h_dom_row, s);
  row_dom.upload(rd.empty() ? std::vector<uint32_t>{0} : rd, s);
  row_pos.upload(rp.empty() ? std::vector<uint32_t>{0} : rp, s);
  ainv.reserve((size_t)std::max(nd, 1u) * MD * MD);
  fails.zero(1, s);
}

// Dense block of domain d (elastic + contact), FP64 Cholesky with the reference's floor and
// diagonal fallback, explicit inverse A^-1 = L^-T L^-1 written as FP32.
__global__ __launch_bounds__(BLOCK) void k_schwarz_build(SellView A, ContactView C, const uint32_t *__restrict__ dom_ptr,
                                                        const uint32_t *__restrict__ dom_row, const uint32_t *__restrict__ row_dom,
                                                        const uint32_t *__restrict__ row_pos, float *__restrict__ ainv, uint32_t *fails) {
  __shared__ double L[MD * MD];
  __shared__ double W[MD * MD];
  __shared__ double diag0[MD];
  __shared__ double red[BLOCK];
  __shared__ int failed;
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const uint32_t base = dom_ptr[d], nv = dom_ptr[d + 1] - base, m = 3 * nv;
  for (uint32_t i = t; i < m * m; i += BLOCK) L[i] = 0.0;
  if (t == 0) failed = 0;
  __syncthreads();
  // elastic blocks of the domain rows (one thread per domain row: no write conflicts)
  for (uint32_t p = t; p < nv; p += BLOCK) {
    uint32_t row = dom_row[base + p];
    uint32_t sbase = A.slice_ptr[row >> 6], lane = row & 63u;
    for (uint32_t k = 0; k < A.row_len[row]; ++k) {
      uint32_t sl = sbase + 64u * k + lane;
      uint32_t c = A.col[sl];
      if (row_dom[c] != d) continue;
      uint32_t q = row_pos[c];
      for (int r = 0; r < 3; ++r)
        for (int cc = 0; cc < 3; ++cc) L[(3 * p + r) * m + 3 * q + cc] = (double)A.val[sell_val_index(sl, 3 * r + cc)];
    }
    // contact: hp g_a g_b^T + lambda w_a w_b (I - n n^T) for partners inside the domain
    if (C.ns) {
      for (uint32_t k = C.inc_ptr[row]; k < C.inc_ptr[row + 1]; ++k) {
        uint32_t e = C.inc[k], s = e >> 2, a = e & 3u;
        double hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
        const float *ga = C.g + 12 * (size_t)s + 3 * a;
        const float *n = C.n + 3 * (size_t)s;
        double wa = C.w[4 * (size_t)s + a];
        for (uint32_t b = 0; b < 4; ++b) {
          uint32_t db = C.dof[4 * (size_t)s + b];
          if (db == NONE_S || row_dom[db] != d) continue;
          uint32_t q = row_pos[db];
          const float *gb = C.g + 12 * (size_t)s + 3 * b;
          double wb = C.w[4 * (size_t)s + b];
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
              L[(3 * p + r) * m + 3 * q + cc] +=
                  hp * (double)ga[r] * (double)gb[cc] + lam * wa * wb * ((r == cc ? 1.0 : 0.0) - (double)n[r] * (double)n[cc]);
        }
      }
    }
  }
  __syncthreads();
  // floor 1e-6 tr/m + 1e-8 (linsolve.rs:232-238)
  double tr = 0.0;
  if (t == 0)
    for (uint32_t k = 0; k < m; ++k) tr += L[k * m + k];
  if (t == 0) red[0] = 1e-6 * (tr / (double)m) + 1e-8;
  __syncthreads();
  double fl = red[0];
  __syncthreads();
  for (uint32_t k = t; k < m; k += BLOCK) { L[k * m + k] += fl; diag0[k] = L[k * m + k]; }
  __syncthreads();
  // left-looking Cholesky, column j (linsolve.rs:80-99 order: sums over k ascending)
  for (uint32_t j = 0; j < m; ++j) {
    if (t == 0) {
      double s = L[j * m + j];
      for (uint32_t k = 0; k < j; ++k) s -= L[j * m + k] * L[j * m + k];
      if (s <= 1e-14) failed = 1;
      else L[j * m + j] = sqrt(s);
    }
    __syncthreads();
    if (failed) break;
    for (uint32_t i = j + 1 + t; i < m; i += BLOCK) {
      double s = L[i * m + j];
      for (uint32_t k = 0; k < j; ++k) s -= L[i * m + k] * L[j * m + k];
      L[i * m + j] = s / L[j * m + j];
    }
    __syncthreads();
  }
  float *out = ainv + (size_t)d * MD * MD;
  if (failed) {
    // diagonal fallback: factor sqrt(a_pp) (or 1) -> inverse 1 / a_pp
    if (t == 0) atomicAdd(fails, 1u);
    for (uint32_t i = t; i < m * m; i += BLOCK) {
      uint32_t p = i / m, q = i % m;
      double v = diag0[p];
      out[p * MD + q] = p == q ? (float)(v > 1e-14 ? 1.0 / v : 1.0) : 0.0f;
    }
    return;
  }
  // W = L^-1 (lower), one column per thread
  for (uint32_t j = t; j < m; j += BLOCK) {
    for (uint32_t i = 0; i < j; ++i) W[i * m + j] = 0.0;
    for (uint32_t i = j; i < m; ++i) {
      double s = i == j ? 1.0 : 0.0;
      for (uint32_t k = j; k < i; ++k) s -= L[i * m + k] * W[k * m + j];
      W[i * m + j] = s / L[i * m + i];
    }
  }
  __syncthreads();
  // A^-1 = W^T W
  for (uint32_t i = t; i < m * m; i += BLOCK) {
    uint32_t p = i / m, q = i % m;
    uint32_t k0 = p > q ? p : q;
    double s = 0.0;
    for (uint32_t k = k0; k < m; ++k) s += W[k * m + p] * W[k * m + q];
    out[p * MD + q] = (float)s;
  }
}

void Pcg::build_precond(const SellMatrix &M, const ContactView &C) {
  ensure(M.nb, C.ns);
  if (sz_) {
    if (sz_->nd) hipLaunchKernelGGL(k_schwarz_build, dim3(sz_->nd), dim3(BLOCK), 0, s_, M.view(), C, (const uint32_t *)sz_->dom_ptr.p,
                                    (const uint32_t *)sz_->dom_row.p, (const uint32_t *)sz_->row_dom.p, (const uint32_t *)sz_->row_pos.p,
                                    sz_->ainv.p, sz_->fails.p);
  } else {
    launch_minv(M, C);
  }
  NGUO_HIP(hipGetLastError());
}

// z = A_d^-1 r over each domain; with `alpha` the PCG update r -= alpha q, d += alpha p is
// folded in first. Partials per domain: r.z -> part[d], r.r -> part[nd + d].
template <bool UPDATE>
__global__ __launch_bounds__(BLOCK) void k_schwarz_apply(uint32_t nd, const uint32_t *__restrict__ dom_ptr, const uint32_t *__restrict__ dom_row,
                                                        const float *__restrict__ ainv, const double *r64, float *__restrict__ r,
                                                        float *__restrict__ z, float *__restrict__ dsol, float *__restrict__ p,
                                                        const float *__restrict__ q, const double *pap_part, uint32_t npap,
                                                        const double *cur, double *part, double *scal, uint32_t *flag, int init) {
  __shared__ float rs[MD];
  __shared__ double lds[4];
  __shared__ double bc;
  if (flag && flag[0]) return;
  const uint32_t d = blockIdx.x, t = threadIdx.x;
  const uint32_t base = dom_ptr[d], m = 3 * (dom_ptr[d + 1] - base);
  float alpha = 0.0f;
  if (UPDATE) {
    double pap = 0, rz = 0;
    {
      double acc = 0;
      for (uint32_t i = t; i < npap; i += BLOCK) acc += pap_part[i];
      pap = block_sum(acc, lds);
      if (t == 0) bc = pap;
      __syncthreads();
      pap = bc;
      __syncthreads();
      acc = 0;
      for (uint32_t i = t; i < nd; i += BLOCK) acc += cur[i];
      rz = block_sum(acc, lds);
      if (t == 0) bc = rz;
      __syncthreads();
      rz = bc;
      __syncthreads();
    }
    if (!(pap > 0.0)) {
      if (d == 0 && t == 0) flag[0] = 2;
      return;
    }
    alpha = (float)(rz / pap);
    if (d == 0 && t == 0) { scal[1] = pap; scal[2] = rz / pap; }
  }
  size_t o = 0;
  float rv = 0.0f;
  if (t < m) {
    uint32_t row = dom_row[base + t / 3];
    o = 3 * (size_t)row + t % 3;
    if (init) {
      rv = (float)r64[o];
      dsol[o] = 0.0f;
    } else {
      rv = r[o] - alpha * q[o];
      dsol[o] += alpha * p[o];
    }
    r[o] = rv;
    rs[t] = rv;
  }
  __syncthreads();
  double rzp = 0, rrp = 0;
  if (t < m) {
    const float *col = ainv + (size_t)d * MD * MD + t;
    float zv = 0.0f;
    for (uint32_t k = 0; k < m; ++k) zv += col[(size_t)k * MD] * rs[k];
    z[o] = zv;
    if (init) p[o] = zv;
    rzp = (double)rv * zv;
    rrp = (double)rv * rv;
  }
  double a = block_sum(rzp, lds);
  double b = block_sum(rrp, lds);
  if (t == 0) { part[d] = a; part[nd + d] = b; }
}

void launch_schwarz_init(hipStream_t s, const Schwarz &sz, const double *r64, float *r, float *z, float *dsol, float *p, double *part) {
  hipLaunchKernelGGL((k_schwarz_apply<false>), dim3(sz.nd), dim3(BLOCK), 0, s, sz.nd, (const uint32_t *)sz.dom_ptr.p,
                     (const uint32_t *)sz.dom_row.p, (const float *)sz.ainv.p, r64, r, z, dsol, p, (const float *)nullptr,
                     (const double *)nullptr, 0u, (const double *)nullptr, part, (double *)nullptr, (uint32_t *)nullptr, 1);
}

void launch_schwarz_update(hipStream_t s, const Schwarz &sz, float *r, float *z, float *dsol, float *p, const float *q, const double *pap_part,
                           uint32_t npap, const double *cur, double *nxt, double *scal, uint32_t *flag) {
  hipLaunchKernelGGL((k_schwarz_apply<true>), dim3(sz.nd), dim3(BLOCK), 0, s, sz.nd, (const uint32_t *)sz.dom_ptr.p,
                     (const uint32_t *)sz.dom_row.p, (const float *)sz.ainv.p, (const double *)nullptr, r, z, dsol, p, q, pap_part, npap, cur,
                     nxt, scal, flag, 0);
}

}  // namespace nguo
