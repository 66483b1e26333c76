// SELL-64 3x3-block SpMV and the FP32 PCG / FP64 iterative-refinement solver.
//
// Reference being replaced: engine-gpu/src/pcg.rs:488-673 Pcg::solve (scalar-CSR SpMV, one
// thread per scalar row, f32 alpha/beta, residual read back every 16 iterations) and the
// CPU oracle engine-core/src/linsolve.rs:350-411 cg_run (f64 block-Jacobi PCG). Here the
// matrix is stored as 3x3 blocks (36 B values + 4 B column per block instead of 9 x 8 B),
// the scalars alpha/beta/rz live on the device in FP64, vectors are FP32, and the true
// residual of each refinement sweep is recomputed in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pcg.h"

namespace nguo {

constexpr uint32_t NONE = 0xffffffffu;

// ---------------------------------------------------------------- pattern
void SellMatrix::build_pattern(uint32_t nblocks, const std::vector<std::vector<uint32_t>> &cols, hipStream_t s) {
  nb = nblocks;
  nslices = (nb + 63) / 64;
  h_slice_ptr.assign(nslices + 1, 0);
  h_row_len.assign(nb, 0);
  nnzb = 0;
  for (uint32_t sl = 0; sl < nslices; ++sl) {
    uint32_t mx = 0;
    for (uint32_t r = sl * 64; r < std::min(nb, sl * 64 + 64); ++r) {
      h_row_len[r] = (uint32_t)cols[r].size();
      mx = std::max(mx, h_row_len[r]);
      nnzb += h_row_len[r];
    }
    h_slice_ptr[sl + 1] = h_slice_ptr[sl] + 64u * mx;
  }
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
__device__ inline void contact_gather(const ContactView &C, uint32_t i, const float *sums, float &y0, float &y1, float &y2) {
  for (uint32_t k = C.inc_ptr[i]; k < C.inc_ptr[i + 1]; ++k) {
    uint32_t e = C.inc[k], s = e >> 2, a = e & 3u;
    float hp = C.coef[2 * s], lam = C.coef[2 * s + 1];
    const float *sm = sums + 4 * (size_t)s;
    float t = hp * sm[0];
    const float *g = C.g + 12 * (size_t)s + 3 * a;
    y0 += t * g[0]; y1 += t * g[1]; y2 += t * g[2];
    if (lam != 0.0f) {
      const float *n = C.n + 3 * (size_t)s;
      float nd = n[0] * sm[1] + n[1] * sm[2] + n[2] * sm[3];
      float pw = lam * C.w[4 * (size_t)s + a];
      y0 += pw * (sm[1] - n[0] * nd); y1 += pw * (sm[2] - n[1] * nd); y2 += pw * (sm[3] - n[2] * nd);
    }
  }
}

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

// y = (A + C) x; part[block] = sum x.y over the block's rows (when part != null).
__global__ __launch_bounds__(BLOCK) void k_spmv(SellView A, ContactView C, const float *__restrict__ x, float *__restrict__ y,
                                               double *part, const uint32_t *flag) {
  __shared__ double lds[4];
  if (flag && flag[0]) return;
  uint32_t row = blockIdx.x * BLOCK + threadIdx.x;
  double dp = 0;
  if (row < A.nb) {
    uint32_t base = A.slice_ptr[row >> 6], len = A.row_len[row], lane = row & 63u;
    float y0 = 0.f, y1 = 0.f, y2 = 0.f;
    for (uint32_t k = 0; k < len; ++k) {
      uint32_t sl = base + 64u * k + lane;
      uint32_t c = A.col[sl];
      const float *v = A.val + (size_t)(base + 64u * k) * 9u + lane;
      float x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
      y0 += v[0] * x0 + v[64] * x1 + v[128] * x2;
      y1 += v[192] * x0 + v[256] * x1 + v[320] * x2;
      y2 += v[384] * x0 + v[448] * x1 + v[512] * x2;
    }
    if (C.ns) contact_gather(C, row, C.sums, y0, y1, y2);
    y[3 * (size_t)row] = y0; y[3 * (size_t)row + 1] = y1; y[3 * (size_t)row + 2] = y2;
    if (part) dp = (double)x[3 * (size_t)row] * y0 + (double)x[3 * (size_t)row + 1] * y1 + (double)x[3 * (size_t)row + 2] * y2;
  }
  if (part) {
    double t = block_sum(dp, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
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
      const float *v = A.val + (size_t)(base + 64u * k) * 9u + lane;
      double x0 = x[3 * (size_t)c], x1 = x[3 * (size_t)c + 1], x2 = x[3 * (size_t)c + 2];
      y0 += (double)v[0] * x0 + (double)v[64] * x1 + (double)v[128] * x2;
      y1 += (double)v[192] * x0 + (double)v[256] * x1 + (double)v[320] * x2;
      y2 += (double)v[384] * x0 + (double)v[448] * x1 + (double)v[512] * x2;
    }
    if (C.ns) contact_gather64(C, row, sums64, y0, y1, y2);
    double r0 = (double)b[3 * (size_t)row] - y0, r1 = (double)b[3 * (size_t)row + 1] - y1, r2 = (double)b[3 * (size_t)row + 2] - y2;
    r[3 * (size_t)row] = r0; r[3 * (size_t)row + 1] = r1; r[3 * (size_t)row + 2] = r2;
    rr = r0 * r0 + r1 * r1 + r2 * r2;
  }
  double t = block_sum(rr, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sums `n` partials in a fixed order (single workgroup).
__device__ inline double reduce_parts(const double *part, uint32_t n, double *lds) {
  double acc = 0;
  for (uint32_t i = threadIdx.x; i < n; i += BLOCK) acc += part[i];
  return block_sum(acc, lds);
}

// scal: 0 rz, 1 pAp, 2 alpha, 3 beta, 4 rr, 5 threshold, 6 rr0.  flag: 0 state, 1 iters, 2 cap.
__global__ __launch_bounds__(BLOCK) void k_alpha(double *scal, uint32_t *flag, const double *part, uint32_t npart) {
  __shared__ double lds[4];
  if (flag[0]) return;
  double pap = reduce_parts(part, npart, lds);
  if (threadIdx.x == 0) {
    scal[1] = pap;
    if (!(pap > 0.0)) { flag[0] = 2; scal[2] = 0.0; }
    else scal[2] = scal[0] / pap;
  }
}

__global__ __launch_bounds__(BLOCK) void k_update(uint32_t nb, const float *__restrict__ p, const float *__restrict__ q,
                                                 float *__restrict__ d, float *__restrict__ r, float *__restrict__ z,
                                                 const float *__restrict__ minv, const double *scal, double *part, const uint32_t *flag) {
  __shared__ double lds[4];
  if (flag[0]) return;
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  float alpha = (float)scal[2];
  double rz = 0, rr = 0;
  if (i < nb) {
    size_t o = 3 * (size_t)i;
    float r0 = r[o] - alpha * q[o], r1 = r[o + 1] - alpha * q[o + 1], r2 = r[o + 2] - alpha * q[o + 2];
    d[o] += alpha * p[o]; d[o + 1] += alpha * p[o + 1]; d[o + 2] += alpha * p[o + 2];
    r[o] = r0; r[o + 1] = r1; r[o + 2] = r2;
    const float *m = minv + 9 * (size_t)i;
    float z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
    float z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
    float z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
    z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
    rz = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
    rr = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
  }
  double a = block_sum(rz, lds);
  double b = block_sum(rr, lds);
  if (threadIdx.x == 0) { part[blockIdx.x] = a; part[gridDim.x + blockIdx.x] = b; }
}

__global__ __launch_bounds__(BLOCK) void k_beta(double *scal, uint32_t *flag, const double *part, uint32_t npart) {
  __shared__ double lds[4];
  if (flag[0]) return;
  double rz_new = reduce_parts(part, npart, lds);
  double rr = reduce_parts(part + npart, npart, lds);
  if (threadIdx.x == 0) {
    uint32_t it = ++flag[1];
    scal[4] = rr;
    scal[3] = rz_new / scal[0];
    scal[0] = rz_new;
    if (rr <= scal[5]) flag[0] = 1;
    else if (it >= flag[2]) flag[0] = 3;
  }
}

__global__ __launch_bounds__(BLOCK) void k_pupdate(uint32_t nb, const float *__restrict__ z, float *__restrict__ p, const double *scal,
                                                  const uint32_t *flag) {
  if (flag[0]) return;
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= 3 * nb) return;
  float beta = (float)scal[3];
  p[i] = z[i] + beta * p[i];
}

// Fused iteration (small systems, <= FUSE_MAX_PARTS workgroups): every workgroup reduces the
// few partials it needs itself, in the same fixed order, so the scalars never round-trip
// through a single-workgroup kernel. Partials of r.z / r.r alternate between two buffers by
// iteration parity: `cur` holds rz_i, the update writes rz_{i+1} into `nxt`.
constexpr uint32_t FUSE_MAX_PARTS = 4096;

__device__ inline double reduce_parts_bcast(const double *part, uint32_t n, double *lds, double *bcast) {
  double s = reduce_parts(part, n, lds);
  if (threadIdx.x == 0) *bcast = s;
  __syncthreads();
  double v = *bcast;
  __syncthreads();
  return v;
}

__global__ __launch_bounds__(BLOCK) void k_update_fused(uint32_t nb, const float *__restrict__ p, const float *__restrict__ q,
                                                       float *__restrict__ d, float *__restrict__ r, float *__restrict__ z,
                                                       const float *__restrict__ minv, const double *pap_part, const double *cur,
                                                       double *nxt, double *scal, uint32_t *flag, uint32_t npart) {
  __shared__ double lds[4];
  __shared__ double bc;
  if (flag[0]) return;
  double pap = reduce_parts_bcast(pap_part, npart, lds, &bc);
  double rz = reduce_parts_bcast(cur, npart, lds, &bc);
  if (!(pap > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 2;
    return;
  }
  float alpha = (float)(rz / pap);
  if (blockIdx.x == 0 && threadIdx.x == 0) { scal[1] = pap; scal[2] = rz / pap; }
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double rzn = 0, rr = 0;
  if (i < nb) {
    size_t o = 3 * (size_t)i;
    float r0 = r[o] - alpha * q[o], r1 = r[o + 1] - alpha * q[o + 1], r2 = r[o + 2] - alpha * q[o + 2];
    d[o] += alpha * p[o]; d[o + 1] += alpha * p[o + 1]; d[o + 2] += alpha * p[o + 2];
    r[o] = r0; r[o + 1] = r1; r[o + 2] = r2;
    const float *m = minv + 9 * (size_t)i;
    float z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
    float z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
    float z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
    z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
    rzn = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
    rr = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
  }
  double a = block_sum(rzn, lds);
  double b = block_sum(rr, lds);
  if (threadIdx.x == 0) { nxt[blockIdx.x] = a; nxt[npart + blockIdx.x] = b; }
}

__global__ __launch_bounds__(BLOCK) void k_pupdate_fused(uint32_t nb, const float *__restrict__ z, float *__restrict__ p, const double *cur,
                                                        const double *nxt, double *scal, uint32_t *flag, uint32_t npart) {
  __shared__ double lds[4];
  __shared__ double bc;
  if (flag[0]) return;
  double rz_old = reduce_parts_bcast(cur, npart, lds, &bc);
  double rz_new = reduce_parts_bcast(nxt, npart, lds, &bc);
  double rr = reduce_parts_bcast(nxt + npart, npart, lds, &bc);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint32_t it = ++flag[1];
    scal[0] = rz_new;
    scal[3] = rz_new / rz_old;
    scal[4] = rr;
    if (rr <= scal[5]) flag[0] = 1;
    else if (it >= flag[2]) flag[0] = 3;
  }
  if (rr <= scal[5]) return;  // converged: p is no longer needed
  float beta = (float)(rz_new / rz_old);
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < nb) {
    size_t o = 3 * (size_t)i;
    p[o] = z[o] + beta * p[o];
    p[o + 1] = z[o + 1] + beta * p[o + 1];
    p[o + 2] = z[o + 2] + beta * p[o + 2];
  }
}

// Start of a sweep: r = (f32) r64, d = 0, z = Minv r, p = z; partials of r.z and r.r.
__global__ __launch_bounds__(BLOCK) void k_init(uint32_t nb, const double *__restrict__ r64, float *__restrict__ r, float *__restrict__ d,
                                               float *__restrict__ z, float *__restrict__ p, const float *__restrict__ minv, double *part) {
  __shared__ double lds[4];
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  double rz = 0, rr = 0;
  if (i < nb) {
    size_t o = 3 * (size_t)i;
    float r0 = (float)r64[o], r1 = (float)r64[o + 1], r2 = (float)r64[o + 2];
    r[o] = r0; r[o + 1] = r1; r[o + 2] = r2;
    d[o] = 0.f; d[o + 1] = 0.f; d[o + 2] = 0.f;
    const float *m = minv + 9 * (size_t)i;
    float z0 = m[0] * r0 + m[1] * r1 + m[2] * r2;
    float z1 = m[3] * r0 + m[4] * r1 + m[5] * r2;
    float z2 = m[6] * r0 + m[7] * r1 + m[8] * r2;
    z[o] = z0; z[o + 1] = z1; z[o + 2] = z2;
    p[o] = z0; p[o + 1] = z1; p[o + 2] = z2;
    rz = (double)r0 * z0 + (double)r1 * z1 + (double)r2 * z2;
    rr = (double)r0 * r0 + (double)r1 * r1 + (double)r2 * r2;
  }
  double a = block_sum(rz, lds);
  double b = block_sum(rr, lds);
  if (threadIdx.x == 0) { part[blockIdx.x] = a; part[gridDim.x + blockIdx.x] = b; }
}

__global__ __launch_bounds__(BLOCK) void k_init_scal(double *scal, uint32_t *flag, const double *part, uint32_t npart, double eta2,
                                                    uint32_t cap) {
  __shared__ double lds[4];
  double rz = reduce_parts(part, npart, lds);
  double rr = reduce_parts(part + npart, npart, lds);
  if (threadIdx.x == 0) {
    scal[0] = rz; scal[4] = rr; scal[6] = rr; scal[5] = eta2 * rr;
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
__global__ void k_accum(uint32_t n, double *x, const float *d) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
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
Pcg::~Pcg() {
  if (h_flag_) (void)hipHostFree(h_flag_);
  if (h_scal_) (void)hipHostFree(h_scal_);
}

void Pcg::ensure(uint32_t nb, uint32_t ns) {
  {
    // pAp partials (G), then two parity buffers of (r.z, r.r) partials of max(G, domains) each
    size_t G = div_up(nb, BLOCK), nz = std::max<size_t>(G, sz_ ? sz_->nd : 0);
    part_.reserve(G + 4 * nz + 2 * (size_t)div_up(3 * (size_t)nb, BLOCK) + 8);
  }
  if (!h_flag_) {
    NGUO_HIP(hipHostMalloc((void **)&h_flag_, 4 * sizeof(uint32_t), hipHostMallocDefault));
    NGUO_HIP(hipHostMalloc((void **)&h_scal_, 8 * sizeof(double), hipHostMallocDefault));
  }
  if (nb > cap_nb_) {
    size_t n3 = 3 * (size_t)nb;
    r_.reserve(n3); z_.reserve(n3); p_.reserve(n3); q_.reserve(n3); d_.reserve(n3); bf_.reserve(n3);
    minv_.reserve(9 * (size_t)nb);
    x64_.reserve(n3); r64_.reserve(n3);
    cap_nb_ = nb;
  }
  scal_.reserve(8);
  flag_.reserve(4);
  if (ns > cap_ns_) {
    sums64_.reserve(4 * (size_t)ns);
    cap_ns_ = ns;
  }
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

void Pcg::apply_precond(const SellMatrix &M, const float *r, float *z) {
  const uint32_t nb = M.nb, n3 = 3 * nb;
  if (sz_) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(div_up(n3, BLOCK)), dim3(BLOCK), 0, s_, n3, r, r64_.p);
    launch_schwarz_init(s_, *sz_, r64_.p, r_.p, z, d_.p, p_.p, part_.p);
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
  const unsigned G = div_up(nb, BLOCK), G3 = div_up(n3, BLOCK), GS = div_up(C.ns, BLOCK);
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
  for (uint32_t sweep = 0; sweep < 12; ++sweep) {
    double res = std::sqrt(rr / bb);
    if (res <= (double)tol || st.iters >= max_iter) break;
    double eta = std::min(0.5, std::max(0.25 * (double)tol / res, 1.0e-5));
    const uint32_t nz = sz_ ? sz_->nd : G;  // partials of r.z / r.r: per domain or per row block
    const bool fused = G <= FUSE_MAX_PARTS && nz <= FUSE_MAX_PARTS;
    NGUO_REQUIRE(fused || !sz_, NGUO_E_UNSUPPORTED, "Schwarz preconditioner: too many domains for the fused iteration");
    double *pap_part = part_.p, *bufA = part_.p + G, *bufB = part_.p + G + 2 * (size_t)std::max(G, nz);
    double *init_part = fused ? bufA : part_.p;
    if (sz_) launch_schwarz_init(s_, *sz_, r64_.p, r_.p, z_.p, d_.p, p_.p, init_part);
    else hipLaunchKernelGGL(k_init, dim3(G), dim3(BLOCK), 0, s_, nb, (const double *)r64_.p, r_.p, d_.p, z_.p, p_.p, (const float *)minv_.p, init_part);
    hipLaunchKernelGGL(k_init_scal, dim3(1), dim3(BLOCK), 0, s_, scal_.p, flag_.p, (const double *)init_part, nz, eta * eta, max_iter - st.iters);
    uint64_t launched = 0;
    while (true) {
      if (fused && sz_) {
        for (uint32_t c = 0; c < chunk; ++c, ++launched) {
          double *cur = (launched & 1) ? bufB : bufA, *nxt = (launched & 1) ? bufA : bufB;
          if (C.ns) hipLaunchKernelGGL((k_contact_sums<float, float>), dim3(GS), dim3(BLOCK), 0, s_, C, (const float *)p_.p, C.sums, (const uint32_t *)flag_.p);
          hipLaunchKernelGGL(k_spmv, dim3(G), dim3(BLOCK), 0, s_, A, C, (const float *)p_.p, q_.p, pap_part, (const uint32_t *)flag_.p);
          launch_schwarz_update(s_, *sz_, r_.p, z_.p, d_.p, p_.p, q_.p, pap_part, G, cur, nxt, scal_.p, flag_.p);
          hipLaunchKernelGGL(k_pupdate_fused, dim3(G), dim3(BLOCK), 0, s_, nb, (const float *)z_.p, p_.p, (const double *)cur, (const double *)nxt,
                             scal_.p, flag_.p, nz);
        }
        NGUO_HIP(hipGetLastError());
        NGUO_HIP(hipMemcpyAsync(h_flag_, flag_.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
        NGUO_HIP(hipStreamSynchronize(s_));
        if (h_flag_[0]) break;
        continue;
      }
      if (fused) {
        for (uint32_t c = 0; c < chunk; ++c, ++launched) {
          double *cur = (launched & 1) ? bufB : bufA, *nxt = (launched & 1) ? bufA : bufB;
          if (C.ns) hipLaunchKernelGGL((k_contact_sums<float, float>), dim3(GS), dim3(BLOCK), 0, s_, C, (const float *)p_.p, C.sums, (const uint32_t *)flag_.p);
          hipLaunchKernelGGL(k_spmv, dim3(G), dim3(BLOCK), 0, s_, A, C, (const float *)p_.p, q_.p, pap_part, (const uint32_t *)flag_.p);
          hipLaunchKernelGGL(k_update_fused, dim3(G), dim3(BLOCK), 0, s_, nb, (const float *)p_.p, (const float *)q_.p, d_.p, r_.p, z_.p,
                             (const float *)minv_.p, (const double *)pap_part, (const double *)cur, nxt, scal_.p, flag_.p, G);
          hipLaunchKernelGGL(k_pupdate_fused, dim3(G), dim3(BLOCK), 0, s_, nb, (const float *)z_.p, p_.p, (const double *)cur, (const double *)nxt,
                             scal_.p, flag_.p, G);
        }
        NGUO_HIP(hipGetLastError());
        NGUO_HIP(hipMemcpyAsync(h_flag_, flag_.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
        NGUO_HIP(hipStreamSynchronize(s_));
        if (h_flag_[0]) break;
        continue;
      }
      for (uint32_t c = 0; c < chunk; ++c) {
        if (C.ns) hipLaunchKernelGGL((k_contact_sums<float, float>), dim3(GS), dim3(BLOCK), 0, s_, C, (const float *)p_.p, C.sums, (const uint32_t *)flag_.p);
        hipLaunchKernelGGL(k_spmv, dim3(G), dim3(BLOCK), 0, s_, A, C, (const float *)p_.p, q_.p, part_.p, (const uint32_t *)flag_.p);
        hipLaunchKernelGGL(k_alpha, dim3(1), dim3(BLOCK), 0, s_, scal_.p, flag_.p, (const double *)part_.p, G);
        hipLaunchKernelGGL(k_update, dim3(G), dim3(BLOCK), 0, s_, nb, (const float *)p_.p, (const float *)q_.p, d_.p, r_.p, z_.p,
                           (const float *)minv_.p, (const double *)scal_.p, part_.p, (const uint32_t *)flag_.p);
        hipLaunchKernelGGL(k_beta, dim3(1), dim3(BLOCK), 0, s_, scal_.p, flag_.p, (const double *)part_.p, G);
        hipLaunchKernelGGL(k_pupdate, dim3(G3), dim3(BLOCK), 0, s_, nb, (const float *)z_.p, p_.p, (const double *)scal_.p, (const uint32_t *)flag_.p);
      }
      NGUO_HIP(hipGetLastError());
      NGUO_HIP(hipMemcpyAsync(h_flag_, flag_.p, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s_));
      NGUO_HIP(hipStreamSynchronize(s_));
      if (h_flag_[0]) break;
    }
    st.iters += h_flag_[1];
    st.sweeps += 1;
    if (h_flag_[0] == 2) st.breakdown = true;
    hipLaunchKernelGGL(k_accum, dim3(G3), dim3(BLOCK), 0, s_, n3, x64_.p, (const float *)d_.p);
    if (C.ns) hipLaunchKernelGGL((k_contact_sums<double, double>), dim3(GS), dim3(BLOCK), 0, s_, C, (const double *)x64_.p, sums64_.p, nullptr);
    hipLaunchKernelGGL(k_residual64, dim3(G), dim3(BLOCK), 0, s_, A, C, (const double *)sums64_.p, b, (const double *)x64_.p, r64_.p, part_.p);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(BLOCK), 0, s_, (const double *)part_.p, G, scal_.p + 7);
    NGUO_HIP(hipMemcpyAsync(h_scal_ + 7, scal_.p + 7, sizeof(double), hipMemcpyDeviceToHost, s_));
    NGUO_HIP(hipStreamSynchronize(s_));
    rr = h_scal_[7];
    if (st.breakdown && h_flag_[1] == 0) break;
  }
  hipLaunchKernelGGL(k_to_f32, dim3(G3), dim3(BLOCK), 0, s_, n3, (const double *)x64_.p, x_out);
  NGUO_HIP(hipGetLastError());
  st.residual = std::sqrt(rr / bb);
  return st;
}

void Pcg::apply(const SellView &A, const ContactView &C, const float *x, float *y) {
  ensure(A.nb, C.ns);
  if (C.ns) hipLaunchKernelGGL((k_contact_sums<float, float>), dim3(div_up(C.ns, BLOCK)), dim3(BLOCK), 0, s_, C, x, C.sums, nullptr);
  hipLaunchKernelGGL(k_spmv, dim3(div_up(A.nb, BLOCK)), dim3(BLOCK), 0, s_, A, C, x, y, nullptr, nullptr);
  NGUO_HIP(hipGetLastError());
}

}  // namespace nguo
