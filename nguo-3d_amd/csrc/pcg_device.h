// Device helpers of the PCG iteration kernels (k_pcg_iter and the block-Jacobi / additive
// Schwarz preconditioner kernels): workgroup sums with a fixed combination order, and the
// redundant per-workgroup reduction of the previous launch's partials.
//
// Each launch stores one partial sum per workgroup; the next launch's workgroups each
// reduce those partials themselves, in the same fixed order, so every workgroup holds
// bitwise the same scalar without a single-workgroup launch or a cross-workgroup handshake
// in between (kernel boundaries order all the memory traffic).
//
// scal: 1 pAp, 2 alpha, 3 beta, 4 rr, 5 threshold (eta^2 rr0), 6 rr0, 7 host-read FP64
//       residual, 8 / 9 rz_j for even / odd j.
// flag: 0 state (0 running, 1 converged, 2 breakdown, 3 iteration cap), 1 iterations done,
//       2 cap.
#pragma once
#include "device_common.h"

namespace nguo {

// Several systems in one launch: the workgroups of system i are [off[i], off[i + 1]) of the
// grid, and a[i] holds its arguments (kernel-argument memory, read with scalar loads).
constexpr uint32_t MAX_BATCH = 8;
template <class ARGS> struct BatchOf {
  ARGS a[MAX_BATCH];
  uint32_t off[MAX_BATCH + 1];
  uint32_t k;
};
__device__ inline uint32_t batch_slot(const uint32_t *off, uint32_t k, uint32_t b) {
  uint32_t i = 0;
  while (i + 1 < k && b >= off[i + 1]) ++i;
  return i;
}

// Iteration launch of one system: the SELL slices' workgroups, then its stencils' (64 a
// workgroup); cur holds the preconditioner's (r.z | r.r) partials, nz each.
struct SpmvArgs {
  SellView A;
  ContactView C;
  const float4 *z, *pold;
  float4 *pnew, *q;
  double *part;
  const double *cur;
  uint32_t nz;
  double *scal;
  uint32_t *flag;
  // host-visible copy of flag[0..2] (state, iterations, cap), written by the first slice's
  // workgroup; set on the last launch of a queued chunk, null otherwise
  uint32_t *hflag = nullptr;
};
struct SchArgs {
  uint32_t nd;
  const uint32_t *dom_ptr, *dom_dof;
  const float *ainv;
  float *r, *z, *dsol, *p;
  const float *q;
  const uint32_t *inc_rng;
  const float4 *contrib;
  const double *pap_part;
  uint32_t npap;
  double *nxt, *scal;
  uint32_t *flag;
};

// Sum of K values over an NT-thread workgroup, broadcast to every thread. Fixed order.
template <int NT, int K> __device__ inline void wg_sum_all(double (&v)[K], double *lds /* [NT / WAVE * K] */) {
  constexpr int NW = NT / WAVE;
  const int w = threadIdx.x / WAVE;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (NW == 1) return;  // one wavefront: its sum is already in every lane
  if ((threadIdx.x & (WAVE - 1)) == 0)
    for (int k = 0; k < K; ++k) lds[w * K + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = lds[k];
    for (int i = 1; i < NW; ++i) s += lds[i * K + k];
    v[k] = s;
  }
  __syncthreads();
}

// v[k] = sum_{i < n[k]} src[k][i] for K partial arrays (src[k] holds at least one entry),
// every thread gets the sums. Each thread adds its strided share in index order. The first U
// strides are loaded with no branch (indices past n[k] re-read entry n[k] - 1 and add
// nothing) and outside any loop, so they join the loads the caller already has in flight (a
// loop would make the compiler drain those first); a launch with up to U NT partials per
// array pays one memory round trip. Further strides follow in a loop. PartSums splits the
// two so a kernel can issue the first strides' loads early and sum them later.
template <int NT, int K, uint32_t U = 4> struct PartSums {
  double x[K][U];
  __device__ inline void load(const double *const (&src)[K], const uint32_t (&n)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t last = n[k] ? n[k] - 1u : 0u;
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) x[k][u] = src[k][min(u * NT + threadIdx.x, last)];
    }
  }
  __device__ inline void sum(const double *const (&src)[K], const uint32_t (&n)[K], double (&v)[K], double *lds) {
    uint32_t nmax = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      nmax = n[k] > nmax ? n[k] : nmax;
      v[k] = 0.0;
#pragma unroll
      for (uint32_t u = 0; u < U; ++u)
        if (u * NT + threadIdx.x < n[k]) v[k] += x[k][u];
    }
    for (uint32_t b = U * NT; b < nmax; b += U * NT) {
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint32_t i = b + u * NT + threadIdx.x;
          x[k][u] = i < n[k] ? src[k][i] : 0.0;
        }
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (uint32_t u = 0; u < U; ++u)
          if (b + u * NT + threadIdx.x < n[k]) v[k] += x[k][u];
    }
    wg_sum_all<NT, K>(v, lds);
  }
};
template <int NT, int K, uint32_t U = 4>
__device__ inline void sum_parts_all(const double *const (&src)[K], const uint32_t (&n)[K], double (&v)[K], double *lds) {
  PartSums<NT, K, U> ps;
  ps.load(src, n);
  ps.sum(src, n, v, lds);
}

}  // namespace nguo
