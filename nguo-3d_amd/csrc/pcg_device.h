// Device helpers of the PCG iteration kernels (k_pcg_spmv and the block-Jacobi / additive
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

struct PreArgs {
  ContactView C;
  const float4 *z, *pold;
  const double *cur;
  uint32_t nz;
  double *scal;
  uint32_t *flag;
};
struct SpmvArgs {
  SellView A;
  ContactView C;
  const float4 *z, *pold;
  float4 *pnew, *q;
  double *part;
  const double *scal;
  const uint32_t *flag;
};
struct SchArgs {
  uint32_t nd;
  const uint32_t *dom_ptr, *dom_dof;
  const float *ainv;
  float *r, *z, *dsol, *p;
  const float *q;
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

// v[k] = sum_{i < n[k]} src[k][i] for K partial arrays, every thread gets the sums. Each
// thread adds its strided share in index order; the loads of up to 4 strides are issued
// before the adds, so a launch with up to 4 NT partials per array pays one memory round trip.
template <int NT, int K>
__device__ inline void sum_parts_all(const double *const (&src)[K], const uint32_t (&n)[K], double (&v)[K], double *lds) {
  constexpr uint32_t U = 4;
  uint32_t nmax = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) { v[k] = 0.0; nmax = n[k] > nmax ? n[k] : nmax; }
  for (uint32_t b = 0; b < nmax; b += U * NT) {
    double x[K][U];
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

}  // namespace nguo
