// Device-side building blocks shared by the kernels of libnguo_hip: wavefront (64-lane)
// and workgroup reductions with a fixed combination order (bitwise-deterministic results
// for a fixed launch shape), and the SELL-64 3x3-block matrix view the PCG runs on.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nguo {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;  // 4 wavefronts per workgroup

template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

// Lane exchanges inside a row of 16 lanes (DPP, no LDS traffic). Controls used here: quad
// permutations 0xB1 (lane ^ 1) and 0x4E (lane ^ 2), 0x141 the half-row mirror (lane i of
// each 8 with lane 7 - i), 0x140 the row mirror (lane i with 15 - i). All lanes must be active.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;
template <int CTRL> __device__ inline float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ inline double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double readlane_d(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane), __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes of a wavefront, every lane gets it (all lanes active). Four
// symmetric exchanges inside each row of 16 (lane ^ 1, lane ^ 2, half-row mirror, row
// mirror) leave the row's sum in all of its lanes: both partners of an exchange add the same
// two values, and a + b == b + a bitwise. The four row sums are then read from lanes 0, 16,
// 32 and 48 and added in a fixed order.
__device__ inline double wave_sum(double v) {
  v += dpp_d<DPP_XOR1>(v);
  v += dpp_d<DPP_XOR2>(v);
  v += dpp_d<DPP_HALF_MIRROR>(v);
  v += dpp_d<DPP_MIRROR>(v);
  return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}
__device__ inline float wave_sum(float v) {
  v += dpp_f<DPP_XOR1>(v);
  v += dpp_f<DPP_XOR2>(v);
  v += dpp_f<DPP_HALF_MIRROR>(v);
  v += dpp_f<DPP_MIRROR>(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return (r0 + r1) + (r2 + r3);
}
template <class T> __device__ inline T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, WAVE); v = v > w ? v : w; }
  return v;
}
template <class T> __device__ inline T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { T w = __shfl_xor(v, o, WAVE); v = v < w ? v : w; }
  return v;
}

// Sum over a 256-thread workgroup; the result is valid in thread 0. Fixed order.
template <class T> __device__ inline T block_sum(T v, T *lds4) {
  v = wave_sum(v);
  int w = threadIdx.x / WAVE;
  if ((threadIdx.x & (WAVE - 1)) == 0) lds4[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

// A 3x3-block matrix in SELL-64 layout: block rows are grouped in slices of 64; inside a
// slice the k-th block of every row sits at slot slice_ptr[s] + 64k + lane, so one wavefront
// reads 64 consecutive column indices. The 9 values of the 64 blocks of a slot group are
// stored as three planes: entries 0-3 as 64 float4, entries 4-7 as 64 float4, entry 8 as
// 64 floats (576 floats, no padding), so a wavefront reads a block with two 16-byte and one
// 4-byte coalesced loads. Padding slots hold col = row and value 0.
struct SellView {
  uint32_t nb;                    // block rows (= free vertices)
  const uint32_t *slice_ptr;      // [nslices + 1], in slots (multiples of 64)
  const uint32_t *row_len;        // [nb] blocks in each row
  const uint32_t *col;            // [slots]
  const float *val;               // [slots * 9], plane layout (see val_index)
  uint32_t ell_w;                 // nonzero: every slice has this width (slice_ptr[s] = 64 s ell_w)
};

__host__ __device__ inline size_t sell_val_index(uint32_t slot, int e) {
  const size_t g = (size_t)(slot & ~63u) * 9u, lane = slot & 63u;
  return e < 4 ? g + 4 * lane + (size_t)e : (e < 8 ? g + 256 + 4 * lane + (size_t)(e - 4) : g + 512 + lane);
}

// The 9 values of the block in slot group `group` (a multiple of 64), lane `lane`.
struct BlockVals {
  float4 a, b;  // entries 0-3, 4-7 (row-major 3x3)
  float c;      // entry 8
};
__device__ inline BlockVals load_block(const float *val, uint32_t group, uint32_t lane) {
  const float *g = val + (size_t)group * 9u;
  BlockVals v;
  v.a = reinterpret_cast<const float4 *>(g)[lane];
  v.b = reinterpret_cast<const float4 *>(g + 256)[lane];
  v.c = g[512 + lane];
  return v;
}

// Matrix-free contact term of the Newton system. Stencil s couples its (up to) four
// vertices with H_s = hp_s g g^T + lambda_s (w w^T) (x) (I - n n^T) where g = [g_0..g_3]
// (distance gradient), w the proximity weights and n the contact normal.
struct ContactView {
  uint32_t ns;                    // stencils (0 = no contact term)
  const uint32_t *dof;            // [4 ns] free-vertex index or 0xffffffff (pinned)
  const float *g;                 // [12 ns]
  const float *w;                 // [4 ns]
  const float *coef;              // [2 ns] hp, lambda
  const float *n;                 // [3 ns]
  float *sums;                    // [4 ns] per-apply: g.p (1) and sum w p (3)
  const uint32_t *inc_ptr;        // [nb + 1] incidence of free vertices
  const uint32_t *inc;            // [..] 4 s + a
  const uint32_t *pos;            // [4 ns] incidence slot of participant (s, a), or ~0u (pinned)
  float4 *contrib;                // [..] per-apply: H_s p restricted to the incidence's vertex
};

}  // namespace nguo
