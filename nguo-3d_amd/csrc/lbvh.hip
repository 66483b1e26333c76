// LBVH build kernels: Morton keys, stable LSD radix sort, Karras internal nodes, bottom-up
// AABB union. See lbvh.h for the reference stages each replaces.
#include <algorithm>

#include "device_common.h"
#include "lbvh.h"

namespace nguo {

// ---------------------------------------------------------------- radix sort
// LSD radix sort, 8-bit digits, stable. A pass over tiles of ITEMS x 256 items:
//   k_radix_hist     per-tile digit counts, stored digit-major (hist[d * tiles + tile]);
//   scan             exclusive scan of those counters = each tile's first output slot per
//                    digit (one workgroup for small inputs, three launches beyond that);
//   k_radix_scatter  ranks every item among the equal digits before it in the tile (chunks of
//                    256 in order, 8-ballot wavefront multisplit), stages the tile in LDS in
//                    (digit, rank) order, then writes it out so that consecutive threads write
//                    consecutive slots of one digit's run.
// Scratch for at least this many items from the first sort on: the per-step sorts (stencil
// and incidence keys) change size every step.
constexpr uint32_t SORT_MIN_CAP = 1u << 14;

template <class K, uint32_t ITEMS>
__global__ __launch_bounds__(BLOCK) void k_radix_hist(uint32_t n, const K *__restrict__ keys, int shift, uint32_t *hist, uint32_t nblocks) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * ITEMS * BLOCK;
  for (uint32_t c = 0; c < ITEMS; ++c) {
    uint32_t i = base + c * BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of one value per thread over a workgroup of NT threads (fixed order); returns
// the prefix, *total receives the workgroup sum.
template <uint32_t NT = BLOCK>
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *wsum /* [NT / WAVE] */, uint32_t *total) {
  const uint32_t lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    uint32_t y = __shfl_up(x, o, WAVE);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == WAVE - 1) wsum[w] = x;
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t q = 0; q < NT / WAVE; ++q) {
    if (q < w) before += wsum[q];
    all += wsum[q];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

// In-place exclusive scan of up to SCAN_EXCL_MAX counters by one workgroup: staged in LDS
// with coalesced loads, each thread sums a contiguous segment, the segment sums are scanned
// across the workgroup by wavefront shuffles, and each thread writes its segment's prefixes.
// 1024 threads: a quarter of the serial segment per thread that 256 would leave.
constexpr uint32_t SCAN_EXCL_MAX = 16384, SCAN_EXCL_THREADS = 1024;
__global__ __launch_bounds__(SCAN_EXCL_THREADS) void k_scan_excl(uint32_t total, uint32_t *a) {
  constexpr uint32_t NT = SCAN_EXCL_THREADS;
  __shared__ uint32_t buf[SCAN_EXCL_MAX];
  __shared__ uint32_t wsum[NT / WAVE];
  for (uint32_t i = threadIdx.x; i < total; i += NT) buf[i] = a[i];
  __syncthreads();
  const uint32_t seg = (total + NT - 1) / NT;
  const uint32_t lo = min(total, threadIdx.x * seg), hi = min(total, lo + seg);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; ++i) s += buf[i];
  uint32_t all;
  uint32_t acc = block_excl_scan<NT>(s, wsum, &all);
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t v = buf[i];
    buf[i] = acc;
    acc += v;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < total; i += NT) a[i] = buf[i];
}

// Large scans: each workgroup scans SCAN_TILE counters in place (exclusive, local) and
// writes its sum; the sums are scanned by one workgroup; each workgroup adds its offset.
__global__ __launch_bounds__(BLOCK) void k_scan_tiles(uint32_t total, uint32_t *a, uint32_t *sums) {
  __shared__ uint32_t wsum[BLOCK / WAVE];
  const uint32_t base = blockIdx.x * SCAN_TILE;
  uint32_t run = 0;
  for (uint32_t c = 0; c < SCAN_TILE / BLOCK; ++c) {
    const uint32_t i = base + c * BLOCK + threadIdx.x;
    const uint32_t v = i < total ? a[i] : 0u;
    uint32_t t;
    const uint32_t e = block_excl_scan(v, wsum, &t);
    if (i < total) a[i] = run + e;
    run += t;
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = run;
}
__global__ __launch_bounds__(BLOCK) void k_scan_add(uint32_t total, uint32_t *a, const uint32_t *sums) {
  const uint32_t off = sums[blockIdx.x], base = blockIdx.x * SCAN_TILE;
  for (uint32_t c = 0; c < SCAN_TILE / BLOCK; ++c) {
    const uint32_t i = base + c * BLOCK + threadIdx.x;
    if (i < total) a[i] += off;
  }
}

template <class K, uint32_t ITEMS>
__global__ __launch_bounds__(BLOCK) void k_radix_scatter(uint32_t n, const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                        K *__restrict__ kout, uint32_t *__restrict__ vout, int shift,
                                                        const uint32_t *__restrict__ hist, uint32_t nblocks) {
  constexpr uint32_t T = ITEMS * BLOCK;
  __shared__ K sk[T];
  __shared__ uint32_t sv[T];
  __shared__ uint32_t gbase[256], lstart[256], dcount[256];
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t wsum[BLOCK / WAVE];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t tile0 = blockIdx.x * T;
  gbase[tid] = hist[(size_t)tid * nblocks + blockIdx.x];
  dcount[tid] = 0;
  const unsigned long long lt = (1ull << lane) - 1ull;
  K key[ITEMS];
  uint32_t val[ITEMS], loc[ITEMS];
#pragma unroll
  for (uint32_t c = 0; c < ITEMS; ++c) {
    const uint32_t i = tile0 + c * BLOCK + tid;
    const bool valid = i < n;
    key[c] = valid ? kin[i] : K(0);
    val[c] = valid ? vin[i] : 0u;
  }
  for (uint32_t c = 0; c < ITEMS; ++c) {
    for (int q = 0; q < 4; ++q) wcnt[q][tid] = 0;
    __syncthreads();
    const bool valid = tile0 + c * BLOCK + tid < n;
    const uint32_t digit = (uint32_t)(key[c] >> shift) & 255u;
    unsigned long long mask = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
      const bool bit = (digit >> b) & 1u;
      const unsigned long long bal = __ballot(bit);
      mask &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(mask & lt);
    if (valid && rank == 0) wcnt[w][digit] = (uint32_t)__popcll(mask);
    __syncthreads();
    {
      uint32_t sacc = dcount[tid];
      for (int q = 0; q < 4; ++q) { uint32_t t = wcnt[q][tid]; wcnt[q][tid] = sacc; sacc += t; }
      dcount[tid] = sacc;
    }
    __syncthreads();
    loc[c] = valid ? wcnt[w][digit] + rank : 0xffffffffu;
    __syncthreads();  // every loc read before the next chunk clears wcnt
  }
  // tile-local start of every digit's run
  uint32_t tot;
  lstart[tid] = block_excl_scan(dcount[tid], wsum, &tot);
  __syncthreads();
#pragma unroll
  for (uint32_t c = 0; c < ITEMS; ++c) {
    if (loc[c] != 0xffffffffu) {
      const uint32_t p = lstart[(uint32_t)(key[c] >> shift) & 255u] + loc[c];
      sk[p] = key[c];
      sv[p] = val[c];
    }
  }
  __syncthreads();
  const uint32_t cnt = min(T, n - tile0);
#pragma unroll
  for (uint32_t c = 0; c < ITEMS; ++c) {
    const uint32_t p = c * BLOCK + tid;
    if (p < cnt) {
      const K k = sk[p];
      const uint32_t d = (uint32_t)(k >> shift) & 255u;
      const uint32_t g = gbase[d] + (p - lstart[d]);
      kout[g] = k;
      vout[g] = sv[p];
    }
  }
}

template <class K, uint32_t ITEMS>
static void radix_passes(uint32_t n, K *keys, uint32_t *vals, K *ktmp, uint32_t *vtmp, uint32_t *hist, uint32_t *sums,
                         const std::vector<int> &shifts, hipStream_t s) {
  const uint32_t nblocks = div_up(n, ITEMS * BLOCK), total = 256u * nblocks;
  K *ka = keys, *kb = ktmp;
  uint32_t *va = vals, *vb = vtmp;
  for (int sh : shifts) {
    hipLaunchKernelGGL((k_radix_hist<K, ITEMS>), dim3(nblocks), dim3(BLOCK), 0, s, n, (const K *)ka, sh, hist, nblocks);
    if (total <= SCAN_EXCL_MAX) {
      hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(SCAN_EXCL_THREADS), 0, s, total, hist);
    } else {
      const uint32_t nt = div_up(total, SCAN_TILE);
      NGUO_REQUIRE(nt <= SCAN_EXCL_MAX, NGUO_E_CAPACITY, "radix sort: too many tiles for the tile-sum scan");
      hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(BLOCK), 0, s, total, hist, sums);
      hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(SCAN_EXCL_THREADS), 0, s, nt, sums);
      hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(BLOCK), 0, s, total, hist, (const uint32_t *)sums);
    }
    hipLaunchKernelGGL((k_radix_scatter<K, ITEMS>), dim3(nblocks), dim3(BLOCK), 0, s, n, (const K *)ka, (const uint32_t *)va, kb, vb, sh,
                       (const uint32_t *)hist, nblocks);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != keys) {
    NGUO_HIP(hipMemcpyAsync(keys, ka, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s));
    NGUO_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  }
  NGUO_HIP(hipGetLastError());
}

template <class K>
static void radix_sort(uint32_t n, K *keys, uint32_t *vals, K *ktmp, uint32_t *vtmp, DevBuf<uint32_t> &hist, DevBuf<uint32_t> &sums,
                       const std::vector<int> &shifts, hipStream_t s) {
  if (n <= 1) return;
  const uint32_t items = n >= LARGE_SORT ? LARGE_ITEMS : SMALL_ITEMS;
  const uint32_t total = 256u * div_up(n, items * BLOCK);
  hist.reserve(std::max<size_t>(total, 256u * div_up(SORT_MIN_CAP, items * BLOCK)));
  sums.reserve(div_up(total, SCAN_TILE) + 1);
  if (items == LARGE_ITEMS) radix_passes<K, LARGE_ITEMS>(n, keys, vals, ktmp, vtmp, hist.p, sums.p, shifts, s);
  else radix_passes<K, SMALL_ITEMS>(n, keys, vals, ktmp, vtmp, hist.p, sums.p, shifts, s);
}

void radix_sort_u32_scratch(uint32_t n, uint32_t *keys, uint32_t *vals, uint32_t *scratch, const std::vector<int> &shifts, hipStream_t s) {
  if (n <= 1) return;
  uint32_t *ktmp = scratch, *vtmp = scratch + n, *hist = vtmp + n, *sums = hist + radix_counters(n);
  if (radix_items(n) == LARGE_ITEMS) radix_passes<uint32_t, LARGE_ITEMS>(n, keys, vals, ktmp, vtmp, hist, sums, shifts, s);
  else radix_passes<uint32_t, SMALL_ITEMS>(n, keys, vals, ktmp, vtmp, hist, sums, shifts, s);
}

void RadixSort::sort_u32(uint32_t n, uint32_t *keys, uint32_t *vals, const std::vector<int> &shifts, hipStream_t s) {
  if (n <= 1) return;
  k32_.reserve(std::max(n, SORT_MIN_CAP));
  v_.reserve(std::max(n, SORT_MIN_CAP));
  radix_sort<uint32_t>(n, keys, vals, k32_.p, v_.p, hist_, sums_, shifts, s);
}

void RadixSort::reserve_u64(uint32_t n) {
  n = std::max(n, SORT_MIN_CAP);
  k64_.reserve(n);
  v_.reserve(n);
  // the small-tile passes use the most counters
  hist_.reserve(256u * (size_t)div_up(n, SMALL_ITEMS * BLOCK));
  sums_.reserve(div_up(256u * (size_t)div_up(n, SMALL_ITEMS * BLOCK), SCAN_TILE) + 1);
}

void RadixSort::sort_u64(uint32_t n, uint64_t *keys, uint32_t *vals, const std::vector<int> &shifts, hipStream_t s) {
  if (n <= 1) return;
  k64_.reserve(std::max(n, SORT_MIN_CAP));
  v_.reserve(std::max(n, SORT_MIN_CAP));
  radix_sort<uint64_t>(n, keys, vals, k64_.p, v_.p, hist_, sums_, shifts, s);
}

// ---------------------------------------------------------------- Morton
__device__ inline uint32_t expand_bits(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
// `as u32` of the reference: truncation, saturating, NaN -> 0.
__device__ inline uint32_t f2u_sat(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}
__device__ inline uint32_t morton_q(float v) { return f2u_sat(clampv(clampv(v, 0.0f, 1.0f) * 1023.0f, 0.0f, 1023.0f)); }

__device__ inline f3 prim_centroid(int kind, uint32_t p, const uint32_t *idx, const float *pos) {
  if (kind == 0) {
    f3 a = ld3(pos, idx[3 * (size_t)p]), b = ld3(pos, idx[3 * (size_t)p + 1]), c = ld3(pos, idx[3 * (size_t)p + 2]);
    return (a + b + c) / 3.0f;
  }
  f3 a = ld3(pos, idx[2 * (size_t)p]), b = ld3(pos, idx[2 * (size_t)p + 1]);
  return (a + b) / 2.0f;
}

// Per-workgroup min/max of the centroids; bounds[6 * block] = lo.xyz, hi.xyz.
__global__ __launch_bounds__(BLOCK) void k_centroid_bounds(int kind, uint32_t n, const uint32_t *ids, const uint32_t *idx, const float *pos,
                                                          float *cent, float *part) {
  __shared__ float lds[6][4];
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  f3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  if (i < n) {
    f3 c = prim_centroid(kind, ids ? ids[i] : i, idx, pos);
    st3(cent, i, c);
    lo = c;
    hi = c;
  }
  float v[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
  for (int k = 0; k < 3; ++k) { v[k] = wave_min(v[k]); v[3 + k] = wave_max(v[3 + k]); }
  int w = threadIdx.x / WAVE;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; ++k) lds[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float r = lds[k][0];
    for (int q = 1; q < 4; ++q) r = k < 3 ? fminf(r, lds[k][q]) : fmaxf(r, lds[k][q]);
    part[6 * (size_t)blockIdx.x + k] = r;
  }
}

__global__ __launch_bounds__(BLOCK) void k_bounds_final(uint32_t nparts, float *part) {
  __shared__ float lds[6][4];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = threadIdx.x; i < nparts; i += BLOCK)
    for (int k = 0; k < 6; ++k) v[k] = k < 3 ? fminf(v[k], part[6 * (size_t)i + k]) : fmaxf(v[k], part[6 * (size_t)i + k]);
  for (int k = 0; k < 3; ++k) { v[k] = wave_min(v[k]); v[3 + k] = wave_max(v[3 + k]); }
  int w = threadIdx.x / WAVE;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; ++k) lds[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    float r = lds[k][0];
    for (int q = 1; q < 4; ++q) r = k < 3 ? fminf(r, lds[k][q]) : fmaxf(r, lds[k][q]);
    part[k] = r;  // final lo/hi in part[0..6)
  }
}

// engine-gpu/src/morton.rs:22-58: code of ((c - lo) * inv), inv = 1/extent (0 if flat).
__global__ __launch_bounds__(BLOCK) void k_morton(uint32_t n, const float *cent, const float *bounds, uint32_t *codes, uint32_t *order) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  f3 lo{bounds[0], bounds[1], bounds[2]}, hi{bounds[3], bounds[4], bounds[5]};
  f3 ext = hi - lo;
  f3 inv{ext.x > 0.0f ? 1.0f / ext.x : 0.0f, ext.y > 0.0f ? 1.0f / ext.y : 0.0f, ext.z > 0.0f ? 1.0f / ext.z : 0.0f};
  f3 d = ld3(cent, i) - lo;
  f3 nrm{d.x * inv.x, d.y * inv.y, d.z * inv.z};
  codes[i] = expand_bits(morton_q(nrm.x)) * 4u + expand_bits(morton_q(nrm.y)) * 2u + expand_bits(morton_q(nrm.z));
  if (order) order[i] = i;
}

void morton_codes(int kind, uint32_t n, const uint32_t *ids, const uint32_t *idx, const float *pos, float *bounds_scratch, float *cent_scratch,
                  uint32_t *codes, hipStream_t s) {
  if (n == 0) return;
  unsigned g = div_up(n, BLOCK);
  hipLaunchKernelGGL(k_centroid_bounds, dim3(g), dim3(BLOCK), 0, s, kind, n, ids, idx, pos, cent_scratch, bounds_scratch);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(BLOCK), 0, s, (uint32_t)g, bounds_scratch);
  hipLaunchKernelGGL(k_morton, dim3(g), dim3(BLOCK), 0, s, n, (const float *)cent_scratch, (const float *)bounds_scratch, codes, (uint32_t *)nullptr);
  NGUO_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- Karras
__device__ inline int kdelta(const uint32_t *codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint32_t x = codes[i] ^ codes[j];
  if (x == 0) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
  return __clz((int)x);
}

// engine-gpu/src/karras.rs:24-88 (Karras 2012, section 4).
__global__ __launch_bounds__(BLOCK) void k_karras(uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right, uint32_t *parent,
                                                 uint32_t *range) {
  int i = (int)(blockIdx.x * BLOCK + threadIdx.x);
  int nn = (int)n;
  if (i >= nn - 1) return;
  int d = (kdelta(codes, nn, i, i + 1) - kdelta(codes, nn, i, i - 1)) < 0 ? -1 : 1;
  int dmin = kdelta(codes, nn, i, i - d);
  int lmax = 2;
  while (kdelta(codes, nn, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(codes, nn, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int first = min(i, j), last = max(i, j);
  int nd = kdelta(codes, nn, first, last);
  int split = first, step = last - first;
  while (true) {
    step = (step + 1) / 2;
    int ns = split + step;
    if (ns < last && kdelta(codes, nn, first, ns) > nd) split = ns;
    if (step <= 1) break;
  }
  int gamma = split;
  uint32_t lc = first == gamma ? ((uint32_t)gamma | LEAF) : (uint32_t)gamma;
  uint32_t rc = last == gamma + 1 ? ((uint32_t)(gamma + 1) | LEAF) : (uint32_t)(gamma + 1);
  left[i] = lc;
  right[i] = rc;
  if (range) { range[2 * i] = (uint32_t)first; range[2 * i + 1] = (uint32_t)last; }
  if (parent) {
    // parent[0 .. n-1) internal, parent[n-1 + leaf] leaves
    parent[(lc & LEAF) ? (n - 1 + (lc & ~LEAF)) : lc] = (uint32_t)i;
    parent[(rc & LEAF) ? (n - 1 + (rc & ~LEAF)) : rc] = (uint32_t)i;
  }
}

void karras_build(uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right, hipStream_t s) {
  if (n < 2) return;
  hipLaunchKernelGGL(k_karras, dim3(div_up(n - 1, BLOCK)), dim3(BLOCK), 0, s, n, codes, left, right, (uint32_t *)nullptr,
                     (uint32_t *)nullptr);
  NGUO_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- AABB
__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t n, uint32_t *a, uint32_t v) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) a[i] = v;
}

// Leaf boxes in sorted order: leaf_box[6 i] = AABB of primitive prim[i].
__global__ __launch_bounds__(BLOCK) void k_leaf_boxes(int kind, uint32_t n, const uint32_t *prim, const uint32_t *idx, const float *pos, float *box) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  uint32_t p = prim[i];
  f3 lo, hi;
  if (kind == 0) {
    f3 a = ld3(pos, idx[3 * (size_t)p]), b = ld3(pos, idx[3 * (size_t)p + 1]), c = ld3(pos, idx[3 * (size_t)p + 2]);
    lo = vmin(vmin(a, b), c);
    hi = vmax(vmax(a, b), c);
  } else {
    f3 a = ld3(pos, idx[2 * (size_t)p]), b = ld3(pos, idx[2 * (size_t)p + 1]);
    lo = vmin(a, b);
    hi = vmax(a, b);
  }
  float *o = box + 6 * (size_t)i;
  o[0] = lo.x; o[1] = lo.y; o[2] = lo.z; o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
}

// Bottom-up union for a general binary tree (one thread per leaf climbs; the second arrival
// at a node merges). Used where leaf ranges are not contiguous.
__global__ __launch_bounds__(BLOCK) void k_aabb_climb(uint32_t n, const uint32_t *left, const uint32_t *right, const uint32_t *parent,
                                                     uint32_t *visit, const float *leaf_box, float *node_box) {
  uint32_t li = blockIdx.x * BLOCK + threadIdx.x;
  if (li >= n) return;
  uint32_t node = parent[n - 1 + li];
  while (true) {
    __threadfence();
    uint32_t prev = atomicAdd(&visit[node], 1u);
    if (prev == 0) return;  // the sibling subtree is not finished; it will continue
    __threadfence();
    uint32_t ch[2] = {left[node], right[node]};
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 2; ++c) {
      const float *src = (ch[c] & LEAF) ? leaf_box + 6 * (size_t)(ch[c] & ~LEAF) : node_box + 6 * (size_t)ch[c];
      float v[6];
      for (int k = 0; k < 6; ++k) v[k] = src[k];
      for (int k = 0; k < 3; ++k) { b[k] = fminf(b[k], v[k]); b[3 + k] = fmaxf(b[3 + k], v[3 + k]); }
    }
    float *o = node_box + 6 * (size_t)node;
    for (int k = 0; k < 6; ++k) o[k] = b[k];
    if (node == 0) return;
    node = parent[node];
  }
}

// Node boxes from leaf ranges. The leaves of a Karras node form one contiguous range
// [first, last] of the sorted order, so its box is the union of those leaf boxes. A sparse
// table level k holds the union of the 2^k leaves starting at each index; a node's box is
// the union of two (overlapping) level-k entries covering its range. Min and max are exact,
// so the result equals the bottom-up climb's, without cross-workgroup fences on the
// critical path: about log2(n) short launches.
__global__ __launch_bounds__(BLOCK) void k_union_level(uint32_t count, uint32_t half, const float *in, float *out) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= count) return;
  const float *a = in + 6 * (size_t)i, *b = in + 6 * ((size_t)i + half);
  float *o = out + 6 * (size_t)i;
  for (int k = 0; k < 3; ++k) { o[k] = fminf(a[k], b[k]); o[3 + k] = fmaxf(a[3 + k], b[3 + k]); }
}

__global__ __launch_bounds__(BLOCK) void k_node_boxes(uint32_t n, const uint32_t *range, const float *leaf_box, const float *table,
                                                     float *node_box) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i + 1 >= n) return;
  const uint32_t a = range[2 * i], b = range[2 * i + 1], len = b - a + 1;
  const int k = 31 - __clz(len);
  const float *lv = k == 0 ? leaf_box : table + (size_t)(k - 1) * 6 * n;
  const float *x = lv + 6 * (size_t)a, *y = lv + 6 * ((size_t)b + 1 - (1u << k));
  float *o = node_box + 6 * (size_t)i;
  for (int c = 0; c < 3; ++c) { o[c] = fminf(x[c], y[c]); o[3 + c] = fmaxf(x[3 + c], y[3 + c]); }
}

void aabb_from_ranges(uint32_t n, const uint32_t *range, const float *leaf_box, float *node_box, DevBuf<float> &table, hipStream_t s) {
  if (n < 2) return;
  int levels = 0;
  while ((2u << levels) <= n) ++levels;  // floor(log2 n)
  table.reserve((size_t)std::max(levels, 1) * 6 * n);
  for (int k = 1; k <= levels; ++k) {
    const uint32_t half = 1u << (k - 1), count = n - (1u << k) + 1;
    const float *in = k == 1 ? leaf_box : table.p + (size_t)(k - 2) * 6 * n;
    hipLaunchKernelGGL(k_union_level, dim3(div_up(count, BLOCK)), dim3(BLOCK), 0, s, count, half, in, table.p + (size_t)(k - 1) * 6 * n);
  }
  hipLaunchKernelGGL(k_node_boxes, dim3(div_up(n - 1, BLOCK)), dim3(BLOCK), 0, s, n, range, leaf_box, (const float *)table.p, node_box);
  NGUO_HIP(hipGetLastError());
}

// Traversal records: internal node i's up to four grandchildren (a child that is a leaf
// stands for itself) in 192 bytes, entry e as (lo.xyz, hi.x) (hi.yz, ref, last leaf slot)
// (a leaf's vertex ids). ref is LEAF | the primitive id for a leaf, else the binary node id,
// whose own record holds its grandchildren; an unused entry has ref TRAV_NONE. A traversal
// then tests four boxes per dependent record load, so its chain of loads is half as long as
// over the binary nodes, and a leaf it meets loads its positions at once.
struct TravSrc {
  int kind;  // 0: triangles (3 vertex ids per primitive), 1: edges (2)
  const uint32_t *idx, *range, *prim;
  const float *leaf_box, *node_box;
};
__device__ inline void trav_entry(uint32_t c, const TravSrc &S, float4 &e0, float4 &e1, float4 &e2) {
  const bool leaf = c & LEAF;
  const float *b = leaf ? S.leaf_box + 6 * (size_t)(c & ~LEAF) : S.node_box + 6 * (size_t)c;
  const uint32_t last = leaf ? c & ~LEAF : S.range[2 * (size_t)c + 1];
  const uint32_t p = leaf ? S.prim[c & ~LEAF] : 0u;
  const uint32_t ref = leaf ? (LEAF | p) : c;
  uint32_t v[3] = {0u, 0u, 0u};
  if (leaf) {
    const uint32_t k = S.kind == 0 ? 3u : 2u;
    for (uint32_t i = 0; i < k; ++i) v[i] = S.idx[k * (size_t)p + i];
  }
  e0 = make_float4(b[0], b[1], b[2], b[3]);
  e1 = make_float4(b[4], b[5], __uint_as_float(ref), __uint_as_float(last));
  e2 = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), 0.0f);
}
__global__ __launch_bounds__(BLOCK) void k_pack_trav(uint32_t n, const uint32_t *left, const uint32_t *right, TravSrc S, float4 *trav) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i + 1 >= n) return;
  float4 e[TRAV_ENTRY * TRAV_WIDTH];
  uint32_t k = 0;
  const uint32_t ch[2] = {left[i], right[i]};
  for (int c = 0; c < 2; ++c) {
    if (ch[c] & LEAF) {
      trav_entry(ch[c], S, e[3 * k], e[3 * k + 1], e[3 * k + 2]);
      ++k;
    } else {
      trav_entry(left[ch[c]], S, e[3 * k], e[3 * k + 1], e[3 * k + 2]);
      trav_entry(right[ch[c]], S, e[3 * k + 3], e[3 * k + 4], e[3 * k + 5]);
      k += 2;
    }
  }
  for (; k < TRAV_WIDTH; ++k) {
    e[3 * k] = make_float4(0.f, 0.f, 0.f, 0.f);
    e[3 * k + 1] = make_float4(0.f, 0.f, __uint_as_float(TRAV_NONE), __uint_as_float(0u));
    e[3 * k + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  static_assert(TRAV_ENTRY == 3, "three float4 per entry");
  float4 *o = trav + TRAV_ENTRY * TRAV_WIDTH * (size_t)i;
  for (uint32_t q = 0; q < TRAV_ENTRY * TRAV_WIDTH; ++q) o[q] = e[q];
}

// slot[k] = the sorted leaf slot holding element k of the build's primitive list.
__global__ __launch_bounds__(BLOCK) void k_slot_of(uint32_t n, const uint32_t *order, uint32_t *slot) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) slot[order[i]] = i;
}

// The traversal frontier: starting from the root's children, internal entries are replaced by
// their two children breadth first while there is room for TRAV_PARTS entries; leaves stay.
// One thread, at most TRAV_PARTS - 2 node expansions.
__global__ void k_frontier(uint32_t n, const uint32_t *left, const uint32_t *right, TravSrc S, float4 *front) {
  if (threadIdx.x != 0) return;
  uint32_t ref[TRAV_PARTS];
  float4 lo[TRAV_PARTS], hi[TRAV_PARTS], vid[TRAV_PARTS];  // hi.w: the entry's last leaf slot
  uint32_t cnt = 0;
  auto entry = [&](uint32_t c, uint32_t o) {
    float4 e0, e1, e2;
    trav_entry(c, S, e0, e1, e2);
    ref[o] = __float_as_uint(e1.z);
    lo[o] = make_float4(e0.x, e0.y, e0.z, 0.0f);
    hi[o] = make_float4(e0.w, e1.x, e1.y, e1.w);
    vid[o] = e2;
  };
  if (n == 1) {
    entry(LEAF | 0u, 0);
    cnt = 1;
  } else if (n >= 2) {
    auto children = [&](uint32_t node, uint32_t o0, uint32_t o1) {
      entry(left[node], o0);
      entry(right[node], o1);
    };
    children(0u, 0, 1);
    cnt = 2;
    bool grew = true;
    while (grew && cnt < TRAV_PARTS) {
      grew = false;
      const uint32_t level = cnt;
      for (uint32_t e = 0; e < level && cnt < TRAV_PARTS; ++e) {
        if (ref[e] & LEAF) continue;
        // entry e becomes the left child, the right child goes to the end
        children(ref[e], e, cnt);
        ++cnt;
        grew = true;
      }
    }
  }
  for (uint32_t k = 0; k < TRAV_PARTS; ++k) {
    const bool used = k < cnt;
    front[TRAV_ENTRY * k] = used ? make_float4(lo[k].x, lo[k].y, lo[k].z, __uint_as_float(ref[k])) : make_float4(0.f, 0.f, 0.f, __uint_as_float(0xffffffffu));
    front[TRAV_ENTRY * k + 1] = used ? hi[k] : make_float4(0.f, 0.f, 0.f, __uint_as_float(0u));
    front[TRAV_ENTRY * k + 2] = used ? vid[k] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

void aabb_build(uint32_t n, const uint32_t *left, const uint32_t *right, uint32_t *parent, uint32_t *visit, const float *leaf_box, float *node_box,
                hipStream_t s) {
  if (n < 2) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, s, n, visit, 0u);
  hipLaunchKernelGGL(k_aabb_climb, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, s, n, left, right, (const uint32_t *)parent, visit, leaf_box, node_box);
  NGUO_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- build
__global__ __launch_bounds__(BLOCK) void k_gather_prim(uint32_t n, const uint32_t *order, const uint32_t *ids, uint32_t *prim) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) prim[i] = ids ? ids[order[i]] : order[i];
}

void Lbvh::refit(int kind, const uint32_t *idx, const float *pos, hipStream_t s) {
  const uint32_t n = n_;
  if (n == 0) return;
  hipLaunchKernelGGL(k_leaf_boxes, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, s, kind, n, (const uint32_t *)prim_.p, idx, pos, leaf_box_.p);
  const TravSrc src{kind, idx, range_.p, prim_.p, leaf_box_.p, node_box_.p};
  if (n >= 2) {
    aabb_from_ranges(n, range_.p, leaf_box_.p, node_box_.p, table_, s);
    hipLaunchKernelGGL(k_pack_trav, dim3(div_up(n - 1, BLOCK)), dim3(BLOCK), 0, s, n, (const uint32_t *)left_.p, (const uint32_t *)right_.p, src,
                       (float4 *)trav_.p);
  }
  hipLaunchKernelGGL(k_frontier, dim3(1), dim3(WAVE), 0, s, n, (const uint32_t *)left_.p, (const uint32_t *)right_.p, src, (float4 *)front_.p);
  NGUO_HIP(hipGetLastError());
}

void Lbvh::build(int kind, uint32_t n, const uint32_t *ids, const uint32_t *idx, const float *pos, hipStream_t s) {
  n_ = n;
  if (n == 0) return;
  unsigned g = div_up(n, BLOCK);
  keys_.reserve(n);
  prim_.reserve(n);
  cent_.reserve(3 * (size_t)n);
  bounds_.reserve(6 * (size_t)g + 6);
  leaf_box_.reserve(6 * (size_t)n);
  left_.reserve(n);
  right_.reserve(n);
  range_.reserve(2 * (size_t)n);
  visit_.reserve(n);
  node_box_.reserve(6 * (size_t)n);
  DevBuf<uint32_t> &order = visit_;  // reused as the sort payload before the climb
  hipLaunchKernelGGL(k_centroid_bounds, dim3(g), dim3(BLOCK), 0, s, kind, n, ids, idx, pos, cent_.p, bounds_.p);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(BLOCK), 0, s, (uint32_t)g, bounds_.p);
  hipLaunchKernelGGL(k_morton, dim3(g), dim3(BLOCK), 0, s, n, (const float *)cent_.p, (const float *)bounds_.p, keys_.p, order.p);
  sorter.sort_u32(n, keys_.p, order.p, {0, 8, 16, 24}, s);
  hipLaunchKernelGGL(k_gather_prim, dim3(g), dim3(BLOCK), 0, s, n, (const uint32_t *)order.p, ids, prim_.p);
  slot_.reserve(n);
  hipLaunchKernelGGL(k_slot_of, dim3(g), dim3(BLOCK), 0, s, n, (const uint32_t *)order.p, slot_.p);
  hipLaunchKernelGGL(k_leaf_boxes, dim3(g), dim3(BLOCK), 0, s, kind, n, (const uint32_t *)prim_.p, idx, pos, leaf_box_.p);
  const TravSrc src{kind, idx, range_.p, prim_.p, leaf_box_.p, node_box_.p};
  if (n >= 2) {
    hipLaunchKernelGGL(k_karras, dim3(div_up(n - 1, BLOCK)), dim3(BLOCK), 0, s, n, (const uint32_t *)keys_.p, left_.p, right_.p,
                       (uint32_t *)nullptr, range_.p);
    aabb_from_ranges(n, range_.p, leaf_box_.p, node_box_.p, table_, s);
    trav_.reserve(4 * TRAV_ENTRY * TRAV_WIDTH * (size_t)n);
    hipLaunchKernelGGL(k_pack_trav, dim3(div_up(n - 1, BLOCK)), dim3(BLOCK), 0, s, n, (const uint32_t *)left_.p, (const uint32_t *)right_.p, src,
                       (float4 *)trav_.p);
  }
  front_.reserve(4 * TRAV_ENTRY * TRAV_PARTS);
  hipLaunchKernelGGL(k_frontier, dim3(1), dim3(WAVE), 0, s, n, (const uint32_t *)left_.p, (const uint32_t *)right_.p, src, (float4 *)front_.p);
  NGUO_HIP(hipGetLastError());
}

}  // namespace nguo
