// C ABI of the LBVH stages on a triangle set, so each stage can be checked bit for bit:
// engine-gpu/src/morton.rs:61-230 (GpuMorton::codes), radix.rs:112-135 (sort_pairs),
// karras.rs:91-262 (GpuKarras::build_tree), bvh_aabb.rs (node AABB propagation).
#include <mutex>

#include "device_common.h"
#include "lbvh.h"

using namespace nguo;

// ---- the allocation-free sort's scratch: one file-scope device array, NGUO_NOALLOC_SORT_MAX
// pairs' worth (the key and value ping-pong halves, the digit counters of the small-tile
// passes, which need the most, and the scan's tile sums: 9.4 MB). Every GPU the library runs
// on holds its own copy until the process ends. In host code `scratch` names a host-side
// stand-in, never the device memory: the copy's device address comes from
// hipGetSymbolAddress, per device, and only that address reaches a kernel.
namespace {
constexpr size_t NOALLOC_WORDS =
    2 * (size_t)NGUO_NOALLOC_SORT_MAX + 256u * ((size_t)NGUO_NOALLOC_SORT_MAX + SMALL_ITEMS * BLOCK - 1) / (SMALL_ITEMS * BLOCK) +
    (256u * (((size_t)NGUO_NOALLOC_SORT_MAX + SMALL_ITEMS * BLOCK - 1) / (SMALL_ITEMS * BLOCK)) + SCAN_TILE - 1) / SCAN_TILE + 1;
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];
constexpr int NOALLOC_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[NOALLOC_DEVICES] = {};
const std::vector<int> kSortShifts{0, 8, 16, 24};  // built once, at load

// This device's copy of `scratch` (the device must be current). Looked up once per device.
uint32_t *scratch_for(int device) {
  NGUO_REQUIRE(device >= 0 && device < NOALLOC_DEVICES, NGUO_E_UNSUPPORTED, "allocation-free sort: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    NGUO_REQUIRE(e == hipSuccess && ptr, NGUO_E_DEVICE,
                 std::string("hipGetSymbolAddress of the sort scratch failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (uint32_t *)ptr;
  }
  return g_scratch_ptr[device];
}
}  // namespace

namespace nguo {
void noalloc_scratch_init(int device) { (void)scratch_for(device); }
}  // namespace nguo

extern "C" {

int nguo_lbvh_sort_pairs_noalloc(nguo_ctx *ctx, uint32_t n, uint32_t *keys_dev, uint32_t *vals_dev) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx, NGUO_E_INVALID, "nguo_lbvh_sort_pairs_noalloc: null context");
    NGUO_REQUIRE(n <= NGUO_NOALLOC_SORT_MAX, NGUO_E_CAPACITY,
                 "nguo_lbvh_sort_pairs_noalloc: " + std::to_string(n) + " pairs exceed the scratch maximum of " +
                     std::to_string(NGUO_NOALLOC_SORT_MAX));
    if (n <= 1) return NGUO_OK;
    NGUO_REQUIRE(keys_dev && vals_dev, NGUO_E_INVALID, "nguo_lbvh_sort_pairs_noalloc: null device buffer");
    static_assert(NOALLOC_WORDS >= radix_scratch_words(NGUO_NOALLOC_SORT_MAX), "scratch below the maximum input's need");
    NGUO_REQUIRE(radix_scratch_words(n) <= NOALLOC_WORDS, NGUO_E_CAPACITY, "nguo_lbvh_sort_pairs_noalloc: scratch too small");
    NGUO_HIP(hipSetDevice(ctx->device));
    uint32_t *buf = scratch_for(ctx->device);
    // every scratch word a pass reads it has written earlier in the same call (counters by the
    // histogram launch, tile sums by the scan, the ping-pong halves by the scatter), so the
    // last call's contents need no clearing
    radix_sort_u32_scratch(n, keys_dev, vals_dev, buf, kSortShifts, ctx->stream);
    NGUO_HIP(hipStreamSynchronize(ctx->stream));
    return NGUO_OK;
  });
}

int nguo_lbvh_morton(nguo_ctx *ctx, uint32_t nv, const float *pos, uint32_t nt, const uint32_t *tris, uint32_t *codes) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && pos && tris && codes, NGUO_E_INVALID, "nguo_lbvh_morton: null argument");
    for (size_t i = 0; i < 3 * (size_t)nt; ++i) NGUO_REQUIRE(tris[i] < nv, NGUO_E_INVALID, "triangle index out of range");
    if (nt == 0) return NGUO_OK;
    hipStream_t s = ctx->stream;
    DevBuf<float> p, b, c;
    DevBuf<uint32_t> t, k;
    p.upload(pos, 3 * (size_t)nv, s);
    t.upload(tris, 3 * (size_t)nt, s);
    b.reserve(6 * (size_t)div_up(nt, 256) + 6);
    c.reserve(3 * (size_t)nt);
    k.reserve(nt);
    morton_codes(0, nt, nullptr, t.p, p.p, b.p, c.p, k.p, s);
    k.download(codes, nt, s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_lbvh_sort_pairs(nguo_ctx *ctx, uint32_t n, uint32_t *keys, uint32_t *vals) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && keys && vals, NGUO_E_INVALID, "nguo_lbvh_sort_pairs: null argument");
    if (n == 0) return NGUO_OK;
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> k, v;
    k.upload(keys, n, s);
    v.upload(vals, n, s);
    RadixSort rs;
    rs.sort_u32(n, k.p, v.p, {0, 8, 16, 24}, s);
    k.download(keys, n, s);
    v.download(vals, n, s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

// Random 32-bit keys (splitmix64 of the index and seed) with index payloads, generated on the
// device; each timed repetition restores the unsorted keys with a device copy (excluded from
// the timing by events around the sort alone).
__global__ void k_bench_keys(uint32_t n, uint64_t seed, uint32_t *keys, uint32_t *vals) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  keys[i] = (uint32_t)z;
  vals[i] = i;
}

int nguo_bench_sort(nguo_ctx *ctx, uint32_t n, uint32_t reps, double *avg_ms, double *bytes) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && avg_ms && bytes && n > 1 && reps > 0, NGUO_E_INVALID, "nguo_bench_sort: bad argument");
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> k0, v0, k, v;
    k0.reserve(n); v0.reserve(n); k.reserve(n); v.reserve(n);
    hipLaunchKernelGGL(k_bench_keys, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, s, n, (uint64_t)7, k0.p, v0.p);
    NGUO_HIP(hipGetLastError());
    RadixSort rs;
    hipEvent_t e0, e1;
    NGUO_HIP(hipEventCreate(&e0));
    NGUO_HIP(hipEventCreate(&e1));
    double total = 0.0;
    for (uint32_t r = 0; r <= reps; ++r) {  // repetition 0 warms up
      NGUO_HIP(hipMemcpyAsync(k.p, k0.p, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
      NGUO_HIP(hipMemcpyAsync(v.p, v0.p, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
      NGUO_HIP(hipEventRecord(e0, s));
      rs.sort_u32(n, k.p, v.p, {0, 8, 16, 24}, s);
      NGUO_HIP(hipEventRecord(e1, s));
      NGUO_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      NGUO_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r) total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = total / reps;
    // per 8-bit pass: histogram reads the keys (4 B), the scatter reads key + payload and
    // writes them (16 B)
    *bytes = 4.0 * 20.0 * (double)n;
    return NGUO_OK;
  });
}

int nguo_lbvh_karras(nguo_ctx *ctx, uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && codes && left && right, NGUO_E_INVALID, "nguo_lbvh_karras: null argument");
    if (n < 2) return NGUO_OK;
    hipStream_t s = ctx->stream;
    DevBuf<uint32_t> c, l, r;
    c.upload(codes, n, s);
    l.reserve(n);
    r.reserve(n);
    karras_build(n, c.p, l.p, r.p, s);
    l.download(left, n - 1, s);
    r.download(right, n - 1, s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

int nguo_lbvh_aabb(nguo_ctx *ctx, uint32_t n, const uint32_t *left, const uint32_t *right, const float *leaf_box, float *node_box) {
  return guarded([&]() -> int {
    NGUO_REQUIRE(ctx && left && right && leaf_box && node_box, NGUO_E_INVALID, "nguo_lbvh_aabb: null argument");
    if (n < 2) return NGUO_OK;
    for (uint32_t i = 0; i + 1 < n; ++i) {
      uint32_t a = left[i], b = right[i];
      NGUO_REQUIRE(((a & LEAF) ? (a & ~LEAF) < n : a < n - 1) && ((b & LEAF) ? (b & ~LEAF) < n : b < n - 1), NGUO_E_INVALID,
                   "nguo_lbvh_aabb: child index out of range");
    }
    hipStream_t s = ctx->stream;
    // parents from the child lists (host), then the device climb
    std::vector<uint32_t> parent(2 * (size_t)n - 1, 0xffffffffu);
    for (uint32_t i = 0; i + 1 < n; ++i)
      for (uint32_t c : {left[i], right[i]}) parent[(c & LEAF) ? (n - 1 + (c & ~LEAF)) : c] = i;
    for (size_t i = 1; i < parent.size(); ++i) {
      NGUO_REQUIRE(parent[i] != 0xffffffffu, NGUO_E_INVALID, "nguo_lbvh_aabb: not a tree");
      // every node must reach the root within n steps (no cycles)
      size_t node = i;
      uint32_t hops = 0;
      while (node != 0 && hops <= n) { node = parent[node]; ++hops; }
      NGUO_REQUIRE(node == 0, NGUO_E_INVALID, "nguo_lbvh_aabb: child lists contain a cycle");
    }
    // leaf ranges bottom-up (children before parents: reverse BFS order from the root);
    // a tree whose subtrees cover contiguous sorted leaves (every Karras tree) takes the
    // range path, any other binary tree the climb
    std::vector<uint32_t> order{0};
    for (size_t k = 0; k < order.size(); ++k)
      for (uint32_t c : {left[order[k]], right[order[k]]})
        if (!(c & LEAF)) order.push_back(c);
    std::vector<uint32_t> range(2 * (size_t)(n - 1));
    bool contiguous = true;
    auto span = [&](uint32_t c, uint32_t &a, uint32_t &b) {
      if (c & LEAF) { a = b = c & ~LEAF; }
      else { a = range[2 * c]; b = range[2 * c + 1]; }
    };
    for (size_t k = order.size(); k-- > 0;) {
      const uint32_t i = order[k];
      uint32_t la, lb2, ra, rb;
      span(left[i], la, lb2);
      span(right[i], ra, rb);
      if (lb2 + 1 != ra) contiguous = false;
      range[2 * i] = la;
      range[2 * i + 1] = rb;
    }
    DevBuf<float> lb, nb, table;
    lb.upload(leaf_box, 6 * (size_t)n, s);
    nb.reserve(6 * (size_t)n);
    if (contiguous) {
      DevBuf<uint32_t> rg;
      rg.upload(range, s);
      aabb_from_ranges(n, rg.p, lb.p, nb.p, table, s);
    } else {
      DevBuf<uint32_t> l, r, p, v;
      l.upload(left, n - 1, s);
      r.upload(right, n - 1, s);
      p.upload(parent, s);
      v.reserve(n);
      aabb_build(n, l.p, r.p, p.p, v.p, lb.p, nb.p, s);
    }
    nb.download(node_box, 6 * (size_t)(n - 1), s);
    NGUO_HIP(hipStreamSynchronize(s));
    return NGUO_OK;
  });
}

}  // extern "C"
