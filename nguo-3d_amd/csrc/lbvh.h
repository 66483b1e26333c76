// LBVH over triangles or edges: 30-bit Morton keys of primitive centroids normalised by the
// centroid bounds, a stable LSD radix sort (8-bit digits, wavefront multisplit ranking in
// LDS), Karras (2012) internal nodes with their leaf ranges, and node boxes as unions over
// those ranges (sparse table of leaf-box unions).
//
// Reference stages (engine-gpu/src/): morton.rs:12-58 + morton.wgsl, radix.rs:102-135 +
// radix.wgsl (serial per-block scatter, n <= 65 536), scan.rs, karras.rs:24-88 + karras.wgsl,
// bvh_aabb.rs:28-55 + bvh_aabb.wgsl (host-driven level passes), bvh_traverse.rs:77-260.
#pragma once
#include <type_traits>
#include <vector>

#include "devmath.h"
#include "nguo_internal.h"

namespace nguo {

constexpr uint32_t LEAF = 0x80000000u;
// entries of a traversal record: a node's grandchildren. (Eight, its great-grandchildren in
// 256 bytes, made the garment drape's step 1% slower than four.)
constexpr uint32_t TRAV_WIDTH = 4;
constexpr uint32_t TRAV_ENTRY = 3;  // float4 per record entry: box and refs, then vertex ids
constexpr uint32_t TRAV_NONE = 0xffffffffu;   // unused record entry (primitive ids are below 2^31)

// Radix sort tiles: items per thread (tiles of 1024 / 4096 items), the input size from which
// the large tile is used, and the counters one workgroup of the large scan covers.
constexpr uint32_t SMALL_ITEMS = 4, LARGE_ITEMS = 16;
constexpr uint32_t LARGE_SORT = 1u << 20;
constexpr uint32_t SCAN_TILE = 16 * BLOCK;

constexpr uint32_t radix_items(uint32_t n) { return n >= LARGE_SORT ? LARGE_ITEMS : SMALL_ITEMS; }
// digit counters of one pass: 256 per tile
constexpr size_t radix_counters(uint32_t n) { return 256u * (((size_t)n + radix_items(n) * BLOCK - 1) / (radix_items(n) * BLOCK)); }
// 32-bit words of scratch radix_sort_u32_scratch needs for n items: the key and value
// ping-pong halves, the digit counters and the large scan's tile sums.
constexpr size_t radix_scratch_words(uint32_t n) {
  return 2 * (size_t)n + radix_counters(n) + (radix_counters(n) + SCAN_TILE - 1) / SCAN_TILE + 1;
}
// The sort of RadixSort::sort_u32 with every intermediate buffer carved from `scratch`
// (radix_scratch_words(n) words); allocates nothing.
void radix_sort_u32_scratch(uint32_t n, uint32_t *keys, uint32_t *vals, uint32_t *scratch, const std::vector<int> &shifts, hipStream_t s);

// Stable ascending sort of (key, value) pairs on the device.
class RadixSort {
 public:
  // keys/vals are sorted in place (ping-pong through internal buffers). `shifts` lists the
  // digit positions to sort by (least significant first), e.g. {0, 8, 16, 24} for u32 keys.
  void sort_u32(uint32_t n, uint32_t *keys, uint32_t *vals, const std::vector<int> &shifts, hipStream_t s);
  void sort_u64(uint32_t n, uint64_t *keys, uint32_t *vals, const std::vector<int> &shifts, hipStream_t s);
  // Sizes the scratch for sorts of up to n u64-keyed pairs (no allocation on later sorts).
  void reserve_u64(uint32_t n);

 private:
  DevBuf<uint32_t> k32_, v_, hist_, sums_;
  DevBuf<uint64_t> k64_;
};

// Flat view of a built tree for device traversal.
struct BvhView {
  uint32_t n;             // leaves
  const uint32_t *left;   // [n-1] child refs (LEAF bit marks a leaf slot)
  const uint32_t *right;  // [n-1]
  const float *node_box;  // [6 (n-1)] lo.xyz hi.xyz
  const float *leaf_box;  // [6 n] in sorted order
  const uint32_t *prim;   // [n] primitive id of sorted leaf slot
  // [TRAV_ENTRY TRAV_WIDTH (n-1)] traversal records: internal node i's up to TRAV_WIDTH
  // grandchildren (a leaf child stands for itself) in one 192-byte record, entry e as
  // (lo.xyz, hi.x) (hi.yz, ref, last sorted leaf slot under it) (vertex ids of a leaf's
  // primitive: three for a triangle, two and 0 for an edge; 0 for a node); ref is LEAF | the
  // primitive id for a leaf (not its slot, as in left / right), else the binary node id;
  // TRAV_NONE for an unused entry
  const float4 *trav;
  // [TRAV_ENTRY TRAV_PARTS] the tree split into disjoint subtrees that together hold every
  // leaf once: entry k = (lo.xyz, ref) (hi.xyz, last leaf slot) (vertex ids as in trav), ref
  // as in trav, ~0u for an unused entry
  const float4 *front;
  const uint32_t *slot;   // [n] sorted leaf slot of element k of the build's primitive list
};

// Subtrees a traversal query is split into (one thread each).
constexpr uint32_t TRAV_PARTS = 8;

class Lbvh {
 public:
  // kind 0: triangles (3 indices per primitive), 1: edges (2 indices). `ids` selects the
  // primitives of this tree (device array of primitive indices into `idx`).
  void build(int kind, uint32_t n, const uint32_t *ids, const uint32_t *idx, const float *pos, hipStream_t s);
  // The last build's topology (leaf order, children, ranges, parts) with every box recomputed
  // from `pos`: a valid tree for any positions, as tight as a build where they moved little
  // since it.
  void refit(int kind, const uint32_t *idx, const float *pos, hipStream_t s);
  BvhView view() const {
    return BvhView{n_, left_.p, right_.p, node_box_.p, leaf_box_.p, prim_.p, (const float4 *)trav_.p, (const float4 *)front_.p, slot_.p};
  }
  uint32_t size() const { return n_; }
  // Stage outputs for parity checks.
  const uint32_t *codes() const { return keys_.p; }   // sorted Morton keys
  RadixSort sorter;

 private:
  uint32_t n_ = 0;
  DevBuf<uint32_t> keys_, prim_, left_, right_, range_, visit_, slot_;
  DevBuf<float> leaf_box_, node_box_, bounds_, cent_, table_, trav_, front_;
};

// Kernels exposed for the stage-by-stage C ABI test entry points.
void morton_codes(int kind, uint32_t n, const uint32_t *ids, const uint32_t *idx, const float *pos, float *bounds_scratch,
                  float *cent_scratch, uint32_t *codes, hipStream_t s);
void karras_build(uint32_t n, const uint32_t *codes, uint32_t *left, uint32_t *right, hipStream_t s);
void aabb_build(uint32_t n, const uint32_t *left, const uint32_t *right, uint32_t *parent, uint32_t *visit, const float *leaf_box,
                float *node_box, hipStream_t s);
// Node boxes of a tree whose internal node i covers the contiguous sorted leaves
// [range[2i], range[2i+1]] (every Karras tree); `table` is scratch.
void aabb_from_ranges(uint32_t n, const uint32_t *range, const float *leaf_box, float *node_box, DevBuf<float> &table, hipStream_t s);

// Stack traversal: calls visit(prim) -- or visit(prim, vid) when the visitor takes the
// primitive's vertex ids (uint3, from the record: the leaf then needs no index load) -- for
// every leaf whose box passes `overlap(box_lo, box_hi)` (a conservative test). The first TRAV_LDS stack entries of a thread live in LDS
// (`lds`: the calling kernel's __shared__ uint32_t[TRAV_LDS * BLOCK], interleaved by thread so
// that a wavefront's pushes and pops fall in distinct banks); deeper entries go to a private
// array. A stack deeper than TRAV_STACK sets *err (reported to the caller as
// NGUO_E_CAPACITY) instead of silently dropping a subtree. Requires BLOCK threads per
// workgroup. Only leaves in sorted slots >= min_slot are visited, and subtrees whose last
// slot lies below it are not entered (a query of a tree against itself visits each pair
// once by passing its own slot + 1).
constexpr int TRAV_STACK = 128;
constexpr int TRAV_LDS = 32;
template <class Visit>
__device__ __forceinline__ void trav_visit(Visit &visit, uint32_t prim, uint3 vid) {
  if constexpr (std::is_invocable_v<Visit &, uint32_t, uint3>) visit(prim, vid);
  else visit(prim);
}

// Depth-first traversal of the subtree under internal node `root`.
template <class Overlap, class Visit>
__device__ inline void bvh_traverse_from(const BvhView &t, uint32_t root, uint32_t *err, uint32_t *lds, Overlap overlap, Visit visit,
                                         uint32_t min_slot = 0) {
  uint32_t spill[TRAV_STACK - TRAV_LDS];
  uint32_t *mine = lds + threadIdx.x;
  mine[0] = root;
  int sp = 1;
  // Each pop loads one record and tests up to four boxes: the grandchildren of the popped
  // binary node. Leaves are visited where they are met. (Holding them for a wavefront-wide
  // visit phase, the "while-while" form, made the garment drape's step 20% slower, its CCD and
  // contact queries about twice as slow: lanes that hold a leaf wait for the deepest descent
  // of the wavefront.)
  while (sp > 0) {
    --sp;
    const uint32_t node = sp < TRAV_LDS ? mine[sp * BLOCK] : spill[sp - TRAV_LDS];
    const float4 *rec = t.trav + TRAV_ENTRY * TRAV_WIDTH * (size_t)node;
    float4 q[TRAV_ENTRY * TRAV_WIDTH];
#pragma unroll
    for (uint32_t k = 0; k < TRAV_ENTRY * TRAV_WIDTH; ++k) q[k] = rec[k];
    // box tests and pushes unrolled (the record stays in registers); the leaves met are
    // visited after, from a mask, through one copy of the visit code
    uint32_t ref[TRAV_WIDTH], leaves = 0;
#pragma unroll
    for (uint32_t c = 0; c < TRAV_WIDTH; ++c) {
      const float4 e0 = q[TRAV_ENTRY * c], e1 = q[TRAV_ENTRY * c + 1];
      ref[c] = __float_as_uint(e1.z);
      const uint32_t last = __float_as_uint(e1.w);
      if (ref[c] == TRAV_NONE || last < min_slot) continue;
      if (!overlap(f3{e0.x, e0.y, e0.z}, f3{e0.w, e1.x, e1.y})) continue;
      if (ref[c] & LEAF) {
        leaves |= 1u << c;
      } else if (sp < TRAV_LDS) {
        mine[sp * BLOCK] = ref[c];
        ++sp;
      } else if (sp < TRAV_STACK) {
        spill[sp - TRAV_LDS] = ref[c];
        ++sp;
      } else {
        atomicOr(err, 1u);
      }
    }
    static_assert(TRAV_WIDTH == 4, "the leaf selection below names four entries");
    while (leaves) {
      const uint32_t c = __builtin_ctz(leaves);
      leaves &= leaves - 1;
      const uint32_t r = c == 0 ? ref[0] : c == 1 ? ref[1] : c == 2 ? ref[2] : ref[3];
      const float4 v = c == 0 ? q[2] : c == 1 ? q[TRAV_ENTRY + 2] : c == 2 ? q[2 * TRAV_ENTRY + 2] : q[3 * TRAV_ENTRY + 2];
      trav_visit(visit, r & ~LEAF, make_uint3(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z)));
    }
  }
}

template <class Overlap, class Visit>
__device__ inline void bvh_traverse(const BvhView &t, uint32_t *err, uint32_t *lds, Overlap overlap, Visit visit) {
  // whole-tree form (no slot bound)
  if (t.n == 0) return;
  if (t.n == 1) {
    // the frontier's first entry is the one leaf
    const float4 a = t.front[0], b = t.front[1], v = t.front[2];
    if (overlap(f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}))
      trav_visit(visit, __float_as_uint(a.w) & ~LEAF, make_uint3(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z)));
    return;
  }
  bvh_traverse_from(t, 0u, err, lds, overlap, visit);
}

// The part `part` (< TRAV_PARTS) of a traversal split over the tree's frontier subtrees: the
// parts of one query visit disjoint leaf sets whose union is what bvh_traverse visits.
template <class Overlap, class Visit>
__device__ inline void bvh_traverse_part(const BvhView &t, uint32_t part, uint32_t *err, uint32_t *lds, Overlap overlap, Visit visit,
                                         uint32_t min_slot = 0) {
  if (t.n == 0) return;
  const float4 a = t.front[TRAV_ENTRY * part], b = t.front[TRAV_ENTRY * part + 1];
  const uint32_t r = __float_as_uint(a.w);
  if (r == 0xffffffffu || __float_as_uint(b.w) < min_slot) return;
  if (!overlap(f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z})) return;
  if (r & LEAF) {
    const float4 v = t.front[TRAV_ENTRY * part + 2];
    trav_visit(visit, r & ~LEAF, make_uint3(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z)));
  }
  else bvh_traverse_from(t, r, err, lds, overlap, visit, min_slot);
}

}  // namespace nguo
