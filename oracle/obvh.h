// Oracle LBVH stages: 30-bit Morton codes of triangle centroids normalised by the
// centroid AABB, the stable key sort, the Karras (2012) internal-node build and the
// bottom-up AABB union. Integer/exact stages: the HIP kernels must match bit for bit.
// TEST INFRASTRUCTURE ONLY (checker for the HIP path).
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "omath.h"

namespace oracle {

constexpr uint32_t LEAF_BIT = 0x80000000u;

// engine-gpu/src/morton.rs:12-19 expand_bits.
inline uint32_t expand_bits(uint32_t v) {
  uint32_t x = v & 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Rust `f32 as u32`: truncation toward zero, saturating, NaN -> 0.
inline uint32_t f32_to_u32_sat(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)v;
}

// engine-gpu/src/morton.rs:22-25 morton_code_cpu.
inline uint32_t morton_code(V3<float> n) {
  auto q = [](float v) { return f32_to_u32_sat(clampv(clampv(v, 0.0f, 1.0f) * 1023.0f, 0.0f, 1023.0f)); };
  return expand_bits(q(n.x)) * 4u + expand_bits(q(n.y)) * 2u + expand_bits(q(n.z));
}

// engine-gpu/src/morton.rs:29-45 centroid_bounds (centroid = (a+b+c)/3, inv = 1/extent or 0).
inline void centroid_bounds(const std::vector<V3<float>> &pos, const std::vector<std::array<uint32_t, 3>> &tris, V3<float> &lo, V3<float> &inv) {
  const float INF = INFINITY;
  lo = splat(INF);
  V3<float> hi = splat(-INF);
  for (auto &t : tris) {
    V3<float> c = (pos[t[0]] + pos[t[1]] + pos[t[2]]) / 3.0f;
    lo = vmin(lo, c);
    hi = vmax(hi, c);
  }
  V3<float> ext = hi - lo;
  inv = V3<float>(ext.x > 0.0f ? 1.0f / ext.x : 0.0f, ext.y > 0.0f ? 1.0f / ext.y : 0.0f, ext.z > 0.0f ? 1.0f / ext.z : 0.0f);
}

// engine-gpu/src/morton.rs:48-58 morton_codes_cpu.
inline std::vector<uint32_t> morton_codes(const std::vector<V3<float>> &pos, const std::vector<std::array<uint32_t, 3>> &tris) {
  V3<float> lo, inv;
  centroid_bounds(pos, tris, lo, inv);
  std::vector<uint32_t> out;
  out.reserve(tris.size());
  for (auto &t : tris) {
    V3<float> c = (pos[t[0]] + pos[t[1]] + pos[t[2]]) / 3.0f;
    V3<float> d = c - lo;
    out.push_back(morton_code(V3<float>(d.x * inv.x, d.y * inv.y, d.z * inv.z)));
  }
  return out;
}

// engine-gpu/src/radix.rs:112-135 sort_pairs semantics: stable ascending sort of (key, index).
inline void sort_pairs(std::vector<uint32_t> &keys, std::vector<uint32_t> &vals) {
  std::vector<uint32_t> order(keys.size());
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  std::vector<uint32_t> k2(keys.size()), v2(keys.size());
  for (size_t i = 0; i < order.size(); ++i) { k2[i] = keys[order[i]]; v2[i] = vals[order[i]]; }
  keys.swap(k2);
  vals.swap(v2);
}

inline int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

// engine-gpu/src/karras.rs:24-88 karras_tree_cpu (LEAF = top bit).
inline void karras_tree(const std::vector<uint32_t> &codes, std::vector<uint32_t> &left, std::vector<uint32_t> &right) {
  int n = (int)codes.size();
  int m = n > 0 ? n - 1 : 0;
  left.assign(m, 0);
  right.assign(m, 0);
  auto delta = [&](int i, int j) -> int {
    if (j < 0 || j >= n) return -1;
    uint32_t x = codes[i] ^ codes[j];
    if (x == 0) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(x);
  };
  for (int i = 0; i < m; ++i) {
    int d = (delta(i, i + 1) - delta(i, i - 1)) < 0 ? -1 : 1;
    int delta_min = delta(i, i - d);
    int l_max = 2;
    while (delta(i, i + l_max * d) > delta_min) l_max *= 2;
    int l = 0;
    for (int t = l_max / 2; t >= 1; t /= 2)
      if (delta(i, i + (l + t) * d) > delta_min) l += t;
    int j = i + l * d;
    int first = std::min(i, j), last = std::max(i, j);
    int node_delta = delta(first, last);
    int split = first, step = last - first;
    while (true) {
      step = (step + 1) / 2;
      int ns = split + step;
      if (ns < last && delta(first, ns) > node_delta) split = ns;
      if (step <= 1) break;
    }
    int gamma = split;
    left[i] = first == gamma ? ((uint32_t)gamma | LEAF_BIT) : (uint32_t)gamma;
    right[i] = last == gamma + 1 ? ((uint32_t)(gamma + 1) | LEAF_BIT) : (uint32_t)(gamma + 1);
  }
}

// engine-gpu/src/bvh_aabb.rs:28-55 propagate_aabb_cpu (union of the subtree's leaves).
inline std::vector<std::array<float, 6>> propagate_aabb(const std::vector<uint32_t> &left, const std::vector<uint32_t> &right,
                                                        const std::vector<std::array<float, 6>> &leaf) {
  size_t n = leaf.size();
  std::vector<std::array<float, 6>> node(n > 0 ? n - 1 : 0);
  if (n < 2) return node;
  auto merge = [](const std::array<float, 6> &a, const std::array<float, 6> &b) {
    return std::array<float, 6>{std::fmin(a[0], b[0]), std::fmin(a[1], b[1]), std::fmin(a[2], b[2]),
                                std::fmax(a[3], b[3]), std::fmax(a[4], b[4]), std::fmax(a[5], b[5])};
  };
  // iterative post-order
  std::vector<std::pair<uint32_t, int>> st;
  st.push_back({0u, 0});
  while (!st.empty()) {
    auto &top = st.back();
    uint32_t i = top.first;
    if (top.second == 0) {
      top.second = 1;
      uint32_t l = left[i], r = right[i];
      if (!(r & LEAF_BIT)) st.push_back({r, 0});
      if (!(l & LEAF_BIT)) st.push_back({l, 0});
    } else {
      uint32_t l = left[i], r = right[i];
      const auto &la = (l & LEAF_BIT) ? leaf[l & ~LEAF_BIT] : node[l];
      const auto &ra = (r & LEAF_BIT) ? leaf[r & ~LEAF_BIT] : node[r];
      node[i] = merge(la, ra);
      st.pop_back();
    }
  }
  return node;
}

}  // namespace oracle
