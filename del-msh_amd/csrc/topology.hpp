// topology.hpp -- Karras (2012) binary radix tree over sorted Morton codes.
//
// One thread per internal node i (0 .. n-2) computes the node's key range and
// split exactly as del-msh-cpu/src/mortons.rs:221-325 defines them, including
// the duplicate-code rule (a node inside a run of equal codes covers the rest
// of the run and splits at its first element), then writes the child links
// of bvhnodes_morton.rs:18-49.  The run end of the duplicate rule is found by
// binary search over the sorted codes instead of the reference's linear scan:
// both return the last index of the run, the search in O(log n).
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {

__device__ __forceinline__ int clz_key(uint32_t v) { return v ? __builtin_clz(v) : 32; }
__device__ __forceinline__ int clz_key(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

template <typename KeyT>
struct KeyTraits {
  static constexpr int kBits = 8 * sizeof(KeyT);
};

// Range [first, last] of internal node idx (1 <= idx <= n-2, or idx == 0).
template <typename KeyT>
__device__ __forceinline__ void node_range(const KeyT* __restrict__ c, uint32_t n, uint32_t idx,
                                           uint32_t& first, uint32_t& last) {
  if (idx == 0) {
    first = 0;
    last = n - 1;
    return;
  }
  const KeyT mc0 = c[idx - 1], mc1 = c[idx], mc2 = c[idx + 1];
  if (mc0 == mc1 && mc1 == mc2) {
    uint32_t lo = idx + 1, hi = n - 1;  // c[lo] == mc1 holds
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (c[mid] == mc1) lo = mid;
      else hi = mid - 1;
    }
    first = idx;
    last = lo;
    return;
  }
  const int64_t i = idx;
  const int dn = clz_key(mc1 ^ mc2), dp = clz_key(mc1 ^ mc0);
  const int64_t d = (dn - dp) > 0 ? 1 : -1;
  const int delta_min = (d > 0) ? dp : dn;  // delta(idx, idx - d)
  int64_t lmax = 2;
  for (;;) {
    const int64_t j = i + lmax * d;
    if (j < 0 || j >= (int64_t)n) break;
    if (clz_key(mc1 ^ c[j]) <= delta_min) break;
    lmax *= 2;
  }
  int64_t l = 0;
  for (int64_t t = lmax / 2; t >= 1; t /= 2) {
    const int64_t j = i + (l + t) * d;
    if (j >= 0 && j < (int64_t)n && clz_key(mc1 ^ c[j]) > delta_min) l += t;
  }
  const int64_t j = i + l * d;
  if (i <= j) {
    first = idx;
    last = (uint32_t)j;
  } else {
    first = (uint32_t)j;
    last = idx;
  }
}

// Split of range [first, last] (first < last), mortons.rs:290-325.
template <typename KeyT>
__device__ __forceinline__ uint32_t node_split(const KeyT* __restrict__ c, uint32_t first,
                                               uint32_t last) {
  const KeyT ms = c[first];
  const int p0 = clz_key(ms ^ c[last]);
  if (p0 == KeyTraits<KeyT>::kBits) return first;
  uint32_t split = first, step = last - first;
  while (step > 1) {
    step = (step + 1) >> 1;
    const uint32_t nw = split + step;
    if (nw >= last) continue;
    if (clz_key(ms ^ c[nw]) > p0) split = nw;
  }
  return split;
}

// Treelets for the refit (optional): a treelet is a subtree of at most k
// leaves whose parent has more than k (or the whole tree when n <= k).  A
// treelet covering leaves [f, l] holds exactly the internal nodes of its
// subtree, and every one of them has an id inside [f, l] (a Karras node sits
// at one end of its own range), so one workgroup can refit it in LDS.
// Entries are (node id, f, l, mesh), appended in any order; at most n of
// them.  Batched builds also append each treelet root to its mesh's own list
// (mesh_roots + root_base, counted in mesh_root_count[mesh]) for the top climb.
// Optionally the nodes above the treelets ("top" nodes, range > k) are
// listed too: top_list[slot] = (node, bit 0: left child is top, bit 1: right
// child is top) and node2slot[node] = slot, for slots < top_cap; top_count
// counts all of them, so a count above top_cap says the list is incomplete.
struct TreeletSink {
  uint32_t k;
  uint32_t* count;  // zeroed by the caller
  uint4* list;      // nullptr: no treelets wanted
  uint32_t cap;
  uint32_t* mesh_root_count;  // batched only, zeroed by the caller
  uint32_t* mesh_roots;
  uint32_t* top_count = nullptr;  // zeroed by the caller; nullptr: no top list
  uint2* top_list = nullptr;
  uint32_t top_cap = 0;
  uint32_t* node2slot = nullptr;
  // Optional: the leaf boxes, written into leaf_aabb (the tree's bvhnode2aabb)
  // as the leaf rows are: from tri_box (6 floats per object) when given, else
  // the point xyz[obj] (the box the refit's leaf_box makes for it).
  float* leaf_aabb = nullptr;
  const float* tri_box = nullptr;
  const float* xyz = nullptr;
};

__device__ __forceinline__ void emit_leaf_box(const TreeletSink& ts, uint32_t leaf, uint32_t obj) {
  const float* src;
  float b[6];
  if (ts.tri_box) {
    src = ts.tri_box + (size_t)obj * 6;
#pragma unroll
    for (int d = 0; d < 6; ++d) b[d] = src[d];
  } else {
    src = ts.xyz + (size_t)obj * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] = src[d] - 0.f;
      b[d + 3] = src[d] + 0.f;
    }
  }
  float2* dst = reinterpret_cast<float2*>(ts.leaf_aabb + (size_t)leaf * 6);
  dst[0] = make_float2(b[0], b[1]);
  dst[1] = make_float2(b[2], b[3]);
  dst[2] = make_float2(b[4], b[5]);
}

__device__ __forceinline__ void emit_treelet(const TreeletSink& ts, uint32_t node, uint32_t f,
                                             uint32_t l, uint32_t mesh, uint32_t root_base) {
  const uint32_t slot = atomicAdd(ts.count, 1u);
  if (slot < ts.cap) ts.list[slot] = make_uint4(node, f, l, mesh);
  if (ts.mesh_roots) {
    const uint32_t r = atomicAdd(ts.mesh_root_count + mesh, 1u);
    ts.mesh_roots[root_base + r] = node;  // at most one root per leaf of the mesh
  }
}

// Internal node i of a tree over n sorted codes: range, split, child links
// (bvhnodes_morton.rs:18-49) and the treelets it starts.  idx2obj[k] - obj_sub
// is the object of sorted position k.
template <typename KeyT>
__device__ __forceinline__ void topology_node(uint32_t n, uint32_t i, uint32_t* __restrict__ bvhnodes,
                                              const KeyT* __restrict__ codes,
                                              const uint32_t* __restrict__ idx2obj,
                                              uint32_t obj_sub, const TreeletSink& ts,
                                              uint32_t mesh, uint32_t root_base) {
  const uint32_t nb = n - 1;
  uint32_t first, last;
  node_range<KeyT>(codes, n, i, first, last);
  const uint32_t split = node_split<KeyT>(codes, first, last);
  if (ts.list) {
    if (last - first + 1 > ts.k) {
      if (ts.top_count) {
        const uint32_t slot = atomicAdd(ts.top_count, 1u);
        if (slot < ts.top_cap) {
          ts.top_list[slot] = make_uint2(i, (split - first + 1 > ts.k ? 1u : 0u) |
                                                (last - split > ts.k ? 2u : 0u));
          ts.node2slot[i] = slot;
        }
      }
      if (split - first + 1 <= ts.k)
        emit_treelet(ts, first == split ? nb + split : split, first, split, mesh, root_base);
      if (last - split <= ts.k)
        emit_treelet(ts, last == split + 1 ? nb + split + 1 : split + 1, split + 1, last, mesh,
                     root_base);
    } else if (i == 0) {
      emit_treelet(ts, 0u, 0u, n - 1, mesh, root_base);
    }
  }
  if (first == split) {
    const uint32_t a = nb + split;
    const uint32_t obj = idx2obj[split] - obj_sub;
    bvhnodes[(size_t)i * 3 + 1] = a;
    bvhnodes[(size_t)a * 3 + 0] = i;
    bvhnodes[(size_t)a * 3 + 1] = obj;
    bvhnodes[(size_t)a * 3 + 2] = DMSH_U32MAX;
    if (ts.leaf_aabb) emit_leaf_box(ts, a, obj);
  } else {
    bvhnodes[(size_t)i * 3 + 1] = split;
    bvhnodes[(size_t)split * 3 + 0] = i;
  }
  if (last == split + 1) {
    const uint32_t b = nb + split + 1;
    const uint32_t obj = idx2obj[split + 1] - obj_sub;
    bvhnodes[(size_t)i * 3 + 2] = b;
    bvhnodes[(size_t)b * 3 + 0] = i;
    bvhnodes[(size_t)b * 3 + 1] = obj;
    bvhnodes[(size_t)b * 3 + 2] = DMSH_U32MAX;
    if (ts.leaf_aabb) emit_leaf_box(ts, b, obj);
  } else {
    bvhnodes[(size_t)i * 3 + 2] = split + 1;
    bvhnodes[(size_t)(split + 1) * 3 + 0] = i;
  }
  if (i == 0) bvhnodes[0] = DMSH_U32MAX;
}

// A tree of one object: the single leaf row (MAX, obj, MAX) and one treelet.
__device__ __forceinline__ void topology_single(uint32_t* __restrict__ bvhnodes, uint32_t obj,
                                                const TreeletSink& ts, uint32_t mesh,
                                                uint32_t root_base) {
  bvhnodes[0] = DMSH_U32MAX;
  bvhnodes[1] = obj;
  bvhnodes[2] = DMSH_U32MAX;
  if (ts.leaf_aabb) emit_leaf_box(ts, 0u, obj);
  if (ts.list) emit_treelet(ts, 0u, 0u, 0u, mesh, root_base);
}

template <typename KeyT>
__global__ __launch_bounds__(256) void k_topology(uint32_t n, uint32_t* __restrict__ bvhnodes,
                                                  const KeyT* __restrict__ codes,
                                                  const uint32_t* __restrict__ idx2obj,
                                                  TreeletSink ts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  topology_node<KeyT>(n, i, bvhnodes, codes, idx2obj, 0u, ts, 0u, 0u);
}

static __global__ void k_topology_single(uint32_t* __restrict__ bvhnodes,
                                         const uint32_t* __restrict__ idx2obj, TreeletSink ts) {
  topology_single(bvhnodes, idx2obj[0], ts, 0u, 0u);
}

template <typename KeyT>
inline int launch_topology(hipStream_t s, uint32_t n, uint32_t* bvhnodes, const KeyT* codes,
                           const uint32_t* idx2obj,
                           TreeletSink ts = TreeletSink{0, nullptr, nullptr, 0, nullptr, nullptr}) {
  if (n == 0) return DMSH_OK;
  if (n == 1) {
    launch("k_topology_single", k_topology_single, dim3(1), dim3(1), 0, s, bvhnodes, idx2obj, ts);
  } else {
    launch("k_topology", k_topology<KeyT>, dim3(ceil_div(n - 1, 256)), dim3(256), 0, s, n, bvhnodes,
           codes, idx2obj, ts);
  }
  return launch_status();
}

}  // namespace dmsh
