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

// Codes [lo, lo + len) copied to LDS, the rest read from memory.  The two
// pointers carry their address spaces, so the two reads stay an LDS read and
// a global read (a select between generic pointers would be a flat load).
template <typename KeyT>
struct CodeWindow {
  const __attribute__((address_space(1))) KeyT* g;
  const __attribute__((address_space(3))) KeyT* s;
  uint32_t lo, len;
  __device__ CodeWindow(const KeyT* global, const KeyT* lds, uint32_t lo_, uint32_t len_)
      : g((const __attribute__((address_space(1))) KeyT*)global),
        s((const __attribute__((address_space(3))) KeyT*)lds),
        lo(lo_),
        len(len_) {}
  __device__ __forceinline__ KeyT operator[](uint32_t j) const {
    const uint32_t o = j - lo;
    if (o < len) return s[o];
    return g[j];
  }
};

// The searches below test "more than p leading bits in common with key k"
// as (c & mask) == (k & mask) with mask = the top p + 1 bits, which is
// clz(k ^ c) > p without a leading-zero count, and run on 32-bit indices
// (n < 2^31).  Every search finds the last index at which a monotone
// predicate holds (along sorted codes the common prefix with a fixed code
// does not grow away from it), exactly as the reference's loops do.
template <typename KeyT>
__device__ __forceinline__ KeyT top_bits_mask(int p) {
  return ~KeyT(0) << (KeyTraits<KeyT>::kBits - 1 - p);  // 0 <= p < kBits
}

// Range [first, last] of internal node idx (1 <= idx <= n-2, or idx == 0).
template <typename KeyT, typename Codes = const KeyT*>
__device__ __forceinline__ void node_range(const Codes& c, uint32_t n, uint32_t idx,
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
  const int dn = clz_key(mc1 ^ mc2), dp = clz_key(mc1 ^ mc0);
  const bool fwd = dn > dp;                 // direction d = +1
  const int delta_min = fwd ? dp : dn;      // delta(idx, idx - d) < kBits
  const KeyT mask = top_bits_mask<KeyT>(delta_min);
  const KeyT key = mc1 & mask;
  const uint32_t xmax = fwd ? n - 1 - idx : idx;  // farthest step inside [0, n)
  uint32_t lmax = 2;
  while (lmax <= xmax && (c[fwd ? idx + lmax : idx - lmax] & mask) == key) lmax <<= 1;
  uint32_t l = 0;
  for (uint32_t t = lmax >> 1; t != 0; t >>= 1) {
    const uint32_t x = l + t;
    if (x <= xmax && (c[fwd ? idx + x : idx - x] & mask) == key) l = x;
  }
  if (fwd) {
    first = idx;
    last = idx + l;
  } else {
    first = idx - l;
    last = idx;
  }
}

// Split of range [first, last] (first < last), mortons.rs:290-325.
template <typename KeyT, typename Codes = const KeyT*>
__device__ __forceinline__ uint32_t node_split(const Codes& c, uint32_t first, uint32_t last) {
  const KeyT ms = c[first];
  const KeyT diff = ms ^ c[last];
  if (diff == 0) return first;
  const KeyT mask = top_bits_mask<KeyT>(clz_key(diff));
  const KeyT key = ms & mask;
  uint32_t split = first, step = last - first;
  while (step > 1) {
    step = (step + 1) >> 1;
    const uint32_t nw = split + step;
    if (nw < last && (c[nw] & mask) == key) split = nw;
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
// listed too: top_list[slot] = (node, left child, right child, bit 0: left
// child is top, bit 1: right child is top) and node2slot[node] = slot, for
// slots < top_cap; top_count
// counts all of them, so a count above top_cap says the list is incomplete.
struct TreeletSink {
  uint32_t k;
  uint32_t* count;  // zeroed by the caller
  uint4* list;      // nullptr: no treelets wanted
  uint32_t cap;
  uint32_t* mesh_root_count;  // batched only, zeroed by the caller
  uint32_t* mesh_roots;
  uint32_t* top_count = nullptr;  // zeroed by the caller; nullptr: no top list
  uint4* top_list = nullptr;
  uint32_t top_cap = 0;
  uint32_t* node2slot = nullptr;
  // Single-tree builds: a treelet entry's last word is 2 * (its parent's top
  // slot) + (1 for a right child), all ones when there is none, in place of
  // the mesh index.
  bool stage_pos = false;
  // Batched builds: per-mesh top lists (used when top_count is null): mesh
  // m's top nodes at slots mesh_top_list[top_base + slot], slot counted in
  // mesh_top_count[m] (zeroed by the caller), mesh_node2slot[top_base +
  // local node] = slot; top_base = the mesh's first internal node over the
  // batch, so a mesh's slots never exceed its internal nodes.
  uint32_t* mesh_top_count = nullptr;
  uint4* mesh_top_list = nullptr;
  uint32_t* mesh_node2slot = nullptr;
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
__device__ __forceinline__ void topology_emit(uint32_t n, uint32_t i, uint32_t first, uint32_t last,
                                              uint32_t split, uint32_t* __restrict__ bvhnodes,
                                              const uint32_t* __restrict__ idx2obj,
                                              uint32_t obj_sub, const TreeletSink& ts,
                                              uint32_t mesh, uint32_t root_base,
                                              uint32_t top_base = 0u) {
  const uint32_t nb = n - 1;
  if (ts.list) {
    if (last - first + 1 > ts.k) {
      uint32_t pos = DMSH_U32MAX;  // 2 * this node's top slot
      if (!ts.top_count && ts.mesh_top_count) {
        const uint32_t slot = atomicAdd(ts.mesh_top_count + mesh, 1u);
        if (slot < nb) {
          ts.mesh_top_list[top_base + slot] =
              make_uint4(i, first == split ? nb + split : split,
                         last == split + 1 ? nb + split + 1 : split + 1,
                         (split - first + 1 > ts.k ? 1u : 0u) | (last - split > ts.k ? 2u : 0u));
          ts.mesh_node2slot[top_base + i] = slot;
        }
      }
      if (ts.top_count) {
        const uint32_t slot = atomicAdd(ts.top_count, 1u);
        if (slot < ts.top_cap) {
          pos = 2u * slot;
          ts.top_list[slot] = make_uint4(i, first == split ? nb + split : split,
                                         last == split + 1 ? nb + split + 1 : split + 1,
                                         (split - first + 1 > ts.k ? 1u : 0u) |
                                             (last - split > ts.k ? 2u : 0u));
          ts.node2slot[i] = slot;
        }
      }
      if (split - first + 1 <= ts.k)
        emit_treelet(ts, first == split ? nb + split : split, first, split,
                     ts.stage_pos ? pos : mesh, root_base);
      if (last - split <= ts.k)
        emit_treelet(ts, last == split + 1 ? nb + split + 1 : split + 1, split + 1, last,
                     ts.stage_pos ? (pos == DMSH_U32MAX ? pos : pos + 1u) : mesh, root_base);
    } else if (i == 0) {
      emit_treelet(ts, 0u, 0u, n - 1, ts.stage_pos ? DMSH_U32MAX : mesh, root_base);
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

template <typename KeyT>
__device__ __forceinline__ void topology_node(uint32_t n, uint32_t i, uint32_t* __restrict__ bvhnodes,
                                              const KeyT* __restrict__ codes,
                                              const uint32_t* __restrict__ idx2obj,
                                              uint32_t obj_sub, const TreeletSink& ts,
                                              uint32_t mesh, uint32_t root_base) {
  uint32_t first, last;
  node_range<KeyT>(codes, n, i, first, last);
  const uint32_t split = node_split<KeyT>(codes, first, last);
  topology_emit(n, i, first, last, split, bvhnodes, idx2obj, obj_sub, ts, mesh, root_base);
}

// A tree of one object: the single leaf row (MAX, obj, MAX) and one treelet.
__device__ __forceinline__ void topology_single(uint32_t* __restrict__ bvhnodes, uint32_t obj,
                                                const TreeletSink& ts, uint32_t mesh,
                                                uint32_t root_base) {
  bvhnodes[0] = DMSH_U32MAX;
  bvhnodes[1] = obj;
  bvhnodes[2] = DMSH_U32MAX;
  if (ts.leaf_aabb) emit_leaf_box(ts, 0u, obj);
  if (ts.list) emit_treelet(ts, 0u, 0u, 0u, ts.stage_pos ? DMSH_U32MAX : mesh, root_base);
}

// One thread per internal node; the block's codes and kTopoHalo either side
// sit in LDS, where nearly every search step lands (ranges beyond the halo
// read the rest from memory).
constexpr uint32_t kTopoThreads = 256;
constexpr uint32_t kTopoHalo = 256;

template <typename KeyT>
__global__ __launch_bounds__(kTopoThreads) void k_topology(uint32_t n, uint32_t* __restrict__ bvhnodes,
                                                           const KeyT* __restrict__ codes,
                                                           const uint32_t* __restrict__ idx2obj,
                                                           TreeletSink ts) {
  __shared__ KeyT s_c[kTopoHalo + kTopoThreads + kTopoHalo + 1];
  const uint32_t base = blockIdx.x * kTopoThreads;
  const uint32_t lo = base > kTopoHalo ? base - kTopoHalo : 0u;
  const uint32_t hi = min(n, base + kTopoThreads + kTopoHalo + 1);
  for (uint32_t j = lo + threadIdx.x; j < hi; j += kTopoThreads) s_c[j - lo] = codes[j];
  __syncthreads();
  const uint32_t i = base + threadIdx.x;
  if (i >= n - 1) return;
  const CodeWindow<KeyT> c(codes, s_c, lo, hi - lo);
  uint32_t first, last;
  node_range<KeyT>(c, n, i, first, last);
  const uint32_t split = node_split<KeyT>(c, first, last);
  topology_emit(n, i, first, last, split, bvhnodes, idx2obj, 0u, ts, 0u, 0u);
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
    launch("k_topology", k_topology<KeyT>, dim3(ceil_div(n - 1, kTopoThreads)), dim3(kTopoThreads),
           0, s, n, bvhnodes, codes, idx2obj, ts);
  }
  return launch_status();
}

}  // namespace dmsh
