// octree.hpp -- quad/oct-tree from the Morton binary radix tree, and the
// sorted-array helpers that prepare its input.
//
// Follows del-msh-cpu/src/quad_oct_tree.rs:1-230 (the reference GPU kernels
// are del-msh-cuda-kernels/cuda/quad_oct_tree.cu:28-233):
//   k_brt_depth       binary radix tree over sorted unique codes (the Karras
//                     range / split of the LBVH topology) and each branch's
//                     depth = (clz(code[split] ^ code[split+1]) - (32 -
//                     num_dim * max_depth)) / num_dim;
//   k_bnode_flags     idx -> the branch holding it as a child, and the flags
//                     "depth differs from the parent's" whose exclusive scan
//                     numbers the tree nodes (onodes);
//   k_octree_make     children / parent links, depths and cell centres of
//                     the onodes and of the points;
//   k_onode_hops +
//   k_onode_sum_level per-node sums of per-point values, each node adding its
//                     children in slot order (a fixed order; del-msh-cpu adds
//                     point by point, quad_oct_tree.rs:420-451, and the
//                     reference GPU kernel with atomics, so neither order is
//                     kept: sums of exactly representable partial sums agree
//                     bit for bit, others to rounding).  The parent link is
//                     entry 0 of the node's 1 + 2^num_dim links (the
//                     reference GPU kernel's num_link; the CPU one reads
//                     entry i*9, right only in 3D).
// Codes must be unique (the reference asserts it when it builds the tree);
// walks that leave the tree set DMSH_DEVERR_INDEX instead of looping.
#pragma once

#include "dmsh_common.hpp"
#include "topology.hpp"

namespace dmsh {
namespace oct {

__global__ __launch_bounds__(256) void k_brt_depth(uint32_t num_bnode,
                                                   const uint32_t* __restrict__ idx2morton,
                                                   uint32_t num_dim, uint32_t max_depth,
                                                   uint32_t* __restrict__ bnodes,
                                                   uint32_t* __restrict__ bnode2depth) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_bnode) return;
  if (i == 0) bnodes[0] = DMSH_U32MAX;
  const uint32_t n = num_bnode + 1;
  uint32_t first, last;
  node_range<uint32_t>(idx2morton, n, i, first, last);
  const uint32_t split = node_split<uint32_t>(idx2morton, first, last);
  if (first == split) {
    bnodes[i * 3 + 1] = num_bnode + split;
  } else {
    bnodes[i * 3 + 1] = split;
    bnodes[split * 3] = i;
  }
  if (last == split + 1) {
    bnodes[i * 3 + 2] = num_bnode + split + 1;
  } else {
    bnodes[i * 3 + 2] = split + 1;
    bnodes[(split + 1) * 3] = i;
  }
  const uint32_t delta = (uint32_t)clz_key(idx2morton[split] ^ idx2morton[split + 1]);
  const uint32_t offset = 32u - num_dim * max_depth;
  bnode2depth[i] = (delta - offset) / num_dim;  // u32 arithmetic, as the reference
}

// flags[i - 1] for branch i >= 1 (num_bnode - 1 entries)
__global__ __launch_bounds__(256) void k_bnode_flags(uint32_t num_bnode,
                                                     const uint32_t* __restrict__ bnodes,
                                                     const uint32_t* __restrict__ bnode2depth,
                                                     uint32_t* __restrict__ flags,
                                                     uint32_t* __restrict__ idx2bnode,
                                                     uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_bnode) return;
#pragma unroll
  for (int c = 1; c <= 2; ++c) {
    const uint32_t ch = bnodes[i * 3 + c];
    if (ch >= num_bnode) {
      const uint32_t idx = ch - num_bnode;
      if (idx <= num_bnode) idx2bnode[idx] = i;
      else raise_error(err, DMSH_DEVERR_INDEX);
    }
  }
  if (i == 0) return;
  const uint32_t p = bnodes[i * 3];
  if (p >= num_bnode) {
    raise_error(err, DMSH_DEVERR_INDEX);
    flags[i - 1] = 0u;
    return;
  }
  flags[i - 1] = bnode2depth[i] != bnode2depth[p] ? 1u : 0u;
}

__device__ __forceinline__ void morton2center(uint32_t morton, uint32_t num_dim, uint32_t depth,
                                              uint32_t max_depth, float* center) {
  const uint32_t sh = (max_depth - depth) * num_dim;
  uint32_t key = sh < 32u ? morton >> sh : 0u;
  for (uint32_t d = 0; d < num_dim; ++d) center[d] = 0.5f;
  for (uint32_t k = 0; k < depth; ++k) {
    for (uint32_t d = 0; d < num_dim; ++d) {
      const uint32_t j = num_dim - d - 1;
      center[j] += (float)(key & 1u);
      center[j] *= 0.5f;
      key >>= 1;
    }
  }
}

// The onode that holds branch b's subtree: climb while the branch shares its
// onode number with its predecessor (the reference's test for a group head).
__device__ __forceinline__ bool onode_of(uint32_t b, uint32_t num_bnode,
                                         const uint32_t* __restrict__ bnodes,
                                         const uint32_t* __restrict__ bnode2onode,
                                         const uint32_t* __restrict__ bnode2depth,
                                         uint32_t& onode, uint32_t& depth) {
  for (uint32_t steps = 0; steps <= num_bnode; ++steps) {
    if (b >= num_bnode) return false;
    if (b == 0 || bnode2onode[b] != bnode2onode[b - 1]) {
      onode = bnode2onode[b];
      depth = bnode2depth[b];
      return true;
    }
    b = bnodes[b * 3];
  }
  return false;
}

// One thread per point i (and, for i < num_bnode, per branch i).
__global__ __launch_bounds__(256) void k_octree_make(
    uint32_t num_vtx, const uint32_t* __restrict__ bnodes, const uint32_t* __restrict__ bnode2onode,
    const uint32_t* __restrict__ bnode2depth, const uint32_t* __restrict__ idx2bnode,
    const uint32_t* __restrict__ idx2morton, uint32_t num_onode, uint32_t max_depth,
    uint32_t num_dim, uint32_t* __restrict__ onodes, uint32_t* __restrict__ onode2depth,
    float* __restrict__ onode2center, uint32_t* __restrict__ idx2onode,
    float* __restrict__ idx2center, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  const uint32_t num_bnode = num_vtx - 1;
  const uint32_t num_child = 1u << num_dim, nlink = num_child + 1;
  {  // the point
    uint32_t op, dp;
    if (!onode_of(idx2bnode[i], num_bnode, bnodes, bnode2onode, bnode2depth, op, dp) ||
        op >= num_onode || dp >= max_depth) {
      raise_error(err, DMSH_DEVERR_INDEX);
    } else {
      const uint32_t key = idx2morton[i];
      const uint32_t c = (key >> ((max_depth - 1 - dp) * num_dim)) & (num_child - 1);
      onodes[(size_t)op * nlink + 1 + c] = i + num_onode;
      idx2onode[i] = op;
      float center[3];
      morton2center(key, num_dim, max_depth, max_depth, center);
      for (uint32_t d = 0; d < num_dim; ++d) idx2center[(size_t)i * num_dim + d] = center[d];
    }
  }
  if (i == num_vtx - 1) return;
  const uint32_t b = i;  // the branch
  if (b == 0) {
    for (uint32_t d = 0; d < num_dim; ++d) onode2center[d] = 0.5f;
    return;
  }
  if (bnode2onode[b - 1] == bnode2onode[b]) return;
  uint32_t op, dp;
  if (!onode_of(bnodes[b * 3], num_bnode, bnodes, bnode2onode, bnode2depth, op, dp) ||
      op >= num_onode || dp >= max_depth || bnode2onode[b] >= num_onode) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t morton = idx2morton[b];
  const uint32_t c = (morton >> ((max_depth - 1 - dp) * num_dim)) & (num_child - 1);
  const uint32_t o = bnode2onode[b];
  onodes[(size_t)op * nlink + 1 + c] = o;
  onodes[(size_t)o * nlink] = op;
  const uint32_t depth = bnode2depth[b];
  float center[3];
  morton2center(morton, num_dim, depth, max_depth, center);
  for (uint32_t d = 0; d < num_dim; ++d) onode2center[(size_t)o * num_dim + d] = center[d];
  onode2depth[o] = depth;
}

constexpr uint32_t kMaxHops = 16;  // node depth grows on every hop, up to 32 / num_dim

// hops[o] = number of parent links from onode o to the root
__global__ __launch_bounds__(256) void k_onode_hops(uint32_t num_onode, uint32_t num_link,
                                                    const uint32_t* __restrict__ onodes,
                                                    uint32_t* __restrict__ hops,
                                                    uint32_t* __restrict__ err) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= num_onode) return;
  uint32_t h = 0, p = onodes[(size_t)o * num_link];
  while (p != DMSH_U32MAX) {
    if (p >= num_onode || h >= kMaxHops) {
      raise_error(err, DMSH_DEVERR_INDEX);
      h = DMSH_U32MAX;
      break;
    }
    ++h;
    p = onodes[(size_t)p * num_link];
  }
  hops[o] = h;
}

// Per-node sums in a fixed order: a node adds its children in slot order (a
// point child contributes its idx2val row, a node child its own sum),
// starting from 0.  Launched once per level, deepest first; a node's
// children are one hop deeper, so their sums are final when it runs.
__global__ __launch_bounds__(256) void k_onode_sum_level(uint32_t num_onode, uint32_t num_vdim,
                                                         uint32_t num_link,
                                                         const uint32_t* __restrict__ onodes,
                                                         const float* __restrict__ idx2val,
                                                         uint32_t num_idx,
                                                         const uint32_t* __restrict__ hops,
                                                         uint32_t level, float* __restrict__ out,
                                                         uint32_t* __restrict__ err) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= num_onode || hops[o] != level) return;
  for (uint32_t d = 0; d < num_vdim; ++d) {
    float acc = 0.f;
    for (uint32_t c = 1; c < num_link; ++c) {
      const uint32_t ch = onodes[(size_t)o * num_link + c];
      if (ch == DMSH_U32MAX) continue;
      if (ch >= num_onode) {
        const uint32_t idx = ch - num_onode;
        if (idx >= num_idx) {
          raise_error(err, DMSH_DEVERR_INDEX);
          continue;
        }
        acc += idx2val[(size_t)idx * num_vdim + d];
      } else {
        acc += out[(size_t)ch * num_vdim + d];
      }
    }
    out[(size_t)o * num_vdim + d] = acc;
  }
}

// ---------------------------------------------------------------------------
// sorted-array helpers (del-msh-cpu/src/array1d.rs:1-15, map_idx.rs:83-115,
// del-msh-dlpack/src/array1d.rs:156-300, offset_array.rs)
// ---------------------------------------------------------------------------
// flags[k] = (val[k + 1] != val[k]), k < n - 1; the exclusive scan of them is
// idx2jdx
// new2val[i] = old2val[new2old[i]] (del-msh-dlpack/src/array1d.rs:17-70); an
// index out of range raises DMSH_DEVERR_INDEX and leaves that entry as it was.
__global__ __launch_bounds__(256) void k_permute_u32(uint32_t n, const uint32_t* __restrict__ old2val,
                                                     const uint32_t* __restrict__ new2old,
                                                     uint32_t* __restrict__ new2val,
                                                     uint32_t* __restrict__ err) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t o = new2old[i];
    if (o >= n) {
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    new2val[i] = old2val[o];
  }
}

__global__ __launch_bounds__(256) void k_change_flags(uint32_t n, const uint32_t* __restrict__ val,
                                                      uint32_t* __restrict__ flags,
                                                      uint32_t* __restrict__ dup) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k + 1 >= n) return;
  const bool same = val[k + 1] == val[k];
  if (flags) flags[k] = same ? 0u : 1u;
  if (dup && same) atomicOr(dup, 1u);
}

// jdx2idx_offset[jdx] = first idx with idx2jdx[idx] >= jdx (jdx <= num_jdx;
// map_idx.rs:83-115) and jdx2val[jdx] = val at that idx, for a non-decreasing
// idx2jdx that starts at 0 and steps by at most 1
__global__ __launch_bounds__(256) void k_unique_scatter(uint32_t n, const uint32_t* __restrict__ val,
                                                        const uint32_t* __restrict__ idx2jdx,
                                                        uint32_t num_jdx,
                                                        uint32_t* __restrict__ jdx2val,
                                                        uint32_t* __restrict__ jdx2idx_offset,
                                                        uint32_t* __restrict__ err) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx > n) return;
  if (idx == n) {
    jdx2idx_offset[num_jdx] = n;
    return;
  }
  const uint32_t j = idx2jdx[idx];
  if (idx == 0 || idx2jdx[idx - 1] != j) {
    if (j >= num_jdx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      return;
    }
    jdx2idx_offset[j] = idx;
    jdx2val[j] = val[idx];
  }
}

// idx2aggval[idx] += sum over j in [off[idx], off[idx + 1]) of kdx2val[jdx2kdx[j]],
// in j order; one thread per (idx, dimension)
__global__ __launch_bounds__(256) void k_offset_aggregate(uint32_t num_idx, uint32_t num_dim,
                                                          const uint32_t* __restrict__ off,
                                                          uint32_t num_jdx,
                                                          const uint32_t* __restrict__ jdx2kdx,
                                                          uint32_t num_kdx,
                                                          const float* __restrict__ kdx2val,
                                                          float* __restrict__ idx2aggval,
                                                          uint32_t* __restrict__ err) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)num_idx * num_dim) return;
  const uint32_t idx = (uint32_t)(t / num_dim), d = (uint32_t)(t % num_dim);
  float acc = idx2aggval[t];
  const uint32_t j0 = off[idx], j1 = off[idx + 1];
  if (j1 > num_jdx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  for (uint32_t j = j0; j < j1; ++j) {
    const uint32_t k = jdx2kdx[j];
    if (k >= num_kdx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    acc += kdx2val[(size_t)k * num_dim + d];
  }
  idx2aggval[t] = acc;
}

}  // namespace oct
}  // namespace dmsh
