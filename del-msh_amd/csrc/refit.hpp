// refit.hpp -- bottom-up AABB refit of an LBVH on gfx950.
//
// One thread per leaf (leaves are rows num_leaf-1 .. 2 num_leaf-2, in sorted
// Morton order, so neighbouring threads touch neighbouring geometry).  The
// leaf box is set from the first vertex and grown by the others
// (del-msh-cpu/src/vtx2xyz.rs:39-61 with margin eps); a thread then climbs to
// the parent and takes the parent's arrival counter.  The first arrival stops;
// the second merges the two children in fixed (left, right) order
// (bvhnode2aabb3.rs:136-141) and continues upward, so the result does not
// depend on which child arrives last.
//
// Memory ordering: the counter update is an agent-scope acq_rel RMW.  Its
// release half publishes the arriving thread's box stores (issued earlier in
// program order) to every XCD; its acquire half makes the sibling's box,
// published by the first arrival, visible to the second.  The reference CUDA
// kernel fences only after the RMW (bvhnode2aabb.cu:50-55), which is not
// enough on a weakly ordered multi-L2 part.
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {

struct Box {
  float v[6];
};

__device__ __forceinline__ void box_set(Box& b, const float* p, float eps) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b.v[d] = p[d] - eps;
    b.v[d + 3] = p[d] + eps;
  }
}
__device__ __forceinline__ void box_add(Box& b, const float* p, float eps) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b.v[d] = fmin_sel(b.v[d], p[d] - eps);
    b.v[d + 3] = fmax_sel(b.v[d + 3], p[d] + eps);
  }
}
__device__ __forceinline__ Box box_merge(const Box& a, const Box& b) {
  Box r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.v[d] = fmin_sel(a.v[d], b.v[d]);
    r.v[d + 3] = fmax_sel(a.v[d + 3], b.v[d + 3]);
  }
  return r;
}
__device__ __forceinline__ Box box_load(const float* __restrict__ aabb, uint32_t node) {
  const float2* p = reinterpret_cast<const float2*>(aabb + (size_t)node * 6);
  const float2 a = p[0], b = p[1], c = p[2];
  return Box{{a.x, a.y, b.x, b.y, c.x, c.y}};
}
__device__ __forceinline__ void box_store(float* __restrict__ aabb, uint32_t node, const Box& b) {
  float2* p = reinterpret_cast<float2*>(aabb + (size_t)node * 6);
  p[0] = make_float2(b.v[0], b.v[1]);
  p[1] = make_float2(b.v[2], b.v[3]);
  p[2] = make_float2(b.v[4], b.v[5]);
}

// TRI: leaves index triangles of tri2vtx; otherwise leaves index points.
template <bool TRI>
__global__ __launch_bounds__(256) void k_refit(float* __restrict__ aabb,
                                               uint32_t* __restrict__ flags, uint32_t num_leaf,
                                               const uint32_t* __restrict__ bvhnodes,
                                               const uint32_t* __restrict__ tri2vtx,
                                               uint32_t num_vtx, const float* __restrict__ xyz,
                                               float eps, uint32_t* __restrict__ err) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= num_leaf) return;
  const uint32_t nb = num_leaf - 1;
  uint32_t node = nb + k;
  const uint32_t* row = bvhnodes + (size_t)node * 3;
  const uint32_t parent0 = row[0], obj = row[1], tag = row[2];
  if (tag != DMSH_U32MAX || obj >= (TRI ? num_leaf : num_vtx)) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  Box mine;
  if (TRI) {
    const uint32_t* tv = tri2vtx + (size_t)obj * 3;
    const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
    if (i0 >= num_vtx || i1 >= num_vtx || i2 >= num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      return;
    }
    box_set(mine, xyz + (size_t)i0 * 3, eps);
    box_add(mine, xyz + (size_t)i1 * 3, eps);
    box_add(mine, xyz + (size_t)i2 * 3, eps);
  } else {
    box_set(mine, xyz + (size_t)obj * 3, 0.f);
  }
  box_store(aabb, node, mine);
  uint32_t parent = parent0;
  // A climb visits at most nb internal nodes; the bound keeps a corrupt
  // (cyclic) parent chain from looping forever.
  for (uint32_t steps = 0; parent != DMSH_U32MAX; ++steps) {
    if (parent >= nb || steps >= nb) {
      raise_error(err, DMSH_DEVERR_REFIT);
      return;
    }
    const uint32_t old = __hip_atomic_fetch_add(flags + parent, 1u, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;  // the sibling subtree finishes this node
    const uint32_t* prow = bvhnodes + (size_t)parent * 3;
    const uint32_t c0 = prow[1], c1 = prow[2];
    if (c0 > 2 * nb || c1 > 2 * nb) {
      raise_error(err, DMSH_DEVERR_REFIT);
      return;
    }
    if (c0 == node) {
      mine = box_merge(mine, box_load(aabb, c1));
    } else if (c1 == node) {
      mine = box_merge(box_load(aabb, c0), mine);
    } else {
      raise_error(err, DMSH_DEVERR_REFIT);
      return;
    }
    box_store(aabb, parent, mine);
    node = parent;
    parent = prow[0];
  }
}

template <bool TRI>
inline int launch_refit(hipStream_t s, float* aabb, uint32_t* flags, uint32_t num_leaf,
                        const uint32_t* bvhnodes, const uint32_t* tri2vtx, uint32_t num_vtx,
                        const float* xyz, float eps, uint32_t* err) {
  if (num_leaf == 0) return DMSH_OK;
  hipLaunchKernelGGL(k_refit<TRI>, dim3(ceil_div(num_leaf, 256)), dim3(256), 0, s, aabb, flags,
                     num_leaf, bvhnodes, tri2vtx, num_vtx, xyz, eps, err);
  return launch_status();
}

}  // namespace dmsh
