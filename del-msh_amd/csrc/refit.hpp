// refit.hpp -- bottom-up AABB refit of an LBVH on gfx950.
//
// Generic refit (any tree in the reference layout, k_refit): one thread per
// leaf (leaves are rows num_leaf-1 .. 2 num_leaf-2, in sorted Morton order,
// so neighbouring threads touch neighbouring geometry).  The leaf box is set
// from the first vertex and grown by the others (del-msh-cpu/src/
// vtx2xyz.rs:39-61 with margin eps); a thread then climbs to the parent and
// takes the parent's arrival counter.  The first arrival stops; the second
// merges the two children in fixed (left, right) order (bvhnode2aabb3.rs:
// 136-141) and continues upward, so the result does not depend on which child
// arrives last.
//
// Memory ordering: the counter update is an agent-scope acq_rel RMW.  Its
// release half publishes the arriving thread's box stores (issued earlier in
// program order) to every XCD; its acquire half makes the sibling's box,
// published by the first arrival, visible to the second.  The reference CUDA
// kernel fences only after the RMW (bvhnode2aabb.cu:50-55), which is not
// enough on a weakly ordered multi-L2 part.
//
// Trees built here take the treelet refit below instead (same merge order,
// same results, no agent-scope atomics), and so does the standalone refit of
// any tree whose node ranges pass the LBVH-layout check (k_leaf_ranges,
// k_list_treelets); the generic climb is its fallback for other trees.
#pragma once

#include <algorithm>

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
// Element i of an LDS box array, addressed with a 24-bit multiply (indices
// here are LDS slots, far below 2^24).
__device__ __forceinline__ const Box& lds_box(const Box* arr, uint32_t i) {
  return *reinterpret_cast<const Box*>(reinterpret_cast<const char*>(arr) +
                                       __umul24(i, (uint32_t)sizeof(Box)));
}
__device__ __forceinline__ Box& lds_box(Box* arr, uint32_t i) {
  return *reinterpret_cast<Box*>(reinterpret_cast<char*>(arr) + __umul24(i, (uint32_t)sizeof(Box)));
}

// box_merge(mine, other) when mine_left, else box_merge(other, mine), with
// the operands selected first so that both orders run the same instructions.
__device__ __forceinline__ Box box_merge_side(const Box& mine, const Box& other, bool mine_left) {
  Box a, b;
#pragma unroll
  for (int d = 0; d < 6; ++d) {
    a.v[d] = mine_left ? mine.v[d] : other.v[d];
    b.v[d] = mine_left ? other.v[d] : mine.v[d];
  }
  return box_merge(a, b);
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

// From leaf row `node` (whose box `mine` is stored), climb through the
// parents' arrival counters; the second arrival merges in (left, right) order.
__device__ __forceinline__ void refit_climb(float* __restrict__ aabb, uint32_t* __restrict__ flags,
                                            uint32_t nb, const uint32_t* __restrict__ bvhnodes,
                                            uint32_t node, uint32_t parent, Box mine,
                                            uint32_t* __restrict__ err) {
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

// What a leaf's object is, and so how its box is made.
constexpr int kLeafPoint = 0;  // point obj of xyz (margin 0)
constexpr int kLeafTri = 1;    // triangle obj of tri2vtx (margin eps)
constexpr int kLeafElem = 2;   // element obj of a CSR mesh: elem2idx over idx2vtx (margin 0)

// One tree (one mesh of a batch): node rows and boxes, and the leaves'
// geometry.  For kLeafElem, tri2vtx is the CSR vertex list idx2vtx.
struct TreeView {
  float* aabb;
  const uint32_t* bvhnodes;
  uint32_t num_leaf;
  const uint32_t* tri2vtx;
  uint32_t num_vtx;
  const float* xyz;
  const float* tri_box = nullptr;  // per-triangle boxes in triangle order, if built
  const uint32_t* elem2idx = nullptr;  // kLeafElem: num_leaf + 1 offsets into tri2vtx
  uint32_t num_idx = 0;
};

// Where a refit also writes the traversal records of the nodes it fits
// (batched builds for the serving cache): NoPack writes none.  A sink has
// leaf_ok(k, p0, p1, p2, tri) for triangle leaf k with its corners,
// leaf_bad(k) for a leaf whose indices do not fit, inner_at(i, c0, c1, b0,
// b1, num_leaf, err) for internal node i with children c0, c1 and their
// boxes, and inner_part(i, part, c0, c1, x0..x3, num_leaf, err) for one
// 16-byte part of that record.
struct NoPack {
  static constexpr bool kOn = false;
};

// Leaf box of leaf k and its parent; returns false (and raises) on a bad
// index.  Points: the point.  Triangles: the box of the three corners grown
// by eps (vtx2xyz.rs:39-61), or, with tri_box (the triangle boxes the
// build's centroid pass wrote, eps = 0), one 24-byte read.  Elements: the box
// of the element's vertices, first vertex first (bvhnode2aabb3.rs:211-228
// with margin 0; bvhnode2aabb.cu:67-122); an empty element is an index error.
template <int LK>
__device__ __forceinline__ bool leaf_box(const TreeView& v, uint32_t k, float eps, Box& box,
                                         uint32_t& parent, uint32_t* err) {
  const uint32_t* row = v.bvhnodes + (size_t)(v.num_leaf - 1 + k) * 3;
  const uint32_t p = row[0], obj = row[1], tag = row[2];
  if (tag != DMSH_U32MAX || obj >= (LK == kLeafPoint ? v.num_vtx : v.num_leaf)) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return false;
  }
  if (LK == kLeafTri && v.tri_box) {
    box = box_load(v.tri_box, obj);
  } else if (LK == kLeafTri) {
    const uint32_t* tv = v.tri2vtx + (size_t)obj * 3;
    const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
    if (i0 >= v.num_vtx || i1 >= v.num_vtx || i2 >= v.num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      return false;
    }
    box_set(box, v.xyz + (size_t)i0 * 3, eps);
    box_add(box, v.xyz + (size_t)i1 * 3, eps);
    box_add(box, v.xyz + (size_t)i2 * 3, eps);
  } else if (LK == kLeafElem) {
    const uint32_t i0 = v.elem2idx[obj], i1 = v.elem2idx[obj + 1];
    if (i0 >= i1 || i1 > v.num_idx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      return false;
    }
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t q = v.tri2vtx[i];
      if (q >= v.num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
        return false;
      }
      if (i == i0) box_set(box, v.xyz + (size_t)q * 3, 0.f);
      else box_add(box, v.xyz + (size_t)q * 3, 0.f);
    }
  } else {
    box_set(box, v.xyz + (size_t)obj * 3, 0.f);
  }
  parent = p;
  return true;
}

// leaf_box of a triangle leaf (no tri_box) that also hands the leaf's
// traversal record to the pack sink from the same loads: the three corners,
// or a degenerate record after an index that does not fit.
template <typename PK>
__device__ __forceinline__ bool leaf_tri_pack(const TreeView& v, uint32_t k, float eps, Box& box,
                                              uint32_t& parent, const PK& pk, uint32_t* err) {
  const uint32_t* row = v.bvhnodes + (size_t)(v.num_leaf - 1 + k) * 3;
  const uint32_t p = row[0], obj = row[1], tag = row[2];
  bool ok = tag == DMSH_U32MAX && obj < v.num_leaf;
  uint32_t i0 = 0, i1 = 0, i2 = 0;
  if (ok) {
    const uint32_t* tv = v.tri2vtx + (size_t)obj * 3;
    i0 = tv[0];
    i1 = tv[1];
    i2 = tv[2];
    ok = i0 < v.num_vtx && i1 < v.num_vtx && i2 < v.num_vtx;
  }
  if (!ok) {
    raise_error(err, DMSH_DEVERR_INDEX);
    pk.leaf_bad(k);
    return false;
  }
  float c[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[0][d] = v.xyz[(size_t)i0 * 3 + d];
    c[1][d] = v.xyz[(size_t)i1 * 3 + d];
    c[2][d] = v.xyz[(size_t)i2 * 3 + d];
  }
  box_set(box, c[0], eps);
  box_add(box, c[1], eps);
  box_add(box, c[2], eps);
  pk.leaf_ok(k, c[0], c[1], c[2], obj);
  parent = p;
  return true;
}

// The generic climb, one thread per leaf (grid-stride).  With `gate`, the
// kernel runs only if *gate is nonzero: the standalone refit sets it for a
// tree that fails the LBVH-layout check.
template <int LK>
__global__ __launch_bounds__(256) void k_refit(TreeView v, uint32_t* __restrict__ flags, float eps,
                                               const uint32_t* __restrict__ gate,
                                               uint32_t* __restrict__ err) {
  if (gate && *gate == 0u) return;
  const uint32_t nb = v.num_leaf - 1;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < v.num_leaf;
       k += gridDim.x * blockDim.x) {
    Box mine;
    uint32_t parent;
    if (!leaf_box<LK>(v, k, eps, mine, parent, err)) continue;
    box_store(v.aabb, nb + k, mine);
    refit_climb(v.aabb, flags, nb, v.bvhnodes, nb + k, parent, mine, err);
  }
}

// ---------------------------------------------------------------------------
// Treelet refit for LBVHs built here (the topology kernel lists the
// treelets: subtrees of at most K leaves whose parent has more than K).
//
// refit_treelet: one workgroup, one treelet.  A treelet over leaves [f, l]
//   keeps its internal nodes at ids inside [f, l], so its node rows, arrival
//   counters and boxes fit in LDS slots (id - f).  Phase 1 builds the leaf
//   boxes (stored to HBM at once, in leaf order) and loads the node rows with
//   coalesced reads; phase 2 climbs from every leaf with LDS counters (the
//   second arrival merges, left child first) up to the treelet root; phase 3
//   writes the internal boxes back in id order.  Every node is written once
//   and no global atomics are used.
// refit_top: one workgroup climbs from a tree's treelet roots through the
//   nodes above them (at most ~2n/K of them for a balanced tree).  Those
//   climbs all run on one CU, so workgroup-scope acq_rel RMWs on the global
//   counters order the box stores and loads (one L1, no cross-XCD traffic) and
//   no L2 write-back is needed.
template <int K>
struct TreeletLds {
  Box leaf[K];
  Box inner[K];
  uint32_t flag[K];
  uint2 child[K];
  uint32_t parent[K];
  uint16_t queue[K];   // CLIMB 2: the nodes whose children are both leaves
  uint16_t queue2[K];  // CLIMB 2: the climbs still running after the first levels
  uint32_t queued, queued2;  // zero between treelets
};

// What the treelet pass leaves for the top merge, in top-slot order, so that
// the one workgroup of the top merge reads it with coalesced loads instead
// of gathering it: kid[e] packs the slots of top slot e's top children (16
// bits each; 0xFFFF for a child below the top, 0xFFFE for a slot out of
// range) and the box of a child below the top (a treelet root; side 0 is
// the left child) component-major, bound d of side s at
// box[(6 s + d) kTopClimbSlots + e], so that consecutive slots' bounds are
// consecutive words.  Kept for slots below kTopClimbSlots.
// The lister of the treelets (topology kernel, k_list_treelets) knows the
// parent's slot when it lists a treelet and puts 2 e + side in the entry's
// last word (all ones: no parent slot).
constexpr int kTopClimbSlots = 4096;

struct TopStage {
  const uint32_t* top_count = nullptr;
  const uint4* top_list = nullptr;
  const uint32_t* node2slot = nullptr;
  uint32_t* kid = nullptr;
  float* box = nullptr;
};

// A treelet's inputs when its leaf boxes are already in v.aabb (PRE): the
// node rows, the leaf rows' parents and the leaf boxes of this thread's
// slots.  None of these loads depends on another, so they can be issued
// together, and a persistent workgroup issues the next treelet's while it
// climbs the current one.
template <int LPT>
struct TreeletIn {
  uint32_t row[LPT][3];
  uint32_t lp[LPT];
  Box lb[LPT];
};

template <int K, int T>
__device__ __forceinline__ void treelet_prefetch(const TreeView& v, uint32_t f, uint32_t l,
                                                 TreeletIn<K / T>& in) {
  const uint32_t nb = v.num_leaf - 1;
  if (f > l || l >= v.num_leaf || l - f >= (uint32_t)K) return;  // refit_treelet raises
  const uint32_t m = l - f + 1;
#pragma unroll
  for (int r = 0; r < K / T; ++r) {
    const uint32_t j = (uint32_t)r * T + threadIdx.x;
    if (j >= m) continue;
    const uint32_t id = f + j;
    if (id < nb) {
      const uint32_t* rw = v.bvhnodes + (size_t)id * 3;
      in.row[r][0] = rw[0];
      in.row[r][1] = rw[1];
      in.row[r][2] = rw[2];
    }
    in.lp[r] = v.bvhnodes[(size_t)(nb + id) * 3];
    in.lb[r] = box_load(v.aabb, nb + id);
  }
}

// PRE: the leaf boxes are already in v.aabb (the kernel before wrote them);
// phase 1 takes this thread's inputs from `pre` (treelet_prefetch).
// CLIMB 0: a level reads the parent's links, then takes its arrival counter
// (workgroup-scope acq_rel), then reads the sibling's box: three dependent
// LDS round trips.  CLIMB 1: the links of the next parent are read one level
// ahead, and the counter, the sibling's box and those links go out together,
// one LDS round trip per level.  The box is read whether or not this is the
// second arrival and used only if it is: LDS operations of a wave complete
// in issue order and the LDS performs them one at a time, so the sibling's
// box store, issued before its own counter update, has been performed when
// that update has; the counter RMW is ordered with the wave's own LDS
// accesses at wavefront scope (no waits), which keeps the compiler from
// moving the box store after it or the box load before it.
// PK: a pack sink (see NoPack): leaf records in phase 1, internal records
// (children's boxes from LDS) in phase 3.
template <int LK, int K, int T, bool PRE = false, int CLIMB = 0, typename PK = NoPack>
__device__ __forceinline__ void refit_treelet(const TreeView& v, uint32_t root, uint32_t f,
                                              uint32_t l, float eps, TreeletLds<K>& s,
                                              uint32_t* __restrict__ err,
                                              const TreeletIn<K / T>* pre = nullptr,
                                              int compact = 4, bool staged = false,
                                              TopStage st = TopStage{},
                                              uint32_t stage_pos = DMSH_U32MAX,
                                              const PK& pk = PK{}) {
  constexpr int LPT = K / T;
  const uint32_t num_leaf = v.num_leaf;
  const uint32_t nb = num_leaf - 1;
  if (f > l || l >= num_leaf || l - f >= (uint32_t)K) {  // uniform over the workgroup
    if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
    return;
  }
  const uint32_t m = l - f + 1;
  uint32_t lpar[LPT];
  // phase 1
  if (PRE) {
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
      const uint32_t j = (uint32_t)r * T + threadIdx.x;
      lpar[r] = DMSH_U32MAX - 1;  // no climb
      if (j >= m) continue;
      const uint32_t id = f + j;
      s.flag[j] = 0u;
      if (id < nb) {
        s.parent[j] = pre->row[r][0];
        s.child[j] = make_uint2(pre->row[r][1], pre->row[r][2]);
      }
      s.leaf[j] = pre->lb[r];
      if (nb + id != root) lpar[r] = pre->lp[r];
    }
  }
#pragma unroll
  for (int r = 0; r < LPT; ++r) {
    if (PRE) break;
    const uint32_t j = (uint32_t)r * T + threadIdx.x;
    lpar[r] = DMSH_U32MAX - 1;  // no climb
    if (j >= m) continue;
    const uint32_t id = f + j;
    s.flag[j] = 0u;
    if (id < nb) {
      const uint32_t* row = v.bvhnodes + (size_t)id * 3;
      s.parent[j] = row[0];
      s.child[j] = make_uint2(row[1], row[2]);
      if (CLIMB == 2) {
        // a child that is a leaf of this treelet has arrived already; a node
        // whose children are both leaves is merged first, from the queue
        // (an id in [f, l] outside the treelet has no leaf child in it)
        const uint32_t lo = nb + f, hi = nb + l;
        const uint32_t nl = (row[1] >= lo && row[1] <= hi ? 1u : 0u) +
                            (row[2] >= lo && row[2] <= hi ? 1u : 0u);
        s.flag[j] = nl;
        if (nl == 2u) s.queue[atomicAdd(&s.queued, 1u)] = (uint16_t)j;
      }
    }
    Box b;
    uint32_t p;
    bool fit;
    if constexpr (PK::kOn) {
      static_assert(LK == kLeafTri, "traversal records are made for triangle leaves");
      fit = leaf_tri_pack(v, id, eps, b, p, pk, err);
    } else {
      fit = leaf_box<LK>(v, id, eps, b, p, err);
    }
    if (fit) {
      s.leaf[j] = b;
      // with a pack sink only the treelet root's box is stored (for the top
      // merge); the traversal records hold every other box
      if (!PK::kOn || nb + id == root) box_store(v.aabb, nb + id, b);
      if (nb + id != root) lpar[r] = p;
    }
  }
  __syncthreads();
  // phase 2: a thread's LPT climbs advance in lock-step, one level per round
  // (a round costs as much as the longest of them, not their sum); the
  // parent's static links are read before its arrival counter, so a level
  // costs the counter and the sibling's box.  Two climbs of one thread that
  // meet at a node arrive in program order; the box each reads was stored in
  // an earlier round.
  uint32_t cnode[LPT], cpar[LPT];
  Box mine[LPT];
  bool act[LPT];
#pragma unroll
  for (int r = 0; r < LPT; ++r) {
    const uint32_t j = (uint32_t)r * T + threadIdx.x;
    cpar[r] = lpar[r];
    act[r] = cpar[r] != DMSH_U32MAX - 1;
    cnode[r] = nb + f + j;
    if (act[r]) mine[r] = s.leaf[j];
  }
  if constexpr (CLIMB == 2) {
    // CLIMB 2: for trees whose layout is known good (built here, or checked
    // by k_list_treelets).  Leaves do not climb: the nodes whose children
    // are both leaves (queued in phase 1) are merged by the first lanes of
    // the workgroup and climb from there, one LDS round trip per level as in
    // CLIMB 1.  After `compact` levels the climbs still running are queued
    // again and continued by the first lanes, so the climbs run on as few
    // waves as they need while the other waves wait at the barrier.
    static_assert(LPT == 1, "one leaf per thread");
    static_assert(!PRE, "the leaf boxes come from phase 1");
    const uint32_t starts = s.queued;
    bool a = threadIdx.x < starts;
    uint32_t node = DMSH_U32MAX, p = DMSH_U32MAX, gp = DMSH_U32MAX;
    uint2 ch = make_uint2(DMSH_U32MAX, DMSH_U32MAX);
    Box me;
    if (a) {
      const uint32_t j = s.queue[threadIdx.x];
      node = f + j;
      const uint2 c = s.child[j];
      me = box_merge(lds_box(s.leaf, c.x - nb - f), lds_box(s.leaf, c.y - nb - f));
      lds_box(s.inner, j) = me;
      if (node == root) {
        a = false;
      } else {
        p = s.parent[j];
        ch = s.child[p - f];
        gp = s.parent[p - f];
      }
    }
    for (int stage = 0; stage < 2; ++stage) {
      const uint32_t limit = (stage == 0 && compact > 0) ? (uint32_t)compact : (uint32_t)K;
      for (uint32_t steps = 0; a && steps < limit; ++steps) {
        const bool left = ch.x == node;
        const uint32_t sib = left ? ch.y : ch.x;
        const uint32_t pj = p - f;
        const uint32_t old = __hip_atomic_fetch_add(&s.flag[pj], 1u, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_WAVEFRONT);
        const Box sb = (sib >= nb) ? lds_box(s.leaf, sib - nb - f) : lds_box(s.inner, sib - f);
        uint2 ch2 = make_uint2(DMSH_U32MAX, DMSH_U32MAX);
        uint32_t gp2 = DMSH_U32MAX;
        if (p != root) {
          ch2 = s.child[gp - f];
          gp2 = s.parent[gp - f];
        }
        if (old == 0) {  // the sibling subtree finishes this node
          a = false;
          break;
        }
        me = box_merge_side(me, sb, left);
        lds_box(s.inner, pj) = me;
        if (p == root) {
          a = false;
          break;
        }
        node = p;
        p = gp;
        ch = ch2;
        gp = gp2;
      }
      if (stage == 1 || compact <= 0 || compact >= K) break;
      if (a) s.queue2[atomicAdd(&s.queued2, 1u)] = (uint16_t)(node - f);
      __syncthreads();
      const uint32_t q = s.queued2;
      a = threadIdx.x < q;
      if (a) {
        node = f + s.queue2[threadIdx.x];
        me = s.inner[node - f];
        p = s.parent[node - f];
        ch = s.child[p - f];
        gp = s.parent[p - f];
      }
    }
  }
  if (CLIMB == 1) {
    uint2 ch[LPT];
    uint32_t gp[LPT];
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
      ch[r] = make_uint2(DMSH_U32MAX, DMSH_U32MAX);
      gp[r] = DMSH_U32MAX;
      if (!act[r]) continue;
      const uint32_t parent = cpar[r];
      if (parent < f || parent > l || parent >= nb) continue;  // raised below
      ch[r] = s.child[parent - f];
      gp[r] = s.parent[parent - f];
    }
    for (uint32_t steps = 0;; ++steps) {
      uint32_t old[LPT];
      Box sb[LPT];
      uint2 ch2[LPT];
      uint32_t gp2[LPT];
#pragma unroll
      for (int r = 0; r < LPT; ++r) {
        if (!act[r]) continue;
        const uint32_t parent = cpar[r], node = cnode[r];
        const uint32_t sib = (ch[r].x == node) ? ch[r].y : ch[r].x;
        const uint32_t si = (sib >= nb) ? sib - nb - f : sib - f;
        if (parent < f || parent > l || parent >= nb || steps >= (uint32_t)K ||
            (ch[r].x != node && ch[r].y != node) || si >= m) {
          raise_error(err, DMSH_DEVERR_REFIT);
          act[r] = false;
          continue;
        }
        const uint32_t pj = parent - f;
        old[r] = __hip_atomic_fetch_add(&s.flag[pj], 1u, __ATOMIC_ACQ_REL,
                                        __HIP_MEMORY_SCOPE_WAVEFRONT);
        sb[r] = (sib >= nb) ? s.leaf[si] : s.inner[si];
        const uint32_t g = gp[r];
        ch2[r] = make_uint2(DMSH_U32MAX, DMSH_U32MAX);
        gp2[r] = DMSH_U32MAX;
        if (parent != root && g >= f && g <= l && g < nb) {
          ch2[r] = s.child[g - f];
          gp2[r] = s.parent[g - f];
        }
      }
      bool more = false;
#pragma unroll
      for (int r = 0; r < LPT; ++r) {
        if (!act[r]) continue;
        if (old[r] == 0) {  // the sibling subtree finishes this node
          act[r] = false;
          continue;
        }
        const uint32_t parent = cpar[r], pj = parent - f, node = cnode[r];
        mine[r] = (ch[r].x == node) ? box_merge(mine[r], sb[r]) : box_merge(sb[r], mine[r]);
        s.inner[pj] = mine[r];
        if (parent == root) {
          act[r] = false;
          continue;
        }
        cnode[r] = parent;
        cpar[r] = gp[r];
        ch[r] = ch2[r];
        gp[r] = gp2[r];
        more = true;
      }
      if (!more) break;
    }
  }
  for (uint32_t steps = 0; CLIMB == 0; ++steps) {
    uint2 ch[LPT];
    uint32_t gp[LPT], old[LPT];
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
      if (!act[r]) continue;
      const uint32_t parent = cpar[r];
      if (parent < f || parent > l || parent >= nb || steps >= (uint32_t)K) {
        raise_error(err, DMSH_DEVERR_REFIT);
        act[r] = false;
        continue;
      }
      const uint32_t pj = parent - f;
      ch[r] = s.child[pj];
      gp[r] = s.parent[pj];
      old[r] =
          __hip_atomic_fetch_add(&s.flag[pj], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    bool more = false;
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
      if (!act[r]) continue;
      if (old[r] == 0) {  // the sibling subtree finishes this node
        act[r] = false;
        continue;
      }
      const uint32_t parent = cpar[r], pj = parent - f, node = cnode[r];
      const uint32_t sib = (ch[r].x == node) ? ch[r].y : ch[r].x;
      const uint32_t si = (sib >= nb) ? sib - nb - f : sib - f;
      if ((ch[r].x != node && ch[r].y != node) || si >= m) {
        raise_error(err, DMSH_DEVERR_REFIT);
        act[r] = false;
        continue;
      }
      const Box sb = (sib >= nb) ? s.leaf[si] : s.inner[si];
      mine[r] = (ch[r].x == node) ? box_merge(mine[r], sb) : box_merge(sb, mine[r]);
      s.inner[pj] = mine[r];
      if (parent == root) {
        act[r] = false;
        continue;
      }
      cnode[r] = parent;
      cpar[r] = gp[r];
      more = true;
    }
    if (!more) break;
  }
  __syncthreads();
  // phase 3: the nodes that received both arrivals are this treelet's
#pragma unroll
  for (int r = 0; r < LPT; ++r) {
    const uint32_t j = (uint32_t)r * T + threadIdx.x;
    if (j < m && f + j < nb && s.flag[j] >= 2u) {
      if (!PK::kOn || f + j == root) box_store(v.aabb, f + j, s.inner[j]);
    }
  }
  if constexpr (PK::kOn) {
    // the internal records of ids f.. in 16-byte parts, consecutive threads
    // on consecutive parts, so that a wave's store is one contiguous
    // kilobyte; a part is made from the two child boxes in LDS (children
    // inside the treelet; slots clamped for a broken tree, whose climb has
    // raised already)
    for (uint32_t c = threadIdx.x; c < 4u * m; c += (uint32_t)T) {
      const uint32_t j = c >> 2, part = c & 3u;
      if (f + j >= nb || s.flag[j] < 2u) continue;
      const uint2 ch = s.child[j];
      const uint32_t s0 = ch.x >= nb ? ch.x - nb - f : ch.x - f;
      const uint32_t s1 = ch.y >= nb ? ch.y - nb - f : ch.y - f;
      const Box* b0 = ch.x >= nb ? &s.leaf[s0 < m ? s0 : 0u] : &s.inner[s0 < m ? s0 : 0u];
      const Box* b1 = ch.y >= nb ? &s.leaf[s1 < m ? s1 : 0u] : &s.inner[s1 < m ? s1 : 0u];
      float x0 = 0.f, x1 = 0.f, x2 = 0.f, x3 = 0.f;
      if (part == 0u) {
        x0 = b0->v[0]; x1 = b0->v[1]; x2 = b0->v[2]; x3 = b0->v[3];
      } else if (part == 1u) {
        x0 = b0->v[4]; x1 = b0->v[5]; x2 = b1->v[0]; x3 = b1->v[1];
      } else if (part == 2u) {
        x0 = b1->v[2]; x1 = b1->v[3]; x2 = b1->v[4]; x3 = b1->v[5];
      }
      pk.inner_part(f + j, part, ch.x, ch.y, x0, x1, x2, x3, num_leaf, err);
    }
  }
  // the root's box into its place in the top merge's staging (the
  // treelet's entry carries 2 * parent slot + side)
  if (staged && threadIdx.x == 0 && stage_pos < 2u * (uint32_t)kTopClimbSlots) {
    const Box b = root >= nb ? s.leaf[root - nb - f] : s.inner[root - f];
    const uint32_t e = stage_pos >> 1, side = stage_pos & 1u;
#pragma unroll
    for (int d = 0; d < 6; ++d) st.box[(side * 6u + (uint32_t)d) * kTopClimbSlots + e] = b.v[d];
  }
  if (CLIMB == 2 && threadIdx.x == 0) {
    s.queued = 0u;
    s.queued2 = 0u;
  }
  __syncthreads();  // LDS is reused by the next treelet
}

// Climbs from roots[e * stride], e < count, through the tree's top nodes.
// PK: a pack sink.  Its records are written after the climbs from the boxes
// they stored (a record stored during a climb would hold up the next level's
// counter RMW, which waits for the wave's stores): the merged nodes are
// listed in s_list (LDS, list_cap entries, *s_count zero on entry and on
// return); those past list_cap are written at the merge.
template <int T, typename PK = NoPack>
__device__ __forceinline__ void refit_top(const TreeView& v, uint32_t* __restrict__ flags,
                                          const uint32_t* __restrict__ roots, uint32_t stride,
                                          uint32_t count, uint32_t* __restrict__ err,
                                          const PK& pk = PK{}, uint32_t* s_list = nullptr,
                                          uint32_t* s_count = nullptr, uint32_t list_cap = 0) {
  const uint32_t nb = v.num_leaf - 1;
  for (uint32_t e = threadIdx.x; e < count; e += T) {
    uint32_t node = roots[(size_t)e * stride];
    if (node > 2 * nb) {
      raise_error(err, DMSH_DEVERR_REFIT);
      continue;
    }
    Box mine = box_load(v.aabb, node);
    uint32_t parent = v.bvhnodes[(size_t)node * 3];
    for (uint32_t steps = 0; parent != DMSH_U32MAX; ++steps) {
      if (parent >= nb || steps >= nb) {
        raise_error(err, DMSH_DEVERR_REFIT);
        break;
      }
      const uint32_t old = __hip_atomic_fetch_add(flags + parent, 1u, __ATOMIC_ACQ_REL,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) break;
      const uint32_t* prow = v.bvhnodes + (size_t)parent * 3;
      const uint32_t a = prow[1], b = prow[2];
      if (a > 2 * nb || b > 2 * nb || (a != node && b != node)) {
        raise_error(err, DMSH_DEVERR_REFIT);
        break;
      }
      const Box sib = box_load(v.aabb, a == node ? b : a);
      if constexpr (PK::kOn) {
        const uint32_t q = list_cap ? atomicAdd(s_count, 1u) : 0u;
        if (q < list_cap) {
          s_list[q] = parent;
        } else if (a == node) {
          pk.inner_at(parent, a, b, mine.v, sib.v, v.num_leaf, err);
        } else {
          pk.inner_at(parent, a, b, sib.v, mine.v, v.num_leaf, err);
        }
      }
      mine = (a == node) ? box_merge(mine, sib) : box_merge(sib, mine);
      box_store(v.aabb, parent, mine);
      node = parent;
      parent = prow[0];
    }
  }
  if constexpr (PK::kOn) {
    if (list_cap) {
      __syncthreads();  // the workgroup's boxes are stored and its list complete
      const uint32_t listed = *s_count < list_cap ? *s_count : list_cap;
      for (uint32_t e = threadIdx.x; e < listed; e += T) {
        const uint32_t parent = s_list[e];
        const uint32_t* prow = v.bvhnodes + (size_t)parent * 3;
        const uint32_t a = prow[1], b = prow[2];  // checked at the merge
        const Box ba = box_load(v.aabb, a), bb = box_load(v.aabb, b);
        pk.inner_at(parent, a, b, ba.v, bb.v, v.num_leaf, err);
      }
      __syncthreads();  // every thread has read the count
      if (threadIdx.x == 0) *s_count = 0u;
      __syncthreads();
    }
  }
}

// Persistent grid over the treelets of one tree.  With a zeroed `next`
// counter the workgroups take treelets from it one at a time (treelet sizes
// vary from 1 to K leaves, so a fixed deal would leave workgroups idle);
// without it, treelet t goes to workgroup t mod grid.  A nonzero *skip (the
// standalone refit's layout check failed) makes the kernel do nothing.
template <int LK, int K, int T, bool PRE, int CLIMB>
__global__ __launch_bounds__(T) __attribute__((amdgpu_num_sgpr(72))) void k_refit_treelet(TreeView v, float eps,
                                                     const uint32_t* __restrict__ treelet_count,
                                                     const uint4* __restrict__ treelets,
                                                     uint32_t* __restrict__ next,
                                                     const uint32_t* __restrict__ skip,
                                                     uint32_t* __restrict__ err, int compact,
                                                     TopStage st) {
  __shared__ TreeletLds<K> s;
  __shared__ uint32_t s_t, s_t2;
  if (skip && *skip) return;
  uint32_t count = *treelet_count;
  if (count > v.num_leaf) count = v.num_leaf;
  if (CLIMB == 2) {
    if (threadIdx.x == 0) {
      s.queued = 0u;
      s.queued2 = 0u;
    }
    __syncthreads();
  }
  const bool staged = st.kid != nullptr;
  if (staged) {
    // the top slots' top-child slots (a gather spread over the first
    // workgroups, so that the top merge's own workgroup need not do it)
    const uint32_t tc = *st.top_count, nb = v.num_leaf - 1;
    const uint32_t e = blockIdx.x * T + threadIdx.x;
    if (tc <= (uint32_t)kTopClimbSlots && e < tc) {
      const uint4 r = st.top_list[e];
      uint32_t sa = 0xFFFFu, sb = 0xFFFFu;
      if (r.w & 1u) sa = r.y < nb && st.node2slot[r.y] < tc ? st.node2slot[r.y] : 0xFFFEu;
      if (r.w & 2u) sb = r.z < nb && st.node2slot[r.z] < tc ? st.node2slot[r.z] : 0xFFFEu;
      st.kid[e] = sa | (sb << 16);
    }
  }
  if constexpr (PRE) {
    // Two treelets in flight per workgroup: the next one's loads are issued
    // before the current one's climb.  Treelet numbers run three deep: t is
    // being climbed, tn's leaves are loading, t3's descriptor is loading, and
    // thread 0 holds the ticket after t3 in a register until the climb is
    // over, so no wait on an atomic or a descriptor sits before a climb.
    // (Tickets come from `next`, or a fixed deal t, t + grid, ...)
    TreeletIn<K / T> cur, nxt;
    uint32_t t, tn, t3;
    if (next) {
      if (threadIdx.x == 0) {
        const uint32_t t0 = atomicAdd(next, 3u);
        s_t = t0;
      }
      __syncthreads();
      t = s_t;
      tn = t + 1u;
      t3 = t + 2u;
      __syncthreads();
    } else {
      t = blockIdx.x;
      tn = t + gridDim.x;
      t3 = tn + gridDim.x;
    }
    if (count == 0u) return;
    // An opaque zero in a vector register, added to the descriptor index: a
    // descriptor whose address the compiler can prove uniform is read with a
    // scalar load, and every LDS wait of the climb (lgkmcnt) would wait for
    // it too.  Descriptors of treelets past the end are never used, so their
    // index is clamped instead of selected.
    uint32_t vzero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(vzero));
    const uint32_t last = count - 1u;
    uint4 d = treelets[(t < count ? t : last) + vzero];
    uint4 dn = treelets[(tn < count ? tn : last) + vzero];
    if (t < count) treelet_prefetch<K, T>(v, d.y, d.z, cur);
    // vmcnt(0) (expcnt / lgkmcnt left alone): everything issued so far is
    // complete on entry, as it is at the end of every iteration, so the loop
    // body waits for nothing it issues before the climb
    __builtin_amdgcn_s_waitcnt(0x0F70);
    while (t < count) {
      if (tn < count) treelet_prefetch<K, T>(v, dn.y, dn.z, nxt);
      const uint4 d3 = treelets[(t3 < count ? t3 : last) + vzero];
      uint32_t t4 = 0u;
      // lane 0 only; the lane id in the address keeps the compiler from
      // turning it into a wave-wide atomic whose result is read at once,
      // so the wait for it (and for the loads before it) comes after the
      // climb, where the ticket is stored
      if (next && threadIdx.x == 0) t4 = atomicAdd(next + __lane_id(), 1u);
      refit_treelet<LK, K, T, true, CLIMB>(v, d.x, d.y, d.z, eps, s, err, &cur, compact, staged,
                                           st, d.w);
      if (next && threadIdx.x == 0) s_t2 = t4;
      __syncthreads();  // thread 0's ticket is visible (also after an early return)
      t4 = next ? s_t2 : t3 + gridDim.x;
      __syncthreads();  // read by everyone before thread 0 writes the next
      t = tn;
      d = dn;
      cur = nxt;
      tn = t3;
      dn = d3;
      t3 = t4;
    }
  } else if (!next) {
    for (uint32_t t = blockIdx.x; t < count; t += gridDim.x) {
      const uint4 tl = treelets[t];
      refit_treelet<LK, K, T, false, CLIMB>(v, tl.x, tl.y, tl.z, eps, s, err, nullptr, compact,
                                            staged, st, tl.w);
    }
  } else {
    // The first treelet of a workgroup is dealt (its block index), the later
    // ones come from `next` (offset by the grid).  Thread 0 asks for the next
    // ticket as a treelet starts and keeps it in a register until the treelet
    // is done, so the wait in the same-address atomic's queue overlaps the
    // treelet's loads instead of preceding them.
    uint32_t t = blockIdx.x;
    while (t < count) {  // uniform over the workgroup
      uint32_t tn = 0u;
      // the lane id in the address keeps the compiler from making it a
      // wave-wide atomic whose result is read at once
      if (threadIdx.x == 0) tn = gridDim.x + atomicAdd(next + __lane_id(), 1u);
      const uint4 tl = treelets[t];
      refit_treelet<LK, K, T, false, CLIMB>(v, tl.x, tl.y, tl.z, eps, s, err, nullptr, compact,
                                            staged, st, tl.w);
      if (threadIdx.x == 0) s_t = tn;
      __syncthreads();
      t = s_t;
      __syncthreads();  // every thread has read s_t before it is reused
    }
  }
}

template <int T>
__global__ __launch_bounds__(T) void k_refit_top(float* __restrict__ aabb,
                                                 uint32_t* __restrict__ flags, uint32_t num_leaf,
                                                 const uint32_t* __restrict__ bvhnodes,
                                                 const uint32_t* __restrict__ treelet_count,
                                                 const uint4* __restrict__ treelets,
                                                 uint32_t* __restrict__ err) {
  const TreeView v{aabb, bvhnodes, num_leaf, nullptr, 0u, nullptr};
  uint32_t count = *treelet_count;
  if (count > num_leaf) count = num_leaf;
  refit_top<T>(v, flags, reinterpret_cast<const uint32_t*>(treelets), 4u, count, err);
}

// The nodes above the treelets, merged in LDS by one workgroup.  The topology
// kernel lists them (top_list, node2slot; see TreeletSink); slot e of this
// kernel is top_list[e], owned by thread e mod T.
//   Phase 1 reads each slot's node row, maps its top children to their slots
//   (kept in registers) and reads the boxes of its other children -- treelet
//   roots or leaves, which the treelet pass wrote.  A node with no top child
//   is finished at once; a node with one keeps the other child's box in its
//   slot until the top child is done.
//   Phase 2 runs in rounds: in round t a thread finishes each of its pending
//   nodes whose top children were finished in an earlier round (done[] < t),
//   merging children in (left, right) order.  A round ends at a barrier, so a
//   box is read only after the round that wrote it; the number of rounds is
//   the height of the top tree, and no atomics are needed.
//   Phase 3 writes the boxes out.
// More than C top nodes (an unbalanced tree) fall back to the climb through
// global counters.
constexpr uint32_t kTopRef = 0x80000000u;
constexpr uint8_t kTopPending = 0xFFu;

// The nodes above the treelets of one tree merged in LDS by barrier rounds,
// by the whole workgroup: `count` top slots listed as (node, left child,
// right child, bit 0 / 1: that child is a top node) with node2slot; a child
// below the top (a treelet root) has its box in v.aabb.  Round t merges the
// slots whose top children finished before round t.  Then every slot's box
// is stored and, with a pack sink, its record written (the children's boxes
// from LDS or v.aabb).  Returns false, having changed nothing, when the list
// does not fit C slots or the rounds a byte counts: the caller climbs then.
template <int C, int T, typename PK = NoPack>
__device__ __forceinline__ bool top_rounds(const TreeView& v, uint32_t count,
                                           const uint4* __restrict__ top_list,
                                           const uint32_t* __restrict__ node2slot,
                                           uint32_t* __restrict__ err, unsigned char* smem,
                                           const PK& pk = PK{}) {
  constexpr int R = (C + T - 1) / T;  // slots per thread
  if (count > (uint32_t)C) return false;
  Box* s_box = reinterpret_cast<Box*>(smem);
  uint8_t* s_done = smem + (size_t)C * sizeof(Box);  // round that finished the slot
  const uint32_t nb = v.num_leaf - 1;
  uint4 rec[R];
  uint2 kid[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    rec[r] = e < count ? top_list[e] : make_uint4(DMSH_U32MAX, 0u, 0u, 0u);
    kid[r] = make_uint2(rec[r].y, rec[r].z);
  }
  bool bad = false;
  uint32_t pending = 0u;  // bit r: slot r T + tid still to merge
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e >= count) continue;
    const bool ta = rec[r].w & 1u, tb = rec[r].w & 2u;
    const uint32_t a = kid[r].x, b = kid[r].y;
    bool ok = rec[r].x < nb && (ta ? a < nb : a <= 2 * nb) && (tb ? b < nb : b <= 2 * nb);
    if (ok && ta) {
      kid[r].x = node2slot[a];
      ok = kid[r].x < count;
      kid[r].x |= kTopRef;
    }
    if (ok && tb) {
      kid[r].y = node2slot[b];
      ok = kid[r].y < count;
      kid[r].y |= kTopRef;
    }
    if (!ok) {
      bad = true;
      continue;
    }
    if (!ta && !tb) {
      s_box[e] = box_merge(box_load(v.aabb, a), box_load(v.aabb, b));
      s_done[e] = 0u;
    } else {
      if (!ta || !tb) s_box[e] = box_load(v.aabb, ta ? b : a);  // the child below the top
      s_done[e] = kTopPending;
      pending |= 1u << r;
    }
  }
  if (__syncthreads_or(bad)) {
    if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
    __syncthreads();  // LDS is reused by the caller's next tree
    return true;
  }
  for (uint32_t t = 1; __syncthreads_or(pending != 0u); ++t) {
    if (t > count || t >= kTopPending) return false;  // uniform; nothing stored yet
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!(pending & (1u << r))) continue;
      const uint32_t e = (uint32_t)r * T + threadIdx.x;
      const uint2 k = kid[r];
      const uint32_t ia = (k.x & kTopRef) ? k.x & ~kTopRef : e;
      const uint32_t ib = (k.y & kTopRef) ? k.y & ~kTopRef : e;
      const bool ra = ia == e || s_done[ia] < t;
      const bool rb = ib == e || s_done[ib] < t;
      if (!(ra && rb)) continue;
      s_box[e] = box_merge(s_box[ia], s_box[ib]);
      s_done[e] = (uint8_t)t;
      pending &= ~(1u << r);
    }
  }
  // the loop's last barrier orders these after every merge
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e >= count) continue;
    box_store(v.aabb, rec[r].x, s_box[e]);
    if constexpr (PK::kOn) {
      const uint32_t a = rec[r].y, b = rec[r].z;
      const Box ba = (kid[r].x & kTopRef) ? s_box[kid[r].x & ~kTopRef] : box_load(v.aabb, a);
      const Box bb = (kid[r].y & kTopRef) ? s_box[kid[r].y & ~kTopRef] : box_load(v.aabb, b);
      pk.inner_at(rec[r].x, a, b, ba.v, bb.v, v.num_leaf, err);
    }
  }
  __syncthreads();  // LDS is reused by the caller's next tree
  return true;
}


// Top nodes merged by climbs through LDS arrival counters (up to C slots):
// slot e keeps its box, its parent's slot (16 bits) and its top children's
// slots (16 bits each, 0xFFFF for a child below the top, whose box the
// slot's own entry holds until the merge), and one byte of a packed counter
// word.  The treelet pass staged each slot's top-child slots and the boxes
// of its children below the top in slot order (TopStage), so the load pass
// is three coalesced reads per slot.  A slot whose children are both below
// the top is finished in the load pass and starts a climb; at each parent
// the arrival that completes its top children merges (left, right) and
// climbs on, one LDS round trip per level.  The climbs run in stages of
// `steps` levels from a queue: the climbs still going at the end of a stage
// are queued again, so every stage runs on as few waves as it has climbs.
template <int C, int T>
__device__ __forceinline__ void top_climb(float* __restrict__ aabb, uint32_t num_leaf,
                                          uint32_t count, const uint4* __restrict__ top_list,
                                          const TopStage& st, unsigned char* __restrict__ smem,
                                          int steps, uint32_t* __restrict__ err) {
  constexpr int R = (C + T - 1) / T;
  constexpr uint32_t kNone = 0xFFFFu;
  Box* s_box = reinterpret_cast<Box*>(smem);
  uint16_t* s_par = reinterpret_cast<uint16_t*>(smem + (size_t)C * sizeof(Box));
  uint32_t* s_kid = reinterpret_cast<uint32_t*>(smem + (size_t)C * (sizeof(Box) + 2));
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(smem + (size_t)C * (sizeof(Box) + 6));
  uint16_t* s_q = reinterpret_cast<uint16_t*>(smem + (size_t)C * (sizeof(Box) + 7));
  uint32_t* s_qn = reinterpret_cast<uint32_t*>(smem + (size_t)C * (sizeof(Box) + 11));
  const uint32_t nb = num_leaf - 1;
  const uint32_t S = steps > 0 ? (uint32_t)steps : (uint32_t)C;
  for (uint32_t w = threadIdx.x; w < (count + 3) / 4; w += T) s_cnt[w] = 0u;
  for (uint32_t e = threadIdx.x; e < count; e += T) s_par[e] = (uint16_t)kNone;
  if (threadIdx.x < 3) s_qn[threadIdx.x] = 0u;
  // every slot's loads are issued before any of them is used (one memory
  // latency for the pass, not one per slot of this thread)
  // (a slot's top children are the ones kid names; its node id is needed
  // only for the final store)
  uint32_t node[R], kid[R];
  Box ba[R], bb[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    const uint32_t ec = e < count ? e : 0u;  // in bounds either way; unused past count
    node[r] = top_list[ec].x;
    kid[r] = st.kid[ec];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      ba[r].v[d] = st.box[(uint32_t)d * C + ec];
      bb[r].v[d] = st.box[(uint32_t)(6 + d) * C + ec];
    }
  }
  Box own[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e >= count) {
      node[r] = DMSH_U32MAX;
      continue;
    }
    const bool ta = (kid[r] & 0xFFFFu) != kNone, tb = (kid[r] >> 16) != kNone;
    own[r] = (!ta && !tb) ? box_merge(ba[r], bb[r]) : (ta ? bb[r] : ba[r]);
  }
  __syncthreads();  // s_par, s_cnt and the queue counters cleared
  bool bad = false;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e >= count) continue;
    const uint32_t sa = kid[r] & kNone, sb = kid[r] >> 16;
    const bool ta = sa != kNone, tb = sb != kNone;
    if (node[r] >= nb || (ta && sa >= count) || (tb && sb >= count)) {
      bad = true;
      continue;
    }
    if (!ta || !tb) s_box[e] = own[r];
    s_kid[e] = kid[r];
    if (ta) s_par[sa] = (uint16_t)e;
    if (tb) s_par[sb] = (uint16_t)e;
    if (!ta && !tb) s_q[atomicAdd(s_qn, 1u)] = (uint16_t)e;
  }
  if (__syncthreads_or(bad)) {
    if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
    return;
  }
  // stage k reads queue k % 2 (length s_qn[k % 3]), queues the climbs still
  // going into queue (k + 1) % 2 and clears the counter stage k + 1 appends to
  for (uint32_t k = 0;; ++k) {
    const uint32_t nq = s_qn[k % 3];
    if (nq == 0u) break;
    if (k > count) {  // more stages than slots: a parent cycle
      if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
      return;
    }
    if (threadIdx.x == 0) s_qn[(k + 2) % 3] = 0u;
    const uint16_t* q = s_q + (size_t)(k & 1u) * C;
    uint16_t* qn = s_q + (size_t)((k + 1) & 1u) * C;
    for (uint32_t i = threadIdx.x; i < nq; i += T) {
      // a climb from slot c (finished) to its parent p; the parent's links
      // are read one level ahead, so the counter, both children's boxes and
      // the next links go out together: one LDS round trip per level.  The
      // counter RMW is ordered at wavefront scope (no wait after it): a
      // wave's LDS operations are performed in issue order, so the box loads
      // issued after it see every box stored before the other child's RMW,
      // which this one observed (as in refit_treelet's climb)
      uint32_t c = q[i];
      uint32_t p = s_par[c];
      if (p == kNone) continue;
      uint32_t kp = s_kid[p], pn = s_par[p];
      bool go = true;
      for (uint32_t j = 0; j < S; ++j) {
        const uint32_t la = kp & kNone, lb = kp >> 16;
        const uint32_t need = (la != kNone ? 1u : 0u) + (lb != kNone ? 1u : 0u);
        const uint32_t sh = 8u * (p & 3u);
        const uint32_t old = (__hip_atomic_fetch_add(s_cnt + (p >> 2), 1u << sh, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_WAVEFRONT) >>
                              sh) &
                             0xFFu;
        const Box A = lds_box(s_box, la != kNone ? la : p);
        const Box B = lds_box(s_box, lb != kNone ? lb : p);
        uint32_t kp2 = 0u, pn2 = kNone;
        if (pn != kNone) {
          kp2 = s_kid[pn];
          pn2 = s_par[pn];
        }
        if (old + 1u < need) {  // the other top child finishes this node
          go = false;
          break;
        }
        lds_box(s_box, p) = box_merge(A, B);
        c = p;
        p = pn;
        kp = kp2;
        pn = pn2;
        if (p == kNone) {  // the top root is done
          go = false;
          break;
        }
      }
      if (go) qn[atomicAdd(s_qn + (k + 1) % 3, 1u)] = (uint16_t)c;
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e < count) box_store(aabb, node[r], s_box[e]);
  }
}

template <int C, int T>
__global__ __launch_bounds__(T) void k_refit_top_lds(
    float* __restrict__ aabb, uint32_t* __restrict__ flags, uint32_t num_leaf,
    const uint32_t* __restrict__ bvhnodes, const uint32_t* __restrict__ top_count,
    const uint4* __restrict__ top_list, const uint32_t* __restrict__ node2slot,
    const uint32_t* __restrict__ treelet_count, const uint4* __restrict__ treelets,
    const uint32_t* __restrict__ skip, uint32_t* __restrict__ err, int rounds, TopStage st,
    int steps) {
  constexpr int R = (C + T - 1) / T;  // slots per thread
  static_assert((size_t)kTopClimbSlots * (sizeof(Box) + 11) + 12 <= (size_t)C * (sizeof(Box) + 1),
                "the climb layout fits in the round layout's LDS");
  __shared__ __align__(16) unsigned char smem[(size_t)C * (sizeof(Box) + 1)];
  if (skip && *skip) return;
  const TreeView v{aabb, bvhnodes, num_leaf, nullptr, 0u, nullptr};
  const uint32_t count = *top_count;
  if (!rounds && st.kid && count <= (uint32_t)kTopClimbSlots) {
    top_climb<kTopClimbSlots, T>(aabb, num_leaf, count, top_list, st, smem, steps, err);
    return;
  }
  if (count > (uint32_t)C) {
    uint32_t nt = *treelet_count;
    if (nt > num_leaf) nt = num_leaf;
    refit_top<T>(v, flags, reinterpret_cast<const uint32_t*>(treelets), 4u, nt, err);
    return;
  }
  Box* s_box = reinterpret_cast<Box*>(smem);
  uint8_t* s_done = smem + (size_t)C * sizeof(Box);  // round that finished the slot, or kTopPending
  const uint32_t nb = num_leaf - 1;
  // phase 1, stage by stage over this thread's slots e = r T + tid
  uint4 rec[R];
  uint2 kid[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    rec[r] = e < count ? top_list[e] : make_uint4(DMSH_U32MAX, 0u, 0u, 0u);
    kid[r] = make_uint2(rec[r].y, rec[r].z);
  }
  bool bad = false;
  uint32_t pending = 0u;  // bit r: slot r T + tid still to merge
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e >= count) continue;
    const bool ta = rec[r].w & 1u, tb = rec[r].w & 2u;
    const uint32_t a = kid[r].x, b = kid[r].y;
    bool ok = rec[r].x < nb && (ta ? a < nb : a <= 2 * nb) && (tb ? b < nb : b <= 2 * nb);
    if (ok && ta) {
      kid[r].x = node2slot[a];
      ok = kid[r].x < count;
      kid[r].x |= kTopRef;
    }
    if (ok && tb) {
      kid[r].y = node2slot[b];
      ok = kid[r].y < count;
      kid[r].y |= kTopRef;
    }
    if (!ok) {
      bad = true;
      continue;
    }
    if (!ta && !tb) {
      s_box[e] = box_merge(box_load(aabb, a), box_load(aabb, b));
      s_done[e] = 0u;
    } else {
      if (!ta || !tb) s_box[e] = box_load(aabb, ta ? b : a);  // the non-top child
      s_done[e] = kTopPending;
      pending |= 1u << r;
    }
  }
  if (__syncthreads_or(bad)) {
    if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
    return;
  }
  // phase 2
  for (uint32_t t = 1; __syncthreads_or(pending != 0u); ++t) {
    if (t > count || t >= kTopPending) {
      // a top tree higher than the done[] rounds can count (or a corrupt
      // one): redo the nodes above the treelets with the global climb, whose
      // arrival counters (the top nodes', all listed here) start at zero
      for (uint32_t e = threadIdx.x; e < count; e += T) {
        const uint32_t node = top_list[e].x;
        if (node < nb) flags[node] = 0u;
      }
      __syncthreads();
      uint32_t nt = *treelet_count;
      if (nt > num_leaf) nt = num_leaf;
      refit_top<T>(v, flags, reinterpret_cast<const uint32_t*>(treelets), 4u, nt, err);
      return;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!(pending & (1u << r))) continue;
      const uint32_t e = (uint32_t)r * T + threadIdx.x;
      const uint2 k = kid[r];
      const uint32_t ia = (k.x & kTopRef) ? k.x & ~kTopRef : e;
      const uint32_t ib = (k.y & kTopRef) ? k.y & ~kTopRef : e;
      const bool ra = ia == e || s_done[ia] < t;
      const bool rb = ib == e || s_done[ib] < t;
      if (!(ra && rb)) continue;
      s_box[e] = box_merge(s_box[ia], s_box[ib]);
      s_done[e] = (uint8_t)t;
      pending &= ~(1u << r);
    }
  }
  // phase 3 (the loop's last barrier orders it after every merge)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t e = (uint32_t)r * T + threadIdx.x;
    if (e < count) box_store(aabb, rec[r].x, s_box[e]);
  }
}

constexpr uint32_t kTreeletLeaves = 1024;  // K
constexpr int kTreeletThreads = 1024;
constexpr uint32_t kTreeletGrid = 1024;    // persistent workgroups, at most
// Two treelet workgroups fit a CU's 160 KB of LDS: the persistent grid is
// what is resident at once (a workgroup past it would start only when a
// resident one has run out of tickets, with its dealt first treelet last).
constexpr uint32_t kTreeletsPerCU = 2;
static_assert(kTreeletsPerCU * (sizeof(TreeletLds<kTreeletLeaves>) + 64) <= 160u * 1024u &&
                  (kTreeletsPerCU + 1) * sizeof(TreeletLds<kTreeletLeaves>) > 160u * 1024u,
              "treelet workgroups per CU");
constexpr int kTopThreads = 1024;
constexpr int kTopSlots = 6400;  // 25 B each: 156 KB of LDS

// flags: num_leaf - 1 zeroed counters; treelets (and, when top_count is not
// null, the top nodes with top_cap = kTopSlots) as listed by the topology
// kernel (or k_list_treelets) with k = kTreeletLeaves.  A nonzero *skip makes
// both kernels do nothing.
template <int LK>
inline int launch_refit_treelets(hipStream_t s, const TreeView& v, uint32_t* flags,
                                 const uint32_t* treelet_count, const uint4* treelets, float eps,
                                 uint32_t* err, const uint32_t* top_count = nullptr,
                                 const uint4* top_list = nullptr,
                                 const uint32_t* node2slot = nullptr,
                                 uint32_t* treelet_next = nullptr,
                                 const uint32_t* skip = nullptr, bool leaves_in_aabb = false,
                                 uint32_t* top_kid = nullptr, float* top_box = nullptr) {
  const uint32_t num_leaf = v.num_leaf;
  if (num_leaf == 0) return DMSH_OK;
  TopStage st;
  if (top_count && top_kid && top_box && num_leaf > kTreeletLeaves)
    st = TopStage{top_count, top_list, node2slot, top_kid, top_box};
  const uint32_t resident = kTreeletsPerCU * (uint32_t)device_cu_count();
  const uint32_t grid = std::min<uint32_t>(std::min<uint32_t>(kTreeletGrid, resident), num_leaf);
  constexpr int K = (int)kTreeletLeaves, Th = kTreeletThreads;
  const int climb = refit_option(kRefitClimb);
  const int compact = refit_option(kRefitCompact);
  if (refit_option(kRefitQueue) == 0) treelet_next = nullptr;
  if (leaves_in_aabb && climb >= 1) {
    launch("k_refit_treelet", (k_refit_treelet<LK, K, Th, true, 1>), dim3(grid), dim3(Th), 0, s,
           v, eps, treelet_count, treelets, treelet_next, skip, err, compact, st);
  } else if (leaves_in_aabb) {
    launch("k_refit_treelet", (k_refit_treelet<LK, K, Th, true, 0>), dim3(grid), dim3(Th), 0, s,
           v, eps, treelet_count, treelets, treelet_next, skip, err, compact, st);
  } else if (climb == 2) {
    launch("k_refit_treelet", (k_refit_treelet<LK, K, Th, false, 2>), dim3(grid), dim3(Th), 0, s,
           v, eps, treelet_count, treelets, treelet_next, skip, err, compact, st);
  } else if (climb == 1) {
    launch("k_refit_treelet", (k_refit_treelet<LK, K, Th, false, 1>), dim3(grid), dim3(Th), 0, s,
           v, eps, treelet_count, treelets, treelet_next, skip, err, compact, st);
  } else {
    launch("k_refit_treelet", (k_refit_treelet<LK, K, Th, false, 0>), dim3(grid), dim3(Th), 0, s,
           v, eps, treelet_count, treelets, treelet_next, skip, err, compact, st);
  }
  int rc = launch_status();
  if (rc != DMSH_OK || num_leaf <= kTreeletLeaves) return rc;
  if (top_count) {
    launch("k_refit_top", k_refit_top_lds<kTopSlots, kTopThreads>, dim3(1), dim3(kTopThreads), 0,
           s, v.aabb, flags, num_leaf, v.bvhnodes, top_count, top_list, node2slot, treelet_count,
           treelets, skip, err, refit_option(kRefitTop), st, compact);
  } else {
    launch("k_refit_top", k_refit_top<kTopThreads>, dim3(1), dim3(kTopThreads), 0, s, v.aabb,
           flags, num_leaf, v.bvhnodes, treelet_count, treelets, err);
  }
  return launch_status();
}

// ---------------------------------------------------------------------------
// Standalone refit of a tree given only its rows (bvhnode2aabb3.rs:68-142,
// bvhnode2aabb.cu:6-122): the treelet kernels need each internal node's leaf
// range [first, last], which the topology kernel knows during a build but a
// stored tree does not carry.
//
// k_leaf_ranges: the first leaf of a node is the first leaf of its left
//   child, so leaf k is the first leaf of every node on the run of left-child
//   links directly above it (and the last leaf of the run of right-child
//   links).  Every internal node lies on exactly one run of each kind, so one
//   walk per leaf writes each first[] and last[] entry once, O(n) in all.
// k_list_treelets: per internal node i with children a, b, checks the layout
//   the treelet kernels rely on -- first[i] = first(a), last[i] = last(b),
//   last(a) + 1 = first(b) (so a subtree's leaves are exactly [first, last])
//   and first[i] <= i <= last[i] (so its internal ids lie in that range, as in
//   every Karras tree) -- and lists the treelets and top nodes as the
//   topology kernel does.  One failed check sets *bad: the treelet kernels
//   then skip and the generic climb (gated on *bad) refits the tree instead.
constexpr uint32_t kPlanHashBlocks = 1024;  // row-hash partials (see k_plan_hash)

struct StandaloneLayout {
  uint32_t* flags;     // num_leaf - 1 arrival counters
  uint32_t* first;     // num_leaf - 1
  uint32_t* last;      // num_leaf - 1
  uint32_t* node2slot; // num_leaf - 1
  uint32_t* counters;  // treelet count, top count, next treelet, bad, plan needed
  uint4* treelets;     // num_leaf
  uint4* top_list;     // kTopSlots: (node, left child, right child, top-child flags)
  unsigned long long* hash;  // -, stamp of the kept plan, tag, -, row-hash partials
  uint32_t* top_kid;   // kTopClimbSlots (TopStage)
  float* top_box;      // 2 kTopClimbSlots boxes (TopStage)
};

inline bool carve_standalone(Arena& a, uint32_t n, StandaloneLayout& L) {
  const size_t nb = n > 1 ? n - 1 : 1;
  L.flags = a.take<uint32_t>(nb);
  L.first = a.take<uint32_t>(nb);
  L.last = a.take<uint32_t>(nb);
  L.node2slot = a.take<uint32_t>(nb);
  L.counters = a.take<uint32_t>(8);
  L.treelets = a.take<uint4>(n ? n : 1);
  L.top_list = a.take<uint4>(kTopSlots);
  L.hash = a.take<unsigned long long>(4 + kPlanHashBlocks);
  L.top_kid = a.take<uint32_t>(kTopClimbSlots);
  L.top_box = a.take<float>((size_t)kTopClimbSlots * 12);
  return a.ok();
}

// The plan -- leaf ranges, treelet and top lists -- depends on the rows
// only, so a workspace that is used again for the same rows keeps it:
// k_plan_hash sums a mixed hash of every (word index, word) of bvhnodes,
// k_plan_gate compares the sum with the stamp kept beside the plan and sets
// the gate of the two plan kernels when they differ (or when the workspace
// has never held a plan, the tree failed the layout check, or its top list
// overflowed).  A workspace fresh from an allocator holds no valid tag, so
// its first call builds the plan.
constexpr unsigned long long kPlanTag = 0x646d73685f706c6eull;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Per word w with value x: (x * A + w * B) with a xor-shift, summed mod 2^64
// (one 32x64 multiply-add per word; the index term keeps a permutation of
// the rows from summing to the same value).  Four words per load where the
// rows are 16-byte aligned.
__device__ __forceinline__ unsigned long long word_hash(unsigned long long w, uint32_t x) {
  unsigned long long v = (unsigned long long)x * 0x9e3779b97f4a7c15ull + w * 0xc2b2ae3d27d4eb4full;
  return v ^ (v >> 29);
}

static __global__ __launch_bounds__(256) void k_plan_hash(const uint32_t* __restrict__ words, size_t num_words,
                                                          unsigned long long* __restrict__ part) {
  __shared__ unsigned long long s_part[4];
  unsigned long long h = 0;
  const size_t stride = (size_t)gridDim.x * 256;
  if ((reinterpret_cast<uintptr_t>(words) & 15u) == 0) {
    const uint4* q = reinterpret_cast<const uint4*>(words);
    const size_t nq = num_words / 4;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nq; i += stride) {
      const uint4 x = q[i];
      const unsigned long long w = 4 * (unsigned long long)i;
      h += word_hash(w, x.x) + word_hash(w + 1, x.y) + word_hash(w + 2, x.z) +
           word_hash(w + 3, x.w);
    }
    for (size_t w = nq * 4 + (size_t)blockIdx.x * 256 + threadIdx.x; w < num_words; w += stride)
      h += word_hash(w, words[w]);
  } else {
    for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < num_words; w += stride)
      h += word_hash(w, words[w]);
  }
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63u) == 0) s_part[threadIdx.x >> 6] = h;
  __syncthreads();
  // one partial per block (the gate adds them): thousands of blocks adding
  // into one word would queue one device-scope atomic behind the other
  if (threadIdx.x == 0) part[blockIdx.x] = s_part[0] + s_part[1] + s_part[2] + s_part[3];
}

static __global__ __launch_bounds__(256) void k_plan_gate(unsigned long long* __restrict__ hash,
                                                          uint32_t* __restrict__ counters,
                                                          uint32_t num_leaf, uint32_t num_part,
                                                          int cache) {
  __shared__ unsigned long long s_part[4];
  unsigned long long h = 0;
  for (uint32_t b = threadIdx.x; b < num_part; b += 256) h += hash[4 + b];
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o);
  if ((threadIdx.x & 63u) == 0) s_part[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x != 0) return;
  h = (s_part[0] + s_part[1] + s_part[2] + s_part[3]) ^ mix64(num_leaf);
  const bool keep = cache && hash[2] == kPlanTag && hash[1] == h && counters[3] == 0u &&
                    counters[1] <= (uint32_t)kTopSlots && counters[0] <= num_leaf;
  counters[4] = keep ? 0u : 1u;
  counters[2] = 0u;  // the treelet kernel's ticket counter
  hash[1] = h;
  hash[2] = kPlanTag;
}

inline size_t standalone_workspace_bytes(uint32_t n) {
  Arena a{nullptr, ~(size_t)0, 0};
  StandaloneLayout L;
  carve_standalone(a, n, L);
  return a.off + 256;
}

// Also zeroes the arrival counters and the four list counters (the next
// kernel is the first to use them).  A thread walks kRangeWalks leaves in
// lock-step, so their dependent row loads overlap.  A walk stops at a link
// out of range; the entries it leaves unwritten fail k_list_treelets' checks
// or the treelet kernels' own bounds checks.
constexpr int kRangeWalks = 1;
constexpr int kListThreads = 1024;
constexpr uint32_t kListGrid = 65535;  // one round of kListThreads nodes per workgroup

// With `boxes`, also stores each leaf's box (leaf_box, as the treelet
// kernels would make it), so the treelet pass reads boxes instead of
// gathering geometry.
template <int LK>
__global__ __launch_bounds__(256) void k_leaf_ranges(TreeView v, float eps, int boxes,
                                                     uint32_t* __restrict__ first,
                                                     uint32_t* __restrict__ last,
                                                     uint32_t* __restrict__ flags,
                                                     uint32_t* __restrict__ counters,
                                                     uint32_t* __restrict__ err) {
  if (counters[4] == 0u) return;  // the kept plan is valid
  const uint32_t num_leaf = v.num_leaf;
  const uint32_t* __restrict__ bvhnodes = v.bvhnodes;
  const uint32_t nb = num_leaf - 1;
  const uint32_t k0 = blockIdx.x * (256u * kRangeWalks) + threadIdx.x;
  if (k0 == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) counters[c] = 0u;  // not counters[4]: read by every block
  }
#pragma unroll
  for (int r = 0; r < kRangeWalks; ++r) {
    const uint32_t k = k0 + 256u * r;
    if (!boxes || k >= num_leaf) continue;
    Box b;
    uint32_t par;
    if (leaf_box<LK>(v, k, eps, b, par, err)) box_store(v.aabb, nb + k, b);
  }
  uint32_t node[kRangeWalks], p[kRangeWalks];
  int side[kRangeWalks];
#pragma unroll
  for (int r = 0; r < kRangeWalks; ++r) {
    const uint32_t k = k0 + 256u * r;
    node[r] = nb + k;
    p[r] = DMSH_U32MAX;
    if (k < nb) flags[k] = 0u;
    if (k < num_leaf) p[r] = bvhnodes[(size_t)node[r] * 3];
  }
  // the side of its parent a leaf is on fixes which run it starts
#pragma unroll
  for (int r = 0; r < kRangeWalks; ++r) {
    side[r] = 0;
    if (p[r] < nb) side[r] = bvhnodes[(size_t)p[r] * 3 + 1] == node[r] ? 1 : 2;
  }
  for (uint32_t steps = 0; steps < nb; ++steps) {
    uint32_t row[kRangeWalks][2];
#pragma unroll
    for (int r = 0; r < kRangeWalks; ++r) {
      if (!side[r]) continue;
      const uint32_t* prow = bvhnodes + (size_t)p[r] * 3;
      row[r][0] = prow[0];
      row[r][1] = prow[side[r]];
    }
    bool more = false;
#pragma unroll
    for (int r = 0; r < kRangeWalks; ++r) {
      if (!side[r]) continue;
      if (row[r][1] != node[r]) {
        side[r] = 0;
        continue;
      }
      (side[r] == 1 ? first : last)[p[r]] = k0 + 256u * r;
      node[r] = p[r];
      p[r] = row[r][0];
      if (p[r] >= nb) side[r] = 0;
      more |= side[r] != 0;
    }
    if (!more) break;
  }
}

// One workgroup of T threads takes T internal nodes per round (grid-stride);
// its treelets and top nodes get their list slots from one 64-bit atomic on
// (top count, treelet count) per round instead of one atomic per entry.
template <int T>
__global__ __launch_bounds__(T) void k_list_treelets(
    uint32_t num_leaf, const uint32_t* __restrict__ bvhnodes, const uint32_t* __restrict__ first,
    const uint32_t* __restrict__ last, uint32_t* __restrict__ counters,
    uint4* __restrict__ treelets, uint4* __restrict__ top_list, uint32_t* __restrict__ node2slot) {
  __shared__ uint32_t s_cnt[2];  // treelets, tops of this round
  __shared__ unsigned long long s_base;
  if (counters[4] == 0u) return;  // the kept plan is valid
  const uint32_t nb = num_leaf - 1;
  const uint32_t K = kTreeletLeaves;
  for (uint32_t base = blockIdx.x * T; base < nb; base += gridDim.x * T) {
    const uint32_t i = base + threadIdx.x;
    uint32_t n_tl = 0, n_top = 0;
    uint4 tl[2];
    uint4 top = make_uint4(0u, 0u, 0u, 0u);
    if (i < nb) {
      const uint32_t* row = bvhnodes + (size_t)i * 3;
      const uint32_t p = row[0], a = row[1], b = row[2];
      const uint32_t f = first[i], l = last[i];
      bool ok = a <= 2 * nb && b <= 2 * nb && (i == 0) == (p == DMSH_U32MAX);
      uint32_t fa = 0, la = 0, fb = 0, lb = 0;
      if (ok) {
        fa = a >= nb ? a - nb : first[a];
        la = a >= nb ? a - nb : last[a];
        fb = b >= nb ? b - nb : first[b];
        lb = b >= nb ? b - nb : last[b];
        ok = f == fa && l == lb && fa <= la && la < nb && la + 1u == fb && fb <= lb && f <= i &&
             i <= l && l <= nb && (i != 0 || (f == 0 && l == nb));
      }
      if (!ok) {
        atomicOr(counters + 3, 1u);
      } else if (l - f + 1 > K) {
        n_top = 1;
        top = make_uint4(i, a, b, (la - fa + 1 > K ? 1u : 0u) | (lb - fb + 1 > K ? 2u : 0u));
        if (la - fa + 1 <= K) tl[n_tl++] = make_uint4(a, fa, la, 0u);  // w: side
        if (lb - fb + 1 <= K) tl[n_tl++] = make_uint4(b, fb, lb, 1u);
      } else if (i == 0) {
        tl[n_tl++] = make_uint4(0u, 0u, nb, DMSH_U32MAX);
      }
    }
    if (threadIdx.x == 0) s_cnt[0] = s_cnt[1] = 0u;
    __syncthreads();
    const uint32_t o_tl = n_tl ? atomicAdd(&s_cnt[0], n_tl) : 0u;
    const uint32_t o_top = n_top ? atomicAdd(&s_cnt[1], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0 && (s_cnt[0] | s_cnt[1]))
      s_base = atomicAdd(reinterpret_cast<unsigned long long*>(counters),
                         ((unsigned long long)s_cnt[1] << 32) | s_cnt[0]);
    __syncthreads();
    const uint32_t b_tl = (uint32_t)s_base, b_top = (uint32_t)(s_base >> 32);
    const uint32_t my_top = b_top + o_top;
    for (uint32_t e = 0; e < n_tl; ++e) {
      uint4 t = tl[e];
      t.w = (t.w != DMSH_U32MAX && n_top && my_top < (uint32_t)kTopSlots) ? 2u * my_top + t.w
                                                                           : DMSH_U32MAX;
      if (b_tl + o_tl + e < num_leaf) treelets[b_tl + o_tl + e] = t;
    }
    if (n_top && b_top + o_top < (uint32_t)kTopSlots) {
      top_list[b_top + o_top] = top;
      node2slot[i] = b_top + o_top;
    }
    __syncthreads();  // s_cnt and s_base are reset next round
  }
}

// ws: standalone_workspace_bytes(v.num_leaf).  Launches: row hash and plan
// gate, ranges and list (gated: idle when the workspace keeps this tree's
// plan), treelets, top (more than kTreeletLeaves leaves), gated generic climb.
template <int LK>
inline int launch_refit_standalone(hipStream_t s, const TreeView& v, float eps, void* ws,
                                   size_t ws_bytes, uint32_t* err) {
  const uint32_t n = v.num_leaf;
  if (n == 0) return DMSH_OK;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  StandaloneLayout L;
  if (!carve_standalone(a, n, L)) return DMSH_ERR_WORKSPACE;
  if (n == 1) {  // one leaf, no internal node: the generic kernel stores its box
    launch("k_refit", k_refit<LK>, dim3(1), dim3(256), 0, s, v, L.flags, eps,
           (const uint32_t*)nullptr, err);
    return launch_status();
  }
  const bool leaves_in_aabb = refit_option(kRefitLeavesInAabb) == 1;
  // leaf boxes written by k_leaf_ranges change with the geometry: no kept plan then
  const int cache = (refit_option(kRefitPlanCache) == 1 && !leaves_in_aabb) ? 1 : 0;
  const size_t num_words = (size_t)(2 * (size_t)n - 1) * 3;
  const uint32_t hash_blocks = std::min<uint32_t>(ceil_div(num_words, 1024), kPlanHashBlocks);
  launch("k_plan_hash", k_plan_hash, dim3(hash_blocks), dim3(256), 0, s, v.bvhnodes, num_words,
         L.hash + 4);
  launch("k_plan_gate", k_plan_gate, dim3(1), dim3(256), 0, s, L.hash, L.counters, n, hash_blocks,
         cache);
  launch("k_leaf_ranges", k_leaf_ranges<LK>, dim3(ceil_div(n, 256 * kRangeWalks)), dim3(256), 0, s,
         v, eps, leaves_in_aabb ? 1 : 0, L.first, L.last, L.flags, L.counters, err);
  int rc = launch_status();
  if (rc != DMSH_OK) return rc;
  launch("k_list_treelets", k_list_treelets<kListThreads>,
         dim3(std::min<uint32_t>(ceil_div(n - 1, kListThreads), kListGrid)), dim3(kListThreads), 0,
         s, n, v.bvhnodes, L.first, L.last, L.counters, L.treelets, L.top_list, L.node2slot);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  rc = launch_refit_treelets<LK>(s, v, L.flags, L.counters + 0, L.treelets, eps, err,
                                 L.counters + 1, L.top_list, L.node2slot, L.counters + 2,
                                 L.counters + 3, leaves_in_aabb, L.top_kid, L.top_box);
  if (rc != DMSH_OK) return rc;
  launch("k_refit", k_refit<LK>, dim3(std::min<uint32_t>(ceil_div(n, 256), 1024)), dim3(256), 0,
         s, v, L.flags, eps, (const uint32_t*)(L.counters + 3), err);
  return launch_status();
}

}  // namespace dmsh
