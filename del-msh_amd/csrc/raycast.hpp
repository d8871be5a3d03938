// raycast.hpp -- per-pixel BVH ray-cast and depth on gfx950.
//
// Semantics are those of del-msh-cpu's recursive first_intersection_ray
// (search_bvh3.rs:103-221) with dis = +inf, not the unordered DFS of the
// reference CUDA kernel:
//   F(node): AABB miss -> none; leaf -> triangle hit with t < inf;
//            both children missed -> none; near = left iff (tl_min, tl_max) <
//            (tr_min, tr_max) lexicographically (else right; a single hit
//            child is near); rn = F(near); if rn and the far entry t > rn.t
//            (or far missed) -> rn; else rf = F(far) and the result is
//            rn if rn.t < rf.t else rf.
// Executed as a flat closest-hit search with the same result: one best hit,
// a stack of 8-byte frames (far child, far entry t) pushed where both children
// are hit (a single hit child is a tail call), a popped far child skipped when
// its entry t exceeds the best t so far, and a triangle hit replacing the best
// when its t <= the best t.  Children are visited near first, so of two hits
// the later one lies in the far subtree of their lowest common ancestor,
// where F's combine keeps far on equal t; a skipped subtree cannot hold a hit
// with t <= the final one.  (The oracle checks this search against the
// recursion on scenes of exact ties: duplicated triangles, rays through grid
// vertices, coplanar soups.)
//
// Layout: before traversal the BVH is packed into
//   inner[i] (64 B, internal node i): both child boxes + both child refs
//   leaf[k]  (48 B, leaf k):          the triangle's three vertices + its id
// so one node expansion is one 64-byte load and one leaf test one 48-byte load
// (the reference layout needs a row load, then dependent box / index / vertex
// gathers).  The values are copies, so the arithmetic is unchanged.
//
// Stack tiers: the main kernel keeps S frames per lane in LDS
// ([depth][thread], conflict-free for 64-bit accesses); a ray that needs more
// is queued and restarted by a kernel with 64 LDS frames, and a ray that needs
// more than that by one with a 4096-frame stack in global memory.  Pixels go
// to waves in 8x8 tiles and tiles to XCDs in contiguous runs, so neighbouring
// rays share BVH nodes in one L2.
#pragma once

#include "dmsh_common.hpp"
#include "mat4.hpp"

namespace dmsh {
namespace rc {

constexpr uint32_t kLeafBit = 0x40000000u;   // child ref: leaf slot
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;
constexpr uint32_t kMaxTri = 0x3FFFFFFFu;    // child refs hold 30-bit indices
constexpr int kTier2Frames = 64;
constexpr uint32_t kDeepFrames = 4096;
constexpr uint32_t kDeepSlots = 256;
constexpr uint32_t kTier2Blocks = 512;

enum : int { kTraceOk = 0, kTraceOverflow = 1, kTraceBadIndex = 2 };

struct alignas(16) InnerNode {
  float4 a;  // c0.lo.xyz, c0.hi.x
  float4 b;  // c0.hi.yz, c1.lo.xy
  float4 c;  // c1.lo.z, c1.hi.xyz
  uint4 d;   // ref c0, ref c1, -, -
};
struct alignas(16) LeafTri {
  float4 a;  // p0.xyz, p1.x
  float4 b;  // p1.yz, p2.xy
  float4 c;  // p2.z, tri id (bits), -, -
};

struct Packed {
  const InnerNode* __restrict__ inner;
  const LeafTri* __restrict__ leaf;
  const float* __restrict__ root_box;  // bvhnode2aabb row 0
  uint32_t num_leaf;
};

// ---------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------
// A child id in the LBVH layout: ids from num_leaf - 1 on are leaves.  The
// child's own row is not read here: every row's shape is checked by the
// thread that packs it (an internal row has a second child, a leaf row
// none), which raises DMSH_DEVERR_INDEX and packs a never-hit node or
// triangle for a row that does not fit its place in the layout.
__device__ __forceinline__ uint32_t child_ref(uint32_t c, uint32_t num_leaf, bool& bad) {
  const uint32_t nb = num_leaf - 1, num_node = 2 * num_leaf - 1;
  if (c >= num_node) {
    bad = true;
    return 0;
  }
  return c >= nb ? (kLeafBit | (c - nb)) : c;
}

// The record of an internal node with children (c0, c1) whose boxes are b0,
// b1 (six floats each, read only when the children fit the layout).
template <typename Box6>
__device__ __forceinline__ void store_inner(InnerNode* __restrict__ inner, uint32_t i,
                                            uint32_t c0, uint32_t c1, const Box6& b0,
                                            const Box6& b1, uint32_t num_leaf,
                                            uint32_t* __restrict__ nonfinite,
                                            uint32_t* __restrict__ err) {
  bool bad = (c1 == DMSH_U32MAX);
  const uint32_t r0 = child_ref(c0, num_leaf, bad);
  const uint32_t r1 = child_ref(c1, num_leaf, bad);
  InnerNode nd;
  if (bad) {
    raise_error(err, DMSH_DEVERR_INDEX);
    // an empty box on both sides: every ray misses this node
    nd.a = make_float4(1.f, 1.f, 1.f, 0.f);
    nd.b = make_float4(0.f, 0.f, 1.f, 1.f);
    nd.c = make_float4(1.f, 0.f, 0.f, 0.f);
    nd.d = make_uint4(kLeafBit, kLeafBit, 0u, 0u);
  } else {
    nd.a = make_float4(b0[0], b0[1], b0[2], b0[3]);
    nd.b = make_float4(b0[4], b0[5], b1[0], b1[1]);
    nd.c = make_float4(b1[2], b1[3], b1[4], b1[5]);
    nd.d = make_uint4(r0, r1, 0u, 0u);
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      fin = fin && fabsf(b0[k]) < __builtin_inff() && fabsf(b1[k]) < __builtin_inff();
    if (!fin) atomicOr(nonfinite, 1u);
  }
  inner[i] = nd;
}

// 16-byte part `part` of the record store_inner writes for children (c0,
// c1): parts 0..2 hold the children's bounds x0..x3 (box 0 bounds 0-3;
// bounds 4, 5 of box 0 and 0, 1 of box 1; box 1 bounds 2-5), part 3 the
// refs.  A record whose children do not fit the layout raises (part 0) and
// is the one every ray misses; a part with a bound that is not finite sets
// *nonfinite.  Writing the four parts from consecutive threads makes a
// wave's store one contiguous kilobyte.
__device__ __forceinline__ uint4 inner_part_bits(uint32_t part, uint32_t c0, uint32_t c1,
                                                 float x0, float x1, float x2, float x3,
                                                 uint32_t num_leaf,
                                                 uint32_t* __restrict__ nonfinite,
                                                 uint32_t* __restrict__ err) {
  bool bad = (c1 == DMSH_U32MAX);
  const uint32_t r0 = child_ref(c0, num_leaf, bad);
  const uint32_t r1 = child_ref(c1, num_leaf, bad);
  if (bad) {
    if (part == 0u) raise_error(err, DMSH_DEVERR_INDEX);
    if (part == 3u) return make_uint4(kLeafBit, kLeafBit, 0u, 0u);
    const float4 e = part == 0u   ? make_float4(1.f, 1.f, 1.f, 0.f)
                     : part == 1u ? make_float4(0.f, 0.f, 1.f, 1.f)
                                  : make_float4(1.f, 0.f, 0.f, 0.f);
    return make_uint4(__float_as_uint(e.x), __float_as_uint(e.y), __float_as_uint(e.z),
                      __float_as_uint(e.w));
  }
  if (part == 3u) return make_uint4(r0, r1, 0u, 0u);
  const bool fin = fabsf(x0) < __builtin_inff() && fabsf(x1) < __builtin_inff() &&
                   fabsf(x2) < __builtin_inff() && fabsf(x3) < __builtin_inff();
  if (!fin) atomicOr(nonfinite, 1u);
  return make_uint4(__float_as_uint(x0), __float_as_uint(x1), __float_as_uint(x2),
                    __float_as_uint(x3));
}

// Internal node i (i + 1 < num_leaf) of one tree into inner[i].
__device__ __forceinline__ void pack_inner(InnerNode* __restrict__ inner,
                                           const uint32_t* __restrict__ bvhnodes,
                                           const float* __restrict__ aabb, uint32_t num_leaf,
                                           uint32_t i, uint32_t* __restrict__ nonfinite,
                                           uint32_t* __restrict__ err) {
  const uint32_t* row = bvhnodes + (size_t)i * 3;
  const uint32_t c0 = row[1], c1 = row[2];
  const uint32_t num_node = 2 * num_leaf - 1;
  // the box pointers are formed only for children inside the tree
  const float* b0 = aabb + (size_t)(c0 < num_node ? c0 : 0u) * 6;
  const float* b1 = aabb + (size_t)(c1 < num_node ? c1 : 0u) * 6;
  store_inner(inner, i, c0, c1, b0, b1, num_leaf, nonfinite, err);
}

// Leaf k of one tree into leaf[k].
__device__ __forceinline__ void pack_leaf(LeafTri* __restrict__ leaf,
                                          const uint32_t* __restrict__ bvhnodes,
                                          const uint32_t* __restrict__ tri2vtx,
                                          const float* __restrict__ xyz, uint32_t num_leaf,
                                          uint32_t num_tri, uint32_t num_vtx, uint32_t k,
                                          uint32_t* __restrict__ err) {
  const uint32_t* row = bvhnodes + (size_t)(num_leaf - 1 + k) * 3;
  const uint32_t tri = row[1];
  LeafTri lt;
  bool ok = row[2] == DMSH_U32MAX && tri < num_tri;
  uint32_t v[3] = {0, 0, 0};
  if (ok) {
    const uint32_t* tv = tri2vtx + (size_t)tri * 3;
    v[0] = tv[0];
    v[1] = tv[1];
    v[2] = tv[2];
    ok = v[0] < num_vtx && v[1] < num_vtx && v[2] < num_vtx;
  }
  if (!ok) {
    raise_error(err, DMSH_DEVERR_INDEX);
    // a degenerate triangle: never hit
    lt.a = make_float4(0.f, 0.f, 0.f, 0.f);
    lt.b = make_float4(0.f, 0.f, 0.f, 0.f);
    lt.c = make_float4(0.f, __uint_as_float(DMSH_U32MAX), 0.f, 0.f);
  } else {
    const float* p0 = xyz + (size_t)v[0] * 3;
    const float* p1 = xyz + (size_t)v[1] * 3;
    const float* p2 = xyz + (size_t)v[2] * 3;
    lt.a = make_float4(p0[0], p0[1], p0[2], p1[0]);
    lt.b = make_float4(p1[1], p1[2], p2[0], p2[1]);
    lt.c = make_float4(p2[2], __uint_as_float(tri), 0.f, 0.f);
  }
  leaf[k] = lt;
}

// The whole tree in one launch: threads [0, num_leaf - 1) pack the internal
// nodes, the next num_leaf the leaves.
static __global__ __launch_bounds__(256) void k_pack_tree(
    InnerNode* __restrict__ inner, LeafTri* __restrict__ leaf,
    const uint32_t* __restrict__ bvhnodes, const float* __restrict__ aabb,
    const uint32_t* __restrict__ tri2vtx, const float* __restrict__ xyz, uint32_t num_leaf,
    uint32_t num_vtx, uint32_t* __restrict__ nonfinite, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nb = num_leaf - 1;
  if (i < nb) {
    pack_inner(inner, bvhnodes, aabb, num_leaf, i, nonfinite, err);
  } else if (i - nb < num_leaf) {
    pack_leaf(leaf, bvhnodes, tri2vtx, xyz, num_leaf, num_leaf, num_vtx, i - nb, err);
  }
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
struct Hit {
  bool has;
  float t;
  uint32_t tri;
};

template <int S>
struct LdsStack {
  uint2* base;
  int stride;
  static constexpr int kCap = S;
  __device__ __forceinline__ void store(int i, uint2 f) { base[i * stride] = f; }
  __device__ __forceinline__ uint2 load(int i) const { return base[i * stride]; }
};

struct GlobalStack {
  uint2* base;
  static constexpr int kCap = (int)kDeepFrames;
  __device__ __forceinline__ void store(int i, uint2 f) { base[i] = f; }
  __device__ __forceinline__ uint2 load(int i) const { return base[i]; }
};

// Rust's tuple `<` on (tmin, tmax).
__device__ __forceinline__ bool tuple_lt(float a0, float a1, float b0, float b1) {
  return (a0 == b0) ? (a1 < b1) : (a0 < b0);
}

__device__ __forceinline__ void leaf_test(const Packed& pk, uint32_t slot, const float org[3],
                                          const float dir[3], bool& has, float& t_out,
                                          uint32_t& tri_out) {
  const LeafTri lt = pk.leaf[slot];
  const float p0[3] = {lt.a.x, lt.a.y, lt.a.z};
  const float p1[3] = {lt.a.w, lt.b.x, lt.b.y};
  const float p2[3] = {lt.b.z, lt.b.w, lt.c.x};
  float t;
  has = tri_intersect(p0, p1, p2, org, dir, true, t, nullptr) && (t < __builtin_inff());
  t_out = has ? t : 0.f;
  tri_out = has ? __float_as_uint(lt.c.y) : DMSH_U32MAX;
}

// Slab test for rays with finite origin and finite reciprocal direction on
// every axis, against a finite box.  No t below can be NaN then, so min / max
// instructions pick the same values as the compare-and-select of slab_hit up
// to the sign of a zero, which no later comparison can see.
__device__ __forceinline__ bool slab_hit_finite(const RaySlab& r, const float lo[3],
                                                const float hi[3], float& tmin_out,
                                                float& tmax_out) {
  const float a0 = (lo[0] - r.o[0]) * r.inv[0], b0 = (hi[0] - r.o[0]) * r.inv[0];
  const float a1 = (lo[1] - r.o[1]) * r.inv[1], b1 = (hi[1] - r.o[1]) * r.inv[1];
  const float a2 = (lo[2] - r.o[2]) * r.inv[2], b2 = (hi[2] - r.o[2]) * r.inv[2];
  const float tmin = fmaxf(fmaxf(fminf(a0, b0), fminf(a1, b1)), fminf(a2, b2));
  const float tmax = fminf(fminf(fmaxf(a0, b0), fmaxf(a1, b1)), fmaxf(a2, b2));
  tmin_out = tmin;
  tmax_out = tmax;
  return !(tmin > tmax) && !(tmax < 0.f);
}

constexpr uint32_t kNoTest = DMSH_U32MAX;

// The traversal loop.  An iteration first pops, if the lane has no node to
// expand, the next far child not skipped by the best hit (ending the lane's
// search when the stack runs out); then expands the lane's node, if any; then
// runs the (single) leaf-test site if the pop or the expansion produced a
// leaf, so a wave whose lanes diverge between these steps pays for the
// triangle test once.
template <bool kFinite, class Stack>
__device__ __forceinline__ int trace_loop(const Packed& pk, const RaySlab& r, const float org[3],
                                          const float dir[3], Stack& stk, Hit& out) {
  bool res_has = false;
  float res_t = 0.f;
  uint32_t res_tri = DMSH_U32MAX;
  uint32_t node = 0;
  int sp = 0;
  bool expand = true;
  for (;;) {
    uint32_t test = kNoTest;
    if (!expand) {
      uint2 f = make_uint2(0u, 0u);
      bool found = false;
      while (sp > 0) {
        f = stk.load(--sp);
        if (!(res_has && __uint_as_float(f.y) > res_t)) {
          found = true;
          break;
        }
      }
      if (!found) break;
      if (f.x & kLeafBit) {
        test = f.x & kIndexMask;
      } else {
        node = f.x;
        expand = true;
      }
    }
    if (expand) {
      const InnerNode nd = pk.inner[node];
      const float l0[3] = {nd.a.x, nd.a.y, nd.a.z}, h0[3] = {nd.a.w, nd.b.x, nd.b.y};
      const float l1[3] = {nd.b.z, nd.b.w, nd.c.x}, h1[3] = {nd.c.y, nd.c.z, nd.c.w};
      float tl0, tl1, tr0, tr1;
      bool hl, hr;
      if (kFinite) {
        hl = slab_hit_finite(r, l0, h0, tl0, tl1);
        hr = slab_hit_finite(r, l1, h1, tr0, tr1);
      } else {
        hl = slab_hit(r, l0, h0, tl0, tl1);
        hr = slab_hit(r, l1, h1, tr0, tr1);
      }
      uint32_t next = kNoTest;
      if (hl && hr) {
        const bool nl = tuple_lt(tl0, tl1, tr0, tr1);
        if (sp == Stack::kCap) return kTraceOverflow;
        stk.store(sp++, make_uint2(nl ? nd.d.y : nd.d.x, __float_as_uint(nl ? tr0 : tl0)));
        next = nl ? nd.d.x : nd.d.y;
      } else if (hl || hr) {
        next = hl ? nd.d.x : nd.d.y;
      }
      if (next == kNoTest) {
        expand = false;  // both children missed
      } else if (next & kLeafBit) {
        test = next & kIndexMask;
        expand = false;
      } else {
        node = next;
      }
    }
    if (test != kNoTest) {
      bool has;
      float t;
      uint32_t tri;
      leaf_test(pk, test, org, dir, has, t, tri);
      if (has && (!res_has || t <= res_t)) {
        res_has = true;
        res_t = t;
        res_tri = tri;
      }
    }
  }
  out.has = res_has;
  out.t = res_t;
  out.tri = res_tri;
  return kTraceOk;
}

// boxes_finite: every packed child box is finite (checked while packing).
template <class Stack>
__device__ __forceinline__ int trace(const Packed& pk, const float org[3], const float dir[3],
                                     Stack& stk, Hit& out, bool boxes_finite) {
  out.has = false;
  out.t = 0.f;
  out.tri = DMSH_U32MAX;
  if (pk.num_leaf == 0) return kTraceOk;
  RaySlab r;
  make_slab(org, dir, r);
  const bool lane_finite = !r.par[0] && !r.par[1] && !r.par[2] &&
                           fabsf(org[0]) < __builtin_inff() && fabsf(org[1]) < __builtin_inff() &&
                           fabsf(org[2]) < __builtin_inff();
  // one path for the whole wave, so the vote must come before any return
  const bool all_finite = __all(boxes_finite && lane_finite);
  {
    const float lo[3] = {pk.root_box[0], pk.root_box[1], pk.root_box[2]};
    const float hi[3] = {pk.root_box[3], pk.root_box[4], pk.root_box[5]};
    float ta, tb;
    if (!slab_hit(r, lo, hi, ta, tb)) return kTraceOk;
  }
  if (pk.num_leaf == 1) {
    bool has;
    float t;
    uint32_t tri;
    leaf_test(pk, 0, org, dir, has, t, tri);
    out.has = has;
    out.t = t;
    out.tri = tri;
    return kTraceOk;
  }
  if (all_finite) return trace_loop<true>(pk, r, org, dir, stk, out);
  return trace_loop<false>(pk, r, org, dir, stk, out);
}

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each
// XCD receives one contiguous run of logical blocks.
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t nblk) {
  const uint32_t per = nblk / 8u, rem = nblk % 8u, x = b % 8u, local = b / 8u;
  const uint32_t start = (x < rem) ? x * (per + 1u) : rem * (per + 1u) + (x - rem) * per;
  return start + local;
}

struct CastArgs {
  uint32_t* pix2tri;
  float* pix2t;
  uint32_t img_w, img_h, tiles_x, num_tiles;
  uint32_t* counts;  // [0]: tier-2 queue length, [1]: tier-3 queue length,
                     // [2]: nonzero when a packed box is not finite
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep_stack;
  uint32_t* err;
};

__device__ __forceinline__ void write_result(const CastArgs& a, uint32_t ip, const Hit& h) {
  a.pix2tri[ip] = h.has ? h.tri : DMSH_U32MAX;
  if (a.pix2t) a.pix2t[ip] = h.has ? h.t : 0.f;
}

__device__ __forceinline__ void load_transform(const float* __restrict__ T, float m[16]) {
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = T[k];
}

// Tier 1: one ray per lane, 8x8 pixel tiles per wave.
template <int S, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_pix2tri(Packed pk, const float* __restrict__ T,
                                                        CastArgs a) {
  __shared__ uint2 s_stack[S][64 * WAVES];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x) * WAVES + wave;
  if (tile >= a.num_tiles) return;
  const uint32_t iw = (tile % a.tiles_x) * 8u + (lane & 7u);
  const uint32_t ih = (tile / a.tiles_x) * 8u + (lane >> 3);
  if (iw >= a.img_w || ih >= a.img_h) return;
  float m[16];
  load_transform(T, m);
  float org[3], dir[3];
  ray_for_pixel(iw, ih, a.img_w, a.img_h, m, org, dir);
  LdsStack<S> stk{&s_stack[0][tid], 64 * WAVES};
  Hit h;
  const int st = trace(pk, org, dir, stk, h, a.counts[2] == 0u);
  const uint32_t ip = ih * a.img_w + iw;
  if (st == kTraceOverflow) {
    a.list2[atomicAdd(a.counts + 0, 1u)] = ip;
    return;
  }
  write_result(a, ip, h);
}

// Tier 2: queued rays, 64 LDS frames per lane, grid-stride over the queue.
static __global__ __launch_bounds__(64) void k_pix2tri_tier2(Packed pk,
                                                             const float* __restrict__ T,
                                                             CastArgs a) {
  __shared__ uint2 s_stack[kTier2Frames][64];
  const uint32_t num_pix = a.img_w * a.img_h;
  uint32_t count = a.counts[0];
  if (count > num_pix) count = num_pix;
  float m[16];
  load_transform(T, m);
  LdsStack<kTier2Frames> stk{&s_stack[0][threadIdx.x], 64};
  const bool boxes_finite = a.counts[2] == 0u;
  for (uint32_t e = blockIdx.x * 64u + threadIdx.x; e < count; e += gridDim.x * 64u) {
    const uint32_t ip = a.list2[e];
    const uint32_t ih = ip / a.img_w, iw = ip - ih * a.img_w;
    float org[3], dir[3];
    ray_for_pixel(iw, ih, a.img_w, a.img_h, m, org, dir);
    Hit h;
    const int st = trace(pk, org, dir, stk, h, boxes_finite);
    if (st == kTraceOverflow) {
      a.list3[atomicAdd(a.counts + 1, 1u)] = ip;
      continue;
    }
    write_result(a, ip, h);
  }
}

// Tier 3: a fixed grid of kDeepSlots threads, each owning a global stack slot.
static __global__ __launch_bounds__(64) void k_pix2tri_tier3(Packed pk,
                                                             const float* __restrict__ T,
                                                             CastArgs a) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthr = gridDim.x * blockDim.x;
  const uint32_t num_pix = a.img_w * a.img_h;
  uint32_t count = a.counts[1];
  if (count > num_pix) count = num_pix;
  float m[16];
  load_transform(T, m);
  GlobalStack stk{a.deep_stack + (size_t)gid * kDeepFrames};
  const bool boxes_finite = a.counts[2] == 0u;
  for (uint32_t e = gid; e < count; e += nthr) {
    const uint32_t ip = a.list3[e];
    const uint32_t ih = ip / a.img_w, iw = ip - ih * a.img_w;
    float org[3], dir[3];
    ray_for_pixel(iw, ih, a.img_w, a.img_h, m, org, dir);
    Hit h;
    const int st = trace(pk, org, dir, stk, h, boxes_finite);
    if (st != kTraceOk) {
      raise_error(a.err, DMSH_DEVERR_STACK);
      h.has = false;
    }
    write_result(a, ip, h);
  }
}

// ---------------------------------------------------------------------------
// depth
// ---------------------------------------------------------------------------

struct TriMesh {
  const uint32_t* __restrict__ tri2vtx;
  const float* __restrict__ xyz;
  uint32_t num_tri, num_vtx;
};

static __global__ __launch_bounds__(256) void k_pix2depth(float* __restrict__ depth,
                                                          const uint32_t* __restrict__ pix2tri,
                                                          TriMesh sc, uint32_t img_w,
                                                          uint32_t img_h,
                                                          const float* __restrict__ T,
                                                          const float* __restrict__ inv,
                                                          uint32_t* __restrict__ err) {
  const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= img_w * img_h) return;
  const uint32_t tri = pix2tri[ip];
  float out = 0.f;
  if (tri != DMSH_U32MAX && inv[16] != 0.f) {
    if (tri >= sc.num_tri) {
      raise_error(err, DMSH_DEVERR_INDEX);
    } else {
      const uint32_t* tv = sc.tri2vtx + (size_t)tri * 3;
      const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
      if (i0 >= sc.num_vtx || i1 >= sc.num_vtx || i2 >= sc.num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
      } else {
        float m[16], w2n[16];
        load_transform(T, m);
        load_transform(inv, w2n);
        const uint32_t ih = ip / img_w, iw = ip - ih * img_w;
        float org[3], dir[3], p0[3], p1[3], p2[3], t;
        ray_for_pixel(iw, ih, img_w, img_h, m, org, dir);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          p0[k] = sc.xyz[(size_t)i0 * 3 + k];
          p1[k] = sc.xyz[(size_t)i1 * 3 + k];
          p2[k] = sc.xyz[(size_t)i2 * 3 + k];
        }
        if (tri_intersect(p0, p1, p2, org, dir, false, t, nullptr)) {
          const float pos[3] = {t * dir[0] + org[0], t * dir[1] + org[1], t * dir[2] + org[2]};
          float ndc[3];
          if (transform_homogeneous(w2n, pos[0], pos[1], pos[2], ndc)) out = (ndc[2] + 1.f) * 0.5f;
        }
      }
    }
  }
  depth[ip] = out;
}

// d transform_homogeneous(m, p) / d p, rows 0..2 (quotient rule):
// J[r][j] = (m[j*4+r] w - x_r m[j*4+3]) / w^2, x_r the numerator of row r.
__device__ __forceinline__ void jacobian_transform3(const float* m, const float p[3],
                                                    float J[3][3]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  const float inv_w2 = 1.f / (w * w);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float q = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r];
#pragma unroll
    for (int j = 0; j < 3; ++j) J[r][j] = (m[j * 4 + r] * w - q * m[j * 4 + 3]) * inv_w2;
  }
}

// Depth backward, one thread per pixel: the line hit (t, bc) of the stored
// triangle, q = sum bc_k p_k, dL/dq = J(world2ndc, q)^T (0, 0, g/2),
// dL/dt = dir . dL/dq, and the hit's sensitivity to the triangle,
// dL/dp_k = (dL/dt / (n . dir)) bc_k n with n = (p1 - p0) x (p2 - p0);
// nine float atomics per covered pixel.
static __global__ __launch_bounds__(256) void k_pix2depth_bwd(
    float* __restrict__ dldw_xyz, const uint32_t* __restrict__ pix2tri, TriMesh sc,
    const float* __restrict__ dldw_depth, uint32_t img_w, uint32_t img_h,
    const float* __restrict__ T, const float* __restrict__ inv, uint32_t* __restrict__ err) {
  const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= img_w * img_h) return;
  const uint32_t tri = pix2tri[ip];
  if (tri == DMSH_U32MAX || inv[16] == 0.f) return;
  if (tri >= sc.num_tri) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t* tv = sc.tri2vtx + (size_t)tri * 3;
  const uint32_t iv[3] = {tv[0], tv[1], tv[2]};
  if (iv[0] >= sc.num_vtx || iv[1] >= sc.num_vtx || iv[2] >= sc.num_vtx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  float m[16], w2n[16];
  load_transform(T, m);
  load_transform(inv, w2n);
  const uint32_t ih = ip / img_w, iw = ip - ih * img_w;
  float org[3], dir[3], p0[3], p1[3], p2[3], t, bc[3];
  ray_for_pixel(iw, ih, img_w, img_h, m, org, dir);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p0[k] = sc.xyz[(size_t)iv[0] * 3 + k];
    p1[k] = sc.xyz[(size_t)iv[1] * 3 + k];
    p2[k] = sc.xyz[(size_t)iv[2] * 3 + k];
  }
  if (!tri_intersect(p0, p1, p2, org, dir, false, t, bc)) return;
  float q[3], J[3][3], dldq[3], e1[3], e2[3], n[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) q[i] = (bc[0] * p0[i] + bc[1] * p1[i]) + bc[2] * p2[i];
  jacobian_transform3(w2n, q, J);
  const float half_g = 0.5f * dldw_depth[ip];
#pragma unroll
  for (int j = 0; j < 3; ++j) dldq[j] = (J[0][j] * 0.f + J[1][j] * 0.f) + J[2][j] * half_g;
  const float dldt = dot3(dir, dldq);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(e1, e2, n);
  const float s = dldt / dot3(n, dir);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float sk = s * bc[k];
#pragma unroll
    for (int i = 0; i < 3; ++i) unsafeAtomicAdd(dldw_xyz + (size_t)iv[k] * 3 + i, sk * n[i]);
  }
}

// ---------------------------------------------------------------------------
// attribute interpolation (pix2tri.rs:176-304)
// ---------------------------------------------------------------------------
// Barycentric coordinates of the line (org, dir) against the plane of the
// triangle: Moller-Trumbore's u, v with no range or epsilon test; also the
// reciprocal determinant and the two cross products the backward reuses.
__device__ __forceinline__ void plane_line_bc(const float p0[3], const float p1[3],
                                              const float p2[3], const float org[3],
                                              const float dir[3], float& u, float& v,
                                              float& inv_det, float pvec[3], float e1xd[3]) {
  float e1[3], e2[3], tvec[3], qvec[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
    tvec[i] = org[i] - p0[i];
  }
  cross3(dir, e2, pvec);
  const float det = dot3(e1, pvec);
  inv_det = 1.f / det;
  u = inv_det * dot3(tvec, pvec);
  cross3(tvec, e1, qvec);
  v = inv_det * dot3(dir, qvec);
  cross3(e1, dir, e1xd);
}

// One thread per pixel: pix2val = bc0 val[i0] + bc1 val[i1] + bc2 val[i2] on
// covered pixels, 0 on the background.
static __global__ __launch_bounds__(256) void k_interpolate(
    float* __restrict__ pix2val, const uint32_t* __restrict__ pix2tri, TriMesh sc,
    uint32_t num_vdim, const float* __restrict__ vtx2val, uint32_t img_w, uint32_t img_h,
    const float* __restrict__ T, uint32_t* __restrict__ err) {
  const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= img_w * img_h) return;
  float* out = pix2val + (size_t)ip * num_vdim;
  const uint32_t tri = pix2tri[ip];
  uint32_t iv[3] = {0u, 0u, 0u};
  bool ok = tri != DMSH_U32MAX;
  if (ok && tri >= sc.num_tri) {
    raise_error(err, DMSH_DEVERR_INDEX);
    ok = false;
  }
  if (ok) {
    const uint32_t* tv = sc.tri2vtx + (size_t)tri * 3;
    iv[0] = tv[0];
    iv[1] = tv[1];
    iv[2] = tv[2];
    if (iv[0] >= sc.num_vtx || iv[1] >= sc.num_vtx || iv[2] >= sc.num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      ok = false;
    }
  }
  if (!ok) {
    for (uint32_t d = 0; d < num_vdim; ++d) out[d] = 0.f;
    return;
  }
  float m[16];
  load_transform(T, m);
  const uint32_t ih = ip / img_w, iw = ip - ih * img_w;
  float org[3], dir[3], p[3][3], u, v, inv_det, pvec[3], c[3];
  ray_for_pixel(iw, ih, img_w, img_h, m, org, dir);
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) p[k][i] = sc.xyz[(size_t)iv[k] * 3 + i];
  plane_line_bc(p[0], p[1], p[2], org, dir, u, v, inv_det, pvec, c);
  const float bc0 = 1.f - u - v;
  for (uint32_t d = 0; d < num_vdim; ++d)
    out[d] = (bc0 * vtx2val[(size_t)iv[0] * num_vdim + d] +
              u * vtx2val[(size_t)iv[1] * num_vdim + d]) +
             v * vtx2val[(size_t)iv[2] * num_vdim + d];
}

// Backward, one thread per covered pixel.  With gu = dL/dbc1 - dL/dbc0 and
// gv = dL/dbc2 - dL/dbc0 (dL/dt = 0), implicit differentiation of
// u e1 + v e2 - t dir = org - p0 gives dL/dp0 = -bc0 lam, dL/dp1 = -u lam,
// dL/dp2 = -v lam with lam = (gu (dir x e2) + gv (e1 x dir)) / det; and
// dL/dval[ik] = dL/dpix bc_k.  Float atomics accumulate both.
static __global__ __launch_bounds__(256) void k_interpolate_bwd(
    float* __restrict__ dldw_xyz, float* __restrict__ dldw_val,
    const uint32_t* __restrict__ pix2tri, TriMesh sc, uint32_t num_vdim,
    const float* __restrict__ vtx2val, const float* __restrict__ dldw_pix2val, uint32_t img_w,
    uint32_t img_h, const float* __restrict__ T, uint32_t* __restrict__ err) {
  const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= img_w * img_h) return;
  const uint32_t tri = pix2tri[ip];
  if (tri == DMSH_U32MAX) return;
  if (tri >= sc.num_tri) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t* tv = sc.tri2vtx + (size_t)tri * 3;
  const uint32_t iv[3] = {tv[0], tv[1], tv[2]};
  if (iv[0] >= sc.num_vtx || iv[1] >= sc.num_vtx || iv[2] >= sc.num_vtx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const float* g = dldw_pix2val + (size_t)ip * num_vdim;
  float gb[3] = {0.f, 0.f, 0.f};
  for (uint32_t d = 0; d < num_vdim; ++d) {
    const float gd = g[d];
#pragma unroll
    for (int k = 0; k < 3; ++k) gb[k] += vtx2val[(size_t)iv[k] * num_vdim + d] * gd;
  }
  const float gu = gb[1] - gb[0], gv = gb[2] - gb[0];
  float m[16];
  load_transform(T, m);
  const uint32_t ih = ip / img_w, iw = ip - ih * img_w;
  float org[3], dir[3], p[3][3], u, v, inv_det, pvec[3], c[3];
  ray_for_pixel(iw, ih, img_w, img_h, m, org, dir);
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 3; ++i) p[k][i] = sc.xyz[(size_t)iv[k] * 3 + i];
  plane_line_bc(p[0], p[1], p[2], org, dir, u, v, inv_det, pvec, c);
  const float bc[3] = {1.f - u - v, u, v};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float lam = (gu * pvec[i] + gv * c[i]) * inv_det;
#pragma unroll
    for (int k = 0; k < 3; ++k) unsafeAtomicAdd(dldw_xyz + (size_t)iv[k] * 3 + i, -(bc[k] * lam));
  }
  for (uint32_t d = 0; d < num_vdim; ++d) {
    const float gd = g[d];
#pragma unroll
    for (int k = 0; k < 3; ++k) unsafeAtomicAdd(dldw_val + (size_t)iv[k] * num_vdim + d, gd * bc[k]);
  }
}

}  // namespace rc
}  // namespace dmsh
