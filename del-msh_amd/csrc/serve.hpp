// serve.hpp -- kernels behind the serving loop: trees packed into
// self-contained device buffers (one per cached mesh) and one ray-cast
// launch over the images of many requests, each naming its tree by address.
//
// Packed tree of N triangles (dmsh_packed_bytes): inner nodes [N-1] (64 B),
// leaves [N] (48 B), the root box (6 floats), then a word that is nonzero
// when a box of the tree is not finite -- the traversal layout of
// raycast.hpp, so a cast reads nothing else.
#pragma once

#include "batch.hpp"

namespace dmsh {
namespace srv {

struct PackedRef {
  rc::InnerNode* inner;
  rc::LeafTri* leaf;
  float* root_box;
};

__host__ __device__ __forceinline__ size_t packed_leaf_offset(uint32_t n) {
  return (size_t)(n > 1 ? n - 1 : 0) * sizeof(rc::InnerNode);
}
__host__ __device__ __forceinline__ size_t packed_box_offset(uint32_t n) {
  return packed_leaf_offset(n) + (size_t)n * sizeof(rc::LeafTri);
}
__host__ __device__ __forceinline__ size_t packed_flag_offset(uint32_t n) {
  return packed_box_offset(n) + 6 * sizeof(float);
}
__host__ __device__ __forceinline__ size_t packed_bytes(uint32_t n) {
  return packed_flag_offset(n) + sizeof(uint32_t);
}

__device__ __forceinline__ PackedRef packed_ref(uint64_t addr, uint32_t n) {
  char* base = reinterpret_cast<char*>(addr);
  return PackedRef{reinterpret_cast<rc::InnerNode*>(base),
                   reinterpret_cast<rc::LeafTri*>(base + packed_leaf_offset(n)),
                   reinterpret_cast<float*>(base + packed_box_offset(n))};
}
__device__ __forceinline__ uint32_t* packed_flag(uint64_t addr, uint32_t n) {
  return reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(addr) + packed_flag_offset(n));
}

// The pack sink of one mesh for the batched refit: its records go straight
// into the mesh's packed buffer, as k_pack_to_buffers writes them.
struct TreePack {
  static constexpr bool kOn = true;
  rc::InnerNode* inner;
  rc::LeafTri* leaf;
  uint32_t* nonfinite;
  // the records rc::pack_leaf writes, from the corners the refit loaded
  __device__ __forceinline__ void leaf_ok(uint32_t k, const float (&p0)[3], const float (&p1)[3],
                                          const float (&p2)[3], uint32_t tri) const {
    rc::LeafTri lt;
    lt.a = make_float4(p0[0], p0[1], p0[2], p1[0]);
    lt.b = make_float4(p1[1], p1[2], p2[0], p2[1]);
    lt.c = make_float4(p2[2], __uint_as_float(tri), 0.f, 0.f);
    leaf[k] = lt;
  }
  __device__ __forceinline__ void leaf_bad(uint32_t k) const {
    rc::LeafTri lt;
    lt.a = make_float4(0.f, 0.f, 0.f, 0.f);
    lt.b = make_float4(0.f, 0.f, 0.f, 0.f);
    lt.c = make_float4(0.f, __uint_as_float(DMSH_U32MAX), 0.f, 0.f);
    leaf[k] = lt;
  }
  // 16-byte part `part` of internal node i's record (rc::inner_part_bits)
  __device__ __forceinline__ void inner_part(uint32_t i, uint32_t part, uint32_t c0, uint32_t c1,
                                             float x0, float x1, float x2, float x3,
                                             uint32_t num_leaf, uint32_t* err) const {
    reinterpret_cast<uint4*>(inner + i)[part] =
        rc::inner_part_bits(part, c0, c1, x0, x1, x2, x3, num_leaf, nonfinite, err);
  }
  __device__ __forceinline__ void inner_at(uint32_t i, uint32_t c0, uint32_t c1,
                                           const float (&b0)[6], const float (&b1)[6],
                                           uint32_t num_leaf, uint32_t* err) const {
    rc::store_inner(inner, i, c0, c1, b0, b1, num_leaf, nonfinite, err);
  }
};

struct PackSrc {
  using PK = TreePack;
  const uint64_t* dst;  // [M] packed buffer addresses
  uint32_t* nonfinite;  // [M], zeroed before the build
  __device__ __forceinline__ TreePack of(uint32_t m, uint32_t n) const {
    const PackedRef p = packed_ref(dst[m], n);
    return TreePack{p.inner, p.leaf, nonfinite + m};
  }
};

// After a build that packed as it refit: each mesh's root box and its
// non-finite flag into its packed buffer.
static __global__ __launch_bounds__(256) void k_pack_roots(uint32_t M,
                                                           const uint32_t* __restrict__ mesh2tri,
                                                           const uint32_t* __restrict__ node_off,
                                                           const float* __restrict__ aabb,
                                                           const uint64_t* __restrict__ dst,
                                                           const uint32_t* __restrict__ nonfinite,
                                                           const uint32_t* __restrict__ bad) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (*bad || m >= M) return;
  const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
  if (n == 0) return;
  const PackedRef p = packed_ref(dst[m], n);
#pragma unroll
  for (int d = 0; d < 6; ++d) p.root_box[d] = aabb[(size_t)node_off[m] * 6 + d];
  *packed_flag(dst[m], n) = nonfinite[m];
}

// The per-mesh non-finite flags into the packed buffers (after the pack).
static __global__ __launch_bounds__(256) void k_pack_flags(uint32_t M,
                                                           const uint32_t* __restrict__ mesh2tri,
                                                           const uint64_t* __restrict__ dst,
                                                           const uint32_t* __restrict__ nonfinite,
                                                           const uint32_t* __restrict__ bad) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (*bad || m >= M) return;
  const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
  if (n > 0) *packed_flag(dst[m], n) = nonfinite[m];
}

// Internal nodes of a CSR batch of trees into their packed buffers; also the
// root boxes and a per-mesh flag for non-finite boxes.
static __global__ __launch_bounds__(256) void k_pack_to_buffers(
    bat::BatchTrees b, uint32_t total_tri, const uint32_t* __restrict__ inner_off,
    const uint64_t* __restrict__ dst, uint32_t* __restrict__ nonfinite,
    uint32_t* __restrict__ err) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (*b.bad || g >= total_tri) return;
  // leaves: one per triangle of the batch
  {
    const uint32_t m = bat::find_segment(b.mesh2tri, b.M, g);
    const TreeView v = bat::tree_of(b, m);
    const uint32_t k = g - b.mesh2tri[m];
    const PackedRef p = packed_ref(dst[m], v.num_leaf);
    rc::pack_leaf(p.leaf, v.bvhnodes, v.tri2vtx, v.xyz, v.num_leaf, v.num_leaf, v.num_vtx, k, err);
    if (k == 0) {
#pragma unroll
      for (int d = 0; d < 6; ++d) p.root_box[d] = v.aabb[d];
    }
  }
  // internal nodes: the first sum (N - 1) threads
  if (g < inner_off[b.M]) {
    const uint32_t m = bat::find_segment(inner_off, b.M, g);
    const TreeView v = bat::tree_of(b, m);
    const PackedRef p = packed_ref(dst[m], v.num_leaf);
    rc::pack_inner(p.inner, v.bvhnodes, v.aabb, v.num_leaf, g - inner_off[m], nonfinite + m, err);
  }
}

// One ray-cast request (the C ABI's dmsh_cast_request).
struct Request {
  uint64_t packed;
  uint32_t num_tri;
  uint32_t nonfinite;
  uint32_t img_w, img_h;
  uint32_t pix_start;
  uint32_t reserved;
  float ndc2world[16];
};
static_assert(sizeof(Request) == 96, "request layout");

struct ReqCast {
  uint32_t R;
  const Request* req;
  const uint32_t* tile_off;  // [R + 1]
  const uint32_t* pix_off;   // [R + 1]
  uint32_t* pix2tri;
  float* pix2t;
  uint32_t* counts;  // [0] tier-2 queue, [1] tier-3 queue, [3] requests rejected,
                     // [16 + 16 x] tile counter of XCD x
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep_stack;
  uint32_t* err;
};

// Tile and pixel offsets of the requests; rejects (counts[3], DMSH_DEVERR_INDEX)
// images that are empty or do not follow one another in pixel order.
static __global__ __launch_bounds__(1024) void k_req_offsets(uint32_t R, uint32_t total_pix,
                                                             const Request* __restrict__ req,
                                                             uint32_t* __restrict__ tile_off,
                                                             uint32_t* __restrict__ pix_off,
                                                             uint32_t* __restrict__ bad,
                                                             uint32_t* __restrict__ err) {
  bat::scan_meshes(R, pix_off, [&](uint32_t r) { return req[r].img_w * req[r].img_h; });
  bat::scan_meshes(R, tile_off, [&](uint32_t r) {
    return ceil_div_u32(req[r].img_w, 8u) * ceil_div_u32(req[r].img_h, 8u);
  });
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < R; r += 1024) {
    const Request& q = req[r];
    // a running sum that wrapped past 2^32 shows up as a decrease
    const bool ok = q.img_w > 0 && q.img_h > 0 && q.pix_start == pix_off[r] &&
                    (uint64_t)q.img_w * q.img_h < (1ull << 32) && q.num_tri < rc::kMaxTri &&
                    pix_off[r + 1] > pix_off[r] && tile_off[r + 1] > tile_off[r];
    if (!ok) {
      atomicOr(bad, 1u);
      raise_error(err, DMSH_DEVERR_INDEX);
    }
  }
  if (threadIdx.x == 0 && pix_off[R] != total_pix) {
    atomicOr(bad, 1u);
    raise_error(err, DMSH_DEVERR_INDEX);
  }
}

__device__ __forceinline__ rc::Packed packed_of(const Request& q) {
  const PackedRef p = packed_ref(q.packed, q.num_tri);
  return rc::Packed{p.inner, p.leaf, p.root_box, q.num_tri};
}

// The general slab test is taken when the request or the tree says so.
__device__ __forceinline__ bool boxes_finite(const Request& q) {
  return q.nonfinite == 0u && (q.num_tri == 0 || *packed_flag(q.packed, q.num_tri) == 0u);
}

__device__ __forceinline__ void req_write(const ReqCast& a, uint32_t gp, const rc::Hit& h) {
  a.pix2tri[gp] = h.has ? h.tri : DMSH_U32MAX;
  if (a.pix2t) a.pix2t[gp] = h.has ? h.t : 0.f;
}

__device__ __forceinline__ uint32_t req_next_tile(uint32_t* counts, uint32_t num_tiles,
                                                  uint32_t& xcd) {
  const uint32_t per = (num_tiles + bat::kXcds - 1) / bat::kXcds;
  for (uint32_t tries = 0; tries < bat::kXcds; ++tries) {
    const uint32_t lo = xcd * per;
    const uint32_t hi = (lo + per < num_tiles) ? lo + per : num_tiles;
    uint32_t t = 0;
    if ((threadIdx.x & 63u) == 0) t = atomicAdd(counts + 16 + 16 * xcd, 1u);
    t = __builtin_amdgcn_readfirstlane(t) + lo;
    if (t < hi) return t;
    xcd = (xcd + 1) % bat::kXcds;
  }
  return num_tiles;
}

template <int S>
__global__ __launch_bounds__(64) void k_req_pix2tri(ReqCast a) {
  __shared__ uint2 s_stack[S][64];
  const uint32_t lane = threadIdx.x;
  if (a.counts[3]) return;
  const uint32_t num_tiles = a.tile_off[a.R];
  uint32_t xcd = blockIdx.x % bat::kXcds;
  for (;;) {
    const uint32_t tile = req_next_tile(a.counts, num_tiles, xcd);
    if (tile >= num_tiles) break;
    const uint32_t r = bat::find_segment(a.tile_off, a.R, tile);
    const Request& q = a.req[r];
    const uint32_t w = q.img_w, h = q.img_h;
    const uint32_t tiles_x = ceil_div_u32(w, 8u), lt = tile - a.tile_off[r];
    const uint32_t iw = (lt % tiles_x) * 8u + (lane & 7u), ih = (lt / tiles_x) * 8u + (lane >> 3);
    if (iw < w && ih < h) {
      float T[16];
      rc::load_transform(q.ndc2world, T);
      float org[3], dir[3];
      ray_for_pixel(iw, ih, w, h, T, org, dir);
      rc::LdsStack<S> stk{&s_stack[0][lane], 64};
      rc::Hit hit;
      const int st = rc::trace(packed_of(q), org, dir, stk, hit, boxes_finite(q));
      const uint32_t gp = q.pix_start + ih * w + iw;
      if (st == rc::kTraceOverflow) a.list2[atomicAdd(a.counts + 0, 1u)] = gp;
      else req_write(a, gp, hit);
    }
  }
}

__device__ __forceinline__ uint32_t req_pixel(const ReqCast& a, uint32_t gp, uint32_t& iw,
                                              uint32_t& ih) {
  const uint32_t r = bat::find_segment(a.pix_off, a.R, gp);
  const uint32_t w = a.req[r].img_w;
  ih = (gp - a.pix_off[r]) / w;
  iw = gp - a.pix_off[r] - ih * w;
  return r;
}

static __global__ __launch_bounds__(64) void k_req_pix2tri_tier2(ReqCast a, uint32_t total_pix) {
  __shared__ uint2 s_stack[rc::kTier2Frames][64];
  if (a.counts[3]) return;
  uint32_t count = a.counts[0];
  if (count > total_pix) count = total_pix;
  rc::LdsStack<rc::kTier2Frames> stk{&s_stack[0][threadIdx.x], 64};
  for (uint32_t e = blockIdx.x * 64u + threadIdx.x; e < count; e += gridDim.x * 64u) {
    const uint32_t gp = a.list2[e];
    uint32_t iw, ih;
    const uint32_t r = req_pixel(a, gp, iw, ih);
    const Request& q = a.req[r];
    float T[16], org[3], dir[3];
    rc::load_transform(q.ndc2world, T);
    ray_for_pixel(iw, ih, q.img_w, q.img_h, T, org, dir);
    rc::Hit hit;
    const int st = rc::trace(packed_of(q), org, dir, stk, hit, boxes_finite(q));
    if (st == rc::kTraceOverflow) a.list3[atomicAdd(a.counts + 1, 1u)] = gp;
    else req_write(a, gp, hit);
  }
}

static __global__ __launch_bounds__(64) void k_req_pix2tri_tier3(ReqCast a, uint32_t total_pix) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthr = gridDim.x * blockDim.x;
  if (a.counts[3]) return;
  uint32_t count = a.counts[1];
  if (count > total_pix) count = total_pix;
  rc::GlobalStack stk{a.deep_stack + (size_t)gid * rc::kDeepFrames};
  for (uint32_t e = gid; e < count; e += nthr) {
    const uint32_t gp = a.list3[e];
    uint32_t iw, ih;
    const uint32_t r = req_pixel(a, gp, iw, ih);
    const Request& q = a.req[r];
    float T[16], org[3], dir[3];
    rc::load_transform(q.ndc2world, T);
    ray_for_pixel(iw, ih, q.img_w, q.img_h, T, org, dir);
    rc::Hit hit;
    const int st = rc::trace(packed_of(q), org, dir, stk, hit, boxes_finite(q));
    if (st != rc::kTraceOk) {
      raise_error(a.err, DMSH_DEVERR_STACK);
      hit.has = false;
    }
    req_write(a, gp, hit);
  }
}

}  // namespace srv
}  // namespace dmsh
