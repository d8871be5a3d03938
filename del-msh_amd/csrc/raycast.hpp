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
// Executed as a loop over an explicit frame stack.  A frame is pushed only
// where both children are hit (a single hit child is a tail call) and holds
// 8 bytes: before the near subtree returns it stores (far child, far entry t);
// after the decision to descend far it stores the near result (tri, t).  The
// stack lives in LDS, lane-interleaved ([depth][thread]) so a wave's pushes
// and pops hit 64 distinct banks; a ray that needs more than S frames is
// queued and re-run by a second kernel with a 4096-frame stack in global
// memory.  Pixels are assigned to waves in 8x8 tiles and tiles to XCDs in
// contiguous runs so that neighbouring rays share BVH nodes in one L2.
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {
namespace rc {

constexpr uint32_t kState1 = 0x80000000u;
constexpr uint32_t kNoneTri = 0x7FFFFFFFu;
constexpr uint32_t kDeepFrames = 4096;
constexpr uint32_t kDeepSlots = 256;

enum : int { kTraceOk = 0, kTraceOverflow = 1, kTraceBadIndex = 2 };

struct Scene {
  const uint32_t* __restrict__ tri2vtx;
  const float* __restrict__ xyz;
  const uint32_t* __restrict__ bvhnodes;
  const float* __restrict__ aabb;
  uint32_t num_tri, num_vtx, num_node;
};

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

__device__ __forceinline__ void load_box(const float* __restrict__ aabb, uint32_t node, float lo[3],
                                         float hi[3]) {
  const float2* p = reinterpret_cast<const float2*>(aabb + (size_t)node * 6);
  const float2 a = p[0], b = p[1], c = p[2];
  lo[0] = a.x;
  lo[1] = a.y;
  lo[2] = b.x;
  hi[0] = b.y;
  hi[1] = c.x;
  hi[2] = c.y;
}

// Rust's tuple `<` on (tmin, tmax).
__device__ __forceinline__ bool tuple_lt(float a0, float a1, float b0, float b1) {
  return (a0 == b0) ? (a1 < b1) : (a0 < b0);
}

template <class Stack>
__device__ __forceinline__ int trace(const Scene& sc, const float org[3], const float dir[3],
                                     Stack& stk, Hit& out) {
  out.has = false;
  out.t = 0.f;
  out.tri = DMSH_U32MAX;
  if (sc.num_node == 0) return kTraceOk;
  RaySlab r;
  make_slab(org, dir, r);
  float lo[3], hi[3], ta, tb;
  load_box(sc.aabb, 0, lo, hi);
  if (!slab_hit(r, lo, hi, ta, tb)) return kTraceOk;

  uint32_t node = 0;
  int sp = 0;
  bool descend = true;
  bool res_has = false;
  float res_t = 0.f;
  uint32_t res_tri = DMSH_U32MAX;
  for (;;) {
    if (descend) {
      if (node >= sc.num_node) return kTraceBadIndex;
      const uint32_t* row = sc.bvhnodes + (size_t)node * 3;
      const uint32_t c0 = row[1], c1 = row[2];
      if (c1 == DMSH_U32MAX) {  // leaf: c0 is the triangle
        if (c0 >= sc.num_tri) return kTraceBadIndex;
        const uint32_t* tv = sc.tri2vtx + (size_t)c0 * 3;
        const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
        if (i0 >= sc.num_vtx || i1 >= sc.num_vtx || i2 >= sc.num_vtx) return kTraceBadIndex;
        float p0[3], p1[3], p2[3], t;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          p0[k] = sc.xyz[(size_t)i0 * 3 + k];
          p1[k] = sc.xyz[(size_t)i1 * 3 + k];
          p2[k] = sc.xyz[(size_t)i2 * 3 + k];
        }
        res_has = tri_intersect(p0, p1, p2, org, dir, true, t, nullptr) && (t < __builtin_inff());
        res_t = res_has ? t : 0.f;
        res_tri = res_has ? c0 : DMSH_U32MAX;
        descend = false;
        continue;
      }
      if (c0 >= sc.num_node || c1 >= sc.num_node) return kTraceBadIndex;
      float l0[3], h0[3], l1[3], h1[3], tl0, tl1, tr0, tr1;
      load_box(sc.aabb, c0, l0, h0);
      load_box(sc.aabb, c1, l1, h1);
      const bool hl = slab_hit(r, l0, h0, tl0, tl1);
      const bool hr = slab_hit(r, l1, h1, tr0, tr1);
      if (hl && hr) {
        const bool nl = tuple_lt(tl0, tl1, tr0, tr1);
        if (sp == Stack::kCap) return kTraceOverflow;
        stk.store(sp++, make_uint2(nl ? c1 : c0, __float_as_uint(nl ? tr0 : tl0)));
        node = nl ? c0 : c1;
      } else if (hl || hr) {
        node = hl ? c0 : c1;
      } else {
        res_has = false;
        res_t = 0.f;
        res_tri = DMSH_U32MAX;
        descend = false;
      }
    } else {
      if (sp == 0) break;
      const uint2 f = stk.load(sp - 1);
      if (!(f.x & kState1)) {
        // near subtree done: prune the far child or descend into it
        if (res_has && __uint_as_float(f.y) > res_t) {
          --sp;
        } else {
          stk.store(sp - 1, make_uint2(kState1 | (res_has ? res_tri : kNoneTri),
                                       __float_as_uint(res_t)));
          node = f.x;
          descend = true;
        }
      } else {
        // far subtree done: combine with the stored near result
        --sp;
        const uint32_t rn_tri = f.x & ~kState1;
        if (rn_tri != kNoneTri) {
          const float rn_t = __uint_as_float(f.y);
          if (!res_has || rn_t < res_t) {
            res_has = true;
            res_t = rn_t;
            res_tri = rn_tri;
          }
        }
      }
    }
  }
  out.has = res_has;
  out.t = res_t;
  out.tri = res_tri;
  return kTraceOk;
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
  uint32_t* ovf_count;
  uint32_t* ovf_list;
  uint2* deep_stack;
  uint32_t* err;
};

__device__ __forceinline__ void write_result(const CastArgs& a, uint32_t ip, const Hit& h) {
  a.pix2tri[ip] = h.has ? h.tri : DMSH_U32MAX;
  if (a.pix2t) a.pix2t[ip] = h.has ? h.t : 0.f;
}

template <int S, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_pix2tri(Scene sc, const float* __restrict__ T,
                                                        CastArgs a) {
  __shared__ uint2 s_stack[S][64 * WAVES];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  const uint32_t tile = xcd_contiguous(blockIdx.x, gridDim.x) * WAVES + wave;
  if (tile >= a.num_tiles) return;
  const uint32_t iw = (tile % a.tiles_x) * 8u + (lane & 7u);
  const uint32_t ih = (tile / a.tiles_x) * 8u + (lane >> 3);
  if (iw >= a.img_w || ih >= a.img_h) return;
  float m[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = T[k];
  float org[3], dir[3];
  ray_for_pixel(iw, ih, a.img_w, a.img_h, m, org, dir);
  LdsStack<S> stk{&s_stack[0][tid], 64 * WAVES};
  Hit h;
  const int st = trace(sc, org, dir, stk, h);
  const uint32_t ip = ih * a.img_w + iw;
  if (st == kTraceOverflow) {
    const uint32_t slot = atomicAdd(a.ovf_count, 1u);
    a.ovf_list[slot] = ip;
    return;
  }
  if (st == kTraceBadIndex) raise_error(a.err, DMSH_DEVERR_INDEX);
  write_result(a, ip, h);
}

// Re-runs the queued rays with a deep stack; a fixed grid of kDeepSlots
// threads strides over the queue, each thread owning one stack slot.
static __global__ __launch_bounds__(64) void k_pix2tri_deep(Scene sc, const float* __restrict__ T,
                                                     CastArgs a) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthr = gridDim.x * blockDim.x;
  const uint32_t num_pix = a.img_w * a.img_h;
  uint32_t count = *a.ovf_count;
  if (count > num_pix) count = num_pix;
  float m[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = T[k];
  GlobalStack stk{a.deep_stack + (size_t)gid * kDeepFrames};
  for (uint32_t e = gid; e < count; e += nthr) {
    const uint32_t ip = a.ovf_list[e];
    const uint32_t ih = ip / a.img_w, iw = ip - ih * a.img_w;
    float org[3], dir[3];
    ray_for_pixel(iw, ih, a.img_w, a.img_h, m, org, dir);
    Hit h;
    const int st = trace(sc, org, dir, stk, h);
    if (st == kTraceOverflow) raise_error(a.err, DMSH_DEVERR_STACK);
    if (st == kTraceBadIndex) raise_error(a.err, DMSH_DEVERR_INDEX);
    if (st != kTraceOk) h.has = false;
    write_result(a, ip, h);
  }
}

// ---------------------------------------------------------------------------
// depth
// ---------------------------------------------------------------------------

// Gauss-Jordan inverse with partial pivoting of a column-major 4x4 (pivot =
// first row with the largest |a|); returns false when singular.
__device__ __forceinline__ bool mat4_inverse(const float* m, float* inv) {
  float a[4][8];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][4 + c] = (r == c) ? 1.f : 0.f;
    }
  }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    float best = fabsf(a[c][c]);
    for (int r = c + 1; r < 4; ++r) {
      if (fabsf(a[r][c]) > best) {
        best = fabsf(a[r][c]);
        piv = r;
      }
    }
    if (best == 0.f) return false;
    if (piv != c) {
      for (int k = 0; k < 8; ++k) {
        const float tmp = a[c][k];
        a[c][k] = a[piv][k];
        a[piv][k] = tmp;
      }
    }
    const float ip = 1.f / a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] = a[c][k] * ip;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const float f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] = a[r][k] - f * a[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][4 + c];
  return true;
}

static __global__ __launch_bounds__(256) void k_pix2depth(float* __restrict__ depth,
                                                   const uint32_t* __restrict__ pix2tri, Scene sc,
                                                   uint32_t img_w, uint32_t img_h,
                                                   const float* __restrict__ T,
                                                   uint32_t* __restrict__ err) {
  __shared__ float s_inv[16];
  __shared__ int s_ok;
  if (threadIdx.x == 0) {
    float m[16];
    for (int k = 0; k < 16; ++k) m[k] = T[k];
    s_ok = mat4_inverse(m, s_inv) ? 1 : 0;
    if (!s_ok && blockIdx.x == 0) raise_error(err, DMSH_DEVERR_SINGULAR);
  }
  __syncthreads();
  const uint32_t ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= img_w * img_h) return;
  const uint32_t tri = pix2tri[ip];
  float out = 0.f;
  if (tri != DMSH_U32MAX && s_ok) {
    if (tri >= sc.num_tri) {
      raise_error(err, DMSH_DEVERR_INDEX);
    } else {
      const uint32_t* tv = sc.tri2vtx + (size_t)tri * 3;
      const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
      if (i0 >= sc.num_vtx || i1 >= sc.num_vtx || i2 >= sc.num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
      } else {
        float m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = T[k];
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
          if (transform_homogeneous(s_inv, pos[0], pos[1], pos[2], ndc)) out = (ndc[2] + 1.f) * 0.5f;
        }
      }
    }
  }
  depth[ip] = out;
}

}  // namespace rc
}  // namespace dmsh
