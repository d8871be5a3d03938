// antialias.hpp -- contour edges and the anti-aliased silhouette (gfx950).
//
// Contour (del-msh-cpu/src/edge2vtx.rs:191-242): an edge is a contour edge
// when its two triangles' unit normals face opposite ways along the view ray
// through the edge midpoint.  Flags are compacted in edge order (block counts,
// one-block scan of the counts, ballot-ranked scatter), so the output order is
// the serial order.
//
// Anti-aliasing (del-msh-cpu/src/antialias.rs:54-159): every contour edge is
// projected to pixels and walked with the DDA of dda.h:6-63 (one step past the
// end); at each walked pixel four neighbour pairs (differentiable_rasterizer.
// cu:15-64) whose first pixel is foreground and second background are
// intersected with the edge, and the crossing ratio rc blends the pair.
//   forward: the serial loop lets the last (edge, step, pair) that writes a
//     pixel win.  The GPU reproduces that exactly in two passes: pass 1 takes
//     the maximum 64-bit key (edge + 1, step * 4 + pair) per written pixel;
//     pass 2 repeats the walk and only the write holding the pixel's key
//     stores its value.
//   backward: one thread per edge accumulates dL/dp of its two vertices over
//     its walk and adds them with six float atomics.
// A walk stops once it has left the image in its direction of travel (no
// pixel after that point can be inside), and edges with a non-finite
// projection are skipped.
#pragma once

#include "dmsh_common.hpp"
#include "mat4.hpp"

namespace dmsh {
namespace aa {

__device__ __forceinline__ void project_xy(const float* m, const float* p, float q[2]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  q[0] = (((m[0] * x + m[4] * y) + m[8] * z) + m[12]) / w;
  q[1] = (((m[1] * x + m[5] * y) + m[9] * z) + m[13]) / w;
}

__device__ __forceinline__ float cross2(float ax, float ay, float bx, float by) {
  return ax * by - ay * bx;
}

// c0 + rc (c1 - c0) = s0 + re (e1 - s0); both in [0, 1].
__device__ __forceinline__ bool intersection_edge2(const float c0[2], const float c1[2],
                                                   const float s0[2], const float e1[2],
                                                   float& rc) {
  const float rx = c1[0] - c0[0], ry = c1[1] - c0[1];
  const float sx = e1[0] - s0[0], sy = e1[1] - s0[1];
  const float d = cross2(rx, ry, sx, sy);
  if (d == 0.f) return false;
  const float ax = s0[0] - c0[0], ay = s0[1] - c0[1];
  const float r0 = cross2(ax, ay, sx, sy) / d;
  const float r1 = cross2(ax, ay, rx, ry) / d;
  if (r0 < 0.f || r0 > 1.f || r1 < 0.f || r1 > 1.f) return false;
  rc = r0;
  return true;
}

__device__ __forceinline__ void dldw_intersection_edge2(const float c0[2], const float c1[2],
                                                        const float s0[2], const float e1[2],
                                                        float dldrc, float dlds0[2],
                                                        float dlde1[2]) {
  const float rx = c1[0] - c0[0], ry = c1[1] - c0[1];
  const float sx = e1[0] - s0[0], sy = e1[1] - s0[1];
  const float ax = s0[0] - c0[0], ay = s0[1] - c0[1];
  const float d = cross2(rx, ry, sx, sy);
  const float rc = cross2(ax, ay, sx, sy) / d;
  const float g = dldrc / d;
  dlds0[0] = ((sy + ay) - rc * ry) * g;
  dlds0[1] = ((-sx - ax) + rc * rx) * g;
  dlde1[0] = (-ay + rc * ry) * g;
  dlde1[1] = (ax - rc * rx) * g;
}

// steps a pixel walk may take before it is abandoned (DMSH_DEVERR_WALK)
constexpr uint32_t kWalkLimit = 1u << 26;

struct Pair {
  uint32_t p0, p1;
  float c0[2], c1[2];
};

// Pair k of the pixel (iw, ih): k = 0 (centre, west|north), 1 (west|north,
// centre), 2 (centre, east|south), 3 (east|south, centre); west/east when
// `horizontal`, else the rows ih - 1 / ih + 1.  Index arithmetic is uint32
// (a west neighbour of column 0 is the previous row's last pixel, the row
// above row 0 wraps past the image and is dropped by the caller).
__device__ __forceinline__ Pair make_pair(uint32_t iw, uint32_t ih, uint32_t img_w,
                                          bool horizontal, int k) {
  const bool far = (k >> 1) != 0;
  const uint32_t pc = ih * img_w + iw;
  const float cx = (float)iw + 0.5f, cy = (float)ih + 0.5f;
  uint32_t pn;
  float nx, ny;
  if (horizontal) {
    pn = far ? ih * img_w + iw + 1u : ih * img_w + iw - 1u;
    nx = far ? (float)iw + 1.5f : (float)iw - 0.5f;
    ny = cy;
  } else {
    pn = far ? (ih + 1u) * img_w + iw : (ih - 1u) * img_w + iw;
    nx = cx;
    ny = far ? (float)ih + 1.5f : (float)ih - 0.5f;
  }
  Pair pr;
  if (k & 1) {
    pr.p0 = pn;
    pr.c0[0] = nx;
    pr.c0[1] = ny;
    pr.p1 = pc;
    pr.c1[0] = cx;
    pr.c1[1] = cy;
  } else {
    pr.p0 = pc;
    pr.c0[0] = cx;
    pr.c0[1] = cy;
    pr.p1 = pn;
    pr.c1[0] = nx;
    pr.c1[1] = ny;
  }
  return pr;
}

// Walks the edge q0 -> q1; calls fn(pair, rc, seq) for every crossed
// foreground/background pair, seq = step * 4 + pair index.  Returns false
// when the walk hit the step cap.
template <class Fn>
__device__ __forceinline__ bool walk(const float q0[2], const float q1[2], uint32_t img_w,
                                     uint32_t img_h, const uint32_t* __restrict__ pix2tri, Fn fn) {
  const float v01x = q1[0] - q0[0], v01y = q1[1] - q0[1];
  if (v01x == 0.f && v01y == 0.f) return true;
  const bool horizontal = fabsf(v01x) < fabsf(v01y);
  float x = q0[0], y = q0[1];
  const float dx = q1[0] - q0[0], dy = q1[1] - q0[1];
  const float step = fmax_sel(fabsf(dx), fabsf(dy));
  const float slope_x = dx / step, slope_y = dy / step;
  const uint32_t total = img_w * img_h;
  const float fw = (float)img_w, fh = (float)img_h;
  for (uint32_t k = 0; k < kWalkLimit; ++k) {
    const float fx = floorf(x), fy = floorf(y);
    if (fx >= 0.f && fy >= 0.f && fx < fw && fy < fh) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const Pair pr = make_pair((uint32_t)fx, (uint32_t)fy, img_w, horizontal, c);
        if (pr.p0 >= total || pr.p1 >= total) continue;
        if (pix2tri[pr.p0] == DMSH_U32MAX || pix2tri[pr.p1] != DMSH_U32MAX) continue;
        float rc;
        if (!intersection_edge2(pr.c0, pr.c1, q1, q0, rc)) continue;
        fn(pr, rc, k * 4u + (uint32_t)c);
      }
    }
    const float eps = 1.0e-3f;
    if (fabsf(x - q0[0]) > fabsf(q1[0] - q0[0]) + eps) return true;
    if (fabsf(y - q0[1]) > fabsf(q1[1] - q0[1]) + eps) return true;
    // left the image for good in the direction of travel: nothing more to do
    if ((slope_x > 0.f && x >= fw) || (slope_x < 0.f && x < 0.f) || (slope_y > 0.f && y >= fh) ||
        (slope_y < 0.f && y < 0.f))
      return true;
    x = x + slope_x;
    y = y + slope_y;
  }
  return false;
}

struct EdgeArgs {
  const uint32_t* __restrict__ cedge2vtx;
  uint32_t num_cedge;
  const float* __restrict__ xyz;
  uint32_t num_vtx;
  const float* __restrict__ T;  // world2pix, column-major
  uint32_t img_w, img_h;
  const uint32_t* __restrict__ pix2tri;
  uint32_t* err;
};

__device__ __forceinline__ bool edge_setup(const EdgeArgs& a, uint32_t e, float m[16],
                                           uint32_t& i0, uint32_t& i1, float q0[2], float q1[2]) {
  i0 = a.cedge2vtx[(size_t)e * 2];
  i1 = a.cedge2vtx[(size_t)e * 2 + 1];
  if (i0 >= a.num_vtx || i1 >= a.num_vtx) {
    raise_error(a.err, DMSH_DEVERR_INDEX);
    return false;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = a.T[k];
  project_xy(m, a.xyz + (size_t)i0 * 3, q0);
  project_xy(m, a.xyz + (size_t)i1 * 3, q1);
  const float big = __builtin_inff();
  return fabsf(q0[0]) < big && fabsf(q0[1]) < big && fabsf(q1[0]) < big && fabsf(q1[1]) < big;
}

static __global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p,
                                                         uint64_t n) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull)
    p[i] = 0ull;
}

static __global__ __launch_bounds__(64) void k_aa_keys(EdgeArgs a,
                                                       unsigned long long* __restrict__ keys) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_cedge) return;
  float m[16], q0[2], q1[2];
  uint32_t i0, i1;
  if (!edge_setup(a, e, m, i0, i1, q0, q1)) return;
  const unsigned long long hi = (unsigned long long)(e + 1) << 32;
  const bool done = walk(q0, q1, a.img_w, a.img_h, a.pix2tri, [&](const Pair& pr, float rc, uint32_t seq) {
    atomicMax(keys + (rc < 0.5f ? pr.p0 : pr.p1), hi | seq);
  });
  if (!done) raise_error(a.err, DMSH_DEVERR_WALK);
}

static __global__ __launch_bounds__(64) void k_aa_write(EdgeArgs a,
                                                        const unsigned long long* __restrict__ keys,
                                                        const float* __restrict__ vin,
                                                        float* __restrict__ vout) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_cedge) return;
  float m[16], q0[2], q1[2];
  uint32_t i0, i1;
  if (!edge_setup(a, e, m, i0, i1, q0, q1)) return;
  const unsigned long long hi = (unsigned long long)(e + 1) << 32;
  walk(q0, q1, a.img_w, a.img_h, a.pix2tri, [&](const Pair& pr, float rc, uint32_t seq) {
    const uint32_t p = rc < 0.5f ? pr.p0 : pr.p1;
    if (keys[p] != (hi | seq)) return;
    const float r0 = rc < 0.5f ? 0.5f + rc : rc - 0.5f;
    vout[p] = r0 * vin[pr.p0] + (1.f - r0) * vin[pr.p1];
  });
}

__device__ __forceinline__ void proj_jacobian_xy(const float* m, const float* p, float J[2][3]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  const float inv_w2 = 1.f / (w * w);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const float q = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r];
#pragma unroll
    for (int j = 0; j < 3; ++j) J[r][j] = (m[j * 4 + r] * w - q * m[j * 4 + 3]) * inv_w2;
  }
}

static __global__ __launch_bounds__(64) void k_aa_bwd(EdgeArgs a, const float* __restrict__ val,
                                                      const float* __restrict__ dldw,
                                                      float* __restrict__ dldw_vtx2xyz) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.num_cedge) return;
  float m[16], q0[2], q1[2];
  uint32_t i0, i1;
  if (!edge_setup(a, e, m, i0, i1, q0, q1)) return;
  float J0[2][3], J1[2][3];
  proj_jacobian_xy(m, a.xyz + (size_t)i0 * 3, J0);
  proj_jacobian_xy(m, a.xyz + (size_t)i1 * 3, J1);
  float g0[3] = {0.f, 0.f, 0.f}, g1[3] = {0.f, 0.f, 0.f};
  const bool done = walk(q0, q1, a.img_w, a.img_h, a.pix2tri, [&](const Pair& pr, float rc, uint32_t) {
    const float dv = val[pr.p0] - val[pr.p1];
    const float dldr0 = rc < 0.5f ? dldw[pr.p0] * dv : dldw[pr.p1] * dv;
    float dldq1[2], dldq0[2];
    dldw_intersection_edge2(pr.c0, pr.c1, q1, q0, dldr0, dldq1, dldq0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      g0[j] += J0[0][j] * dldq0[0] + J0[1][j] * dldq0[1];
      g1[j] += J1[0][j] * dldq1[0] + J1[1][j] * dldq1[1];
    }
  });
  if (!done) raise_error(a.err, DMSH_DEVERR_WALK);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    unsafeAtomicAdd(dldw_vtx2xyz + (size_t)i0 * 3 + j, g0[j]);
    unsafeAtomicAdd(dldw_vtx2xyz + (size_t)i1 * 3 + j, g1[j]);
  }
}

// ---------------------------------------------------------------------------
// contour
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tri_unit_normal(const float* p0, const float* p1, const float* p2,
                                                float n[3]) {
  float e1[3], e2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(e1, e2, n);
  const float len = sqrtf(dot3(n, n));
  const float inv = 1.f / len;
#pragma unroll
  for (int i = 0; i < 3; ++i) n[i] = n[i] * inv;
}

struct ContourArgs {
  const uint32_t* __restrict__ tri2vtx;
  uint32_t num_tri;
  const float* __restrict__ xyz;
  uint32_t num_vtx;
  const float* __restrict__ w2n;  // column-major
  const float* __restrict__ n2w;  // inverse, computed on the device (+ flag at [16])
  uint32_t num_edge;
  const uint32_t* __restrict__ edge2vtx;
  const uint32_t* __restrict__ edge2tri;
  uint32_t* err;
};

__device__ __forceinline__ bool contour_flag(const ContourArgs& a, uint32_t e) {
  if (a.n2w[16] == 0.f) return false;
  const uint32_t i0 = a.edge2vtx[(size_t)e * 2], i1 = a.edge2vtx[(size_t)e * 2 + 1];
  const uint32_t t0 = a.edge2tri[(size_t)e * 2], t1 = a.edge2tri[(size_t)e * 2 + 1];
  if (i0 >= a.num_vtx || i1 >= a.num_vtx || t0 >= a.num_tri || t1 >= a.num_tri) {
    raise_error(a.err, DMSH_DEVERR_INDEX);
    return false;
  }
  const uint32_t* ta = a.tri2vtx + (size_t)t0 * 3;
  const uint32_t* tb = a.tri2vtx + (size_t)t1 * 3;
  if (ta[0] >= a.num_vtx || ta[1] >= a.num_vtx || ta[2] >= a.num_vtx || tb[0] >= a.num_vtx ||
      tb[1] >= a.num_vtx || tb[2] >= a.num_vtx) {
    raise_error(a.err, DMSH_DEVERR_INDEX);
    return false;
  }
  float mid[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    mid[i] = (a.xyz[(size_t)i0 * 3 + i] + a.xyz[(size_t)i1 * 3 + i]) * 0.5f;
  float ndc[3], q0[3], q1[3], dir[3], n0[3], n1[3];
  if (!transform_homogeneous(a.w2n, mid[0], mid[1], mid[2], ndc)) ndc[0] = ndc[1] = ndc[2] = __builtin_nanf("");
  if (!transform_homogeneous(a.n2w, ndc[0], ndc[1], 1.f, q0)) q0[0] = q0[1] = q0[2] = __builtin_nanf("");
  if (!transform_homogeneous(a.n2w, ndc[0], ndc[1], -1.f, q1)) q1[0] = q1[1] = q1[2] = __builtin_nanf("");
#pragma unroll
  for (int i = 0; i < 3; ++i) dir[i] = q1[i] - q0[i];
  tri_unit_normal(a.xyz + (size_t)ta[0] * 3, a.xyz + (size_t)ta[1] * 3, a.xyz + (size_t)ta[2] * 3, n0);
  tri_unit_normal(a.xyz + (size_t)tb[0] * 3, a.xyz + (size_t)tb[1] * 3, a.xyz + (size_t)tb[2] * 3, n1);
  const bool f0 = dot3(n0, dir) > 0.f;
  const bool f1 = dot3(n1, dir) > 0.f;
  return f0 != f1;
}

// pass 1: per-block count of contour edges
static __global__ __launch_bounds__(256) void k_contour_count(ContourArgs a,
                                                              uint32_t* __restrict__ flags,
                                                              uint32_t* __restrict__ block_count) {
  __shared__ uint32_t s_w[4];
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  const bool f = e < a.num_edge && contour_flag(a, e);
  if (e < a.num_edge) flags[e] = f ? 1u : 0u;
  const uint64_t m = __ballot(f);
  if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) block_count[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// pass 2: one block turns the block counts into exclusive offsets (+ total)
static __global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* __restrict__ block_count,
                                                             uint32_t num_blocks,
                                                             uint32_t* __restrict__ total) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < num_blocks; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < num_blocks ? block_count[i] : 0u;
    uint32_t inc = v;
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t u = __shfl_up(inc, d, 64);
      if (lane >= (uint32_t)d) inc += u;
    }
    if (lane == 63u) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t pre = s_carry;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) pre += s_w[w];
    if (i < num_blocks) block_count[i] = pre + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = pre + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

// pass 3: scatter the selected vertex pairs in edge order
static __global__ __launch_bounds__(256) void k_contour_scatter(
    const uint32_t* __restrict__ edge2vtx, uint32_t num_edge, const uint32_t* __restrict__ flags,
    const uint32_t* __restrict__ block_offset, uint32_t* __restrict__ cedge2vtx) {
  __shared__ uint32_t s_w[4];
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  const bool f = e < num_edge && flags[e] != 0u;
  const uint64_t m = __ballot(f);
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) s_w[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (!f) return;
  uint32_t pos = block_offset[blockIdx.x] + mask_rank(m);
  for (uint32_t w = 0; w < wave; ++w) pos += s_w[w];
  cedge2vtx[(size_t)pos * 2] = edge2vtx[(size_t)e * 2];
  cedge2vtx[(size_t)pos * 2 + 1] = edge2vtx[(size_t)e * 2 + 1];
}

}  // namespace aa
}  // namespace dmsh
