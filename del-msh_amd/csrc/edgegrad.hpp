// edgegrad.hpp -- edge-gradient backward of the rasterized silhouette (gfx950).
//
// For every pair of 4-neighbour pixels whose triangles differ, the pixel
// centres are tested against the other pixel's projected triangle (2D
// barycentric signs; background and degenerate triangles count as inside).
// When exactly one centre lies inside the other triangle, the triangle that
// does not cover it receives dL/dp = sum_k (g0+g1)/2 (v0-v1) through the
// barycentric weights of its own pixel centre and the perspective Jacobian
// of pixel y (vertical pairs) or pixel x (horizontal pairs); see
// del-msh-cpu/src/rasterized_edge_gradient.rs:1-74 and :434-603.
//
// One thread owns pixel (iw, ih) and handles the pair below it and the pair
// to its right, so a 64-wide row of threads reads pix2tri / pix2val rows
// coalesced; the scatter to the three vertices uses hardware float atomics
// (global_atomic_add_f32), so the summation order is not deterministic and
// results agree with the serial CPU loop to rounding only.
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {
namespace eg {

__device__ __forceinline__ float tri2_area(const float p0[2], const float p1[2],
                                           const float p2[2]) {
  return 0.5f * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

__device__ __forceinline__ void project_xy(const float* m, const float* p, float q[2]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  q[0] = (((m[0] * x + m[4] * y) + m[8] * z) + m[12]) / w;
  q[1] = (((m[1] * x + m[5] * y) + m[9] * z) + m[13]) / w;
}

struct Ctx {
  const uint32_t* __restrict__ tri2vtx;
  const float* __restrict__ xyz;
  uint32_t num_tri, num_vtx;
  float m[16];
  uint32_t* err;
};

// 1: barycentrics of pc in itri; 0: none (background / degenerate); -1: bad index.
__device__ __forceinline__ int bary(const Ctx& c, const float pc[2], uint32_t itri, float b[3]) {
  if (itri == DMSH_U32MAX) return 0;
  if (itri >= c.num_tri) return -1;
  const uint32_t* tv = c.tri2vtx + (size_t)itri * 3;
  const uint32_t i0 = tv[0], i1 = tv[1], i2 = tv[2];
  if (i0 >= c.num_vtx || i1 >= c.num_vtx || i2 >= c.num_vtx) return -1;
  float p0[2], p1[2], p2[2];
  project_xy(c.m, c.xyz + (size_t)i0 * 3, p0);
  project_xy(c.m, c.xyz + (size_t)i1 * 3, p1);
  project_xy(c.m, c.xyz + (size_t)i2 * 3, p2);
  const float area = tri2_area(p0, p1, p2);
  if (area == 0.f) return 0;
  const float inv_area = 1.f / area;
  b[0] = tri2_area(pc, p1, p2) * inv_area;
  b[1] = tri2_area(p0, pc, p2) * inv_area;
  b[2] = 1.f - b[0] - b[1];
  return 1;
}

__device__ __forceinline__ bool inside(int has, const float b[3]) {
  if (has <= 0) return true;
  return (b[0] >= 0.f && b[1] >= 0.f && b[2] >= 0.f) || (b[0] <= 0.f && b[1] <= 0.f && b[2] <= 0.f);
}

__device__ __forceinline__ bool proj_grad(const float* m, const float* xyz, int axis, float d[3]) {
  const float q = m[axis] * xyz[0] + m[axis + 4] * xyz[1] + m[axis + 8] * xyz[2] + m[axis + 12];
  const float w = m[3] * xyz[0] + m[7] * xyz[1] + m[11] * xyz[2] + m[15];
  if (fabsf(w) <= FLT_EPSILON) return false;
  const float inv_w2 = 1.f / (w * w);
  d[0] = (m[axis] * w - q * m[3]) * inv_w2;
  d[1] = (m[axis + 4] * w - q * m[7]) * inv_w2;
  d[2] = (m[axis + 8] * w - q * m[11]) * inv_w2;
  return true;
}

__device__ __forceinline__ void edge_pair(const Ctx& c, const uint32_t* __restrict__ pix2tri,
                                          uint32_t nv, const float* __restrict__ pix2val,
                                          const float* __restrict__ dpix, size_t ip0, size_t ip1,
                                          const float pc0[2], const float pc1[2], int axis,
                                          float* __restrict__ dldw) {
  const uint32_t t0 = pix2tri[ip0], t1 = pix2tri[ip1];
  if (t0 == t1) return;
  float b[3];
  const int h0 = bary(c, pc0, t1, b);
  const bool in0 = inside(h0, b);
  const int h1 = bary(c, pc1, t0, b);
  const bool in1 = inside(h1, b);
  if (h0 < 0 || h1 < 0) {
    raise_error(c.err, DMSH_DEVERR_INDEX);
    return;
  }
  if (!in0 && !in1) return;
  if (in0 && in1) return;  // intersecting triangles: not handled by the reference either
  float dldpa = 0.f;
  for (uint32_t k = 0; k < nv; ++k) {
    const float v0 = pix2val[ip0 * nv + k], v1 = pix2val[ip1 * nv + k];
    const float g0 = dpix[ip0 * nv + k], g1 = dpix[ip1 * nv + k];
    dldpa += (g0 + g1) * 0.5f * (v0 - v1);
  }
  const uint32_t itri = in1 ? t1 : t0;
  const float* pc = in1 ? pc1 : pc0;
  if (bary(c, pc, itri, b) != 1) return;
  const uint32_t* tv = c.tri2vtx + (size_t)itri * 3;
#pragma unroll
  for (int inode = 0; inode < 3; ++inode) {
    const uint32_t iv = tv[inode];
    float d[3];
    if (!proj_grad(c.m, c.xyz + (size_t)iv * 3, axis, d)) continue;
    unsafeAtomicAdd(dldw + (size_t)iv * 3 + 0, b[inode] * d[0] * dldpa);
    unsafeAtomicAdd(dldw + (size_t)iv * 3 + 1, b[inode] * d[1] * dldpa);
    unsafeAtomicAdd(dldw + (size_t)iv * 3 + 2, b[inode] * d[2] * dldpa);
  }
}

// block (64, 4): threadIdx.x runs along a pixel row.
static __global__ __launch_bounds__(256) void k_edge_grad_bwd(
    const uint32_t* __restrict__ tri2vtx, uint32_t num_tri, const float* __restrict__ xyz,
    uint32_t num_vtx, float* __restrict__ dldw, const float* __restrict__ T, uint32_t img_w,
    uint32_t img_h, const uint32_t* __restrict__ pix2tri, uint32_t nv,
    const float* __restrict__ pix2val, const float* __restrict__ dpix, uint32_t* err) {
  const uint32_t iw = blockIdx.x * 64u + threadIdx.x;
  const uint32_t ih = blockIdx.y * 4u + threadIdx.y;
  if (iw >= img_w || ih >= img_h) return;
  Ctx c;
  c.tri2vtx = tri2vtx;
  c.xyz = xyz;
  c.num_tri = num_tri;
  c.num_vtx = num_vtx;
  c.err = err;
#pragma unroll
  for (int k = 0; k < 16; ++k) c.m[k] = T[k];
  const size_t ip = (size_t)ih * img_w + iw;
  const float pc[2] = {(float)iw + 0.5f, (float)ih + 0.5f};
  if (ih + 1 < img_h) {
    const float pcs[2] = {(float)iw + 0.5f, (float)(ih + 1) + 0.5f};
    edge_pair(c, pix2tri, nv, pix2val, dpix, ip, ip + img_w, pc, pcs, 1, dldw);
  }
  if (iw + 1 < img_w) {
    const float pce[2] = {(float)(iw + 1) + 0.5f, (float)ih + 0.5f};
    edge_pair(c, pix2tri, nv, pix2val, dpix, ip, ip + 1, pc, pce, 0, dldw);
  }
}

// ---------------------------------------------------------------------------
// Ordered variant: every vertex sums its terms in the serial loop's order, so
// the result equals del-msh-cpu's bit for bit (rasterized_edge_gradient.rs:
// 449-602 adds term after term into dldw_vtx2xyz; binary32 addition is not
// associative, so only that order gives those bits).
//   pair order s: first the vertical pairs (iw, ih0)-(iw, ih0 + 1) with iw
//   outer and ih0 inner (pixel-y Jacobian), s = iw (H - 1) + ih0; then the
//   horizontal pairs (iw0, ih)-(iw0 + 1, ih) with iw0 outer and ih inner
//   (pixel-x Jacobian), s = W (H - 1) + iw0 H + ih.
//   k_eg_count: terms per pair (0..3, one per vertex slot whose projection
//   gradient exists), in pair order;  exclusive scan -> record offsets and
//   the record count on the device;  k_eg_records: record r = (vertex id,
//   r, the term's three floats), records in pair order and, inside a pair,
//   in vertex-slot order (the loop's order);  a stable radix sort by vertex
//   id keeps that order inside each vertex;  k_eg_sum: one thread per vertex
//   adds its run to dldw_vtx2xyz left to right.
// ---------------------------------------------------------------------------
struct PairRef {
  size_t ip0, ip1;
  float pc0[2], pc1[2];
  int axis;
};

__device__ __forceinline__ PairRef pair_of(uint32_t s, uint32_t img_w, uint32_t img_h) {
  PairRef r;
  const uint32_t num_vert = img_w * (img_h - 1u);
  if (s < num_vert) {
    const uint32_t iw = s / (img_h - 1u), ih0 = s - iw * (img_h - 1u);
    r.ip0 = (size_t)ih0 * img_w + iw;
    r.ip1 = r.ip0 + img_w;
    r.pc0[0] = (float)iw + 0.5f;
    r.pc0[1] = (float)ih0 + 0.5f;
    r.pc1[0] = (float)iw + 0.5f;
    r.pc1[1] = (float)(ih0 + 1u) + 0.5f;
    r.axis = 1;
  } else {
    const uint32_t t = s - num_vert;
    const uint32_t iw0 = t / img_h, ih = t - iw0 * img_h;
    r.ip0 = (size_t)ih * img_w + iw0;
    r.ip1 = r.ip0 + 1;
    r.pc0[0] = (float)iw0 + 0.5f;
    r.pc0[1] = (float)ih + 0.5f;
    r.pc1[0] = (float)(iw0 + 1u) + 0.5f;
    r.pc1[1] = (float)ih + 0.5f;
    r.axis = 0;
  }
  return r;
}

// The pair's terms, as edge_pair computes them: returns the mask of vertex
// slots of the receiving triangle `itri` that get a term; term[k] holds slot
// k's three floats.
__device__ __forceinline__ uint32_t pair_terms(const Ctx& c, const uint32_t* __restrict__ pix2tri,
                                               uint32_t nv, const float* __restrict__ pix2val,
                                               const float* __restrict__ dpix, const PairRef& r,
                                               uint32_t& itri, float term[3][3]) {
  const uint32_t t0 = pix2tri[r.ip0], t1 = pix2tri[r.ip1];
  if (t0 == t1) return 0u;
  float b[3];
  const int h0 = bary(c, r.pc0, t1, b);
  const bool in0 = inside(h0, b);
  const int h1 = bary(c, r.pc1, t0, b);
  const bool in1 = inside(h1, b);
  if (h0 < 0 || h1 < 0) {
    raise_error(c.err, DMSH_DEVERR_INDEX);
    return 0u;
  }
  if (in0 == in1) return 0u;  // neither: shared edge; both: intersection
  float dldpa = 0.f;
  for (uint32_t k = 0; k < nv; ++k) {
    const float v0 = pix2val[r.ip0 * nv + k], v1 = pix2val[r.ip1 * nv + k];
    const float g0 = dpix[r.ip0 * nv + k], g1 = dpix[r.ip1 * nv + k];
    dldpa += (g0 + g1) * 0.5f * (v0 - v1);
  }
  itri = in1 ? t1 : t0;
  if (bary(c, in1 ? r.pc1 : r.pc0, itri, b) != 1) return 0u;
  const uint32_t* tv = c.tri2vtx + (size_t)itri * 3;
  uint32_t mask = 0u;
#pragma unroll
  for (int inode = 0; inode < 3; ++inode) {
    float d[3];
    if (!proj_grad(c.m, c.xyz + (size_t)tv[inode] * 3, r.axis, d)) continue;
    term[inode][0] = b[inode] * d[0] * dldpa;
    term[inode][1] = b[inode] * d[1] * dldpa;
    term[inode][2] = b[inode] * d[2] * dldpa;
    mask |= 1u << inode;
  }
  return mask;
}

__device__ __forceinline__ Ctx make_ctx(const uint32_t* tri2vtx, uint32_t num_tri,
                                        const float* xyz, uint32_t num_vtx, const float* T,
                                        uint32_t* err) {
  Ctx c;
  c.tri2vtx = tri2vtx;
  c.xyz = xyz;
  c.num_tri = num_tri;
  c.num_vtx = num_vtx;
  c.err = err;
#pragma unroll
  for (int k = 0; k < 16; ++k) c.m[k] = T[k];
  return c;
}

static __global__ __launch_bounds__(256) void k_eg_count(
    const uint32_t* __restrict__ tri2vtx, uint32_t num_tri, const float* __restrict__ xyz,
    uint32_t num_vtx, const float* __restrict__ T, uint32_t img_w, uint32_t img_h,
    const uint32_t* __restrict__ pix2tri, uint32_t nv, const float* __restrict__ pix2val,
    const float* __restrict__ dpix, uint32_t num_pair, uint32_t* __restrict__ cnt,
    uint32_t* err) {
  const Ctx c = make_ctx(tri2vtx, num_tri, xyz, num_vtx, T, err);
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < num_pair;
       s += gridDim.x * blockDim.x) {
    uint32_t itri;
    float term[3][3];
    const uint32_t mask = pair_terms(c, pix2tri, nv, pix2val, dpix, pair_of(s, img_w, img_h),
                                     itri, term);
    cnt[s] = (uint32_t)__builtin_popcount(mask);
  }
}

static __global__ __launch_bounds__(256) void k_eg_records(
    const uint32_t* __restrict__ tri2vtx, uint32_t num_tri, const float* __restrict__ xyz,
    uint32_t num_vtx, const float* __restrict__ T, uint32_t img_w, uint32_t img_h,
    const uint32_t* __restrict__ pix2tri, uint32_t nv, const float* __restrict__ pix2val,
    const float* __restrict__ dpix, uint32_t num_pair, const uint32_t* __restrict__ off,
    uint32_t* __restrict__ rec_vtx, uint32_t* __restrict__ rec_idx,
    float* __restrict__ rec_term, uint32_t* err) {
  // off [num_pair + 1]: the exclusive scan of k_eg_count's counts
  const Ctx c = make_ctx(tri2vtx, num_tri, xyz, num_vtx, T, err);
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < num_pair;
       s += gridDim.x * blockDim.x) {
    if (off[s + 1] == off[s]) continue;
    uint32_t itri;
    float term[3][3];
    const uint32_t mask = pair_terms(c, pix2tri, nv, pix2val, dpix, pair_of(s, img_w, img_h),
                                     itri, term);
    uint32_t r = off[s];
#pragma unroll
    for (int inode = 0; inode < 3; ++inode) {
      if (!((mask >> inode) & 1u)) continue;
      rec_vtx[r] = tri2vtx[(size_t)itri * 3 + inode];
      rec_idx[r] = r;
      rec_term[(size_t)r * 3 + 0] = term[inode][0];
      rec_term[(size_t)r * 3 + 1] = term[inode][1];
      rec_term[(size_t)r * 3 + 2] = term[inode][2];
      ++r;
    }
  }
}

// One thread per run of equal vertex ids in the sorted records; the run is in
// pair order, so the additions are the serial loop's.
static __global__ __launch_bounds__(256) void k_eg_sum(const uint32_t* __restrict__ num_rec,
                                                       const uint32_t* __restrict__ vtx_sorted,
                                                       const uint32_t* __restrict__ idx_sorted,
                                                       const float* __restrict__ rec_term,
                                                       uint32_t num_vtx, uint32_t cap,
                                                       float* __restrict__ dldw, uint32_t* err) {
  const uint32_t n = *num_rec < cap ? *num_rec : cap;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t v = vtx_sorted[j];
    if (j > 0 && vtx_sorted[j - 1] == v) continue;
    if (v >= num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    float s0 = dldw[(size_t)v * 3 + 0], s1 = dldw[(size_t)v * 3 + 1], s2 = dldw[(size_t)v * 3 + 2];
    uint32_t q = j;
    // four records per round: their loads are issued together, the sums stay in order
    while (q + 4 <= n && vtx_sorted[q + 3] == v) {
      uint32_t id[4];
      float t[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u) id[u] = idx_sorted[q + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        t[u][0] = rec_term[(size_t)id[u] * 3 + 0];
        t[u][1] = rec_term[(size_t)id[u] * 3 + 1];
        t[u][2] = rec_term[(size_t)id[u] * 3 + 2];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        s0 += t[u][0];
        s1 += t[u][1];
        s2 += t[u][2];
      }
      q += 4;
    }
    for (; q < n && vtx_sorted[q] == v; ++q) {
      const uint32_t id = idx_sorted[q];
      s0 += rec_term[(size_t)id * 3 + 0];
      s1 += rec_term[(size_t)id * 3 + 1];
      s2 += rec_term[(size_t)id * 3 + 2];
    }
    dldw[(size_t)v * 3 + 0] = s0;
    dldw[(size_t)v * 3 + 1] = s1;
    dldw[(size_t)v * 3 + 2] = s2;
  }
}

// ---------------------------------------------------------------------------
// Smoothed variant on the staggered grid of pixel edges
// (rasterized_edge_gradient.rs:77-432): hedge (ih, iw) lies between pixels
// (iw, ih) and (iw, ih + 1), [H-1][W]; vedge (ih, iw) between (iw, ih) and
// (iw + 1, ih), [H][W-1].  Type 0: same triangle; 1: shared edge; 2 / 3: the
// first / second pixel's triangle is in front; 4: intersection.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint8_t pair_type(const Ctx& c, uint32_t t0, uint32_t t1,
                                             const float pc0[2], const float pc1[2]) {
  if (t0 == t1) return 0;
  float b[3];
  const int h0 = bary(c, pc0, t1, b);
  const bool in0 = inside(h0, b);
  const int h1 = bary(c, pc1, t0, b);
  const bool in1 = inside(h1, b);
  if (h0 < 0 || h1 < 0) {
    raise_error(c.err, DMSH_DEVERR_INDEX);
    return 0;
  }
  if (!in0 && !in1) return 1;
  if (in0 && !in1) return 2;
  if (!in0 && in1) return 3;
  return 4;
}

__device__ __forceinline__ float pair_dldr(uint32_t nv, const float* __restrict__ pix2val,
                                           const float* __restrict__ dpix, size_t ip0,
                                           size_t ip1) {
  float r = 0.f;
  for (uint32_t k = 0; k < nv; ++k) {
    const float v0 = pix2val[ip0 * nv + k], v1 = pix2val[ip1 * nv + k];
    const float g0 = dpix[ip0 * nv + k], g1 = dpix[ip1 * nv + k];
    r += (g0 + g1) * 0.5f * (v0 - v1);
  }
  return r;
}

// block (64, 4): the thread of pixel (iw, ih) does the hedge below it and
// the vedge to its right.
static __global__ __launch_bounds__(256) void k_edge_type_grad(
    const uint32_t* __restrict__ tri2vtx, uint32_t num_tri, const float* __restrict__ xyz,
    uint32_t num_vtx, const float* __restrict__ T, uint32_t img_w, uint32_t img_h,
    const uint32_t* __restrict__ pix2tri, uint32_t nv, const float* __restrict__ pix2val,
    const float* __restrict__ dpix, uint8_t* __restrict__ htype, float* __restrict__ hdldr,
    uint8_t* __restrict__ vtype, float* __restrict__ vdldr, uint32_t* err) {
  const uint32_t iw = blockIdx.x * 64u + threadIdx.x;
  const uint32_t ih = blockIdx.y * 4u + threadIdx.y;
  if (iw >= img_w || ih >= img_h) return;
  Ctx c;
  c.tri2vtx = tri2vtx;
  c.xyz = xyz;
  c.num_tri = num_tri;
  c.num_vtx = num_vtx;
  c.err = err;
#pragma unroll
  for (int k = 0; k < 16; ++k) c.m[k] = T[k];
  const size_t ip = (size_t)ih * img_w + iw;
  const float pc[2] = {(float)iw + 0.5f, (float)ih + 0.5f};
  if (ih + 1 < img_h) {
    const float pcs[2] = {(float)iw + 0.5f, (float)(ih + 1) + 0.5f};
    const size_t e = (size_t)ih * img_w + iw;
    htype[e] = pair_type(c, pix2tri[ip], pix2tri[ip + img_w], pc, pcs);
    hdldr[e] = pair_dldr(nv, pix2val, dpix, ip, ip + img_w);
  }
  if (iw + 1 < img_w) {
    const float pce[2] = {(float)(iw + 1) + 0.5f, (float)ih + 0.5f};
    const size_t e = (size_t)ih * (img_w - 1) + iw;
    vtype[e] = pair_type(c, pix2tri[ip], pix2tri[ip + 1], pc, pce);
    vdldr[e] = pair_dldr(nv, pix2val, dpix, ip, ip + 1);
  }
}

__device__ __forceinline__ bool fixed_type(uint8_t t) { return t == 2 || t == 3; }

// Red-black smoothing of both grids (rasterized_edge_gradient.cu:139-275):
// per iteration the cells with (iw + ih) even, then odd, each taking the
// average of its admitted neighbours, summed north, south, west, east
// (hedges) or west, east, north, south (vedges); cells of type 2 / 3 stay
// fixed.  The two grids do not read each other's values, only types.
// It runs K half-steps per launch: a workgroup loads a tile of one grid with
// a halo of K cells into LDS, together with each cell's links (which
// neighbours enter its average, in summation order; worked out from the
// types once per call by k_smooth_links), runs the half-steps in LDS and
// writes the tile back to the other buffer.  After s half-steps every cell at
// least s cells inside the loaded region holds exactly the value the global
// iteration gives, so the tile itself (K cells inside) is exact after up to
// K half-steps.  Launches alternate between two buffers because tiles read
// each other's cells as halo.
constexpr int kSmoothK = 16;
constexpr int kSmoothTR = 32;
constexpr int kSmoothTC = 64;
constexpr int kSmoothThreads = 1024;
constexpr uint8_t kLinkFixed = 0x10;

__device__ __forceinline__ int clampi(int x, int lo, int hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

// link bits of hedge (r, c): bit0 north, bit1 south, bit2 west, bit3 east
__device__ __forceinline__ uint8_t hedge_links(uint32_t W, uint32_t H, const uint8_t* ht,
                                               const uint8_t* vt, int r, int c) {
  const int R = (int)H - 1, C = (int)W;
  if (fixed_type(ht[r * C + c])) return kLinkFixed;
  uint8_t m = 0;
  const int vw = (int)W - 1;
  if (r > 0 && ht[(r - 1) * C + c] != 2) m |= 1;
  if (r + 1 < R && ht[(r + 1) * C + c] != 3) m |= 2;
  if (c > 0 && !fixed_type(vt[r * vw + c - 1]) && !fixed_type(vt[(r + 1) * vw + c - 1])) m |= 4;
  if (c + 1 < C && !fixed_type(vt[r * vw + c]) && !fixed_type(vt[(r + 1) * vw + c])) m |= 8;
  return m;
}

// link bits of vedge (r, c): bit0 west, bit1 east, bit2 north, bit3 south
__device__ __forceinline__ uint8_t vedge_links(uint32_t W, uint32_t H, const uint8_t* ht,
                                               const uint8_t* vt, int r, int c) {
  const int R = (int)H, C = (int)W - 1;
  if (fixed_type(vt[r * C + c])) return kLinkFixed;
  uint8_t m = 0;
  if (c > 0 && vt[r * C + c - 1] != 2) m |= 1;
  if (c + 1 < C && vt[r * C + c + 1] != 3) m |= 2;
  if (r > 0 && !fixed_type(ht[(r - 1) * (int)W + c]) && !fixed_type(ht[(r - 1) * (int)W + c + 1]))
    m |= 4;
  if (r + 1 < R && !fixed_type(ht[r * (int)W + c]) && !fixed_type(ht[r * (int)W + c + 1])) m |= 8;
  return m;
}

// links of every hedge (first nh cells) and vedge (next nv cells)
static __global__ __launch_bounds__(256) void k_smooth_links(
    uint32_t W, uint32_t H, const uint8_t* __restrict__ ht, const uint8_t* __restrict__ vt,
    uint8_t* __restrict__ hm, uint8_t* __restrict__ vm) {
  const uint32_t nh = (H - 1) * W, n = nh + H * (W - 1);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (i < nh) {
      hm[i] = hedge_links(W, H, ht, vt, (int)(i / W), (int)(i % W));
    } else {
      const uint32_t j = i - nh;
      vm[j] = vedge_links(W, H, ht, vt, (int)(j / (W - 1)), (int)(j % (W - 1)));
    }
  }
}

// One cell's update in a half-step: the LDS addresses of its four terms in
// summation order (a disabled link reads the word that holds +0), where the
// result goes (cells that keep their value write to a sink word of their own)
// and 1 / (number of links) in binary64.
struct SmoothCell {
  uint32_t a0, a1, a2, a3, out;
  double inv;
};

// acc / n for n in 1..4, correctly rounded to binary32: for n = 1, 2, 4 the
// binary64 product is exact; for n = 3, acc / 3 is never a binary32 midpoint
// and lies at least 2^-26 (relative) from every one, while RN64(acc *
// RN64(1/3)) is within 2^-52 of it, so the one rounding to binary32 lands on
// the correctly rounded quotient (subnormal results included: they are
// multiples of 2^-149 / 3 off the 2^-150 midpoint grid).
__device__ __forceinline__ float smooth_div(float acc, double inv) {
  return (float)((double)acc * inv);
}

template <int NPT>
__device__ __forceinline__ void smooth_half_step(float* sv, const SmoothCell (&cell)[NPT]) {
  float a[NPT][4];
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    a[k][0] = sv[cell[k].a0];
    a[k][1] = sv[cell[k].a1];
    a[k][2] = sv[cell[k].a2];
    a[k][3] = sv[cell[k].a3];
  }
#pragma unroll
  for (int k = 0; k < NPT; ++k) {
    const float acc = (((0.f + a[k][0]) + a[k][1]) + a[k][2]) + a[k][3];
    sv[cell[k].out] = smooth_div(acc, cell[k].inv);
  }
}

// blockIdx.x < tiles_h: a tile of the hedge grid, else of the vedge grid.
// src / dst: the two buffers of each grid (cells outside the grid load as 0);
// hm / vm: the links.  The tile sits in LDS as two arrays [RR][CC / 2], one
// per colour ((r + c) & 1 of the global cell): cell (lr, lc) is entry
// (lr, lc >> 1) of its colour's array.  Its north and south neighbours are
// entries (lr -/+ 1, lc >> 1) of the other array, and its west and east ones
// (lr, (lc >> 1) - 1 + p) and (lr, (lc >> 1) + p) with p = lc & 1, so the
// threads of a wave, which own consecutive entries, read consecutive words.
// After the arrays come the +0 word and 2 * RR * (CC / 2) sink words.  Every
// thread owns the same NPT entries of each colour in every half-step and
// keeps their SmoothCell in registers; a cell that keeps its value (fixed,
// outside the grid, no links, or on the border ring of the tile) reads +0 and
// writes its sink word.  All reads of a half-step are issued before its
// writes: a cell's neighbours have the other colour, so no cell read in a
// half-step is written in it.  Adding the +0 of a disabled link leaves the
// running sum (it starts at +0 and can never be -0) exactly as skipping the
// term would.
template <int K, int TR, int TC, int T>
__global__ __launch_bounds__(T) void k_smooth_tiled(
    uint32_t W, uint32_t H, const uint8_t* __restrict__ hm, const uint8_t* __restrict__ vm,
    const float* __restrict__ hsrc, float* __restrict__ hdst, const float* __restrict__ vsrc,
    float* __restrict__ vdst, uint32_t tiles_h, uint32_t step0, uint32_t nsteps) {
  constexpr int RR = TR + 2 * K, CC = TC + 2 * K, HC = CC / 2, SZ = RR * HC;
  constexpr int NPT = SZ / T, NL = RR * CC / T, ZERO = 2 * SZ, SINK = 2 * SZ + 1;
  static_assert(CC % 2 == 0 && SZ % T == 0 && (RR * CC) % T == 0,
                "colour entries split evenly over the threads");
  __shared__ float sv[4 * SZ + 1];
  const bool hedge = blockIdx.x < tiles_h;
  const int R = hedge ? (int)H - 1 : (int)H;
  const int C = hedge ? (int)W : (int)W - 1;
  const int tcols = (C + TC - 1) / TC;
  const int t = hedge ? (int)blockIdx.x : (int)(blockIdx.x - tiles_h);
  const int r0 = (t / tcols) * TR - K, c0 = (t % tcols) * TC - K;
  const float* src = hedge ? hsrc : vsrc;
  float* dst = hedge ? hdst : vdst;
  const uint8_t* msk = hedge ? hm : vm;
  float v[NL];
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int i = (int)threadIdx.x + k * T;
    const int r = r0 + i / CC, c = c0 + i % CC;
    const float x = src[clampi(r, 0, R - 1) * C + clampi(c, 0, C - 1)];
    v[k] = (r >= 0 && c >= 0 && r < R && c < C) ? x : 0.f;
  }
  SmoothCell cell[2][NPT];
#pragma unroll
  for (int color = 0; color < 2; ++color) {
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
      const int j = (int)threadIdx.x + k * T;
      const int lr = j / HC, h = j % HC;
      const int p = (color + r0 + c0 + lr) & 1;
      const int lc = 2 * h + p;
      const int r = r0 + lr, c = c0 + lc;
      const uint32_t m = msk[clampi(r, 0, R - 1) * C + clampi(c, 0, C - 1)];
      const bool live = lr >= 1 && lc >= 1 && lr < RR - 1 && lc < CC - 1 && r >= 0 && c >= 0 &&
                        r < R && c < C && !(m & kLinkFixed) && (m & 15u);
      const uint32_t own = (uint32_t)(color * SZ + j);
      // the same entry of the other colour's array, then its four neighbours
      const int b = (int)own + (color ? -SZ : SZ);
      const int ns0 = b - HC, ns1 = b + HC, we0 = b + p - 1, we1 = b + p;
      const int t0 = hedge ? ns0 : we0, t1 = hedge ? ns1 : we1;
      const int t2 = hedge ? we0 : ns0, t3 = hedge ? we1 : ns1;
      SmoothCell& e = cell[color][k];
      e.a0 = (live && (m & 1u)) ? (uint32_t)t0 : (uint32_t)ZERO;
      e.a1 = (live && (m & 2u)) ? (uint32_t)t1 : (uint32_t)ZERO;
      e.a2 = (live && (m & 4u)) ? (uint32_t)t2 : (uint32_t)ZERO;
      e.a3 = (live && (m & 8u)) ? (uint32_t)t3 : (uint32_t)ZERO;
      e.out = live ? own : (uint32_t)SINK + own;
      const uint32_t n = __builtin_popcount(m & 15u);
      e.inv = n == 2 ? 0.5 : (n == 3 ? 1.0 / 3.0 : (n == 4 ? 0.25 : 1.0));
    }
  }
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int i = (int)threadIdx.x + k * T;
    const int lr = i / CC, lc = i % CC;
    sv[((r0 + c0 + lr + lc) & 1) * SZ + lr * HC + (lc >> 1)] = v[k];
  }
  if (threadIdx.x == 0) sv[ZERO] = 0.f;
  __syncthreads();
  uint32_t s = 0;
  if ((step0 & 1u) && s < nsteps) {
    smooth_half_step(sv, cell[1]);
    __syncthreads();
    ++s;
  }
  for (; s + 2 <= nsteps; s += 2) {
    smooth_half_step(sv, cell[0]);
    __syncthreads();
    smooth_half_step(sv, cell[1]);
    __syncthreads();
  }
  if (s < nsteps) {
    smooth_half_step(sv, cell[0]);
    __syncthreads();
  }
  for (int i = threadIdx.x; i < TR * TC; i += T) {
    const int lr = K + i / TC, lc = K + i % TC;
    const int r = r0 + lr, c = c0 + lc;
    if (r < R && c < C) dst[r * C + c] = sv[((r + c) & 1) * SZ + lr * HC + (lc >> 1)];
  }
}

static __global__ __launch_bounds__(256) void k_copy_f32(const float* __restrict__ a,
                                                        float* __restrict__ b, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Rust's `f32 as i32`: NaN -> 0, saturating
__device__ __forceinline__ int f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483647.f) return 2147483647;
  if (x <= -2147483648.f) return -2147483647 - 1;
  return (int)x;
}

__device__ __forceinline__ float clamp01(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

// One thread per vertex: x-velocity bilinear from the vedges (sampled at
// (iw + 1, ih + 0.5)), y-velocity from the hedges (at (iw + 0.5, ih + 1)).
// Bilinear read of the staggered pixel-edge grids at point (px, py):
// (x velocity from the vedges, y velocity from the hedges),
// rasterized_edge_gradient.rs:202-248.
__device__ __forceinline__ void interp_staggered(uint32_t W, uint32_t H,
                                                 const float* __restrict__ hvy,
                                                 const float* __restrict__ vvx, float px, float py,
                                                 float& vx, float& vy) {
  {
    const float gx = px - 1.0f, gy = py - 0.5f;
    const int ix0 = clampi(f2i_sat(floorf(gx)), 0, (int)W - 2);
    const int iy0 = clampi(f2i_sat(floorf(gy)), 0, (int)H - 2);
    const int ix1 = ix0 + 1 < (int)W - 2 ? ix0 + 1 : (int)W - 2;
    const int iy1 = iy0 + 1 < (int)H - 1 ? iy0 + 1 : (int)H - 1;
    const float tx = clamp01(gx - (float)ix0), ty = clamp01(gy - (float)iy0);
    const int w = (int)W - 1;
    vx = (((1.f - tx) * (1.f - ty) * vvx[iy0 * w + ix0] + tx * (1.f - ty) * vvx[iy0 * w + ix1]) +
          (1.f - tx) * ty * vvx[iy1 * w + ix0]) +
         tx * ty * vvx[iy1 * w + ix1];
  }
  {
    const float gx = px - 0.5f, gy = py - 1.0f;
    const int ix0 = clampi(f2i_sat(floorf(gx)), 0, (int)W - 2);
    const int iy0 = clampi(f2i_sat(floorf(gy)), 0, (int)H - 2);
    const int ix1 = ix0 + 1 < (int)W - 1 ? ix0 + 1 : (int)W - 1;
    const int iy1 = iy0 + 1 < (int)H - 2 ? iy0 + 1 : (int)H - 2;
    const float tx = clamp01(gx - (float)ix0), ty = clamp01(gy - (float)iy0);
    vy = (((1.f - tx) * (1.f - ty) * hvy[iy0 * W + ix0] + tx * (1.f - ty) * hvy[iy0 * W + ix1]) +
          (1.f - tx) * ty * hvy[iy1 * W + ix0]) +
         tx * ty * hvy[iy1 * W + ix1];
  }
}

static __global__ __launch_bounds__(256) void k_interpolate_staggered(
    uint32_t W, uint32_t H, const float* __restrict__ hvy, const float* __restrict__ vvx,
    uint32_t num_vtx, const float* __restrict__ vtx2xy, float* __restrict__ vtx2velo) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  float vx, vy;
  interp_staggered(W, H, hvy, vvx, vtx2xy[(size_t)i * 2], vtx2xy[(size_t)i * 2 + 1], vx, vy);
  vtx2velo[(size_t)i * 2] = vx;
  vtx2velo[(size_t)i * 2 + 1] = vy;
}

// The tail of the smoothed edge-gradient backward, per vertex (the Python
// chain of RasterizedEdgeGradient/torch.py:237-274 in one pass): q = M [x, 1]
// (M = world2pix, column-major T), u = (q0, q1) / q3, g = dL/du read from the
// smoothed pixel-edge grids at u, dL/dq = (g / q3, 0, -(g . u) / q3),
// dL/dx = dL/dq M[:, :3].
static __global__ __launch_bounds__(256) void k_smoothed_vertex_grad(
    uint32_t W, uint32_t H, const float* __restrict__ hvy, const float* __restrict__ vvx,
    uint32_t num_vtx, const float* __restrict__ xyz, const float* __restrict__ T,
    float* __restrict__ dldx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  const float x0 = xyz[(size_t)i * 3], x1 = xyz[(size_t)i * 3 + 1], x2 = xyz[(size_t)i * 3 + 2];
  float q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = ((T[k] * x0 + T[4 + k] * x1) + T[8 + k] * x2) + T[12 + k];
  const float inv_w = 1.f / q[3];
  const float u0 = q[0] * inv_w, u1 = q[1] * inv_w;
  float g0, g1;
  interp_staggered(W, H, hvy, vvx, u0, u1, g0, g1);
  const float d0 = g0 * inv_w, d1 = g1 * inv_w, d3 = -(g0 * u0 + g1 * u1) * inv_w;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    dldx[(size_t)i * 3 + j] = (d0 * T[j * 4] + d1 * T[j * 4 + 1]) + d3 * T[j * 4 + 3];
}

}  // namespace eg
}  // namespace dmsh
