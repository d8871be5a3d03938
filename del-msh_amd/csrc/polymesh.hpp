// polymesh.hpp -- mixed polyhedron meshes in CSR form (elem2idx_offset,
// idx2vtx): element centroids, point location through the element BVH and
// shape-function interpolation.
//
// Point location follows del-msh-cpu/src/polyhedron_mesh.rs:470-575: a
// depth-first walk of the BVH, left child first, into boxes that contain the
// point (min <= q <= max on every axis), stopping at the first element that
// contains it.  The walk needs no stack: it climbs through the parent links
// of the reference layout (after a left child comes the right one, after a
// right child the parent's sibling), so it visits the nodes of the
// recursion in the recursion's order at any tree depth.
//
// Parametric coordinates of the point in an element (del-geo's
// tet/pyramid/prism/hex parametric_coord_for_origin is not vendored; this is
// the written-down rule the oracle shares).  Vertices are first shifted by
// -q, so the point is the origin.
//   tet (4):     barycentric coordinates by volume ratios, b_k = det of the
//                tet with vertex k replaced by the origin / det of the tet;
//                inside iff b0, b1, b2 and 1 - b0 - b1 - b2 are all > 0
//                (polyhedron_mesh.rs:89-100); returns (b0, b1, b2).
//   pyramid (5), prism (6), hex (8): Newton's method on
//                x(r,s,t) = sum_i N_i(r,s,t) v_i = 0 with the shape
//                functions below, from (1/2, 1/2, 1/4), (1/3, 1/3, 1/2) and
//                (0, 0, 0), each step J d = -x solved by Cramer's rule,
//                stopping when max |d| < 1e-6 (at most 20 steps; a zero or
//                non-finite det, or no convergence, is a miss); inside with
//                a margin of 1e-5 in parameter space: pyramid r, s, t in
//                [0, 1]; prism r, s >= 0, r + s <= 1, t in [0, 1]; hex
//                r, s, t in [-1, 1].
// The shape functions and their operation order are those of the reference
// GPU interpolation kernel (polyhedron_mesh.cu:112-177):
//   pyramid N = ((1-r)(1-s))(1-t), (r(1-s))(1-t), (rs)(1-t), ((1-r)s)(1-t), t
//   prism   N = ((1-r)-s)(1-t), r(1-t), s(1-t), ((1-r)-s)t, rt, st
//   hex     N_i = ((1/8 (1 +- r))(1 +- s))(1 +- t), nodes (-,-,-) (+,-,-)
//           (+,+,-) (-,+,-) (-,-,+) (+,-,+) (+,+,+) (-,+,+).
#pragma once

#include "dmsh_common.hpp"
#include "morton.hpp"

namespace dmsh {
namespace poly {

struct V3 {
  float x, y, z;
};

__device__ __forceinline__ V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 neg3(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// a . (b x c)
__device__ __forceinline__ float det3(V3 a, V3 b, V3 c) {
  const float cx = b.y * c.z - b.z * c.y;
  const float cy = b.z * c.x - b.x * c.z;
  const float cz = b.x * c.y - b.y * c.x;
  return (a.x * cx + a.y * cy) + a.z * cz;
}

constexpr int kNewtonSteps = 20;
constexpr float kNewtonTol = 1e-6f;
constexpr float kInsideTol = 1e-5f;

__device__ __forceinline__ bool tet_param(const V3* p, float pco[3]) {
  const V3 e1 = sub3(p[1], p[0]), e2 = sub3(p[2], p[0]), e3 = sub3(p[3], p[0]);
  const float vol = det3(e1, e2, e3);
  if (vol == 0.f) return false;
  const V3 o = neg3(p[0]);
  const float b0 = det3(p[1], p[2], p[3]) / vol;
  const float b1 = det3(o, e2, e3) / vol;
  const float b2 = det3(e1, o, e3) / vol;
  if (!(b0 > 0.f && b1 > 0.f && b2 > 0.f && ((1.f - b0) - b1) - b2 > 0.f)) return false;
  pco[0] = b0;
  pco[1] = b1;
  pco[2] = b2;
  return true;
}

// shape functions (n nodes) and their r, s, t derivatives at (r, s, t)
__device__ __forceinline__ void shape_pyramid(float r, float s, float t, float* N, float* dr,
                                              float* ds, float* dt) {
  const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
  N[0] = (rm * sm) * tm;
  N[1] = (r * sm) * tm;
  N[2] = (r * s) * tm;
  N[3] = (rm * s) * tm;
  N[4] = t;
  if (!dr) return;
  dr[0] = -(sm * tm);
  dr[1] = sm * tm;
  dr[2] = s * tm;
  dr[3] = -(s * tm);
  dr[4] = 0.f;
  ds[0] = -(rm * tm);
  ds[1] = -(r * tm);
  ds[2] = r * tm;
  ds[3] = rm * tm;
  ds[4] = 0.f;
  dt[0] = -(rm * sm);
  dt[1] = -(r * sm);
  dt[2] = -(r * s);
  dt[3] = -(rm * s);
  dt[4] = 1.f;
}

__device__ __forceinline__ void shape_prism(float r, float s, float t, float* N, float* dr,
                                            float* ds, float* dt) {
  const float rs = (1.f - r) - s, tm = 1.f - t;
  N[0] = rs * tm;
  N[1] = r * tm;
  N[2] = s * tm;
  N[3] = rs * t;
  N[4] = r * t;
  N[5] = s * t;
  if (!dr) return;
  dr[0] = -tm;
  dr[1] = tm;
  dr[2] = 0.f;
  dr[3] = -t;
  dr[4] = t;
  dr[5] = 0.f;
  ds[0] = -tm;
  ds[1] = 0.f;
  ds[2] = tm;
  ds[3] = -t;
  ds[4] = 0.f;
  ds[5] = t;
  dt[0] = -rs;
  dt[1] = -r;
  dt[2] = -s;
  dt[3] = rs;
  dt[4] = r;
  dt[5] = s;
}

__device__ __forceinline__ void shape_hex(float r, float s, float t, float* N, float* dr,
                                          float* ds, float* dt) {
  const float e = 0.125f;
  const float sr[8] = {-1.f, 1.f, 1.f, -1.f, -1.f, 1.f, 1.f, -1.f};
  const float ss[8] = {-1.f, -1.f, 1.f, 1.f, -1.f, -1.f, 1.f, 1.f};
  const float st[8] = {-1.f, -1.f, -1.f, -1.f, 1.f, 1.f, 1.f, 1.f};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float fr = sr[i] > 0.f ? 1.f + r : 1.f - r;
    const float fs = ss[i] > 0.f ? 1.f + s : 1.f - s;
    const float ft = st[i] > 0.f ? 1.f + t : 1.f - t;
    N[i] = ((e * fr) * fs) * ft;
    if (dr) {
      dr[i] = ((e * sr[i]) * fs) * ft;
      ds[i] = ((e * fr) * ss[i]) * ft;
      dt[i] = ((e * fr) * fs) * st[i];
    }
  }
}

template <int NN>
__device__ __forceinline__ void shape(float r, float s, float t, float* N, float* dr, float* ds,
                                      float* dt) {
  if (NN == 5) shape_pyramid(r, s, t, N, dr, ds, dt);
  else if (NN == 6) shape_prism(r, s, t, N, dr, ds, dt);
  else shape_hex(r, s, t, N, dr, ds, dt);
}

template <int NN>
__device__ __forceinline__ bool newton_param(const V3* p, float r, float s, float t,
                                             float pco[3]) {
  bool converged = false;
  for (int it = 0; it < kNewtonSteps; ++it) {
    float N[NN], dr[NN], ds[NN], dt[NN];
    shape<NN>(r, s, t, N, dr, ds, dt);
    V3 x{0.f, 0.f, 0.f}, jr{0.f, 0.f, 0.f}, js{0.f, 0.f, 0.f}, jt{0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < NN; ++i) {
      x = V3{x.x + N[i] * p[i].x, x.y + N[i] * p[i].y, x.z + N[i] * p[i].z};
      jr = V3{jr.x + dr[i] * p[i].x, jr.y + dr[i] * p[i].y, jr.z + dr[i] * p[i].z};
      js = V3{js.x + ds[i] * p[i].x, js.y + ds[i] * p[i].y, js.z + ds[i] * p[i].z};
      jt = V3{jt.x + dt[i] * p[i].x, jt.y + dt[i] * p[i].y, jt.z + dt[i] * p[i].z};
    }
    const float d = det3(jr, js, jt);
    if (!(d != 0.f) || !(fabsf(d) < __builtin_inff())) return false;
    const V3 f = neg3(x);
    const float dr0 = det3(f, js, jt) / d;
    const float ds0 = det3(jr, f, jt) / d;
    const float dt0 = det3(jr, js, f) / d;
    r = r + dr0;
    s = s + ds0;
    t = t + dt0;
    const float m = fmaxf(fmaxf(fabsf(dr0), fabsf(ds0)), fabsf(dt0));
    if (!(m == m)) return false;
    if (m < kNewtonTol) {
      converged = true;
      break;
    }
  }
  if (!converged) return false;
  const float lo = -kInsideTol, hi = 1.f + kInsideTol;
  bool inside;
  if (NN == 5) {
    inside = r >= lo && r <= hi && s >= lo && s <= hi && t >= lo && t <= hi;
  } else if (NN == 6) {
    inside = r >= lo && s >= lo && r + s <= hi && t >= lo && t <= hi;
  } else {
    inside = r >= -hi && r <= hi && s >= -hi && s <= hi && t >= -hi && t <= hi;
  }
  if (!inside) return false;
  pco[0] = r;
  pco[1] = s;
  pco[2] = t;
  return true;
}

// p: the element's vertices minus the query point
__device__ __forceinline__ bool param_coord(const V3* p, uint32_t n, float pco[3]) {
  switch (n) {
    case 4: return tet_param(p, pco);
    case 5: return newton_param<5>(p, 0.5f, 0.5f, 0.25f, pco);
    case 6: return newton_param<6>(p, 1.f / 3.f, 1.f / 3.f, 0.5f, pco);
    case 8: return newton_param<8>(p, 0.f, 0.f, 0.f, pco);
    default: return false;
  }
}

// elem2center (elem2center.rs:59-89 / elem2center.cu): the mean of the
// element's vertices, summed in order and scaled by 1 / n (0 for n == 0)
static __global__ __launch_bounds__(256) void k_elem2center(uint32_t num_elem,
                                                            const uint32_t* __restrict__ elem2idx,
                                                            uint32_t num_idx,
                                                            const uint32_t* __restrict__ idx2vtx,
                                                            uint32_t num_vtx, uint32_t num_dim,
                                                            const float* __restrict__ xyz,
                                                            float* __restrict__ cog,
                                                            uint32_t* __restrict__ err) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= num_elem) return;
  const uint32_t i0 = elem2idx[e], i1 = elem2idx[e + 1];
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  if (i0 > i1 || i1 > num_idx) {
    raise_error(err, DMSH_DEVERR_INDEX);
  } else {
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t v = idx2vtx[i];
      if (v >= num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
        continue;
      }
      for (uint32_t d = 0; d < num_dim; ++d) acc[d] += xyz[(size_t)v * num_dim + d];
    }
  }
  const uint32_t n = i1 > i0 ? i1 - i0 : 0u;
  const float ratio = n == 0 ? 0.f : 1.f / (float)n;
  for (uint32_t d = 0; d < num_dim; ++d) cog[(size_t)e * num_dim + d] = acc[d] * ratio;
}

struct MeshView {
  uint32_t num_elem;
  const uint32_t* elem2idx;
  uint32_t num_idx;
  const uint32_t* idx2vtx;
  uint32_t num_vtx;
  const float* xyz;
};

// 30-bit Morton code of each query point in the unit-cube fit of the root
// box (aabb row 0), and the identity permutation: sorting by these codes puts
// nearby queries in one wave, so its lanes walk the same branches and test
// the same elements.  Only the processing order changes, not any result.
static __global__ __launch_bounds__(256) void k_query_codes(uint32_t num_wtx,
                                                            const float* __restrict__ wtx2xyz,
                                                            const float* __restrict__ aabb,
                                                            uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ idx) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= num_wtx) return;
  float ext = fmax_sel(fmax_sel(aabb[3] - aabb[0], aabb[4] - aabb[1]), aabb[5] - aabb[2]);
  const float s = ext > 0.f ? 1.f / ext : 1.f;
  float T[16] = {s, 0.f, 0.f, 0.f, 0.f, s, 0.f, 0.f, 0.f, 0.f, s, 0.f,
                 -aabb[0] * s, -aabb[1] * s, -aabb[2] * s, 1.f};
  keys[w] = morton3_transformed<uint32_t>(T, wtx2xyz[(size_t)w * 3], wtx2xyz[(size_t)w * 3 + 1],
                                          wtx2xyz[(size_t)w * 3 + 2]);
  idx[w] = w;
}

// one thread per query point; with `order`, thread i takes query order[i]
static __global__ __launch_bounds__(256) void k_search_elem_contain_points(
    MeshView m, const uint32_t* __restrict__ bvhnodes, const float* __restrict__ aabb,
    uint32_t num_wtx, const float* __restrict__ wtx2xyz, uint32_t* __restrict__ wtx2elem,
    float* __restrict__ wtx2param, const uint32_t* __restrict__ order,
    uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_wtx) return;
  const uint32_t w = order ? order[i] : i;
  if (w >= num_wtx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const V3 q{wtx2xyz[(size_t)w * 3], wtx2xyz[(size_t)w * 3 + 1], wtx2xyz[(size_t)w * 3 + 2]};
  uint32_t found = DMSH_U32MAX;
  float pco[3] = {0.f, 0.f, 0.f};
  const uint32_t num_node = m.num_elem ? 2 * m.num_elem - 1 : 0;
  uint32_t node = 0;
  bool down = num_node > 0;
  const uint64_t max_steps = 2ull * num_node + 2;
  for (uint64_t steps = 0; num_node > 0; ++steps) {
    if (steps > max_steps || node >= num_node) {
      raise_error(err, DMSH_DEVERR_INDEX);
      break;
    }
    const uint32_t* row = bvhnodes + (size_t)node * 3;
    if (down) {
      const float* b = aabb + (size_t)node * 6;
      const bool in = b[0] <= q.x && q.x <= b[3] && b[1] <= q.y && q.y <= b[4] && b[2] <= q.z &&
                      q.z <= b[5];
      if (in) {
        if (row[2] == DMSH_U32MAX) {
          const uint32_t e = row[1];
          if (e >= m.num_elem) {
            raise_error(err, DMSH_DEVERR_INDEX);
            break;
          }
          const uint32_t i0 = m.elem2idx[e], i1 = m.elem2idx[e + 1];
          const uint32_t n = i1 - i0;
          if (i1 < i0 || i1 > m.num_idx) {
            raise_error(err, DMSH_DEVERR_INDEX);
            break;
          }
          if (n == 4 || n == 5 || n == 6 || n == 8) {
            // all eight slots by constant index (registers, not scratch)
            V3 p[8];
            bool ok = true;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
              p[k] = V3{0.f, 0.f, 0.f};
              if (k < n) {
                const uint32_t v = m.idx2vtx[i0 + k];
                if (v >= m.num_vtx) {
                  ok = false;
                } else {
                  const float* x = m.xyz + (size_t)v * 3;
                  p[k] = V3{x[0] - q.x, x[1] - q.y, x[2] - q.z};
                }
              }
            }
            if (!ok) {
              raise_error(err, DMSH_DEVERR_INDEX);
              break;
            }
            float c[3];
            if (param_coord(p, n, c)) {
              found = e;
              pco[0] = c[0];
              pco[1] = c[1];
              pco[2] = c[2];
              break;
            }
          }
          down = false;  // leaf done: climb
          continue;
        }
        node = row[1];  // left child first
        continue;
      }
      down = false;
      continue;
    }
    // climbing from `node`
    const uint32_t par = row[0];
    if (par == DMSH_U32MAX) break;  // back at the root: not found
    if (par >= num_node) {
      raise_error(err, DMSH_DEVERR_INDEX);
      break;
    }
    const uint32_t* prow = bvhnodes + (size_t)par * 3;
    if (prow[1] == node) {
      node = prow[2];
      down = true;
    } else {
      node = par;
    }
  }
  wtx2elem[w] = found;
  wtx2param[(size_t)w * 3] = pco[0];
  wtx2param[(size_t)w * 3 + 1] = pco[1];
  wtx2param[(size_t)w * 3 + 2] = pco[2];
}

// wtx2val[w] = sum_i N_i(param) vtx2val[node i] (0 outside the mesh or for
// other element sizes); polyhedron_mesh.cu:112-177
static __global__ __launch_bounds__(256) void k_interpolate_at_points(
    MeshView m, uint32_t num_wtx, uint32_t num_vdim, const float* __restrict__ vtx2val,
    const uint32_t* __restrict__ wtx2elem, const float* __restrict__ wtx2param,
    float* __restrict__ out, uint32_t* __restrict__ err) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= num_wtx) return;
  float* o = out + (size_t)w * num_vdim;
  for (uint32_t d = 0; d < num_vdim; ++d) o[d] = 0.f;
  const uint32_t e = wtx2elem[w];
  if (e == DMSH_U32MAX) return;
  if (e >= m.num_elem) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t i0 = m.elem2idx[e], i1 = m.elem2idx[e + 1];
  if (i1 < i0 || i1 > m.num_idx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t n = i1 - i0;
  const float r = wtx2param[(size_t)w * 3], s = wtx2param[(size_t)w * 3 + 1],
              t = wtx2param[(size_t)w * 3 + 2];
  float sf[8];
  switch (n) {
    case 4:
      sf[0] = r;
      sf[1] = s;
      sf[2] = t;
      sf[3] = ((1.f - r) - s) - t;
      break;
    case 5: shape_pyramid(r, s, t, sf, nullptr, nullptr, nullptr); break;
    case 6: shape_prism(r, s, t, sf, nullptr, nullptr, nullptr); break;
    case 8: shape_hex(r, s, t, sf, nullptr, nullptr, nullptr); break;
    default: return;
  }
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t v = m.idx2vtx[i0 + k];
    if (v >= m.num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    const float* val = vtx2val + (size_t)v * num_vdim;
    for (uint32_t d = 0; d < num_vdim; ++d) o[d] += sf[k] * val[d];
  }
}

}  // namespace poly
}  // namespace dmsh
