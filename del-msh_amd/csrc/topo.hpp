// topo.hpp -- mesh topology on the device: elements around each vertex,
// vertices around each vertex, edges from those, the two triangles of each
// edge, and the graph Laplacian (Jacobi smoothing, product).
//
// Layouts are the reference's offset arrays: vtx2idx[num_vtx + 1] (exclusive
// prefix of the per-vertex counts) over a flat idx2* array.
//
//   vtx2elem  the (vertex, flat slot) pairs of elem2vtx are sorted by vertex
//             with the stable radix argsort, so each vertex lists its elements
//             in ascending order exactly as the serial counting pass of
//             del-msh-cpu/src/vtx2elem.rs:58-96 does; the offsets come from
//             the boundaries of the sorted keys.
//   vtx2vtx   one thread per vertex walks its elements (ascending) and nodes
//             and keeps first appearances (vtx2vtx.rs:11-76), into a
//             per-vertex slice of the workspace; a scan of the counts gives
//             the offsets, and a second kernel packs the lists.
//   edges     edge2vtx_from_vtx2vtx (edge2vtx.rs:16-36), one thread per entry;
//             edge2elem (edge2elem.rs:20-70), one thread per edge.
//   Laplacian vtx2vtx.rs:210-280, one thread per vertex, f32 in the CPU's
//             order: rhs + sum of lambda * neighbour, times 1 / (1 + lambda
//             valence).
#pragma once

#include "dmsh_common.hpp"
#include "sort.hpp"

namespace dmsh {
namespace topo {

// ---------------------------------------------------------------------------
// exclusive scan of n u32 counts into out[0..n] (out[n] = total), 3 launches
// ---------------------------------------------------------------------------
constexpr uint32_t kScanTile = 1024;  // 256 threads x 4

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += u;
  }
  return v;
}

static __global__ __launch_bounds__(256) void k_scan_tile_sums(const uint32_t* __restrict__ in,
                                                               uint32_t n,
                                                               uint32_t* __restrict__ sums) {
  __shared__ uint32_t s_w[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4u;
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) v += (base + k < n) ? in[base + k] : 0u;
  v = wave_incl_scan(v);
  if ((threadIdx.x & 63u) == 63u) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// one workgroup: tile sums -> exclusive tile offsets; out[n] = total
static __global__ __launch_bounds__(1024) void k_scan_tile_offsets(uint32_t* __restrict__ sums,
                                                                  uint32_t num_tiles,
                                                                  uint32_t* __restrict__ total) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < num_tiles; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < num_tiles ? sums[i] : 0u;
    const uint32_t inc = wave_incl_scan(v);
    if ((threadIdx.x & 63u) == 63u) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t pre = s_carry;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) pre += s_w[w];
    if (i < num_tiles) sums[i] = pre + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = pre + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = s_carry;
}

static __global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in,
                                                           uint32_t n,
                                                           const uint32_t* __restrict__ offs,
                                                           uint32_t* __restrict__ out) {
  __shared__ uint32_t s_w[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4u;
  uint32_t v[4], t = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    t += v[k];
  }
  const uint32_t inc = wave_incl_scan(t);
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 63u) s_w[wave] = inc;
  __syncthreads();
  uint32_t pre = offs[blockIdx.x] + inc - t;
  for (uint32_t w = 0; w < wave; ++w) pre += s_w[w];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += v[k];
  }
}

inline uint32_t scan_tiles(uint32_t n) { return n ? ceil_div(n, kScanTile) : 1u; }

// `in` and `out` may alias.  `sums` holds scan_tiles(n) words.
inline int launch_exclusive_scan(hipStream_t s, uint32_t n, const uint32_t* in, uint32_t* out,
                                 uint32_t* sums) {
  const uint32_t nt = scan_tiles(n);
  if (n == 0) return sortk::launch_fill(s, out, 1, 0u);
  launch("k_scan_tile_sums", k_scan_tile_sums, dim3(nt), dim3(256), 0, s, in, n, sums);
  int rc = launch_status();
  if (rc != DMSH_OK) return rc;
  launch("k_scan_tile_offsets", k_scan_tile_offsets, dim3(1), dim3(1024), 0, s, sums, nt, out + n);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_scan_apply", k_scan_apply, dim3(nt), dim3(256), 0, s, in, n, sums, out);
  return launch_status();
}

// ---------------------------------------------------------------------------
// vertex -> elements
// ---------------------------------------------------------------------------
// keys = elem2vtx (flat); an index out of range raises and is clamped so the
// sort and the offsets stay in bounds
static __global__ __launch_bounds__(256) void k_vtx_keys(const uint32_t* __restrict__ elem2vtx,
                                                         uint32_t nflat, uint32_t num_vtx,
                                                         uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ err) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nflat;
       k += gridDim.x * blockDim.x) {
    uint32_t v = elem2vtx[k];
    if (v >= num_vtx) {
      raise_error(err, DMSH_DEVERR_INDEX);
      v = num_vtx - 1;
    }
    keys[k] = v;
  }
}

// vtx2idx[v] = first sorted position whose key is >= v (v <= num_vtx)
static __global__ __launch_bounds__(256) void k_vtx2elem_offsets(const uint32_t* __restrict__ keys,
                                                                 uint32_t nflat, uint32_t num_vtx,
                                                                 uint32_t* __restrict__ vtx2idx) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nflat) return;
  const uint32_t lo = k == 0 ? 0u : keys[k - 1] + 1u;
  const uint32_t hi = k < nflat ? keys[k] : num_vtx;
  for (uint32_t v = lo; v <= hi; ++v) vtx2idx[v] = k;
}

static __global__ __launch_bounds__(256) void k_idx2elem(const uint32_t* __restrict__ slot,
                                                         uint32_t nflat, uint32_t num_node,
                                                         uint32_t* __restrict__ idx2elem) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nflat;
       k += gridDim.x * blockDim.x)
    idx2elem[k] = slot[k] / num_node;
}

// ---------------------------------------------------------------------------
// vertex -> vertices
// ---------------------------------------------------------------------------
// Vertex i's unique neighbours go to buff[vtx2jdx[i] * m ...] (m = num_node,
// or num_node - 1 without the vertex itself: the most one element adds).
// One wave per vertex (four per workgroup of 256).  The candidates, element
// jdx2elem[jdx] node k for jdx in [j0, j1) and k in [0, num_node), are taken
// 64 at a time in that order, one per lane.  A lane's candidate is new if it
// is not the vertex itself (unless is_self), not in the list so far and not
// held by a lower lane of the same chunk; new ones are appended in lane order,
// which keeps the list in order of first appearance.  The first kV2VLocal
// entries of the list are also kept in LDS, where the lanes' membership scans
// read them as broadcasts; longer lists are scanned in buff beyond that.
constexpr uint32_t kV2VLocal = 1024;
constexpr uint32_t kV2VWaves = 4;

static __global__ __launch_bounds__(256) void k_vtx2vtx_count(
    uint32_t num_vtx, const uint32_t* __restrict__ elem2vtx, uint32_t num_elem, uint32_t num_node,
    uint32_t is_self, const uint32_t* __restrict__ vtx2jdx, const uint32_t* __restrict__ jdx2elem,
    uint32_t* __restrict__ buff, uint32_t* __restrict__ counts, uint32_t* __restrict__ err) {
  __shared__ uint32_t loc[kV2VWaves * kV2VLocal];
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const uint32_t i = blockIdx.x * kV2VWaves + w;
  if (i >= num_vtx) return;
  uint32_t* sl = loc + w * kV2VLocal;
  const uint32_t m = is_self ? num_node : num_node - 1u;
  const uint32_t j0 = vtx2jdx[i], j1 = vtx2jdx[i + 1];
  const uint32_t cap = (j1 - j0) * m;
  const uint32_t ncand = (j1 - j0) * num_node;
  uint32_t* b = buff + (size_t)j0 * m;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < ncand; base += 64) {
    const uint32_t c = base + lane;
    bool valid = c < ncand;
    uint32_t cand = 0xffffffffu;
    if (valid) {
      const uint32_t e = jdx2elem[j0 + c / num_node];
      if (e >= num_elem) {
        raise_error(err, DMSH_DEVERR_INDEX);
        valid = false;
      } else {
        cand = elem2vtx[(size_t)e * num_node + c % num_node];
        valid = is_self || cand != i;
      }
    }
    bool dup = false;
    const uint32_t nl = cnt < kV2VLocal ? cnt : kV2VLocal;
    for (uint32_t q = 0; q < nl; ++q) dup |= sl[q] == cand;
    if (cnt > kV2VLocal) {
      __threadfence();  // the wave's own earlier stores to b, seen by all its lanes
      for (uint32_t q = kV2VLocal; q < cnt; ++q) dup |= b[q] == cand;
    }
    const uint32_t nc = ncand - base < 64u ? ncand - base : 64u;
    for (uint32_t l = 0; l + 1 < nc; ++l) {
      const uint32_t x = (uint32_t)__shfl((int)cand, (int)l);
      dup |= l < lane && x == cand;
    }
    // an invalid lane holds 0xffffffff, which a valid lane equals only for an
    // out-of-range index, one the vertex-key pass has already flagged
    const bool fresh = valid && !dup;
    const uint64_t mask = __ballot(fresh);
    const uint32_t pos = cnt + (uint32_t)__popcll(mask & below);
    if (fresh) {
      if (pos >= cap) {  // only an element holding i more than its nodes allow
        raise_error(err, DMSH_DEVERR_INDEX);
      } else {
        if (pos < kV2VLocal) sl[pos] = cand;
        b[pos] = cand;
      }
    }
    cnt += (uint32_t)__popcll(mask);
    if (cnt > cap) cnt = cap;
  }
  if (lane == 0) counts[i] = cnt;
}

static __global__ __launch_bounds__(256) void k_vtx2vtx_fill(uint32_t num_vtx, uint32_t m,
                                                             const uint32_t* __restrict__ vtx2jdx,
                                                             const uint32_t* __restrict__ buff,
                                                             const uint32_t* __restrict__ vtx2idx,
                                                             uint32_t* __restrict__ idx2vtx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  const uint32_t* b = buff + (size_t)vtx2jdx[i] * m;
  const uint32_t i0 = vtx2idx[i], i1 = vtx2idx[i + 1];
  for (uint32_t idx = i0; idx < i1; ++idx) idx2vtx[idx] = b[idx - i0];
}

// ---------------------------------------------------------------------------
// edges
// ---------------------------------------------------------------------------
// the vertex whose slice holds entry idx: the last i with vtx2idx[i] <= idx
__device__ __forceinline__ uint32_t owner_of(const uint32_t* __restrict__ vtx2idx,
                                             uint32_t num_vtx, uint32_t idx) {
  uint32_t lo = 0, hi = num_vtx;  // vtx2idx[lo] <= idx < vtx2idx[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (vtx2idx[mid] <= idx) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Entry idx of vertex i -> edge (i, idx2vtx[idx]).  The offsets must be
// non-decreasing with vtx2idx[0] = 0 and vtx2idx[num_vtx] = num_idx (checked).
static __global__ __launch_bounds__(256) void k_edge2vtx(uint32_t num_vtx,
                                                         const uint32_t* __restrict__ vtx2idx,
                                                         uint32_t num_idx,
                                                         const uint32_t* __restrict__ idx2vtx,
                                                         uint32_t* __restrict__ edge2vtx,
                                                         uint32_t* __restrict__ err) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= num_idx) return;
  if (vtx2idx[0] != 0u || vtx2idx[num_vtx] != num_idx) {
    if (idx == 0) raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t i = owner_of(vtx2idx, num_vtx, idx);
  edge2vtx[(size_t)idx * 2] = i;
  edge2vtx[(size_t)idx * 2 + 1] = idx2vtx[idx];
}

static __global__ __launch_bounds__(256) void k_edge2elem(
    uint32_t num_edge, const uint32_t* __restrict__ edge2vtx, uint32_t num_tri,
    const uint32_t* __restrict__ tri2vtx, uint32_t num_vtx, const uint32_t* __restrict__ vtx2idx,
    uint32_t num_idx, const uint32_t* __restrict__ idx2tri, uint32_t* __restrict__ edge2tri,
    uint32_t* __restrict__ err) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= num_edge) return;
  const uint32_t i0 = edge2vtx[(size_t)e * 2], i1 = edge2vtx[(size_t)e * 2 + 1];
  if (i0 >= num_vtx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const uint32_t a = vtx2idx[i0], b = vtx2idx[i0 + 1];
  if (a > b || b > num_idx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  uint32_t cnt = 0;
  for (uint32_t idx = a; idx < b && cnt < 2u; ++idx) {
    const uint32_t t = idx2tri[idx];
    if (t >= num_tri) {
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    const uint32_t j0 = tri2vtx[(size_t)t * 3], j1 = tri2vtx[(size_t)t * 3 + 1],
                   j2 = tri2vtx[(size_t)t * 3 + 2];
    const uint32_t hits = (i0 == j0) + (i0 == j1) + (i0 == j2);
    if (hits != 1u) {  // the listed triangle must hold i0 exactly once
      raise_error(err, DMSH_DEVERR_INDEX);
      continue;
    }
    const bool adjacent = (i0 == j0) ? (j1 == i1 || j2 == i1)
                        : (i0 == j1) ? (j2 == i1 || j0 == i1)
                                     : (j0 == i1 || j1 == i1);
    if (!adjacent) continue;
    edge2tri[(size_t)e * 2 + cnt] = t;
    ++cnt;
  }
}

// ---------------------------------------------------------------------------
// graph Laplacian
// ---------------------------------------------------------------------------
// next = (rhs + sum_j lambda prev_j) * (1 / (1 + lambda valence))
static __global__ __launch_bounds__(256) void k_laplacian_jacobi(
    uint32_t num_vtx, const uint32_t* __restrict__ vtx2idx, uint32_t num_idx,
    const uint32_t* __restrict__ idx2vtx, float lambda, uint32_t num_dim,
    const float* __restrict__ prev, float* __restrict__ next, const float* __restrict__ rhs,
    uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  const uint32_t a = vtx2idx[i], b = vtx2idx[i + 1];
  if (a > b || b > num_idx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const float valence = (float)(b - a);
  const float inv_dia = 1.f / (1.f + lambda * valence);
  for (uint32_t d = 0; d < num_dim; ++d) {
    float acc = rhs[(size_t)i * num_dim + d];
    for (uint32_t idx = a; idx < b; ++idx) {
      const uint32_t j = idx2vtx[idx];
      if (j >= num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
        continue;
      }
      acc += lambda * prev[(size_t)j * num_dim + d];
    }
    next[(size_t)i * num_dim + d] = acc * inv_dia;
  }
}

// lhs = valence rhs_i - sum_j rhs_j
static __global__ __launch_bounds__(256) void k_graph_laplacian(
    uint32_t num_vtx, const uint32_t* __restrict__ vtx2idx, uint32_t num_idx,
    const uint32_t* __restrict__ idx2vtx, uint32_t num_dim, const float* __restrict__ rhs,
    float* __restrict__ lhs, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_vtx) return;
  const uint32_t a = vtx2idx[i], b = vtx2idx[i + 1];
  if (a > b || b > num_idx) {
    raise_error(err, DMSH_DEVERR_INDEX);
    return;
  }
  const float valence = (float)(b - a);
  for (uint32_t d = 0; d < num_dim; ++d) {
    float acc = valence * rhs[(size_t)i * num_dim + d];
    for (uint32_t idx = a; idx < b; ++idx) {
      const uint32_t j = idx2vtx[idx];
      if (j >= num_vtx) {
        raise_error(err, DMSH_DEVERR_INDEX);
        continue;
      }
      acc -= rhs[(size_t)j * num_dim + d];
    }
    lhs[(size_t)i * num_dim + d] = acc;
  }
}

}  // namespace topo
}  // namespace dmsh
