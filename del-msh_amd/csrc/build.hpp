// build.hpp -- device-resident LBVH build of a triangle mesh or a point set.
//
//   init        bounding-box keys, histogram / ticket / look-back / refit
//               counters (one fill kernel: no memset nodes, graph safe)
//   centroid    triangle centroids ((0 + p0 + p1 + p2) * (1/3), as
//               del-msh-cpu/src/elem2center.rs:8-37) and the bounding box of
//               the centroids, reduced through order-preserving integer keys so
//               the box does not depend on reduction order; also each
//               triangle's box (the refit's leaf box, in triangle order, so the
//               refit reads one 24-byte record per leaf instead of gathering
//               the indices and three vertices)
//   morton      unit-cube fit of that box (Mat44/torch.py:120-140 formula),
//               Morton code per element, and the digit histograms of the
//               radix passes (the sort's histogram read is fused in here)
//   sort        onesweep digit passes (sort.hpp), values = element index
//   topology    Karras tree (topology.hpp)
//   refit       bottom-up boxes (refit.hpp)
#pragma once

#include "morton.hpp"
#include "refit.hpp"
#include "sort.hpp"
#include "topology.hpp"

namespace dmsh {
namespace bld {

// The centroid box is reduced into kBoxCopies copies of (min keys, max keys),
// one per 256-byte line (block b takes copy b mod kBoxCopies): device-scope
// atomics on one word are performed one after the other, ~14 ns apart, so
// 2048 blocks on one set of six words would queue for tens of microseconds.
// Min and max over order keys do not depend on order, so the copies'
// reduction gives the same box.
constexpr uint32_t kBoxCopies = 64;
constexpr uint32_t kBoxStride = 64;  // words between copies
static_assert(kBoxCopies == 64, "k_morton_hist reduces the copies with one wave");

// keys6[c*kBoxStride + 0..2] = min order keys (init 0xFFFFFFFF),
// keys6[c*kBoxStride + 3..5] = max (init 0) for each copy c, then
// `zero_count` words starting at `zero` are cleared.
static __global__ __launch_bounds__(256) void k_build_init(uint32_t* __restrict__ keys6,
                                                    uint32_t* __restrict__ zero,
                                                    size_t zero_count) {
  const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i0 < 6 * kBoxCopies) keys6[(i0 / 6) * kBoxStride + i0 % 6] = (i0 % 6 < 3) ? 0xFFFFFFFFu : 0u;
  for (size_t i = i0; i < zero_count; i += (size_t)gridDim.x * blockDim.x) zero[i] = 0u;
}

constexpr int kCentroidItems = 4;  // elements whose loads a thread issues together

template <bool TRI>
__global__ __launch_bounds__(256) void k_centroid_aabb(uint32_t n,
                                                       const uint32_t* __restrict__ tri2vtx,
                                                       uint32_t num_vtx,
                                                       const float* __restrict__ xyz,
                                                       float* __restrict__ cntr,
                                                       float* __restrict__ tri_box,
                                                       uint32_t* __restrict__ keys6,
                                                       uint32_t* __restrict__ err) {
  __shared__ uint32_t s_red[4][6];
  uint32_t kmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, kmax[3] = {0u, 0u, 0u};
  // grid-stride: a bounded number of blocks keeps the 6 global atomics cheap;
  // each round issues the loads of kCentroidItems elements before using them
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n;
       i0 += stride * kCentroidItems) {
    uint32_t v[kCentroidItems][3];
#pragma unroll
    for (int u = 0; u < kCentroidItems; ++u) {
      const uint32_t i = i0 + (uint32_t)u * stride;
      v[u][0] = v[u][1] = v[u][2] = 0u;
      if (TRI && i < n) {
        const uint32_t* tv = tri2vtx + (size_t)i * 3;
        v[u][0] = tv[0];
        v[u][1] = tv[1];
        v[u][2] = tv[2];
      }
    }
    float x[kCentroidItems][9];
#pragma unroll
    for (int u = 0; u < kCentroidItems; ++u) {
      const uint32_t i = i0 + (uint32_t)u * stride;
      if (i >= n) continue;
      if (TRI) {
        if (v[u][0] >= num_vtx || v[u][1] >= num_vtx || v[u][2] >= num_vtx) {
          raise_error(err, DMSH_DEVERR_INDEX);
          v[u][0] = v[u][1] = v[u][2] = 0;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int d = 0; d < 3; ++d) x[u][3 * k + d] = xyz[(size_t)v[u][k] * 3 + d];
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) x[u][d] = xyz[(size_t)i * 3 + d];
      }
    }
#pragma unroll
    for (int u = 0; u < kCentroidItems; ++u) {
      const uint32_t i = i0 + (uint32_t)u * stride;
      if (i >= n) continue;
      float c[3];
      if (TRI) {
        const float ratio = 1.f / 3.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float cog = 0.f;
          cog += x[u][d];
          cog += x[u][3 + d];
          cog += x[u][6 + d];
          c[d] = cog * ratio;
          cntr[(size_t)i * 3 + d] = c[d];
        }
        Box b;
        box_set(b, &x[u][0], 0.f);
        box_add(b, &x[u][3], 0.f);
        box_add(b, &x[u][6], 0.f);
        box_store(tri_box, i, b);
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) c[d] = x[u][d];
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const uint32_t key = f32_order_key(c[d]);
        kmin[d] = key < kmin[d] ? key : kmin[d];
        kmax[d] = key > kmax[d] ? key : kmax[d];
      }
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint32_t a = __shfl_xor(kmin[d], off, 64), b = __shfl_xor(kmax[d], off, 64);
      kmin[d] = a < kmin[d] ? a : kmin[d];
      kmax[d] = b > kmax[d] ? b : kmax[d];
    }
  }
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      s_red[wave][d] = kmin[d];
      s_red[wave][d + 3] = kmax[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const uint32_t d = threadIdx.x;
    const uint32_t nw = blockDim.x >> 6;
    uint32_t v = s_red[0][d];
    for (uint32_t w = 1; w < nw; ++w) {
      const uint32_t u = s_red[w][d];
      v = (d < 3) ? (u < v ? u : v) : (u > v ? u : v);
    }
    uint32_t* box = keys6 + (blockIdx.x % kBoxCopies) * kBoxStride;
    if (d < 3) atomicMin(box + d, v);
    else atomicMax(box + d, v);
  }
}

// Column-major unit-cube fit of the reduced box: s = 1 / max extent,
// translation -center * s + 0.5.
__device__ __forceinline__ void fit_transform(const uint32_t* __restrict__ keys6, float T[16]) {
  float mn[3], mx[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    mn[d] = f32_from_order_key(keys6[d]);
    mx[d] = f32_from_order_key(keys6[d + 3]);
  }
  float ext = mx[0] - mn[0];
  ext = fmax_sel(ext, mx[1] - mn[1]);
  ext = fmax_sel(ext, mx[2] - mn[2]);
  const float s = 1.f / ext;
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = 0.f;
  T[0] = s;
  T[5] = s;
  T[10] = s;
  T[15] = 1.f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = (mn[d] + mx[d]) * 0.5f;
    T[12 + d] = (-c) * s + 0.5f;
  }
}

template <typename KeyT>
__global__ __launch_bounds__(256) void k_morton_hist(uint32_t n, const float* __restrict__ pts,
                                                     const uint32_t* __restrict__ keys6,
                                                     KeyT* __restrict__ codes, uint32_t first,
                                                     uint32_t num_passes, uint32_t key_bits,
                                                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[sortk::kMaxPasses][sortk::kRadix];
  __shared__ uint32_t s_k6[6];
  for (uint32_t e = threadIdx.x; e < sortk::kMaxPasses * sortk::kRadix; e += blockDim.x)
    (&s_h[0][0])[e] = 0;
  if (threadIdx.x < 64) {  // lane c reads copy c; the wave reduces them
    uint32_t k6[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) k6[d] = keys6[threadIdx.x * kBoxStride + d];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        const uint32_t v = __shfl_xor(k6[d], off, 64);
        k6[d] = d < 3 ? (v < k6[d] ? v : k6[d]) : (v > k6[d] ? v : k6[d]);
      }
    }
    if (threadIdx.x == 0) {
#pragma unroll
      for (int d = 0; d < 6; ++d) s_k6[d] = k6[d];
    }
  }
  __syncthreads();
  float T[16];
  {
    uint32_t k6[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) k6[d] = s_k6[d];
    fit_transform(k6, T);
  }
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n;
       i0 += stride * kCentroidItems) {
    float x[kCentroidItems][3];
#pragma unroll
    for (int u = 0; u < kCentroidItems; ++u) {
      const uint32_t i = i0 + (uint32_t)u * stride;
#pragma unroll
      for (int d = 0; d < 3; ++d) x[u][d] = i < n ? pts[(size_t)i * 3 + d] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kCentroidItems; ++u) {
      const uint32_t i = i0 + (uint32_t)u * stride;
      if (i >= n) continue;
      const KeyT code = morton3_transformed<KeyT>(T, x[u][0], x[u][1], x[u][2]);
      codes[i] = code;
      sortk::hist_accumulate(s_h, code, first, num_passes, key_bits);
    }
  }
  __syncthreads();
  sortk::hist_flush(s_h, first, num_passes, hist);
}

template <typename KeyT>
inline uint32_t key_bits_of() {
  return sizeof(KeyT) == 4 ? 30u : 63u;
}

struct BuildLayout {
  uint32_t* keys6;
  sortk::Layout sort;
  uint32_t* flags;
  uint32_t* treelet_count;
  uint4* treelets;
  uint32_t* top_count;
  uint32_t* treelet_next;
  uint4* top_list;
  uint32_t* node2slot;
  uint32_t* top_kid;
  float* top_box;
  float* cntr;
  float* tri_box;
  void* codes;
  uint32_t* idx;
  uint32_t* zero_begin;
  size_t zero_words;
};

template <typename KeyT>
inline bool carve(Arena& a, uint32_t n, bool tri, BuildLayout& L) {
  const uint32_t kb = key_bits_of<KeyT>();
  L.keys6 = a.take<uint32_t>(kBoxCopies * kBoxStride);
  // hist, tickets, status and the refit counters are one contiguous zero run
  L.sort.hist = a.take<uint32_t>(sortk::kMaxPasses * sortk::kRadix + sortk::kTicketWords);
  L.sort.tickets = L.sort.hist + sortk::kMaxPasses * sortk::kRadix;
  const size_t status_words = sortk::status_words<KeyT>(n, kb);
  L.sort.status = L.sort.tickets + sortk::kTicketWords;
  a.off += sizeof(uint32_t) * status_words;  // contiguous after the tickets
  L.flags = L.sort.status + status_words;
  a.off += sizeof(uint32_t) * (n > 1 ? n - 1 : 1);
  L.treelet_count = L.flags + (n > 1 ? n - 1 : 1);
  L.top_count = L.treelet_count + 1;
  L.treelet_next = L.top_count + 1;
  a.off += 3 * sizeof(uint32_t);
  L.zero_begin = L.sort.hist;
  L.zero_words = (size_t)(L.treelet_next - L.sort.hist) + 1;
  L.sort.zero_words = 0;
  L.treelets = a.take<uint4>(n);
  L.top_list = a.take<uint4>(kTopSlots);
  L.node2slot = a.take<uint32_t>(n);
  L.top_kid = a.take<uint32_t>(kTopClimbSlots);
  L.top_box = a.take<float>((size_t)kTopClimbSlots * 12);
  L.cntr = tri ? a.take<float>((size_t)n * 3) : nullptr;
  L.tri_box = tri ? a.take<float>((size_t)n * 6) : nullptr;
  L.codes = a.take<KeyT>(n);
  L.idx = a.take<uint32_t>(n);
  L.sort.alt_keys = a.take<KeyT>(n);
  L.sort.alt_vals = a.take<uint32_t>(n);
  L.sort.groups = a.take<uint32_t>(sortk::group_capacity(n));
  return a.ok();
}

template <typename KeyT>
inline size_t workspace_bytes(uint32_t n) {
  Arena a{nullptr, 0, 0};
  BuildLayout L;
  carve<KeyT>(a, n, true, L);
  return a.off + 256;
}

template <typename KeyT, bool TRI>
inline int build(hipStream_t s, uint32_t n, const uint32_t* tri2vtx, uint32_t num_vtx,
                 const float* xyz, uint32_t* bvhnodes, float* aabb, void* ws, size_t ws_bytes) {
  if (n == 0) return DMSH_OK;
  if (!xyz || !bvhnodes || !aabb || (TRI && !tri2vtx)) return DMSH_ERR_INVALID;
  if (n >= (1u << 30)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < workspace_bytes<KeyT>(n)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  BuildLayout L;
  if (!carve<KeyT>(a, n, TRI, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  const uint32_t kb = key_bits_of<KeyT>();
  const uint32_t P = sortk::num_passes(kb);
  int rc;
  launch("k_build_init", k_build_init, dim3(std::max<uint32_t>(2u, std::min<uint32_t>(ceil_div(L.zero_words, 256), 2048))),
                     dim3(256), 0, s, L.keys6, L.zero_begin, L.zero_words);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_centroid_aabb", k_centroid_aabb<TRI>, dim3(std::min<uint32_t>(ceil_div(n, 256), 2048)),
                     dim3(256), 0, s, n, tri2vtx, TRI ? num_vtx : n, xyz, L.cntr, L.tri_box,
                     L.keys6, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  KeyT* codes = static_cast<KeyT*>(L.codes);
  const uint32_t mh_blocks = std::min<uint32_t>(sortk::num_tiles<KeyT>(n), 2048);
  launch("k_morton_hist", k_morton_hist<KeyT>, dim3(mh_blocks), dim3(256), 0, s, n,
         TRI ? L.cntr : xyz, L.keys6, codes, sortk::first_hist_pass(P), P, kb, L.sort.hist);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  // P is even for 30- and 63-bit codes: the passes end back in `codes`.
  rc = sortk::run_passes<KeyT>(s, n, kb, L.sort, codes, codes, L.idx, err, true);
  if (rc != DMSH_OK) return rc;
  // the topology pass also writes every leaf's box, so the treelet pass
  // loads boxes instead of gathering them
  const bool leaves_in_aabb = refit_option(kRefitLeavesInAabb) == 1;
  TreeletSink ts{kTreeletLeaves, L.treelet_count, L.treelets, n, nullptr, nullptr,
                 L.top_count, L.top_list, (uint32_t)kTopSlots, L.node2slot};
  ts.stage_pos = true;
  if (leaves_in_aabb) {
    ts.leaf_aabb = aabb;
    ts.tri_box = TRI ? L.tri_box : nullptr;
    ts.xyz = xyz;
  }
  rc = launch_topology<KeyT>(s, n, bvhnodes, codes, L.idx, ts);
  if (rc != DMSH_OK) return rc;
  const TreeView v{aabb, bvhnodes, n, tri2vtx, num_vtx, xyz, L.tri_box};
  return launch_refit_treelets<TRI ? kLeafTri : kLeafPoint>(
      s, v, L.flags, L.treelet_count, L.treelets, 0.f, err, L.top_count, L.top_list, L.node2slot,
      L.treelet_next, nullptr, leaves_in_aabb, L.top_kid, L.top_box);
}

}  // namespace bld
}  // namespace dmsh
