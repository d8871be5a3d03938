// batch.hpp -- LBVH build and ray-cast of many meshes of uneven size, each
// stage one launch over all meshes (CSR offsets), no per-mesh launches.
//
// Meshes are concatenated: triangles of mesh m are [mesh2tri[m], mesh2tri[m+1])
// with vertex indices local to the mesh's vertices [mesh2vtx[m], mesh2vtx[m+1]).
// Each mesh's tree is exactly the tree the single-mesh build gives it (30-bit
// codes): per-mesh centroid boxes and unit-cube fits, then one stable radix
// sort of composite keys (mesh << 30 | code) -- stable by (mesh, code, index),
// i.e. every mesh's segment is its own stable sort -- then the Karras
// topology and the treelet refit run on each segment with local indices.
// Output rows of mesh m start at node_off[m] = sum_{j<m} (2 N_j - 1).
//
// The ray-cast hands 8x8 pixel tiles of all meshes to a persistent grid of
// waves through per-XCD counters (each XCD starts on its own contiguous run of
// tiles and steals from the others when it runs out), so meshes of every size
// share the chip evenly and neighbouring tiles stay in one L2.
#pragma once

#include "build.hpp"
#include "raycast.hpp"

namespace dmsh {
namespace bat {

constexpr uint32_t kMeshKeyShift = 30;
constexpr uint32_t kXcds = 8;

// The segment holding element x (x < off[M]): the last m with off[m] <= x,
// which skips empty segments (repeated offsets).
__device__ __forceinline__ uint32_t find_segment(const uint32_t* __restrict__ off, uint32_t M,
                                                 uint32_t x) {
  uint32_t lo = 0, hi = M;  // off[lo] <= x < off[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Exclusive scan over meshes of value(m) into out[0..M] (out[M] = total), by
// one workgroup of 1024 threads.
template <class F>
__device__ __forceinline__ void scan_meshes(uint32_t M, uint32_t* __restrict__ out, F value) {
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base <= M; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < M ? value(i) : 0u;
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
    if (i <= M) out[i] = pre + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = pre + inc;
    __syncthreads();
  }
}

__device__ __forceinline__ uint32_t seg_len(const uint32_t* __restrict__ off, uint32_t m) {
  return off[m + 1] - off[m];
}

// An offset array fits its data: first entry 0, non-decreasing, last entry
// `total`.  Otherwise *bad is set (every later kernel of the call then does
// nothing) and DMSH_DEVERR_INDEX is raised.
__device__ __forceinline__ void check_offsets(const uint32_t* __restrict__ off, uint32_t M,
                                              uint32_t total, uint32_t* __restrict__ bad,
                                              uint32_t* __restrict__ err) {
  bool ok = true;
  for (uint32_t m = threadIdx.x; m < M; m += blockDim.x) ok = ok && off[m + 1] >= off[m];
  if (threadIdx.x == 0) ok = ok && off[0] == 0u && off[M] == total;
  if (!ok) {
    atomicOr(bad, 1u);
    raise_error(err, DMSH_DEVERR_INDEX);
  }
}

// inner_off[m] = sum max(N_j - 1, 0); node_off[m] = sum max(2 N_j - 1, 0).
static __global__ __launch_bounds__(1024) void k_batch_offsets(
    uint32_t M, uint32_t total_tri, uint32_t total_vtx, const uint32_t* __restrict__ mesh2tri,
    const uint32_t* __restrict__ mesh2vtx, uint32_t* __restrict__ inner_off,
    uint32_t* __restrict__ node_off, uint32_t* __restrict__ bad, uint32_t* __restrict__ err) {
  check_offsets(mesh2tri, M, total_tri, bad, err);
  check_offsets(mesh2vtx, M, total_vtx, bad, err);
  scan_meshes(M, inner_off, [&](uint32_t m) {
    const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
    return n > 0 ? n - 1 : 0u;
  });
  if (node_off) {
    scan_meshes(M, node_off, [&](uint32_t m) {
      const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
      return n > 0 ? 2 * n - 1 : 0u;
    });
  }
}

// A mesh's centroid box is accumulated in kBoxCopies copies of its six
// order keys, each on its own 128-byte line (a wave adds into copy wave %
// kBoxCopies), so that the waves of a large mesh do not queue their device
// atomics on one line; k_batch_box_reduce then folds the copies into keys6.
constexpr uint32_t kBoxCopies = 16;
constexpr uint32_t kBoxCopyWords = 32;

// The box copies of every mesh (min keys 0xFFFFFFFF, max keys 0), then zero
// words.
static __global__ __launch_bounds__(256) void k_batch_init(uint32_t M, uint32_t* __restrict__ copies,
                                                           uint32_t* __restrict__ zero,
                                                           size_t zero_count) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  for (size_t i = i0; i < (size_t)M * kBoxCopies * 6; i += stride)
    copies[(i / 6) * kBoxCopyWords + i % 6] = (i % 6 < 3) ? 0xFFFFFFFFu : 0u;
  for (size_t i = i0; i < zero_count; i += stride) zero[i] = 0u;
}

// keys6[m][d] = the fold of mesh m's box copies.
static __global__ __launch_bounds__(256) void k_batch_box_reduce(
    const uint32_t* __restrict__ bad, uint32_t M, const uint32_t* __restrict__ copies,
    uint32_t* __restrict__ keys6) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (*bad || i >= M * 6u) return;
  const uint32_t m = i / 6u, d = i % 6u;
  const uint32_t* c = copies + (size_t)m * kBoxCopies * kBoxCopyWords + d;
  uint32_t k = c[0];
#pragma unroll
  for (uint32_t j = 1; j < kBoxCopies; ++j) {
    const uint32_t x = c[j * kBoxCopyWords];
    k = d < 3 ? (x < k ? x : k) : (x > k ? x : k);
  }
  keys6[i] = k;
}

constexpr uint32_t kCentroidChunks = 16;  // 64-triangle chunks per wave

__device__ __forceinline__ void flush_box(uint32_t* __restrict__ copies, uint32_t m,
                                          uint32_t copy, const uint32_t (&kmin)[3],
                                          const uint32_t (&kmax)[3]) {
  uint32_t* c = copies + ((size_t)m * kBoxCopies + copy) * kBoxCopyWords;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    atomicMin(c + d, kmin[d]);
    atomicMax(c + 3 + d, kmax[d]);
  }
}

// Triangle centroids ((0 + p0 + p1 + p2) * (1/3)) and per-mesh centroid
// boxes through order keys.  Each wave takes kCentroidChunks consecutive
// chunks of 64 triangles.  Triangles are ordered by mesh, so the lanes of a
// chunk form runs of one mesh; a segmented min/max over the lanes leaves each
// run's box in its first lane.  The chunk's last run is carried into the next
// chunk (it may continue there), every other run, and the carry where the
// next chunk starts with another mesh or the wave ends, goes out as six
// atomics.  So the atomics per mesh are bounded by the waves that touch it.
static __global__ __launch_bounds__(256) void k_batch_centroid(
    const uint32_t* __restrict__ bad, uint32_t total_tri, uint32_t M,
    const uint32_t* __restrict__ mesh2tri, const uint32_t* __restrict__ mesh2vtx,
    const uint32_t* __restrict__ tri2vtx, const float* __restrict__ xyz, float* __restrict__ cntr,
    uint32_t* __restrict__ copies, uint32_t* __restrict__ err) {
  if (*bad) return;
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t copy = wave % kBoxCopies;
  const uint32_t w0 = wave * 64u * kCentroidChunks;
  if (w0 >= total_tri) return;
  uint32_t m = find_segment(mesh2tri, M, w0 + lane < total_tri ? w0 + lane : total_tri - 1);
  uint32_t carry = kNone;  // wave-uniform: the mesh of the carried run
  uint32_t cmin[3] = {kNone, kNone, kNone}, cmax[3] = {0u, 0u, 0u};
  for (uint32_t ch = 0; ch < kCentroidChunks; ++ch) {
    const uint32_t base = w0 + ch * 64u;
    if (base >= total_tri) break;
    const uint32_t t = base + lane;
    const bool valid = t < total_tri;
    uint32_t mm = kNone;
    uint32_t kmin[3] = {kNone, kNone, kNone}, kmax[3] = {0u, 0u, 0u};
    if (valid) {
      while (mesh2tri[m + 1] <= t) ++m;
      mm = m;
      const uint32_t vb = mesh2vtx[m], nv = mesh2vtx[m + 1] - vb;
      const uint32_t* tv = tri2vtx + (size_t)t * 3;
      const uint32_t v[3] = {tv[0], tv[1], tv[2]};
      float c[3] = {0.f, 0.f, 0.f};
      if (v[0] >= nv || v[1] >= nv || v[2] >= nv) {
        raise_error(err, DMSH_DEVERR_INDEX);
      } else {
        const float ratio = 1.f / 3.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float cog = 0.f;
          cog += xyz[((size_t)vb + v[0]) * 3 + d];
          cog += xyz[((size_t)vb + v[1]) * 3 + d];
          cog += xyz[((size_t)vb + v[2]) * 3 + d];
          c[d] = cog * ratio;
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cntr[(size_t)t * 3 + d] = c[d];
        kmin[d] = kmax[d] = f32_order_key(c[d]);
      }
    }
    // segmented reduction towards each run's first lane (invalid lanes, at
    // the end, form a run of kNone)
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t om = (uint32_t)__shfl_down((int)mm, off, 64);
      const bool same = lane + off < 64 && om == mm;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const uint32_t a = (uint32_t)__shfl_down((int)kmin[d], off, 64);
        const uint32_t b = (uint32_t)__shfl_down((int)kmax[d], off, 64);
        if (same) {
          kmin[d] = a < kmin[d] ? a : kmin[d];
          kmax[d] = b > kmax[d] ? b : kmax[d];
        }
      }
    }
    const uint32_t prev = (uint32_t)__shfl_up((int)mm, 1, 64);
    const bool head = mm != kNone && (lane == 0 || prev != mm);
    const uint64_t heads = __ballot(head);
    const uint32_t last = 63u - (uint32_t)__clzll(heads);  // first lane of the last run
    const uint32_t first_m = (uint32_t)__shfl((int)mm, 0, 64);
    if (carry != kNone && carry == first_m) {  // the carried run continues here
      if (lane == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          kmin[d] = cmin[d] < kmin[d] ? cmin[d] : kmin[d];
          kmax[d] = cmax[d] > kmax[d] ? cmax[d] : kmax[d];
        }
      }
    } else if (carry != kNone && lane == 0) {
      flush_box(copies, carry, copy, cmin, cmax);
    }
    if (head && lane != last) flush_box(copies, mm, copy, kmin, kmax);
    carry = (uint32_t)__shfl((int)mm, (int)last, 64);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cmin[d] = (uint32_t)__shfl((int)kmin[d], (int)last, 64);
      cmax[d] = (uint32_t)__shfl((int)kmax[d], (int)last, 64);
    }
  }
  if (carry != kNone && lane == 0) flush_box(copies, carry, copy, cmin, cmax);
}

// Composite key (mesh << 30 | 30-bit code under the mesh's unit-cube fit) and
// the digit histograms of every sort pass.
static __global__ __launch_bounds__(256) void k_batch_morton_hist(
    const uint32_t* __restrict__ bad, uint32_t total_tri, uint32_t M,
    const uint32_t* __restrict__ mesh2tri, const float* __restrict__ cntr,
    const uint32_t* __restrict__ keys6, uint64_t* __restrict__ codes, uint32_t num_passes,
    uint32_t key_bits, uint32_t* __restrict__ hist) {
  if (*bad) return;  // the sort then orders whatever the keys hold, in bounds
  __shared__ uint32_t s_h[sortk::kMaxPasses][sortk::kRadix];
  for (uint32_t e = threadIdx.x; e < sortk::kMaxPasses * sortk::kRadix; e += blockDim.x)
    (&s_h[0][0])[e] = 0;
  __syncthreads();
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total_tri;
       t += gridDim.x * blockDim.x) {
    const uint32_t m = find_segment(mesh2tri, M, t);
    float T[16];
    bld::fit_transform(keys6 + (size_t)m * 6, T);
    const float* p = cntr + (size_t)t * 3;
    const uint32_t code = morton3_transformed<uint32_t>(T, p[0], p[1], p[2]);
    const uint64_t key = ((uint64_t)m << kMeshKeyShift) | code;
    codes[t] = key;
    sortk::hist_accumulate(s_h, key, 0u, num_passes, key_bits);
  }
  __syncthreads();
  sortk::hist_flush(s_h, 0u, num_passes, hist);
}

// The segmented sort of the composite keys: the triangles are ordered by
// mesh, so a mesh's keys keep its range and only the 30-bit code needs
// ordering: three passes over code bits [6, 30) scatter inside the meshes,
// the finish orders the low 6 bits.  Up to kSegMaxMeshes meshes (the offset
// tables take 2 x 3 KB per mesh).
constexpr uint32_t kSegLow = 6;          // code bits the finish orders
constexpr uint32_t kSegMaxMeshes = 8192;
constexpr uint32_t kSegWindow = 4;       // meshes a block counts in LDS
constexpr uint32_t kSegTables = sortk::kTopPasses * sortk::kRadix;  // words per mesh

// Composite keys (as k_batch_morton_hist) and per mesh the histograms of the
// segmented passes' digits, hseg[p][m][d] for code bits [6 + 8p, 14 + 8p).
// Each block takes one contiguous range of triangles in rounds of 256, so it
// meets the meshes in order: a window of kSegWindow consecutive meshes counts
// in LDS and is flushed to hseg when a round reaches past it (then it starts
// at the round's first mesh); a round spanning more meshes than the window
// (meshes of fewer than 64 triangles) counts the rest with global atomics.
static __global__ __launch_bounds__(256) void k_batch_morton_seghist(
    const uint32_t* __restrict__ bad, uint32_t total_tri, uint32_t M,
    const uint32_t* __restrict__ mesh2tri, const float* __restrict__ cntr,
    const uint32_t* __restrict__ keys6, uint64_t* __restrict__ codes,
    uint32_t* __restrict__ hseg) {
  if (*bad) return;  // the sort then orders whatever the keys hold, in bounds
  constexpr uint32_t kP = sortk::kTopPasses, kR = sortk::kRadix, kW = kSegWindow * kP * kR;
  __shared__ uint32_t s_h[kW];
  const uint32_t per = ceil_div_u32(ceil_div_u32(total_tri, gridDim.x), 256u) * 256u;
  const uint32_t t0 = blockIdx.x * per;
  if (t0 >= total_tri) return;  // uniform over the block
  const uint32_t t1 = total_tri - t0 < per ? total_tri : t0 + per;
  for (uint32_t e = threadIdx.x; e < kW; e += blockDim.x) s_h[e] = 0u;
  uint32_t wbase = find_segment(mesh2tri, M, t0);  // the window's first mesh (uniform)
  auto flush = [&]() {
    for (uint32_t e = threadIdx.x; e < kW; e += blockDim.x) {
      const uint32_t v = s_h[e];
      const uint32_t w = e / (kP * kR), q = (e / kR) % kP, d = e % kR;
      if (v && wbase + w < M) atomicAdd(hseg + ((size_t)q * M + wbase + w) * kR + d, v);
      s_h[e] = 0u;
    }
  };
  __syncthreads();
  for (uint32_t r0 = t0; r0 < t1; r0 += blockDim.x) {
    const uint32_t rl = t1 - r0 < blockDim.x ? t1 - 1 : r0 + blockDim.x - 1;
    const uint32_t mlast = find_segment(mesh2tri, M, rl);  // uniform
    if (mlast - wbase >= kSegWindow) {
      __syncthreads();  // the window's counts are complete
      flush();
      wbase = find_segment(mesh2tri, M, r0);
      __syncthreads();  // zeroed before the round counts
    }
    const uint32_t t = r0 + threadIdx.x;
    if (t < t1) {
      const uint32_t m = find_segment(mesh2tri, M, t);
      float T[16];
      bld::fit_transform(keys6 + (size_t)m * 6, T);
      const float* p = cntr + (size_t)t * 3;
      const uint32_t code = morton3_transformed<uint32_t>(T, p[0], p[1], p[2]);
      codes[t] = ((uint64_t)m << kMeshKeyShift) | code;
      const uint32_t w = m - wbase;
#pragma unroll
      for (uint32_t q = 0; q < kP; ++q) {
        const uint32_t d = (code >> (kSegLow + q * sortk::kRadixBits)) & (kR - 1u);
        if (w < kSegWindow) atomicAdd(&s_h[(w * kP + q) * kR + d], 1u);
        else atomicAdd(hseg + ((size_t)q * M + m) * kR + d, 1u);
      }
    }
  }
  __syncthreads();
  flush();
}

// Scatter offsets of the segmented passes (sortk::SegOff), per pass p, mesh
// m and digit d: off[p][m][d] = mesh2tri[m] + (triangles of mesh m with a
// smaller digit) - (triangles of meshes before m with digit d).  Meshes in
// chunks of kSegChunk, one block per (chunk, pass), a wave per kPerWave
// meshes with four digits per lane and every load issued first.
// k_batch_seg_rows counts the last term only over the meshes of its own
// chunk (the waves before from LDS, the wave's own earlier meshes from
// registers) and stores the chunk's per-digit totals; with more than one
// chunk, k_batch_seg_cols then subtracts the totals of the chunks before.
// After a data error every offset is zero (the sort then stays in bounds).
constexpr uint32_t kSegChunk = 64;
constexpr uint32_t kSegRowThreads = 1024;

static __global__ __launch_bounds__(kSegRowThreads) void k_batch_seg_rows(
    const uint32_t* __restrict__ bad, uint32_t M, const uint32_t* __restrict__ mesh2tri,
    const uint32_t* __restrict__ hseg, uint32_t* __restrict__ off,
    uint32_t* __restrict__ chunk_sum) {
  constexpr uint32_t kR = sortk::kRadix, kWaves = kSegRowThreads / 64;
  constexpr uint32_t kPerWave = kSegChunk / kWaves;
  __shared__ uint32_t s_cs[kWaves][kR];
  const uint32_t c = blockIdx.x, p = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const bool ok = *bad == 0u;
  const uint32_t* H = hseg + (size_t)p * M * kR;
  uint32_t* O = off + (size_t)p * M * kR;
  const uint32_t mw = c * kSegChunk + wave * kPerWave;  // the wave's first mesh
  uint32_t h[kPerWave][4];
#pragma unroll
  for (uint32_t i = 0; i < kPerWave; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      h[i][j] = (ok && mw + i < M) ? H[(size_t)(mw + i) * kR + 4 * lane + j] : 0u;
  }
  // per mesh: mesh2tri[m] + smaller digits of the mesh - this wave's earlier
  // meshes with the digit
  uint32_t r[kPerWave][4], cs[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (uint32_t i = 0; i < kPerWave; ++i) {
    const uint32_t t = h[i][0] + h[i][1] + h[i][2] + h[i][3];
    uint32_t run = sortk::wave_inclusive_scan(t) - t + ((ok && mw + i < M) ? mesh2tri[mw + i] : 0u);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      r[i][j] = run - cs[j];
      run += h[i][j];
      cs[j] += h[i][j];
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) s_cs[wave][4 * lane + j] = cs[j];
  __syncthreads();
  uint32_t before[4] = {0u, 0u, 0u, 0u};  // the chunk's earlier waves
  for (uint32_t w = 0; w < wave; ++w) {
#pragma unroll
    for (int j = 0; j < 4; ++j) before[j] += s_cs[w][4 * lane + j];
  }
#pragma unroll
  for (uint32_t i = 0; i < kPerWave; ++i) {
    if (mw + i < M) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        O[(size_t)(mw + i) * kR + 4 * lane + j] = ok ? r[i][j] - before[j] : 0u;
    }
  }
  const uint32_t d = threadIdx.x;
  if (gridDim.x > 1 && d < kR) {
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) sum += s_cs[w][d];
    chunk_sum[((size_t)p * gridDim.x + c) * kR + d] = sum;
  }
}

static __global__ __launch_bounds__(256) void k_batch_seg_cols(
    const uint32_t* __restrict__ bad, uint32_t M, uint32_t* __restrict__ off,
    const uint32_t* __restrict__ chunk_sum) {
  constexpr uint32_t kR = sortk::kRadix;
  const uint32_t c = blockIdx.x, p = blockIdx.y, d = threadIdx.x;
  if (*bad || c == 0) return;  // the offsets stay zero / need nothing before
  const uint32_t* S = chunk_sum + (size_t)p * gridDim.x * kR;
  uint32_t col = 0;  // triangles with digit d in the chunks before
  for (uint32_t k = 0; k < c; ++k) col += S[(size_t)k * kR + d];
  uint32_t* O = off + (size_t)p * M * kR;
  const uint32_t m0 = c * kSegChunk;
  const uint32_t cnt = M - m0 < kSegChunk ? M - m0 : kSegChunk;
  uint32_t o[kSegChunk];
#pragma unroll
  for (uint32_t i = 0; i < kSegChunk; ++i) o[i] = i < cnt ? O[(size_t)(m0 + i) * kR + d] : 0u;
#pragma unroll
  for (uint32_t i = 0; i < kSegChunk; ++i)
    if (i < cnt) O[(size_t)(m0 + i) * kR + d] = o[i] - col;
}

// The codes of one mesh (index 0 = the mesh's first) through a window over
// all meshes' codes.
template <typename KeyT>
struct MeshCodes {
  CodeWindow<KeyT> w;
  uint32_t tb;
  __device__ MeshCodes(const CodeWindow<KeyT>& w_, uint32_t tb_) : w(w_), tb(tb_) {}
  __device__ __forceinline__ KeyT operator[](uint32_t j) const { return w[tb + j]; }
};

// Internal node g of all meshes (g < inner_off[M]), a thread each.  Node g
// of mesh m searches the codes around mesh2tri[m] + (g - inner_off[m]), so a
// block's nodes search one nearly contiguous run of codes: that run and
// kTopoHalo codes either side sit in LDS (kWin codes at most; any other
// code is read from memory), as in k_topology.
static __global__ __launch_bounds__(kTopoThreads) void k_batch_topology(
    const uint32_t* __restrict__ bad, uint32_t total_tri, uint32_t M,
    const uint32_t* __restrict__ mesh2tri, const uint32_t* __restrict__ inner_off,
    const uint32_t* __restrict__ node_off, uint32_t* __restrict__ bvhnodes,
    const uint64_t* __restrict__ codes, const uint32_t* __restrict__ idx, TreeletSink ts) {
  constexpr uint32_t kWin = 2 * kTopoHalo + 2 * kTopoThreads + 1;
  __shared__ uint64_t s_c[kWin];
  if (*bad) return;
  const uint32_t total_inner = inner_off[M];
  const uint32_t base = blockIdx.x * kTopoThreads;
  if (base >= total_inner || base >= total_tri) return;  // uniform over the block
  const uint32_t last_g = (total_inner - base < kTopoThreads ? total_inner : base + kTopoThreads) - 1u;
  const uint32_t m_first = find_segment(inner_off, M, base);
  const uint32_t m_last = find_segment(inner_off, M, last_g);
  const uint32_t a_lo = base + mesh2tri[m_first] - inner_off[m_first];
  const uint32_t a_hi = last_g + mesh2tri[m_last] - inner_off[m_last];
  const uint32_t lo = a_lo > kTopoHalo ? a_lo - kTopoHalo : 0u;
  uint32_t hi = total_tri - a_hi > kTopoHalo + 2u ? a_hi + kTopoHalo + 2u : total_tri;
  if (hi - lo > kWin) hi = lo + kWin;
  for (uint32_t j = lo + threadIdx.x; j < hi; j += kTopoThreads) s_c[j - lo] = codes[j];
  __syncthreads();
  const uint32_t g = base + threadIdx.x;
  if (g > last_g) return;
  const uint32_t m = find_segment(inner_off, M, g);
  const uint32_t tb = mesh2tri[m], n = mesh2tri[m + 1] - tb;
  const MeshCodes<uint64_t> c(CodeWindow<uint64_t>(codes, s_c, lo, hi - lo), tb);
  const uint32_t i = g - inner_off[m];
  uint32_t first, last;
  node_range<uint64_t>(c, n, i, first, last);
  const uint32_t split = node_split<uint64_t>(c, first, last);
  topology_emit(n, i, first, last, split, bvhnodes + (size_t)node_off[m] * 3, idx + tb, tb, ts, m,
                tb, inner_off[m]);
}

// Meshes of one triangle: the single leaf row and its treelet.
static __global__ __launch_bounds__(256) void k_batch_singles(
    const uint32_t* __restrict__ bad, uint32_t M, const uint32_t* __restrict__ mesh2tri,
    const uint32_t* __restrict__ node_off, uint32_t* __restrict__ bvhnodes,
    const uint32_t* __restrict__ idx, TreeletSink ts) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (*bad || m >= M) return;
  const uint32_t tb = mesh2tri[m];
  if (mesh2tri[m + 1] - tb != 1) return;
  topology_single(bvhnodes + (size_t)node_off[m] * 3, idx[tb] - tb, ts, m, tb);
}

struct BatchTrees {
  const uint32_t* bad;  // nonzero: the offsets were rejected, do nothing
  uint32_t M;
  const uint32_t* mesh2tri;
  const uint32_t* mesh2vtx;
  const uint32_t* node_off;
  float* aabb;
  const uint32_t* bvhnodes;
  const uint32_t* tri2vtx;
  const float* xyz;
};

__device__ __forceinline__ TreeView tree_of(const BatchTrees& b, uint32_t m) {
  const uint32_t tb = b.mesh2tri[m], vb = b.mesh2vtx[m];
  return TreeView{b.aabb + (size_t)b.node_off[m] * 6, b.bvhnodes + (size_t)b.node_off[m] * 3,
                  b.mesh2tri[m + 1] - tb, b.tri2vtx + (size_t)tb * 3, b.mesh2vtx[m + 1] - vb,
                  b.xyz + (size_t)vb * 3};
}

// Per-mesh pack sinks of the batched refit (see NoPack): a source's
// of(m, n) is mesh m's sink; NoPackSrc packs nothing.
struct NoPackSrc {
  using PK = NoPack;
  __device__ __forceinline__ NoPack of(uint32_t, uint32_t) const { return NoPack{}; }
};

// CLIMB: the treelet climb variant (refit_treelet), 2 by default as for a
// single mesh: climbs start at the nodes whose children are both leaves.
template <int K, int T, int CLIMB, typename Src = NoPackSrc>
__global__ __launch_bounds__(T) __attribute__((amdgpu_num_sgpr(72))) void k_batch_refit_treelet(
    BatchTrees b, uint32_t total_tri, const uint32_t* __restrict__ count_p,
    const uint4* __restrict__ treelets, uint32_t* __restrict__ err, Src src) {
  __shared__ TreeletLds<K> s;
  if (*b.bad) return;
  if (CLIMB == 2) {
    if (threadIdx.x == 0) {
      s.queued = 0u;
      s.queued2 = 0u;
    }
    __syncthreads();
  }
  uint32_t count = *count_p;
  if (count > total_tri) count = total_tri;
  for (uint32_t t = blockIdx.x; t < count; t += gridDim.x) {
    const uint4 tl = treelets[t];
    if (tl.w >= b.M) {
      if (threadIdx.x == 0) raise_error(err, DMSH_DEVERR_REFIT);
      continue;
    }
    const TreeView v = tree_of(b, tl.w);
    refit_treelet<kLeafTri, K, T, false, CLIMB, typename Src::PK>(
        v, tl.x, tl.y, tl.z, 0.f, s, err, nullptr, 0, false, TopStage{}, DMSH_U32MAX,
        src.of(tl.w, v.num_leaf));
  }
}

// One workgroup per mesh with more than K triangles (grid-stride over
// meshes): the mesh's top nodes, listed by the topology kernel (top_count[m]
// slots at top_list / node2slot + inner_off[m]), merged in LDS by rounds
// (top_rounds) when they fit kBatchTopSlots, else climbed from the mesh's
// treelet roots through global counters (refit_top).
constexpr int kBatchTopSlots = 4096;

template <int K, int T, typename Src = NoPackSrc>
__global__ __launch_bounds__(T) void k_batch_refit_top(
    BatchTrees b, const uint32_t* __restrict__ inner_off, uint32_t* __restrict__ flags,
    const uint32_t* __restrict__ root_count, const uint32_t* __restrict__ roots,
    uint32_t* __restrict__ err, Src src, const uint32_t* __restrict__ top_count,
    const uint4* __restrict__ top_list, const uint32_t* __restrict__ node2slot) {
  constexpr uint32_t kList = Src::PK::kOn ? 4096u : 1u;  // merged nodes listed for the pack
  __shared__ uint32_t s_list[kList];
  __shared__ uint32_t s_count;
  __shared__ __align__(16) unsigned char s_top[(size_t)kBatchTopSlots * (sizeof(Box) + 1)];
  if (*b.bad) return;
  if (threadIdx.x == 0) s_count = 0u;
  __syncthreads();
  for (uint32_t m = blockIdx.x; m < b.M; m += gridDim.x) {
    const uint32_t n = seg_len(b.mesh2tri, m);
    if (n <= (uint32_t)K) continue;
    const TreeView v = tree_of(b, m);
    if (top_count && top_rounds<kBatchTopSlots, T, typename Src::PK>(
                         v, top_count[m], top_list + inner_off[m], node2slot + inner_off[m], err,
                         s_top, src.of(m, n)))
      continue;
    uint32_t count = root_count[m];
    if (count > n) count = n;
    refit_top<T, typename Src::PK>(v, flags + inner_off[m], roots + b.mesh2tri[m], 1u, count, err,
                                   src.of(m, n), s_list, &s_count, Src::PK::kOn ? kList : 0u);
  }
}

// ---------------------------------------------------------------------------
// ray-cast
// ---------------------------------------------------------------------------

// inner_off as for the build, tile_off[m] = sum of 8x8 tiles of the images
// before m.  Rejects (sets *bad, raises DMSH_DEVERR_INDEX) offsets that do not
// fit their arrays, a node offset that is not the build's, and an image
// whose pixel count is not a multiple of its width.
static __global__ __launch_bounds__(1024) void k_batch_cast_offsets(
    uint32_t M, uint32_t total_tri, uint32_t total_vtx, uint32_t total_pix,
    const uint32_t* __restrict__ mesh2tri, const uint32_t* __restrict__ mesh2vtx,
    const uint32_t* __restrict__ node_off, const uint32_t* __restrict__ pix_off,
    const uint32_t* __restrict__ widths, uint32_t* __restrict__ inner_off,
    uint32_t* __restrict__ tile_off, uint32_t* __restrict__ bad, uint32_t* __restrict__ err) {
  check_offsets(mesh2tri, M, total_tri, bad, err);
  check_offsets(mesh2vtx, M, total_vtx, bad, err);
  check_offsets(pix_off, M, total_pix, bad, err);
  for (uint32_t m = threadIdx.x; m < M; m += 1024) {
    const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
    const uint32_t p = pix_off[m + 1] - pix_off[m], w = widths[m];
    const bool rows_ok = node_off[m + 1] - node_off[m] == (n > 0 ? 2 * n - 1 : 0u);
    const bool img_ok = p == 0 || (w != 0 && p % w == 0);
    if (!rows_ok || !img_ok) {
      atomicOr(bad, 1u);
      raise_error(err, DMSH_DEVERR_INDEX);
    }
  }
  if (threadIdx.x == 0 && node_off[0] != 0u) {
    atomicOr(bad, 1u);
    raise_error(err, DMSH_DEVERR_INDEX);
  }
  scan_meshes(M, inner_off, [&](uint32_t m) {
    const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
    return n > 0 ? n - 1 : 0u;
  });
  scan_meshes(M, tile_off, [&](uint32_t m) {
    const uint32_t p = pix_off[m + 1] - pix_off[m], w = widths[m];
    if (p == 0 || w == 0 || p % w != 0) return 0u;
    return ceil_div_u32(w, 8u) * ceil_div_u32(p / w, 8u);
  });
}

static __global__ __launch_bounds__(256) void k_batch_pack_inner(
    BatchTrees b, uint32_t total_tri, const uint32_t* __restrict__ inner_off,
    rc::InnerNode* __restrict__ inner, uint32_t* __restrict__ nonfinite,
    uint32_t* __restrict__ err) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (*b.bad || g >= total_tri || g >= inner_off[b.M]) return;
  const uint32_t m = find_segment(inner_off, b.M, g);
  const TreeView v = tree_of(b, m);
  rc::pack_inner(inner + inner_off[m], v.bvhnodes, v.aabb, v.num_leaf, g - inner_off[m], nonfinite,
                 err);
}

static __global__ __launch_bounds__(256) void k_batch_pack_leaf(BatchTrees b, uint32_t total_tri,
                                                                rc::LeafTri* __restrict__ leaf,
                                                                uint32_t* __restrict__ err) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (*b.bad || t >= total_tri) return;
  const uint32_t m = find_segment(b.mesh2tri, b.M, t);
  const TreeView v = tree_of(b, m);
  const uint32_t tb = b.mesh2tri[m];
  rc::pack_leaf(leaf + tb, v.bvhnodes, v.tri2vtx, v.xyz, v.num_leaf, v.num_leaf, v.num_vtx, t - tb,
                err);
}

struct BatchCast {
  uint32_t M;
  const uint32_t* mesh2tri;
  const uint32_t* node_off;
  const uint32_t* inner_off;
  const uint32_t* tile_off;
  const uint32_t* pix_off;
  const uint32_t* widths;
  const rc::InnerNode* inner;
  const rc::LeafTri* leaf;
  const float* aabb;
  const float* transforms;  // [M][16] ndc2world, column-major
  uint32_t* pix2tri;
  float* pix2t;
  uint32_t* counts;  // [0] tier-2 queue, [1] tier-3 queue, [2] non-finite box flag,
                     // [3] offsets rejected, [16 + 16 x] tile counter of XCD x
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep_stack;
  uint32_t* err;
};

__device__ __forceinline__ rc::Packed packed_of(const BatchCast& a, uint32_t m) {
  const uint32_t tb = a.mesh2tri[m];
  return rc::Packed{a.inner + a.inner_off[m], a.leaf + tb, a.aabb + (size_t)a.node_off[m] * 6,
                    a.mesh2tri[m + 1] - tb};
}

__device__ __forceinline__ void batch_write(const BatchCast& a, uint32_t gp, const rc::Hit& h) {
  a.pix2tri[gp] = h.has ? h.tri : DMSH_U32MAX;
  if (a.pix2t) a.pix2t[gp] = h.has ? h.t : 0.f;
}

// Next tile for the calling wave (the same value in every lane), or
// num_tiles when none is left.
__device__ __forceinline__ uint32_t next_tile(const BatchCast& a, uint32_t num_tiles,
                                              uint32_t& xcd) {
  const uint32_t per = (num_tiles + kXcds - 1) / kXcds;
  for (uint32_t tries = 0; tries < kXcds; ++tries) {
    const uint32_t lo = xcd * per;
    const uint32_t hi = (lo + per < num_tiles) ? lo + per : num_tiles;
    uint32_t t = 0;
    if ((threadIdx.x & 63u) == 0) t = atomicAdd(a.counts + 16 + 16 * xcd, 1u);
    t = __builtin_amdgcn_readfirstlane(t) + lo;
    if (t < hi) return t;
    xcd = (xcd + 1) % kXcds;
  }
  return num_tiles;
}

// Tier 1 over all meshes: persistent workgroups of W waves, each wave taking
// its own tiles (no barrier between them; W is the workgroup size the
// autotuner picks).
template <int S, int W>
__global__ __launch_bounds__(64 * W) void k_batch_pix2tri(BatchCast a) {
  __shared__ uint2 s_stack_all[W][S][64];
  uint2 (*s_stack)[64] = s_stack_all[threadIdx.x >> 6];
  const uint32_t lane = threadIdx.x & 63u;
  if (a.counts[3]) return;  // offsets rejected
  const uint32_t num_tiles = a.tile_off[a.M];
  const bool boxes_finite = a.counts[2] == 0u;
  uint32_t xcd = blockIdx.x % kXcds;
  for (;;) {
    const uint32_t tile = next_tile(a, num_tiles, xcd);
    if (tile >= num_tiles) break;
    const uint32_t m = find_segment(a.tile_off, a.M, tile);
    const uint32_t w = a.widths[m], p0 = a.pix_off[m], h = (a.pix_off[m + 1] - p0) / w;
    const uint32_t tiles_x = ceil_div_u32(w, 8u), lt = tile - a.tile_off[m];
    const uint32_t iw = (lt % tiles_x) * 8u + (lane & 7u), ih = (lt / tiles_x) * 8u + (lane >> 3);
    if (iw < w && ih < h) {
      float T[16];
      rc::load_transform(a.transforms + (size_t)m * 16, T);
      float org[3], dir[3];
      ray_for_pixel(iw, ih, w, h, T, org, dir);
      rc::LdsStack<S> stk{&s_stack[0][lane], 64};
      rc::Hit hit;
      const int st = rc::trace(packed_of(a, m), org, dir, stk, hit, boxes_finite);
      const uint32_t gp = p0 + ih * w + iw;
      if (st == rc::kTraceOverflow) a.list2[atomicAdd(a.counts + 0, 1u)] = gp;
      else batch_write(a, gp, hit);
    }
  }
}

__device__ __forceinline__ void batch_pixel(const BatchCast& a, uint32_t gp, uint32_t& m,
                                            uint32_t& iw, uint32_t& ih, uint32_t& w,
                                            uint32_t& h) {
  m = find_segment(a.pix_off, a.M, gp);
  w = a.widths[m];
  const uint32_t p0 = a.pix_off[m];
  h = (a.pix_off[m + 1] - p0) / w;
  ih = (gp - p0) / w;
  iw = gp - p0 - ih * w;
}

static __global__ __launch_bounds__(64) void k_batch_pix2tri_tier2(BatchCast a,
                                                                   uint32_t total_pix) {
  __shared__ uint2 s_stack[rc::kTier2Frames][64];
  if (a.counts[3]) return;
  uint32_t count = a.counts[0];
  if (count > total_pix) count = total_pix;
  const bool boxes_finite = a.counts[2] == 0u;
  rc::LdsStack<rc::kTier2Frames> stk{&s_stack[0][threadIdx.x], 64};
  for (uint32_t e = blockIdx.x * 64u + threadIdx.x; e < count; e += gridDim.x * 64u) {
    const uint32_t gp = a.list2[e];
    uint32_t m, iw, ih, w, h;
    batch_pixel(a, gp, m, iw, ih, w, h);
    float T[16], org[3], dir[3];
    rc::load_transform(a.transforms + (size_t)m * 16, T);
    ray_for_pixel(iw, ih, w, h, T, org, dir);
    rc::Hit hit;
    const int st = rc::trace(packed_of(a, m), org, dir, stk, hit, boxes_finite);
    if (st == rc::kTraceOverflow) a.list3[atomicAdd(a.counts + 1, 1u)] = gp;
    else batch_write(a, gp, hit);
  }
}

static __global__ __launch_bounds__(64) void k_batch_pix2tri_tier3(BatchCast a,
                                                                   uint32_t total_pix) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nthr = gridDim.x * blockDim.x;
  if (a.counts[3]) return;
  uint32_t count = a.counts[1];
  if (count > total_pix) count = total_pix;
  const bool boxes_finite = a.counts[2] == 0u;
  rc::GlobalStack stk{a.deep_stack + (size_t)gid * rc::kDeepFrames};
  for (uint32_t e = gid; e < count; e += nthr) {
    const uint32_t gp = a.list3[e];
    uint32_t m, iw, ih, w, h;
    batch_pixel(a, gp, m, iw, ih, w, h);
    float T[16], org[3], dir[3];
    rc::load_transform(a.transforms + (size_t)m * 16, T);
    ray_for_pixel(iw, ih, w, h, T, org, dir);
    rc::Hit hit;
    const int st = rc::trace(packed_of(a, m), org, dir, stk, hit, boxes_finite);
    if (st != rc::kTraceOk) {
      raise_error(a.err, DMSH_DEVERR_STACK);
      hit.has = false;
    }
    batch_write(a, gp, hit);
  }
}

}  // namespace bat
}  // namespace dmsh
