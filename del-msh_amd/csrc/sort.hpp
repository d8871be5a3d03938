// sort.hpp -- stable LSD radix argsort for gfx950 (u32 / u64 keys, u32 values).
//
// Design (one pass per 8-bit digit, Onesweep-style):
//   * k_hist_all reads the keys once and builds the digit histogram of every
//     pass it runs (LDS counters, then one global atomic per bin and block).
//   * k_onesweep, one launch per digit: a block takes the next tile from an
//     atomic ticket (so tiles are started in order), ranks its elements with
//     wave64 ballots (one ballot per digit bit gives each lane the mask of
//     lanes holding the same digit; popcount below the lane is its rank),
//     publishes per-digit tile counts and resolves its global offset by
//     decoupled look-back over the preceding tiles' published counts, then
//     stages the tile in LDS in digit order and writes it out coalesced.
// Stability: ranks follow element order inside a wave round, wave rounds go
// in element order, waves and tiles are prefixed in order.  The look-back
// spin is bounded; a block that gives up raises DMSH_DEVERR_SORT_PROGRESS and
// still exits, so a launch always drains.
//
// Keys of five or more digits (63-bit Morton codes, batch keys) take
// kTopPasses passes over their top digits only, which leaves them grouped by
// top bits with each group in element order; k_sort_finish then orders every
// group in LDS by the whole key, ties by position (a group is a handful of
// elements for Morton codes of spread-out geometry).  The result is the same
// stable order as the full LSD sort.  A group too long for the finish's LDS
// window (heavily repeated top bits) is listed instead, and k_sort_groups,
// one launch that does nothing when the list is empty, sorts each listed
// group by its low digits in one workgroup.
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {
namespace sortk {

constexpr int kThreads = 256;  // 4 waves
constexpr int kWideThreads = 512;             // wide tiles (see k_onesweep)
constexpr uint32_t kWideTilesFrom = 1u << 21;  // elements from which the default takes them
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 8;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kLookBack = 8;
constexpr uint32_t kTopPasses = 3;      // global passes before the finish
constexpr uint32_t kFinishHalo = 1024;  // LDS window margin of the finish
// One ticket per launched pass (kMaxPasses at most); the last word counts
// the groups the finish leaves to k_sort_groups.
constexpr uint32_t kTicketWords = 2 * kMaxPasses;

// Groups longer than kFinishHalo: at most this many in n elements.
inline uint32_t group_capacity(uint32_t n) { return n / (kFinishHalo + 1) + 1; }

// Keys of five or more digits take the top passes and the finish (with four,
// three passes and the finish cost about as much as four plain passes).
__host__ __device__ inline bool uses_finish(uint32_t num_passes) {
  return num_passes > kTopPasses + 1;
}

// The first digit whose histogram the key producers of a sort that may
// finish in LDS accumulate: with the finish, only the top digits'
// (k_sort_groups counts the low digits of its groups itself).
__host__ __device__ inline uint32_t first_hist_pass(uint32_t num_passes) {
  return uses_finish(num_passes) ? num_passes - kTopPasses : 0u;
}

// Passes that take a status region: the top passes when the finish is
// used, else every digit.
inline uint32_t status_passes(uint32_t num_passes) {
  return uses_finish(num_passes) ? kTopPasses : num_passes;
}

template <typename KeyT>
struct Tune {
  static constexpr int kItems = 16;
};

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += u;
  }
  return v;
}

// Exclusive scan of one value per thread over the block (s_wsum: one word
// per wave).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wsum) {
  const uint32_t inc = wave_inclusive_scan(v);
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 63u) s_wsum[wave] = inc;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave; ++w) base += s_wsum[w];
  __syncthreads();
  return base + inc - v;
}

__host__ __device__ __forceinline__ uint32_t pass_bits(uint32_t key_bits, uint32_t p) {
  const uint32_t lo = p * kRadixBits;
  const uint32_t rem = key_bits > lo ? key_bits - lo : 0u;
  return rem < (uint32_t)kRadixBits ? rem : (uint32_t)kRadixBits;
}

template <typename KeyT>
__device__ __forceinline__ uint32_t digit_of(KeyT k, uint32_t shift, uint32_t mask) {
  return (uint32_t)(k >> shift) & mask;
}

static __global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, size_t count,
                                                  uint32_t value) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = value;
}

template <typename KeyT>
__global__ __launch_bounds__(256) void k_copy_keys(const KeyT* __restrict__ a, KeyT* __restrict__ b,
                                                   uint32_t n, const uint32_t* __restrict__ n_dev) {
  if (n_dev) n = *n_dev;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = a[i];
}

// Accumulates the histograms of passes [first, num_passes) into the LDS bins
// of this block.
template <typename KeyT>
__device__ __forceinline__ void hist_accumulate(uint32_t (*s_h)[kRadix], KeyT k, uint32_t first,
                                                uint32_t num_passes, uint32_t key_bits) {
  for (uint32_t p = first; p < num_passes; ++p) {
    const uint32_t mask = (1u << pass_bits(key_bits, p)) - 1u;
    atomicAdd(&s_h[p][digit_of(k, p * kRadixBits, mask)], 1u);
  }
}

__device__ __forceinline__ void hist_flush(uint32_t (*s_h)[kRadix], uint32_t first,
                                           uint32_t num_passes, uint32_t* __restrict__ hist) {
  for (uint32_t e = first * kRadix + threadIdx.x; e < num_passes * kRadix; e += blockDim.x) {
    const uint32_t v = (&s_h[0][0])[e];
    if (v) atomicAdd(hist + e, v);
  }
}

template <typename KeyT>
__global__ __launch_bounds__(256) void k_hist_all(const KeyT* __restrict__ keys, uint32_t n,
                                                  const uint32_t* __restrict__ n_dev,
                                                  uint32_t first, uint32_t num_passes,
                                                  uint32_t key_bits, uint32_t* __restrict__ hist) {
  if (n_dev) n = *n_dev;
  __shared__ uint32_t s_h[kMaxPasses][kRadix];
  for (uint32_t e = threadIdx.x; e < kMaxPasses * kRadix; e += blockDim.x) (&s_h[0][0])[e] = 0;
  __syncthreads();
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    hist_accumulate(s_h, keys[i], first, num_passes, key_bits);
  __syncthreads();
  hist_flush(s_h, first, num_passes, hist);
}

// Scatter offsets of a segmented pass (keys = segment << shift | code, the
// segments contiguous and in order in the input): an element lands at the
// number of elements with its digit before it over all segments (what the
// look-back counts) + off[segment * kRadix + digit].  Segment indices above
// `last` (keys the producer left unwritten after a data error) read row
// `last`; the positions stay in range because the offsets are then zero.
struct SegOff {
  const uint32_t* off;  // [last + 1][kRadix]
  uint32_t shift;
  uint32_t last;
};

// One digit pass.  IOTA: values are the element indices (first pass).  With
// `n_dev`, the element
// count is read from the device (the grid is sized for a capacity); blocks
// whose ticket lies past the last tile exit before publishing anything, and
// only earlier tiles are ever looked back at, so they wait on nobody.
// THREADS: 256 (4 waves) or 512 (8 waves, for large inputs: half the tiles,
// so half the same-address tickets and a look-back chain half as long).
// SEG: a segmented pass (see SegOff); `hist` is not read.
template <typename KeyT, int ITEMS, bool IOTA, int THREADS = kThreads, bool SEG = false>
__global__ __launch_bounds__(THREADS) void k_onesweep(
    const KeyT* __restrict__ kin, const uint32_t* __restrict__ vin, KeyT* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, const uint32_t* __restrict__ n_dev, uint32_t shift,
    uint32_t dbits, const uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
    uint32_t* __restrict__ ticket, uint32_t* __restrict__ err, SegOff seg) {
  static_assert(THREADS >= kRadix && THREADS % 64 == 0, "one thread per digit");
  if (n_dev) n = *n_dev;
  constexpr int kTile = THREADS * ITEMS;
  constexpr uint32_t W = THREADS / 64;
  __shared__ uint32_t s_wcnt[W][kRadix];
  __shared__ uint32_t s_excl[kRadix];
  __shared__ uint32_t s_gbase[kRadix];
  __shared__ uint32_t s_wsum[W];
  __shared__ uint32_t s_tile;
  __shared__ KeyT s_key[kTile];
  __shared__ uint32_t s_val[kTile];

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  const uint32_t mask = (1u << dbits) - 1u;
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  for (uint32_t e = tid; e < W * kRadix; e += THREADS) (&s_wcnt[0][0])[e] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t tile_base = tile * (uint32_t)kTile;
  if (tile_base >= n) return;  // past the data (a capacity-sized grid)
  const uint32_t wave_base = tile_base + wave * 64u * ITEMS;

  KeyT key[ITEMS];
  uint32_t val[ITEMS], dig[ITEMS], pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t idx = wave_base + (uint32_t)i * 64u + lane;
    const bool valid = idx < n;
    key[i] = valid ? kin[idx] : (KeyT)0;
    val[i] = valid ? (IOTA ? idx : vin[idx]) : 0u;
    dig[i] = valid ? digit_of(key[i], shift, mask) : 0xFFFFFFFFu;
  }
  // Rank inside the wave: rounds in element order, lanes in element order.
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = dig[i] != 0xFFFFFFFFu;
    uint64_t m = __ballot(valid);
    for (uint32_t b = 0; b < dbits; ++b) {
      const bool bit = (dig[i] >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    if (valid) {
      const uint32_t rank = mask_rank(m);
      const uint32_t old = s_wcnt[wave][dig[i]];
      pos[i] = old + rank;
      if (rank == 0) s_wcnt[wave][dig[i]] = old + (uint32_t)__popcll(m);
    }
  }
  __syncthreads();
  // Per digit (thread d < kRadix): wave prefixes and the tile count.
  const uint32_t d = tid;
  const bool has_d = d < (uint32_t)kRadix;
  uint32_t cnt = 0;
  if (has_d) {
#pragma unroll
    for (uint32_t w = 0; w < W; ++w) {
      const uint32_t c = s_wcnt[w][d];
      s_wcnt[w][d] = cnt;
      cnt += c;
    }
  }
  uint32_t* my_status = status + (size_t)tile * kRadix + d;
  if (has_d)
    __hip_atomic_store(my_status, (tile == 0 ? kFlagPrefix : kFlagAgg) | cnt, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t excl = block_exclusive_scan(cnt, s_wsum);
  uint32_t goff = 0;
  if constexpr (!SEG) goff = block_exclusive_scan(has_d ? hist[d] : 0u, s_wsum);
  // Decoupled look-back over the preceding tiles for digit d, kLookBack
  // status words per round trip: tiles that start together all publish their
  // counts at about the same time and then walk back over each other, so a
  // walk of k tiles costs k / kLookBack memory latencies instead of k.
  uint32_t prefix = 0;
  if (has_d && tile > 0) {
    uint32_t tt = tile - 1, spins = 0;  // next tile to read
    bool done = false;
    while (!done) {
      uint32_t v[kLookBack];
#pragma unroll
      for (int b = 0; b < kLookBack; ++b)
        v[b] = (uint32_t)b <= tt ? __hip_atomic_load(status + (size_t)(tt - b) * kRadix + d,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                 : 0u;
      uint32_t used = 0;  // words consumed this round, in tile order
#pragma unroll
      for (int b = 0; b < kLookBack; ++b) {
        if (done || used != (uint32_t)b) continue;
        const uint32_t flag = v[b] & ~kValueMask;
        if (flag == 0u) continue;  // not published yet: read again from this tile
        prefix += v[b] & kValueMask;
        ++used;
        if (flag == kFlagPrefix || tt == (uint32_t)b) done = true;
      }
      if (done) break;
      if (used < (uint32_t)kLookBack) {
        if (++spins > kSpinLimit) {
          raise_error(err, DMSH_DEVERR_SORT_PROGRESS);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      tt -= used;
    }
    __hip_atomic_store(my_status, kFlagPrefix | (prefix + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (has_d) {
    s_excl[d] = excl;
    s_gbase[d] = goff + prefix - excl;  // modular; final positions are in range
  }
  __syncthreads();
  // Stage in digit order.
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (dig[i] != 0xFFFFFFFFu) {
      const uint32_t lp = s_excl[dig[i]] + s_wcnt[wave][dig[i]] + pos[i];
      if (lp < (uint32_t)kTile) {
        s_key[lp] = key[i];
        s_val[lp] = val[i];
      }
    }
  }
  __syncthreads();
  const uint32_t tile_n = (n - tile_base) < (uint32_t)kTile ? (n - tile_base) : (uint32_t)kTile;
  for (uint32_t j = tid; j < tile_n; j += THREADS) {
    const KeyT k = s_key[j];
    const uint32_t dk = digit_of(k, shift, mask);
    uint32_t g = s_gbase[dk] + j;
    if constexpr (SEG) {
      const KeyT sk = k >> seg.shift;
      const uint32_t m = sk < (KeyT)seg.last ? (uint32_t)sk : seg.last;
      g += seg.off[(size_t)m * kRadix + dk];
    }
    if (g < n) {
      kout[g] = k;
      vout[g] = s_val[j];
    } else {
      raise_error(err, DMSH_DEVERR_SORT_PROGRESS);
    }
  }
}

// The finish.  Input: keys (and element indices) sorted stably by their bits
// at and above `shift` (masked to the sort's key bits).  Each block loads
// kFinishHalo keys either side of its 4096 elements into LDS and marks the
// group starts (equal bits above `shift` form a group) in an LDS bit set;
// each element finds the ends of its group from that set, at most
// kFinishHalo - 1 slots down and kFinishHalo up, and lands at group start +
// the number of group members whose masked key is smaller, or equal and
// earlier.  An element whose group end lies beyond that span writes nothing;
// the first element of such a group lists it (groups[*group_count++]) for
// k_sort_groups.
constexpr int kFinishThreads = 1024;
constexpr int kFinishItems = 4;

template <typename KeyT>
__global__ __launch_bounds__(kFinishThreads) void k_sort_finish(const KeyT* __restrict__ kin,
                                                     const uint32_t* __restrict__ vin,
                                                     KeyT* __restrict__ kout,
                                                     uint32_t* __restrict__ vout, uint32_t n,
                                                     uint32_t shift, KeyT kmask,
                                                     uint32_t* __restrict__ group_count,
                                                     uint32_t* __restrict__ groups,
                                                     uint32_t group_cap) {
  constexpr uint32_t kTW = (uint32_t)kFinishThreads * kFinishItems;
  constexpr uint32_t kG = kFinishHalo;
  constexpr uint32_t kWin = kTW + 2 * kG;
  constexpr int kLoads = (int)(kWin / kFinishThreads);
  static_assert(kWin % kFinishThreads == 0, "the window is whole rounds of the block");
  __shared__ KeyT s_k[kWin];  // keys as stored (the mask applies when compared)
  const int64_t w0 = (int64_t)blockIdx.x * kTW - kG;  // element index of slot 0
  const uint32_t lo = w0 < 0 ? (uint32_t)(-w0) : 0u;  // first slot holding an element
  const int64_t rem = (int64_t)n - w0;
  const uint32_t hi = rem < (int64_t)kWin ? (uint32_t)rem : kWin;  // one past the last
  // every load of the window and of this thread's values is issued up front
  KeyT kr[kLoads];
#pragma unroll
  for (int q = 0; q < kLoads; ++q) {
    const uint32_t j = (uint32_t)q * kFinishThreads + threadIdx.x;
    kr[q] = (j >= lo && j < hi) ? kin[w0 + j] : (KeyT)0;
  }
  uint32_t vr[kFinishItems];
#pragma unroll
  for (int r = 0; r < kFinishItems; ++r) {
    const uint32_t li = kG + (uint32_t)r * kFinishThreads + threadIdx.x;
    vr[r] = li < hi ? vin[w0 + li] : 0u;
  }
#pragma unroll
  for (int q = 0; q < kLoads; ++q) s_k[(uint32_t)q * kFinishThreads + threadIdx.x] = kr[q];
  // group starts as a bit set over the window: slot j starts a group if it is
  // the first element of the sort or its top bits differ from slot j - 1's;
  // slot hi is marked when the sort ends there (kWin / 32 + 1 words)
  __shared__ uint32_t s_start[kWin / 32 + 1];
  const uint32_t lane = threadIdx.x & 63u;
  if (threadIdx.x == 0) s_start[kWin / 32] = (w0 + (int64_t)kWin == (int64_t)n) ? 1u : 0u;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kLoads; ++q) {
    const uint32_t j = (uint32_t)q * kFinishThreads + threadIdx.x;
    bool st;
    if (j < lo || j > hi) st = false;
    else if (j == lo) st = w0 + (int64_t)j == 0;
    else if (j == hi) st = true;  // hi < kWin here: the sort ends at slot hi
    else st = ((s_k[j] & kmask) >> shift) != ((s_k[j - 1] & kmask) >> shift);
    const uint64_t m = __ballot(st);
    if (lane == 0) {
      s_start[j >> 5] = (uint32_t)m;
      s_start[(j >> 5) + 1] = (uint32_t)(m >> 32);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kFinishItems; ++r) {
    const uint32_t li = kG + (uint32_t)r * kFinishThreads + threadIdx.x;
    if (li < hi) {
      const KeyT full = s_k[li];
      const KeyT key = full & kmask;
      // a: the last group start <= li, looked for down to li - (kG - 1)
      uint32_t w = li >> 5;
      uint32_t bits = s_start[w] & (0xFFFFFFFFu >> (31u - (li & 31u)));
      const uint32_t wlo = (li - (kG - 1)) >> 5;
      while (bits == 0u && w > wlo) bits = s_start[--w];
      const bool found_a = bits != 0u;
      const uint32_t a = found_a ? w * 32u + 31u - (uint32_t)__clz(bits) : 0u;
      // b: the first group start > li (or the end of the sort), up to li + kG
      w = (li + 1) >> 5;
      bits = s_start[w] & (0xFFFFFFFFu << ((li + 1) & 31u));
      const uint32_t whi = (li + kG) >> 5;
      while (bits == 0u && w < whi) bits = s_start[++w];
      const bool found_b = bits != 0u;
      const uint32_t b = found_b ? w * 32u + (uint32_t)__ffs(bits) - 1u : 0u;
      if (!found_a || !found_b || a + (kG - 1) < li || b > li + kG) {
        // a group end lies outside the searched span: the group is longer
        // than kG, so its first element sees it too and lists it
        if (found_a && a == li) {
          const uint32_t slot = atomicAdd(group_count, 1u);
          if (slot < group_cap) groups[slot] = (uint32_t)(w0 + li);
        }
      } else {
        uint32_t rank = 0;
#pragma unroll 4
        for (uint32_t m = a; m < b; ++m) {
          const KeyT km = s_k[m] & kmask;
          rank += (km < key || (km == key && m < li)) ? 1u : 0u;
        }
        const size_t dst = (size_t)(w0 + a) + rank;
        kout[dst] = full;
        vout[dst] = vr[r];
      }
    }
  }
}

// The groups the finish listed (more than kFinishHalo elements with equal
// bits at and above `shift`), one workgroup per group.  The group runs from
// its listed first element to the first later element with other top bits;
// it is sorted by the digits below `shift`, least significant first, each
// pass stable in element order (ranks inside a wave by ballots, then waves
// and chunks in order), alternating between the group's range of the
// top-sorted buffer (kin, vin) and of the output (kout, vout), with a copy
// back to the output after an even number of passes.  Only the group's
// ranges are touched; every other element was placed by the finish.
constexpr int kGroupThreads = 1024;
constexpr int kGroupItems = 4;
constexpr uint32_t kGroupGrid = 256;

template <typename KeyT>
__global__ __launch_bounds__(kGroupThreads) void k_sort_groups(
    KeyT* __restrict__ kin, uint32_t* __restrict__ vin, KeyT* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, uint32_t shift, KeyT kmask,
    const uint32_t* __restrict__ group_count, const uint32_t* __restrict__ groups,
    uint32_t group_cap) {
  constexpr uint32_t W = kGroupThreads / 64;
  constexpr uint32_t kChunk = (uint32_t)kGroupThreads * kGroupItems;
  __shared__ uint32_t s_hist[kMaxPasses][kRadix];
  __shared__ uint32_t s_wcnt[W][kRadix];
  __shared__ uint32_t s_base[kRadix];
  __shared__ uint32_t s_tot[kRadix];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_end;
  uint32_t count = *group_count;
  if (count > group_cap) count = group_cap;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
  // digits covering the bits below the group's common bits (a digit that
  // straddles `shift` is constant above it inside the group)
  const uint32_t low = (shift + kRadixBits - 1) / kRadixBits;
  for (uint32_t g = blockIdx.x; g < count; g += gridDim.x) {
    const uint32_t a = groups[g];
    if (a >= n) continue;  // uniform over the workgroup
    const KeyT top = (kin[a] & kmask) >> shift;
    if (tid == 0) s_end = n;
    for (uint32_t e = tid; e < low * (uint32_t)kRadix; e += kGroupThreads) (&s_hist[0][0])[e] = 0u;
    __syncthreads();
    for (uint32_t base = a + kFinishHalo + 1; base < n; base += kGroupThreads) {
      const uint32_t i = base + tid;
      const bool diff = i < n && ((kin[i] & kmask) >> shift) != top;
      if (diff) atomicMin(&s_end, i);
      if (__syncthreads_or(diff)) break;
    }
    const uint32_t end = s_end;
    for (uint32_t i = a + tid; i < end; i += kGroupThreads) {
      const KeyT k = kin[i];
      for (uint32_t p = 0; p < low; ++p)
        atomicAdd(&s_hist[p][digit_of(k, p * kRadixBits, (uint32_t)kRadix - 1u)], 1u);
    }
    for (uint32_t p = 0; p < low; ++p) {
      KeyT* sk = (p & 1u) ? kout : kin;
      uint32_t* sv = (p & 1u) ? vout : vin;
      KeyT* dk = (p & 1u) ? kin : kout;
      uint32_t* dv = (p & 1u) ? vin : vout;
      __syncthreads();  // histograms complete; the previous pass's stores visible
      uint32_t v = 0, inc = 0;
      if (tid < (uint32_t)kRadix) {
        v = s_hist[p][tid];
        inc = wave_inclusive_scan(v);
        if (lane == 63u) s_wsum[wave] = inc;
      }
      for (uint32_t e = tid; e < W * kRadix; e += kGroupThreads) (&s_wcnt[0][0])[e] = 0u;
      __syncthreads();
      if (tid < (uint32_t)kRadix) {
        uint32_t b = 0;
        for (uint32_t w = 0; w < wave; ++w) b += s_wsum[w];
        s_base[tid] = a + b + inc - v;
      }
      for (uint32_t cb = a; cb < end; cb += kChunk) {
        KeyT key[kGroupItems];
        uint32_t val[kGroupItems], dig[kGroupItems], pos[kGroupItems];
#pragma unroll
        for (int i = 0; i < kGroupItems; ++i) {
          const uint32_t idx = cb + wave * 64u * kGroupItems + (uint32_t)i * 64u + lane;
          const bool valid = idx < end;
          key[i] = valid ? sk[idx] : (KeyT)0;
          val[i] = valid ? sv[idx] : 0u;
          dig[i] = valid ? digit_of(key[i], p * kRadixBits, (uint32_t)kRadix - 1u) : 0xFFFFFFFFu;
        }
        // ranks inside the wave: rounds in element order, lanes in element order
#pragma unroll
        for (int i = 0; i < kGroupItems; ++i) {
          const bool valid = dig[i] != 0xFFFFFFFFu;
          uint64_t m = __ballot(valid);
          for (uint32_t b = 0; b < (uint32_t)kRadixBits; ++b) {
            const bool bit = (dig[i] >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
          }
          if (valid) {
            const uint32_t rank = mask_rank(m);
            const uint32_t old = s_wcnt[wave][dig[i]];
            pos[i] = old + rank;
            if (rank == 0) s_wcnt[wave][dig[i]] = old + (uint32_t)__popcll(m);
          }
        }
        __syncthreads();
        if (tid < (uint32_t)kRadix) {
          uint32_t run = 0;
          for (uint32_t w = 0; w < W; ++w) {
            const uint32_t c = s_wcnt[w][tid];
            s_wcnt[w][tid] = run;
            run += c;
          }
          s_tot[tid] = run;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kGroupItems; ++i) {
          if (dig[i] == 0xFFFFFFFFu) continue;
          const uint32_t out = s_base[dig[i]] + s_wcnt[wave][dig[i]] + pos[i];
          if (out >= a && out < end) {
            dk[out] = key[i];
            dv[out] = val[i];
          }
        }
        __syncthreads();
        if (tid < (uint32_t)kRadix) {
          s_base[tid] += s_tot[tid];
          for (uint32_t w = 0; w < W; ++w) s_wcnt[w][tid] = 0u;
        }
        __syncthreads();
      }
    }
    if (low % 2u == 0u) {  // the last pass wrote (kin, vin)
      __syncthreads();
      for (uint32_t i = a + tid; i < end; i += kGroupThreads) {
        kout[i] = kin[i];
        vout[i] = vin[i];
      }
    }
    __syncthreads();  // LDS and s_end are reused by the next group
  }
}

template <typename KeyT>
inline uint32_t num_tiles(uint32_t n) {
  return ceil_div(n, (uint64_t)kThreads * Tune<KeyT>::kItems);
}

// Elements per thread of a digit pass: the default (Tune), or 8 / 12 / 16 for
// callers that tune it (the batched build); a workspace used with more than
// one value is sized for the smallest, which has the most tiles.
constexpr int kMinItems = 8;
inline bool valid_items(int items) { return items == 8 || items == 12 || items == 16; }
inline uint32_t num_tiles_items(uint32_t n, int items) {
  return ceil_div(n, (uint64_t)kThreads * (uint64_t)items);
}

inline uint32_t num_passes(uint32_t key_bits) { return ceil_div(key_bits, kRadixBits); }

// Status words of every pass that may run (see status_passes).
template <typename KeyT>
inline size_t status_words(uint32_t n, uint32_t key_bits) {
  return (size_t)status_passes(num_passes(key_bits)) * num_tiles<KeyT>(n) * kRadix;
}

// Workspace: hist[8][256] + tickets[16] + status[passes][tiles][256] +
// alt keys + alt values + the finish's group list.
template <typename KeyT>
inline size_t workspace_bytes(uint32_t n, uint32_t key_bits) {
  Arena a{nullptr, 0, 0};
  a.take<uint32_t>(kMaxPasses * kRadix + kTicketWords);
  a.take<uint32_t>(status_words<KeyT>(n, key_bits));
  a.take<KeyT>(n);
  a.take<uint32_t>(n);
  a.take<uint32_t>(group_capacity(n));
  return a.off + 256;
}

struct Layout {
  uint32_t* hist;     // [kMaxPasses][kRadix], then tickets[kTicketWords]
  uint32_t* tickets;  // the last word counts the groups left by the finish
  uint32_t* status;   // [status_passes][tiles][kRadix]
  size_t zero_words;  // hist + tickets + status are contiguous from `hist`
  void* alt_keys;
  uint32_t* alt_vals;
  uint32_t* groups;   // [group_capacity(n)]: first elements of those groups
};

template <typename KeyT>
inline bool carve(Arena& a, uint32_t n, uint32_t key_bits, Layout& L) {
  L.hist = a.take<uint32_t>(kMaxPasses * kRadix + kTicketWords);
  L.tickets = L.hist + kMaxPasses * kRadix;
  const size_t sw = status_words<KeyT>(n, key_bits);
  L.status = a.take<uint32_t>(sw);
  L.zero_words = (size_t)(L.status - L.hist) + sw;
  L.alt_keys = a.take<KeyT>(n);
  L.alt_vals = a.take<uint32_t>(n);
  L.groups = a.take<uint32_t>(group_capacity(n));
  return a.ok();
}

inline int launch_fill(hipStream_t s, uint32_t* p, size_t count, uint32_t value) {
  if (count == 0) return DMSH_OK;
  const uint32_t blocks = (uint32_t)std::min<size_t>(ceil_div(count, 256), 4096);
  launch("k_fill_u32", k_fill_u32, dim3(blocks), dim3(256), 0, s, p, count, value);
  return launch_status();
}

// One digit pass over (sk, sv) into (dk, dv); `slot` picks its status region
// and ticket.  iota: the values are the element indices.
// SEG: a segmented pass over the 8 bits at `shift` (digit then only names
// the status region's histogram slot, which is not read).
template <typename KeyT, int ITEMS, bool SEG = false>
inline int launch_pass_items(hipStream_t s, uint32_t n, uint32_t key_bits, const Layout& L,
                             uint32_t digit, uint32_t slot, bool iota, const KeyT* sk,
                             const uint32_t* sv, KeyT* dk, uint32_t* dv, uint32_t* err,
                             const uint32_t* n_dev, SegOff seg = SegOff{nullptr, 0u, 0u},
                             uint32_t seg_shift = 0u) {
  const uint32_t tiles = num_tiles_items(n, ITEMS);
  const uint32_t shift = SEG ? seg_shift : digit * kRadixBits;
  const uint32_t dbits = SEG ? (uint32_t)kRadixBits : pass_bits(key_bits, digit);
  // each pass's status region is sized for 256-thread tiles
  uint32_t* st = L.status + (size_t)slot * tiles * kRadix;
  const int wide_opt = refit_option(kSortWideTiles);
  const bool wide = wide_opt == 2 || (wide_opt == 1 && n >= kWideTilesFrom);
  if (wide) {
    const uint32_t wt = ceil_div(n, (uint64_t)kWideThreads * ITEMS);
    if (iota) {
      launch("k_onesweep", (k_onesweep<KeyT, ITEMS, true, kWideThreads, SEG>), dim3(wt),
             dim3(kWideThreads), 0, s, sk, sv, dk, dv, n, n_dev, shift, dbits,
             L.hist + digit * kRadix, st, L.tickets + slot, err, seg);
    } else {
      launch("k_onesweep", (k_onesweep<KeyT, ITEMS, false, kWideThreads, SEG>), dim3(wt),
             dim3(kWideThreads), 0, s, sk, sv, dk, dv, n, n_dev, shift, dbits,
             L.hist + digit * kRadix, st, L.tickets + slot, err, seg);
    }
  } else if (iota) {
    launch("k_onesweep", (k_onesweep<KeyT, ITEMS, true, kThreads, SEG>), dim3(tiles), dim3(kThreads), 0, s, sk,
           sv, dk, dv, n, n_dev, shift, dbits, L.hist + digit * kRadix, st, L.tickets + slot, err,
           seg);
  } else {
    launch("k_onesweep", (k_onesweep<KeyT, ITEMS, false, kThreads, SEG>), dim3(tiles), dim3(kThreads), 0, s, sk,
           sv, dk, dv, n, n_dev, shift, dbits, L.hist + digit * kRadix, st, L.tickets + slot, err,
           seg);
  }
  return launch_status();
}

// items: elements per thread (0: the default of KeyT).
template <typename KeyT>
inline int launch_pass(hipStream_t s, uint32_t n, uint32_t key_bits, const Layout& L,
                       uint32_t digit, uint32_t slot, bool iota, const KeyT* sk,
                       const uint32_t* sv, KeyT* dk, uint32_t* dv, uint32_t* err,
                       const uint32_t* n_dev = nullptr, int items = 0) {
  switch (items ? items : Tune<KeyT>::kItems) {
    case 8:
      return launch_pass_items<KeyT, 8>(s, n, key_bits, L, digit, slot, iota, sk, sv, dk, dv,
                                        err, n_dev);
    case 12:
      return launch_pass_items<KeyT, 12>(s, n, key_bits, L, digit, slot, iota, sk, sv, dk, dv,
                                         err, n_dev);
    case 16:
      return launch_pass_items<KeyT, 16>(s, n, key_bits, L, digit, slot, iota, sk, sv, dk, dv,
                                         err, n_dev);
    default:
      return DMSH_ERR_INVALID;
  }
}

// Sorts the keys in `first_src` (values: the element indices).  With `finish`
// (allowed for uses_finish(P) keys) the histograms of digits
// [first_hist_pass(P), P) must already be in L.hist and the others zero;
// without it, every digit's.  L.tickets / L.status must be zero.  The sorted keys and permutation end in (dst_keys, dst_vals).
//   Plain (up to four digits, or no finish): one pass per digit, alternating
//   so that the last lands in dst; `first_src` may alias dst_keys only when
//   the number of passes is even.
//   With the finish: top passes first_src -> alt -> dst -> alt, the finish
//   alt -> dst, and k_sort_groups for the groups the finish leaves (idle when
//   there are none).  `first_src` must not be the alternate key buffer.
//   With `n_dev` (plain path only) the element count is read from the device
//   and n is the capacity the grids are sized for; with `first_vals` the
//   first pass carries those values instead of the element indices.
template <typename KeyT>
inline int run_passes(hipStream_t s, uint32_t n, uint32_t key_bits, const Layout& L,
                      const KeyT* first_src, KeyT* dst_keys, uint32_t* dst_vals, uint32_t* err,
                      bool finish, const uint32_t* n_dev = nullptr,
                      const uint32_t* first_vals = nullptr, int items = 0) {
  const uint32_t P = num_passes(key_bits);
  KeyT* alt_k = static_cast<KeyT*>(L.alt_keys);
  uint32_t* alt_v = L.alt_vals;
  int rc;
  if (!finish || !uses_finish(P)) {
    const KeyT* src_k = first_src;
    const uint32_t* src_v = first_vals;
    for (uint32_t p = 0; p < P; ++p) {
      const bool to_user = ((P - 1 - p) % 2) == 0;
      KeyT* dk = to_user ? dst_keys : alt_k;
      uint32_t* dv = to_user ? dst_vals : alt_v;
      rc = launch_pass<KeyT>(s, n, key_bits, L, p, p, p == 0 && first_vals == nullptr, src_k,
                             src_v, dk, dv, err, n_dev, items);
      if (rc != DMSH_OK) return rc;
      src_k = dk;
      src_v = dv;
    }
    return DMSH_OK;
  }
  if (first_src == alt_k || n_dev || first_vals || items) return DMSH_ERR_INVALID;
  const uint32_t d0 = P - kTopPasses;
  if ((rc = launch_pass<KeyT>(s, n, key_bits, L, d0, 0, true, first_src, nullptr, alt_k, alt_v,
                              err)) != DMSH_OK)
    return rc;
  if ((rc = launch_pass<KeyT>(s, n, key_bits, L, d0 + 1, 1, false, alt_k, alt_v, dst_keys,
                              dst_vals, err)) != DMSH_OK)
    return rc;
  if ((rc = launch_pass<KeyT>(s, n, key_bits, L, d0 + 2, 2, false, dst_keys, dst_vals, alt_k,
                              alt_v, err)) != DMSH_OK)
    return rc;
  uint32_t* group_count = L.tickets + kTicketWords - 1;
  const uint32_t cap = group_capacity(n);
  const KeyT kmask = key_bits >= 8 * sizeof(KeyT) ? ~(KeyT)0 : (((KeyT)1 << key_bits) - 1);
  launch("k_sort_finish", k_sort_finish<KeyT>,
         dim3(ceil_div(n, (uint64_t)kFinishThreads * kFinishItems)), dim3(kFinishThreads), 0, s,
         alt_k, alt_v, dst_keys, dst_vals, n, d0 * kRadixBits, kmask, group_count, L.groups, cap);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_sort_groups", k_sort_groups<KeyT>, dim3(std::min<uint32_t>(cap, kGroupGrid)),
         dim3(kGroupThreads), 0, s, alt_k, alt_v, dst_keys, dst_vals, n, d0 * kRadixBits, kmask,
         (const uint32_t*)group_count, (const uint32_t*)L.groups, cap);
  rc = launch_status();
  return rc;
}

// Segmented sort of composite keys (segment << seg_shift | code of
// code_bits bits; the segments contiguous and in order, so a segment's
// elements keep its range): three passes over the code's top 24 bits, from
// lo = code_bits - 24 up, each scattering inside the segments through its
// offsets (seg_off + p * seg_rows * kRadix, see SegOff), then the finish on
// the groups of equal bits at and above lo (segment included) and
// k_sort_groups.  The keys start in and end in `keys`, the element indices
// end in `vals`.  L.tickets / L.status must be zero; L.hist is not read.
template <typename KeyT, int ITEMS>
inline int run_seg_passes_items(hipStream_t s, uint32_t n, uint32_t key_bits, uint32_t code_bits,
                                const Layout& L, const uint32_t* seg_off, uint32_t seg_rows,
                                uint32_t seg_shift, KeyT* keys, uint32_t* vals, uint32_t* err) {
  if (code_bits < (uint32_t)(kTopPasses * kRadixBits) || seg_rows == 0 || !L.groups)
    return DMSH_ERR_INVALID;
  const uint32_t lo = code_bits - (uint32_t)(kTopPasses * kRadixBits);
  KeyT* alt_k = static_cast<KeyT*>(L.alt_keys);
  uint32_t* alt_v = L.alt_vals;
  const size_t rows = (size_t)seg_rows * kRadix;
  const SegOff s0{seg_off, seg_shift, seg_rows - 1}, s1{seg_off + rows, seg_shift, seg_rows - 1},
      s2{seg_off + 2 * rows, seg_shift, seg_rows - 1};
  int rc;
  if ((rc = launch_pass_items<KeyT, ITEMS, true>(s, n, key_bits, L, 0, 0, true, keys, nullptr,
                                                 alt_k, alt_v, err, nullptr, s0, lo)) != DMSH_OK)
    return rc;
  if ((rc = launch_pass_items<KeyT, ITEMS, true>(s, n, key_bits, L, 1, 1, false, alt_k, alt_v,
                                                 keys, vals, err, nullptr, s1, lo + 8)) != DMSH_OK)
    return rc;
  if ((rc = launch_pass_items<KeyT, ITEMS, true>(s, n, key_bits, L, 2, 2, false, keys, vals,
                                                 alt_k, alt_v, err, nullptr, s2, lo + 16)) !=
      DMSH_OK)
    return rc;
  uint32_t* group_count = L.tickets + kTicketWords - 1;
  const uint32_t cap = group_capacity(n);
  const KeyT kmask = key_bits >= 8 * sizeof(KeyT) ? ~(KeyT)0 : (((KeyT)1 << key_bits) - 1);
  launch("k_sort_finish", k_sort_finish<KeyT>,
         dim3(ceil_div(n, (uint64_t)kFinishThreads * kFinishItems)), dim3(kFinishThreads), 0, s,
         alt_k, alt_v, keys, vals, n, lo, kmask, group_count, L.groups, cap);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_sort_groups", k_sort_groups<KeyT>, dim3(std::min<uint32_t>(cap, kGroupGrid)),
         dim3(kGroupThreads), 0, s, alt_k, alt_v, keys, vals, n, lo, kmask,
         (const uint32_t*)group_count, (const uint32_t*)L.groups, cap);
  return launch_status();
}

// items: elements per thread of the passes (8, 12 or 16).
template <typename KeyT>
inline int run_seg_passes(hipStream_t s, uint32_t n, uint32_t key_bits, uint32_t code_bits,
                          const Layout& L, const uint32_t* seg_off, uint32_t seg_rows,
                          uint32_t seg_shift, KeyT* keys, uint32_t* vals, uint32_t* err,
                          int items) {
  switch (items) {
    case 8:
      return run_seg_passes_items<KeyT, 8>(s, n, key_bits, code_bits, L, seg_off, seg_rows,
                                           seg_shift, keys, vals, err);
    case 12:
      return run_seg_passes_items<KeyT, 12>(s, n, key_bits, code_bits, L, seg_off, seg_rows,
                                            seg_shift, keys, vals, err);
    case 16:
      return run_seg_passes_items<KeyT, 16>(s, n, key_bits, code_bits, L, seg_off, seg_rows,
                                            seg_shift, keys, vals, err);
    default:
      return DMSH_ERR_INVALID;
  }
}

// Full argsort: keys_inout (sorted on return), idx_out (permutation).
template <typename KeyT>
inline int argsort(hipStream_t s, uint32_t n, KeyT* keys_inout, uint32_t* idx_out,
                   uint32_t key_bits, void* ws, size_t ws_bytes) {
  if (n == 0) return DMSH_OK;
  if (!keys_inout || !idx_out) return DMSH_ERR_INVALID;
  if (n >= (1u << 30)) return DMSH_ERR_INVALID;
  if (key_bits == 0 || key_bits > 8 * sizeof(KeyT)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < workspace_bytes<KeyT>(n, key_bits)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  Layout L;
  if (!carve<KeyT>(a, n, key_bits, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  int rc = launch_fill(s, L.hist, L.zero_words, 0u);
  if (rc != DMSH_OK) return rc;
  const uint32_t P = num_passes(key_bits);
  const uint32_t hist_blocks = std::min<uint32_t>(num_tiles<KeyT>(n), 2048);
  launch("k_hist_all", k_hist_all<KeyT>, dim3(hist_blocks), dim3(256), 0, s, keys_inout, n,
         (const uint32_t*)nullptr, first_hist_pass(P), P, key_bits, L.hist);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  const KeyT* first_src = keys_inout;
  if (P % 2 == 1 && !uses_finish(P)) {
    // An odd pass count would end in the scratch buffer: start from a copy.
    KeyT* alt_k = static_cast<KeyT*>(L.alt_keys);
    launch("k_copy_keys", k_copy_keys<KeyT>, dim3(std::min<uint32_t>(ceil_div(n, 256), 4096)),
                       dim3(256), 0, s, keys_inout, alt_k, n, (const uint32_t*)nullptr);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    first_src = alt_k;
  }
  return run_passes<KeyT>(s, n, key_bits, L, first_src, keys_inout, idx_out, err, true);
}

// Stable sort of (keys, values) pairs whose count lives on the device
// (*n_dev <= n_cap): keys_inout and vals_inout hold the sorted pairs on
// return.  Keys of at most four digits (the plain passes); the grids and the
// workspace (workspace_bytes(n_cap, key_bits)) are sized for n_cap.
template <typename KeyT>
inline int sort_pairs_device_count(hipStream_t s, uint32_t n_cap, const uint32_t* n_dev,
                                   KeyT* keys_inout, uint32_t* vals_inout, uint32_t key_bits,
                                   void* ws, size_t ws_bytes) {
  if (n_cap == 0) return DMSH_OK;
  if (!keys_inout || !vals_inout || !n_dev) return DMSH_ERR_INVALID;
  if (n_cap >= (1u << 30)) return DMSH_ERR_INVALID;
  if (key_bits == 0 || key_bits > 8 * sizeof(KeyT)) return DMSH_ERR_INVALID;
  const uint32_t P = num_passes(key_bits);
  if (uses_finish(P)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < workspace_bytes<KeyT>(n_cap, key_bits)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  Layout L;
  if (!carve<KeyT>(a, n_cap, key_bits, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  int rc = launch_fill(s, L.hist, L.zero_words, 0u);
  if (rc != DMSH_OK) return rc;
  const uint32_t hist_blocks = std::min<uint32_t>(num_tiles<KeyT>(n_cap), 2048);
  launch("k_hist_all", k_hist_all<KeyT>, dim3(hist_blocks), dim3(256), 0, s, keys_inout, n_cap,
         n_dev, 0u, P, key_bits, L.hist);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  const KeyT* first_src = keys_inout;
  const uint32_t* first_vals = vals_inout;
  if (P % 2 == 1) {
    // an odd pass count would end in the scratch buffers: start from copies
    KeyT* alt_k = static_cast<KeyT*>(L.alt_keys);
    launch("k_copy_keys", k_copy_keys<KeyT>, dim3(std::min<uint32_t>(ceil_div(n_cap, 256), 4096)),
           dim3(256), 0, s, (const KeyT*)keys_inout, alt_k, n_cap, n_dev);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    launch("k_copy_keys", k_copy_keys<uint32_t>,
           dim3(std::min<uint32_t>(ceil_div(n_cap, 256), 4096)), dim3(256), 0, s,
           (const uint32_t*)vals_inout, L.alt_vals, n_cap, n_dev);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    first_src = alt_k;
    first_vals = L.alt_vals;
  }
  return run_passes<KeyT>(s, n_cap, key_bits, L, first_src, keys_inout, vals_inout, err, false,
                          n_dev, first_vals);
}

}  // namespace sortk
}  // namespace dmsh
