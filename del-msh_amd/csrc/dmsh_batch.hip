// dmsh_batch.hip -- C ABI: LBVH build and ray-cast of many meshes per launch.
#include <algorithm>
#include <atomic>
#include <cstring>

#include "batch.hpp"
#include "serve.hpp"

using namespace dmsh;

namespace {

constexpr uint32_t kCastCounts = 256;  // see bat::BatchCast::counts

std::atomic<int> g_batch_stack{20};      // LDS frames per lane of the batched tier 1
std::atomic<int> g_batch_wgs_per_cu{16};  // persistent ray-cast waves per CU
std::atomic<int> g_batch_sort_items{16};  // elements per thread of the digit passes
std::atomic<int> g_batch_sort_seg{1};     // segmented passes + finish (up to kSegMaxMeshes)
std::atomic<int> g_batch_cast_waves{1};   // waves per ray-cast workgroup
std::atomic<int> g_batch_treelet_grid{1024};
// a mesh's top merge: one workgroup, a thread per treelet root at a time
constexpr int kBatchTopThreads = 1024;

// key bits of the composite (mesh << 30 | code): 30 + the bits of M - 1
inline uint32_t batch_key_bits(uint32_t M) {
  uint32_t mb = 0;
  while (mb < 34 && ((uint64_t)1 << mb) < M) ++mb;
  return bat::kMeshKeyShift + mb;
}

struct BatchBuildLayout {
  uint32_t* keys6;       // [M][6]
  uint32_t* box_copies;  // [M][kBoxCopies][kBoxCopyWords]
  sortk::Layout sort;
  uint32_t* flags;          // [total_tri] top-climb counters (>= internal nodes)
  uint32_t* treelet_count;  // 1
  uint32_t* root_count;     // [M]
  uint32_t* bad;            // 1: offsets rejected
  uint32_t* top_count;      // [M] top nodes listed per mesh (zeroed with the rest)
  uint32_t* hseg;     // [3][M][256] or null (more than kSegMaxMeshes meshes)
  uint32_t* seg_off;  // [3][M][256] or null
  uint32_t* seg_chunk;  // [3][chunks of kSegChunk meshes][256] or null
  uint32_t* zero_begin;
  size_t zero_words;
  uint32_t* inner_off;  // [M + 1]
  uint4* treelets;      // [total_tri]
  uint4* top_list;      // [total_tri]: mesh m's top nodes from inner_off[m] on
  uint32_t* node2slot;  // [total_tri]: mesh m's local node -> top slot, from inner_off[m]
  uint32_t* roots;      // [total_tri]
  float* cntr;          // [total_tri][3]
  uint64_t* codes;
  uint32_t* idx;
};

inline bool carve_batch_build(Arena& a, uint32_t M, uint32_t n, BatchBuildLayout& L) {
  const uint32_t kb = batch_key_bits(M);
  const size_t nn = n ? n : 1;
  L.keys6 = a.take<uint32_t>((size_t)M * 6);
  L.box_copies = a.take<uint32_t>((size_t)M * bat::kBoxCopies * bat::kBoxCopyWords);
  L.sort.hist = a.take<uint32_t>(sortk::kMaxPasses * sortk::kRadix + sortk::kTicketWords);
  L.sort.tickets = L.sort.hist + sortk::kMaxPasses * sortk::kRadix;
  L.sort.status = L.sort.tickets + sortk::kTicketWords;
  // sized for the smallest tile, so that batch.sort_items may change
  // between the size query and the build, and for one pass per digit (the
  // batched sort takes no LDS finish, whatever the number of digits)
  const size_t status_words =
      (size_t)sortk::num_passes(kb) * sortk::num_tiles_items(n, sortk::kMinItems) * sortk::kRadix;
  a.off += sizeof(uint32_t) * status_words;
  L.flags = L.sort.status + status_words;
  a.off += sizeof(uint32_t) * nn;
  L.treelet_count = L.flags + nn;
  a.off += sizeof(uint32_t);
  L.root_count = L.treelet_count + 1;
  a.off += sizeof(uint32_t) * M;
  L.bad = L.root_count + M;
  a.off += sizeof(uint32_t);
  L.top_count = L.bad + 1;
  a.off += sizeof(uint32_t) * M;
  // per-mesh digit histograms of the segmented sort (zeroed with the rest)
  const bool seg = M <= bat::kSegMaxMeshes;
  L.hseg = seg ? L.top_count + M : nullptr;
  if (seg) a.off += sizeof(uint32_t) * (size_t)M * bat::kSegTables;
  L.zero_begin = L.sort.hist;
  L.zero_words =
      (size_t)((seg ? L.hseg + (size_t)M * bat::kSegTables : L.top_count + M) - L.sort.hist);
  L.sort.zero_words = 0;
  L.inner_off = a.take<uint32_t>((size_t)M + 1);
  L.treelets = a.take<uint4>(nn);
  L.top_list = a.take<uint4>(nn);
  L.node2slot = a.take<uint32_t>(nn);
  L.roots = a.take<uint32_t>(nn);
  L.cntr = a.take<float>(nn * 3);
  L.codes = a.take<uint64_t>(nn);
  L.idx = a.take<uint32_t>(nn);
  L.sort.alt_keys = a.take<uint64_t>(nn);
  L.sort.alt_vals = a.take<uint32_t>(nn);
  L.seg_off = seg ? a.take<uint32_t>((size_t)M * bat::kSegTables) : nullptr;
  L.seg_chunk = seg ? a.take<uint32_t>((size_t)ceil_div(M, bat::kSegChunk) * bat::kSegTables)
                    : nullptr;
  L.sort.groups = seg ? a.take<uint32_t>(sortk::group_capacity(nn)) : nullptr;
  return a.ok();
}

struct BatchCastLayout {
  uint32_t* counts;  // [kCastCounts]
  uint32_t* inner_off;
  uint32_t* tile_off;
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep;
  rc::InnerNode* inner;
  rc::LeafTri* leaf;
};

inline bool carve_batch_cast(Arena& a, uint32_t M, uint32_t n, uint32_t num_pix,
                             BatchCastLayout& L) {
  L.counts = a.take<uint32_t>(kCastCounts);
  L.inner_off = a.take<uint32_t>((size_t)M + 1);
  L.tile_off = a.take<uint32_t>((size_t)M + 1);
  L.list2 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.list3 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.deep = a.take<uint2>((size_t)rc::kDeepSlots * rc::kDeepFrames);
  L.inner = a.take<rc::InnerNode>(n ? n : 1);
  L.leaf = a.take<rc::LeafTri>(n ? n : 1);
  return a.ok();
}

template <int S>
void launch_batch_tier1(hipStream_t s, uint32_t waves, const bat::BatchCast& a) {
  switch (g_batch_cast_waves.load()) {
    case 2:
      launch("k_batch_pix2tri", bat::k_batch_pix2tri<S, 2>, dim3(ceil_div(waves, 2)), dim3(128), 0,
             s, a);
      break;
    case 4:
      launch("k_batch_pix2tri", bat::k_batch_pix2tri<S, 4>, dim3(ceil_div(waves, 4)), dim3(256), 0,
             s, a);
      break;
    default:
      launch("k_batch_pix2tri", bat::k_batch_pix2tri<S, 1>, dim3(waves), dim3(64), 0, s, a);
      break;
  }
}

// Offsets of a batch of built trees (as for the batched cast, without images).
__global__ __launch_bounds__(1024) void k_pack_offsets(uint32_t M, uint32_t total_tri,
                                                       uint32_t total_vtx,
                                                       const uint32_t* __restrict__ mesh2tri,
                                                       const uint32_t* __restrict__ mesh2vtx,
                                                       const uint32_t* __restrict__ node_off,
                                                       uint32_t* __restrict__ inner_off,
                                                       uint32_t* __restrict__ bad,
                                                       uint32_t* __restrict__ err) {
  bat::check_offsets(mesh2tri, M, total_tri, bad, err);
  bat::check_offsets(mesh2vtx, M, total_vtx, bad, err);
  for (uint32_t m = threadIdx.x; m < M; m += 1024) {
    const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
    if (node_off[m + 1] - node_off[m] != (n > 0 ? 2 * n - 1 : 0u) || (m == 0 && node_off[0] != 0u)) {
      atomicOr(bad, 1u);
      raise_error(err, DMSH_DEVERR_INDEX);
    }
  }
  bat::scan_meshes(M, inner_off, [&](uint32_t m) {
    const uint32_t n = mesh2tri[m + 1] - mesh2tri[m];
    return n > 0 ? n - 1 : 0u;
  });
}

struct ReqLayout {
  uint32_t* counts;
  uint32_t* tile_off;
  uint32_t* pix_off;
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep;
};

inline bool carve_req(Arena& a, uint32_t R, uint32_t num_pix, ReqLayout& L) {
  L.counts = a.take<uint32_t>(kCastCounts);
  L.tile_off = a.take<uint32_t>((size_t)R + 1);
  L.pix_off = a.take<uint32_t>((size_t)R + 1);
  L.list2 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.list3 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.deep = a.take<uint2>((size_t)rc::kDeepSlots * rc::kDeepFrames);
  return a.ok();
}

template <int S>
void launch_req_tier1(hipStream_t s, uint32_t grid, const srv::ReqCast& a) {
  launch("k_req_pix2tri", srv::k_req_pix2tri<S>, dim3(grid), dim3(64), 0, s, a);
}

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 256;
  return cus > 0 ? cus : 256;
}

}  // namespace

extern "C" {

int dmsh_batch_set_tuning(const char* key, int value) {
  if (!key) return DMSH_ERR_INVALID;
  if (std::strcmp(key, "batch.stack") == 0) {
    if (value != 16 && value != 20 && value != 24 && value != 32) return DMSH_ERR_INVALID;
    g_batch_stack.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "batch.wgs_per_cu") == 0) {
    if (value < 1 || value > 64) return DMSH_ERR_INVALID;
    g_batch_wgs_per_cu.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "batch.treelet_grid") == 0) {
    if (value < 1 || value > 65536) return DMSH_ERR_INVALID;
    g_batch_treelet_grid.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "batch.sort_items") == 0) {
    if (!sortk::valid_items(value)) return DMSH_ERR_INVALID;
    g_batch_sort_items.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "batch.cast_waves") == 0) {
    if (value != 1 && value != 2 && value != 4) return DMSH_ERR_INVALID;
    g_batch_cast_waves.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "batch.sort_seg") == 0) {
    if (value != 0 && value != 1) return DMSH_ERR_INVALID;
    g_batch_sort_seg.store(value);
    return DMSH_OK;
  }
  return DMSH_ERR_INVALID;
}

int dmsh_batch_get_tuning(const char* key) {
  if (!key) return -1;
  if (std::strcmp(key, "batch.stack") == 0) return g_batch_stack.load();
  if (std::strcmp(key, "batch.wgs_per_cu") == 0) return g_batch_wgs_per_cu.load();
  if (std::strcmp(key, "batch.treelet_grid") == 0) return g_batch_treelet_grid.load();
  if (std::strcmp(key, "batch.sort_items") == 0) return g_batch_sort_items.load();
  if (std::strcmp(key, "batch.cast_waves") == 0) return g_batch_cast_waves.load();
  if (std::strcmp(key, "batch.sort_seg") == 0) return g_batch_sort_seg.load();
  return -1;
}

size_t dmsh_batch_build_workspace_bytes(uint32_t num_mesh, uint32_t total_tri) {
  Arena a{nullptr, ~(size_t)0, 0};
  BatchBuildLayout L;
  carve_batch_build(a, num_mesh, total_tri, L);
  return a.off + 256;
}

extern "C++" {
namespace {

// The treelet and top refits of a batched build; Src: where the nodes'
// traversal records go as they are fitted (bat::NoPackSrc: nowhere).
template <typename Src>
int launch_batch_refit(hipStream_t s, const bat::BatchTrees& bt, const BatchBuildLayout& L,
                       uint32_t M, uint32_t n, uint32_t* err, Src src) {
  const uint32_t tgrid = std::min<uint32_t>((uint32_t)g_batch_treelet_grid.load(), n);
  if (refit_option(kRefitClimb) == 2) {
    launch("k_refit_treelet",
           (bat::k_batch_refit_treelet<(int)kTreeletLeaves, kTreeletThreads, 2, Src>),
           dim3(tgrid), dim3(kTreeletThreads), 0, s, bt, n, L.treelet_count, L.treelets, err,
           src);
  } else {
    launch("k_refit_treelet",
           (bat::k_batch_refit_treelet<(int)kTreeletLeaves, kTreeletThreads, 0, Src>),
           dim3(tgrid), dim3(kTreeletThreads), 0, s, bt, n, L.treelet_count, L.treelets, err,
           src);
  }
  int rc = launch_status();
  if (rc != DMSH_OK) return rc;
  launch("k_refit_top", (bat::k_batch_refit_top<(int)kTreeletLeaves, kBatchTopThreads, Src>),
         dim3(std::min<uint32_t>(M, 1024)), dim3(kBatchTopThreads), 0, s, bt, L.inner_off,
         L.flags, L.root_count, L.roots, err, src, (const uint32_t*)L.top_count,
         (const uint4*)L.top_list, (const uint32_t*)L.node2slot);
  return launch_status();
}

// The batched build; with `packed` (and `nonfinite`) the refit also writes
// every mesh's packed traversal buffer, as dmsh_batch_pack would.
int batch_build(void* stream, uint32_t num_mesh, uint32_t total_tri, uint32_t total_vtx,
                const uint32_t* mesh2tri, const uint32_t* mesh2vtx, const uint32_t* tri2vtx,
                const float* vtx2xyz, uint32_t* mesh2node, uint32_t* bvhnodes,
                float* bvhnode2aabb, void* ws, size_t ws_bytes, const uint64_t* packed,
                uint32_t* nonfinite) {
  if (num_mesh == 0 || !mesh2tri || !mesh2vtx || !mesh2node) return DMSH_ERR_INVALID;
  if (total_tri >= (1u << 30)) return DMSH_ERR_INVALID;
  if (total_tri > 0 && (!tri2vtx || !vtx2xyz || !bvhnodes || !bvhnode2aabb || total_vtx == 0))
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_batch_build_workspace_bytes(num_mesh, total_tri))
    return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  BatchBuildLayout L;
  if (!carve_batch_build(ar, num_mesh, total_tri, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  const uint32_t M = num_mesh, n = total_tri;
  const uint32_t kb = batch_key_bits(M);
  const uint32_t P = sortk::num_passes(kb);
  int rc;
  if (packed && (rc = sortk::launch_fill(s, nonfinite, M, 0u)) != DMSH_OK) return rc;
  launch("k_batch_init", bat::k_batch_init, dim3(std::min<uint32_t>(ceil_div(L.zero_words, 256), 2048)),
         dim3(256), 0, s, M, L.box_copies, L.zero_begin, L.zero_words);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_batch_offsets", bat::k_batch_offsets, dim3(1), dim3(1024), 0, s, M, n, total_vtx,
         mesh2tri, mesh2vtx, L.inner_off, mesh2node, L.bad, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  if (n == 0) return DMSH_OK;
  launch("k_batch_centroid", bat::k_batch_centroid,
         dim3(ceil_div(n, 4u * 64u * bat::kCentroidChunks)), dim3(256), 0, s, L.bad, n, M, mesh2tri, mesh2vtx, tri2vtx, vtx2xyz, L.cntr, L.box_copies, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_batch_box_reduce", bat::k_batch_box_reduce, dim3(ceil_div((uint64_t)M * 6, 256)),
         dim3(256), 0, s, (const uint32_t*)L.bad, M, (const uint32_t*)L.box_copies, L.keys6);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  if (L.seg_off && g_batch_sort_seg.load() == 1) {
    // segmented: the mesh bits never move a key, so three passes over the
    // code's top 24 bits inside the meshes and the finish on the low 6
    launch("k_morton_hist", bat::k_batch_morton_seghist,
           dim3(std::min<uint32_t>(ceil_div(n, 2048), 2048)), dim3(256), 0, s, L.bad, n, M,
           mesh2tri, L.cntr, L.keys6, L.codes, L.hseg);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    const dim3 sgrid(ceil_div(M, bat::kSegChunk), sortk::kTopPasses);
    launch("k_seg_offsets", bat::k_batch_seg_rows, sgrid, dim3(bat::kSegRowThreads), 0, s, L.bad,
           M, mesh2tri, L.hseg, L.seg_off, L.seg_chunk);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    if (sgrid.x > 1) {  // with one chunk the rows kernel finished the offsets
      launch("k_seg_offsets", bat::k_batch_seg_cols, sgrid, dim3(256), 0, s, L.bad, M,
             L.seg_off, (const uint32_t*)L.seg_chunk);
      if ((rc = launch_status()) != DMSH_OK) return rc;
    }
    rc = sortk::run_seg_passes<uint64_t>(s, n, kb, bat::kMeshKeyShift, L.sort, L.seg_off, M,
                                         bat::kMeshKeyShift, L.codes, L.idx, err,
                                         g_batch_sort_items.load());
  } else {
    // one pass per digit of the composite key (the top digits are mostly the
    // mesh index, so groups of equal top bits are too long for the finish);
    // with an even number of passes the sort returns to the buffer it starts
    // from, with an odd number it starts from the alternate key buffer
    uint64_t* morton_out = (P % 2 == 0) ? L.codes : static_cast<uint64_t*>(L.sort.alt_keys);
    launch("k_morton_hist", bat::k_batch_morton_hist,
           dim3(std::min<uint32_t>(sortk::num_tiles<uint64_t>(n), 2048)), dim3(256), 0, s, L.bad,
           n, M, mesh2tri, L.cntr, L.keys6, morton_out, P, kb, L.sort.hist);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    rc = sortk::run_passes<uint64_t>(s, n, kb, L.sort, morton_out, L.codes, L.idx, err, false,
                                     nullptr, nullptr, g_batch_sort_items.load());
  }
  if (rc != DMSH_OK) return rc;
  TreeletSink ts{kTreeletLeaves, L.treelet_count, L.treelets, n, L.root_count, L.roots};
  ts.mesh_top_count = L.top_count;
  ts.mesh_top_list = L.top_list;
  ts.mesh_node2slot = L.node2slot;
  launch("k_topology", bat::k_batch_topology, dim3(ceil_div(n, kTopoThreads)), dim3(kTopoThreads),
         0, s, L.bad, n,
         M, mesh2tri, L.inner_off, mesh2node, bvhnodes, L.codes, L.idx, ts);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_batch_singles", bat::k_batch_singles, dim3(ceil_div(M, 256)), dim3(256), 0, s, L.bad,
         M, mesh2tri, mesh2node, bvhnodes, L.idx, ts);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  const bat::BatchTrees bt{L.bad,    M,        mesh2tri, mesh2vtx, mesh2node,
                           bvhnode2aabb, bvhnodes, tri2vtx, vtx2xyz};
  if (!packed) return launch_batch_refit(s, bt, L, M, n, err, bat::NoPackSrc{});
  if ((rc = launch_batch_refit(s, bt, L, M, n, err, srv::PackSrc{packed, nonfinite})) != DMSH_OK)
    return rc;
  launch("k_pack_roots", srv::k_pack_roots, dim3(ceil_div(M, 256)), dim3(256), 0, s, M, mesh2tri,
         (const uint32_t*)mesh2node, (const float*)bvhnode2aabb, packed,
         (const uint32_t*)nonfinite, (const uint32_t*)L.bad);
  return launch_status();
}

}  // namespace
}  // extern "C++"

int dmsh_batch_build_trimesh3(void* stream, uint32_t num_mesh, uint32_t total_tri,
                              uint32_t total_vtx, const uint32_t* mesh2tri,
                              const uint32_t* mesh2vtx, const uint32_t* tri2vtx,
                              const float* vtx2xyz, uint32_t* mesh2node, uint32_t* bvhnodes,
                              float* bvhnode2aabb, void* ws, size_t ws_bytes) {
  return batch_build(stream, num_mesh, total_tri, total_vtx, mesh2tri, mesh2vtx, tri2vtx, vtx2xyz,
                     mesh2node, bvhnodes, bvhnode2aabb, ws, ws_bytes, nullptr, nullptr);
}

int dmsh_batch_build_pack_trimesh3(void* stream, uint32_t num_mesh, uint32_t total_tri,
                                   uint32_t total_vtx, const uint32_t* mesh2tri,
                                   const uint32_t* mesh2vtx, const uint32_t* tri2vtx,
                                   const float* vtx2xyz, uint32_t* mesh2node, uint32_t* bvhnodes,
                                   float* bvhnode2aabb, const uint64_t* packed,
                                   uint32_t* nonfinite, void* ws, size_t ws_bytes) {
  if (!packed || !nonfinite) return DMSH_ERR_INVALID;
  if (total_tri >= rc::kMaxTri) return DMSH_ERR_INVALID;
  return batch_build(stream, num_mesh, total_tri, total_vtx, mesh2tri, mesh2vtx, tri2vtx, vtx2xyz,
                     mesh2node, bvhnodes, bvhnode2aabb, ws, ws_bytes, packed, nonfinite);
}

size_t dmsh_batch_pix2tri_workspace_bytes(uint32_t num_mesh, uint32_t total_tri,
                                          uint32_t total_pix) {
  Arena a{nullptr, ~(size_t)0, 0};
  BatchCastLayout L;
  carve_batch_cast(a, num_mesh, total_tri, total_pix, L);
  return a.off + 256;
}

int dmsh_batch_pix2tri(void* stream, uint32_t num_mesh, uint32_t total_tri, uint32_t total_vtx,
                       uint32_t total_pix, const uint32_t* mesh2tri, const uint32_t* mesh2vtx,
                       const uint32_t* tri2vtx, const float* vtx2xyz, const uint32_t* mesh2node,
                       const uint32_t* bvhnodes, const float* bvhnode2aabb,
                       const uint32_t* mesh2pix, const uint32_t* mesh2width,
                       const float* transforms_ndc2world, uint32_t* pix2tri, float* pix2t,
                       void* ws, size_t ws_bytes) {
  if (num_mesh == 0 || !mesh2tri || !mesh2vtx || !mesh2node || !mesh2pix || !mesh2width ||
      !transforms_ndc2world)
    return DMSH_ERR_INVALID;
  if (total_tri >= rc::kMaxTri || total_pix == 0) return total_pix == 0 ? DMSH_OK : DMSH_ERR_INVALID;
  if (!pix2tri) return DMSH_ERR_INVALID;
  if (total_tri > 0 && (!tri2vtx || !vtx2xyz || !bvhnodes || !bvhnode2aabb))
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_batch_pix2tri_workspace_bytes(num_mesh, total_tri, total_pix))
    return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  BatchCastLayout L;
  if (!carve_batch_cast(ar, num_mesh, total_tri, total_pix, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  const uint32_t M = num_mesh, n = total_tri;
  int rc = sortk::launch_fill(s, L.counts, kCastCounts, 0u);
  if (rc != DMSH_OK) return rc;
  launch("k_batch_cast_offsets", bat::k_batch_cast_offsets, dim3(1), dim3(1024), 0, s, M, n,
         total_vtx, total_pix, mesh2tri, mesh2vtx, mesh2node, mesh2pix, mesh2width, L.inner_off,
         L.tile_off, L.counts + 3, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  const bat::BatchTrees bt{L.counts + 3, M, mesh2tri, mesh2vtx, mesh2node,
                           const_cast<float*>(bvhnode2aabb), bvhnodes, tri2vtx, vtx2xyz};
  if (n > 0) {
    launch("k_pack_inner", bat::k_batch_pack_inner, dim3(ceil_div(n, 256)), dim3(256), 0, s, bt,
           n, L.inner_off, L.inner, L.counts + 2, err);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    launch("k_pack_leaf", bat::k_batch_pack_leaf, dim3(ceil_div(n, 256)), dim3(256), 0, s, bt, n,
           L.leaf, err);
    if ((rc = launch_status()) != DMSH_OK) return rc;
  }
  bat::BatchCast a;
  a.M = M;
  a.mesh2tri = mesh2tri;
  a.node_off = mesh2node;
  a.inner_off = L.inner_off;
  a.tile_off = L.tile_off;
  a.pix_off = mesh2pix;
  a.widths = mesh2width;
  a.inner = L.inner;
  a.leaf = L.leaf;
  a.aabb = bvhnode2aabb;
  a.transforms = transforms_ndc2world;
  a.pix2tri = pix2tri;
  a.pix2t = pix2t;
  a.counts = L.counts;
  a.list2 = L.list2;
  a.list3 = L.list3;
  a.deep_stack = L.deep;
  a.err = err;
  const uint32_t grid = (uint32_t)device_cus() * (uint32_t)g_batch_wgs_per_cu.load();
  switch (g_batch_stack.load()) {
    case 16: launch_batch_tier1<16>(s, grid, a); break;
    case 20: launch_batch_tier1<20>(s, grid, a); break;
    case 32: launch_batch_tier1<32>(s, grid, a); break;
    default: launch_batch_tier1<24>(s, grid, a); break;
  }
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_batch_pix2tri_tier2", bat::k_batch_pix2tri_tier2, dim3(rc::kTier2Blocks), dim3(64), 0,
         s, a, total_pix);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_batch_pix2tri_tier3", bat::k_batch_pix2tri_tier3, dim3(rc::kDeepSlots / 64), dim3(64),
         0, s, a, total_pix);
  return launch_status();
}

size_t dmsh_packed_bytes(uint32_t num_tri) { return srv::packed_bytes(num_tri); }

size_t dmsh_batch_pack_workspace_bytes(uint32_t num_mesh) {
  Arena a{nullptr, ~(size_t)0, 0};
  a.take<uint32_t>((size_t)num_mesh + 1);
  a.take<uint32_t>(64);
  return a.off + 256;
}

int dmsh_batch_pack(void* stream, uint32_t num_mesh, uint32_t total_tri, uint32_t total_vtx,
                    const uint32_t* mesh2tri, const uint32_t* mesh2vtx, const uint32_t* tri2vtx,
                    const float* vtx2xyz, const uint32_t* mesh2node, const uint32_t* bvhnodes,
                    const float* bvhnode2aabb, const uint64_t* packed, uint32_t* nonfinite,
                    void* ws, size_t ws_bytes) {
  if (num_mesh == 0 || !mesh2tri || !mesh2vtx || !mesh2node || !packed || !nonfinite)
    return DMSH_ERR_INVALID;
  if (total_tri >= rc::kMaxTri) return DMSH_ERR_INVALID;
  if (total_tri > 0 && (!tri2vtx || !vtx2xyz || !bvhnodes || !bvhnode2aabb))
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_batch_pack_workspace_bytes(num_mesh)) return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  uint32_t* inner_off = ar.take<uint32_t>((size_t)num_mesh + 1);
  uint32_t* bad = ar.take<uint32_t>(64);
  if (!ar.ok()) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  int rc = sortk::launch_fill(s, bad, 1, 0u);
  if (rc != DMSH_OK) return rc;
  if ((rc = sortk::launch_fill(s, nonfinite, num_mesh, 0u)) != DMSH_OK) return rc;
  launch("k_pack_offsets", k_pack_offsets, dim3(1), dim3(1024), 0, s, num_mesh, total_tri,
         total_vtx, mesh2tri, mesh2vtx, mesh2node, inner_off, bad, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  if (total_tri == 0) return DMSH_OK;
  const bat::BatchTrees bt{bad, num_mesh, mesh2tri, mesh2vtx, mesh2node,
                           const_cast<float*>(bvhnode2aabb), bvhnodes, tri2vtx, vtx2xyz};
  launch("k_pack_to_buffers", srv::k_pack_to_buffers, dim3(ceil_div(total_tri, 256)), dim3(256),
         0, s, bt, total_tri, inner_off, packed, nonfinite, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_pack_flags", srv::k_pack_flags, dim3(ceil_div(num_mesh, 256)), dim3(256), 0, s,
         num_mesh, mesh2tri, packed, nonfinite, bad);
  return launch_status();
}

size_t dmsh_cast_requests_workspace_bytes(uint32_t num_req, uint32_t total_pix) {
  Arena a{nullptr, ~(size_t)0, 0};
  ReqLayout L;
  carve_req(a, num_req, total_pix, L);
  return a.off + 256;
}

int dmsh_cast_requests(void* stream, uint32_t num_req, const dmsh_cast_request* requests,
                       uint32_t total_pix, uint32_t* pix2tri, float* pix2t, void* ws,
                       size_t ws_bytes) {
  if (num_req == 0 || total_pix == 0) return DMSH_OK;
  if (!requests || !pix2tri) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_cast_requests_workspace_bytes(num_req, total_pix))
    return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  ReqLayout L;
  if (!carve_req(ar, num_req, total_pix, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  int rc = sortk::launch_fill(s, L.counts, kCastCounts, 0u);
  if (rc != DMSH_OK) return rc;
  const srv::Request* req = reinterpret_cast<const srv::Request*>(requests);
  launch("k_req_offsets", srv::k_req_offsets, dim3(1), dim3(1024), 0, s, num_req, total_pix, req,
         L.tile_off, L.pix_off, L.counts + 3, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  srv::ReqCast a{num_req, req, L.tile_off, L.pix_off, pix2tri, pix2t, L.counts,
                 L.list2, L.list3, L.deep, err};
  const uint32_t grid = (uint32_t)device_cus() * (uint32_t)g_batch_wgs_per_cu.load();
  switch (g_batch_stack.load()) {
    case 16: launch_req_tier1<16>(s, grid, a); break;
    case 20: launch_req_tier1<20>(s, grid, a); break;
    case 32: launch_req_tier1<32>(s, grid, a); break;
    default: launch_req_tier1<24>(s, grid, a); break;
  }
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_req_pix2tri_tier2", srv::k_req_pix2tri_tier2, dim3(rc::kTier2Blocks), dim3(64), 0, s,
         a, total_pix);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_req_pix2tri_tier3", srv::k_req_pix2tri_tier3, dim3(rc::kDeepSlots / 64), dim3(64), 0,
         s, a, total_pix);
  return launch_status();
}

}  // extern "C"
