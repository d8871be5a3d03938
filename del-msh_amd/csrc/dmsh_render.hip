// dmsh_render.hip -- C ABI: ray-cast (pix2tri + hit t), depth, edge-gradient backward.
#include <algorithm>
#include <atomic>
#include <cstring>

#include "edgegrad.hpp"
#include "raycast.hpp"
#include "sort.hpp"
#include "topo.hpp"

using namespace dmsh;

namespace {

// Defaults from a sweep on the 2^21-triangle / 2048^2-ray workload (MI355X),
// whose deepest ray needs 24 frames: 16 frames x 1 wave 3.36 ms (some rays
// restart in tier 2), 24 x 1 3.37 ms (none do), 24 x 2 3.53 ms, 32 x 1
// 3.98 ms, 32 x 4 5.12 ms.
std::atomic<int> g_cast_stack{20};  // LDS frames per lane (tier 1)
std::atomic<int> g_cast_waves{1};   // waves per workgroup (tier 1)

struct CastLayout {
  uint32_t* counts;
  uint32_t* list2;
  uint32_t* list3;
  uint2* deep;
  rc::InnerNode* inner;
  rc::LeafTri* leaf;
};

inline bool carve_cast(Arena& a, uint32_t num_tri, uint32_t num_pix, CastLayout& L) {
  L.counts = a.take<uint32_t>(64);
  L.list2 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.list3 = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.deep = a.take<uint2>((size_t)rc::kDeepSlots * rc::kDeepFrames);
  L.inner = a.take<rc::InnerNode>(num_tri > 1 ? num_tri - 1 : 1);
  L.leaf = a.take<rc::LeafTri>(num_tri ? num_tri : 1);
  return a.ok();
}

template <int S, int WAVES>
void launch_tier1(hipStream_t s, const rc::Packed& pk, const float* T, const rc::CastArgs& a) {
  launch("k_pix2tri", (rc::k_pix2tri<S, WAVES>), dim3(ceil_div(a.num_tiles, WAVES)),
                     dim3(64 * WAVES), 0, s, pk, T, a);
}

template <int S>
bool launch_tier1_waves(int waves, hipStream_t s, const rc::Packed& pk, const float* T,
                        const rc::CastArgs& a) {
  switch (waves) {
    case 1: launch_tier1<S, 1>(s, pk, T, a); return true;
    case 2: launch_tier1<S, 2>(s, pk, T, a); return true;
    case 4: launch_tier1<S, 4>(s, pk, T, a); return true;
    default: return false;
  }
}

bool launch_tier1_any(int stack, int waves, hipStream_t s, const rc::Packed& pk, const float* T,
                      const rc::CastArgs& a) {
  switch (stack) {
    case 12: return launch_tier1_waves<12>(waves, s, pk, T, a);
    case 16: return launch_tier1_waves<16>(waves, s, pk, T, a);
    case 20: return launch_tier1_waves<20>(waves, s, pk, T, a);
    case 24: return launch_tier1_waves<24>(waves, s, pk, T, a);
    case 32: return launch_tier1_waves<32>(waves, s, pk, T, a);
    default: return false;
  }
}

}  // namespace

extern "C" {

int dmsh_set_tuning(const char* key, int value) {
  if (!key) return DMSH_ERR_INVALID;
  if (std::strcmp(key, "pix2tri.stack") == 0) {
    if (value != 12 && value != 16 && value != 20 && value != 24 && value != 32)
      return DMSH_ERR_INVALID;
    g_cast_stack.store(value);
    return DMSH_OK;
  }
  if (std::strcmp(key, "pix2tri.waves") == 0) {
    if (value != 1 && value != 2 && value != 4) return DMSH_ERR_INVALID;
    g_cast_waves.store(value);
    return DMSH_OK;
  }
  return set_refit_option(key, value) ? DMSH_OK : DMSH_ERR_INVALID;
}

int dmsh_get_tuning(const char* key) {
  if (!key) return -1;
  if (std::strcmp(key, "pix2tri.stack") == 0) return g_cast_stack.load();
  if (std::strcmp(key, "pix2tri.waves") == 0) return g_cast_waves.load();
  return get_refit_option(key);
}

size_t dmsh_pix2tri_workspace_bytes(uint32_t num_tri, uint32_t img_w, uint32_t img_h) {
  Arena a{nullptr, 0, 0};
  CastLayout L;
  carve_cast(a, num_tri, img_w * img_h, L);
  return a.off + 256;
}

int dmsh_pix2tri(void* stream, uint32_t* pix2tri, float* pix2t, uint32_t num_tri,
                 const uint32_t* tri2vtx, uint32_t num_vtx, const float* vtx2xyz, uint32_t img_w,
                 uint32_t img_h, const float* transform_ndc2world, const uint32_t* bvhnodes,
                 const float* bvhnode2aabb, void* ws, size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || !pix2tri || !transform_ndc2world) return DMSH_ERR_INVALID;
  if (num_tri >= rc::kMaxTri) return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz || !bvhnodes || !bvhnode2aabb)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_pix2tri_workspace_bytes(num_tri, img_w, img_h))
    return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  CastLayout L;
  if (!carve_cast(ar, num_tri, (uint32_t)num_pix, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  int st = sortk::launch_fill(s, L.counts, 3, 0u);
  if (st != DMSH_OK) return st;
  if (num_tri > 0) {
    launch("k_pack_tree", rc::k_pack_tree, dim3(ceil_div(2 * (uint64_t)num_tri - 1, 256)),
           dim3(256), 0, s, L.inner, L.leaf, bvhnodes, bvhnode2aabb, tri2vtx, vtx2xyz, num_tri,
           num_vtx, L.counts + 2, err);
    if ((st = launch_status()) != DMSH_OK) return st;
  }
  rc::Packed pk{L.inner, L.leaf, bvhnode2aabb, num_tri};
  rc::CastArgs a;
  a.pix2tri = pix2tri;
  a.pix2t = pix2t;
  a.img_w = img_w;
  a.img_h = img_h;
  a.tiles_x = ceil_div(img_w, 8);
  a.num_tiles = a.tiles_x * ceil_div(img_h, 8);
  a.counts = L.counts;
  a.list2 = L.list2;
  a.list3 = L.list3;
  a.deep_stack = L.deep;
  a.err = err;
  if (!launch_tier1_any(g_cast_stack.load(), g_cast_waves.load(), s, pk, transform_ndc2world, a))
    return DMSH_ERR_INVALID;
  if ((st = launch_status()) != DMSH_OK) return st;
  launch("k_pix2tri_tier2", rc::k_pix2tri_tier2, dim3(rc::kTier2Blocks), dim3(64), 0, s, pk,
                     transform_ndc2world, a);
  if ((st = launch_status()) != DMSH_OK) return st;
  launch("k_pix2tri_tier3", rc::k_pix2tri_tier3, dim3(rc::kDeepSlots / 64), dim3(64), 0, s, pk,
                     transform_ndc2world, a);
  return launch_status();
}

size_t dmsh_pix2depth_workspace_bytes(void) { return 256; }

int dmsh_pix2depth(void* stream, float* pix2depth, const uint32_t* pix2tri, uint32_t num_tri,
                   const uint32_t* tri2vtx, uint32_t num_vtx, const float* vtx2xyz,
                   uint32_t img_w, uint32_t img_h, const float* transform_ndc2world, void* ws,
                   size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || !pix2depth || !pix2tri || !transform_ndc2world)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_pix2depth_workspace_bytes()) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  float* inv = static_cast<float*>(ws);
  launch("k_mat4_inverse", k_mat4_inverse, dim3(1), dim3(1), 0, s, transform_ndc2world, inv, err);
  int st = launch_status();
  if (st != DMSH_OK) return st;
  rc::TriMesh sc{tri2vtx, vtx2xyz, num_tri, num_vtx};
  launch("k_pix2depth", rc::k_pix2depth, dim3(ceil_div(num_pix, 256)), dim3(256), 0, s, pix2depth,
                     pix2tri, sc, img_w, img_h, transform_ndc2world, inv, err);
  return launch_status();
}

int dmsh_pix2depth_bwd_wrt_vtx2xyz(void* stream, float* dldw_vtx2xyz, const uint32_t* pix2tri,
                                   uint32_t num_tri, const uint32_t* tri2vtx, uint32_t num_vtx,
                                   const float* vtx2xyz, const float* dldw_pix2depth,
                                   uint32_t img_w, uint32_t img_h,
                                   const float* transform_ndc2world, void* ws, size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || !dldw_vtx2xyz || !pix2tri || !dldw_pix2depth ||
      !transform_ndc2world)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_pix2depth_workspace_bytes()) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  float* inv = static_cast<float*>(ws);
  launch("k_mat4_inverse", k_mat4_inverse, dim3(1), dim3(1), 0, s, transform_ndc2world, inv, err);
  int st = launch_status();
  if (st != DMSH_OK) return st;
  rc::TriMesh sc{tri2vtx, vtx2xyz, num_tri, num_vtx};
  launch("k_pix2depth_bwd", rc::k_pix2depth_bwd, dim3(ceil_div(num_pix, 256)), dim3(256), 0, s,
         dldw_vtx2xyz, pix2tri, sc, dldw_pix2depth, img_w, img_h, transform_ndc2world, inv, err);
  return launch_status();
}

int dmsh_interpolate(void* stream, const uint32_t* pix2tri, uint32_t img_w, uint32_t img_h,
                     uint32_t num_tri, const uint32_t* tri2vtx, uint32_t num_vtx,
                     const float* vtx2xyz, uint32_t num_vdim, const float* vtx2val,
                     const float* transform_ndc2world, float* pix2val) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || num_vdim == 0 || !pix2tri || !transform_ndc2world || !pix2val)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  if (num_vtx > 0 && !vtx2val) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  rc::TriMesh sc{tri2vtx, vtx2xyz, num_tri, num_vtx};
  launch("k_interpolate", rc::k_interpolate, dim3(ceil_div(num_pix, 256)), dim3(256), 0,
         as_stream(stream), pix2val, pix2tri, sc, num_vdim, vtx2val, img_w, img_h,
         transform_ndc2world, err);
  return launch_status();
}

int dmsh_interpolate_bwd(void* stream, const uint32_t* pix2tri, uint32_t img_w, uint32_t img_h,
                         uint32_t num_tri, const uint32_t* tri2vtx, uint32_t num_vtx,
                         const float* vtx2xyz, uint32_t num_vdim, const float* vtx2val,
                         const float* transform_ndc2world, const float* dldw_pix2val,
                         float* dldw_vtx2xyz, float* dldw_vtx2val) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || num_vdim == 0 || !pix2tri || !transform_ndc2world ||
      !dldw_pix2val)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  if (num_vtx > 0 && (!vtx2val || !dldw_vtx2xyz || !dldw_vtx2val)) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  rc::TriMesh sc{tri2vtx, vtx2xyz, num_tri, num_vtx};
  launch("k_interpolate_bwd", rc::k_interpolate_bwd, dim3(ceil_div(num_pix, 256)), dim3(256), 0,
         as_stream(stream), dldw_vtx2xyz, dldw_vtx2val, pix2tri, sc, num_vdim, vtx2val,
         dldw_pix2val, img_w, img_h, transform_ndc2world, err);
  return launch_status();
}

int dmsh_rasterized_edge_gradient_bwd(void* stream, const uint32_t* tri2vtx, uint32_t num_tri,
                                      const float* vtx2xyz, uint32_t num_vtx,
                                      float* dldw_vtx2xyz, const float* transform_world2pix,
                                      uint32_t img_w, uint32_t img_h, const uint32_t* pix2tri,
                                      uint32_t num_vdim, const float* pix2val,
                                      const float* dldw_pix2val) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || num_vdim == 0) return DMSH_ERR_INVALID;
  if (!tri2vtx || !vtx2xyz || !dldw_vtx2xyz || !transform_world2pix || !pix2tri || !pix2val ||
      !dldw_pix2val)
    return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_edge_grad_bwd", eg::k_edge_grad_bwd, dim3(ceil_div(img_w, 64), ceil_div(img_h, 4)),
                     dim3(64, 4), 0, as_stream(stream), tri2vtx, num_tri, vtx2xyz, num_vtx,
                     dldw_vtx2xyz, transform_world2pix, img_w, img_h, pix2tri, num_vdim, pix2val,
                     dldw_pix2val, err);
  return launch_status();
}


namespace {
// Workspace of the ordered edge-gradient backward for P pixel pairs:
// offsets [P + 1], scan tile sums, then per record slot (3 per pair) the
// vertex id, the record index and the term, and the sort's workspace.
struct EgLayout {
  uint32_t* off;
  uint32_t* sums;
  uint32_t* rec_vtx;
  uint32_t* rec_idx;
  float* rec_term;
  void* sort_ws;
  size_t sort_bytes;
};

uint32_t eg_num_pairs(uint32_t img_w, uint32_t img_h) {
  return img_w * (img_h - 1u) + (img_w - 1u) * img_h;
}

uint32_t eg_key_bits(uint32_t num_vtx) {
  uint32_t b = 1;
  while (b < 32 && (num_vtx - 1u) >> b) ++b;
  return b;
}

bool eg_carve(Arena& a, uint32_t num_pair, uint32_t num_vtx, EgLayout& L) {
  const uint32_t cap = 3u * num_pair;
  L.off = a.take<uint32_t>((size_t)num_pair + 1);
  L.sums = a.take<uint32_t>(topo::scan_tiles(num_pair));
  L.rec_vtx = a.take<uint32_t>(cap);
  L.rec_idx = a.take<uint32_t>(cap);
  L.rec_term = a.take<float>((size_t)cap * 3);
  L.sort_bytes = sortk::workspace_bytes<uint32_t>(cap, eg_key_bits(num_vtx));
  L.sort_ws = a.take<char>(L.sort_bytes);
  return a.ok();
}

bool eg_size_ok(uint32_t img_w, uint32_t img_h) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  // 3 record slots per pair must stay below the sort's 2^30 elements
  return num_pix < (1ull << 27);
}
}  // namespace

size_t dmsh_rasterized_edge_gradient_bwd_workspace_bytes(uint32_t img_w, uint32_t img_h,
                                                         uint32_t num_vtx) {
  if (img_w == 0 || img_h == 0 || num_vtx == 0 || !eg_size_ok(img_w, img_h)) return 256;
  Arena a{nullptr, 0, 0};
  EgLayout L;
  eg_carve(a, eg_num_pairs(img_w, img_h), num_vtx, L);
  return a.off + 256;
}

int dmsh_rasterized_edge_gradient_bwd_ordered(
    void* stream, const uint32_t* tri2vtx, uint32_t num_tri, const float* vtx2xyz,
    uint32_t num_vtx, float* dldw_vtx2xyz, const float* transform_world2pix, uint32_t img_w,
    uint32_t img_h, const uint32_t* pix2tri, uint32_t num_vdim, const float* pix2val,
    const float* dldw_pix2val, void* ws, size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0 || num_vtx == 0) return DMSH_OK;
  if (!eg_size_ok(img_w, img_h) || num_vdim == 0) return DMSH_ERR_INVALID;
  if (!tri2vtx || !vtx2xyz || !dldw_vtx2xyz || !transform_world2pix || !pix2tri || !pix2val ||
      !dldw_pix2val)
    return DMSH_ERR_INVALID;
  const uint32_t num_pair = eg_num_pairs(img_w, img_h);
  if (num_pair == 0) return DMSH_OK;
  if (!ws || ws_bytes < dmsh_rasterized_edge_gradient_bwd_workspace_bytes(img_w, img_h, num_vtx))
    return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  EgLayout L;
  if (!eg_carve(a, num_pair, num_vtx, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  const uint32_t cap = 3u * num_pair;
  const uint32_t blocks = std::min<uint32_t>(ceil_div(num_pair, 256), 16384);
  launch("k_eg_count", eg::k_eg_count, dim3(blocks), dim3(256), 0, s, tri2vtx, num_tri, vtx2xyz,
         num_vtx, transform_world2pix, img_w, img_h, pix2tri, num_vdim, pix2val, dldw_pix2val,
         num_pair, L.off, err);
  int rc = launch_status();
  if (rc != DMSH_OK) return rc;
  // in place: off[s] = records before pair s, off[num_pair] = all records
  if ((rc = topo::launch_exclusive_scan(s, num_pair, L.off, L.off, L.sums)) != DMSH_OK) return rc;
  launch("k_eg_records", eg::k_eg_records, dim3(blocks), dim3(256), 0, s, tri2vtx, num_tri,
         vtx2xyz, num_vtx, transform_world2pix, img_w, img_h, pix2tri, num_vdim, pix2val,
         dldw_pix2val, num_pair, L.off, L.rec_vtx, L.rec_idx, L.rec_term, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  rc = sortk::sort_pairs_device_count<uint32_t>(s, cap, L.off + num_pair, L.rec_vtx, L.rec_idx,
                                                eg_key_bits(num_vtx), L.sort_ws, L.sort_bytes);
  if (rc != DMSH_OK) return rc;
  launch("k_eg_sum", eg::k_eg_sum, dim3(std::min<uint32_t>(ceil_div(cap, 256), 16384)),
         dim3(256), 0, s, (const uint32_t*)(L.off + num_pair), (const uint32_t*)L.rec_vtx,
         (const uint32_t*)L.rec_idx, (const float*)L.rec_term, num_vtx, cap, dldw_vtx2xyz, err);
  return launch_status();
}

int dmsh_edge_gradient_and_type(void* stream, const uint32_t* tri2vtx, uint32_t num_tri,
                                const float* vtx2xyz, uint32_t num_vtx,
                                const float* transform_world2pix, uint32_t img_w, uint32_t img_h,
                                const uint32_t* pix2tri, uint32_t num_vdim, const float* pix2val,
                                const float* dldw_pix2val, uint8_t* hedge2type,
                                float* hedge2dldr, uint8_t* vedge2type, float* vedge2dldr) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (img_w < 2 || img_h < 2 || num_pix >= (1ull << 32) || num_vdim == 0) return DMSH_ERR_INVALID;
  if (!transform_world2pix || !pix2tri || !pix2val || !dldw_pix2val || !hedge2type ||
      !hedge2dldr || !vedge2type || !vedge2dldr)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_edge_type_grad", eg::k_edge_type_grad, dim3(ceil_div(img_w, 64), ceil_div(img_h, 4)),
         dim3(64, 4), 0, as_stream(stream), tri2vtx, num_tri, vtx2xyz, num_vtx,
         transform_world2pix, img_w, img_h, pix2tri, num_vdim, pix2val, dldw_pix2val, hedge2type,
         hedge2dldr, vedge2type, vedge2dldr, err);
  return launch_status();
}

namespace {
struct SmoothGrids {
  uint32_t tiles_h, tiles_v, nh, nv;
};
inline SmoothGrids smooth_grids(uint32_t W, uint32_t H) {
  SmoothGrids g;
  g.nh = (H - 1) * W;
  g.nv = H * (W - 1);
  g.tiles_h = ceil_div(H - 1, eg::kSmoothTR) * ceil_div(W, eg::kSmoothTC);
  g.tiles_v = ceil_div(H, eg::kSmoothTR) * ceil_div(W - 1, eg::kSmoothTC);
  return g;
}
}  // namespace

size_t dmsh_smooth_gradient_workspace_bytes(uint32_t img_w, uint32_t img_h) {
  if (img_w < 2 || img_h < 2) return 256;
  const SmoothGrids g = smooth_grids(img_w, img_h);
  Arena a{nullptr, ~(size_t)0, 0};
  a.take<float>(g.nh);
  a.take<float>(g.nv);
  a.take<uint8_t>(g.nh);
  a.take<uint8_t>(g.nv);
  return a.off + 256;
}

int dmsh_smooth_gradient(void* stream, uint32_t img_w, uint32_t img_h, const uint8_t* hedge2type,
                         const uint8_t* vedge2type, uint32_t num_iter, float* hedge2dldr,
                         float* vedge2dldr, void* ws, size_t ws_bytes) {
  if (img_w < 2 || img_h < 2 || (uint64_t)img_w * img_h >= (1ull << 31)) return DMSH_ERR_INVALID;
  if (!hedge2type || !vedge2type || !hedge2dldr || !vedge2dldr) return DMSH_ERR_INVALID;
  if (num_iter == 0) return DMSH_OK;
  if (!ws || ws_bytes < dmsh_smooth_gradient_workspace_bytes(img_w, img_h))
    return DMSH_ERR_WORKSPACE;
  const SmoothGrids g = smooth_grids(img_w, img_h);
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  float* h2 = a.take<float>(g.nh);
  float* v2 = a.take<float>(g.nv);
  uint8_t* hm = a.take<uint8_t>(g.nh);
  uint8_t* vm = a.take<uint8_t>(g.nv);
  if (!a.ok()) return DMSH_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  int rc;
  launch("k_smooth_links", eg::k_smooth_links,
         dim3(std::min<uint32_t>(ceil_div(g.nh + g.nv, 256), 4096)), dim3(256), 0, s, img_w,
         img_h, hedge2type, vedge2type, hm, vm);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  const uint32_t steps = 2u * num_iter;
  const uint32_t launches = ceil_div(steps, (uint32_t)eg::kSmoothK);
  // launches alternate buffers and must end in the caller's arrays
  float *hs = hedge2dldr, *vs = vedge2dldr, *hd = h2, *vd = v2;
  if (launches % 2 == 1) {
    launch("k_copy_f32", eg::k_copy_f32, dim3(std::min<uint32_t>(ceil_div(g.nh, 256), 2048)),
           dim3(256), 0, s, hedge2dldr, h2, g.nh);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    launch("k_copy_f32", eg::k_copy_f32, dim3(std::min<uint32_t>(ceil_div(g.nv, 256), 2048)),
           dim3(256), 0, s, vedge2dldr, v2, g.nv);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    hs = h2;
    vs = v2;
    hd = hedge2dldr;
    vd = vedge2dldr;
  }
  for (uint32_t step0 = 0; step0 < steps; step0 += eg::kSmoothK) {
    const uint32_t n = std::min<uint32_t>(eg::kSmoothK, steps - step0);
    launch("k_smooth_tiled",
           (eg::k_smooth_tiled<eg::kSmoothK, eg::kSmoothTR, eg::kSmoothTC, eg::kSmoothThreads>),
           dim3(g.tiles_h + g.tiles_v), dim3(eg::kSmoothThreads), 0, s, img_w, img_h, hm, vm, hs,
           hd, vs, vd, g.tiles_h, step0, n);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    std::swap(hs, hd);
    std::swap(vs, vd);
  }
  return DMSH_OK;
}

int dmsh_interpolate_staggered_grid(void* stream, uint32_t img_w, uint32_t img_h,
                                    const float* hedge2vy, const float* vedge2vx,
                                    uint32_t num_vtx, const float* vtx2xy, float* vtx2velo) {
  if (num_vtx == 0) return DMSH_OK;
  if (img_w < 2 || img_h < 2 || !hedge2vy || !vedge2vx || !vtx2xy || !vtx2velo)
    return DMSH_ERR_INVALID;
  launch("k_interpolate_staggered", eg::k_interpolate_staggered, dim3(ceil_div(num_vtx, 256)),
         dim3(256), 0, as_stream(stream), img_w, img_h, hedge2vy, vedge2vx, num_vtx, vtx2xy,
         vtx2velo);
  return launch_status();
}

int dmsh_smoothed_vertex_gradient(void* stream, uint32_t img_w, uint32_t img_h,
                                  const float* hedge2dldr, const float* vedge2dldr,
                                  uint32_t num_vtx, const float* vtx2xyz,
                                  const float* transform_world2pix, float* dldw_vtx2xyz) {
  if (num_vtx == 0) return DMSH_OK;
  if (img_w < 2 || img_h < 2 || !hedge2dldr || !vedge2dldr || !vtx2xyz || !transform_world2pix ||
      !dldw_vtx2xyz)
    return DMSH_ERR_INVALID;
  launch("k_smoothed_vertex_grad", eg::k_smoothed_vertex_grad, dim3(ceil_div(num_vtx, 256)),
         dim3(256), 0, as_stream(stream), img_w, img_h, hedge2dldr, vedge2dldr, num_vtx, vtx2xyz,
         transform_world2pix, dldw_vtx2xyz);
  return launch_status();
}

}  // extern "C"
