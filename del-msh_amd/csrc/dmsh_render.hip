// dmsh_render.hip -- C ABI: ray-cast (pix2tri + hit t), depth, edge-gradient backward.
#include <algorithm>

#include "edgegrad.hpp"
#include "raycast.hpp"
#include "sort.hpp"

using namespace dmsh;

namespace {

constexpr int kCastStack = 16;  // LDS frames per lane
constexpr int kCastWaves = 2;   // waves per workgroup

struct CastLayout {
  uint32_t* ovf_count;
  uint32_t* ovf_list;
  uint2* deep;
};

inline bool carve_cast(Arena& a, uint32_t num_pix, CastLayout& L) {
  L.ovf_count = a.take<uint32_t>(64);
  L.ovf_list = a.take<uint32_t>(num_pix ? num_pix : 1);
  L.deep = a.take<uint2>((size_t)rc::kDeepSlots * rc::kDeepFrames);
  return a.ok();
}

inline bool scene_ok(uint32_t num_tri, const uint32_t* tri2vtx, const float* vtx2xyz,
                     const uint32_t* bvhnodes, const float* aabb) {
  if (num_tri >= rc::kNoneTri) return false;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz || !bvhnodes || !aabb)) return false;
  return true;
}

}  // namespace

extern "C" {

size_t dmsh_pix2tri_workspace_bytes(uint32_t img_w, uint32_t img_h) {
  Arena a{nullptr, 0, 0};
  CastLayout L;
  carve_cast(a, img_w * img_h, L);
  return a.off + 256;
}

int dmsh_pix2tri(void* stream, uint32_t* pix2tri, float* pix2t, uint32_t num_tri,
                 const uint32_t* tri2vtx, uint32_t num_vtx, const float* vtx2xyz, uint32_t img_w,
                 uint32_t img_h, const float* transform_ndc2world, const uint32_t* bvhnodes,
                 const float* bvhnode2aabb, void* ws, size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || !pix2tri || !transform_ndc2world) return DMSH_ERR_INVALID;
  if (!scene_ok(num_tri, tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_pix2tri_workspace_bytes(img_w, img_h)) return DMSH_ERR_WORKSPACE;
  Arena ar{static_cast<char*>(ws), ws_bytes, 0};
  CastLayout L;
  if (!carve_cast(ar, (uint32_t)num_pix, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  int st = sortk::launch_fill(s, L.ovf_count, 1, 0u);
  if (st != DMSH_OK) return st;
  rc::Scene sc{tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, num_tri, num_vtx,
               num_tri ? 2 * num_tri - 1 : 0};
  rc::CastArgs a;
  a.pix2tri = pix2tri;
  a.pix2t = pix2t;
  a.img_w = img_w;
  a.img_h = img_h;
  a.tiles_x = ceil_div(img_w, 8);
  a.num_tiles = a.tiles_x * ceil_div(img_h, 8);
  a.ovf_count = L.ovf_count;
  a.ovf_list = L.ovf_list;
  a.deep_stack = L.deep;
  a.err = err;
  hipLaunchKernelGGL((rc::k_pix2tri<kCastStack, kCastWaves>), dim3(ceil_div(a.num_tiles, kCastWaves)),
                     dim3(64 * kCastWaves), 0, s, sc, transform_ndc2world, a);
  if ((st = launch_status()) != DMSH_OK) return st;
  hipLaunchKernelGGL(rc::k_pix2tri_deep, dim3(rc::kDeepSlots / 64), dim3(64), 0, s, sc,
                     transform_ndc2world, a);
  return launch_status();
}

int dmsh_pix2depth(void* stream, float* pix2depth, const uint32_t* pix2tri, uint32_t num_tri,
                   const uint32_t* tri2vtx, uint32_t num_vtx, const float* vtx2xyz,
                   uint32_t img_w, uint32_t img_h, const float* transform_ndc2world) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || !pix2depth || !pix2tri || !transform_ndc2world)
    return DMSH_ERR_INVALID;
  if (num_tri > 0 && (!tri2vtx || !vtx2xyz)) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  rc::Scene sc{tri2vtx, vtx2xyz, nullptr, nullptr, num_tri, num_vtx, 0};
  hipLaunchKernelGGL(rc::k_pix2depth, dim3(ceil_div(num_pix, 256)), dim3(256), 0,
                     as_stream(stream), pix2depth, pix2tri, sc, img_w, img_h, transform_ndc2world,
                     err);
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
  hipLaunchKernelGGL(eg::k_edge_grad_bwd, dim3(ceil_div(img_w, 64), ceil_div(img_h, 4)),
                     dim3(64, 4), 0, as_stream(stream), tri2vtx, num_tri, vtx2xyz, num_vtx,
                     dldw_vtx2xyz, transform_world2pix, img_w, img_h, pix2tri, num_vdim, pix2val,
                     dldw_pix2val, err);
  return launch_status();
}

}  // extern "C"
