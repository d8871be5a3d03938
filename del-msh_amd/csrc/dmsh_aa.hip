// dmsh_aa.hip -- C ABI: contour edges and the anti-aliased silhouette.
#include <algorithm>

#include "antialias.hpp"

using namespace dmsh;

namespace {

struct ContourLayout {
  float* n2w;  // 16 + ok flag
  uint32_t* flags;
  uint32_t* block_count;
};

inline uint32_t contour_blocks(uint32_t num_edge) { return ceil_div(num_edge, 256); }

inline bool carve_contour(Arena& a, uint32_t num_edge, ContourLayout& L) {
  L.n2w = a.take<float>(17);
  L.flags = a.take<uint32_t>(num_edge ? num_edge : 1);
  L.block_count = a.take<uint32_t>(num_edge ? contour_blocks(num_edge) : 1);
  return a.ok();
}

}  // namespace

extern "C" {

size_t dmsh_contour_workspace_bytes(uint32_t num_edge) {
  Arena a{nullptr, ~(size_t)0, 0};
  ContourLayout L;
  carve_contour(a, num_edge, L);
  return a.off + 256;
}

int dmsh_contour_for_triangle_mesh(void* stream, uint32_t num_tri, const uint32_t* tri2vtx,
                                   uint32_t num_vtx, const float* vtx2xyz,
                                   const float* transform_world2ndc, uint32_t num_edge,
                                   const uint32_t* edge2vtx, const uint32_t* edge2tri,
                                   uint32_t* cedge2vtx, uint32_t* num_cedge, void* ws,
                                   size_t ws_bytes) {
  if (!num_cedge || !transform_world2ndc) return DMSH_ERR_INVALID;
  if (num_edge > 0 && (!tri2vtx || !vtx2xyz || !edge2vtx || !edge2tri || !cedge2vtx))
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_contour_workspace_bytes(num_edge)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  ContourLayout L;
  if (!carve_contour(a, num_edge, L)) return DMSH_ERR_WORKSPACE;
  launch("k_mat4_inverse", k_mat4_inverse, dim3(1), dim3(1), 0, s, transform_world2ndc, L.n2w, err);
  int st = launch_status();
  if (st != DMSH_OK) return st;
  const uint32_t nb = num_edge ? contour_blocks(num_edge) : 0;
  if (nb) {
    aa::ContourArgs ca{tri2vtx, num_tri, vtx2xyz, num_vtx, transform_world2ndc, L.n2w,
                       num_edge, edge2vtx, edge2tri, err};
    launch("k_contour_count", aa::k_contour_count, dim3(nb), dim3(256), 0, s, ca, L.flags,
                       L.block_count);
    if ((st = launch_status()) != DMSH_OK) return st;
  }
  // also writes the total when there is no edge (0)
  launch("k_scan_counts", aa::k_scan_counts, dim3(1), dim3(1024), 0, s, L.block_count, nb, num_cedge);
  if ((st = launch_status()) != DMSH_OK) return st;
  if (nb) {
    launch("k_contour_scatter", aa::k_contour_scatter, dim3(nb), dim3(256), 0, s, edge2vtx, num_edge,
                       L.flags, L.block_count, cedge2vtx);
    st = launch_status();
  }
  return st;
}

size_t dmsh_antialias_workspace_bytes(uint32_t img_w, uint32_t img_h) {
  return (size_t)img_w * img_h * sizeof(unsigned long long) + 256;
}

int dmsh_antialias_fwd(void* stream, uint32_t num_cedge, const uint32_t* cedge2vtx,
                       uint32_t num_vtx, const float* vtx2xyz, const float* transform_world2pix,
                       uint32_t img_w, uint32_t img_h, const uint32_t* pix2tri,
                       const float* pix2vin, float* pix2vout, void* ws, size_t ws_bytes) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0 || num_cedge == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || num_cedge == DMSH_U32MAX) return DMSH_ERR_INVALID;
  if (!cedge2vtx || !vtx2xyz || !transform_world2pix || !pix2tri || !pix2vin || !pix2vout)
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_antialias_workspace_bytes(img_w, img_h)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  auto* keys = static_cast<unsigned long long*>(ws);
  launch("k_zero_u64", aa::k_zero_u64, dim3(std::min<uint32_t>(ceil_div(num_pix, 256), 4096)),
                     dim3(256), 0, s, keys, num_pix);
  int st = launch_status();
  if (st != DMSH_OK) return st;
  aa::EdgeArgs ea{cedge2vtx, num_cedge, vtx2xyz, num_vtx, transform_world2pix,
                  img_w, img_h, pix2tri, err};
  launch("k_aa_keys", aa::k_aa_keys, dim3(ceil_div(num_cedge, 64)), dim3(64), 0, s, ea, keys);
  if ((st = launch_status()) != DMSH_OK) return st;
  launch("k_aa_write", aa::k_aa_write, dim3(ceil_div(num_cedge, 64)), dim3(64), 0, s, ea, keys,
                     pix2vin, pix2vout);
  return launch_status();
}

int dmsh_antialias_bwd(void* stream, uint32_t num_cedge, const uint32_t* cedge2vtx,
                       uint32_t num_vtx, const float* vtx2xyz, float* dldw_vtx2xyz,
                       const float* transform_world2pix, uint32_t img_w, uint32_t img_h,
                       const float* pix2val, const float* dldw_pix2val, const uint32_t* pix2tri) {
  const uint64_t num_pix = (uint64_t)img_w * img_h;
  if (num_pix == 0 || num_cedge == 0) return DMSH_OK;
  if (num_pix >= (1ull << 32) || num_cedge == DMSH_U32MAX) return DMSH_ERR_INVALID;
  if (!cedge2vtx || !vtx2xyz || !dldw_vtx2xyz || !transform_world2pix || !pix2val ||
      !dldw_pix2val || !pix2tri)
    return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  aa::EdgeArgs ea{cedge2vtx, num_cedge, vtx2xyz, num_vtx, transform_world2pix,
                  img_w, img_h, pix2tri, err};
  launch("k_aa_bwd", aa::k_aa_bwd, dim3(ceil_div(num_cedge, 64)), dim3(64), 0, as_stream(stream),
                     ea, pix2val, dldw_pix2val, dldw_vtx2xyz);
  return launch_status();
}

}  // extern "C"
