// dmsh_polymesh.hip -- C ABI: mixed polyhedron meshes in CSR form (element
// centroids, element-BVH refit, point location, interpolation at points).
#include <algorithm>

#include "dmsh.h"
#include "polymesh.hpp"
#include "refit.hpp"
#include "sort.hpp"

using namespace dmsh;

namespace {
inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, ceil_div(n, 256))); }
// query sets at least this large are searched in Morton order
constexpr uint32_t kSortedSearchMin = 1u << 16;
}  // namespace

extern "C" {

int dmsh_polyhedron_elem2center(void* stream, uint32_t num_elem, const uint32_t* elem2idx_offset,
                                uint32_t num_idx, const uint32_t* idx2vtx, uint32_t num_vtx,
                                uint32_t num_dim, const float* vtx2xyz, float* elem2center) {
  if (num_elem == 0) return DMSH_OK;
  if (num_dim == 0 || num_dim > 4 || !elem2idx_offset || !elem2center ||
      (num_idx > 0 && (!idx2vtx || !vtx2xyz)))
    return DMSH_ERR_INVALID;
  launch("k_elem2center", poly::k_elem2center, grid_for(num_elem), dim3(256), 0,
         as_stream(stream), num_elem, elem2idx_offset, num_idx, idx2vtx, num_vtx, num_dim,
         vtx2xyz, elem2center, device_error_word());
  return launch_status();
}

size_t dmsh_polyhedron_refit_workspace_bytes(uint32_t num_elem) {
  return standalone_workspace_bytes(num_elem);
}

int dmsh_bvhnode2aabb_polyhedron(void* stream, uint32_t num_bvhnode, const uint32_t* bvhnodes,
                                 uint32_t num_elem, const uint32_t* elem2idx_offset,
                                 uint32_t num_idx, const uint32_t* idx2vtx, uint32_t num_vtx,
                                 const float* vtx2xyz, float* bvhnode2aabb, void* ws,
                                 size_t ws_bytes) {
  if (num_elem == 0) return num_bvhnode == 0 ? DMSH_OK : DMSH_ERR_INVALID;
  if (num_elem >= (1u << 31) || num_bvhnode != 2 * num_elem - 1) return DMSH_ERR_INVALID;
  if (!bvhnodes || !elem2idx_offset || !idx2vtx || !vtx2xyz || !bvhnode2aabb)
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_polyhedron_refit_workspace_bytes(num_elem))
    return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  const TreeView v{bvhnode2aabb, bvhnodes, num_elem, idx2vtx, num_vtx, vtx2xyz, nullptr,
                   elem2idx_offset, num_idx};
  return launch_refit_standalone<kLeafElem>(as_stream(stream), v, 0.f, ws, ws_bytes, err);
}

size_t dmsh_polyhedron_search_workspace_bytes(uint32_t num_wtx) {
  Arena a{nullptr, ~(size_t)0, 0};
  a.take<uint32_t>(num_wtx ? num_wtx : 1);  // query codes
  a.take<uint32_t>(num_wtx ? num_wtx : 1);  // processing order
  a.take<char>(sortk::workspace_bytes<uint32_t>(num_wtx ? num_wtx : 1, 30));
  return a.off + 256;
}

int dmsh_polyhedron_search_elem_contain_points(void* stream, uint32_t num_elem,
                                               const uint32_t* elem2idx_offset, uint32_t num_idx,
                                               const uint32_t* idx2vtx, uint32_t num_vtx,
                                               const float* vtx2xyz, const uint32_t* bvhnodes,
                                               const float* bvhnode2aabb, uint32_t num_wtx,
                                               const float* wtx2xyz, uint32_t* wtx2elem,
                                               float* wtx2param, void* ws, size_t ws_bytes) {
  if (num_wtx == 0) return DMSH_OK;
  if (!wtx2xyz || !wtx2elem || !wtx2param || !elem2idx_offset ||
      (num_elem > 0 && (!bvhnodes || !bvhnode2aabb || !idx2vtx || !vtx2xyz)))
    return DMSH_ERR_INVALID;
  if (num_elem >= (1u << 31) || num_wtx >= (1u << 30)) return DMSH_ERR_INVALID;
  hipStream_t s = as_stream(stream);
  const uint32_t* order = nullptr;
  // Large query sets are processed in Morton order of the query points
  // (workspace given); small ones, or no workspace, in the given order.
  if (ws && num_elem > 0 && num_wtx >= kSortedSearchMin) {
    if (ws_bytes < dmsh_polyhedron_search_workspace_bytes(num_wtx)) return DMSH_ERR_WORKSPACE;
    Arena a{static_cast<char*>(ws), ws_bytes, 0};
    uint32_t* keys = a.take<uint32_t>(num_wtx);
    uint32_t* idx = a.take<uint32_t>(num_wtx);
    const size_t sort_bytes = sortk::workspace_bytes<uint32_t>(num_wtx, 30);
    char* sort_ws = a.take<char>(sort_bytes);
    if (!a.ok()) return DMSH_ERR_WORKSPACE;
    launch("k_query_codes", poly::k_query_codes, grid_for(num_wtx), dim3(256), 0, s, num_wtx,
           wtx2xyz, bvhnode2aabb, keys, idx);
    int rc = launch_status();
    if (rc != DMSH_OK) return rc;
    rc = sortk::argsort<uint32_t>(s, num_wtx, keys, idx, 30, sort_ws, sort_bytes);
    if (rc != DMSH_OK) return rc;
    order = idx;
  }
  const poly::MeshView m{num_elem, elem2idx_offset, num_idx, idx2vtx, num_vtx, vtx2xyz};
  launch("k_search_elem_contain_points", poly::k_search_elem_contain_points, grid_for(num_wtx),
         dim3(256), 0, s, m, bvhnodes, bvhnode2aabb, num_wtx, wtx2xyz, wtx2elem, wtx2param,
         order, device_error_word());
  return launch_status();
}

int dmsh_polyhedron_interpolate_at_points(void* stream, uint32_t num_wtx, uint32_t num_elem,
                                          const uint32_t* elem2idx_offset, uint32_t num_idx,
                                          const uint32_t* idx2vtx, uint32_t num_vtx,
                                          uint32_t num_vdim, const float* vtx2val,
                                          const uint32_t* wtx2elem, const float* wtx2param,
                                          float* wtx2val) {
  if (num_wtx == 0 || num_vdim == 0) return DMSH_OK;
  if (!wtx2elem || !wtx2param || !wtx2val || !elem2idx_offset ||
      (num_elem > 0 && (!idx2vtx || !vtx2val)))
    return DMSH_ERR_INVALID;
  const poly::MeshView m{num_elem, elem2idx_offset, num_idx, idx2vtx, num_vtx, nullptr};
  launch("k_interpolate_at_points", poly::k_interpolate_at_points, grid_for(num_wtx), dim3(256),
         0, as_stream(stream), m, num_wtx, num_vdim, vtx2val, wtx2elem, wtx2param, wtx2val,
         device_error_word());
  return launch_status();
}

}  // extern "C"
