// dmsh_octree.hip -- C ABI: quad/oct-tree from sorted Morton codes, and the
// sorted-array / offset-array helpers around it.
#include <algorithm>

#include "dmsh.h"
#include "octree.hpp"
#include "sort.hpp"
#include "topo.hpp"

using namespace dmsh;

namespace {
inline dim3 grid_for(uint64_t n) { return dim3((uint32_t)std::max<uint64_t>(1, ceil_div(n, 256))); }
}  // namespace

extern "C" {

size_t dmsh_octree_bnodes_workspace_bytes(uint32_t num_idx) {
  Arena a{nullptr, ~(size_t)0, 0};
  a.take<uint32_t>(num_idx ? num_idx : 1);                   // flags
  a.take<uint32_t>(topo::scan_tiles(num_idx ? num_idx : 1));  // scan tile sums
  return a.off + 256;
}

int dmsh_octree_bnodes(void* stream, uint32_t num_idx, const uint32_t* idx2morton,
                       uint32_t num_dim, uint32_t max_depth, uint32_t* bnodes,
                       uint32_t* bnode2depth, uint32_t* bnode2onode, uint32_t* idx2bnode, void* ws,
                       size_t ws_bytes) {
  if (num_idx < 2 || (num_dim != 2 && num_dim != 3) || max_depth == 0 ||
      num_dim * max_depth > 32u)
    return DMSH_ERR_INVALID;
  if (!idx2morton || !bnodes || !bnode2depth || !bnode2onode || !idx2bnode)
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_octree_bnodes_workspace_bytes(num_idx)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  uint32_t* flags = a.take<uint32_t>(num_idx);
  uint32_t* sums = a.take<uint32_t>(topo::scan_tiles(num_idx));
  if (!a.ok()) return DMSH_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  uint32_t* err = device_error_word();
  const uint32_t num_bnode = num_idx - 1;
  int rc = sortk::launch_fill(s, idx2bnode, num_idx, DMSH_U32MAX);
  if (rc != DMSH_OK) return rc;
  launch("k_brt_depth", oct::k_brt_depth, grid_for(num_bnode), dim3(256), 0, s, num_bnode,
         idx2morton, num_dim, max_depth, bnodes, bnode2depth);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  launch("k_bnode_flags", oct::k_bnode_flags, grid_for(num_bnode), dim3(256), 0, s, num_bnode,
         bnodes, bnode2depth, flags, idx2bnode, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  // bnode2onode[i] = flags[0] + ... + flags[i - 1], i < num_bnode
  return topo::launch_exclusive_scan(s, num_bnode - 1, flags, bnode2onode, sums);
}

int dmsh_octree_make_tree(void* stream, uint32_t num_idx, const uint32_t* bnodes,
                          const uint32_t* bnode2onode, const uint32_t* bnode2depth,
                          const uint32_t* idx2bnode, const uint32_t* idx2morton, uint32_t num_dim,
                          uint32_t max_depth, uint32_t num_onode, uint32_t* onodes,
                          uint32_t* onode2depth, float* onode2center, uint32_t* idx2onode,
                          float* idx2center) {
  if (num_idx < 2 || (num_dim != 2 && num_dim != 3) || max_depth == 0 ||
      num_dim * max_depth > 32u || num_onode == 0 || num_onode > num_idx)
    return DMSH_ERR_INVALID;
  if (!bnodes || !bnode2onode || !bnode2depth || !idx2bnode || !idx2morton || !onodes ||
      !onode2depth || !onode2center || !idx2onode || !idx2center)
    return DMSH_ERR_INVALID;
  hipStream_t s = as_stream(stream);
  const uint32_t nlink = (1u << num_dim) + 1u;
  int rc = sortk::launch_fill(s, onodes, (size_t)num_onode * nlink, DMSH_U32MAX);
  if (rc != DMSH_OK) return rc;
  if ((rc = sortk::launch_fill(s, onode2depth, num_onode, 0u)) != DMSH_OK) return rc;
  launch("k_octree_make", oct::k_octree_make, grid_for(num_idx), dim3(256), 0, s, num_idx, bnodes,
         bnode2onode, bnode2depth, idx2bnode, idx2morton, num_onode, max_depth, num_dim, onodes,
         onode2depth, onode2center, idx2onode, idx2center, device_error_word());
  return launch_status();
}

size_t dmsh_octree_aggregate_workspace_bytes(uint32_t num_onode) {
  return sizeof(uint32_t) * (size_t)(num_onode ? num_onode : 1) + 256;
}

int dmsh_octree_aggregate(void* stream, uint32_t num_idx, uint32_t num_vdim, const float* idx2val,
                          const uint32_t* idx2onode, uint32_t num_link, uint32_t num_onode,
                          const uint32_t* onodes, float* onode2aggval, void* ws,
                          size_t ws_bytes) {
  if (num_onode == 0 || num_vdim == 0) return DMSH_OK;
  if (num_link < 2 || !idx2onode || !onodes || !onode2aggval || (num_idx > 0 && !idx2val))
    return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_octree_aggregate_workspace_bytes(num_onode))
    return DMSH_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  uint32_t* hops = static_cast<uint32_t*>(ws);
  uint32_t* err = device_error_word();
  launch("k_onode_hops", oct::k_onode_hops, grid_for(num_onode), dim3(256), 0, s, num_onode,
         num_link, onodes, hops, err);
  int rc = launch_status();
  for (int level = (int)oct::kMaxHops; level >= 0 && rc == DMSH_OK; --level) {
    launch("k_onode_sum_level", oct::k_onode_sum_level, grid_for(num_onode), dim3(256), 0, s,
           num_onode, num_vdim, num_link, onodes, idx2val, num_idx, hops, (uint32_t)level,
           onode2aggval, err);
    rc = launch_status();
  }
  return rc;
}

int dmsh_permute_u32(void* stream, uint32_t n, const uint32_t* old2val, const uint32_t* new2old,
                     uint32_t* new2val) {
  if (n == 0) return DMSH_OK;
  if (!old2val || !new2old || !new2val) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_permute_u32", oct::k_permute_u32, dim3(std::min<uint32_t>(ceil_div(n, 256), 8192)),
         dim3(256), 0, as_stream(stream), n, old2val, new2old, new2val, err);
  return launch_status();
}

int dmsh_has_duplicate_sorted_u32(void* stream, uint32_t n, const uint32_t* idx2val,
                                  uint32_t* flag) {
  if (!flag || (n > 0 && !idx2val)) return DMSH_ERR_INVALID;
  hipStream_t s = as_stream(stream);
  int rc = sortk::launch_fill(s, flag, 1, 0u);
  if (rc != DMSH_OK || n < 2) return rc;
  launch("k_change_flags", oct::k_change_flags, grid_for(n - 1), dim3(256), 0, s, n, idx2val,
         (uint32_t*)nullptr, flag);
  return launch_status();
}

size_t dmsh_unique_for_sorted_workspace_bytes(uint32_t n) {
  Arena a{nullptr, ~(size_t)0, 0};
  a.take<uint32_t>(n ? n : 1);
  a.take<uint32_t>(topo::scan_tiles(n ? n : 1));
  return a.off + 256;
}

int dmsh_unique_for_sorted_u32(void* stream, uint32_t n, const uint32_t* idx2val,
                               uint32_t* idx2jdx, void* ws, size_t ws_bytes) {
  if (n == 0) return DMSH_OK;
  if (!idx2val || !idx2jdx) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_unique_for_sorted_workspace_bytes(n)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  uint32_t* flags = a.take<uint32_t>(n);
  uint32_t* sums = a.take<uint32_t>(topo::scan_tiles(n));
  if (!a.ok()) return DMSH_ERR_WORKSPACE;
  hipStream_t s = as_stream(stream);
  if (n >= 2) {
    launch("k_change_flags", oct::k_change_flags, grid_for(n - 1), dim3(256), 0, s, n, idx2val,
           flags, (uint32_t*)nullptr);
    int rc = launch_status();
    if (rc != DMSH_OK) return rc;
  }
  return topo::launch_exclusive_scan(s, n - 1, flags, idx2jdx, sums);
}

int dmsh_unique_jdx2val_jdx2idx(void* stream, uint32_t n, const uint32_t* idx2val,
                                const uint32_t* idx2jdx, uint32_t num_jdx, uint32_t* jdx2val,
                                uint32_t* jdx2idx_offset) {
  if (!jdx2idx_offset || (num_jdx > 0 && !jdx2val) || (n > 0 && (!idx2val || !idx2jdx)))
    return DMSH_ERR_INVALID;
  launch("k_unique_scatter", oct::k_unique_scatter, grid_for((uint64_t)n + 1), dim3(256), 0,
         as_stream(stream), n, idx2val, idx2jdx, num_jdx, jdx2val, jdx2idx_offset,
         device_error_word());
  return launch_status();
}

int dmsh_offset_array_aggregate(void* stream, uint32_t num_idx, const uint32_t* idx2jdx_offset,
                                uint32_t num_jdx, const uint32_t* jdx2kdx, uint32_t num_kdx,
                                uint32_t num_dim, const float* kdx2val, float* idx2aggval) {
  if (num_idx == 0 || num_dim == 0) return DMSH_OK;
  if (!idx2jdx_offset || !idx2aggval || (num_jdx > 0 && (!jdx2kdx || !kdx2val)))
    return DMSH_ERR_INVALID;
  launch("k_offset_aggregate", oct::k_offset_aggregate, grid_for((uint64_t)num_idx * num_dim),
         dim3(256), 0, as_stream(stream), num_idx, num_dim, idx2jdx_offset, num_jdx, jdx2kdx,
         num_kdx, kdx2val, idx2aggval, device_error_word());
  return launch_status();
}

}  // extern "C"
