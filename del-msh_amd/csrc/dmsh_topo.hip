// dmsh_topo.hip -- C ABI: mesh topology (vertex -> elements / vertices,
// edges, edge -> triangles), the graph Laplacian, and an exclusive scan.
#include <algorithm>

#include "dmsh.h"
#include "topo.hpp"

using namespace dmsh;

namespace {

inline uint32_t key_bits_for(uint32_t num_vtx) {
  uint32_t b = 1;
  while (b < 32 && ((num_vtx - 1) >> b) != 0u) ++b;
  return b;
}

struct Vtx2ElemLayout {
  uint32_t* keys;     // nflat
  uint32_t* slot;     // nflat
  char* sort_ws;
  size_t sort_bytes;
};

inline bool carve_vtx2elem(Arena& a, uint32_t nflat, uint32_t num_vtx, Vtx2ElemLayout& L) {
  L.keys = a.take<uint32_t>(nflat ? nflat : 1);
  L.slot = a.take<uint32_t>(nflat ? nflat : 1);
  L.sort_bytes = sortk::workspace_bytes<uint32_t>(nflat ? nflat : 1, key_bits_for(num_vtx));
  L.sort_ws = a.take<char>(L.sort_bytes);
  return a.ok();
}

// vtx2idx [num_vtx + 1] and idx2elem [nflat] of a uniform mesh
int vtx2elem(hipStream_t s, uint32_t num_elem, uint32_t num_node, const uint32_t* elem2vtx,
             uint32_t num_vtx, uint32_t* vtx2idx, uint32_t* idx2elem, const Vtx2ElemLayout& L,
             uint32_t* err) {
  const uint32_t nflat = num_elem * num_node;
  int rc;
  if (nflat > 0) {
    launch("k_vtx_keys", topo::k_vtx_keys, dim3(std::min<uint32_t>(ceil_div(nflat, 256), 4096)),
           dim3(256), 0, s, elem2vtx, nflat, num_vtx, L.keys, err);
    if ((rc = launch_status()) != DMSH_OK) return rc;
    rc = sortk::argsort<uint32_t>(s, nflat, L.keys, L.slot, key_bits_for(num_vtx), L.sort_ws,
                                  L.sort_bytes);
    if (rc != DMSH_OK) return rc;
  }
  launch("k_vtx2elem_offsets", topo::k_vtx2elem_offsets, dim3(ceil_div(nflat + 1, 256)),
         dim3(256), 0, s, L.keys, nflat, num_vtx, vtx2idx);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  if (nflat == 0) return DMSH_OK;
  launch("k_idx2elem", topo::k_idx2elem, dim3(std::min<uint32_t>(ceil_div(nflat, 256), 4096)),
         dim3(256), 0, s, L.slot, nflat, num_node, idx2elem);
  return launch_status();
}

bool mesh_args_ok(uint32_t num_elem, uint32_t num_node, const uint32_t* elem2vtx,
                  uint32_t num_vtx) {
  if (num_node == 0 || num_vtx == 0) return false;
  if ((uint64_t)num_elem * num_node >= (1ull << 30)) return false;
  return num_elem == 0 || elem2vtx != nullptr;
}

struct Vtx2VtxLayout {
  Vtx2ElemLayout ve;
  uint32_t* vtx2jdx;   // num_vtx + 1
  uint32_t* jdx2elem;  // nflat
  uint32_t* buff;      // nflat * m
  uint32_t* counts;    // num_vtx
  uint32_t* sums;      // scan tiles
};

inline uint32_t neighbours_per_elem(uint32_t num_node, int is_self) {
  return is_self ? num_node : num_node - 1u;
}

inline bool carve_vtx2vtx(Arena& a, uint32_t num_elem, uint32_t num_node, uint32_t num_vtx,
                          int is_self, Vtx2VtxLayout& L) {
  const uint32_t nflat = num_elem * num_node;
  const size_t m = neighbours_per_elem(num_node, is_self);
  L.vtx2jdx = a.take<uint32_t>((size_t)num_vtx + 1);
  L.jdx2elem = a.take<uint32_t>(nflat ? nflat : 1);
  L.buff = a.take<uint32_t>(std::max<size_t>((size_t)nflat * m, 1));
  L.counts = a.take<uint32_t>(num_vtx);
  L.sums = a.take<uint32_t>(topo::scan_tiles(num_vtx));
  return carve_vtx2elem(a, nflat, num_vtx, L.ve) && a.ok();
}

}  // namespace

extern "C" {

size_t dmsh_exclusive_scan_workspace_bytes(uint32_t n) {
  return sizeof(uint32_t) * (size_t)topo::scan_tiles(n) + 256;
}

int dmsh_exclusive_scan_u32(void* stream, uint32_t n, const uint32_t* in, uint32_t* out, void* ws,
                            size_t ws_bytes) {
  if (!out || (n > 0 && !in)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_exclusive_scan_workspace_bytes(n)) return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  uint32_t* sums = a.take<uint32_t>(topo::scan_tiles(n));
  if (!a.ok()) return DMSH_ERR_WORKSPACE;
  return topo::launch_exclusive_scan(as_stream(stream), n, in, out, sums);
}

size_t dmsh_vtx2elem_workspace_bytes(uint32_t num_elem, uint32_t num_node, uint32_t num_vtx) {
  Arena a{nullptr, ~(size_t)0, 0};
  Vtx2ElemLayout L;
  carve_vtx2elem(a, num_elem * num_node, num_vtx ? num_vtx : 1, L);
  return a.off + 256;
}

int dmsh_vtx2elem_from_uniform_mesh(void* stream, uint32_t num_elem, uint32_t num_node,
                                    const uint32_t* elem2vtx, uint32_t num_vtx,
                                    uint32_t* vtx2idx, uint32_t* idx2elem, void* ws,
                                    size_t ws_bytes) {
  if (!mesh_args_ok(num_elem, num_node, elem2vtx, num_vtx) || !vtx2idx) return DMSH_ERR_INVALID;
  if (num_elem > 0 && !idx2elem) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_vtx2elem_workspace_bytes(num_elem, num_node, num_vtx))
    return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  Vtx2ElemLayout L;
  if (!carve_vtx2elem(a, num_elem * num_node, num_vtx, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  return vtx2elem(as_stream(stream), num_elem, num_node, elem2vtx, num_vtx, vtx2idx, idx2elem, L,
                  err);
}

size_t dmsh_vtx2vtx_workspace_bytes(uint32_t num_elem, uint32_t num_node, uint32_t num_vtx,
                                    int is_self) {
  Arena a{nullptr, ~(size_t)0, 0};
  Vtx2VtxLayout L;
  carve_vtx2vtx(a, num_elem, num_node ? num_node : 1, num_vtx ? num_vtx : 1, is_self, L);
  return a.off + 256;
}

int dmsh_vtx2vtx_count(void* stream, uint32_t num_elem, uint32_t num_node,
                       const uint32_t* elem2vtx, uint32_t num_vtx, int is_self,
                       uint32_t* vtx2idx, void* ws, size_t ws_bytes) {
  if (!mesh_args_ok(num_elem, num_node, elem2vtx, num_vtx) || !vtx2idx) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_vtx2vtx_workspace_bytes(num_elem, num_node, num_vtx, is_self))
    return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  Vtx2VtxLayout L;
  if (!carve_vtx2vtx(a, num_elem, num_node, num_vtx, is_self, L)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  int rc = vtx2elem(s, num_elem, num_node, elem2vtx, num_vtx, L.vtx2jdx, L.jdx2elem, L.ve, err);
  if (rc != DMSH_OK) return rc;
  launch("k_vtx2vtx_count", topo::k_vtx2vtx_count, dim3(ceil_div(num_vtx, topo::kV2VWaves)),
         dim3(256), 0, s,
         num_vtx, elem2vtx, num_elem, num_node, (uint32_t)(is_self != 0), L.vtx2jdx, L.jdx2elem,
         L.buff, L.counts, err);
  if ((rc = launch_status()) != DMSH_OK) return rc;
  return topo::launch_exclusive_scan(s, num_vtx, L.counts, vtx2idx, L.sums);
}

int dmsh_vtx2vtx_fill(void* stream, uint32_t num_elem, uint32_t num_node, uint32_t num_vtx,
                      int is_self, const uint32_t* vtx2idx, uint32_t* idx2vtx, void* ws,
                      size_t ws_bytes) {
  if (num_node == 0 || num_vtx == 0 || !vtx2idx) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_vtx2vtx_workspace_bytes(num_elem, num_node, num_vtx, is_self))
    return DMSH_ERR_WORKSPACE;
  Arena a{static_cast<char*>(ws), ws_bytes, 0};
  Vtx2VtxLayout L;
  if (!carve_vtx2vtx(a, num_elem, num_node, num_vtx, is_self, L)) return DMSH_ERR_WORKSPACE;
  if (!idx2vtx) return DMSH_ERR_INVALID;
  launch("k_vtx2vtx_fill", topo::k_vtx2vtx_fill, dim3(ceil_div(num_vtx, 256)), dim3(256), 0,
         as_stream(stream), num_vtx, neighbours_per_elem(num_node, is_self), L.vtx2jdx, L.buff,
         vtx2idx, idx2vtx);
  return launch_status();
}

int dmsh_edge2vtx_from_vtx2vtx(void* stream, uint32_t num_vtx, const uint32_t* vtx2idx,
                               uint32_t num_idx, const uint32_t* idx2vtx, uint32_t* edge2vtx) {
  if (num_vtx == 0 || !vtx2idx) return DMSH_ERR_INVALID;
  if (num_idx == 0) return DMSH_OK;
  if (!idx2vtx || !edge2vtx) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_edge2vtx", topo::k_edge2vtx, dim3(ceil_div(num_idx, 256)), dim3(256), 0,
         as_stream(stream), num_vtx, vtx2idx, num_idx, idx2vtx, edge2vtx, err);
  return launch_status();
}

int dmsh_edge2elem_from_edge2vtx_of_tri2vtx(void* stream, uint32_t num_edge,
                                            const uint32_t* edge2vtx, uint32_t num_tri,
                                            const uint32_t* tri2vtx, uint32_t num_vtx,
                                            const uint32_t* vtx2idx, uint32_t num_idx,
                                            const uint32_t* idx2tri, uint32_t* edge2tri) {
  if (num_edge == 0) return DMSH_OK;
  if (!edge2vtx || !tri2vtx || !vtx2idx || !edge2tri || num_vtx == 0) return DMSH_ERR_INVALID;
  if (num_idx > 0 && !idx2tri) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_edge2elem", topo::k_edge2elem, dim3(ceil_div(num_edge, 256)), dim3(256), 0,
         as_stream(stream), num_edge, edge2vtx, num_tri, tri2vtx, num_vtx, vtx2idx, num_idx,
         idx2tri, edge2tri, err);
  return launch_status();
}

int dmsh_laplacian_smoothing(void* stream, uint32_t num_vtx, const uint32_t* vtx2idx,
                             uint32_t num_idx, const uint32_t* idx2vtx, float lambda,
                             uint32_t num_dim, float* vtx2lhs, const float* vtx2rhs,
                             uint32_t num_iter, float* vtx2lhstmp) {
  if (num_vtx == 0 || num_iter == 0) return DMSH_OK;
  if (!vtx2idx || !vtx2lhs || !vtx2rhs || !vtx2lhstmp || num_dim == 0) return DMSH_ERR_INVALID;
  if (num_idx > 0 && !idx2vtx) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  const dim3 grid(ceil_div(num_vtx, 256));
  for (uint32_t it = 0; it < num_iter; ++it) {
    launch("k_laplacian_jacobi", topo::k_laplacian_jacobi, grid, dim3(256), 0, s, num_vtx, vtx2idx,
           num_idx, idx2vtx, lambda, num_dim, vtx2lhs, vtx2lhstmp, vtx2rhs, err);
    int rc = launch_status();
    if (rc != DMSH_OK) return rc;
    launch("k_laplacian_jacobi", topo::k_laplacian_jacobi, grid, dim3(256), 0, s, num_vtx, vtx2idx,
           num_idx, idx2vtx, lambda, num_dim, vtx2lhstmp, vtx2lhs, vtx2rhs, err);
    if ((rc = launch_status()) != DMSH_OK) return rc;
  }
  return DMSH_OK;
}

int dmsh_multiply_graph_laplacian(void* stream, uint32_t num_vtx, const uint32_t* vtx2idx,
                                  uint32_t num_idx, const uint32_t* idx2vtx, uint32_t num_dim,
                                  const float* vtx2rhs, float* vtx2lhs) {
  if (num_vtx == 0) return DMSH_OK;
  if (!vtx2idx || !vtx2lhs || !vtx2rhs || num_dim == 0) return DMSH_ERR_INVALID;
  if (num_idx > 0 && !idx2vtx) return DMSH_ERR_INVALID;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  launch("k_graph_laplacian", topo::k_graph_laplacian, dim3(ceil_div(num_vtx, 256)), dim3(256), 0,
         as_stream(stream), num_vtx, vtx2idx, num_idx, idx2vtx, num_dim, vtx2rhs, vtx2lhs, err);
  return launch_status();
}

}  // extern "C"
