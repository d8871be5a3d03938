// dmsh_build.hip -- C ABI: Morton codes, argsort, topology, refit, whole build.
#include <algorithm>
#include <mutex>

#include "build.hpp"

using namespace dmsh;

// Scratch of the allocation-free build: one array per GPU image, held until
// the process ends.  Host code never passes this name to a kernel (on the
// host it is only a stand-in for the symbol); it uses the device address
// that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[DMSH_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kScratchDevices = 64;
void* g_scratch_ptr[kScratchDevices];
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device, looked up on the
// first call on that device (with that device current) and kept.
void* scratch_of_current_device() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int dmsh_vtx2morton_u32(void* stream, uint32_t num_vtx, const float* vtx2co, uint32_t num_dim,
                        const float* transform_co2unit, uint32_t* vtx2morton) {
  if (num_vtx == 0) return DMSH_OK;
  if (!vtx2co || !transform_co2unit || !vtx2morton) return DMSH_ERR_INVALID;
  hipStream_t s = as_stream(stream);
  if (num_dim == 3) {
    launch("k_vtx2morton3", k_vtx2morton3<uint32_t>, dim3(ceil_div(num_vtx, 256)), dim3(256), 0, s,
                       num_vtx, vtx2co, transform_co2unit, vtx2morton);
  } else if (num_dim == 2) {
    launch("k_vtx2morton2", k_vtx2morton2, dim3(ceil_div(num_vtx, 256)), dim3(256), 0, s, num_vtx,
                       vtx2co, transform_co2unit, vtx2morton);
  } else {
    return DMSH_ERR_INVALID;
  }
  return launch_status();
}

int dmsh_vtx2morton_u64(void* stream, uint32_t num_vtx, const float* vtx2xyz,
                        const float* transform_co2unit, uint64_t* vtx2morton) {
  if (num_vtx == 0) return DMSH_OK;
  if (!vtx2xyz || !transform_co2unit || !vtx2morton) return DMSH_ERR_INVALID;
  launch("k_vtx2morton3", k_vtx2morton3<uint64_t>, dim3(ceil_div(num_vtx, 256)), dim3(256), 0,
                     as_stream(stream), num_vtx, vtx2xyz, transform_co2unit, vtx2morton);
  return launch_status();
}

size_t dmsh_argsort_workspace_bytes(uint32_t n, uint32_t key_bytes) {
  if (key_bytes == 8) return sortk::workspace_bytes<uint64_t>(n, 64);
  return sortk::workspace_bytes<uint32_t>(n, 32);
}

int dmsh_argsort_u32(void* stream, uint32_t n, uint32_t* keys_inout, uint32_t* idx_out,
                     uint32_t key_bits, void* ws, size_t ws_bytes) {
  return sortk::argsort<uint32_t>(as_stream(stream), n, keys_inout, idx_out, key_bits, ws,
                                  ws_bytes);
}

int dmsh_argsort_u64(void* stream, uint32_t n, uint64_t* keys_inout, uint32_t* idx_out,
                     uint32_t key_bits, void* ws, size_t ws_bytes) {
  return sortk::argsort<uint64_t>(as_stream(stream), n, keys_inout, idx_out, key_bits, ws,
                                  ws_bytes);
}

int dmsh_bvhnodes_from_sorted_mortons_u32(void* stream, uint32_t n, uint32_t* bvhnodes,
                                          const uint32_t* idx2morton, const uint32_t* idx2obj) {
  if (n == 0) return DMSH_OK;
  if (!bvhnodes || !idx2morton || !idx2obj || n >= (1u << 31)) return DMSH_ERR_INVALID;
  return launch_topology<uint32_t>(as_stream(stream), n, bvhnodes, idx2morton, idx2obj);
}

int dmsh_bvhnodes_from_sorted_mortons_u64(void* stream, uint32_t n, uint32_t* bvhnodes,
                                          const uint64_t* idx2morton, const uint32_t* idx2obj) {
  if (n == 0) return DMSH_OK;
  if (!bvhnodes || !idx2morton || !idx2obj || n >= (1u << 31)) return DMSH_ERR_INVALID;
  return launch_topology<uint64_t>(as_stream(stream), n, bvhnodes, idx2morton, idx2obj);
}

size_t dmsh_refit_workspace_bytes(uint32_t num_leaf) {
  return standalone_workspace_bytes(num_leaf);
}

static int refit_common(void* stream, bool tri, float* aabb, uint32_t num_bvhnode,
                        const uint32_t* bvhnodes, uint32_t num_leaf, const uint32_t* tri2vtx,
                        uint32_t num_vtx, const float* xyz, float eps, void* ws,
                        size_t ws_bytes) {
  if (num_leaf == 0) return num_bvhnode == 0 ? DMSH_OK : DMSH_ERR_INVALID;
  if (num_leaf >= (1u << 31) || num_bvhnode != 2 * num_leaf - 1) return DMSH_ERR_INVALID;
  if (!aabb || !bvhnodes || !xyz || (tri && !tri2vtx)) return DMSH_ERR_INVALID;
  if (!ws || ws_bytes < dmsh_refit_workspace_bytes(num_leaf)) return DMSH_ERR_WORKSPACE;
  uint32_t* err = device_error_word();
  if (!err) return DMSH_ERR_HIP;
  hipStream_t s = as_stream(stream);
  if (tri) {
    const TreeView v{aabb, bvhnodes, num_leaf, tri2vtx, num_vtx, xyz};
    return launch_refit_standalone<kLeafTri>(s, v, eps, ws, ws_bytes, err);
  }
  const TreeView v{aabb, bvhnodes, num_leaf, nullptr, num_leaf, xyz};
  return launch_refit_standalone<kLeafPoint>(s, v, 0.f, ws, ws_bytes, err);
}

int dmsh_bvhnode2aabb_trimesh3(void* stream, float* bvhnode2aabb, uint32_t num_bvhnode,
                               const uint32_t* bvhnodes, uint32_t num_tri,
                               const uint32_t* tri2vtx, uint32_t num_vtx, const float* vtx2xyz,
                               float eps, void* ws, size_t ws_bytes) {
  return refit_common(stream, true, bvhnode2aabb, num_bvhnode, bvhnodes, num_tri, tri2vtx,
                      num_vtx, vtx2xyz, eps, ws, ws_bytes);
}

int dmsh_bvhnode2aabb_points(void* stream, float* bvhnode2aabb, uint32_t num_bvhnode,
                             const uint32_t* bvhnodes, uint32_t num_pnt, const float* vtx2xyz,
                             void* ws, size_t ws_bytes) {
  return refit_common(stream, false, bvhnode2aabb, num_bvhnode, bvhnodes, num_pnt, nullptr,
                      num_pnt, vtx2xyz, 0.f, ws, ws_bytes);
}

size_t dmsh_build_workspace_bytes(uint32_t num_elem, uint32_t morton_bits) {
  if (morton_bits == 63) return bld::workspace_bytes<uint64_t>(num_elem);
  return bld::workspace_bytes<uint32_t>(num_elem);
}

int dmsh_build_trimesh3(void* stream, uint32_t num_tri, const uint32_t* tri2vtx, uint32_t num_vtx,
                        const float* vtx2xyz, uint32_t morton_bits, uint32_t* bvhnodes,
                        float* bvhnode2aabb, void* ws, size_t ws_bytes) {
  hipStream_t s = as_stream(stream);
  if (morton_bits == 30)
    return bld::build<uint32_t, true>(s, num_tri, tri2vtx, num_vtx, vtx2xyz, bvhnodes,
                                      bvhnode2aabb, ws, ws_bytes);
  if (morton_bits == 63)
    return bld::build<uint64_t, true>(s, num_tri, tri2vtx, num_vtx, vtx2xyz, bvhnodes,
                                      bvhnode2aabb, ws, ws_bytes);
  return DMSH_ERR_INVALID;
}

int dmsh_build_trimesh3_noalloc(void* stream, uint32_t num_tri, const uint32_t* tri2vtx,
                                uint32_t num_vtx, const float* vtx2xyz, uint32_t morton_bits,
                                uint32_t* bvhnodes, float* bvhnode2aabb) {
  if (morton_bits != 30 && morton_bits != 63) return DMSH_ERR_INVALID;
  if (num_tri == 0) return DMSH_OK;
  if (num_tri >= (1u << 30)) return DMSH_ERR_INVALID;
  // sized before anything is looked up or enqueued
  if (dmsh_build_workspace_bytes(num_tri, morton_bits) > (size_t)DMSH_NOALLOC_SCRATCH_BYTES)
    return DMSH_ERR_TOO_LARGE;
  void* ws = scratch_of_current_device();
  if (!ws) return DMSH_ERR_HIP;
  // the build clears what it needs zeroed (counters, histograms, tile status)
  // with its own kernel on every call
  return dmsh_build_trimesh3(stream, num_tri, tri2vtx, num_vtx, vtx2xyz, morton_bits, bvhnodes,
                             bvhnode2aabb, ws, (size_t)DMSH_NOALLOC_SCRATCH_BYTES);
}

int dmsh_build_points(void* stream, uint32_t num_pnt, const float* vtx2xyz, uint32_t morton_bits,
                      uint32_t* bvhnodes, float* bvhnode2aabb, void* ws, size_t ws_bytes) {
  hipStream_t s = as_stream(stream);
  if (morton_bits == 30)
    return bld::build<uint32_t, false>(s, num_pnt, nullptr, num_pnt, vtx2xyz, bvhnodes,
                                       bvhnode2aabb, ws, ws_bytes);
  if (morton_bits == 63)
    return bld::build<uint64_t, false>(s, num_pnt, nullptr, num_pnt, vtx2xyz, bvhnodes,
                                       bvhnode2aabb, ws, ws_bytes);
  return DMSH_ERR_INVALID;
}

}  // extern "C"
