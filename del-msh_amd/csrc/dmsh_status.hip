// dmsh_status.hip -- status strings and the sticky device error word.
#include <mutex>

#include "dmsh_common.hpp"

// One word per device image: data errors found inside kernels are OR-ed in.
static __device__ uint32_t g_dmsh_device_errors;

namespace dmsh {

static constexpr int kMaxDevices = 64;
static uint32_t* g_err_ptr[kMaxDevices];
static std::mutex g_err_mutex;

uint32_t* device_error_word() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_err_mutex);
  if (!g_err_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_dmsh_device_errors)) != hipSuccess) return nullptr;
    g_err_ptr[dev] = static_cast<uint32_t*>(p);
  }
  return g_err_ptr[dev];
}

}  // namespace dmsh

extern "C" {

const char* dmsh_status_string(int status) {
  switch (status) {
    case DMSH_OK: return "ok";
    case DMSH_ERR_INVALID: return "invalid argument";
    case DMSH_ERR_HIP: return "HIP runtime error";
    case DMSH_ERR_WORKSPACE: return "workspace too small";
    default: return "unknown status";
  }
}

int dmsh_abi_version(void) { return 1; }

int dmsh_get_device_errors(uint32_t* bits_out) {
  if (!bits_out) return DMSH_ERR_INVALID;
  uint32_t* p = dmsh::device_error_word();
  if (!p) return DMSH_ERR_HIP;
  if (hipDeviceSynchronize() != hipSuccess) return DMSH_ERR_HIP;
  uint32_t v = 0;
  if (hipMemcpy(&v, p, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return DMSH_ERR_HIP;
  const uint32_t zero = 0;
  if (hipMemcpy(p, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return DMSH_ERR_HIP;
  *bits_out = v;
  return DMSH_OK;
}

}  // extern "C"
