// dmsh_status.hip -- status strings and the sticky device error word.
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dmsh_common.hpp"

// One word per device image: data errors found inside kernels are OR-ed in.
static __device__ uint32_t g_dmsh_device_errors;

namespace dmsh {

static constexpr int kMaxDevices = 64;
static uint32_t* g_err_ptr[kMaxDevices];
static std::mutex g_err_mutex;

// refit options: "refit.leaves_in_aabb" (0: the treelet kernel builds the
// leaf boxes; 1: the kernel before writes them), "refit.climb" (0: each
// level waits for its links, then its arrival counter, then the sibling's
// box; 1: one LDS round trip per level; 2: as 1 for trees of known-good
// layout, climbs starting at the nodes whose children are both leaves),
// "refit.queue" (1: workgroups take treelets from a shared counter; 0:
// treelet t goes to workgroup t mod grid), "refit.top" (0: the nodes above
// the treelets climb through LDS counters when there are at most 4096 of
// them; 1: they merge in barrier rounds), "refit.plan_cache" (1: a
// standalone refit whose workspace already holds the plan of the same rows
// skips the plan kernels; 0: always rebuild it), "refit.compact" (levels a
// climb runs before the climbs still going are queued again, 0..1023; 0:
// never), and the sort's "sort.wide_tiles" (digit passes in 512-thread
// tiles: 0 never, 1 from 2^21 elements, 2 always)
static std::atomic<int> g_refit_option[kRefitOptionCount] = {{0}, {2}, {1}, {0}, {1}, {0}, {1}};
static const char* const kRefitOptionKey[kRefitOptionCount] = {
    "refit.leaves_in_aabb", "refit.climb", "refit.queue", "refit.top", "refit.plan_cache",
    "refit.compact", "sort.wide_tiles"};
static const int kRefitOptionMax[kRefitOptionCount] = {1, 2, 1, 1, 1, 1023, 2};

int refit_option(RefitOption which) { return g_refit_option[which].load(); }

bool set_refit_option(const char* key, int value) {
  for (int i = 0; i < kRefitOptionCount; ++i) {
    if (std::strcmp(key, kRefitOptionKey[i]) == 0) {
      if (value < 0 || value > kRefitOptionMax[i]) return false;
      g_refit_option[i].store(value);
      return true;
    }
  }
  return false;
}

int get_refit_option(const char* key) {
  for (int i = 0; i < kRefitOptionCount; ++i)
    if (std::strcmp(key, kRefitOptionKey[i]) == 0) return g_refit_option[i].load();
  return -1;
}

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

int device_cu_count() {
  static std::atomic<int> cus[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
  int c = cus[dev].load(std::memory_order_relaxed);
  if (c > 0) return c;
  if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
    c = 256;
  cus[dev].store(c, std::memory_order_relaxed);
  return c;
}

// ---------------------------------------------------------------------------
// kernel timing
// ---------------------------------------------------------------------------
namespace {
struct EventPair {
  hipEvent_t start, stop;
};
std::atomic<bool> g_timing{false};
std::mutex g_timing_mutex;
std::map<std::string, std::vector<hipEvent_t>> g_open;     // started, not stopped
std::map<std::string, std::vector<EventPair>> g_recorded;  // complete launches

void destroy_all_events() {
  for (auto& kv : g_open)
    for (hipEvent_t e : kv.second) (void)hipEventDestroy(e);
  for (auto& kv : g_recorded)
    for (EventPair& p : kv.second) {
      (void)hipEventDestroy(p.start);
      (void)hipEventDestroy(p.stop);
    }
  g_open.clear();
  g_recorded.clear();
}
}  // namespace

bool timing_on() { return g_timing.load(std::memory_order_relaxed); }

void timing_mark(hipStream_t s, const char* label, bool begin) {
  std::lock_guard<std::mutex> lock(g_timing_mutex);
  if (begin) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    if (hipEventRecord(e, s) != hipSuccess) {
      (void)hipEventDestroy(e);
      return;
    }
    g_open[label].push_back(e);
    return;
  }
  auto it = g_open.find(label);
  if (it == g_open.end() || it->second.empty()) return;
  hipEvent_t start = it->second.back();
  it->second.pop_back();
  hipEvent_t stop = nullptr;
  if (hipEventCreate(&stop) != hipSuccess || hipEventRecord(stop, s) != hipSuccess) {
    if (stop) (void)hipEventDestroy(stop);
    (void)hipEventDestroy(start);
    return;
  }
  g_recorded[label].push_back(EventPair{start, stop});
}

}  // namespace dmsh

extern "C" {

int dmsh_timing_enable(int on) {
  std::lock_guard<std::mutex> lock(dmsh::g_timing_mutex);
  if (on) dmsh::destroy_all_events();
  dmsh::g_timing.store(on != 0);
  return DMSH_OK;
}

int dmsh_timing_query(const char* label, uint32_t* launches, double* total_ms, double* max_ms) {
  if (!label || !launches || !total_ms) return DMSH_ERR_INVALID;
  std::lock_guard<std::mutex> lock(dmsh::g_timing_mutex);
  *launches = 0;
  *total_ms = 0.0;
  if (max_ms) *max_ms = 0.0;
  auto it = dmsh::g_recorded.find(label);
  if (it == dmsh::g_recorded.end()) return DMSH_OK;
  double total = 0.0, mx = 0.0;
  for (const dmsh::EventPair& p : it->second) {
    if (hipEventSynchronize(p.stop) != hipSuccess) return DMSH_ERR_HIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.start, p.stop) != hipSuccess) return DMSH_ERR_HIP;
    total += ms;
    mx = ms > mx ? ms : mx;
  }
  *launches = (uint32_t)it->second.size();
  *total_ms = total;
  if (max_ms) *max_ms = mx;
  return DMSH_OK;
}

const char* dmsh_status_string(int status) {
  switch (status) {
    case DMSH_OK: return "ok";
    case DMSH_ERR_INVALID: return "invalid argument";
    case DMSH_ERR_HIP: return "HIP runtime error";
    case DMSH_ERR_WORKSPACE: return "workspace too small";
    case DMSH_ERR_TOO_LARGE:
      return "input too large for the allocation-free entry point's scratch "
             "(DMSH_NOALLOC_SCRATCH_BYTES)";
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
