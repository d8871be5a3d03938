// dmsh_common.hpp -- device helpers shared by the libdmsh HIP kernels (gfx950).
//
// Floating-point contract: every kernel is compiled with -ffp-contract=off and
// IEEE-correct binary32 division, so each expression below is rounded in the
// order written.  The geometry primitives restate the del-geo-core functions
// that del-msh-cpu calls (that crate is not vendored in the reference); the
// CPU oracle implements the same formula set independently.
//   transform_homogeneous  ((m3 x + m7 y) + m11 z) + m15 as w, w == 0 -> none
//   pixel ray              x = 2(iw+.5)/W - 1, y = 1 - 2(ih+.5)/H, z = +1 -> -1
//   ray-AABB               slab test, reciprocal form (Williams et al. 2005),
//                          non-finite reciprocal = parallel axis
//   ray-triangle           Moller-Trumbore 1997, |det| < FLT_EPSILON = miss,
//                          ray hit needs t > FLT_EPSILON
//   min / max              compare-and-select, ties keep the first operand
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstddef>
#include <cstdint>

#include "dmsh.h"

#define DMSH_U32MAX 0xFFFFFFFFu

namespace dmsh {

// ---------------------------------------------------------------------------
// host-side plumbing
// ---------------------------------------------------------------------------
inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Launch-error check after a kernel launch; returns a dmsh status code.
inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? DMSH_OK : DMSH_ERR_HIP;
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
__host__ __device__ __forceinline__ uint32_t ceil_div_u32(uint32_t a, uint32_t b) {
  return a / b + (a % b != 0u);
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.
struct Arena {
  char* base;
  size_t cap;
  size_t off;
  template <typename T>
  T* take(size_t count) {
    off = align_up(off, 256);
    T* p = reinterpret_cast<T*>(base + off);
    off += sizeof(T) * count;
    return p;
  }
  bool ok() const { return off <= cap; }
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fmin_sel(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_sel(float a, float b) { return (b > a) ? b : a; }

// Total order on binary32 (-0 below +0) for order-independent min/max
// reductions through integer atomics.
__device__ __forceinline__ uint32_t f32_order_key(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float f32_from_order_key(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  return __uint_as_float(u);
}

// transform_homogeneous for a column-major 4x4; returns false when w == 0.
__device__ __forceinline__ bool transform_homogeneous(const float* m, float x, float y,
                                                      float z, float out[3]) {
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  if (w == 0.f) return false;
  out[0] = (((m[0] * x + m[4] * y) + m[8] * z) + m[12]) / w;
  out[1] = (((m[1] * x + m[5] * y) + m[9] * z) + m[13]) / w;
  out[2] = (((m[2] * x + m[6] * y) + m[10] * z) + m[14]) / w;
  return true;
}

// Pixel ray (origin on NDC z = +1, direction to NDC z = -1).
__device__ __forceinline__ void ray_for_pixel(uint32_t iw, uint32_t ih, uint32_t img_w,
                                              uint32_t img_h, const float* T, float org[3],
                                              float dir[3]) {
  const float x0 = 2.f * ((float)iw + 0.5f) / (float)img_w - 1.f;
  const float y0 = 1.f - 2.f * ((float)ih + 0.5f) / (float)img_h;
  float q0[3], q1[3];
  if (!transform_homogeneous(T, x0, y0, 1.f, q0)) { q0[0] = q0[1] = q0[2] = __builtin_nanf(""); }
  if (!transform_homogeneous(T, x0, y0, -1.f, q1)) { q1[0] = q1[1] = q1[2] = __builtin_nanf(""); }
  org[0] = q0[0];
  org[1] = q0[1];
  org[2] = q0[2];
  dir[0] = q1[0] - q0[0];
  dir[1] = q1[1] - q0[1];
  dir[2] = q1[2] - q0[2];
}

// Per-ray constants of the slab test.
struct RaySlab {
  float o[3];
  float inv[3];
  bool par[3];  // axis with a non-finite reciprocal: treated as parallel
};

__device__ __forceinline__ void make_slab(const float org[3], const float dir[3], RaySlab& r) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    r.o[i] = org[i];
    r.inv[i] = 1.f / dir[i];
    r.par[i] = !(fabsf(r.inv[i]) < __builtin_inff());
  }
}

// Slab test against an AABB (lo[3], hi[3]); returns hit and (tmin, tmax).
__device__ __forceinline__ bool slab_hit(const RaySlab& r, const float lo[3], const float hi[3],
                                         float& tmin_out, float& tmax_out) {
  float tmin = -__builtin_inff(), tmax = __builtin_inff();
  bool miss = false;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (r.par[i]) {
      miss = miss || (r.o[i] < lo[i] || r.o[i] > hi[i]);
    } else {
      const float t0 = (lo[i] - r.o[i]) * r.inv[i];
      const float t1 = (hi[i] - r.o[i]) * r.inv[i];
      tmin = fmax_sel(tmin, fmin_sel(t0, t1));
      tmax = fmin_sel(tmax, fmax_sel(t0, t1));
    }
  }
  tmin_out = tmin;
  tmax_out = tmax;
  return !miss && !(tmin > tmax) && !(tmax < 0.f);
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Moller-Trumbore; `ray` requires t > FLT_EPSILON.  bc may be null.
__device__ __forceinline__ bool tri_intersect(const float p0[3], const float p1[3],
                                              const float p2[3], const float org[3],
                                              const float dir[3], bool ray, float& t_out,
                                              float* bc) {
  float e1[3], e2[3], pvec[3], tvec[3], qvec[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(dir, e2, pvec);
  const float det = dot3(e1, pvec);
  if (det > -FLT_EPSILON && det < FLT_EPSILON) return false;
  const float inv_det = 1.f / det;
#pragma unroll
  for (int i = 0; i < 3; ++i) tvec[i] = org[i] - p0[i];
  const float u = inv_det * dot3(tvec, pvec);
  if (u < 0.f || u > 1.f) return false;
  cross3(tvec, e1, qvec);
  const float v = inv_det * dot3(dir, qvec);
  if (v < 0.f || u + v > 1.f) return false;
  const float t = inv_det * dot3(e2, qvec);
  if (ray && !(t > FLT_EPSILON)) return false;
  t_out = t;
  if (bc) {
    bc[0] = 1.f - u - v;
    bc[1] = u;
    bc[2] = v;
  }
  return true;
}

// Wavefront helpers (wave64 on gfx950).
__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

}  // namespace dmsh

namespace dmsh {
// Per-device sticky error word (a __device__ variable in dmsh_status.hip);
// returns its device address for the current device, or nullptr on failure.
uint32_t* device_error_word();
// Compute units of the current device (cached per device; 256 if unknown).
int device_cu_count();

// Launch options of the build and refit kernels (process-wide, set through
// dmsh_set_tuning; results do not depend on them).
enum RefitOption {
  kRefitLeavesInAabb = 0,
  kRefitClimb = 1,
  kRefitQueue = 2,
  kRefitTop = 3,
  kRefitPlanCache = 4,
  kRefitCompact = 5,
  kSortWideTiles = 6,
  kRefitOptionCount = 7
};
int refit_option(RefitOption which);
bool set_refit_option(const char* key, int value);
int get_refit_option(const char* key);

// Opt-in kernel timing (dmsh_timing_enable): while on, a TimedScope records a
// HIP event on the launch stream before and after the kernels it encloses,
// filed under its label.  Off, it costs one relaxed atomic load.
bool timing_on();
void timing_mark(hipStream_t s, const char* label, bool begin);
struct TimedScope {
  hipStream_t s;
  const char* label;
  bool on;
  TimedScope(hipStream_t s_, const char* label_) : s(s_), label(label_), on(timing_on()) {
    if (on) timing_mark(s, label, true);
  }
  ~TimedScope() {
    if (on) timing_mark(s, label, false);
  }
  TimedScope(const TimedScope&) = delete;
  TimedScope& operator=(const TimedScope&) = delete;
};

// Every kernel launch of the library goes through here, filed under `label`
// (the kernel's name) when timing is on.
template <typename... KArgs, typename... Args>
inline void launch(const char* label, void (*kernel)(KArgs...), dim3 grid, dim3 block,
                   uint32_t shmem, hipStream_t s, Args&&... args) {
  TimedScope ts(s, label);
  hipLaunchKernelGGL(kernel, grid, block, shmem, s, static_cast<KArgs>(args)...);
}

__device__ __forceinline__ void raise_error(uint32_t* err, uint32_t bits) {
  if (err) atomicOr(err, bits);
}
}  // namespace dmsh
