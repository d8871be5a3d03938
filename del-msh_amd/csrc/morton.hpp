// morton.hpp -- Morton codes on gfx950.
//
// 30-bit (10 bits/axis) and 2D 32-bit codes follow del-msh-cpu/src/mortons.rs:3-102;
// the 63-bit code (21 bits/axis, u64) is an extension with the same
// construction.  Quantization is `clamp(x * 2^b, 0, 2^b - 1) as u32` with
// NaN -> 0, as Rust's clamp + saturating cast does.
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {

__device__ __forceinline__ uint32_t quantize(float v, float hi) {
  if (v != v) return 0u;
  v = (v < 0.f) ? 0.f : v;
  v = (v > hi) ? hi : v;
  return (uint32_t)v;
}

// 16 -> 32 bits, one zero between bits.
__device__ __forceinline__ uint32_t spread2(uint32_t x) {
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}
// 10 -> 30 bits, two zeros between bits.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  return (x | (x << 2)) & 0x09249249u;
}
// 21 -> 63 bits, two zeros between bits.
__device__ __forceinline__ uint64_t spread3_u64(uint32_t v) {
  uint64_t x = v & 0x1FFFFFu;
  x = (x | (x << 32)) & 0x001F00000000FFFFull;
  x = (x | (x << 16)) & 0x001F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ __forceinline__ uint32_t morton2_u32(float x, float y) {
  return spread2(quantize(x * 65536.f, 65535.f)) * 2u + spread2(quantize(y * 65536.f, 65535.f));
}
__device__ __forceinline__ uint32_t morton3_u32(float x, float y, float z) {
  return spread3(quantize(x * 1024.f, 1023.f)) * 4u + spread3(quantize(y * 1024.f, 1023.f)) * 2u +
         spread3(quantize(z * 1024.f, 1023.f));
}
__device__ __forceinline__ uint64_t morton3_u64(float x, float y, float z) {
  const float s = 2097152.f, hi = 2097151.f;
  return spread3_u64(quantize(x * s, hi)) * 4ull + spread3_u64(quantize(y * s, hi)) * 2ull +
         spread3_u64(quantize(z * s, hi));
}

template <typename KeyT>
__device__ __forceinline__ KeyT morton3(float x, float y, float z);
template <>
__device__ __forceinline__ uint32_t morton3<uint32_t>(float x, float y, float z) {
  return morton3_u32(x, y, z);
}
template <>
__device__ __forceinline__ uint64_t morton3<uint64_t>(float x, float y, float z) {
  return morton3_u64(x, y, z);
}

// Morton code of a 3D point under a column-major 4x4 transform.
template <typename KeyT>
__device__ __forceinline__ KeyT morton3_transformed(const float* T, float x, float y, float z) {
  float q[3];
  if (!transform_homogeneous(T, x, y, z, q)) q[0] = q[1] = q[2] = __builtin_nanf("");
  return morton3<KeyT>(q[0], q[1], q[2]);
}

// 2D: column-major 3x3 transform.
__device__ __forceinline__ uint32_t morton2_transformed(const float* T, float x, float y) {
  const float w = (T[2] * x + T[5] * y) + T[8];
  float qx = __builtin_nanf(""), qy = __builtin_nanf("");
  if (w != 0.f) {
    qx = ((T[0] * x + T[3] * y) + T[6]) / w;
    qy = ((T[1] * x + T[4] * y) + T[7]) / w;
  }
  return morton2_u32(qx, qy);
}

// One thread per point; the AoS [n][3] reads of a wave cover one contiguous
// 768-byte span, so each dword load instruction is fully coalesced.
template <typename KeyT>
__global__ __launch_bounds__(256) void k_vtx2morton3(uint32_t n, const float* __restrict__ co,
                                                     const float* __restrict__ T,
                                                     KeyT* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float m[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = T[k];
  const float* p = co + (size_t)i * 3;
  out[i] = morton3_transformed<KeyT>(m, p[0], p[1], p[2]);
}

static __global__ __launch_bounds__(256) void k_vtx2morton2(uint32_t n, const float* __restrict__ co,
                                                     const float* __restrict__ T,
                                                     uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float m[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) m[k] = T[k];
  const float* p = co + (size_t)i * 2;
  out[i] = morton2_transformed(m, p[0], p[1]);
}

}  // namespace dmsh
