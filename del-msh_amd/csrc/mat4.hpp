// mat4.hpp -- 4x4 inverse on the device (keeps the host free of a sync).
#pragma once

#include "dmsh_common.hpp"

namespace dmsh {

// Gauss-Jordan inverse with partial pivoting of a column-major 4x4 (pivot =
// first row with the largest |a|); returns false when singular.
__device__ __forceinline__ bool mat4_inverse(const float* m, float* inv) {
  float a[4][8];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][4 + c] = (r == c) ? 1.f : 0.f;
    }
  }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    float best = fabsf(a[c][c]);
    for (int r = c + 1; r < 4; ++r) {
      if (fabsf(a[r][c]) > best) {
        best = fabsf(a[r][c]);
        piv = r;
      }
    }
    if (best == 0.f) return false;
    if (piv != c) {
      for (int k = 0; k < 8; ++k) {
        const float tmp = a[c][k];
        a[c][k] = a[piv][k];
        a[piv][k] = tmp;
      }
    }
    const float ip = 1.f / a[c][c];
    for (int k = 0; k < 8; ++k) a[c][k] = a[c][k] * ip;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const float f = a[r][c];
      for (int k = 0; k < 8; ++k) a[r][k] = a[r][k] - f * a[c][k];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) inv[c * 4 + r] = a[r][4 + c];
  return true;
}

// One thread: inverse of the transform into inv[0..15], flag in inv[16].
static __global__ void k_mat4_inverse(const float* __restrict__ T, float* __restrict__ inv,
                                      uint32_t* __restrict__ err) {
  float m[16], r[16];
  for (int k = 0; k < 16; ++k) m[k] = T[k];
  const bool ok = mat4_inverse(m, r);
  for (int k = 0; k < 16; ++k) inv[k] = r[k];
  inv[16] = ok ? 1.f : 0.f;
  if (!ok) raise_error(err, DMSH_DEVERR_SINGULAR);
}

}  // namespace dmsh
