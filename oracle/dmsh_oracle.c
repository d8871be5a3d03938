/*
 * dmsh_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see dmsh_oracle.h).
 *
 * Plain C restatement of the del-msh-cpu functions on the LBVH / ray-cast /
 * edge-gradient path.  Reference files are cited as del-msh-cpu/src/<file>:line.
 * Build with -ffp-contract=off: every float expression below is evaluated in
 * the order written, in binary32.
 *
 * Formula set for the del-geo-core primitives (that crate is not vendored in
 * the reference, so these are restated from published methods):
 *   F1 transform_homogeneous: w = ((m3 x + m7 y) + m11 z) + m15; w == 0 -> none;
 *      out_r = (((m_r x + m_{4+r} y) + m_{8+r} z) + m_{12+r}) / w.
 *   F2 pixel ray: x = 2 (iw + .5) / W - 1, y = 1 - 2 (ih + .5) / H;
 *      org = F1(T, (x,y,+1)), dir = F1(T, (x,y,-1)) - org.
 *   F3 ray-AABB slab test (Kay & Kajiya 1986; reciprocal form of Williams et
 *      al. 2005): inv = 1/d per axis; an axis whose inv is not finite is a
 *      parallel axis (miss iff o outside [lo,hi]); otherwise t0 = (lo-o) inv,
 *      t1 = (hi-o) inv; tmin = max(tmin, min(t0,t1)), tmax = min(tmax,
 *      max(t0,t1)) from (-inf, +inf); miss iff tmin > tmax or tmax < 0.
 *   F4 ray-triangle: Moller & Trumbore 1997 (non-culling branch), |det| <
 *      FLT_EPSILON -> miss, 0 <= u, v; u + v <= 1; ray hit needs t > FLT_EPSILON.
 *   F5 min/max are compare-and-select: min(a,b) = b < a ? b : a,
 *      max(a,b) = b > a ? b : a (ties keep the first operand).
 */
#include "dmsh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define U32MAX 0xFFFFFFFFu

static inline float fmin_sel(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_sel(float a, float b) { return (b > a) ? b : a; }

/* ======================================================================= */
/* Morton codes                                                             */
/* ======================================================================= */

/* mortons.rs:3-8 */
uint32_t orc_expand_bits2(uint32_t x) {
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}

/* Rust `(v).clamp(lo, hi) as u32`: NaN stays NaN through clamp and the
 * saturating cast maps NaN to 0. */
static inline uint32_t clamp_to_u32(float v, float hi) {
  if (v != v) return 0u;
  if (v < 0.f) v = 0.f;
  if (v > hi) v = hi;
  return (uint32_t)v;
}

/* mortons.rs:20-27 */
uint32_t orc_morton_code2(float x, float y) {
  uint32_t ix = clamp_to_u32(x * 65536.f, 65535.f);
  uint32_t iy = clamp_to_u32(y * 65536.f, 65535.f);
  ix = orc_expand_bits2(ix);
  iy = orc_expand_bits2(iy);
  return ix * 2u + iy;
}

/* mortons.rs:76-81 */
uint32_t orc_expand_bits3(uint32_t x) {
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  return (x | (x << 2)) & 0x09249249u;
}

/* mortons.rs:94-102 */
uint32_t orc_morton_code3(float x, float y, float z) {
  uint32_t ix = clamp_to_u32(x * 1024.f, 1023.f);
  uint32_t iy = clamp_to_u32(y * 1024.f, 1023.f);
  uint32_t iz = clamp_to_u32(z * 1024.f, 1023.f);
  ix = orc_expand_bits3(ix);
  iy = orc_expand_bits3(iy);
  iz = orc_expand_bits3(iz);
  return ix * 4u + iy * 2u + iz;
}

/* 63-bit extension: the same construction with 21 bits per axis.  Written
 * as a plain bit loop (one source bit at a time) so that it is independent
 * of the magic-mask form the GPU kernel uses. */
uint64_t orc_expand_bits3_u64(uint32_t x) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) {
    if ((x >> b) & 1u) r |= (uint64_t)1 << (3 * b);
  }
  return r;
}

uint64_t orc_morton_code3_u64(float x, float y, float z) {
  const float scale = 2097152.f; /* 2^21 */
  const float hi = 2097151.f;    /* 2^21 - 1, exact in binary32 */
  uint64_t ix = orc_expand_bits3_u64(clamp_to_u32(x * scale, hi));
  uint64_t iy = orc_expand_bits3_u64(clamp_to_u32(y * scale, hi));
  uint64_t iz = orc_expand_bits3_u64(clamp_to_u32(z * scale, hi));
  return ix * 4u + iy * 2u + iz;
}

/* ======================================================================= */
/* geometry formula set                                                     */
/* ======================================================================= */

/* F1: del-geo-core mat4_col_major::transform_homogeneous */
int orc_transform_homogeneous(const float m[16], const float p[3], float out[3]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  if (w == 0.f) return 0;
  for (int r = 0; r < 3; ++r) {
    out[r] = (((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r]) / w;
  }
  return 1;
}

/* mat3_col_major::transform_homogeneous, 2D points */
static int transform_homogeneous2(const float m[9], const float p[2], float out[2]) {
  const float x = p[0], y = p[1];
  const float w = (m[2] * x + m[5] * y) + m[8];
  if (w == 0.f) return 0;
  for (int r = 0; r < 2; ++r) out[r] = ((m[r] * x + m[3 + r] * y) + m[6 + r]) / w;
  return 1;
}

/* mortons.rs:171-212 */
int orc_vtx2morton(size_t num_vtx, const float *vtx2co, int num_dim,
                   const float *transform, uint32_t *vtx2morton) {
  if (num_dim == 2) {
    for (size_t i = 0; i < num_vtx; ++i) {
      float q[2];
      if (!transform_homogeneous2(transform, vtx2co + i * 2, q)) { q[0] = NAN; q[1] = NAN; }
      vtx2morton[i] = orc_morton_code2(q[0], q[1]);
    }
    return 0;
  }
  if (num_dim == 3) {
    for (size_t i = 0; i < num_vtx; ++i) {
      float q[3];
      if (!orc_transform_homogeneous(transform, vtx2co + i * 3, q)) { q[0] = q[1] = q[2] = NAN; }
      vtx2morton[i] = orc_morton_code3(q[0], q[1], q[2]);
    }
    return 0;
  }
  return -1;
}

int orc_vtx2morton_u64(size_t num_vtx, const float *vtx2xyz,
                       const float *transform, uint64_t *vtx2morton) {
  for (size_t i = 0; i < num_vtx; ++i) {
    float q[3];
    if (!orc_transform_homogeneous(transform, vtx2xyz + i * 3, q)) { q[0] = q[1] = q[2] = NAN; }
    vtx2morton[i] = orc_morton_code3_u64(q[0], q[1], q[2]);
  }
  return 0;
}

/* F2: del-geo-core ray_from_transform_ndc2world_and_pixel_coordinates, in the
 * form the reference GPU kernel spells out (ray_for_pixel.h:3-19). */
void orc_ray_for_pixel(uint32_t iw, uint32_t ih, uint32_t img_w, uint32_t img_h,
                       const float T[16], float org[3], float dir[3]) {
  const float x0 = 2.f * ((float)iw + 0.5f) / (float)img_w - 1.f;
  const float y0 = 1.f - 2.f * ((float)ih + 0.5f) / (float)img_h;
  const float a[3] = {x0, y0, +1.f};
  const float b[3] = {x0, y0, -1.f};
  float q0[3], q1[3];
  if (!orc_transform_homogeneous(T, a, q0)) { q0[0] = q0[1] = q0[2] = NAN; }
  if (!orc_transform_homogeneous(T, b, q1)) { q1[0] = q1[1] = q1[2] = NAN; }
  for (int i = 0; i < 3; ++i) {
    org[i] = q0[i];
    dir[i] = q1[i] - q0[i];
  }
}

/* F3: slab test.  The reciprocal is recomputed here from dir so that the
 * function stands alone; 1/d is correctly rounded so this equals caching it. */
int orc_ray_aabb(const float aabb[6], const float org[3], const float dir[3],
                 float *tmin_out, float *tmax_out) {
  float tmin = -INFINITY, tmax = INFINITY;
  for (int i = 0; i < 3; ++i) {
    const float inv = 1.f / dir[i];
    if (!(fabsf(inv) < INFINITY)) {
      if (org[i] < aabb[i] || org[i] > aabb[i + 3]) return 0;
      continue;
    }
    const float t0 = (aabb[i] - org[i]) * inv;
    const float t1 = (aabb[i + 3] - org[i]) * inv;
    tmin = fmax_sel(tmin, fmin_sel(t0, t1));
    tmax = fmin_sel(tmax, fmax_sel(t0, t1));
  }
  if (tmin > tmax) return 0;
  if (tmax < 0.f) return 0;
  *tmin_out = tmin;
  *tmax_out = tmax;
  return 1;
}

static inline float dot3(const float a[3], const float b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline void cross3(const float a[3], const float b[3], float c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* F4: Moller-Trumbore; `ray` selects the t > FLT_EPSILON requirement */
static int mt_intersect(const float p0[3], const float p1[3], const float p2[3],
                        const float org[3], const float dir[3], int ray,
                        float *t_out, float bc[3]) {
  float e1[3], e2[3], pvec[3], tvec[3], qvec[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(dir, e2, pvec);
  const float det = dot3(e1, pvec);
  if (det > -FLT_EPSILON && det < FLT_EPSILON) return 0;
  const float inv_det = 1.f / det;
  for (int i = 0; i < 3; ++i) tvec[i] = org[i] - p0[i];
  const float u = inv_det * dot3(tvec, pvec);
  if (u < 0.f || u > 1.f) return 0;
  cross3(tvec, e1, qvec);
  const float v = inv_det * dot3(dir, qvec);
  if (v < 0.f || u + v > 1.f) return 0;
  const float t = inv_det * dot3(e2, qvec);
  if (ray && !(t > FLT_EPSILON)) return 0;
  *t_out = t;
  if (bc) {
    bc[0] = 1.f - u - v;
    bc[1] = u;
    bc[2] = v;
  }
  return 1;
}

int orc_ray_tri(const float p0[3], const float p1[3], const float p2[3],
                const float org[3], const float dir[3], float *t, float bc[3]) {
  return mt_intersect(p0, p1, p2, org, dir, 1, t, bc);
}
int orc_line_tri(const float p0[3], const float p1[3], const float p2[3],
                 const float org[3], const float dir[3], float *t, float bc[3]) {
  return mt_intersect(p0, p1, p2, org, dir, 0, t, bc);
}

/* mat4_col_major::try_inverse_with_pivot, restated as Gauss-Jordan
 * elimination with partial pivoting on the augmented matrix [A | I]
 * (A[r][c] = m[c*4+r]); pivot = first row with the largest |A[r][c]|. */
int orc_mat4_inverse(const float m[16], float inv[16]) {
  float a[4][8];
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < 4; ++c) {
      a[r][c] = m[c * 4 + r];
      a[r][4 + c] = (r == c) ? 1.f : 0.f;
    }
  }
  for (int c = 0; c < 4; ++c) {
    int piv = c;
    float best = fabsf(a[c][c]);
    for (int r = c + 1; r < 4; ++r) {
      if (fabsf(a[r][c]) > best) { best = fabsf(a[r][c]); piv = r; }
    }
    if (best == 0.f) return 0;
    if (piv != c) {
      for (int k = 0; k < 8; ++k) { float tmp = a[c][k]; a[c][k] = a[piv][k]; a[piv][k] = tmp; }
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
  return 1;
}

/* elem2center.rs:8-37: cog starts at +0 and adds the nodes in order, then
 * is scaled by ratio = 1/num_node. */
void orc_tri2centroid(size_t num_tri, const uint32_t *tri2vtx,
                      const float *vtx2xyz, float *tri2cntr) {
  const float ratio = 1.f / 3.f;
  for (size_t t = 0; t < num_tri; ++t) {
    for (int d = 0; d < 3; ++d) {
      float cog = 0.f;
      for (int k = 0; k < 3; ++k) cog += vtx2xyz[(size_t)tri2vtx[t * 3 + k] * 3 + d];
      tri2cntr[t * 3 + d] = cog * ratio;
    }
  }
}

/* Total order on binary32 that puts -0 below +0 (used for the bounding-box
 * reduction so the result does not depend on reduction order). */
static inline uint32_t f32_order_key(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* Mat44/torch.py:120-140 from_fit_vtx2xyz_into_unit_cube:
 * s = 1 / max(extent), translation = -center * s + 0.5 */
void orc_fit_into_unit_cube(size_t num_pnt, const float *p, float out[16]) {
  float mn[3], mx[3];
  for (int d = 0; d < 3; ++d) { mn[d] = p[d]; mx[d] = p[d]; }
  for (size_t i = 1; i < num_pnt; ++i) {
    for (int d = 0; d < 3; ++d) {
      const float v = p[i * 3 + d];
      if (f32_order_key(v) < f32_order_key(mn[d])) mn[d] = v;
      if (f32_order_key(v) > f32_order_key(mx[d])) mx[d] = v;
    }
  }
  float ext = mx[0] - mn[0];
  for (int d = 1; d < 3; ++d) ext = fmax_sel(ext, mx[d] - mn[d]);
  const float s = 1.f / ext;
  memset(out, 0, sizeof(float) * 16);
  out[0] = s;
  out[5] = s;
  out[10] = s;
  out[15] = 1.f;
  for (int d = 0; d < 3; ++d) {
    const float c = (mn[d] + mx[d]) * 0.5f;
    out[12 + d] = (-c) * s + 0.5f;
  }
}

/* ======================================================================= */
/* stable argsort                                                           */
/* ======================================================================= */

/* Bottom-up merge sort of indices by key: stable, like Rust's slice
 * `sort_by_key` used at mortons.rs:143. */
#define DEFINE_ARGSORT(NAME, KT)                                                  \
  void NAME(size_t n, const KT *keys, uint32_t *idx2obj, KT *idx2key) {          \
    uint32_t *a = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));            \
    uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));            \
    for (size_t i = 0; i < n; ++i) a[i] = (uint32_t)i;                          \
    for (size_t w = 1; w < n; w *= 2) {                                          \
      for (size_t lo = 0; lo < n; lo += 2 * w) {                                 \
        size_t mid = lo + w < n ? lo + w : n;                                    \
        size_t hi = lo + 2 * w < n ? lo + 2 * w : n;                             \
        size_t i = lo, j = mid, k = lo;                                          \
        while (i < mid && j < hi) {                                              \
          if (keys[a[j]] < keys[a[i]]) b[k++] = a[j++];                          \
          else b[k++] = a[i++];                                                  \
        }                                                                        \
        while (i < mid) b[k++] = a[i++];                                         \
        while (j < hi) b[k++] = a[j++];                                          \
      }                                                                          \
      uint32_t *tmp = a; a = b; b = tmp;                                         \
    }                                                                            \
    for (size_t i = 0; i < n; ++i) {                                             \
      idx2obj[i] = a[i];                                                         \
      if (idx2key) idx2key[i] = keys[a[i]];                                      \
    }                                                                            \
    free(a);                                                                     \
    free(b);                                                                     \
  }
DEFINE_ARGSORT(orc_argsort_u32, uint32_t)
DEFINE_ARGSORT(orc_argsort_u64, uint64_t)

/* ======================================================================= */
/* Karras topology                                                          */
/* ======================================================================= */

static inline int64_t clz32(uint32_t v) { return v ? (int64_t)__builtin_clz(v) : 32; }
static inline int64_t clz64(uint64_t v) { return v ? (int64_t)__builtin_clzll(v) : 64; }

/* mortons.rs:216-218 `delta`, mortons.rs:221-288 range,
 * mortons.rs:290-325 split; generated for u32 (30-bit codes, nbits 32) and
 * u64 (63-bit extension, nbits 64) keys. */
#define DEFINE_TOPOLOGY(SUF, KT, CLZ, NBITS)                                      \
  static inline int64_t delta_##SUF(size_t i0, size_t i1, const KT *c) {          \
    return CLZ(c[i0] ^ c[i1]);                                                    \
  }                                                                               \
  void orc_range_##SUF(const KT *c, size_t n, size_t idx1, size_t *first,          \
                       size_t *last) {                                            \
    if (idx1 == 0) { *first = 0; *last = n - 1; return; }                        \
    if (idx1 == n - 1) { *first = n - 1; *last = n - 1; return; }                \
    const KT mc0 = c[idx1 - 1], mc1 = c[idx1], mc2 = c[idx1 + 1];                 \
    if (mc0 == mc1 && mc1 == mc2) {                                               \
      size_t jdx = idx1 + 1;                                                      \
      while (jdx < n - 1) {                                                       \
        jdx += 1;                                                                 \
        if (c[jdx] != mc1) { *first = idx1; *last = jdx - 1; return; }           \
      }                                                                           \
      *first = idx1; *last = jdx; return;                                         \
    }                                                                             \
    int64_t d = delta_##SUF(idx1, idx1 + 1, c) - delta_##SUF(idx1, idx1 - 1, c);  \
    d = d > 0 ? 1 : -1;                                                           \
    const int64_t delta_min = delta_##SUF(idx1, (size_t)((int64_t)idx1 - d), c);  \
    int64_t lmax = 2;                                                             \
    for (;;) {                                                                    \
      const int64_t jdx = (int64_t)idx1 + lmax * d;                               \
      if (jdx < 0 || jdx >= (int64_t)n) break;                                    \
      if (delta_##SUF(idx1, (size_t)jdx, c) <= delta_min) break;                  \
      lmax *= 2;                                                                  \
    }                                                                             \
    int64_t l = 0;                                                                \
    for (int64_t t = lmax / 2; t >= 1; t /= 2) {                                  \
      const int64_t jdx = (int64_t)idx1 + (l + t) * d;                            \
      if (jdx >= 0 && jdx < (int64_t)n &&                                         \
          delta_##SUF(idx1, (size_t)jdx, c) > delta_min) {                        \
        l += t;                                                                   \
      }                                                                           \
    }                                                                             \
    const size_t jdx = (size_t)((int64_t)idx1 + l * d);                           \
    if (idx1 <= jdx) { *first = idx1; *last = jdx; }                              \
    else { *first = jdx; *last = idx1; }                                          \
  }                                                                               \
  size_t orc_split_##SUF(const KT *c, size_t i_start, size_t i_end) {             \
    if (i_start == i_end) return (size_t)-1;                                      \
    const KT mc_start = c[i_start];                                               \
    const int64_t nbitcommon0 = CLZ(mc_start ^ c[i_end]);                         \
    if (nbitcommon0 == NBITS) return i_start;                                     \
    size_t split = i_start;                                                       \
    size_t step = i_end - i_start;                                                \
    while (step > 1) {                                                            \
      step = (step + 1) / 2;                                                      \
      const size_t inew = split + step;                                           \
      if (inew >= i_end) continue;                                                \
      if (CLZ(mc_start ^ c[inew]) > nbitcommon0) split = inew;                    \
    }                                                                             \
    return split;                                                                 \
  }                                                                               \
  int orc_bvhnodes_##SUF(uint32_t *bvhnodes, const uint32_t *idx2obj,             \
                         const KT *c, size_t n) {                                 \
    if (n == 0) return -1;                                                        \
    bvhnodes[0] = U32MAX;                                                         \
    const size_t num_branch = n - 1;                                              \
    for (size_t ib = 0; ib < num_branch; ++ib) {                                  \
      size_t r0, r1;                                                              \
      orc_range_##SUF(c, n, ib, &r0, &r1);                                        \
      const size_t isplit = orc_split_##SUF(c, r0, r1);                           \
      if (isplit == (size_t)-1) return -2;                                        \
      if (r0 == isplit) {                                                         \
        const size_t a = num_branch + isplit;                                     \
        bvhnodes[ib * 3 + 1] = (uint32_t)a;                                       \
        bvhnodes[a * 3] = (uint32_t)ib;                                           \
        bvhnodes[a * 3 + 1] = idx2obj[isplit];                                    \
        bvhnodes[a * 3 + 2] = U32MAX;                                             \
      } else {                                                                    \
        bvhnodes[ib * 3 + 1] = (uint32_t)isplit;                                  \
        bvhnodes[isplit * 3] = (uint32_t)ib;                                      \
      }                                                                           \
      if (r1 == isplit + 1) {                                                     \
        const size_t b = num_branch + isplit + 1;                                 \
        bvhnodes[ib * 3 + 2] = (uint32_t)b;                                       \
        bvhnodes[b * 3] = (uint32_t)ib;                                           \
        bvhnodes[b * 3 + 1] = idx2obj[isplit + 1];                                \
        bvhnodes[b * 3 + 2] = U32MAX;                                             \
      } else {                                                                    \
        bvhnodes[ib * 3 + 2] = (uint32_t)(isplit + 1);                            \
        bvhnodes[(isplit + 1) * 3] = (uint32_t)ib;                                \
      }                                                                           \
    }                                                                             \
    if (n == 1) { bvhnodes[1] = idx2obj[0]; bvhnodes[2] = U32MAX; }              \
    return 0;                                                                     \
  }
DEFINE_TOPOLOGY(u32, uint32_t, clz32, 32)
DEFINE_TOPOLOGY(u64, uint64_t, clz64, 64)

/* mortons.rs:330-352 */
int orc_check_morton_code_range_split_u32(const uint32_t *c, size_t n) {
  if (n == 0) return 1;
  for (size_t i = 0; i + 1 < n; ++i)
    if (c[i] > c[i + 1]) return 2;
  for (size_t ini = 0; ini + 1 < n; ++ini) {
    size_t r0, r1, a0, a1, b0, b1;
    orc_range_u32(c, n, ini, &r0, &r1);
    const size_t isplit = orc_split_u32(c, r0, r1);
    orc_range_u32(c, n, isplit, &a0, &a1);
    orc_range_u32(c, n, isplit + 1, &b0, &b1);
    if (r0 != a0) return 3;
    if (r1 != b1) return 4;
    const size_t last1 = (isplit == r0) ? isplit : a1;
    const size_t first1 = (isplit + 1 == r1) ? isplit + 1 : b0;
    if (last1 + 1 != first1) return 5;
  }
  return 0;
}

static void increment_leaf(const uint32_t *bn, size_t inode, size_t num_idx,
                           size_t *flag) {
  const size_t c0 = bn[inode * 3 + 1];
  if (c0 >= num_idx - 1) flag[c0 - (num_idx - 1)] += 1;
  else increment_leaf(bn, c0, num_idx, flag);
  const size_t c1 = bn[inode * 3 + 2];
  if (c1 >= num_idx - 1) flag[c1 - (num_idx - 1)] += 1;
  else increment_leaf(bn, c1, num_idx, flag);
}

/* mortons.rs:384-433 (checks the branch part of bvhnodes, rows 0..n-2) */
int orc_check_binary_radix_tree_u32(const uint32_t *bn, const uint32_t *c, size_t n) {
  if (n < 2) return 0;
  size_t *flag = (size_t *)calloc(n, sizeof(size_t));
  increment_leaf(bn, 0, n, flag);
  int ret = 0;
  for (size_t i = 0; i < n; ++i)
    if (flag[i] != 1) { ret = 1; break; }
  free(flag);
  if (ret) return ret;
  for (size_t ib = 0; ib + 1 < n; ++ib) {
    const size_t il = bn[ib * 3 + 1];
    const size_t isplit = il >= n - 1 ? il - (n - 1) : il;
    size_t r0, r1;
    orc_range_u32(c, n, ib, &r0, &r1);
    if (orc_split_u32(c, r0, r1) != isplit) return 2;
  }
  return 0;
}

/* bvhnodes.rs:5-36 mark_child / check_bvh_topology (iterative walk) */
int orc_check_bvh_topology(const uint32_t *bn, size_t num_leaf) {
  if (num_leaf == 0) return 1;
  const size_t num_node = 2 * num_leaf - 1;
  size_t *cnt = (size_t *)calloc(num_leaf, sizeof(size_t));
  size_t *stack = (size_t *)malloc(sizeof(size_t) * (num_node + 1));
  size_t sp = 0;
  int ret = 0;
  stack[sp++] = 0;
  while (sp > 0 && !ret) {
    const size_t nd = stack[--sp];
    if (nd >= num_node) { ret = 2; break; }
    if (bn[nd * 3 + 2] == U32MAX) {
      const size_t obj = bn[nd * 3 + 1];
      if (obj >= num_leaf) { ret = 3; break; }
      cnt[obj] += 1;
      continue;
    }
    if (sp + 2 > num_node + 1) { ret = 4; break; }
    stack[sp++] = bn[nd * 3 + 2];
    stack[sp++] = bn[nd * 3 + 1];
  }
  for (size_t i = 0; i < num_leaf && !ret; ++i)
    if (cnt[i] != 1) ret = 5;
  for (size_t i = 0; i < num_leaf && !ret; ++i)
    if (bn[(num_leaf - 1 + i) * 3 + 2] != U32MAX) ret = 6;
  free(cnt);
  free(stack);
  return ret;
}

/* ======================================================================= */
/* refit                                                                    */
/* ======================================================================= */

/* aabb3::set_as_cube + aabb3::add_point (vtx2xyz.rs:39-61, eps margin) */
static inline void aabb_set_point(float *aabb, const float *p, float eps) {
  for (int d = 0; d < 3; ++d) { aabb[d] = p[d] - eps; aabb[d + 3] = p[d] + eps; }
}
static inline void aabb_add_point(float *aabb, const float *p, float eps) {
  for (int d = 0; d < 3; ++d) {
    aabb[d] = fmin_sel(aabb[d], p[d] - eps);
    aabb[d + 3] = fmax_sel(aabb[d + 3], p[d] + eps);
  }
}
/* aabb3::from_two_aabbs */
static inline void aabb_merge(float *out, const float *a, const float *b) {
  for (int d = 0; d < 3; ++d) {
    out[d] = fmin_sel(a[d], b[d]);
    out[d + 3] = fmax_sel(a[d + 3], b[d + 3]);
  }
}

typedef struct {
  float *aabb;
  const uint32_t *bn;
  const uint32_t *tri2vtx; /* NULL: leaves store point indices */
  const float *xyz;
  float eps;
  int err;
} refit_ctx;

/* bvhnode2aabb3.rs:68-142 (triangles) / :6-63 (points): recursion into
 * child 1 then child 2, then merge. */
static void refit_rec(refit_ctx *cx, size_t node) {
  const uint32_t *bn = cx->bn;
  float *box = cx->aabb + node * 6;
  if (bn[node * 3 + 2] == U32MAX) {
    const size_t e = bn[node * 3 + 1];
    if (cx->tri2vtx) {
      const uint32_t *v = cx->tri2vtx + e * 3;
      aabb_set_point(box, cx->xyz + (size_t)v[0] * 3, cx->eps);
      aabb_add_point(box, cx->xyz + (size_t)v[1] * 3, cx->eps);
      aabb_add_point(box, cx->xyz + (size_t)v[2] * 3, cx->eps);
    } else {
      aabb_set_point(box, cx->xyz + e * 3, 0.f);
    }
    return;
  }
  const size_t c0 = bn[node * 3 + 1], c1 = bn[node * 3 + 2];
  if (bn[c0 * 3] != node || bn[c1 * 3] != node) { cx->err = 1; return; }
  refit_rec(cx, c0);
  refit_rec(cx, c1);
  aabb_merge(box, cx->aabb + c0 * 6, cx->aabb + c1 * 6);
}

int orc_bvhnode2aabb_trimesh3(float *aabb, const uint32_t *bn, size_t num_node,
                              const uint32_t *tri2vtx, const float *vtx2xyz,
                              float eps) {
  if (num_node == 0) return -1;
  refit_ctx cx = {aabb, bn, tri2vtx, vtx2xyz, eps, 0};
  refit_rec(&cx, 0);
  return cx.err ? -2 : 0;
}

int orc_bvhnode2aabb_points(float *aabb, const uint32_t *bn, size_t num_node,
                            const float *vtx2xyz) {
  if (num_node == 0) return -1;
  refit_ctx cx = {aabb, bn, NULL, vtx2xyz, 0.f, 0};
  refit_rec(&cx, 0);
  return cx.err ? -2 : 0;
}

/* ======================================================================= */
/* ray-cast                                                                 */
/* ======================================================================= */

typedef struct {
  const float *org, *dir;
  const uint32_t *tri2vtx;
  const float *xyz;
  const uint32_t *bn;
  const float *aabb;
  orc_trace_stats st;
  int frames, max_frames; /* nesting of both-children-hit nodes on the path */
} cast_ctx;

typedef struct {
  float t;
  uint32_t tri;
  float bc[3];
} hit_rec;

/* Rust tuple `<` on (tmin, tmax): lexicographic with partial_cmp. */
static inline int tuple_lt(float a0, float a1, float b0, float b1) {
  if (a0 == b0) return a1 < b1;
  return a0 < b0;
}

/* search_bvh3.rs:103-221 first_intersection_ray with dis = +inf */
static int fir(cast_ctx *cx, size_t node, hit_rec *out) {
  float tn0, tn1;
  if (!orc_ray_aabb(cx->aabb + node * 6, cx->org, cx->dir, &tn0, &tn1)) return 0;
  const uint32_t *bn = cx->bn;
  if (bn[node * 3 + 2] == U32MAX) {
    const uint32_t tri = bn[node * 3 + 1];
    const uint32_t *v = cx->tri2vtx + (size_t)tri * 3;
    float t, bc[3];
    cx->st.tris_tested++;
    if (!orc_ray_tri(cx->xyz + (size_t)v[0] * 3, cx->xyz + (size_t)v[1] * 3,
                     cx->xyz + (size_t)v[2] * 3, cx->org, cx->dir, &t, bc))
      return 0;
    if (!(t < INFINITY)) return 0;
    out->t = t;
    out->tri = tri;
    memcpy(out->bc, bc, sizeof(bc));
    return 1;
  }
  cx->st.nodes_visited++; /* an internal node whose two child boxes are tested */
  const size_t left = bn[node * 3 + 1], right = bn[node * 3 + 2];
  float tl0, tl1, tr0, tr1;
  const int hl = orc_ray_aabb(cx->aabb + left * 6, cx->org, cx->dir, &tl0, &tl1);
  const int hr = orc_ray_aabb(cx->aabb + right * 6, cx->org, cx->dir, &tr0, &tr1);
  int near_is_left;
  if (hl && hr) near_is_left = tuple_lt(tl0, tl1, tr0, tr1);
  else if (!hl && !hr) return 0;
  else near_is_left = hl;
  const size_t nnear = near_is_left ? left : right;
  const size_t nfar = near_is_left ? right : left;
  hit_rec rn, rf;
  const int both = hl && hr;
  if (both && ++cx->frames > cx->max_frames) cx->max_frames = cx->frames;
  const int has_n = fir(cx, nnear, &rn);
  if (has_n) {
    float tf0, tf1;
    if (orc_ray_aabb(cx->aabb + nfar * 6, cx->org, cx->dir, &tf0, &tf1)) {
      if (tf0 > rn.t) { cx->frames -= both; *out = rn; return 1; }
    } else {
      cx->frames -= both;
      *out = rn;
      return 1;
    }
  }
  const int has_f = fir(cx, nfar, &rf);
  cx->frames -= both;
  if (has_n && has_f) { *out = (rn.t < rf.t) ? rn : rf; return 1; }
  if (has_n) { *out = rn; return 1; }
  if (has_f) { *out = rf; return 1; }
  return 0;
}

int orc_first_intersection_ray(const float org[3], const float dir[3],
                               const uint32_t *tri2vtx, const float *vtx2xyz,
                               const uint32_t *bvhnodes, size_t num_node,
                               const float *aabb, float *t, uint32_t *i_tri,
                               float bc[3], orc_trace_stats *stats) {
  if (num_node == 0) return 0;
  cast_ctx cx;
  memset(&cx, 0, sizeof(cx));
  cx.org = org;
  cx.dir = dir;
  cx.tri2vtx = tri2vtx;
  cx.xyz = vtx2xyz;
  cx.bn = bvhnodes;
  cx.aabb = aabb;
  hit_rec h;
  const int hit = fir(&cx, 0, &h);
  if (stats) {
    stats->nodes_visited += cx.st.nodes_visited;
    stats->tris_tested += cx.st.tris_tested;
    stats->frames_hist[cx.max_frames < 63 ? cx.max_frames : 63] += 1;
  }
  if (!hit) return 0;
  *t = h.t;
  *i_tri = h.tri;
  if (bc) memcpy(bc, h.bc, sizeof(h.bc));
  return 1;
}

/* The same search as fir() without per-subtree results: one best hit, a
 * stack of far children with their entry t, a far child skipped when its
 * entry t exceeds the best t so far, and a hit replacing the best when its
 * t <= the best t.  Children are visited near first, so of two hits the later
 * one lies in the far subtree of their lowest common ancestor, where fir()'s
 * combine keeps far on equal t; and a skipped subtree cannot hold a hit with
 * t <= the final one.  Returns -1 if the stack (cap frames) overflows. */
static int fir_flat(cast_ctx *cx, hit_rec *out, size_t *stk_node, float *stk_t, int cap) {
  float t0, t1;
  if (!orc_ray_aabb(cx->aabb, cx->org, cx->dir, &t0, &t1)) return 0;
  const uint32_t *bn = cx->bn;
  int has = 0, sp = 0;
  size_t node = 0;
  for (;;) {
    if (bn[node * 3 + 2] == U32MAX) {
      const uint32_t tri = bn[node * 3 + 1];
      const uint32_t *v = cx->tri2vtx + (size_t)tri * 3;
      float t, bc[3];
      cx->st.tris_tested++;
      if (orc_ray_tri(cx->xyz + (size_t)v[0] * 3, cx->xyz + (size_t)v[1] * 3,
                      cx->xyz + (size_t)v[2] * 3, cx->org, cx->dir, &t, bc) &&
          t < INFINITY && (!has || t <= out->t)) {
        has = 1;
        out->t = t;
        out->tri = tri;
        memcpy(out->bc, bc, sizeof(bc));
      }
    } else {
      cx->st.nodes_visited++;
      const size_t left = bn[node * 3 + 1], right = bn[node * 3 + 2];
      float tl0, tl1, tr0, tr1;
      const int hl = orc_ray_aabb(cx->aabb + left * 6, cx->org, cx->dir, &tl0, &tl1);
      const int hr = orc_ray_aabb(cx->aabb + right * 6, cx->org, cx->dir, &tr0, &tr1);
      if (hl && hr) {
        const int nl = tuple_lt(tl0, tl1, tr0, tr1);
        if (sp == cap) return -1;
        stk_node[sp] = nl ? right : left;
        stk_t[sp] = nl ? tr0 : tl0;
        ++sp;
        if (sp > cx->max_frames) cx->max_frames = sp;
        node = nl ? left : right;
        continue;
      }
      if (hl || hr) {
        node = hl ? left : right;
        continue;
      }
    }
    for (;;) {  /* pop the next far child that may still hold a hit */
      if (sp == 0) return has;
      --sp;
      if (!(has && stk_t[sp] > out->t)) break;
    }
    node = stk_node[sp];
  }
}

int orc_pix2tri_flat(uint32_t *pix2tri, float *pix2t, size_t num_tri, const uint32_t *tri2vtx,
                     const float *vtx2xyz, const uint32_t *bvhnodes, const float *aabb,
                     uint32_t img_w, uint32_t img_h, const float T[16], int num_threads,
                     orc_trace_stats *stats) {
  enum { CAP = 4096 };
  const int64_t num_pix = (int64_t)img_w * (int64_t)img_h;
  uint64_t nv = 0, nt = 0, hist[64];
  memset(hist, 0, sizeof(hist));
  int rc = 0;
#ifdef _OPENMP
  const int nth = num_threads > 0 ? num_threads : 1;
#pragma omp parallel num_threads(nth) reduction(+ : nv, nt, hist[:64]) reduction(min : rc)
#endif
  {
    size_t *stk_node = (size_t *)malloc(sizeof(size_t) * CAP);
    float *stk_t = (float *)malloc(sizeof(float) * CAP);
    if (!stk_node || !stk_t) rc = -1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
    for (int64_t ip = 0; ip < num_pix; ++ip) {
      const uint32_t ih = (uint32_t)(ip / img_w);
      const uint32_t iw = (uint32_t)(ip - (int64_t)ih * img_w);
      cast_ctx cx;
      memset(&cx, 0, sizeof(cx));
      float org[3], dir[3];
      orc_ray_for_pixel(iw, ih, img_w, img_h, T, org, dir);
      cx.org = org;
      cx.dir = dir;
      cx.tri2vtx = tri2vtx;
      cx.xyz = vtx2xyz;
      cx.bn = bvhnodes;
      cx.aabb = aabb;
      hit_rec h;
      int hit = 0;
      if (num_tri && stk_node && stk_t) hit = fir_flat(&cx, &h, stk_node, stk_t, CAP);
      if (hit < 0) rc = -1;
      pix2tri[ip] = hit > 0 ? h.tri : U32MAX;
      if (pix2t) pix2t[ip] = hit > 0 ? h.t : 0.f;
      nv += cx.st.nodes_visited;
      nt += cx.st.tris_tested;
      hist[cx.max_frames < 63 ? cx.max_frames : 63] += 1;
    }
    free(stk_node);
    free(stk_t);
  }
  (void)num_threads;
  if (stats) {
    stats->nodes_visited += nv;
    stats->tris_tested += nt;
    for (int k = 0; k < 64; ++k) stats->frames_hist[k] += hist[k];
  }
  return rc;
}

/* pix2tri.rs:3-48 (rayon par_iter over pixels -> OpenMP parallel for) */
int orc_pix2tri(uint32_t *pix2tri, float *pix2t, size_t num_tri,
                const uint32_t *tri2vtx, const float *vtx2xyz,
                const uint32_t *bvhnodes, const float *aabb, uint32_t img_w,
                uint32_t img_h, const float T[16], int num_threads,
                orc_trace_stats *stats) {
  const size_t num_node = num_tri ? 2 * num_tri - 1 : 0;
  const int64_t num_pix = (int64_t)img_w * (int64_t)img_h;
  uint64_t nv = 0, nt = 0, hist[64];
  memset(hist, 0, sizeof(hist));
#ifdef _OPENMP
  const int nth = num_threads > 0 ? num_threads : 1;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nth) reduction(+ : nv, nt, hist[:64])
#endif
  for (int64_t ip = 0; ip < num_pix; ++ip) {
    const uint32_t ih = (uint32_t)(ip / img_w);
    const uint32_t iw = (uint32_t)(ip - (int64_t)ih * img_w);
    float org[3], dir[3], t = 0.f;
    uint32_t tri = U32MAX;
    orc_ray_for_pixel(iw, ih, img_w, img_h, T, org, dir);
    orc_trace_stats st;
    memset(&st, 0, sizeof(st));
    if (!orc_first_intersection_ray(org, dir, tri2vtx, vtx2xyz, bvhnodes, num_node,
                                    aabb, &t, &tri, NULL, &st)) {
      tri = U32MAX;
      t = 0.f;
    }
    nv += st.nodes_visited;
    nt += st.tris_tested;
    for (int k = 0; k < 64; ++k) hist[k] += st.frames_hist[k];
    pix2tri[ip] = tri;
    if (pix2t) pix2t[ip] = t;
  }
  (void)num_threads;
  if (stats) {
    stats->nodes_visited += nv;
    stats->tris_tested += nt;
    for (int k = 0; k < 64; ++k) stats->frames_hist[k] += hist[k];
  }
  return 0;
}

/* pix2depth.rs:107-146: world2ndc = inverse(ndc2world); recompute the hit of
 * the stored triangle with the line test; depth = (ndc_z + 1) / 2, 0 on miss. */
int orc_pix2depth(float *pix2depth, const uint32_t *pix2tri,
                  const uint32_t *tri2vtx, const float *vtx2xyz, uint32_t img_w,
                  uint32_t img_h, const float T[16]) {
  float w2n[16];
  if (!orc_mat4_inverse(T, w2n)) return -1;
  const size_t num_pix = (size_t)img_w * img_h;
  for (size_t ip = 0; ip < num_pix; ++ip) {
    const uint32_t tri = pix2tri[ip];
    if (tri == U32MAX) { pix2depth[ip] = 0.f; continue; }
    const uint32_t iw = (uint32_t)(ip % img_w), ih = (uint32_t)(ip / img_w);
    float org[3], dir[3], t, pos[3], ndc[3];
    orc_ray_for_pixel(iw, ih, img_w, img_h, T, org, dir);
    const uint32_t *v = tri2vtx + (size_t)tri * 3;
    if (!orc_line_tri(vtx2xyz + (size_t)v[0] * 3, vtx2xyz + (size_t)v[1] * 3,
                      vtx2xyz + (size_t)v[2] * 3, org, dir, &t, NULL)) {
      pix2depth[ip] = 0.f; /* the reference unwraps here */
      continue;
    }
    for (int i = 0; i < 3; ++i) pos[i] = t * dir[i] + org[i]; /* vec3::axpy */
    if (!orc_transform_homogeneous(w2n, pos, ndc)) { pix2depth[ip] = 0.f; continue; }
    pix2depth[ip] = (ndc[2] + 1.f) * 0.5f;
  }
  return 0;
}

/* mat4_col_major::jacobian_transform: d transform_homogeneous(m, p) / d p,
 * J[r][j] = (m[j*4+r] w - x_r m[j*4+3]) / w^2 with x_r the homogeneous
 * numerator of row r (quotient rule). */
static void jacobian_transform3(const float m[16], const float *p, float J[3][3]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  const float inv_w2 = 1.f / (w * w);
  for (int r = 0; r < 3; ++r) {
    const float q = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r];
    for (int j = 0; j < 3; ++j) J[r][j] = (m[j * 4 + r] * w - q * m[j * 4 + 3]) * inv_w2;
  }
}

/* pix2depth.rs:27-56 (Depth::bwd) as driven by trimesh3_raycast.rs:25-71
 * (bwd_continuous).  Per covered pixel: the line hit (t, bc) of the stored
 * triangle, q = bc0 p0 + bc1 p1 + bc2 p2, dL/dq = J(world2ndc, q)^T (0,0,g/2),
 * dL/dt = dir . dL/dq, and then
 * F6 tri3::intersection_against_line_bwd_wrt_tri with dL/du = dL/dv = 0:
 *    dL/dp_k = (dL/dt / (n . dir)) bc_k n,  n = (p1 - p0) x (p2 - p0)
 *    (moving p_k by d moves the plane at the hit point by bc_k d, and only
 *    its normal component shifts the hit along the line).
 * Accumulates into dldw_vtx2xyz; returns -1 when ndc2world is singular. */
int orc_pix2depth_bwd(float *dldw_vtx2xyz, const uint32_t *pix2tri, const uint32_t *tri2vtx,
                      const float *vtx2xyz, const float *dldw_pix2depth, uint32_t img_w,
                      uint32_t img_h, const float T[16]) {
  float w2n[16];
  if (!orc_mat4_inverse(T, w2n)) return -1;
  const size_t num_pix = (size_t)img_w * img_h;
  for (size_t ip = 0; ip < num_pix; ++ip) {
    const uint32_t tri = pix2tri[ip];
    if (tri == U32MAX) continue;
    const uint32_t iw = (uint32_t)(ip % img_w), ih = (uint32_t)(ip / img_w);
    float org[3], dir[3], t, bc[3];
    orc_ray_for_pixel(iw, ih, img_w, img_h, T, org, dir);
    const uint32_t *v = tri2vtx + (size_t)tri * 3;
    const float *p0 = vtx2xyz + (size_t)v[0] * 3;
    const float *p1 = vtx2xyz + (size_t)v[1] * 3;
    const float *p2 = vtx2xyz + (size_t)v[2] * 3;
    if (!orc_line_tri(p0, p1, p2, org, dir, &t, bc)) continue; /* the reference unwraps */
    float q[3], J[3][3], dldq[3], e1[3], e2[3], n[3];
    for (int i = 0; i < 3; ++i) q[i] = (bc[0] * p0[i] + bc[1] * p1[i]) + bc[2] * p2[i];
    jacobian_transform3(w2n, q, J);
    const float half_g = 0.5f * dldw_pix2depth[ip];
    for (int j = 0; j < 3; ++j) dldq[j] = (J[0][j] * 0.f + J[1][j] * 0.f) + J[2][j] * half_g;
    const float dldt = dot3(dir, dldq);
    for (int i = 0; i < 3; ++i) {
      e1[i] = p1[i] - p0[i];
      e2[i] = p2[i] - p0[i];
    }
    cross3(e1, e2, n);
    const float s = dldt / dot3(n, dir);
    for (int k = 0; k < 3; ++k) {
      const float sk = s * bc[k];
      for (int i = 0; i < 3; ++i) dldw_vtx2xyz[(size_t)v[k] * 3 + i] += sk * n[i];
    }
  }
  return 0;
}

/* ======================================================================= */
/* edge-gradient backward (rasterized_edge_gradient.rs)                     */
/* ======================================================================= */

/* tri2::area: signed area, 0.5 * ((p1-p0) x (p2-p0)) */
static inline float tri2_area(const float p0[2], const float p1[2], const float p2[2]) {
  return 0.5f * ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]));
}

/* transform_homogeneous(...).xy(); the reference unwraps (w != 0 assumed) */
static inline void project_xy(const float m[16], const float *p, float q[2]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  q[0] = (((m[0] * x + m[4] * y) + m[8] * z) + m[12]) / w;
  q[1] = (((m[1] * x + m[5] * y) + m[9] * z) + m[13]) / w;
}

/* rasterized_edge_gradient.rs:1-35 fn_barycentric with tri2::barycentric_coords
 * (signed-area ratios; degenerate -> none). Returns 1 with b, 0 for none. */
static int fn_barycentric(const uint32_t *tri2vtx, const float *xyz,
                          const float pc[2], uint32_t itri, const float m[16],
                          float b[3]) {
  if (itri == U32MAX) return 0;
  const uint32_t *v = tri2vtx + (size_t)itri * 3;
  float p0[2], p1[2], p2[2];
  project_xy(m, xyz + (size_t)v[0] * 3, p0);
  project_xy(m, xyz + (size_t)v[1] * 3, p1);
  project_xy(m, xyz + (size_t)v[2] * 3, p2);
  const float area = tri2_area(p0, p1, p2);
  if (area == 0.f) return 0;
  const float inv_area = 1.f / area;
  b[0] = tri2_area(pc, p1, p2) * inv_area;
  b[1] = tri2_area(p0, pc, p2) * inv_area;
  b[2] = 1.f - b[0] - b[1];
  return 1;
}

/* rasterized_edge_gradient.rs:37-48 fn_inside */
static int fn_inside(int has_b, const float b[3]) {
  if (!has_b) return 1;
  return (b[0] >= 0.f && b[1] >= 0.f && b[2] >= 0.f) ||
         (b[0] <= 0.f && b[1] <= 0.f && b[2] <= 0.f);
}

/* rasterized_edge_gradient.rs:52-74 fn_projection_gradient */
static int fn_projection_gradient(const float m[16], const float *xyz, int axis,
                                  float d[3]) {
  const float q = m[axis] * xyz[0] + m[axis + 4] * xyz[1] + m[axis + 8] * xyz[2] + m[axis + 12];
  const float w = m[3] * xyz[0] + m[7] * xyz[1] + m[11] * xyz[2] + m[15];
  if (fabsf(w) <= FLT_EPSILON) return 0;
  const float inv_w2 = 1.f / (w * w);
  d[0] = (m[axis] * w - q * m[3]) * inv_w2;
  d[1] = (m[axis + 4] * w - q * m[7]) * inv_w2;
  d[2] = (m[axis + 8] * w - q * m[11]) * inv_w2;
  return 1;
}

static void edge_pair(const uint32_t *tri2vtx, const float *xyz, float *dldw,
                      const float m[16], const uint32_t *pix2tri, uint32_t nv,
                      const float *pix2val, const float *dpix, size_t ipix0,
                      size_t ipix1, const float pc0[2], const float pc1[2],
                      int axis) {
  const uint32_t itri0 = pix2tri[ipix0], itri1 = pix2tri[ipix1];
  if (itri0 == itri1) return;
  float b[3];
  const int in0 = fn_inside(fn_barycentric(tri2vtx, xyz, pc0, itri1, m, b), b);
  const int in1 = fn_inside(fn_barycentric(tri2vtx, xyz, pc1, itri0, m, b), b);
  if (!in0 && !in1) return;
  float dldpa = 0.f;
  for (uint32_t k = 0; k < nv; ++k) {
    const float val0 = pix2val[ipix0 * nv + k], val1 = pix2val[ipix1 * nv + k];
    const float dv0 = dpix[ipix0 * nv + k], dv1 = dpix[ipix1 * nv + k];
    dldpa += (dv0 + dv1) * 0.5f * (val0 - val1);
  }
  if (in0 && in1) return; /* intersection: `todo` in the reference */
  uint32_t itri;
  const float *pc;
  if (in1) { itri = itri1; pc = pc1; } /* only tri1 receives the gradient */
  else { itri = itri0; pc = pc0; }     /* only tri0 receives the gradient */
  if (!fn_barycentric(tri2vtx, xyz, pc, itri, m, b)) return; /* unwrap in ref */
  for (int inode = 0; inode < 3; ++inode) {
    const size_t iv = tri2vtx[(size_t)itri * 3 + inode];
    float d[3];
    if (!fn_projection_gradient(m, xyz + iv * 3, axis, d)) continue;
    dldw[iv * 3 + 0] += b[inode] * d[0] * dldpa;
    dldw[iv * 3 + 1] += b[inode] * d[1] * dldpa;
    dldw[iv * 3 + 2] += b[inode] * d[2] * dldpa;
  }
}

/* rasterized_edge_gradient.rs:434-603 bwd: vertically adjacent pairs
 * (pixel-y Jacobian) with iw outer / ih0 inner, then horizontally adjacent
 * pairs (pixel-x Jacobian) with iw0 outer / ih inner. */
void orc_edge_gradient_bwd(const uint32_t *tri2vtx, const float *xyz,
                           float *dldw, const float m[16], uint32_t img_w,
                           uint32_t img_h, const uint32_t *pix2tri, uint32_t nv,
                           const float *pix2val, const float *dpix) {
  for (uint32_t iw = 0; iw < img_w; ++iw) {
    for (uint32_t ih0 = 0; ih0 + 1 < img_h; ++ih0) {
      const uint32_t ih1 = ih0 + 1;
      const float pc0[2] = {(float)iw + 0.5f, (float)ih0 + 0.5f};
      const float pc1[2] = {(float)iw + 0.5f, (float)ih1 + 0.5f};
      edge_pair(tri2vtx, xyz, dldw, m, pix2tri, nv, pix2val, dpix,
                (size_t)ih0 * img_w + iw, (size_t)ih1 * img_w + iw, pc0, pc1, 1);
    }
  }
  for (uint32_t iw0 = 0; iw0 + 1 < img_w; ++iw0) {
    for (uint32_t ih = 0; ih < img_h; ++ih) {
      const uint32_t iw1 = iw0 + 1;
      const float pc0[2] = {(float)iw0 + 0.5f, (float)ih + 0.5f};
      const float pc1[2] = {(float)iw1 + 0.5f, (float)ih + 0.5f};
      edge_pair(tri2vtx, xyz, dldw, m, pix2tri, nv, pix2val, dpix,
                (size_t)ih * img_w + iw0, (size_t)ih * img_w + iw1, pc0, pc1, 0);
    }
  }
}

/* ======================================================================= */
/* contour and anti-aliasing                                                */
/* ======================================================================= */

/* tri3::unit_normal: cross(p1 - p0, p2 - p0) scaled by 1 / |.| */
static void tri_unit_normal(const float *p0, const float *p1, const float *p2, float n[3]) {
  float e1[3], e2[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
  }
  cross3(e1, e2, n);
  const float len = sqrtf(dot3(n, n));
  const float inv = 1.f / len;
  for (int i = 0; i < 3; ++i) n[i] = n[i] * inv;
}

/* mat4_col_major::ray_from_transform_world2ndc: the view ray through a world
 * point -- project it to NDC, then take the NDC segment z = +1 -> -1 at that
 * (x, y) back to world space. */
static void ray_through_point(const float w2n[16], const float n2w[16], const float pos[3],
                              float dir[3]) {
  float ndc[3], q0[3], q1[3];
  if (!orc_transform_homogeneous(w2n, pos, ndc)) ndc[0] = ndc[1] = ndc[2] = NAN;
  const float a[3] = {ndc[0], ndc[1], 1.f}, b[3] = {ndc[0], ndc[1], -1.f};
  if (!orc_transform_homogeneous(n2w, a, q0)) q0[0] = q0[1] = q0[2] = NAN;
  if (!orc_transform_homogeneous(n2w, b, q1)) q1[0] = q1[1] = q1[2] = NAN;
  for (int i = 0; i < 3; ++i) dir[i] = q1[i] - q0[i];
}

/* edge2vtx.rs:191-242 */
long orc_contour_for_triangle_mesh(const uint32_t *tri2vtx, const float *xyz,
                                   const float w2n[16], size_t num_edge,
                                   const uint32_t *edge2vtx, const uint32_t *edge2tri,
                                   uint32_t *cedge2vtx) {
  float n2w[16];
  if (!orc_mat4_inverse(w2n, n2w)) return -1;
  long cnt = 0;
  for (size_t e = 0; e < num_edge; ++e) {
    const uint32_t i0 = edge2vtx[e * 2], i1 = edge2vtx[e * 2 + 1];
    float mid[3], dir[3], n0[3], n1[3];
    for (int i = 0; i < 3; ++i) mid[i] = (xyz[(size_t)i0 * 3 + i] + xyz[(size_t)i1 * 3 + i]) * 0.5f;
    ray_through_point(w2n, n2w, mid, dir);
    const uint32_t *ta = tri2vtx + (size_t)edge2tri[e * 2] * 3;
    const uint32_t *tb = tri2vtx + (size_t)edge2tri[e * 2 + 1] * 3;
    tri_unit_normal(xyz + (size_t)ta[0] * 3, xyz + (size_t)ta[1] * 3, xyz + (size_t)ta[2] * 3, n0);
    tri_unit_normal(xyz + (size_t)tb[0] * 3, xyz + (size_t)tb[1] * 3, xyz + (size_t)tb[2] * 3, n1);
    const int f0 = dot3(n0, dir) > 0.f;
    const int f1 = dot3(n1, dir) > 0.f;
    if (f0 == f1) continue;
    cedge2vtx[cnt * 2] = i0;
    cedge2vtx[cnt * 2 + 1] = i1;
    ++cnt;
  }
  return cnt;
}

static inline float cross2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

/* parametric segment-segment intersection: rc = (a x s) / (r x s),
 * re = (a x r) / (r x s) with r = c1 - c0, s = e1 - s0, a = s0 - c0 */
int orc_intersection_edge2(const float c0[2], const float c1[2], const float s0[2],
                           const float e1[2], float *rc, float *re) {
  const float rx = c1[0] - c0[0], ry = c1[1] - c0[1];
  const float sx = e1[0] - s0[0], sy = e1[1] - s0[1];
  const float d = cross2(rx, ry, sx, sy);
  if (d == 0.f) return 0;
  const float ax = s0[0] - c0[0], ay = s0[1] - c0[1];
  const float r0 = cross2(ax, ay, sx, sy) / d;
  const float r1 = cross2(ax, ay, rx, ry) / d;
  if (r0 < 0.f || r0 > 1.f || r1 < 0.f || r1 > 1.f) return 0;
  *rc = r0;
  *re = r1;
  return 1;
}

/* d rc / d s0 = (dN/ds0 - rc dD/ds0) / D, d rc / d e1 likewise, with
 * N = a x s, D = r x s (derived from the formula above) */
void orc_dldw_intersection_edge2(const float c0[2], const float c1[2], const float s0[2],
                                 const float e1[2], float dldrc, float dlds0[2],
                                 float dlde1[2]) {
  const float rx = c1[0] - c0[0], ry = c1[1] - c0[1];
  const float sx = e1[0] - s0[0], sy = e1[1] - s0[1];
  const float ax = s0[0] - c0[0], ay = s0[1] - c0[1];
  const float d = cross2(rx, ry, sx, sy);
  const float rc = cross2(ax, ay, sx, sy) / d;
  const float g = dldrc / d;
  dlds0[0] = ((sy + ay) - rc * ry) * g;
  dlds0[1] = ((-sx - ax) + rc * rx) * g;
  dlde1[0] = (-ay + rc * ry) * g;
  dlde1[1] = (ax - rc * rx) * g;
}

/* one (pix0, c0, pix1, c1) neighbour pair, differentiable_rasterizer.cu:15-64 */
typedef struct {
  uint32_t p0, p1;
  float c0[2], c1[2];
} aa_pair;

static void aa_pairs(uint32_t ipix, uint32_t img_w, int horizontal, aa_pair out[4]) {
  const uint32_t iw = ipix % img_w, ih = ipix / img_w;
  const uint32_t pix[5] = {ih * img_w + iw, ih * img_w + iw - 1u, ih * img_w + iw + 1u,
                           (ih - 1u) * img_w + iw, (ih + 1u) * img_w + iw};
  const float pos[5][2] = {{(float)iw + 0.5f, (float)ih + 0.5f},
                           {(float)iw - 0.5f, (float)ih + 0.5f},
                           {(float)iw + 1.5f, (float)ih + 0.5f},
                           {(float)iw + 0.5f, (float)ih - 0.5f},
                           {(float)iw + 0.5f, (float)ih + 1.5f}};
  static const int hz[4][2] = {{0, 1}, {1, 0}, {0, 2}, {2, 0}};
  static const int vt[4][2] = {{0, 3}, {3, 0}, {0, 4}, {4, 0}};
  for (int k = 0; k < 4; ++k) {
    const int a = horizontal ? hz[k][0] : vt[k][0], b = horizontal ? hz[k][1] : vt[k][1];
    out[k].p0 = pix[a];
    out[k].p1 = pix[b];
    out[k].c0[0] = pos[a][0];
    out[k].c0[1] = pos[a][1];
    out[k].c1[0] = pos[b][0];
    out[k].c1[1] = pos[b][1];
  }
}

/* Walks the pixels of the projected edge q0 -> q1 as dda.h:6-63 does (one
 * step past the end so the line stays connected) and calls fn on every
 * foreground/background neighbour pair crossed by the edge.
 * Two exits that change no result: the walk stops once it has left the image
 * in its direction of travel (every later pixel is outside too), and edges
 * whose projection is not finite are skipped (their walk would not end).
 * A walk longer than ORC_WALK_LIMIT steps is abandoned; returns 0 then. */
#define ORC_WALK_LIMIT (1u << 26)
typedef void (*aa_visit_fn)(void *ctx, const aa_pair *pr, float rc);

static int aa_walk(const float q0[2], const float q1[2], uint32_t img_w, uint32_t img_h,
                   const uint32_t *pix2tri, aa_visit_fn fn, void *ctx) {
  if (!(isfinite(q0[0]) && isfinite(q0[1]) && isfinite(q1[0]) && isfinite(q1[1]))) return 1;
  const float v01x = q1[0] - q0[0], v01y = q1[1] - q0[1];
  if (v01x == 0.f && v01y == 0.f) return 1;
  const int horizontal = fabsf(v01x) < fabsf(v01y);
  float x = q0[0], y = q0[1];
  const float dx = q1[0] - q0[0], dy = q1[1] - q0[1];
  const float step = fmax_sel(fabsf(dx), fabsf(dy));
  const float slope_x = dx / step, slope_y = dy / step;
  const uint32_t total = img_w * img_h;
  const float fw = (float)img_w, fh = (float)img_h;
  for (uint32_t k = 0; k < ORC_WALK_LIMIT; ++k) {
    const float fx = floorf(x), fy = floorf(y);
    if (fx >= 0.f && fy >= 0.f && fx < fw && fy < fh) {
      aa_pair prs[4];
      aa_pairs((uint32_t)fy * img_w + (uint32_t)fx, img_w, horizontal, prs);
      for (int k = 0; k < 4; ++k) {
        const aa_pair *pr = &prs[k];
        if (pr->p0 >= total || pr->p1 >= total) continue;
        if (pix2tri[pr->p0] == U32MAX || pix2tri[pr->p1] != U32MAX) continue;
        float rc, re;
        if (!orc_intersection_edge2(pr->c0, pr->c1, q1, q0, &rc, &re)) continue;
        fn(ctx, pr, rc);
      }
    }
    const float eps = 1.0e-3f;
    if (fabsf(x - q0[0]) > fabsf(q1[0] - q0[0]) + eps) return 1;
    if (fabsf(y - q0[1]) > fabsf(q1[1] - q0[1]) + eps) return 1;
    if ((slope_x > 0.f && x >= fw) || (slope_x < 0.f && x < 0.f) || (slope_y > 0.f && y >= fh) ||
        (slope_y < 0.f && y < 0.f))
      return 1;
    x = x + slope_x;
    y = y + slope_y;
  }
  return 0;
}

typedef struct {
  const float *vin;
  float *vout;
} aa_fwd_ctx;

static void aa_fwd_visit(void *vctx, const aa_pair *pr, float rc) {
  aa_fwd_ctx *c = (aa_fwd_ctx *)vctx;
  if (rc < 0.5f) {
    const float r0 = 0.5f + rc;
    c->vout[pr->p0] = r0 * c->vin[pr->p0] + (1.f - r0) * c->vin[pr->p1];
  } else {
    const float r0 = rc - 0.5f;
    c->vout[pr->p1] = r0 * c->vin[pr->p0] + (1.f - r0) * c->vin[pr->p1];
  }
}

/* antialias.rs:54-99 */
long orc_antialias(size_t num_cedge, const uint32_t *cedge2vtx, const float *xyz,
                   const float m[16], uint32_t img_w, uint32_t img_h, const uint32_t *pix2tri,
                   const float *pix2vin, float *pix2vout) {
  aa_fwd_ctx c = {pix2vin, pix2vout};
  long abandoned = 0;
  for (size_t e = 0; e < num_cedge; ++e) {
    float q0[2], q1[2];
    project_xy(m, xyz + (size_t)cedge2vtx[e * 2] * 3, q0);
    project_xy(m, xyz + (size_t)cedge2vtx[e * 2 + 1] * 3, q1);
    abandoned += !aa_walk(q0, q1, img_w, img_h, pix2tri, aa_fwd_visit, &c);
  }
  return abandoned;
}

/* 2x3 xy-rows of the Jacobian of transform_homogeneous at p
 * (mat4_col_major::jacobian_transform + to_mat2x3_col_major_xy):
 * J[r][j] = (m[j*4+r] w - q_r m[j*4+3]) / w^2 */
static void proj_jacobian_xy(const float m[16], const float *p, float J[2][3]) {
  const float x = p[0], y = p[1], z = p[2];
  const float w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15];
  const float inv_w2 = 1.f / (w * w);
  for (int r = 0; r < 2; ++r) {
    const float q = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r];
    for (int j = 0; j < 3; ++j) J[r][j] = (m[j * 4 + r] * w - q * m[j * 4 + 3]) * inv_w2;
  }
}

typedef struct {
  const float *val, *dldw;
  const float *q0, *q1;
  float J0[2][3], J1[2][3];
  float *g0, *g1; /* dL/dp0, dL/dp1 accumulators (the vertex rows) */
} aa_bwd_ctx;

static void aa_bwd_visit(void *vctx, const aa_pair *pr, float rc) {
  aa_bwd_ctx *c = (aa_bwd_ctx *)vctx;
  const float dv = c->val[pr->p0] - c->val[pr->p1];
  const float dldr0 = rc < 0.5f ? c->dldw[pr->p0] * dv : c->dldw[pr->p1] * dv;
  float dldq1[2], dldq0[2];
  orc_dldw_intersection_edge2(pr->c0, pr->c1, c->q1, c->q0, dldr0, dldq1, dldq0);
  for (int j = 0; j < 3; ++j) {
    c->g0[j] += c->J0[0][j] * dldq0[0] + c->J0[1][j] * dldq0[1];
    c->g1[j] += c->J1[0][j] * dldq1[0] + c->J1[1][j] * dldq1[1];
  }
}

/* antialias.rs:102-159 */
long orc_bwd_antialias(size_t num_cedge, const uint32_t *cedge2vtx, const float *xyz,
                       float *dldw_vtx2xyz, const float m[16], uint32_t img_w, uint32_t img_h,
                       const float *pix2val, const float *dldw_pix2val, const uint32_t *pix2tri) {
  long abandoned = 0;
  for (size_t e = 0; e < num_cedge; ++e) {
    const uint32_t i0 = cedge2vtx[e * 2], i1 = cedge2vtx[e * 2 + 1];
    float q0[2], q1[2];
    project_xy(m, xyz + (size_t)i0 * 3, q0);
    project_xy(m, xyz + (size_t)i1 * 3, q1);
    aa_bwd_ctx c;
    c.val = pix2val;
    c.dldw = dldw_pix2val;
    c.q0 = q0;
    c.q1 = q1;
    proj_jacobian_xy(m, xyz + (size_t)i0 * 3, c.J0);
    proj_jacobian_xy(m, xyz + (size_t)i1 * 3, c.J1);
    c.g0 = dldw_vtx2xyz + (size_t)i0 * 3;
    c.g1 = dldw_vtx2xyz + (size_t)i1 * 3;
    abandoned += !aa_walk(q0, q1, img_w, img_h, pix2tri, aa_bwd_visit, &c);
  }
  return abandoned;
}

/* ========================================================================
 * Mesh topology (vertex-to-element, vertex-to-vertex, edges)
 * ===================================================================== */

/* vtx2elem.rs:58-96 */
int orc_vtx2elem(size_t num_elem, size_t num_node, const uint32_t *elem2vtx, size_t num_vtx,
                 uint32_t *vtx2idx, uint32_t *idx2elem) {
  for (size_t v = 0; v <= num_vtx; ++v) vtx2idx[v] = 0;
  for (size_t e = 0; e < num_elem; ++e)
    for (size_t k = 0; k < num_node; ++k) {
      const uint32_t v = elem2vtx[e * num_node + k];
      if (v >= num_vtx) return -1;
      vtx2idx[v + 1] += 1;
    }
  for (size_t v = 0; v < num_vtx; ++v) vtx2idx[v + 1] += vtx2idx[v];
  for (size_t e = 0; e < num_elem; ++e)
    for (size_t k = 0; k < num_node; ++k) {
      const uint32_t v = elem2vtx[e * num_node + k];
      idx2elem[vtx2idx[v]] = (uint32_t)e;
      vtx2idx[v] += 1;
    }
  for (size_t v = num_vtx; v >= 1; --v) vtx2idx[v] = vtx2idx[v - 1];
  vtx2idx[0] = 0;
  return 0;
}

/* vtx2vtx.rs:11-76 (a per-vertex "last visited by" flag removes repeats) */
size_t orc_vtx2vtx(size_t num_elem, size_t num_node, const uint32_t *elem2vtx, size_t num_vtx,
                   const uint32_t *vtx2jdx, const uint32_t *jdx2elem, int is_self,
                   uint32_t *vtx2idx, uint32_t *idx2vtx) {
  (void)num_elem;
  size_t *flag = (size_t *)malloc(sizeof(size_t) * (num_vtx ? num_vtx : 1));
  for (size_t v = 0; v < num_vtx; ++v) flag[v] = (size_t)-1;
  vtx2idx[0] = 0;
  size_t cnt = 0;
  for (size_t i = 0; i < num_vtx; ++i) {
    if (!is_self) flag[i] = i;
    for (uint32_t jdx = vtx2jdx[i]; jdx < vtx2jdx[i + 1]; ++jdx) {
      const size_t e = jdx2elem[jdx];
      for (size_t k = 0; k < num_node; ++k) {
        const uint32_t j = elem2vtx[e * num_node + k];
        if (flag[j] == i) continue;
        flag[j] = i;
        if (idx2vtx) idx2vtx[cnt] = j;
        ++cnt;
      }
    }
    vtx2idx[i + 1] = (uint32_t)cnt;
  }
  free(flag);
  return cnt;
}

/* edge2vtx.rs:16-36 */
void orc_edge2vtx_from_vtx2vtx(size_t num_vtx, const uint32_t *vtx2idx, const uint32_t *idx2vtx,
                               uint32_t *edge2vtx) {
  for (size_t i = 0; i < num_vtx; ++i)
    for (uint32_t idx = vtx2idx[i]; idx < vtx2idx[i + 1]; ++idx) {
      edge2vtx[(size_t)idx * 2] = (uint32_t)i;
      edge2vtx[(size_t)idx * 2 + 1] = idx2vtx[idx];
    }
}

/* edge2elem.rs:20-70 */
int orc_edge2elem(size_t num_edge, const uint32_t *edge2vtx, const uint32_t *tri2vtx,
                  const uint32_t *vtx2idx, const uint32_t *idx2tri, uint32_t *edge2tri) {
  for (size_t e = 0; e < num_edge; ++e) {
    const uint32_t i0 = edge2vtx[e * 2], i1 = edge2vtx[e * 2 + 1];
    int cnt = 0;
    for (uint32_t idx = vtx2idx[i0]; idx < vtx2idx[i0 + 1]; ++idx) {
      const uint32_t t = idx2tri[idx];
      const uint32_t j0 = tri2vtx[(size_t)t * 3], j1 = tri2vtx[(size_t)t * 3 + 1],
                     j2 = tri2vtx[(size_t)t * 3 + 2];
      int adjacent;
      if (i0 == j0 && i0 != j1 && i0 != j2) adjacent = (j1 == i1) || (j2 == i1);
      else if (i0 != j0 && i0 == j1 && i0 != j2) adjacent = (j2 == i1) || (j0 == i1);
      else if (i0 != j0 && i0 != j1 && i0 == j2) adjacent = (j0 == i1) || (j1 == i1);
      else return -1;
      if (!adjacent) continue;
      edge2tri[e * 2 + (size_t)cnt] = t;
      if (++cnt == 2) break;
    }
  }
  return 0;
}

/* vtx2vtx.rs:226-240: one Jacobi update of vertex i from prev */
static void lap_update(size_t i, const uint32_t *vtx2idx, const uint32_t *idx2vtx, float lambda,
                       size_t num_dim, const float *vtx2rhs, const float *prev, float *next) {
  float buff[64];
  for (size_t d = 0; d < num_dim; ++d) buff[d] = vtx2rhs[i * num_dim + d];
  for (uint32_t idx = vtx2idx[i]; idx < vtx2idx[i + 1]; ++idx) {
    const size_t j = idx2vtx[idx];
    for (size_t d = 0; d < num_dim; ++d) buff[d] += lambda * prev[j * num_dim + d];
  }
  const float valence = (float)(vtx2idx[i + 1] - vtx2idx[i]);
  const float inv_dia = 1.f / (1.f + lambda * valence);
  for (size_t d = 0; d < num_dim; ++d) next[i * num_dim + d] = buff[d] * inv_dia;
}

/* vtx2vtx.rs:210-252 (num_dim <= 64) */
void orc_laplacian_smoothing(size_t num_vtx, const uint32_t *vtx2idx, const uint32_t *idx2vtx,
                             float lambda, size_t num_dim, float *vtx2lhs, const float *vtx2rhs,
                             size_t num_iter, float *vtx2lhstmp) {
  for (size_t it = 0; it < num_iter; ++it) {
    for (size_t i = 0; i < num_vtx; ++i)
      lap_update(i, vtx2idx, idx2vtx, lambda, num_dim, vtx2rhs, vtx2lhs, vtx2lhstmp);
    for (size_t i = 0; i < num_vtx; ++i)
      lap_update(i, vtx2idx, idx2vtx, lambda, num_dim, vtx2rhs, vtx2lhstmp, vtx2lhs);
  }
}

/* vtx2vtx.rs:254-280 */
void orc_multiply_graph_laplacian(size_t num_vtx, const uint32_t *vtx2idx,
                                  const uint32_t *idx2vtx, size_t num_dim, const float *vtx2rhs,
                                  float *vtx2lhs) {
  for (size_t i = 0; i < num_vtx; ++i) {
    const float valence = (float)(vtx2idx[i + 1] - vtx2idx[i]);
    for (size_t d = 0; d < num_dim; ++d) vtx2lhs[i * num_dim + d] = valence * vtx2rhs[i * num_dim + d];
    for (uint32_t idx = vtx2idx[i]; idx < vtx2idx[i + 1]; ++idx) {
      const size_t j = idx2vtx[idx];
      for (size_t d = 0; d < num_dim; ++d) vtx2lhs[i * num_dim + d] -= vtx2rhs[j * num_dim + d];
    }
  }
}

/* ========================================================================
 * Attribute interpolation over pix2tri
 * ===================================================================== */

/* F7: barycentric coordinates of the line (org, dir) against the plane of
 * (p0, p1, p2) -- Moller-Trumbore's u, v with no range or epsilon test
 * (tri3::intersection_plane_of_tri3_against_line as used at
 * pix2tri.rs:176-229).  Also returns inv_det and the two cross products the
 * backward reuses. */
static void plane_line_bc(const float p0[3], const float p1[3], const float p2[3],
                          const float org[3], const float dir[3], float *u_out, float *v_out,
                          float *inv_det_out, float pvec[3], float e1xd[3]) {
  float e1[3], e2[3], tvec[3], qvec[3];
  for (int i = 0; i < 3; ++i) {
    e1[i] = p1[i] - p0[i];
    e2[i] = p2[i] - p0[i];
    tvec[i] = org[i] - p0[i];
  }
  cross3(dir, e2, pvec);
  const float det = dot3(e1, pvec);
  const float inv_det = 1.f / det;
  *u_out = inv_det * dot3(tvec, pvec);
  cross3(tvec, e1, qvec);
  *v_out = inv_det * dot3(dir, qvec);
  *inv_det_out = inv_det;
  cross3(e1, dir, e1xd);
}

/* pix2tri.rs:176-229: pix2val[p] = bc0 val[i0] + bc1 val[i1] + bc2 val[i2]
 * for covered pixels, 0 elsewhere */
int orc_interpolate(float *pix2val, const uint32_t *pix2tri, uint32_t img_w, uint32_t img_h,
                    const uint32_t *tri2vtx, size_t num_tri, const float *vtx2xyz,
                    size_t num_vtx, uint32_t num_vdim, const float *vtx2val, const float T[16]) {
  const size_t np = (size_t)img_w * img_h;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (size_t ip = 0; ip < np; ++ip) {
    float *out = pix2val + ip * num_vdim;
    const uint32_t it = pix2tri[ip];
    if (it == U32MAX) {
      for (uint32_t d = 0; d < num_vdim; ++d) out[d] = 0.f;
      continue;
    }
    if (it >= num_tri) { bad = 1; continue; }
    const uint32_t i0 = tri2vtx[(size_t)it * 3], i1 = tri2vtx[(size_t)it * 3 + 1],
                   i2 = tri2vtx[(size_t)it * 3 + 2];
    if (i0 >= num_vtx || i1 >= num_vtx || i2 >= num_vtx) { bad = 1; continue; }
    float org[3], dir[3], u, v, inv_det, pvec[3], c[3];
    orc_ray_for_pixel((uint32_t)(ip % img_w), (uint32_t)(ip / img_w), img_w, img_h, T, org, dir);
    plane_line_bc(vtx2xyz + (size_t)i0 * 3, vtx2xyz + (size_t)i1 * 3, vtx2xyz + (size_t)i2 * 3,
                  org, dir, &u, &v, &inv_det, pvec, c);
    const float bc0 = 1.f - u - v;
    for (uint32_t d = 0; d < num_vdim; ++d)
      out[d] = (bc0 * vtx2val[(size_t)i0 * num_vdim + d] + u * vtx2val[(size_t)i1 * num_vdim + d]) +
               v * vtx2val[(size_t)i2 * num_vdim + d];
  }
  return bad ? -1 : 0;
}

/* pix2tri.rs:231-304 (serial; accumulates).  F8: with gu = dL/du, gv = dL/dv
 * (dL/dbc1 - dL/dbc0, dL/dbc2 - dL/dbc0) and dL/dt = 0, the gradient of the
 * plane intersection with respect to the corners is dL/dp0 = -bc0 lam,
 * dL/dp1 = -u lam, dL/dp2 = -v lam, lam = (gu (dir x e2) + gv (e1 x dir)) / det
 * (implicit differentiation of u e1 + v e2 - t dir = org - p0). */
int orc_interpolate_bwd(float *dldw_vtx2xyz, float *dldw_vtx2val, const uint32_t *pix2tri,
                        uint32_t img_w, uint32_t img_h, const uint32_t *tri2vtx, size_t num_tri,
                        const float *vtx2xyz, size_t num_vtx, uint32_t num_vdim,
                        const float *vtx2val, const float T[16], const float *dldw_pix2val) {
  const size_t np = (size_t)img_w * img_h;
  for (size_t ip = 0; ip < np; ++ip) {
    const uint32_t it = pix2tri[ip];
    if (it == U32MAX) continue;
    if (it >= num_tri) return -1;
    const uint32_t iv[3] = {tri2vtx[(size_t)it * 3], tri2vtx[(size_t)it * 3 + 1],
                            tri2vtx[(size_t)it * 3 + 2]};
    if (iv[0] >= num_vtx || iv[1] >= num_vtx || iv[2] >= num_vtx) return -1;
    const float *g = dldw_pix2val + ip * num_vdim;
    float gb[3] = {0.f, 0.f, 0.f};
    for (uint32_t d = 0; d < num_vdim; ++d)
      for (int k = 0; k < 3; ++k) gb[k] += vtx2val[(size_t)iv[k] * num_vdim + d] * g[d];
    const float gu = gb[1] - gb[0], gv = gb[2] - gb[0];
    float org[3], dir[3], u, v, inv_det, pvec[3], c[3];
    orc_ray_for_pixel((uint32_t)(ip % img_w), (uint32_t)(ip / img_w), img_w, img_h, T, org, dir);
    plane_line_bc(vtx2xyz + (size_t)iv[0] * 3, vtx2xyz + (size_t)iv[1] * 3,
                  vtx2xyz + (size_t)iv[2] * 3, org, dir, &u, &v, &inv_det, pvec, c);
    const float bc[3] = {1.f - u - v, u, v};
    for (int i = 0; i < 3; ++i) {
      const float lam = (gu * pvec[i] + gv * c[i]) * inv_det;
      for (int k = 0; k < 3; ++k) dldw_vtx2xyz[(size_t)iv[k] * 3 + i] += -(bc[k] * lam);
    }
    for (uint32_t d = 0; d < num_vdim; ++d)
      for (int k = 0; k < 3; ++k) dldw_vtx2val[(size_t)iv[k] * num_vdim + d] += g[d] * bc[k];
  }
  return 0;
}

/* ========================================================================
 * Smoothed edge gradient on the staggered pixel-edge grid
 * ===================================================================== */

static uint8_t edge_type(const uint32_t *tri2vtx, const float *xyz, const float m[16],
                         uint32_t itri0, uint32_t itri1, const float pc0[2], const float pc1[2]) {
  if (itri0 == itri1) return 0;
  float b[3];
  const int in0 = fn_inside(fn_barycentric(tri2vtx, xyz, pc0, itri1, m, b), b);
  const int in1 = fn_inside(fn_barycentric(tri2vtx, xyz, pc1, itri0, m, b), b);
  if (!in0 && !in1) return 1; /* shared edge */
  if (in0 && !in1) return 2;  /* tri0 in front */
  if (!in0 && in1) return 3;  /* tri1 in front */
  return 4;                   /* intersection */
}

static float edge_dldr(uint32_t nv, const float *pix2val, const float *dpix, size_t ip0,
                       size_t ip1) {
  float r = 0.f;
  for (uint32_t k = 0; k < nv; ++k) {
    const float v0 = pix2val[ip0 * nv + k], v1 = pix2val[ip1 * nv + k];
    const float d0 = dpix[ip0 * nv + k], d1 = dpix[ip1 * nv + k];
    r += (d0 + d1) * 0.5f * (v0 - v1);
  }
  return r;
}

/* rasterized_edge_gradient.rs:77-200: hedge (ih0, iw) between pixels
 * (iw, ih0) and (iw, ih0+1), [H-1][W]; vedge (ih, iw0) between (iw0, ih) and
 * (iw0+1, ih), [H][W-1]. */
void orc_edge_gradient_and_type(const uint32_t *tri2vtx, const float *xyz, const float m[16],
                                uint32_t img_w, uint32_t img_h, const uint32_t *pix2tri,
                                uint32_t nv, const float *pix2val, const float *dpix,
                                uint8_t *hedge2type, float *hedge2dldr, uint8_t *vedge2type,
                                float *vedge2dldr) {
  for (uint32_t iw = 0; iw < img_w; ++iw)
    for (uint32_t ih0 = 0; ih0 + 1 < img_h; ++ih0) {
      const size_t ip0 = (size_t)ih0 * img_w + iw, ip1 = ip0 + img_w;
      const float pc0[2] = {(float)iw + 0.5f, (float)ih0 + 0.5f};
      const float pc1[2] = {(float)iw + 0.5f, (float)(ih0 + 1) + 0.5f};
      const size_t e = (size_t)ih0 * img_w + iw;
      hedge2type[e] = edge_type(tri2vtx, xyz, m, pix2tri[ip0], pix2tri[ip1], pc0, pc1);
      hedge2dldr[e] = edge_dldr(nv, pix2val, dpix, ip0, ip1);
    }
  for (uint32_t iw0 = 0; iw0 + 1 < img_w; ++iw0)
    for (uint32_t ih = 0; ih < img_h; ++ih) {
      const size_t ip0 = (size_t)ih * img_w + iw0, ip1 = ip0 + 1;
      const float pc0[2] = {(float)iw0 + 0.5f, (float)ih + 0.5f};
      const float pc1[2] = {(float)(iw0 + 1) + 0.5f, (float)ih + 0.5f};
      const size_t e = (size_t)ih * (img_w - 1) + iw0;
      vedge2type[e] = edge_type(tri2vtx, xyz, m, pix2tri[ip0], pix2tri[ip1], pc0, pc1);
      vedge2dldr[e] = edge_dldr(nv, pix2val, dpix, ip0, ip1);
    }
}

static int fixed_type(uint8_t t) { return t == 2 || t == 3; }

/* one hedge cell update; returns 1 and the new value when it has neighbours */
static int hedge_avg(uint32_t W, uint32_t H, const uint8_t *ht, const float *hv, const uint8_t *vt,
                     uint32_t ih, uint32_t iw, float *out) {
  float s = 0.f;
  uint32_t n = 0;
  if (ih > 0 && ht[(size_t)(ih - 1) * W + iw] != 2) { s += hv[(size_t)(ih - 1) * W + iw]; ++n; }
  if (ih + 1 < H - 1 && ht[(size_t)(ih + 1) * W + iw] != 3) { s += hv[(size_t)(ih + 1) * W + iw]; ++n; }
  if (iw > 0 && !fixed_type(vt[(size_t)ih * (W - 1) + iw - 1]) &&
      !fixed_type(vt[(size_t)(ih + 1) * (W - 1) + iw - 1])) {
    s += hv[(size_t)ih * W + iw - 1];
    ++n;
  }
  if (iw + 1 < W && !fixed_type(vt[(size_t)ih * (W - 1) + iw]) &&
      !fixed_type(vt[(size_t)(ih + 1) * (W - 1) + iw])) {
    s += hv[(size_t)ih * W + iw + 1];
    ++n;
  }
  if (!n) return 0;
  *out = s / (float)n;
  return 1;
}

static int vedge_avg(uint32_t W, uint32_t H, const uint8_t *vt, const float *vv, const uint8_t *ht,
                     uint32_t ih, uint32_t iw, float *out) {
  const uint32_t w = W - 1;
  float s = 0.f;
  uint32_t n = 0;
  if (iw > 0 && vt[(size_t)ih * w + iw - 1] != 2) { s += vv[(size_t)ih * w + iw - 1]; ++n; }
  if (iw + 1 < w && vt[(size_t)ih * w + iw + 1] != 3) { s += vv[(size_t)ih * w + iw + 1]; ++n; }
  if (ih > 0 && !fixed_type(ht[(size_t)(ih - 1) * W + iw]) &&
      !fixed_type(ht[(size_t)(ih - 1) * W + iw + 1])) {
    s += vv[(size_t)(ih - 1) * w + iw];
    ++n;
  }
  if (ih + 1 < H && !fixed_type(ht[(size_t)ih * W + iw]) && !fixed_type(ht[(size_t)ih * W + iw + 1])) {
    s += vv[(size_t)(ih + 1) * w + iw];
    ++n;
  }
  if (!n) return 0;
  *out = s / (float)n;
  return 1;
}

/* rasterized_edge_gradient.rs:250-432: in-place sweeps in index order
 * (Gauss-Seidel), hedges then vedges, cells of type 2 / 3 held fixed */
void orc_smooth_gradient_serial(uint32_t W, uint32_t H, const uint8_t *ht, const uint8_t *vt,
                                uint32_t num_iter, float *hv, float *vv) {
  for (uint32_t it = 0; it < num_iter; ++it)
    for (uint32_t ih = 0; ih + 1 < H; ++ih)
      for (uint32_t iw = 0; iw < W; ++iw) {
        if (fixed_type(ht[(size_t)ih * W + iw])) continue;
        float v;
        if (hedge_avg(W, H, ht, hv, vt, ih, iw, &v)) hv[(size_t)ih * W + iw] = v;
      }
  for (uint32_t it = 0; it < num_iter; ++it)
    for (uint32_t ih = 0; ih < H; ++ih)
      for (uint32_t iw = 0; iw + 1 < W; ++iw) {
        if (fixed_type(vt[(size_t)ih * (W - 1) + iw])) continue;
        float v;
        if (vedge_avg(W, H, vt, vv, ht, ih, iw, &v)) vv[(size_t)ih * (W - 1) + iw] = v;
      }
}

/* the GPU path of the reference (rasterized_edge_gradient.cu:139-275 as
 * launched at del-msh-dlpack/src/rasterized_edge_gradient.rs:336-380): per
 * iteration the cells with (iw + ih) even, then odd, each from the values
 * before that half-step (a cell's neighbours have the other colour) */
void orc_smooth_gradient_red_black(uint32_t W, uint32_t H, const uint8_t *ht, const uint8_t *vt,
                                   uint32_t num_iter, float *hv, float *vv) {
  for (uint32_t it = 0; it < num_iter; ++it)
    for (uint32_t color = 0; color < 2; ++color)
      for (uint32_t ih = 0; ih + 1 < H; ++ih)
        for (uint32_t iw = 0; iw < W; ++iw) {
          if (((iw + ih) & 1u) != color || fixed_type(ht[(size_t)ih * W + iw])) continue;
          float v;
          if (hedge_avg(W, H, ht, hv, vt, ih, iw, &v)) hv[(size_t)ih * W + iw] = v;
        }
  for (uint32_t it = 0; it < num_iter; ++it)
    for (uint32_t color = 0; color < 2; ++color)
      for (uint32_t ih = 0; ih < H; ++ih)
        for (uint32_t iw = 0; iw + 1 < W; ++iw) {
          if (((iw + ih) & 1u) != color || fixed_type(vt[(size_t)ih * (W - 1) + iw])) continue;
          float v;
          if (vedge_avg(W, H, vt, vv, ht, ih, iw, &v)) vv[(size_t)ih * (W - 1) + iw] = v;
        }
}

static inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
/* Rust's `f32 as i32`: NaN -> 0, saturating */
static inline int f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483647.f) return 2147483647;
  if (x <= -2147483648.f) return (-2147483647 - 1);
  return (int)x;
}
static inline float clampf01(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

/* rasterized_edge_gradient.rs:202-248: x-velocity bilinear from vedges at
 * (iw0 + 1, ih + 0.5), y-velocity from hedges at (iw + 0.5, ih0 + 1) */
void orc_interpolate_staggered_grid(uint32_t W, uint32_t H, const float *hedge2vy,
                                    const float *vedge2vx, size_t num_vtx, const float *vtx2xy,
                                    float *vtx2velo) {
  for (size_t i = 0; i < num_vtx; ++i) {
    const float px = vtx2xy[i * 2], py = vtx2xy[i * 2 + 1];
    {
      const float gx = px - 1.0f, gy = py - 0.5f;
      const int ix0 = clampi(f2i_sat(floorf(gx)), 0, (int)W - 2);
      const int iy0 = clampi(f2i_sat(floorf(gy)), 0, (int)H - 2);
      const int ix1 = ix0 + 1 < (int)W - 2 ? ix0 + 1 : (int)W - 2;
      const int iy1 = iy0 + 1 < (int)H - 1 ? iy0 + 1 : (int)H - 1;
      const float tx = clampf01(gx - (float)ix0), ty = clampf01(gy - (float)iy0);
      const int w = (int)W - 1;
      vtx2velo[i * 2] = (((1.f - tx) * (1.f - ty) * vedge2vx[iy0 * w + ix0] +
                          tx * (1.f - ty) * vedge2vx[iy0 * w + ix1]) +
                         (1.f - tx) * ty * vedge2vx[iy1 * w + ix0]) +
                        tx * ty * vedge2vx[iy1 * w + ix1];
    }
    {
      const float gx = px - 0.5f, gy = py - 1.0f;
      const int ix0 = clampi(f2i_sat(floorf(gx)), 0, (int)W - 2);
      const int iy0 = clampi(f2i_sat(floorf(gy)), 0, (int)H - 2);
      const int ix1 = ix0 + 1 < (int)W - 1 ? ix0 + 1 : (int)W - 1;
      const int iy1 = iy0 + 1 < (int)H - 2 ? iy0 + 1 : (int)H - 2;
      const float tx = clampf01(gx - (float)ix0), ty = clampf01(gy - (float)iy0);
      vtx2velo[i * 2 + 1] = (((1.f - tx) * (1.f - ty) * hedge2vy[iy0 * W + ix0] +
                              tx * (1.f - ty) * hedge2vy[iy0 * W + ix1]) +
                             (1.f - tx) * ty * hedge2vy[iy1 * W + ix0]) +
                            tx * ty * hedge2vy[iy1 * W + ix1];
    }
  }
}


/* =========================================================================
 * Quad/oct-tree (del-msh-cpu/src/quad_oct_tree.rs) and sorted-array helpers
 * (array1d.rs:1-15, map_idx.rs:83-115, offset_array aggregate)
 * ========================================================================= */

/* quad_oct_tree.rs:1-49 binary_radix_tree_and_depth and :77-109
 * bnode2onode_and_idx2bnode (idx2bnode must come filled with U32MAX, as the
 * reference's wrapper does) */
int orc_octree_bnodes(size_t num_idx, const uint32_t *idx2morton, uint32_t num_dim,
                      uint32_t max_depth, uint32_t *bnodes, uint32_t *bnode2depth,
                      uint32_t *bnode2onode, uint32_t *idx2bnode) {
  if (num_idx < 2) return -1;
  const size_t num_bnode = num_idx - 1;
  bnodes[0] = U32MAX;
  for (size_t ib = 0; ib < num_bnode; ++ib) {
    size_t r0, r1;
    orc_range_u32(idx2morton, num_idx, ib, &r0, &r1);
    const size_t sp = orc_split_u32(idx2morton, r0, r1);
    if (sp == (size_t)-1) return -2;
    if (r0 == sp) {
      bnodes[ib * 3 + 1] = (uint32_t)(num_bnode + sp);
    } else {
      bnodes[ib * 3 + 1] = (uint32_t)sp;
      bnodes[sp * 3] = (uint32_t)ib;
    }
    if (r1 == sp + 1) {
      bnodes[ib * 3 + 2] = (uint32_t)(num_bnode + sp + 1);
    } else {
      bnodes[ib * 3 + 2] = (uint32_t)(sp + 1);
      bnodes[(sp + 1) * 3] = (uint32_t)ib;
    }
    const uint32_t delta = (uint32_t)clz32(idx2morton[sp] ^ idx2morton[sp + 1]);
    const uint32_t offset = 32u - num_dim * max_depth;
    bnode2depth[ib] = (delta - offset) / num_dim;
  }
  uint32_t *isonode = (uint32_t *)calloc(num_bnode, sizeof(uint32_t));
  if (!isonode) return -3;
  for (size_t ib = 0; ib < num_bnode; ++ib) {
    if (bnodes[ib * 3 + 1] >= num_bnode) idx2bnode[bnodes[ib * 3 + 1] - num_bnode] = (uint32_t)ib;
    if (bnodes[ib * 3 + 2] >= num_bnode) idx2bnode[bnodes[ib * 3 + 2] - num_bnode] = (uint32_t)ib;
    if (ib == 0) continue;
    const uint32_t p = bnodes[ib * 3];
    if (bnode2depth[ib] != bnode2depth[p]) isonode[ib - 1] = 1;
  }
  bnode2onode[0] = 0;
  for (size_t ib = 1; ib < num_bnode; ++ib) bnode2onode[ib] = bnode2onode[ib - 1] + isonode[ib - 1];
  free(isonode);
  return 0;
}

/* quad_oct_tree.rs:51-63 */
static void morton2center(uint32_t morton, uint32_t num_dim, uint32_t depth, uint32_t max_depth,
                          float *center) {
  const uint32_t sh = (max_depth - depth) * num_dim;
  uint32_t key = sh < 32u ? morton >> sh : 0u;
  for (uint32_t d = 0; d < num_dim; ++d) center[d] = 0.5f;
  for (uint32_t k = 0; k < depth; ++k)
    for (uint32_t d = 0; d < num_dim; ++d) {
      const uint32_t j = num_dim - d - 1;
      center[j] += (float)(key & 1u);
      center[j] *= 0.5f;
      key >>= 1;
    }
}

void orc_morton2center(uint32_t morton, uint32_t num_dim, uint32_t depth, uint32_t max_depth,
                       float *center) {
  morton2center(morton, num_dim, depth, max_depth, center);
}

/* quad_oct_tree.rs:111-227 make_tree_from_binary_radix_tree (onodes come
 * filled with U32MAX and onode2depth with 0, as the reference's wrapper
 * does); returns -1 on a duplicated child (the reference's assert) */
int orc_octree_make_tree(size_t num_vtx, const uint32_t *bnodes, const uint32_t *bnode2onode,
                         const uint32_t *bnode2depth, const uint32_t *idx2bnode,
                         const uint32_t *idx2morton, size_t num_onode, uint32_t max_depth,
                         uint32_t num_dim, uint32_t *onodes, uint32_t *onode2depth,
                         float *onode2center, uint32_t *idx2onode, float *idx2center) {
  const uint32_t num_child = 1u << num_dim;
  const size_t nlink = num_child + 1;
  for (size_t i = 0; i < num_vtx; ++i) {
    {
      size_t cur = idx2bnode[i];
      uint32_t op, dp;
      for (;;) {
        if (cur == 0 || bnode2onode[cur] != bnode2onode[cur - 1]) {
          op = bnode2onode[cur];
          dp = bnode2depth[cur];
          break;
        }
        cur = bnodes[cur * 3];
      }
      if (dp >= max_depth) return -2;
      const uint32_t key = idx2morton[i];
      const uint32_t c = (key >> ((max_depth - 1 - dp) * num_dim)) & (num_child - 1);
      if (onodes[op * nlink + 1 + c] != U32MAX) return -1;
      onodes[op * nlink + 1 + c] = (uint32_t)(i + num_onode);
      idx2onode[i] = op;
      float center[3];
      morton2center(key, num_dim, max_depth, max_depth, center);
      for (uint32_t d = 0; d < num_dim; ++d) idx2center[i * num_dim + d] = center[d];
    }
    if (i + 1 < num_vtx) {
      const size_t ib = i;
      if (ib == 0) {
        for (uint32_t d = 0; d < num_dim; ++d) onode2center[d] = 0.5f;
        continue;
      }
      if (bnode2onode[ib - 1] == bnode2onode[ib]) continue;
      size_t par = bnodes[ib * 3];
      uint32_t op, dp;
      for (;;) {
        if (par == 0 || bnode2onode[par] != bnode2onode[par - 1]) {
          op = bnode2onode[par];
          dp = bnode2depth[par];
          break;
        }
        par = bnodes[par * 3];
      }
      if (dp >= max_depth) return -2;
      const uint32_t morton = idx2morton[ib];
      const uint32_t c = (morton >> ((max_depth - 1 - dp) * num_dim)) & (num_child - 1);
      if (onodes[op * nlink + 1 + c] != U32MAX) return -1;
      const uint32_t o = bnode2onode[ib];
      onodes[op * nlink + 1 + c] = o;
      onodes[(size_t)o * nlink] = op;
      const uint32_t depth = bnode2depth[ib];
      float center[3];
      morton2center(morton, num_dim, depth, max_depth, center);
      for (uint32_t d = 0; d < num_dim; ++d) onode2center[(size_t)o * num_dim + d] = center[d];
      onode2depth[o] = depth;
    }
  }
  return 0;
}

/* quad_oct_tree.rs:420-451 aggregate: per point, climb and add (parent link
 * = entry 0 of num_link links) */
void orc_octree_aggregate(size_t num_idx, uint32_t num_vdim, const float *idx2val,
                          const uint32_t *idx2onode, uint32_t num_link, size_t num_onode,
                          const uint32_t *onodes, float *onode2aggval) {
  memset(onode2aggval, 0, sizeof(float) * num_onode * num_vdim);
  for (size_t idx = 0; idx < num_idx; ++idx) {
    uint32_t o = idx2onode[idx];
    while (o != U32MAX && o < num_onode) {
      for (uint32_t d = 0; d < num_vdim; ++d)
        onode2aggval[(size_t)o * num_vdim + d] += idx2val[idx * num_vdim + d];
      o = onodes[(size_t)o * num_link];
    }
  }
}

/* array1d.rs:1-15 */
void orc_unique_for_sorted(size_t n, const uint32_t *val, uint32_t *idx2jdx) {
  if (n == 0) return;
  idx2jdx[0] = 0;
  for (size_t i = 0; i + 1 < n; ++i) idx2jdx[i + 1] = val[i] == val[i + 1] ? 0u : 1u;
  for (size_t i = 0; i + 1 < n; ++i) idx2jdx[i + 1] += idx2jdx[i];
}

/* map_idx.rs:83-115 inverse (bisection), then jdx2val[j] = val[offset[j]] */
void orc_unique_jdx2val_jdx2idx(size_t n, const uint32_t *val, const uint32_t *idx2jdx,
                                size_t num_jdx, uint32_t *jdx2val, uint32_t *jdx2idx_offset) {
  for (size_t j = 0; j <= num_jdx; ++j) {
    size_t i0 = 0, i2 = n;
    uint32_t res;
    for (;;) {
      if (i2 - i0 == 1) {
        res = (size_t)idx2jdx[i0] >= j ? (uint32_t)i0 : (uint32_t)i2;
        break;
      }
      const size_t i1 = (i0 + i2) / 2;
      if ((size_t)idx2jdx[i1] < j) i0 = i1;
      else i2 = i1;
    }
    jdx2idx_offset[j] = res;
  }
  for (size_t j = 0; j < num_jdx; ++j) jdx2val[j] = val[jdx2idx_offset[j]];
}

/* del-msh-dlpack/src/offset_array.rs (CPU branch): accumulates */
void orc_offset_array_aggregate(size_t num_idx, const uint32_t *off, const uint32_t *jdx2kdx,
                                uint32_t num_dim, const float *kdx2val, float *idx2aggval) {
  for (size_t i = 0; i < num_idx; ++i)
    for (uint32_t j = off[i]; j < off[i + 1]; ++j) {
      const size_t k = jdx2kdx[j];
      for (uint32_t d = 0; d < num_dim; ++d) idx2aggval[i * num_dim + d] += kdx2val[k * num_dim + d];
    }
}

/* =========================================================================
 * Mixed polyhedron meshes in CSR form (del-msh-cpu/src/polyhedron_mesh.rs,
 * elem2center.rs, bvhnode2aabb3.rs).  The pyramid / prism / hex parametric
 * coordinates follow the written rule of the kernels (del-geo's solver is
 * not vendored): Newton's method from a fixed start, Cramer's rule, at most
 * 20 steps, converged when max |d| < 1e-6, inside with a 1e-5 margin.
 * ========================================================================= */

/* elem2center.rs:59-89 */
void orc_polyhedron_elem2center(size_t num_elem, const uint32_t *off, const uint32_t *idx2vtx,
                                uint32_t num_dim, const float *xyz, float *cog) {
  for (size_t e = 0; e < num_elem; ++e) {
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t i = off[e]; i < off[e + 1]; ++i)
      for (uint32_t d = 0; d < num_dim; ++d) acc[d] += xyz[(size_t)idx2vtx[i] * num_dim + d];
    const uint32_t n = off[e + 1] - off[e];
    const float ratio = n == 0 ? 0.f : 1.f / (float)n;
    for (uint32_t d = 0; d < num_dim; ++d) cog[e * num_dim + d] = acc[d] * ratio;
  }
}

/* bvhnode2aabb3.rs:197-259 (recursive, children first, margin 0) */
static void poly_refit_rec(float *aabb, size_t node, const uint32_t *bn, const uint32_t *off,
                           const uint32_t *idx2vtx, const float *xyz) {
  float *b = aabb + node * 6;
  if (bn[node * 3 + 2] == U32MAX) {
    const uint32_t e = bn[node * 3 + 1];
    for (uint32_t i = off[e]; i < off[e + 1]; ++i) {
      const float *p = xyz + (size_t)idx2vtx[i] * 3;
      for (int d = 0; d < 3; ++d) {
        if (i == off[e]) {
          b[d] = p[d];
          b[d + 3] = p[d];
        } else {
          b[d] = fmin_sel(b[d], p[d]);
          b[d + 3] = fmax_sel(b[d + 3], p[d]);
        }
      }
    }
    return;
  }
  const size_t c0 = bn[node * 3 + 1], c1 = bn[node * 3 + 2];
  poly_refit_rec(aabb, c0, bn, off, idx2vtx, xyz);
  poly_refit_rec(aabb, c1, bn, off, idx2vtx, xyz);
  for (int d = 0; d < 3; ++d) {
    b[d] = fmin_sel(aabb[c0 * 6 + d], aabb[c1 * 6 + d]);
    b[d + 3] = fmax_sel(aabb[c0 * 6 + d + 3], aabb[c1 * 6 + d + 3]);
  }
}

void orc_polyhedron_refit(float *aabb, const uint32_t *bvhnodes, size_t num_elem,
                          const uint32_t *off, const uint32_t *idx2vtx, const float *xyz) {
  if (num_elem) poly_refit_rec(aabb, 0, bvhnodes, off, idx2vtx, xyz);
}

typedef struct { float x, y, z; } v3f;
static inline v3f v3sub(v3f a, v3f b) { v3f r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline v3f v3neg(v3f a) { v3f r = {-a.x, -a.y, -a.z}; return r; }
/* a . (b x c) */
static inline float v3det(v3f a, v3f b, v3f c) {
  const float cx = b.y * c.z - b.z * c.y;
  const float cy = b.z * c.x - b.x * c.z;
  const float cz = b.x * c.y - b.y * c.x;
  return (a.x * cx + a.y * cy) + a.z * cz;
}

/* polyhedron_mesh.rs:89-100 with barycentrics by volume ratios */
static int tet_param(const v3f *p, float pco[3]) {
  const v3f e1 = v3sub(p[1], p[0]), e2 = v3sub(p[2], p[0]), e3 = v3sub(p[3], p[0]);
  const float vol = v3det(e1, e2, e3);
  if (vol == 0.f) return 0;
  const v3f o = v3neg(p[0]);
  const float b0 = v3det(p[1], p[2], p[3]) / vol;
  const float b1 = v3det(o, e2, e3) / vol;
  const float b2 = v3det(e1, o, e3) / vol;
  if (!(b0 > 0.f && b1 > 0.f && b2 > 0.f && ((1.f - b0) - b1) - b2 > 0.f)) return 0;
  pco[0] = b0;
  pco[1] = b1;
  pco[2] = b2;
  return 1;
}

/* shape functions with the operation order of polyhedron_mesh.cu:134-163 */
void orc_shape(int nn, float r, float s, float t, float *N, float *dr, float *ds, float *dt) {
  if (nn == 5) {
    const float rm = 1.f - r, sm = 1.f - s, tm = 1.f - t;
    N[0] = (rm * sm) * tm; N[1] = (r * sm) * tm; N[2] = (r * s) * tm; N[3] = (rm * s) * tm;
    N[4] = t;
    if (!dr) return;
    dr[0] = -(sm * tm); dr[1] = sm * tm; dr[2] = s * tm; dr[3] = -(s * tm); dr[4] = 0.f;
    ds[0] = -(rm * tm); ds[1] = -(r * tm); ds[2] = r * tm; ds[3] = rm * tm; ds[4] = 0.f;
    dt[0] = -(rm * sm); dt[1] = -(r * sm); dt[2] = -(r * s); dt[3] = -(rm * s); dt[4] = 1.f;
  } else if (nn == 6) {
    const float rs = (1.f - r) - s, tm = 1.f - t;
    N[0] = rs * tm; N[1] = r * tm; N[2] = s * tm; N[3] = rs * t; N[4] = r * t; N[5] = s * t;
    if (!dr) return;
    dr[0] = -tm; dr[1] = tm; dr[2] = 0.f; dr[3] = -t; dr[4] = t; dr[5] = 0.f;
    ds[0] = -tm; ds[1] = 0.f; ds[2] = tm; ds[3] = -t; ds[4] = 0.f; ds[5] = t;
    dt[0] = -rs; dt[1] = -r; dt[2] = -s; dt[3] = rs; dt[4] = r; dt[5] = s;
  } else {
    static const float sr[8] = {-1.f, 1.f, 1.f, -1.f, -1.f, 1.f, 1.f, -1.f};
    static const float ss[8] = {-1.f, -1.f, 1.f, 1.f, -1.f, -1.f, 1.f, 1.f};
    static const float st[8] = {-1.f, -1.f, -1.f, -1.f, 1.f, 1.f, 1.f, 1.f};
    const float e = 0.125f;
    for (int i = 0; i < 8; ++i) {
      const float fr = sr[i] > 0.f ? 1.f + r : 1.f - r;
      const float fs = ss[i] > 0.f ? 1.f + s : 1.f - s;
      const float ft = st[i] > 0.f ? 1.f + t : 1.f - t;
      N[i] = ((e * fr) * fs) * ft;
      if (dr) {
        dr[i] = ((e * sr[i]) * fs) * ft;
        ds[i] = ((e * fr) * ss[i]) * ft;
        dt[i] = ((e * fr) * fs) * st[i];
      }
    }
  }
}

static int newton_param(int nn, const v3f *p, float r, float s, float t, float pco[3]) {
  int converged = 0;
  for (int it = 0; it < 20; ++it) {
    float N[8], dr[8], ds[8], dt[8];
    orc_shape(nn, r, s, t, N, dr, ds, dt);
    v3f x = {0.f, 0.f, 0.f}, jr = {0.f, 0.f, 0.f}, js = {0.f, 0.f, 0.f}, jt = {0.f, 0.f, 0.f};
    for (int i = 0; i < nn; ++i) {
      x.x = x.x + N[i] * p[i].x; x.y = x.y + N[i] * p[i].y; x.z = x.z + N[i] * p[i].z;
      jr.x = jr.x + dr[i] * p[i].x; jr.y = jr.y + dr[i] * p[i].y; jr.z = jr.z + dr[i] * p[i].z;
      js.x = js.x + ds[i] * p[i].x; js.y = js.y + ds[i] * p[i].y; js.z = js.z + ds[i] * p[i].z;
      jt.x = jt.x + dt[i] * p[i].x; jt.y = jt.y + dt[i] * p[i].y; jt.z = jt.z + dt[i] * p[i].z;
    }
    const float d = v3det(jr, js, jt);
    if (!(d != 0.f) || !(fabsf(d) < INFINITY)) return 0;
    const v3f f = v3neg(x);
    const float dr0 = v3det(f, js, jt) / d;
    const float ds0 = v3det(jr, f, jt) / d;
    const float dt0 = v3det(jr, js, f) / d;
    r = r + dr0;
    s = s + ds0;
    t = t + dt0;
    const float m = fmaxf(fmaxf(fabsf(dr0), fabsf(ds0)), fabsf(dt0));
    if (!(m == m)) return 0;
    if (m < 1e-6f) { converged = 1; break; }
  }
  if (!converged) return 0;
  const float lo = -1e-5f, hi = 1.f + 1e-5f;
  int inside;
  if (nn == 5) inside = r >= lo && r <= hi && s >= lo && s <= hi && t >= lo && t <= hi;
  else if (nn == 6) inside = r >= lo && s >= lo && r + s <= hi && t >= lo && t <= hi;
  else inside = r >= -hi && r <= hi && s >= -hi && s <= hi && t >= -hi && t <= hi;
  if (!inside) return 0;
  pco[0] = r;
  pco[1] = s;
  pco[2] = t;
  return 1;
}

/* polyhedron_mesh.rs:79-130: the element's vertices shifted by -q */
int orc_param_coord(const float *q, const uint32_t *node2vtx, uint32_t n, const float *xyz,
                    float pco[3]) {
  v3f p[8];
  if (n != 4 && n != 5 && n != 6 && n != 8) return 0;
  for (uint32_t k = 0; k < n; ++k) {
    const float *x = xyz + (size_t)node2vtx[k] * 3;
    p[k].x = x[0] - q[0];
    p[k].y = x[1] - q[1];
    p[k].z = x[2] - q[2];
  }
  if (n == 4) return tet_param(p, pco);
  if (n == 5) return newton_param(5, p, 0.5f, 0.5f, 0.25f, pco);
  if (n == 6) return newton_param(6, p, 1.f / 3.f, 1.f / 3.f, 0.5f, pco);
  return newton_param(8, p, 0.f, 0.f, 0.f, pco);
}

/* polyhedron_mesh.rs:470-522 recursion */
static void poly_search_rec(const float *q, const uint32_t *bn, const float *aabb,
                            const uint32_t *off, const uint32_t *idx2vtx, const float *xyz,
                            size_t node, uint32_t *best, float *pco) {
  if (*best != U32MAX) return;
  const float *b = aabb + node * 6;
  if (!(b[0] <= q[0] && q[0] <= b[3] && b[1] <= q[1] && q[1] <= b[4] && b[2] <= q[2] &&
        q[2] <= b[5]))
    return;
  if (bn[node * 3 + 2] == U32MAX) {
    const uint32_t e = bn[node * 3 + 1];
    float c[3];
    if (orc_param_coord(q, idx2vtx + off[e], off[e + 1] - off[e], xyz, c)) {
      *best = e;
      pco[0] = c[0];
      pco[1] = c[1];
      pco[2] = c[2];
    }
    return;
  }
  poly_search_rec(q, bn, aabb, off, idx2vtx, xyz, bn[node * 3 + 1], best, pco);
  poly_search_rec(q, bn, aabb, off, idx2vtx, xyz, bn[node * 3 + 2], best, pco);
}

/* polyhedron_mesh.rs:524-575 */
void orc_polyhedron_search(size_t num_elem, const uint32_t *off, const uint32_t *idx2vtx,
                           const float *xyz, const uint32_t *bvhnodes, const float *aabb,
                           size_t num_wtx, const float *wtx2xyz, uint32_t *wtx2elem,
                           float *wtx2param) {
  for (size_t w = 0; w < num_wtx; ++w) {
    uint32_t best = U32MAX;
    float pco[3] = {0.f, 0.f, 0.f};
    if (num_elem) poly_search_rec(wtx2xyz + w * 3, bvhnodes, aabb, off, idx2vtx, xyz, 0, &best, pco);
    wtx2elem[w] = best;
    wtx2param[w * 3] = pco[0];
    wtx2param[w * 3 + 1] = pco[1];
    wtx2param[w * 3 + 2] = pco[2];
  }
}

/* polyhedron_mesh.rs:868-907 with polyhedron_mesh.cu:127-176's shape
 * functions; accumulates onto wtx2val (zeroed by the caller) */
void orc_polyhedron_interpolate(size_t num_wtx, const uint32_t *off, const uint32_t *idx2vtx,
                                uint32_t num_vdim, const float *vtx2val, const uint32_t *wtx2elem,
                                const float *wtx2param, float *wtx2val) {
  for (size_t w = 0; w < num_wtx; ++w) {
    const uint32_t e = wtx2elem[w];
    if (e == U32MAX) continue;
    const uint32_t n = off[e + 1] - off[e];
    const float r = wtx2param[w * 3], s = wtx2param[w * 3 + 1], t = wtx2param[w * 3 + 2];
    float sf[8];
    if (n == 4) {
      sf[0] = r; sf[1] = s; sf[2] = t; sf[3] = ((1.f - r) - s) - t;
    } else if (n == 5 || n == 6 || n == 8) {
      orc_shape((int)n, r, s, t, sf, NULL, NULL, NULL);
    } else {
      continue;
    }
    for (uint32_t k = 0; k < n; ++k) {
      const float *val = vtx2val + (size_t)idx2vtx[off[e] + k] * num_vdim;
      for (uint32_t d = 0; d < num_vdim; ++d) wtx2val[w * num_vdim + d] += sf[k] * val[d];
    }
  }
}

/* The kernels' order for the octree aggregate: each node adds its children
 * in slot order (point child: its value row; node child: its sum), from 0. */
static void oct_sum_rec(uint32_t o, size_t num_onode, uint32_t num_link, const uint32_t *onodes,
                        uint32_t num_vdim, const float *idx2val, float *out) {
  for (uint32_t c = 1; c < num_link; ++c) {
    const uint32_t ch = onodes[(size_t)o * num_link + c];
    if (ch != U32MAX && ch < num_onode) oct_sum_rec(ch, num_onode, num_link, onodes, num_vdim, idx2val, out);
  }
  for (uint32_t d = 0; d < num_vdim; ++d) {
    float acc = 0.f;
    for (uint32_t c = 1; c < num_link; ++c) {
      const uint32_t ch = onodes[(size_t)o * num_link + c];
      if (ch == U32MAX) continue;
      acc += ch >= num_onode ? idx2val[(size_t)(ch - num_onode) * num_vdim + d]
                             : out[(size_t)ch * num_vdim + d];
    }
    out[(size_t)o * num_vdim + d] = acc;
  }
}

void orc_octree_aggregate_tree(size_t num_onode, uint32_t num_link, const uint32_t *onodes,
                               uint32_t num_vdim, const float *idx2val, float *out) {
  if (num_onode) oct_sum_rec(0, num_onode, num_link, onodes, num_vdim, idx2val, out);
}
