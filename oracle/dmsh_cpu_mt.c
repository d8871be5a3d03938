/*
 * dmsh_cpu_mt.c -- all-cores variant of the CPU LBVH build (triangle mesh,
 * 63-bit codes) for the benchmark's CPU baseline.  TEST / BASELINE
 * INFRASTRUCTURE ONLY, like the rest of oracle/: nothing in the product may
 * call it.
 *
 * del-msh-cpu builds on one thread.  This variant runs each stage of the same
 * pipeline on OpenMP threads and gives the same tree and boxes bit for bit:
 *   centroid, Morton codes  per element, same arithmetic
 *                           (elem2center.rs:8-37, mortons.rs:171-212)
 *   unit-cube fit           per-thread min / max over order keys, combined
 *                           (order-independent; Mat44/torch.py:120-140)
 *   stable argsort          LSD radix sort, 8-bit digits, per-thread
 *                           histograms over contiguous chunks prefixed in
 *                           thread order (stable, so the same permutation as
 *                           the serial merge sort; mortons.rs:139-146)
 *   Karras topology         one node per iteration (each writes its own row
 *                           and its children's parent words; mortons.rs:
 *                           216-325, bvhnodes_morton.rs:10-50)
 *   refit                   leaf boxes in parallel, then climbs through
 *                           per-node arrival counters (acq_rel); the second
 *                           arrival merges (left, right) as bvhnode2aabb3.rs:
 *                           136-141 does, so the order of arrival does not
 *                           change a bit.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dmsh_oracle.h"

#define U32MAX 0xFFFFFFFFu

static inline float fmin_sel(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_sel(float a, float b) { return (b > a) ? b : a; }

static inline uint32_t order_key(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
static inline float from_order_key(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Stable LSD radix argsort of 64-bit keys over `bits` low bits. */
static void radix_argsort_u64(size_t n, const uint64_t *keys, uint32_t *idx, uint64_t *sk,
                              int bits, int nth) {
  uint64_t *ka = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  uint64_t *kb = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  uint32_t *va = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t *vb = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  size_t *cnt = (size_t *)malloc(sizeof(size_t) * 256 * (size_t)nth);
#pragma omp parallel for num_threads(nth) schedule(static)
  for (size_t i = 0; i < n; ++i) {
    ka[i] = keys[i];
    va[i] = (uint32_t)i;
  }
  for (int shift = 0; shift < bits; shift += 8) {
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      const size_t lo = n * (size_t)t / (size_t)nth, hi = n * (size_t)(t + 1) / (size_t)nth;
      size_t *c = cnt + 256 * (size_t)t;
      memset(c, 0, sizeof(size_t) * 256);
      for (size_t i = lo; i < hi; ++i) ++c[(ka[i] >> shift) & 255u];
#pragma omp barrier
#pragma omp single
      {
        size_t sum = 0;
        for (int d = 0; d < 256; ++d)
          for (int u = 0; u < nth; ++u) {
            const size_t v = cnt[256 * (size_t)u + d];
            cnt[256 * (size_t)u + d] = sum;
            sum += v;
          }
      }
      for (size_t i = lo; i < hi; ++i) {
        const size_t pos = c[(ka[i] >> shift) & 255u]++;
        kb[pos] = ka[i];
        vb[pos] = va[i];
      }
    }
    uint64_t *tk = ka; ka = kb; kb = tk;
    uint32_t *tv = va; va = vb; vb = tv;
  }
#pragma omp parallel for num_threads(nth) schedule(static)
  for (size_t i = 0; i < n; ++i) {
    idx[i] = va[i];
    sk[i] = ka[i];
  }
  free(ka); free(kb); free(va); free(vb); free(cnt);
}

static inline void box_of_tri(const uint32_t *tri2vtx, const float *xyz, size_t t, float *b) {
  const float *p0 = xyz + (size_t)tri2vtx[t * 3] * 3;
  const float *p1 = xyz + (size_t)tri2vtx[t * 3 + 1] * 3;
  const float *p2 = xyz + (size_t)tri2vtx[t * 3 + 2] * 3;
  for (int d = 0; d < 3; ++d) {
    b[d] = p0[d] - 0.f;
    b[d + 3] = p0[d] + 0.f;
  }
  for (int d = 0; d < 3; ++d) {
    b[d] = fmin_sel(b[d], p1[d] - 0.f);
    b[d + 3] = fmax_sel(b[d + 3], p1[d] + 0.f);
  }
  for (int d = 0; d < 3; ++d) {
    b[d] = fmin_sel(b[d], p2[d] - 0.f);
    b[d + 3] = fmax_sel(b[d + 3], p2[d] + 0.f);
  }
}

int orc_build_trimesh3_u64_mt(size_t num_tri, const uint32_t *tri2vtx, const float *vtx2xyz,
                              int num_threads, uint32_t *bvhnodes, float *aabb) {
  const size_t n = num_tri;
  if (n == 0) return -1;
  const int nth = num_threads > 0 ? num_threads : 1;
  float *cntr = (float *)malloc(sizeof(float) * 3 * n);
  uint64_t *codes = (uint64_t *)malloc(sizeof(uint64_t) * n);
  uint64_t *sk = (uint64_t *)malloc(sizeof(uint64_t) * n);
  uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * n);
  uint32_t kmin[3] = {U32MAX, U32MAX, U32MAX}, kmax[3] = {0, 0, 0};
  const float ratio = 1.f / 3.f;
#pragma omp parallel num_threads(nth)
  {
    uint32_t lmin[3] = {U32MAX, U32MAX, U32MAX}, lmax[3] = {0, 0, 0};
#pragma omp for schedule(static)
    for (size_t t = 0; t < n; ++t) {
      for (int d = 0; d < 3; ++d) {
        float cog = 0.f;
        for (int k = 0; k < 3; ++k) cog += vtx2xyz[(size_t)tri2vtx[t * 3 + k] * 3 + d];
        const float c = cog * ratio;
        cntr[t * 3 + d] = c;
        const uint32_t key = order_key(c);
        if (key < lmin[d]) lmin[d] = key;
        if (key > lmax[d]) lmax[d] = key;
      }
    }
#pragma omp critical
    for (int d = 0; d < 3; ++d) {
      if (lmin[d] < kmin[d]) kmin[d] = lmin[d];
      if (lmax[d] > kmax[d]) kmax[d] = lmax[d];
    }
  }
  /* the fit of orc_fit_into_unit_cube from the reduced box */
  float mn[3], mx[3], T[16];
  for (int d = 0; d < 3; ++d) {
    mn[d] = from_order_key(kmin[d]);
    mx[d] = from_order_key(kmax[d]);
  }
  float ext = mx[0] - mn[0];
  for (int d = 1; d < 3; ++d) ext = fmax_sel(ext, mx[d] - mn[d]);
  const float s = 1.f / ext;
  memset(T, 0, sizeof T);
  T[0] = s;
  T[5] = s;
  T[10] = s;
  T[15] = 1.f;
  for (int d = 0; d < 3; ++d) {
    const float c = (mn[d] + mx[d]) * 0.5f;
    T[12 + d] = (-c) * s + 0.5f;
  }
#pragma omp parallel for num_threads(nth) schedule(static)
  for (size_t i = 0; i < n; ++i) {
    float q[3];
    if (!orc_transform_homogeneous(T, cntr + i * 3, q)) q[0] = q[1] = q[2] = NAN;
    codes[i] = orc_morton_code3_u64(q[0], q[1], q[2]);
  }
  radix_argsort_u64(n, codes, idx, sk, 63, nth);
  const size_t nb = n - 1;
  bvhnodes[0] = U32MAX;
  if (n == 1) {
    bvhnodes[1] = idx[0];
    bvhnodes[2] = U32MAX;
  }
  int bad = 0;
#pragma omp parallel for num_threads(nth) schedule(dynamic, 4096) reduction(| : bad)
  for (size_t ib = 0; ib < nb; ++ib) {
    size_t r0, r1;
    orc_range_u64(sk, n, ib, &r0, &r1);
    const size_t sp = orc_split_u64(sk, r0, r1);
    if (sp == (size_t)-1) {
      bad |= 1;
      continue;
    }
    if (r0 == sp) {
      const size_t a = nb + sp;
      bvhnodes[ib * 3 + 1] = (uint32_t)a;
      bvhnodes[a * 3] = (uint32_t)ib;
      bvhnodes[a * 3 + 1] = idx[sp];
      bvhnodes[a * 3 + 2] = U32MAX;
    } else {
      bvhnodes[ib * 3 + 1] = (uint32_t)sp;
      bvhnodes[sp * 3] = (uint32_t)ib;
    }
    if (r1 == sp + 1) {
      const size_t b = nb + sp + 1;
      bvhnodes[ib * 3 + 2] = (uint32_t)b;
      bvhnodes[b * 3] = (uint32_t)ib;
      bvhnodes[b * 3 + 1] = idx[sp + 1];
      bvhnodes[b * 3 + 2] = U32MAX;
    } else {
      bvhnodes[ib * 3 + 2] = (uint32_t)(sp + 1);
      bvhnodes[(sp + 1) * 3] = (uint32_t)ib;
    }
  }
  bvhnodes[0] = U32MAX;
  /* refit: leaf boxes, then climbs through arrival counters */
  uint32_t *flags = (uint32_t *)calloc(nb ? nb : 1, sizeof(uint32_t));
#pragma omp parallel for num_threads(nth) schedule(static) reduction(| : bad)
  for (size_t k = 0; k < n; ++k) {
    const size_t leaf = nb + k;
    float *box = aabb + leaf * 6;
    box_of_tri(tri2vtx, vtx2xyz, bvhnodes[leaf * 3 + 1], box);
    size_t node = leaf;
    uint32_t parent = bvhnodes[leaf * 3];
    while (parent != U32MAX) {
      if (__atomic_fetch_add(flags + parent, 1u, __ATOMIC_ACQ_REL) == 0) break;
      const size_t c0 = bvhnodes[(size_t)parent * 3 + 1], c1 = bvhnodes[(size_t)parent * 3 + 2];
      if (c0 != node && c1 != node) {
        bad |= 2;
        break;
      }
      const float *a = aabb + c0 * 6, *b = aabb + c1 * 6;
      float *out = aabb + (size_t)parent * 6;
      for (int d = 0; d < 3; ++d) {
        out[d] = fmin_sel(a[d], b[d]);
        out[d + 3] = fmax_sel(a[d + 3], b[d + 3]);
      }
      node = parent;
      parent = bvhnodes[(size_t)parent * 3];
    }
  }
  free(flags);
  free(cntr);
  free(codes);
  free(sk);
  free(idx);
  return bad ? -2 : 0;
}
