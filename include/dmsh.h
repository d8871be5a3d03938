/*
 * dmsh.h -- C ABI of libdmsh, the MI355X (gfx950) implementation of the
 * del-msh LBVH / ray-cast / differentiable-rasterizer hot path.
 *
 * This ABI replaces the launch plumbing that del-msh reaches through
 * del-msh-cudarc-sys + del-cudarc-sys (cuModuleLoadData / cuLaunchKernel on
 * nvcc fatbins, del-msh-cudarc-sys/src/lib.rs:7-27) with host launchers that
 * take plain device pointers, sizes and a HIP stream.  Conventions:
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream); every
 *     entry point only enqueues work on it and returns without synchronizing.
 *   - all arrays are device pointers owned by the caller; layouts are the
 *     reference ones (row-major [n][k], uint32 indices, float32 values,
 *     4x4 transforms column-major as the Python layer passes them: M.T).
 *   - entry points that need scratch take (ws, ws_bytes); the matching
 *     *_workspace_bytes() gives the size.  Nothing is allocated inside.
 *   - return value: DMSH_OK or an error code; arguments are validated on the
 *     host before anything is enqueued.  Data errors found on the device
 *     (an index out of range in bvhnodes / tri2vtx, a traversal stack
 *     overflow) set sticky bits readable with dmsh_get_device_errors().
 */
#ifndef DMSH_H
#define DMSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMSH_OK 0
#define DMSH_ERR_INVALID 1   /* bad size / null pointer / unsupported option */
#define DMSH_ERR_HIP 2       /* a HIP runtime call or launch failed          */
#define DMSH_ERR_WORKSPACE 3 /* workspace smaller than *_workspace_bytes()    */
#define DMSH_ERR_TOO_LARGE 4 /* input beyond an allocation-free entry point's
                                fixed scratch                               */

/* device error bits (dmsh_get_device_errors) */
#define DMSH_DEVERR_INDEX 0x1u          /* index out of range in input data   */
#define DMSH_DEVERR_STACK 0x2u          /* ray traversal needed > 4096 frames */
#define DMSH_DEVERR_SORT_PROGRESS 0x4u  /* radix-sort look-back gave up       */
#define DMSH_DEVERR_REFIT 0x8u          /* refit climb found a broken tree    */
#define DMSH_DEVERR_SINGULAR 0x10u      /* transform could not be inverted    */
#define DMSH_DEVERR_WALK 0x20u          /* antialias pixel walk > 2^26 steps  */

const char *dmsh_status_string(int status);
int dmsh_abi_version(void);

/* Reads and clears the sticky device error bits of the current device
 * (synchronizes the device). */
int dmsh_get_device_errors(uint32_t *bits_out);

/* Process-wide launch tuning (no reference counterpart; set by the
 * autotuner).  Keys: "pix2tri.stack" (LDS frames per lane: 12, 16, 20, 24,
 * 32; default 20), "pix2tri.waves" (waves per workgroup: 1, 2, 4), and the
 * refit's "refit.leaves_in_aabb" (0, 1), "refit.climb" (0, 1, 2; default 2),
 * "refit.queue" (0, 1), "refit.top" (0, 1), "refit.plan_cache" (0, 1),
 * "refit.compact" (levels climbed before the remaining climbs move to the
 * first lanes, 0..1023; default 0: never), and the sort's "sort.wide_tiles"
 * (digit passes in 512-thread tiles: 0 never, 1 from 2^21 elements (default),
 * 2 always).  get returns -1 for an unknown key. */
int dmsh_set_tuning(const char *key, int value);
int dmsh_get_tuning(const char *key);

/* Kernel timing with HIP events on the launch stream (no reference
 * counterpart; used by benchmarks and the autotuner).  enable(1) clears what
 * was recorded and starts recording, enable(0) stops.  While on, each kernel
 * launch of the entry points below is filed under the kernel's name
 * ("k_pix2tri", "k_onesweep", ...).  query synchronizes the recorded events
 * of a label and returns the launch count, the summed and the longest
 * duration (max_ms may be NULL).  Keep it off under graph capture. */
int dmsh_timing_enable(int on);
int dmsh_timing_query(const char *label, uint32_t *launches, double *total_ms,
                      double *max_ms);

/* ---------------------------------------------------------------------
 * Morton codes
 * Replaces kernel `vtx2morton` (del-msh-cuda-kernels/cuda/mortons.cu:7-36)
 * launched by mortons_vtx2morton_from_vtx2co (del-msh-dlpack/src/mortons.rs:17-77);
 * CPU counterpart del-msh-cpu/src/mortons.rs:171-212.
 * vtx2co [num_vtx][num_dim] (num_dim 2 or 3), transform (num_dim+1)^2
 * column-major.  2D: 16 bits/axis; 3D: 10 bits/axis (30-bit codes).
 * ------------------------------------------------------------------- */
int dmsh_vtx2morton_u32(void *stream, uint32_t num_vtx, const float *vtx2co,
                        uint32_t num_dim, const float *transform_co2unit,
                        uint32_t *vtx2morton);
/* 63-bit extension (21 bits/axis, 3D only); no reference counterpart. */
int dmsh_vtx2morton_u64(void *stream, uint32_t num_vtx, const float *vtx2xyz,
                        const float *transform_co2unit, uint64_t *vtx2morton);

/* ---------------------------------------------------------------------
 * Stable argsort of keys (LSD radix sort, `key_bits` low bits significant).
 * Replaces the GPU branch of array1d_argsort
 * (del-msh-dlpack/src/array1d.rs:99-154, thrust sort_u32_by_u32_key at :146);
 * CPU counterpart: the stable `sort_by_key` of del-msh-cpu/src/mortons.rs:143.
 * On return keys_inout holds the sorted keys and idx_out the permutation
 * (idx_out[j] = original index of the j-th smallest key; ties keep input order).
 * n < 2^30.
 * ------------------------------------------------------------------- */
size_t dmsh_argsort_workspace_bytes(uint32_t n, uint32_t key_bytes);
int dmsh_argsort_u32(void *stream, uint32_t n, uint32_t *keys_inout,
                     uint32_t *idx_out, uint32_t key_bits, void *ws,
                     size_t ws_bytes);
int dmsh_argsort_u64(void *stream, uint32_t n, uint64_t *keys_inout,
                     uint32_t *idx_out, uint32_t key_bits, void *ws,
                     size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Karras LBVH topology from sorted codes.
 * Replaces kernel `kernel_MortonCode_BVHTopology`
 * (del-msh-cuda-kernels/cuda/bvhnodes_morton.cu:28-72) launched by
 * mortons_make_bvhnodes_from_sorted_mortons (del-msh-dlpack/src/mortons.rs:79-135);
 * CPU counterpart del-msh-cpu/src/bvhnodes_morton.rs:10-50.
 * bvhnodes [2n-1][3] = (parent, left|obj, right|UINT32_MAX); root 0 has
 * parent UINT32_MAX; leaves are rows n-1 .. 2n-2.  n == 1 writes the single
 * leaf row (UINT32_MAX, idx2obj[0], UINT32_MAX).
 * ------------------------------------------------------------------- */
int dmsh_bvhnodes_from_sorted_mortons_u32(void *stream, uint32_t n,
                                          uint32_t *bvhnodes,
                                          const uint32_t *idx2morton,
                                          const uint32_t *idx2obj);
int dmsh_bvhnodes_from_sorted_mortons_u64(void *stream, uint32_t n,
                                          uint32_t *bvhnodes,
                                          const uint64_t *idx2morton,
                                          const uint32_t *idx2obj);

/* ---------------------------------------------------------------------
 * Bottom-up AABB refit.
 * Replaces kernel `from_trimesh3` (del-msh-cuda-kernels/cuda/bvhnode2aabb.cu:6-65)
 * launched by bvhnode2aabb_update_aabb (del-msh-dlpack/src/bvhnode2aabb.rs:13-108);
 * CPU counterpart del-msh-cpu/src/bvhnode2aabb3.rs:68-142 (triangles) and
 * :6-63 (points).  bvhnode2aabb [num_bvhnode][6] = (min xyz, max xyz).
 * The tree must be an LBVH layout (leaves in rows num_leaf-1 .. 2 num_leaf-2).
 * A tree whose internal nodes also lie inside their own leaf ranges (every
 * Morton-code LBVH) is refit subtree by subtree in LDS; any other tree in
 * that layout by the per-leaf climb through arrival counters.  Both give
 * the same boxes.  The workspace (about 32 bytes per leaf) holds the
 * counters, the leaf ranges and the subtree lists -- the tree's plan -- with
 * a hash of the rows they were made from: a later call with the same
 * workspace and unchanged rows keeps the plan and skips its two kernels
 * (tuning key "refit.plan_cache" = 0 rebuilds it every call).  Calls that
 * share a workspace must run in order (one stream).
 * ------------------------------------------------------------------- */
size_t dmsh_refit_workspace_bytes(uint32_t num_leaf);
int dmsh_bvhnode2aabb_trimesh3(void *stream, float *bvhnode2aabb,
                               uint32_t num_bvhnode, const uint32_t *bvhnodes,
                               uint32_t num_tri, const uint32_t *tri2vtx,
                               uint32_t num_vtx, const float *vtx2xyz,
                               float eps, void *ws, size_t ws_bytes);
int dmsh_bvhnode2aabb_points(void *stream, float *bvhnode2aabb,
                             uint32_t num_bvhnode, const uint32_t *bvhnodes,
                             uint32_t num_pnt, const float *vtx2xyz, void *ws,
                             size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Whole LBVH build of a triangle mesh / a point set, device resident:
 * centroid -> unit-cube fit -> Morton (30-bit u32 or 63-bit u64) -> stable
 * radix argsort -> Karras topology -> refit.  Replaces the Python driver
 * TriMesh3.torch.make_bvhnodes_bvhnode2aabb
 * (del-msh-dlpack/del_msh_dlpack/TriMesh3/torch.py:233-285), which chains
 * the five launches above through torch; CPU counterpart
 * del-msh-cpu/src/bvhnodes_morton.rs:151-189 + bvhnode2aabb3.rs:168-195.
 * morton_bits: 30 or 63.
 * ------------------------------------------------------------------- */
size_t dmsh_build_workspace_bytes(uint32_t num_elem, uint32_t morton_bits);
int dmsh_build_trimesh3(void *stream, uint32_t num_tri, const uint32_t *tri2vtx,
                        uint32_t num_vtx, const float *vtx2xyz,
                        uint32_t morton_bits, uint32_t *bvhnodes,
                        float *bvhnode2aabb, void *ws, size_t ws_bytes);
/* Allocation-free variant of dmsh_build_trimesh3.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as dmsh_build_trimesh3 without (ws, ws_bytes):
 * the workspace is a fixed DMSH_NOALLOC_SCRATCH_BYTES (64 MiB) device array
 * inside the library, one per GPU, held until the process ends.  Meshes
 * whose dmsh_build_workspace_bytes(num_tri, morton_bits) exceeds it (above
 * 769,642 triangles with 63-bit codes, 867,093 with 30-bit codes) are
 * rejected with DMSH_ERR_TOO_LARGE before anything is enqueued.  It calls no
 * hipMalloc / hipFree.  Every call on a GPU shares that GPU's scratch: calls
 * that overlap on one GPU (on different streams, or from different host
 * threads) race; calls in order on one stream do not. */
#define DMSH_NOALLOC_SCRATCH_BYTES (64u << 20)
int dmsh_build_trimesh3_noalloc(void *stream, uint32_t num_tri, const uint32_t *tri2vtx,
                                uint32_t num_vtx, const float *vtx2xyz, uint32_t morton_bits,
                                uint32_t *bvhnodes, float *bvhnode2aabb);
int dmsh_build_points(void *stream, uint32_t num_pnt, const float *vtx2xyz,
                      uint32_t morton_bits, uint32_t *bvhnodes,
                      float *bvhnode2aabb, void *ws, size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Per-pixel BVH ray-cast with the ordered near-first semantics of
 * del-msh-cpu (search_bvh3.rs:103-221, pix2tri.rs:3-48).
 * Replaces kernel `pix_to_tri` (del-msh-cuda-kernels/cuda/pix2tri.cu:12-76)
 * launched by pix2tri_by_raycast (del-msh-dlpack/src/pix2tri.rs:15-94).
 * pix2tri [img_h][img_w]: hit triangle or UINT32_MAX.  pix2t (nullable):
 * ray coefficient of the hit (0 on a miss); the ray runs from NDC z = +1
 * (t = 0) to NDC z = -1 (t = 1).
 * ------------------------------------------------------------------- */
size_t dmsh_pix2tri_workspace_bytes(uint32_t num_tri, uint32_t img_w,
                                    uint32_t img_h);
int dmsh_pix2tri(void *stream, uint32_t *pix2tri, float *pix2t,
                 uint32_t num_tri, const uint32_t *tri2vtx, uint32_t num_vtx,
                 const float *vtx2xyz, uint32_t img_w, uint32_t img_h,
                 const float *transform_ndc2world, const uint32_t *bvhnodes,
                 const float *bvhnode2aabb, void *ws, size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Depth of the stored hit, (ndc_z + 1) / 2, 0 on background.
 * Replaces kernel `fwd` (del-msh-cuda-kernels/cuda/pix2depth.cu:11-47)
 * launched by pix2depth_update (del-msh-dlpack/src/pix2depth.rs:14-97); the
 * inverse transform is computed on the device (no host round trip).
 * CPU counterpart del-msh-cpu/src/pix2depth.rs:107-146.
 * ------------------------------------------------------------------- */
size_t dmsh_pix2depth_workspace_bytes(void);
int dmsh_pix2depth(void *stream, float *pix2depth, const uint32_t *pix2tri,
                   uint32_t num_tri, const uint32_t *tri2vtx, uint32_t num_vtx,
                   const float *vtx2xyz, uint32_t img_w, uint32_t img_h,
                   const float *transform_ndc2world, void *ws,
                   size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Depth backward (accumulates dL/d(vertex) into dldw_vtx2xyz).  Replaces
 * kernel `bwd_wrt_vtx2xyz` (del-msh-cuda-kernels/cuda/pix2depth.cu:49-97)
 * launched by pix2depth_bwd_wrt_vtx2xyz
 * (del-msh-dlpack/src/pix2depth.rs:101-190); CPU counterpart
 * del-msh-cpu/src/pix2depth.rs:27-56 driven by
 * del-msh-cpu/src/trimesh3_raycast.rs:25-71.  Same workspace as
 * dmsh_pix2depth (the inverse transform, computed on the device).
 * ------------------------------------------------------------------- */
int dmsh_pix2depth_bwd_wrt_vtx2xyz(void *stream, float *dldw_vtx2xyz,
                                   const uint32_t *pix2tri, uint32_t num_tri,
                                   const uint32_t *tri2vtx, uint32_t num_vtx,
                                   const float *vtx2xyz,
                                   const float *dldw_pix2depth, uint32_t img_w,
                                   uint32_t img_h,
                                   const float *transform_ndc2world, void *ws,
                                   size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Smoothed edge gradient (the backward of RasterizedEdgeGradient.
 * AutogradWithSmooth, del_msh_dlpack/RasterizedEdgeGradient/torch.py:
 * 237-274).  Replaces rasterized_edge_gradient.cu:43-326
 * (hedge/vedge_gradient_and_type, smooth_hedge/vedge_red_black,
 * interpolate_staggered_grid) as launched by del-msh-dlpack/src/
 * rasterized_edge_gradient.rs:153-450.  Pixel-edge grids: hedge (ih, iw)
 * between pixels (iw, ih) and (iw, ih+1), [H-1][W]; vedge (ih, iw) between
 * (iw, ih) and (iw+1, ih), [H][W-1]; types 0 same triangle, 1 shared edge,
 * 2 / 3 first / second pixel's triangle in front, 4 intersection;
 * dldr = sum_k (g0 + g1) * 0.5 * (v0 - v1) (rasterized_edge_gradient.rs:
 * 77-200).  dmsh_smooth_gradient runs num_iter red-black sweeps of both
 * grids in place (cells of type 2 / 3 stay fixed) as the reference GPU path
 * does; del-msh-cpu's serial sweeps (rasterized_edge_gradient.rs:250-432)
 * converge to the same fixed point.  img_w, img_h >= 2.  The smoothing
 * needs a device workspace of dmsh_smooth_gradient_workspace_bytes (a second
 * buffer of each grid and the per-cell links); it runs 16 half-sweeps per
 * launch in LDS tiles and gives bit for bit the values of the sweeps run one
 * colour at a time over the whole grid.
 * ------------------------------------------------------------------- */
int dmsh_edge_gradient_and_type(void *stream, const uint32_t *tri2vtx, uint32_t num_tri,
                                const float *vtx2xyz, uint32_t num_vtx,
                                const float *transform_world2pix, uint32_t img_w, uint32_t img_h,
                                const uint32_t *pix2tri, uint32_t num_vdim, const float *pix2val,
                                const float *dldw_pix2val, uint8_t *hedge2type,
                                float *hedge2dldr, uint8_t *vedge2type, float *vedge2dldr);
size_t dmsh_smooth_gradient_workspace_bytes(uint32_t img_w, uint32_t img_h);
int dmsh_smooth_gradient(void *stream, uint32_t img_w, uint32_t img_h, const uint8_t *hedge2type,
                         const uint8_t *vedge2type, uint32_t num_iter, float *hedge2dldr,
                         float *vedge2dldr, void *ws, size_t ws_bytes);
/* rasterized_edge_gradient.rs:202-248: vtx2velo [num_vtx][2] */
int dmsh_interpolate_staggered_grid(void *stream, uint32_t img_w, uint32_t img_h,
                                    const float *hedge2vy, const float *vedge2vx,
                                    uint32_t num_vtx, const float *vtx2xy, float *vtx2velo);
/* The tail of AutogradWithSmooth.backward (RasterizedEdgeGradient/torch.py:
 * 237-274: project the vertices, read the smoothed grids there, chain through
 * the projection) in one kernel: dldw_vtx2xyz [num_vtx][3] from the smoothed
 * hedge2dldr / vedge2dldr, vtx2xyz [num_vtx][3] and the column-major
 * transform_world2pix (4x4). */
int dmsh_smoothed_vertex_gradient(void *stream, uint32_t img_w, uint32_t img_h,
                                  const float *hedge2dldr, const float *vedge2dldr,
                                  uint32_t num_vtx, const float *vtx2xyz,
                                  const float *transform_world2pix, float *dldw_vtx2xyz);

/* ---------------------------------------------------------------------
 * Attribute interpolation over pix2tri and its backward.  Replaces
 * pix2tri.cu:78-182 (interpolate / interpolate_bwd) as launched by
 * del-msh-dlpack/src/pix2tri.rs:98-271; results follow del-msh-cpu/src/
 * pix2tri.rs:176-304: per covered pixel the barycentric coordinates of the
 * pixel's line against the plane of its triangle (u, v of the
 * Moller-Trumbore form, no range test), pix2val = bc0 val0 + bc1 val1 +
 * bc2 val2 ([H][W][num_vdim], 0 on the background).  The backward
 * accumulates into dldw_vtx2xyz [num_vtx][3] and dldw_vtx2val
 * [num_vtx][num_vdim] (zero them first) with float atomics.
 * transform_ndc2world: 16 floats column-major.
 * ------------------------------------------------------------------- */
int dmsh_interpolate(void *stream, const uint32_t *pix2tri, uint32_t img_w, uint32_t img_h,
                     uint32_t num_tri, const uint32_t *tri2vtx, uint32_t num_vtx,
                     const float *vtx2xyz, uint32_t num_vdim, const float *vtx2val,
                     const float *transform_ndc2world, float *pix2val);
int dmsh_interpolate_bwd(void *stream, const uint32_t *pix2tri, uint32_t img_w, uint32_t img_h,
                         uint32_t num_tri, const uint32_t *tri2vtx, uint32_t num_vtx,
                         const float *vtx2xyz, uint32_t num_vdim, const float *vtx2val,
                         const float *transform_ndc2world, const float *dldw_pix2val,
                         float *dldw_vtx2xyz, float *dldw_vtx2val);

/* ---------------------------------------------------------------------
 * Edge-gradient backward of the rasterized silhouette (accumulates into
 * dldw_vtx2xyz).  Replaces kernels `bwd_hedge` / `bwd_vedge`
 * (del-msh-cuda-kernels/cuda/rasterized_edge_gradient.cu:387-520) launched
 * by rasterized_edge_gradient_bwd
 * (del-msh-dlpack/src/rasterized_edge_gradient.rs:27-150); CPU counterpart
 * del-msh-cpu/src/rasterized_edge_gradient.rs:434-603.
 * transform_world2pix: 16 floats column-major.
 *   dmsh_rasterized_edge_gradient_bwd: one launch, the terms added with
 *   float atomics (as the reference GPU kernels do): equal to the CPU loop
 *   to rounding, summation order not fixed.
 *   dmsh_rasterized_edge_gradient_bwd_ordered: each vertex adds its terms in
 *   the CPU loop's order (vertical pixel pairs column by column, then the
 *   horizontal pairs), so dldw_vtx2xyz equals del-msh-cpu's bit for bit and
 *   does not change from run to run.  Workspace: about 56 bytes per pixel
 *   pair (2 W H pairs); img_w * img_h < 2^27.
 * ------------------------------------------------------------------- */
int dmsh_rasterized_edge_gradient_bwd(void *stream, const uint32_t *tri2vtx,
                                      uint32_t num_tri, const float *vtx2xyz,
                                      uint32_t num_vtx, float *dldw_vtx2xyz,
                                      const float *transform_world2pix,
                                      uint32_t img_w, uint32_t img_h,
                                      const uint32_t *pix2tri,
                                      uint32_t num_vdim, const float *pix2val,
                                      const float *dldw_pix2val);
size_t dmsh_rasterized_edge_gradient_bwd_workspace_bytes(uint32_t img_w, uint32_t img_h,
                                                         uint32_t num_vtx);
int dmsh_rasterized_edge_gradient_bwd_ordered(void *stream, const uint32_t *tri2vtx,
                                              uint32_t num_tri, const float *vtx2xyz,
                                              uint32_t num_vtx, float *dldw_vtx2xyz,
                                              const float *transform_world2pix, uint32_t img_w,
                                              uint32_t img_h, const uint32_t *pix2tri,
                                              uint32_t num_vdim, const float *pix2val,
                                              const float *dldw_pix2val, void *ws,
                                              size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Contour edges: the edges whose two triangles face opposite ways along the
 * view ray through the edge midpoint, in edge order.  Replaces kernel
 * `edge2vtx_contour_set_flag` (del-msh-cuda-kernels/cuda/edge2vtx.cu:28-75)
 * plus the flag compaction, launched by edge2vtx_contour_for_triangle_mesh
 * (del-msh-dlpack/src/edge2vtx.rs:19-115); CPU counterpart
 * del-msh-cpu/src/edge2vtx.rs:191-242.  The inverse of transform_world2ndc
 * (16 floats column-major) is computed on the device.  cedge2vtx needs room
 * for num_edge pairs; the number written lands in *num_cedge (device
 * memory), so reading it is the caller's one sync.
 * ------------------------------------------------------------------- */
size_t dmsh_contour_workspace_bytes(uint32_t num_edge);
int dmsh_contour_for_triangle_mesh(void *stream, uint32_t num_tri,
                                   const uint32_t *tri2vtx, uint32_t num_vtx,
                                   const float *vtx2xyz,
                                   const float *transform_world2ndc,
                                   uint32_t num_edge, const uint32_t *edge2vtx,
                                   const uint32_t *edge2tri, uint32_t *cedge2vtx,
                                   uint32_t *num_cedge, void *ws,
                                   size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Anti-aliased silhouette, forward: blends pix2vin across every contour
 * edge into pix2vout, which the caller has initialised (normally a copy of
 * pix2vin); a pixel written by several edges keeps the value of the last
 * write in serial (edge, step, pair) order, so the result is deterministic.
 * Replaces kernel `antialias_fwd`
 * (del-msh-cuda-kernels/cuda/differentiable_rasterizer.cu:69-123) launched
 * by differentiable_antialias_fwd
 * (del-msh-dlpack/src/differentiable_antialias.rs:103-182); CPU counterpart
 * del-msh-cpu/src/antialias.rs:54-99.  transform_world2pix: 16 floats
 * column-major.
 * ------------------------------------------------------------------- */
size_t dmsh_antialias_workspace_bytes(uint32_t img_w, uint32_t img_h);
int dmsh_antialias_fwd(void *stream, uint32_t num_cedge,
                       const uint32_t *cedge2vtx, uint32_t num_vtx,
                       const float *vtx2xyz, const float *transform_world2pix,
                       uint32_t img_w, uint32_t img_h, const uint32_t *pix2tri,
                       const float *pix2vin, float *pix2vout, void *ws,
                       size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Anti-aliased silhouette, backward (accumulates into dldw_vtx2xyz).
 * Replaces kernel `antialias_bwd`
 * (del-msh-cuda-kernels/cuda/differentiable_rasterizer.cu:126-196) launched
 * by differentiable_antialias_bwd
 * (del-msh-dlpack/src/differentiable_antialias.rs:16-101); CPU counterpart
 * del-msh-cpu/src/antialias.rs:102-159.
 * ------------------------------------------------------------------- */
int dmsh_antialias_bwd(void *stream, uint32_t num_cedge,
                       const uint32_t *cedge2vtx, uint32_t num_vtx,
                       const float *vtx2xyz, float *dldw_vtx2xyz,
                       const float *transform_world2pix, uint32_t img_w,
                       uint32_t img_h, const float *pix2val,
                       const float *dldw_pix2val, const uint32_t *pix2tri);

/* ---------------------------------------------------------------------
 * Batched build and ray-cast: many triangle meshes of uneven size, each
 * stage one launch over all of them.  No reference counterpart (the
 * reference builds and casts one mesh per call,
 * del-msh-dlpack/del_msh_dlpack/TriMesh3/torch.py:233-285 and
 * Pix2Tri/torch.py:8-61); per mesh the results equal what
 * dmsh_build_trimesh3 (30-bit codes) and dmsh_pix2tri return for that mesh.
 *
 * Meshes are concatenated in CSR form (all offsets device arrays of
 * num_mesh + 1 u32, first entry 0): triangles [mesh2tri[m], mesh2tri[m+1])
 * with vertex indices local to vertices [mesh2vtx[m], mesh2vtx[m+1]).
 * Empty meshes are allowed.  total_tri < 2^30.  Offsets that do not fit
 * the totals (first entry 0, non-decreasing, last entry = total), a
 * mesh2node other than the build's, or an image whose pixel count is not a
 * multiple of its width set DMSH_DEVERR_INDEX and the call then writes
 * nothing past the offsets check.
 * Build output: mesh2node[m] (written, num_mesh + 1 entries) = the first
 * row of mesh m in bvhnodes / bvhnode2aabb, which hold
 * sum_m (2 N_m - 1) rows in the single-mesh layout with local indices.
 * Ray-cast: image of mesh m = pixels [mesh2pix[m], mesh2pix[m+1]) of
 * pix2tri / pix2t (nullable), mesh2width[m] pixels wide, row 0 at the top,
 * cast with transforms_ndc2world + 16 m (column-major).
 * Tuning keys (results do not depend on them): "batch.stack" (16, 20, 24,
 * 32; default 20),
 * "batch.wgs_per_cu" (persistent ray-cast workgroups per CU, 1..64),
 * "batch.treelet_grid" (refit workgroups, 1..65536),
 * "batch.sort_items" (elements per thread of the digit passes: 8, 12, 16),
 * "batch.cast_waves" (waves per ray-cast workgroup: 1, 2, 4),
 * "batch.sort_seg" (1, the default: up to 8192 meshes, the composite keys
 * are sorted inside their meshes by three passes over the code's top 24 bits
 * and an LDS finish on the low 6; 0: one pass per digit of the whole key).
 * ------------------------------------------------------------------- */
int dmsh_batch_set_tuning(const char *key, int value);
int dmsh_batch_get_tuning(const char *key);
size_t dmsh_batch_build_workspace_bytes(uint32_t num_mesh, uint32_t total_tri);
int dmsh_batch_build_trimesh3(void *stream, uint32_t num_mesh, uint32_t total_tri,
                              uint32_t total_vtx, const uint32_t *mesh2tri,
                              const uint32_t *mesh2vtx, const uint32_t *tri2vtx,
                              const float *vtx2xyz, uint32_t *mesh2node,
                              uint32_t *bvhnodes, float *bvhnode2aabb, void *ws,
                              size_t ws_bytes);
size_t dmsh_batch_pix2tri_workspace_bytes(uint32_t num_mesh, uint32_t total_tri,
                                          uint32_t total_pix);
int dmsh_batch_pix2tri(void *stream, uint32_t num_mesh, uint32_t total_tri,
                       uint32_t total_vtx, uint32_t total_pix,
                       const uint32_t *mesh2tri,
                       const uint32_t *mesh2vtx, const uint32_t *tri2vtx,
                       const float *vtx2xyz, const uint32_t *mesh2node,
                       const uint32_t *bvhnodes, const float *bvhnode2aabb,
                       const uint32_t *mesh2pix, const uint32_t *mesh2width,
                       const float *transforms_ndc2world, uint32_t *pix2tri,
                       float *pix2t, void *ws, size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Serving: cached trees and one ray-cast launch over many requests.  No
 * reference counterpart (the serving loop that drives these lives in
 * del_msh_dlpack/Serve).
 *
 * dmsh_batch_pack packs the trees of a CSR batch (as built by
 * dmsh_batch_build_trimesh3, same offsets) into one self-contained buffer
 * per mesh of dmsh_packed_bytes(N_m) bytes at device address packed[m];
 * nonfinite[m] (written) is nonzero when a box of that tree is not finite.
 * A packed tree is all a ray-cast reads, so it is what a cache keeps.
 *
 * dmsh_cast_requests ray-casts requests[0..num_req) (a device array) into
 * pix2tri / pix2t (nullable): request r fills pixels [pix_start,
 * pix_start + img_w img_h) with pix_start the running sum of the earlier
 * requests' pixel counts (so total_pix is the sum of all of them); per
 * pixel the result equals dmsh_pix2tri on the request's mesh and camera.
 * Tuning keys as for the batched ray-cast.
 * ------------------------------------------------------------------- */
typedef struct dmsh_cast_request {
  uint64_t packed;    /* device address of a tree packed by dmsh_batch_pack */
  uint32_t num_tri;   /* triangles of that tree                             */
  uint32_t nonfinite; /* as dmsh_batch_pack wrote it for the tree            */
  uint32_t img_w, img_h;
  uint32_t pix_start;
  uint32_t reserved;
  float ndc2world[16]; /* column-major */
} dmsh_cast_request;   /* 96 bytes */

size_t dmsh_packed_bytes(uint32_t num_tri);
size_t dmsh_batch_pack_workspace_bytes(uint32_t num_mesh);
int dmsh_batch_pack(void *stream, uint32_t num_mesh, uint32_t total_tri,
                    uint32_t total_vtx, const uint32_t *mesh2tri,
                    const uint32_t *mesh2vtx, const uint32_t *tri2vtx,
                    const float *vtx2xyz, const uint32_t *mesh2node,
                    const uint32_t *bvhnodes, const float *bvhnode2aabb,
                    const uint64_t *packed, uint32_t *nonfinite, void *ws,
                    size_t ws_bytes);
/* The batched build (as dmsh_batch_build_trimesh3, workspace
 * dmsh_batch_build_workspace_bytes) whose refit writes every mesh's packed
 * buffer and nonfinite flag, byte for byte what dmsh_batch_pack gives after
 * the build; no reference counterpart (the serving cache).  mesh2node and
 * bvhnodes are the build's; bvhnode2aabb is scratch: on return only the rows
 * of each mesh's root, treelet roots and the nodes above them hold boxes. */
int dmsh_batch_build_pack_trimesh3(void *stream, uint32_t num_mesh,
                                   uint32_t total_tri, uint32_t total_vtx,
                                   const uint32_t *mesh2tri,
                                   const uint32_t *mesh2vtx,
                                   const uint32_t *tri2vtx,
                                   const float *vtx2xyz, uint32_t *mesh2node,
                                   uint32_t *bvhnodes, float *bvhnode2aabb,
                                   const uint64_t *packed, uint32_t *nonfinite,
                                   void *ws, size_t ws_bytes);
size_t dmsh_cast_requests_workspace_bytes(uint32_t num_req, uint32_t total_pix);
int dmsh_cast_requests(void *stream, uint32_t num_req,
                       const dmsh_cast_request *requests, uint32_t total_pix,
                       uint32_t *pix2tri, float *pix2t, void *ws,
                       size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Mesh topology: offset arrays (vtx2idx[num_vtx + 1] over a flat idx2*).
 * Replaces the launches of del-msh-cudarc-sys/src/vtx2elem.rs:2-68
 * (uniform_mesh.cu vtx2nelem / fill_idx2elem + offset_array::sort),
 * del-msh-cudarc-sys/src/vtx2vtx.rs:2-84 (vtx2vtx.cu
 * vtx2nvtx_from_uniform_mesh / idx2vtx_from_vtx2buff_for_uniform_mesh),
 * del-msh-dlpack/src/edge2vtx.rs:117-174 (edge2vtx.cu edge2vtx_from_vtx2vtx),
 * del-msh-dlpack/src/edge2elem.rs:17-83 (edge2elem.cu), and
 * del-msh-dlpack/src/vtx2vtx.rs:23-190 (vtx2vtx.cu laplacian_smoothing /
 * multiply_graph_laplacian).  Results equal del-msh-cpu's: elements
 * ascending per vertex (vtx2elem.rs:58-96), neighbours in first-appearance
 * order (vtx2vtx.rs:11-76), edge2tri as edge2elem.rs:20-70 (entries with no
 * second triangle are left untouched), Laplacian sweeps in f32 in the CPU's
 * operation order (vtx2vtx.rs:210-280).
 *
 * Vertex-to-vertex is two calls on one workspace: dmsh_vtx2vtx_count writes
 * vtx2idx (the caller reads vtx2idx[num_vtx] to size idx2vtx, as the
 * reference does) and keeps the lists in ws; dmsh_vtx2vtx_fill, with the
 * same sizes, is_self and workspace, writes idx2vtx.
 * Out-of-range indices set DMSH_DEVERR_INDEX.
 * ------------------------------------------------------------------- */
size_t dmsh_exclusive_scan_workspace_bytes(uint32_t n);
/* out[0..n] = exclusive prefix sums of in[0..n), out[n] = total; in may alias out */
int dmsh_exclusive_scan_u32(void *stream, uint32_t n, const uint32_t *in, uint32_t *out,
                            void *ws, size_t ws_bytes);
size_t dmsh_vtx2elem_workspace_bytes(uint32_t num_elem, uint32_t num_node, uint32_t num_vtx);
int dmsh_vtx2elem_from_uniform_mesh(void *stream, uint32_t num_elem, uint32_t num_node,
                                    const uint32_t *elem2vtx, uint32_t num_vtx,
                                    uint32_t *vtx2idx, uint32_t *idx2elem, void *ws,
                                    size_t ws_bytes);
size_t dmsh_vtx2vtx_workspace_bytes(uint32_t num_elem, uint32_t num_node, uint32_t num_vtx,
                                    int is_self);
int dmsh_vtx2vtx_count(void *stream, uint32_t num_elem, uint32_t num_node,
                       const uint32_t *elem2vtx, uint32_t num_vtx, int is_self,
                       uint32_t *vtx2idx, void *ws, size_t ws_bytes);
int dmsh_vtx2vtx_fill(void *stream, uint32_t num_elem, uint32_t num_node, uint32_t num_vtx,
                      int is_self, const uint32_t *vtx2idx, uint32_t *idx2vtx, void *ws,
                      size_t ws_bytes);
int dmsh_edge2vtx_from_vtx2vtx(void *stream, uint32_t num_vtx, const uint32_t *vtx2idx,
                               uint32_t num_idx, const uint32_t *idx2vtx, uint32_t *edge2vtx);
int dmsh_edge2elem_from_edge2vtx_of_tri2vtx(void *stream, uint32_t num_edge,
                                            const uint32_t *edge2vtx, uint32_t num_tri,
                                            const uint32_t *tri2vtx, uint32_t num_vtx,
                                            const uint32_t *vtx2idx, uint32_t num_idx,
                                            const uint32_t *idx2tri, uint32_t *edge2tri);
int dmsh_laplacian_smoothing(void *stream, uint32_t num_vtx, const uint32_t *vtx2idx,
                             uint32_t num_idx, const uint32_t *idx2vtx, float lambda,
                             uint32_t num_dim, float *vtx2lhs, const float *vtx2rhs,
                             uint32_t num_iter, float *vtx2lhstmp);
int dmsh_multiply_graph_laplacian(void *stream, uint32_t num_vtx, const uint32_t *vtx2idx,
                                  uint32_t num_idx, const uint32_t *idx2vtx, uint32_t num_dim,
                                  const float *vtx2rhs, float *vtx2lhs);

/* ---------------------------------------------------------------------
 * Quad/oct-tree from sorted unique Morton codes.  Replaces
 * quad_oct_tree.cu:28-233 (binary_radix_tree_and_depth,
 * bnode2isonode_and_idx2bnode, make_tree_from_binary_radix_tree, aggregate)
 * as launched by del-msh-dlpack/src/quad_oct_tree.rs; results follow
 * del-msh-cpu/src/quad_oct_tree.rs:1-230.  num_dim 2 or 3, num_dim *
 * max_depth <= 32 (max_depth is 32 / num_dim by default in the reference).
 * dmsh_octree_bnodes writes bnodes [num_idx-1][3], bnode2depth,
 * bnode2onode [num_idx-1] (the caller reads bnode2onode[num_idx-2] + 1 =
 * num_onode, as the reference does) and idx2bnode [num_idx];
 * dmsh_octree_make_tree fills onodes [num_onode][1+2^num_dim] (entry 0 the
 * parent, then the children; a point child is idx + num_onode; empty slots
 * 0xFFFFFFFF), onode2depth, onode2center [num_onode][num_dim], idx2onode,
 * idx2center [num_idx][num_dim].  dmsh_octree_aggregate writes every
 * node's sum of idx2val over its points, each node adding its children in
 * slot order (a fixed order, so the same bits on every run; del-msh-cpu adds
 * point by point and the reference GPU with atomics, so sums agree with
 * theirs bit for bit when the partial sums are exact, else to rounding).
 * ------------------------------------------------------------------- */
size_t dmsh_octree_bnodes_workspace_bytes(uint32_t num_idx);
int dmsh_octree_bnodes(void *stream, uint32_t num_idx, const uint32_t *idx2morton,
                       uint32_t num_dim, uint32_t max_depth, uint32_t *bnodes,
                       uint32_t *bnode2depth, uint32_t *bnode2onode, uint32_t *idx2bnode, void *ws,
                       size_t ws_bytes);
int dmsh_octree_make_tree(void *stream, uint32_t num_idx, const uint32_t *bnodes,
                          const uint32_t *bnode2onode, const uint32_t *bnode2depth,
                          const uint32_t *idx2bnode, const uint32_t *idx2morton, uint32_t num_dim,
                          uint32_t max_depth, uint32_t num_onode, uint32_t *onodes,
                          uint32_t *onode2depth, float *onode2center, uint32_t *idx2onode,
                          float *idx2center);
size_t dmsh_octree_aggregate_workspace_bytes(uint32_t num_onode);
int dmsh_octree_aggregate(void *stream, uint32_t num_idx, uint32_t num_vdim, const float *idx2val,
                          const uint32_t *idx2onode, uint32_t num_link, uint32_t num_onode,
                          const uint32_t *onodes, float *onode2aggval, void *ws,
                          size_t ws_bytes);

/* ---------------------------------------------------------------------
 * Sorted-array and offset-array helpers.  Replace del-msh-dlpack/src/
 * array1d.rs:156-300 (array1d_has_duplicate_sorted_array,
 * array1d_unique_for_sorted_array, array1d_unique_jdx2val_jdx2idx; GPU side
 * in the external del-cudarc) and offset_array.rs (offset_array_aggregate);
 * results follow del-msh-cpu/src/array1d.rs:1-15 and map_idx.rs:83-115.
 * dmsh_has_duplicate_sorted_u32 sets the device word *flag to 0 or 1.
 * dmsh_unique_for_sorted_u32: idx2jdx[i] = number of value changes in
 * idx2val[0..i].  dmsh_unique_jdx2val_jdx2idx: jdx2idx_offset [num_jdx+1],
 * jdx2val [num_jdx].  dmsh_offset_array_aggregate adds, per idx, the rows
 * kdx2val[jdx2kdx[j]] for j in [off[idx], off[idx+1]) in j order onto
 * idx2aggval (the reference's wrapper zero-fills it first).
 * ------------------------------------------------------------------- */
/* new2val[i] = old2val[new2old[i]].  Replaces the GPU branch of
 * array1d_permute (del-msh-dlpack/src/array1d.rs:17-70, del_cudarc_sys
 * array1d::permute); an index >= n raises DMSH_DEVERR_INDEX. */
int dmsh_permute_u32(void *stream, uint32_t n, const uint32_t *old2val, const uint32_t *new2old,
                     uint32_t *new2val);
int dmsh_has_duplicate_sorted_u32(void *stream, uint32_t n, const uint32_t *idx2val,
                                  uint32_t *flag);
size_t dmsh_unique_for_sorted_workspace_bytes(uint32_t n);
int dmsh_unique_for_sorted_u32(void *stream, uint32_t n, const uint32_t *idx2val,
                               uint32_t *idx2jdx, void *ws, size_t ws_bytes);
int dmsh_unique_jdx2val_jdx2idx(void *stream, uint32_t n, const uint32_t *idx2val,
                                const uint32_t *idx2jdx, uint32_t num_jdx, uint32_t *jdx2val,
                                uint32_t *jdx2idx_offset);
int dmsh_offset_array_aggregate(void *stream, uint32_t num_idx, const uint32_t *idx2jdx_offset,
                                uint32_t num_jdx, const uint32_t *jdx2kdx, uint32_t num_kdx,
                                uint32_t num_dim, const float *kdx2val, float *idx2aggval);

/* ---------------------------------------------------------------------
 * Mixed polyhedron meshes in CSR form (elem2idx_offset [num_elem+1],
 * idx2vtx [num_idx]; tet 4, pyramid 5, prism 6, hex 8 nodes).  Replaces
 * elem2center.cu (from_polygon_mesh_as_points), bvhnode2aabb.cu:67-122
 * (from_polyhedron_mesh) and polyhedron_mesh.cu:49-177
 * (search_elem_contain_points, interpolate_values_at_points) as launched by
 * del-msh-dlpack/src/polyhedron_mesh.rs; results follow del-msh-cpu/src/
 * elem2center.rs:59-89, bvhnode2aabb3.rs:197-259 and polyhedron_mesh.rs:
 * 79-155,470-575,868-907.  Point location walks the element BVH left child
 * first and reports the first element containing the point (wtx2elem
 * 0xFFFFFFFF and zero parameters if none).  The pyramid / prism / hex
 * parametric coordinates are those of a fixed Newton iteration (del-geo's
 * own solver is not available), shared with the test oracle.
 * ------------------------------------------------------------------- */
int dmsh_polyhedron_elem2center(void *stream, uint32_t num_elem, const uint32_t *elem2idx_offset,
                                uint32_t num_idx, const uint32_t *idx2vtx, uint32_t num_vtx,
                                uint32_t num_dim, const float *vtx2xyz, float *elem2center);
size_t dmsh_polyhedron_refit_workspace_bytes(uint32_t num_elem);
int dmsh_bvhnode2aabb_polyhedron(void *stream, uint32_t num_bvhnode, const uint32_t *bvhnodes,
                                 uint32_t num_elem, const uint32_t *elem2idx_offset,
                                 uint32_t num_idx, const uint32_t *idx2vtx, uint32_t num_vtx,
                                 const float *vtx2xyz, float *bvhnode2aabb, void *ws,
                                 size_t ws_bytes);
/* ws (optional, dmsh_polyhedron_search_workspace_bytes(num_wtx)): with it,
 * sets of 65536 or more query points are searched in the Morton order of the
 * points in the root box (only the order of work changes, not any result);
 * ws == NULL searches in the given order. */
size_t dmsh_polyhedron_search_workspace_bytes(uint32_t num_wtx);
int dmsh_polyhedron_search_elem_contain_points(void *stream, uint32_t num_elem,
                                               const uint32_t *elem2idx_offset, uint32_t num_idx,
                                               const uint32_t *idx2vtx, uint32_t num_vtx,
                                               const float *vtx2xyz, const uint32_t *bvhnodes,
                                               const float *bvhnode2aabb, uint32_t num_wtx,
                                               const float *wtx2xyz, uint32_t *wtx2elem,
                                               float *wtx2param, void *ws, size_t ws_bytes);
int dmsh_polyhedron_interpolate_at_points(void *stream, uint32_t num_wtx, uint32_t num_elem,
                                          const uint32_t *elem2idx_offset, uint32_t num_idx,
                                          const uint32_t *idx2vtx, uint32_t num_vtx,
                                          uint32_t num_vdim, const float *vtx2val,
                                          const uint32_t *wtx2elem, const float *wtx2param,
                                          float *wtx2val);

#ifdef __cplusplus
}
#endif
#endif /* DMSH_H */
