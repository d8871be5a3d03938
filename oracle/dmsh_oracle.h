/*
 * dmsh_oracle.h -- CPU oracle for the del-msh LBVH / ray-cast / edge-gradient
 * hot path.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark load this library as the checker.
 * Nothing in the product (libdmsh, the Python drop-in) may call it.
 *
 * Every function restates one function of del-msh-cpu (Rust) in plain C and
 * cites the reference file:line it follows.  The float primitives that
 * del-msh-cpu takes from the un-vendored crate del-geo-core are restated from
 * the published methods named beside each one; all of them are evaluated
 * left to right in IEEE binary32 with no FMA contraction (-ffp-contract=off).
 *
 * Index type: uint32_t, as in the Python entry points (torch.uint32).
 * Missing triangle / no hit: UINT32_MAX.
 */
#ifndef DMSH_ORACLE_H
#define DMSH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Morton codes (del-msh-cpu/src/mortons.rs) ------------------------- */
uint32_t orc_expand_bits2(uint32_t x);                 /* mortons.rs:3-8     */
uint32_t orc_morton_code2(float x, float y);           /* mortons.rs:20-27   */
uint32_t orc_expand_bits3(uint32_t x);                 /* mortons.rs:76-81   */
uint32_t orc_morton_code3(float x, float y, float z);  /* mortons.rs:94-102  */
/* 63-bit extension (no reference counterpart): 21 bits per axis. */
uint64_t orc_expand_bits3_u64(uint32_t x);
uint64_t orc_morton_code3_u64(float x, float y, float z);

/* mortons.rs:171-212; transform is column-major (num_dim+1)^2.
 * Returns 0, or -1 on a bad num_dim. */
int orc_vtx2morton(size_t num_vtx, const float *vtx2co, int num_dim,
                   const float *transform_co2unit, uint32_t *vtx2morton);
int orc_vtx2morton_u64(size_t num_vtx, const float *vtx2xyz,
                       const float *transform_co2unit, uint64_t *vtx2morton);

/* ---- stable argsort (mortons.rs:139-146: `sort_by_key`, stable) -------- */
void orc_argsort_u32(size_t n, const uint32_t *keys, uint32_t *idx2obj,
                     uint32_t *idx2key);
void orc_argsort_u64(size_t n, const uint64_t *keys, uint32_t *idx2obj,
                     uint64_t *idx2key);

/* ---- Karras topology (mortons.rs:216-325, bvhnodes_morton.rs:10-50) ---- */
void orc_range_u32(const uint32_t *idx2morton, size_t n, size_t idx,
                   size_t *first, size_t *last);        /* mortons.rs:221-288 */
size_t orc_split_u32(const uint32_t *idx2morton, size_t first,
                     size_t last);                      /* mortons.rs:290-325 */
void orc_range_u64(const uint64_t *idx2morton, size_t n, size_t idx,
                   size_t *first, size_t *last);
size_t orc_split_u64(const uint64_t *idx2morton, size_t first, size_t last);
/* bvhnodes [2n-1][3] = (parent, left|obj, right|UINT32_MAX) */
int orc_bvhnodes_u32(uint32_t *bvhnodes, const uint32_t *idx2obj,
                     const uint32_t *idx2morton, size_t n);
int orc_bvhnodes_u64(uint32_t *bvhnodes, const uint32_t *idx2obj,
                     const uint64_t *idx2morton, size_t n);

/* Validators; return 0 when the invariant holds, a nonzero code otherwise.
 * mortons.rs:330-352, mortons.rs:384-433, bvhnodes.rs:22-36 */
int orc_check_morton_code_range_split_u32(const uint32_t *idx2morton, size_t n);
int orc_check_binary_radix_tree_u32(const uint32_t *bvhnodes,
                                    const uint32_t *idx2morton, size_t n);
int orc_check_bvh_topology(const uint32_t *bvhnodes, size_t num_leaf);

/* ---- refit (bvhnode2aabb3.rs) ------------------------------------------ */
/* bvhnode2aabb3.rs:68-142 with elem2vtx = tri2vtx, num_noel = 3, no CCD.
 * aabb [num_node][6] = (xmin,ymin,zmin,xmax,ymax,zmax). */
int orc_bvhnode2aabb_trimesh3(float *bvhnode2aabb, const uint32_t *bvhnodes,
                              size_t num_node, const uint32_t *tri2vtx,
                              const float *vtx2xyz, float eps);
/* bvhnode2aabb3.rs:6-63 (points, no CCD) */
int orc_bvhnode2aabb_points(float *bvhnode2aabb, const uint32_t *bvhnodes,
                            size_t num_node, const float *vtx2xyz);

/* ---- geometry formula set (del-geo-core restated) ---------------------- */
/* mat4_col_major::transform_homogeneous; returns 0 when w == 0 */
int orc_transform_homogeneous(const float m[16], const float p[3], float out[3]);
/* mat4_col_major::ray_from_transform_ndc2world_and_pixel_coordinates */
void orc_ray_for_pixel(uint32_t iw, uint32_t ih, uint32_t img_w, uint32_t img_h,
                       const float ndc2world[16], float org[3], float dir[3]);
/* aabb3 slab test; returns 1 and (tmin,tmax) on hit */
int orc_ray_aabb(const float aabb[6], const float org[3], const float dir[3],
                 float *tmin, float *tmax);
/* tri3::intersection_against_ray (Moller-Trumbore); returns 1 on hit */
int orc_ray_tri(const float p0[3], const float p1[3], const float p2[3],
                const float org[3], const float dir[3], float *t, float bc[3]);
/* tri3::intersection_against_line (no t >= 0 requirement) */
int orc_line_tri(const float p0[3], const float p1[3], const float p2[3],
                 const float org[3], const float dir[3], float *t, float bc[3]);
/* mat4_col_major::try_inverse_with_pivot; returns 0 when singular */
int orc_mat4_inverse(const float m[16], float inv[16]);
/* centroid of triangles, elem2center.rs:8-37 (sum * (1/3)) */
void orc_tri2centroid(size_t num_tri, const uint32_t *tri2vtx,
                      const float *vtx2xyz, float *tri2cntr);
/* Mat44/torch.py:120-140 fit into the unit cube; out is column-major */
void orc_fit_into_unit_cube(size_t num_pnt, const float *vtx2xyz, float out[16]);

/* ---- ray-cast (search_bvh3.rs, pix2tri.rs, pix2depth.rs) --------------- */
typedef struct {
  uint64_t nodes_visited; /* internal nodes expanded (both child boxes tested) */
  uint64_t tris_tested;   /* leaves reached (ray-triangle tests performed)     */
  uint64_t frames_hist[64]; /* rays by the deepest nesting of nodes whose two
                               children were both hit (clamped to 63): the
                               frame-stack depth an iterative traversal needs */
} orc_trace_stats;

/* search_bvh3.rs:103-221 from the root with dis = +inf.
 * Returns 1 on hit with (t, tri, bc). */
int orc_first_intersection_ray(const float org[3], const float dir[3],
                               const uint32_t *tri2vtx, const float *vtx2xyz,
                               const uint32_t *bvhnodes, size_t num_node,
                               const float *bvhnode2aabb, float *t,
                               uint32_t *i_tri, float bc[3],
                               orc_trace_stats *stats);
/* pix2tri.rs:3-48; pix2t (may be NULL) receives the hit ray coefficient.
 * num_threads <= 1 runs serially (the reference runs it on a rayon pool). */
int orc_pix2tri(uint32_t *pix2tri, float *pix2t, size_t num_tri,
                const uint32_t *tri2vtx, const float *vtx2xyz,
                const uint32_t *bvhnodes, const float *bvhnode2aabb,
                uint32_t img_w, uint32_t img_h, const float ndc2world[16],
                int num_threads, orc_trace_stats *stats);
/* the same image by a flat closest-hit search (one best hit, far children on
 * a stack, ties to the later hit); equal to orc_pix2tri by construction, and
 * checked to be so */
int orc_pix2tri_flat(uint32_t *pix2tri, float *pix2t, size_t num_tri, const uint32_t *tri2vtx,
                     const float *vtx2xyz, const uint32_t *bvhnodes, const float *aabb,
                     uint32_t img_w, uint32_t img_h, const float T[16], int num_threads,
                     orc_trace_stats *stats);
/* pix2depth.rs:107-146 */
int orc_pix2depth(float *pix2depth, const uint32_t *pix2tri,
                  const uint32_t *tri2vtx, const float *vtx2xyz,
                  uint32_t img_w, uint32_t img_h, const float ndc2world[16]);
/* pix2depth.rs:27-56 + trimesh3_raycast.rs:25-71 (accumulates) */
int orc_pix2depth_bwd(float *dldw_vtx2xyz, const uint32_t *pix2tri,
                      const uint32_t *tri2vtx, const float *vtx2xyz,
                      const float *dldw_pix2depth, uint32_t img_w, uint32_t img_h,
                      const float ndc2world[16]);

/* ---- contour and anti-aliasing (edge2vtx.rs, antialias.rs) ------------- */
/* edge2vtx.rs:191-242: edges whose two triangles face opposite ways along
 * the view ray through the edge midpoint.  Writes the selected edges' vertex
 * pairs in edge order into cedge2vtx (capacity num_edge) and returns their
 * count, or -1 when world2ndc is singular. */
long orc_contour_for_triangle_mesh(const uint32_t *tri2vtx, const float *vtx2xyz,
                                   const float world2ndc[16], size_t num_edge,
                                   const uint32_t *edge2vtx, const uint32_t *edge2tri,
                                   uint32_t *cedge2vtx);
/* edge2::intersection_edge2: c0 + rc (c1 - c0) = s0 + re (e1 - s0); returns 1
 * with (rc, re) when both lie in [0, 1] */
int orc_intersection_edge2(const float c0[2], const float c1[2], const float s0[2],
                           const float e1[2], float *rc, float *re);
/* edge2::dldw_intersection_edge2 (dldre = 0): dL/ds0 and dL/de1 of dL/drc */
void orc_dldw_intersection_edge2(const float c0[2], const float c1[2], const float s0[2],
                                 const float e1[2], float dldrc, float dlds0[2],
                                 float dlde1[2]);
/* antialias.rs:54-99 (pixel walk of dda.h:6-63, neighbour pairs of
 * differentiable_rasterizer.cu:15-64); pix2vout must start as a copy of pix2vin */
long orc_antialias(size_t num_cedge, const uint32_t *cedge2vtx, const float *vtx2xyz,
                   const float world2pix[16], uint32_t img_w, uint32_t img_h,
                   const uint32_t *pix2tri, const float *pix2vin, float *pix2vout);
/* antialias.rs:102-159 (accumulates into dldw_vtx2xyz).  Both return the
 * number of edges whose pixel walk was abandoned after 2^26 steps. */
long orc_bwd_antialias(size_t num_cedge, const uint32_t *cedge2vtx, const float *vtx2xyz,
                        float *dldw_vtx2xyz, const float world2pix[16], uint32_t img_w,
                        uint32_t img_h, const float *pix2val, const float *dldw_pix2val,
                       const uint32_t *pix2tri);

/* ---- differentiable rasterizer ----------------------------------------- */
/* rasterized_edge_gradient.rs:434-603 (accumulates into dldw_vtx2xyz) */
void orc_edge_gradient_bwd(const uint32_t *tri2vtx, const float *vtx2xyz,
                           float *dldw_vtx2xyz, const float world2pix[16],
                           uint32_t img_w, uint32_t img_h,
                           const uint32_t *pix2tri, uint32_t num_vdim,
                           const float *pix2val, const float *dldw_pix2val);

/* ---- attribute interpolation ------------------------------------------- */
/* pix2tri.rs:176-229 (F7 plane barycentrics); returns -1 on a bad index */
int orc_interpolate(float *pix2val, const uint32_t *pix2tri, uint32_t img_w, uint32_t img_h,
                    const uint32_t *tri2vtx, size_t num_tri, const float *vtx2xyz,
                    size_t num_vtx, uint32_t num_vdim, const float *vtx2val,
                    const float ndc2world[16]);
/* pix2tri.rs:231-304 (F8; accumulates) */
int orc_interpolate_bwd(float *dldw_vtx2xyz, float *dldw_vtx2val, const uint32_t *pix2tri,
                        uint32_t img_w, uint32_t img_h, const uint32_t *tri2vtx, size_t num_tri,
                        const float *vtx2xyz, size_t num_vtx, uint32_t num_vdim,
                        const float *vtx2val, const float ndc2world[16],
                        const float *dldw_pix2val);

/* ---- smoothed edge gradient (staggered pixel-edge grid) ----------------- */
/* rasterized_edge_gradient.rs:77-200; types 0 same triangle, 1 shared edge,
 * 2 / 3 the first / second pixel's triangle in front, 4 intersection */
void orc_edge_gradient_and_type(const uint32_t *tri2vtx, const float *vtx2xyz,
                                const float world2pix[16], uint32_t img_w, uint32_t img_h,
                                const uint32_t *pix2tri, uint32_t num_vdim,
                                const float *pix2val, const float *dldw_pix2val,
                                uint8_t *hedge2type, float *hedge2dldr, uint8_t *vedge2type,
                                float *vedge2dldr);
/* rasterized_edge_gradient.rs:250-432 (serial, del-msh-cpu) */
void orc_smooth_gradient_serial(uint32_t img_w, uint32_t img_h, const uint8_t *hedge2type,
                                const uint8_t *vedge2type, uint32_t num_iter, float *hedge2dldr,
                                float *vedge2dldr);
/* rasterized_edge_gradient.cu:139-275 (red-black, the reference GPU path) */
void orc_smooth_gradient_red_black(uint32_t img_w, uint32_t img_h, const uint8_t *hedge2type,
                                   const uint8_t *vedge2type, uint32_t num_iter,
                                   float *hedge2dldr, float *vedge2dldr);
/* rasterized_edge_gradient.rs:202-248 (img_w, img_h >= 2) */
void orc_interpolate_staggered_grid(uint32_t img_w, uint32_t img_h, const float *hedge2vy,
                                    const float *vedge2vx, size_t num_vtx, const float *vtx2xy,
                                    float *vtx2velo);

/* ---- mesh topology ------------------------------------------------------- */
/* vtx2elem.rs:58-96: elements around each vertex (offsets vtx2idx[num_vtx+1],
 * idx2elem[num_elem*num_node], elements ascending per vertex).  Returns 0, or
 * -1 on a vertex index >= num_vtx (the reference asserts). */
int orc_vtx2elem(size_t num_elem, size_t num_node, const uint32_t *elem2vtx, size_t num_vtx,
                 uint32_t *vtx2idx, uint32_t *idx2elem);
/* vtx2vtx.rs:11-76 from_uniform_mesh_with_vtx2elem: vertices around each
 * vertex in first-appearance order over the vertex's elements (ascending) and
 * their nodes; is_self keeps the vertex itself.  Pass idx2vtx = NULL to count
 * only.  Returns the number of entries (vtx2idx[num_vtx]). */
size_t orc_vtx2vtx(size_t num_elem, size_t num_node, const uint32_t *elem2vtx, size_t num_vtx,
                   const uint32_t *vtx2jdx, const uint32_t *jdx2elem, int is_self,
                   uint32_t *vtx2idx, uint32_t *idx2vtx);
/* edge2vtx.rs:16-36 from_vtx2vtx: edge (i, j) for every entry j of vertex i */
void orc_edge2vtx_from_vtx2vtx(size_t num_vtx, const uint32_t *vtx2idx, const uint32_t *idx2vtx,
                               uint32_t *edge2vtx);
/* edge2elem.rs:20-70: the first two triangles around edge2vtx[e][0] (in
 * vtx2idx/idx2tri order) that also hold edge2vtx[e][1]; entries not found
 * are left as they are.  Returns 0, or -1 where a listed triangle does not
 * hold the edge's first vertex exactly once (the reference panics). */
int orc_edge2elem(size_t num_edge, const uint32_t *edge2vtx, const uint32_t *tri2vtx,
                  const uint32_t *vtx2idx, const uint32_t *idx2tri, uint32_t *edge2tri);
/* vtx2vtx.rs:210-252: num_iter Jacobi sweeps of [I + lambda L] x = rhs,
 * each sweep vtx2lhs -> vtx2lhstmp -> vtx2lhs */
void orc_laplacian_smoothing(size_t num_vtx, const uint32_t *vtx2idx, const uint32_t *idx2vtx,
                             float lambda, size_t num_dim, float *vtx2lhs, const float *vtx2rhs,
                             size_t num_iter, float *vtx2lhstmp);
/* vtx2vtx.rs:254-280: lhs = L rhs, L = [.., -1, .., valence, .., -1, ..] */
void orc_multiply_graph_laplacian(size_t num_vtx, const uint32_t *vtx2idx,
                                  const uint32_t *idx2vtx, size_t num_dim, const float *vtx2rhs,
                                  float *vtx2lhs);

/* quad/oct-tree (del-msh-cpu/src/quad_oct_tree.rs) and sorted-array helpers */
int orc_octree_bnodes(size_t num_idx, const uint32_t *idx2morton, uint32_t num_dim,
                      uint32_t max_depth, uint32_t *bnodes, uint32_t *bnode2depth,
                      uint32_t *bnode2onode, uint32_t *idx2bnode);
void orc_morton2center(uint32_t morton, uint32_t num_dim, uint32_t depth, uint32_t max_depth,
                       float *center);
int orc_octree_make_tree(size_t num_vtx, const uint32_t *bnodes, const uint32_t *bnode2onode,
                         const uint32_t *bnode2depth, const uint32_t *idx2bnode,
                         const uint32_t *idx2morton, size_t num_onode, uint32_t max_depth,
                         uint32_t num_dim, uint32_t *onodes, uint32_t *onode2depth,
                         float *onode2center, uint32_t *idx2onode, float *idx2center);
void orc_octree_aggregate(size_t num_idx, uint32_t num_vdim, const float *idx2val,
                          const uint32_t *idx2onode, uint32_t num_link, size_t num_onode,
                          const uint32_t *onodes, float *onode2aggval);
void orc_octree_aggregate_tree(size_t num_onode, uint32_t num_link, const uint32_t *onodes,
                               uint32_t num_vdim, const float *idx2val, float *out);
void orc_unique_for_sorted(size_t n, const uint32_t *val, uint32_t *idx2jdx);
void orc_unique_jdx2val_jdx2idx(size_t n, const uint32_t *val, const uint32_t *idx2jdx,
                                size_t num_jdx, uint32_t *jdx2val, uint32_t *jdx2idx_offset);
void orc_offset_array_aggregate(size_t num_idx, const uint32_t *off, const uint32_t *jdx2kdx,
                                uint32_t num_dim, const float *kdx2val, float *idx2aggval);

/* mixed polyhedron meshes in CSR form (polyhedron_mesh.rs, elem2center.rs,
 * bvhnode2aabb3.rs) */
void orc_polyhedron_elem2center(size_t num_elem, const uint32_t *off, const uint32_t *idx2vtx,
                                uint32_t num_dim, const float *xyz, float *cog);
void orc_polyhedron_refit(float *aabb, const uint32_t *bvhnodes, size_t num_elem,
                          const uint32_t *off, const uint32_t *idx2vtx, const float *xyz);
void orc_shape(int nn, float r, float s, float t, float *N, float *dr, float *ds, float *dt);
int orc_param_coord(const float *q, const uint32_t *node2vtx, uint32_t n, const float *xyz,
                    float pco[3]);
void orc_polyhedron_search(size_t num_elem, const uint32_t *off, const uint32_t *idx2vtx,
                           const float *xyz, const uint32_t *bvhnodes, const float *aabb,
                           size_t num_wtx, const float *wtx2xyz, uint32_t *wtx2elem,
                           float *wtx2param);
void orc_polyhedron_interpolate(size_t num_wtx, const uint32_t *off, const uint32_t *idx2vtx,
                                uint32_t num_vdim, const float *vtx2val, const uint32_t *wtx2elem,
                                const float *wtx2param, float *wtx2val);

/* ---- all-cores variant of the build (dmsh_cpu_mt.c; baseline only) ----
 * centroid -> fit -> 63-bit Morton -> stable argsort -> topology -> refit
 * (eps 0) on num_threads OpenMP threads; the same tree and boxes as the
 * serial stages.  Returns 0, or a negative value on a malformed tree. */
int orc_build_trimesh3_u64_mt(size_t num_tri, const uint32_t *tri2vtx, const float *vtx2xyz,
                              int num_threads, uint32_t *bvhnodes, float *aabb);

#ifdef __cplusplus
}
#endif
#endif
