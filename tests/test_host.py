"""Host-side pieces that need no GPU: camera matrices and the library's
timing registry."""

import math

import numpy as np
import torch

import del_msh_dlpack.Mat44.torch as Mat44
from del_msh_dlpack import _lib


def _ndc(m, p):
    q = m.double().numpy() @ np.append(np.asarray(p, np.float64), 1.0)
    return q[:3] / q[3]


def test_camera_perspective_blender_depth_convention():
    """Reference test_pix2depth.py:23 uses (1, 30, 2, 6, True): the camera
    looks down -z, z = -near maps to NDC z = +1 and z = -far to -1."""
    m = Mat44.camera_perspective_blender(1.0, 30.0, 2.0, 6.0, True)
    assert math.isclose(_ndc(m, (0, 0, -2.0))[2], 1.0, abs_tol=1e-6)
    assert math.isclose(_ndc(m, (0, 0, -6.0))[2], -1.0, abs_tol=1e-6)
    # focus = cot(atan2(18, lens)) = lens / 18; x_ndc = focus * x / |z|
    p = _ndc(m, (0.3, -0.2, -4.0))
    assert math.isclose(p[0], 30.0 / 18.0 * 0.3 / 4.0, rel_tol=1e-5)
    assert math.isclose(p[1], 30.0 / 18.0 * -0.2 / 4.0, rel_tol=1e-5)
    # aspect > 1 scales y by the aspect, aspect < 1 scales x by 1/aspect
    m2 = Mat44.camera_perspective_blender(2.0, 30.0, 2.0, 6.0, True)
    assert math.isclose(_ndc(m2, (0, 0.3, -4.0))[1], 2.0 * 30.0 / 18.0 * 0.3 / 4.0, rel_tol=1e-5)
    m3 = Mat44.camera_perspective_blender(0.5, 30.0, 2.0, 6.0, True)
    assert math.isclose(_ndc(m3, (0.3, 0, -4.0))[0], 2.0 * 30.0 / 18.0 * 0.3 / 4.0, rel_tol=1e-5)


def test_camera_perspective_blender_other_direction():
    """Without proj_direction: the matrix for (-far, -near) times diag(-1,-1,-1,1),
    worked by hand for (1, 30, 2, 6) to [[f,0,0,0],[0,f,0,0],[0,0,-2,-6],[0,0,-1,0]]
    with f = 30/18: still looking down -z, depth reversed (near -> -1)."""
    m = Mat44.camera_perspective_blender(1.0, 30.0, 2.0, 6.0, False)
    f = 30.0 / 18.0
    want = np.array([[f, 0, 0, 0], [0, f, 0, 0], [0, 0, -2, -6], [0, 0, -1, 0]])
    assert np.allclose(m.double().numpy(), want, rtol=1e-6, atol=1e-6)
    assert math.isclose(_ndc(m, (0, 0, -2.0))[2], -1.0, abs_tol=1e-6)
    assert math.isclose(_ndc(m, (0, 0, -6.0))[2], 1.0, abs_tol=1e-6)


def test_perspective_near_plane_is_ndc_plus_one():
    m = Mat44.perspective(45.0, 1.0, 1.0, 5.0, dtype=torch.float64)
    assert math.isclose(_ndc(m, (0, 0, -1.0))[2], 1.0, abs_tol=1e-12)
    assert math.isclose(_ndc(m, (0, 0, -5.0))[2], -1.0, abs_tol=1e-12)
    f = 1.0 / math.tan(math.radians(22.5))
    assert math.isclose(_ndc(m, (0.5, 0, -2.0))[0], f * 0.25, rel_tol=1e-12)


def test_timing_registry_without_launches():
    _lib.timing_enable(True)
    assert _lib.timing_query("k_pix2tri") == (0, 0.0, 0.0)
    _lib.timing_enable(False)
    assert _lib.timing_query("no_such_kernel") == (0, 0.0, 0.0)


def test_edge_gradient_bwd_checks_transform_before_any_launch():
    """A float64 or 3x3 world-to-pixel matrix is refused on the host (it would
    otherwise be read as 16 float32 values)."""
    import pytest

    import del_msh_dlpack.RasterizedEdgeGradient.torch as RasterizedEdgeGradient

    tri = torch.zeros((1, 3), dtype=torch.uint32)
    vtx = torch.zeros((3, 3), dtype=torch.float32)
    pix = torch.zeros((4, 4), dtype=torch.uint32)
    val = torch.zeros((4, 4, 1), dtype=torch.float32)
    for bad in (torch.eye(4, dtype=torch.float64), torch.eye(3, dtype=torch.float32)):
        with pytest.raises(AssertionError):
            RasterizedEdgeGradient.bwd(tri, vtx, bad, pix, val, val.clone())
    # a well-formed call on host tensors gets as far as the device check
    with pytest.raises(NotImplementedError):
        RasterizedEdgeGradient.bwd(tri, vtx, torch.eye(4), pix, val, val.clone())


# the pyo3 names the reference registers (del-msh-dlpack/src/*.rs add_functions)
# on the hot path and the LBVH-machinery modules this build covers
NATIVE_NAMES = [
    "get_cuda_driver_version", "get_ptx_compiler_version",
    "mortons_vtx2morton_from_vtx2co", "mortons_make_bvhnodes_from_sorted_mortons",
    "array1d_permute", "array1d_argsort", "array1d_has_duplicate_sorted_array",
    "array1d_unique_for_sorted_array", "array1d_unique_jdx2val_jdx2idx",
    "bvhnode2aabb_update_aabb", "pix2tri_by_raycast", "pix2tri_interpolate",
    "pix2tri_interpolate_bwd", "pix2depth_update", "pix2depth_bwd_wrt_vtx2xyz",
    "rasterized_edge_gradient_bwd", "rasterized_edge_gradient_edge_gradient_and_type",
    "rasterized_edge_gradient_smooth_gradient", "rasterized_edge_gradient_interpolate",
    "differentiable_antialias_fwd", "differentiable_antialias_bwd",
    "edge2vtx_contour_for_triangle_mesh", "edge2vtx_from_vtx2vtx",
    "edge2elem_from_edge2vtx_of_tri2vtx_with_vtx2vtx", "vtx2elem_from_uniform_mesh",
    "vtx2vtx_from_uniform_mesh", "vtx2vtx_laplacian_smoothing", "vtx2vtx_multiply_graph_laplacian",
    "quad_oct_tree_bnodes_and_bnode2depth_and_bnode2onode_and_idx2bnode",
    "quad_oct_tree_make_tree_from_binary_radix_tree", "quad_oct_tree_aggregate",
    "offset_array_aggregate", "polyhedron_mesh_elem2center",
    "polyhedron_mesh_bvhnode2aabb_from_bvhnodes", "polyhedron_mesh_search_elem_contain_points",
    "polyhedron_mesh_interpolate_values_at_points",
]


def test_native_module_names_and_host_arrays():
    """The native-level names exist; host arrays are refused (no CPU path in
    this build) before anything is launched; the version queries return the
    values of a build without CUDA."""
    import pytest
    from torch.utils.dlpack import to_dlpack

    import del_msh_dlpack.del_msh_dlpack as N

    missing = [n for n in NATIVE_NAMES if not callable(getattr(N, n, None))]
    assert not missing, missing
    assert issubclass(N.PanicException, BaseException)
    assert not issubclass(N.PanicException, Exception)
    assert N.get_cuda_driver_version() == (0xFFFFFFFF, 0xFFFFFFFF)
    assert N.get_ptx_compiler_version() == (-1, -1, -1)
    keys = torch.arange(8, dtype=torch.int64).to(torch.uint32)
    idx = torch.empty(8, dtype=torch.uint32)
    with pytest.raises(NotImplementedError):
        N.array1d_argsort(to_dlpack(keys), to_dlpack(idx), 0)
    with pytest.raises(NotImplementedError):
        N.array1d_permute(keys, keys, idx, 0)


def test_tuning_keys_accept_and_reject_without_a_gpu():
    """The launch parameters the autotuner and the benchmarks set: valid
    values are kept, invalid ones refused with DMSH_ERR_INVALID (host only)."""
    L = _lib.lib()
    batch = {b"batch.sort_items": ((8, 12, 16), (0, 4, 10, 32)),
             b"batch.cast_waves": ((1, 2, 4), (0, 3, 8)),
             b"batch.sort_seg": ((0, 1), (2, -1)),
             b"batch.stack": ((16, 20, 24, 32), (8, 17)),
             b"batch.treelet_grid": ((1, 1024), (0, 70000))}
    for key, (good, bad) in batch.items():
        old = L.dmsh_batch_get_tuning(key)
        try:
            for v in good:
                assert L.dmsh_batch_set_tuning(key, v) == 0
                assert L.dmsh_batch_get_tuning(key) == v
            for v in bad:
                assert L.dmsh_batch_set_tuning(key, v) != 0
                assert L.dmsh_batch_get_tuning(key) == good[-1]
        finally:
            assert L.dmsh_batch_set_tuning(key, old) == 0
    for key in (b"refit.plan_cache", b"refit.top", b"refit.queue", b"refit.leaves_in_aabb"):
        old = L.dmsh_get_tuning(key)
        assert old in (0, 1)
        assert L.dmsh_set_tuning(key, 1 - old) == 0 and L.dmsh_get_tuning(key) == 1 - old
        assert L.dmsh_set_tuning(key, 2) != 0 and L.dmsh_get_tuning(key) == 1 - old
        assert L.dmsh_set_tuning(key, old) == 0
    for key, top in ((b"refit.climb", 2), (b"refit.compact", 1023), (b"sort.wide_tiles", 2)):
        old = L.dmsh_get_tuning(key)
        try:
            for v in (0, top):
                assert L.dmsh_set_tuning(key, v) == 0 and L.dmsh_get_tuning(key) == v
            assert L.dmsh_set_tuning(key, top + 1) != 0 and L.dmsh_get_tuning(key) == top
            assert L.dmsh_set_tuning(key, -1) != 0
        finally:
            assert L.dmsh_set_tuning(key, old) == 0
    assert L.dmsh_set_tuning(b"refit.no_such_key", 1) != 0
