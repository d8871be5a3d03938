"""TriMesh3 (reference: del_msh_dlpack/TriMesh3/torch.py)."""

import numpy as np
import torch

from .. import _lib, primitives
from ..util_torch import assert_shape_dtype_device


def make_tri2centroid(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor) -> torch.Tensor:
    """Centroids of triangles, (a + b + c) / 3 (reference TriMesh3/torch.py:8-24)."""
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    idx = tri2vtx.long()
    return (vtx2xyz[idx[:, 0]] + vtx2xyz[idx[:, 1]] + vtx2xyz[idx[:, 2]]) / 3.0


def torus(major_raidus: float, minor_radius: float, ndiv_major: int, ndiv_minor: int):
    """Torus mesh on the host (CPU tensors), as the reference returns it."""
    tri, vtx = primitives.torus_zup(major_raidus, minor_radius, ndiv_major, ndiv_minor)
    return torch.from_numpy(tri), torch.from_numpy(vtx)


def sphere(raidus: float, ndiv_longtitude: int, ndiv_latitude: int):
    """UV sphere with the pole on +y, on the host (CPU tensors)."""
    tri, vtx = primitives.sphere_yup(raidus, ndiv_longtitude, ndiv_latitude)
    return torch.from_numpy(tri), torch.from_numpy(vtx)


def make_bvhnodes_bvhnode2aabb(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                               morton_bits: int = 30):
    """LBVH of a triangle mesh: centroids -> unit-cube fit -> Morton codes ->
    stable argsort -> Karras topology -> refit, all on the device in one call.

    morton_bits: 30 (uint32 codes, as the reference) or 63 (uint64 extension).
    Returns:
        bvhnodes: (2*num_tri-1, 3) uint32
        bvhnode2aabb: (2*num_tri-1, 6) float32
    """
    vtx2xyz = vtx2xyz.detach()
    num_vtx = vtx2xyz.shape[0]
    num_tri = tri2vtx.shape[0]
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz)
    num_node = max(2 * num_tri - 1, 0)
    bvhnodes = torch.empty((num_node, 3), dtype=torch.uint32, device=device)
    bvhnode2aabb = torch.empty((num_node, 6), dtype=torch.float32, device=device)
    ws = _lib.workspace(_lib.lib().dmsh_build_workspace_bytes(num_tri, morton_bits), device)
    _lib.check(
        _lib.lib().dmsh_build_trimesh3(
            _lib.stream_of(device), num_tri, _lib.ptr(tri2vtx), num_vtx, _lib.ptr(vtx2xyz),
            morton_bits, _lib.ptr(bvhnodes), _lib.ptr(bvhnode2aabb), _lib.ptr(ws), ws.numel()),
        "dmsh_build_trimesh3",
    )
    return bvhnodes, bvhnode2aabb


def make_bvhnodes_bvhnode2aabb_for_points(vtx2xyz: torch.Tensor, morton_bits: int = 30):
    """LBVH over a point set (leaves index points), del-msh-cpu
    bvhnodes_morton.rs:130-149 + bvhnode2aabb3.rs:6-63."""
    vtx2xyz = vtx2xyz.detach()
    num_pnt = vtx2xyz.shape[0]
    device = vtx2xyz.device
    assert_shape_dtype_device(vtx2xyz, (num_pnt, 3), torch.float32, device)
    _lib.require_gpu(vtx2xyz)
    num_node = max(2 * num_pnt - 1, 0)
    bvhnodes = torch.empty((num_node, 3), dtype=torch.uint32, device=device)
    bvhnode2aabb = torch.empty((num_node, 6), dtype=torch.float32, device=device)
    ws = _lib.workspace(_lib.lib().dmsh_build_workspace_bytes(num_pnt, morton_bits), device)
    _lib.check(
        _lib.lib().dmsh_build_points(
            _lib.stream_of(device), num_pnt, _lib.ptr(vtx2xyz), morton_bits,
            _lib.ptr(bvhnodes), _lib.ptr(bvhnode2aabb), _lib.ptr(ws), ws.numel()),
        "dmsh_build_points",
    )
    return bvhnodes, bvhnode2aabb


def make_bvhnode2aabb_from_bvhnodes(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                                    bvhnodes: torch.Tensor) -> torch.Tensor:
    """Refit only (the topology is kept)."""
    num_node = bvhnodes.shape[0]
    bvhnode2aabb = torch.empty((num_node, 6), dtype=torch.float32, device=bvhnodes.device)
    vtx2xyz1 = torch.zeros((0, 3), dtype=torch.float32, device=bvhnodes.device)
    from . import make_bvhnode2aabb_from_bvhnodes as native

    native(tri2vtx, vtx2xyz.detach(), vtx2xyz1, bvhnodes, bvhnode2aabb,
           _lib.stream_of(bvhnodes.device))
    return bvhnode2aabb


def make_edge2vtx(tri2vtx: torch.Tensor, num_vtx: int) -> torch.Tensor:
    """Edges of a triangle mesh from its vertex-to-vertex adjacency
    (TriMesh3/torch.py:288-302 in the reference): each mesh edge appears once
    from each end, (i, j) and (j, i)."""
    assert tri2vtx.ndim == 2 and tri2vtx.shape[1] == 3
    assert tri2vtx.is_contiguous()
    assert tri2vtx.dtype == torch.uint32
    from ..Vtx2Vtx.torch import from_uniform_mesh
    from ..Edge2Vtx.torch import from_vtx2vtx

    vtx2idx_offset, idx2vtx = from_uniform_mesh(tri2vtx, num_vtx, False)
    edge2vtx = torch.empty((idx2vtx.shape[0], 2), dtype=torch.uint32, device=tri2vtx.device)
    from_vtx2vtx(vtx2idx_offset, idx2vtx, edge2vtx)
    return edge2vtx


def make_edge2tri(tri2vtx: torch.Tensor, num_vtx: int, edge2vtx: torch.Tensor) -> torch.Tensor:
    """The two triangles of each edge (TriMesh3/torch.py:305-325 in the
    reference).  Like the reference, the output starts uninitialised: an edge
    with fewer than two triangles leaves the missing entries as they were."""
    assert tri2vtx.ndim == 2 and tri2vtx.shape[1] == 3
    assert tri2vtx.is_contiguous()
    assert tri2vtx.dtype == torch.uint32
    assert edge2vtx.ndim == 2 and edge2vtx.shape[1] == 2
    assert edge2vtx.is_contiguous()
    assert edge2vtx.dtype == torch.uint32
    assert tri2vtx.device == edge2vtx.device
    from ..Vtx2Elem.torch import from_uniform_mesh
    from ..Edge2Elem.torch import from_edge2vtx_of_tri2vtx_with_vtx2vtx

    vtx2jdx_offset, jdx2tri = from_uniform_mesh(tri2vtx, num_vtx)
    edge2tri = torch.empty((edge2vtx.shape[0], 2), dtype=torch.uint32, device=tri2vtx.device)
    from_edge2vtx_of_tri2vtx_with_vtx2vtx(edge2vtx, tri2vtx, vtx2jdx_offset, jdx2tri, edge2tri)
    return edge2tri
