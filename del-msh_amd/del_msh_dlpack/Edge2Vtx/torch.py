"""Edge2Vtx (reference: del_msh_dlpack/Edge2Vtx/torch.py:7-73)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def from_vtx2vtx(vtx2idx_offset: torch.Tensor, idx2vtx: torch.Tensor, edge2vtx: torch.Tensor):
    """One edge (i, j) per entry j of vertex i of a vertex-to-vertex adjacency
    (del-msh-cpu/src/edge2vtx.rs:16-36), written into edge2vtx in place.  A
    full adjacency (Vtx2Vtx.from_uniform_mesh) lists each mesh edge once from
    each end."""
    num_vtx = vtx2idx_offset.shape[0] - 1
    num_idx = idx2vtx.shape[0]
    num_edge = edge2vtx.shape[0]
    device = vtx2idx_offset.device
    assert num_edge == num_idx
    assert_shape_dtype_device(vtx2idx_offset, (num_vtx + 1,), torch.uint32, device)
    assert_shape_dtype_device(idx2vtx, (num_idx,), torch.uint32, device)
    assert_shape_dtype_device(edge2vtx, (num_edge, 2), torch.uint32, device)
    _lib.require_gpu(vtx2idx_offset, idx2vtx, edge2vtx)
    from . import from_vtx2vtx as native

    native(vtx2idx_offset, idx2vtx, edge2vtx, _lib.stream_of(device))


def contour_for_triangle_mesh(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                              transform_world2ndc: torch.Tensor, edge2vtx: torch.Tensor,
                              edge2tri: torch.Tensor) -> torch.Tensor:
    """Edges on the silhouette seen through transform_world2ndc.

    Args:
        tri2vtx: (num_tri, 3) uint32; vtx2xyz: (num_vtx, 3) float32
        transform_world2ndc: (4, 4) float32, row-major
        edge2vtx, edge2tri: (num_edge, 2) uint32 (the two triangles of each edge)
    Returns:
        cedge2vtx: (num_contour, 2) uint32, in edge order
    """
    vtx2xyz = vtx2xyz.detach()
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    num_edge = edge2vtx.shape[0]
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_world2ndc, (4, 4), torch.float32, device)
    assert_shape_dtype_device(edge2vtx, (num_edge, 2), torch.uint32, device)
    assert_shape_dtype_device(edge2tri, (num_edge, 2), torch.uint32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri)
    from . import contour_for_triangle_mesh as native

    return native(tri2vtx, vtx2xyz, transform_world2ndc.T.contiguous().flatten(), edge2vtx,
                  edge2tri, _lib.stream_of(device))
