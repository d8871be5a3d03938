"""Edge2Elem (reference: del_msh_dlpack/Edge2Elem/torch.py:5-52)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def from_edge2vtx_of_tri2vtx_with_vtx2vtx(edge2vtx: torch.Tensor, tri2vtx: torch.Tensor,
                                          vtx2idx_offset: torch.Tensor, idx2vtx: torch.Tensor,
                                          edge2tri: torch.Tensor):
    """The (at most two) triangles sharing each edge, found through the
    vertex-to-triangle adjacency (vtx2idx_offset, idx2vtx) of the edge's
    first vertex (del-msh-cpu/src/edge2elem.rs:20-70).  edge2tri is written
    in place; an edge with fewer than two triangles keeps its other entries.
    """
    num_edge = edge2vtx.shape[0]
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2idx_offset.shape[0] - 1
    num_idx = idx2vtx.shape[0]
    device = edge2vtx.device
    assert_shape_dtype_device(edge2vtx, (num_edge, 2), torch.uint32, device)
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2idx_offset, (num_vtx + 1,), torch.uint32, device)
    assert_shape_dtype_device(idx2vtx, (num_idx,), torch.uint32, device)
    assert_shape_dtype_device(edge2tri, (num_edge, 2), torch.uint32, device)
    _lib.require_gpu(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri)
    from . import from_edge2vtx_of_tri2vtx_with_vtx2vtx as native

    native(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri, _lib.stream_of(device))
