"""Vtx2Elem (reference: del_msh_dlpack/Vtx2Elem/torch.py:6-24)."""

import torch

from .. import _lib


def from_uniform_mesh(elem2vtx: torch.Tensor, num_vtx: int):
    """Elements around each vertex of a uniform mesh (elements ascending per
    vertex, as del-msh-cpu/src/vtx2elem.rs:58-96).

    Args:
        elem2vtx: (num_elem, num_node) uint32
        num_vtx: number of vertices
    Returns:
        vtx2idx: (num_vtx + 1,) uint32 offsets; idx2elem: (num_elem * num_node,) uint32
    """
    assert len(elem2vtx.shape) == 2
    assert elem2vtx.dtype == torch.uint32
    assert elem2vtx.is_contiguous()
    _lib.require_gpu(elem2vtx)
    from . import from_uniform_mesh as native

    return native(elem2vtx, num_vtx, _lib.stream_of(elem2vtx.device))
