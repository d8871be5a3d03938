"""Mid-level vertex-to-element adjacency (reference: native function
vtx2elem_from_uniform_mesh at del-msh-dlpack/src/vtx2elem.rs:14-66)."""

import torch as _torch

from .. import _lib
from ..util_torch import as_tensor


def from_uniform_mesh(elem2vtx, num_vtx: int, stream_ptr=0):
    """Returns (vtx2idx [num_vtx+1], idx2elem [num_elem*num_node]) uint32
    device tensors: the elements around each vertex, ascending."""
    elem2vtx = as_tensor(elem2vtx)
    dev = _lib.require_gpu(elem2vtx)
    num_elem, num_node = elem2vtx.shape
    stream = stream_ptr or _lib.stream_of(dev)
    L = _lib.lib()
    vtx2idx = _torch.empty((num_vtx + 1,), dtype=_torch.uint32, device=dev)
    idx2elem = _torch.empty((num_elem * num_node,), dtype=_torch.uint32, device=dev)
    ws = _lib.workspace(L.dmsh_vtx2elem_workspace_bytes(num_elem, num_node, num_vtx), dev)
    _lib.check(L.dmsh_vtx2elem_from_uniform_mesh(stream, num_elem, num_node, _lib.ptr(elem2vtx),
                                                 num_vtx, _lib.ptr(vtx2idx), _lib.ptr(idx2elem),
                                                 _lib.ptr(ws), ws.numel()),
               "dmsh_vtx2elem_from_uniform_mesh")
    return vtx2idx, idx2elem
