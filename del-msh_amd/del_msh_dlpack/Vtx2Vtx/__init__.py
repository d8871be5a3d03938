"""Mid-level vertex-to-vertex adjacency and graph Laplacian (reference:
native functions of del-msh-dlpack/src/vtx2vtx.rs:23-248)."""

import torch as _torch

from .. import _lib
from ..util_torch import as_tensor


def from_uniform_mesh(elem2vtx, num_vtx: int, is_self: bool, stream_ptr=0):
    """Returns (vtx2idx [num_vtx+1], idx2vtx) uint32 device tensors: the
    vertices around each vertex in first-appearance order.  One
    device-to-host read of the entry count sizes idx2vtx, as the reference's
    `vtx2idx.last()` does."""
    elem2vtx = as_tensor(elem2vtx)
    dev = _lib.require_gpu(elem2vtx)
    num_elem, num_node = elem2vtx.shape
    stream = stream_ptr or _lib.stream_of(dev)
    L = _lib.lib()
    s = int(bool(is_self))
    vtx2idx = _torch.empty((num_vtx + 1,), dtype=_torch.uint32, device=dev)
    ws = _lib.workspace(L.dmsh_vtx2vtx_workspace_bytes(num_elem, num_node, num_vtx, s), dev)
    _lib.check(L.dmsh_vtx2vtx_count(stream, num_elem, num_node, _lib.ptr(elem2vtx), num_vtx, s,
                                    _lib.ptr(vtx2idx), _lib.ptr(ws), ws.numel()),
               "dmsh_vtx2vtx_count")
    num_idx = int(vtx2idx[num_vtx:].view(_torch.int32).cpu()[0])
    idx2vtx = _torch.empty((num_idx,), dtype=_torch.uint32, device=dev)
    if num_idx > 0:
        _lib.check(L.dmsh_vtx2vtx_fill(stream, num_elem, num_node, num_vtx, s, _lib.ptr(vtx2idx),
                                       _lib.ptr(idx2vtx), _lib.ptr(ws), ws.numel()),
                   "dmsh_vtx2vtx_fill")
    return vtx2idx, idx2vtx


def laplacian_smoothing(vtx2idx, idx2vtx, lambda0: float, vtx2lhs, vtx2rhs, num_iter: int,
                        vtx2lhstmp, stream_ptr=0):
    """num_iter Jacobi sweeps of [I + lambda L] lhs = rhs, in place on vtx2lhs."""
    vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp = map(
        as_tensor, (vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp))
    dev = _lib.require_gpu(vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_laplacian_smoothing(
        stream, vtx2idx.shape[0] - 1, _lib.ptr(vtx2idx), idx2vtx.shape[0], _lib.ptr(idx2vtx),
        float(lambda0), vtx2lhs.shape[1], _lib.ptr(vtx2lhs), _lib.ptr(vtx2rhs), int(num_iter),
        _lib.ptr(vtx2lhstmp)), "dmsh_laplacian_smoothing")


def multiply_graph_laplacian(vtx2idx, idx2vtx, vtx2rhs, vtx2lhs, stream_ptr=0):
    """vtx2lhs = L vtx2rhs, L = [.., -1, .., valence, .., -1, ..]."""
    vtx2idx, idx2vtx, vtx2rhs, vtx2lhs = map(as_tensor, (vtx2idx, idx2vtx, vtx2rhs, vtx2lhs))
    dev = _lib.require_gpu(vtx2idx, idx2vtx, vtx2rhs, vtx2lhs)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_multiply_graph_laplacian(
        stream, vtx2idx.shape[0] - 1, _lib.ptr(vtx2idx), idx2vtx.shape[0], _lib.ptr(idx2vtx),
        vtx2rhs.shape[1], _lib.ptr(vtx2rhs), _lib.ptr(vtx2lhs)), "dmsh_multiply_graph_laplacian")
