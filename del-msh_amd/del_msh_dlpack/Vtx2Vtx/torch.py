"""Vtx2Vtx (reference: del_msh_dlpack/Vtx2Vtx/torch.py:6-131)."""

import torch

from .. import _lib


def from_uniform_mesh(elem2vtx: torch.Tensor, num_vtx: int, is_self: bool):
    """Vertices around each vertex of a uniform mesh (first-appearance order
    over the vertex's elements and their nodes, del-msh-cpu/src/vtx2vtx.rs:
    11-76); is_self keeps the vertex itself."""
    assert len(elem2vtx.shape) == 2
    assert elem2vtx.dtype == torch.uint32
    _lib.require_gpu(elem2vtx)
    from . import from_uniform_mesh as native

    return native(elem2vtx, num_vtx, is_self, _lib.stream_of(elem2vtx.device))


def laplacian_smoothing(vtx2idx: torch.Tensor, idx2vtx: torch.Tensor, lambda0: float,
                        vtx2lhs: torch.Tensor, vtx2rhs: torch.Tensor, num_iter: int,
                        vtx2lhstmp: torch.Tensor | None):
    """Solve [I + lambda * L] {vtx2lhs} = {vtx2rhs} with num_iter Jacobi
    sweep pairs, L = [.., -1, .., valence, .., -1, ..] (in place on vtx2lhs)."""
    device = vtx2idx.device
    num_vtx = vtx2idx.shape[0] - 1
    if vtx2lhstmp is None:
        vtx2lhstmp = torch.zeros_like(vtx2lhs)
    assert len(vtx2idx.shape) == 1 and len(idx2vtx.shape) == 1
    assert vtx2idx.dtype == torch.uint32 and idx2vtx.dtype == torch.uint32
    assert len(vtx2lhs.shape) == 2 and vtx2lhs.shape[0] == num_vtx
    assert vtx2lhs.shape == vtx2rhs.shape == vtx2lhstmp.shape
    assert vtx2lhs.dtype == vtx2rhs.dtype == vtx2lhstmp.dtype == torch.float32
    assert num_iter >= 0
    for t in (idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp):
        assert t.device == device
    _lib.require_gpu(vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp)
    from . import laplacian_smoothing as native

    native(vtx2idx, idx2vtx, lambda0, vtx2lhs, vtx2rhs, num_iter, vtx2lhstmp,
           _lib.stream_of(device))


def multiply_graph_laplacian(vtx2idx: torch.Tensor, idx2vtx: torch.Tensor,
                             vtx2rhs: torch.Tensor) -> torch.Tensor:
    num_vtx = vtx2idx.shape[0] - 1
    device = vtx2idx.device
    assert vtx2idx.shape == (num_vtx + 1,) and vtx2idx.dtype == torch.uint32
    assert idx2vtx.dtype == torch.uint32 and idx2vtx.device == device
    assert vtx2rhs.shape[0] == num_vtx and vtx2rhs.dtype == torch.float32
    assert vtx2rhs.device == device
    vtx2lhs = torch.zeros_like(vtx2rhs)
    _lib.require_gpu(vtx2idx, idx2vtx, vtx2rhs)
    from . import multiply_graph_laplacian as native

    native(vtx2idx, idx2vtx, vtx2rhs.contiguous(), vtx2lhs, _lib.stream_of(device))
    return vtx2lhs


class GraphLaplacian(torch.autograd.Function):
    """y = L x; L is symmetric, so dL/dx = L dL/dy."""

    @staticmethod
    def forward(ctx, vtx2idx, idx2vtx, x):
        ctx.save_for_backward(vtx2idx, idx2vtx)
        return multiply_graph_laplacian(vtx2idx.detach(), idx2vtx.detach(), x.detach())

    @staticmethod
    def backward(ctx, dw_lx):
        vtx2idx, idx2vtx = ctx.saved_tensors
        dw_x = multiply_graph_laplacian(vtx2idx.detach(), idx2vtx.detach(),
                                        dw_lx.detach().contiguous())
        return None, None, dw_x
