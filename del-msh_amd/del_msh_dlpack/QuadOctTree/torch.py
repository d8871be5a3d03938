"""Quad/oct-tree over sorted unique Morton codes (reference:
del_msh_dlpack/QuadOctTree/torch.py:1-153; results follow
del-msh-cpu/src/quad_oct_tree.rs)."""

import torch

from .. import _lib


def bnodes_and_bnode2depth_and_bnode2onode_and_idx2bnode(idx2morton: torch.Tensor, num_dim,
                                                         max_depth: int | None = None):
    if max_depth is None:
        max_depth = 32 // num_dim
    num_idx = idx2morton.shape[0]
    num_bnode = num_idx - 1
    device = idx2morton.device
    assert idx2morton.dtype == torch.uint32
    dev = _lib.require_gpu(idx2morton)
    bnodes = torch.empty((num_bnode, 3), device=device, dtype=torch.uint32)
    bnode2depth = torch.empty((num_bnode,), device=device, dtype=torch.uint32)
    bnode2onode = torch.empty((num_bnode,), device=device, dtype=torch.uint32)
    idx2bnode = torch.empty((num_idx,), device=device, dtype=torch.uint32)
    from . import bnodes_and_bnode2depth_and_bnode2onode

    bnodes_and_bnode2depth_and_bnode2onode(idx2morton, num_dim, max_depth, bnodes, bnode2depth,
                                           bnode2onode, idx2bnode, _lib.stream_of(dev))
    return bnodes, bnode2depth, bnode2onode, idx2bnode


def make_tree_from_binary_radix_tree(bnodes: torch.Tensor, bnode2onode: torch.Tensor,
                                     bnode2depth: torch.Tensor, idx2bnode: torch.Tensor,
                                     idx2morton: torch.Tensor, num_dim: int,
                                     max_depth: int | None = None):
    num_bnode = bnodes.shape[0]
    num_idx = num_bnode + 1
    if max_depth is None:
        max_depth = 32 // num_dim
    nlink = 1 + 2 ** num_dim
    num_onode = int(bnode2onode[-1:].view(torch.int32).item()) + 1
    device = bnodes.device
    dev = _lib.require_gpu(bnodes)
    onodes = torch.empty((num_onode, nlink), device=device, dtype=torch.uint32)
    onode2depth = torch.empty((num_onode,), device=device, dtype=torch.uint32)
    onode2center = torch.empty((num_onode, num_dim), device=device, dtype=torch.float32)
    idx2onode = torch.empty((num_idx,), device=device, dtype=torch.uint32)
    idx2center = torch.empty((num_idx, num_dim), device=device, dtype=torch.float32)
    from . import make_tree_from_binary_radix_tree as native

    native(bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton, num_dim, max_depth, onodes,
           onode2depth, onode2center, idx2onode, idx2center, _lib.stream_of(dev))
    return onodes, onode2depth, onode2center, idx2onode, idx2center


class QuadOctTree:
    def __init__(self):
        pass

    def construct_from_idx2morton(self, idx2morton: torch.Tensor, num_dim: int,
                                  is_save_intermediate=False):
        self.num_dim = num_dim
        bnodes, bnode2depth, bnode2onode, idx2bnode = \
            bnodes_and_bnode2depth_and_bnode2onode_and_idx2bnode(idx2morton, num_dim)
        (self.onodes, self.onode2depth, self.onode2center, self.idx2onode,
         self.idx2center) = make_tree_from_binary_radix_tree(bnodes, bnode2onode, bnode2depth,
                                                             idx2bnode, idx2morton, num_dim)
        if is_save_intermediate:
            self.bnodes = bnodes
            self.bnode2depth = bnode2depth
            self.bnode2onode = bnode2onode
            self.idx2bnode = idx2bnode

    def aggregate(self, idx2val) -> torch.Tensor:
        """Per node, the sum of idx2val over the node's points, added in
        ascending point order."""
        num_vdim = idx2val.shape[1]
        num_idx = self.idx2onode.shape[0]
        num_onode = self.onodes.shape[0]
        device = self.onodes.device
        assert idx2val.shape == (num_idx, num_vdim) and idx2val.device == device
        assert idx2val.dtype == torch.float32
        onode2aggval = torch.zeros(size=(num_onode, num_vdim), dtype=torch.float32, device=device)
        from . import aggregate as native

        native(idx2val, self.idx2onode, self.onodes, onode2aggval, _lib.stream_of(device))
        return onode2aggval
