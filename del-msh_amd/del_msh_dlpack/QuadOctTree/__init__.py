"""Mid-level quad/oct-tree construction and aggregation (reference:
del-msh-dlpack/src/quad_oct_tree.rs)."""

from .. import _lib
from ..util_torch import as_tensor


def bnodes_and_bnode2depth_and_bnode2onode(idx2morton, num_dim, max_depth, bnodes, bnode2depth,
                                           bnode2onode, idx2bnode, stream_ptr=0):
    """Binary radix tree over sorted unique codes, branch depths, onode
    numbering and idx -> branch (idx2bnode is filled here)."""
    idx2morton = as_tensor(idx2morton)
    outs = [as_tensor(t) for t in (bnodes, bnode2depth, bnode2onode, idx2bnode)]
    dev = _lib.require_gpu(idx2morton, *outs)
    n = idx2morton.shape[0]
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_octree_bnodes_workspace_bytes(n), dev)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(L.dmsh_octree_bnodes(stream, n, _lib.ptr(idx2morton), num_dim, max_depth,
                                    *[_lib.ptr(t) for t in outs], _lib.ptr(ws), ws.numel()),
               "dmsh_octree_bnodes")


def make_tree_from_binary_radix_tree(bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton,
                                     num_dim, max_depth, onodes, onode2depth, onode2center,
                                     idx2onode, idx2center, stream_ptr=0):
    ins = [as_tensor(t) for t in (bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton)]
    outs = [as_tensor(t) for t in (onodes, onode2depth, onode2center, idx2onode, idx2center)]
    dev = _lib.require_gpu(*ins, *outs)
    n = ins[4].shape[0]
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_octree_make_tree(stream, n, *[_lib.ptr(t) for t in ins], num_dim,
                                                max_depth, outs[0].shape[0],
                                                *[_lib.ptr(t) for t in outs]),
               "dmsh_octree_make_tree")


def aggregate(idx2val, idx2onode, onodes, onode2aggval, stream_ptr=0):
    idx2val, idx2onode = as_tensor(idx2val), as_tensor(idx2onode)
    onodes, onode2aggval = as_tensor(onodes), as_tensor(onode2aggval)
    dev = _lib.require_gpu(idx2val, idx2onode, onodes, onode2aggval)
    num_onode, num_link = onodes.shape
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_octree_aggregate_workspace_bytes(num_onode), dev)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(L.dmsh_octree_aggregate(stream, idx2val.shape[0], idx2val.shape[1],
                                       _lib.ptr(idx2val), _lib.ptr(idx2onode), num_link, num_onode,
                                       _lib.ptr(onodes), _lib.ptr(onode2aggval), _lib.ptr(ws),
                                       ws.numel()), "dmsh_octree_aggregate")
