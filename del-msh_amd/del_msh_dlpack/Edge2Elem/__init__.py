"""Mid-level edge-to-triangle adjacency (reference: native function
edge2elem_from_edge2vtx_of_tri2vtx_with_vtx2vtx at
del-msh-dlpack/src/edge2elem.rs:17-83)."""

from .. import _lib
from ..util_torch import as_tensor


def from_edge2vtx_of_tri2vtx_with_vtx2vtx(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri,
                                          stream_ptr=0):
    """Writes the two triangles of each edge into edge2tri (in place);
    vtx2idx_offset / idx2vtx is the vertex-to-triangle adjacency."""
    edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri = map(
        as_tensor, (edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri))
    dev = _lib.require_gpu(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_edge2elem_from_edge2vtx_of_tri2vtx(
        stream, edge2vtx.shape[0], _lib.ptr(edge2vtx), tri2vtx.shape[0], _lib.ptr(tri2vtx),
        vtx2idx_offset.shape[0] - 1, _lib.ptr(vtx2idx_offset), idx2vtx.shape[0],
        _lib.ptr(idx2vtx), _lib.ptr(edge2tri)), "dmsh_edge2elem_from_edge2vtx_of_tri2vtx")
