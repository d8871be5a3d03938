"""Mid-level edge functions (reference: native functions
edge2vtx_contour_for_triangle_mesh and edge2vtx_from_vtx2vtx at
del-msh-dlpack/src/edge2vtx.rs:19-115 and 117-174)."""

import torch as _torch

from .. import _lib
from ..util_torch import as_tensor


def contour_for_triangle_mesh(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri,
                              stream_ptr=0):
    """transform_world2ndc is 16 floats column-major.  Returns a new
    (num_contour, 2) uint32 tensor, in edge order (one device-to-host read of
    the count, as the reference's compaction does)."""
    tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri = map(
        as_tensor, (tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri))
    dev = _lib.require_gpu(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri)
    num_edge = edge2vtx.shape[0]
    stream = stream_ptr or _lib.stream_of(dev)
    cedge2vtx = _torch.empty((max(num_edge, 1), 2), dtype=_torch.uint32, device=dev)
    count = _torch.empty((1,), dtype=_torch.uint32, device=dev)
    ws = _lib.workspace(_lib.lib().dmsh_contour_workspace_bytes(num_edge), dev)
    _lib.check(
        _lib.lib().dmsh_contour_for_triangle_mesh(
            stream, tri2vtx.shape[0], _lib.ptr(tri2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz),
            _lib.ptr(transform_world2ndc), num_edge, _lib.ptr(edge2vtx), _lib.ptr(edge2tri),
            _lib.ptr(cedge2vtx), _lib.ptr(count), _lib.ptr(ws), ws.numel()),
        "dmsh_contour_for_triangle_mesh",
    )
    n = int(count.view(_torch.int32).cpu()[0])
    return cedge2vtx[:n].clone()


def contour_for_triangle_mesh_into(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri,
                                   cedge2vtx, count, ws, stream_ptr=0):
    """As contour_for_triangle_mesh, without the read-back: the contour edges
    go to the front of cedge2vtx (at least (num_edge, 2) uint32) and their
    number to count (one uint32 on the device); ws is a workspace of at least
    dmsh_contour_workspace_bytes(num_edge) bytes.  Nothing synchronizes, so a
    caller can enqueue many meshes and read all the counts at once."""
    args = tuple(map(as_tensor, (tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri,
                                 cedge2vtx, count, ws)))
    tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri, cedge2vtx, count, ws = args
    dev = _lib.require_gpu(*args)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(
        _lib.lib().dmsh_contour_for_triangle_mesh(
            stream, tri2vtx.shape[0], _lib.ptr(tri2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz),
            _lib.ptr(transform_world2ndc), edge2vtx.shape[0], _lib.ptr(edge2vtx),
            _lib.ptr(edge2tri), _lib.ptr(cedge2vtx), _lib.ptr(count), _lib.ptr(ws), ws.numel()),
        "dmsh_contour_for_triangle_mesh",
    )


def from_vtx2vtx(vtx2idx_offset, idx2vtx, edge2vtx, stream_ptr=0):
    """edge2vtx[idx] = (i, idx2vtx[idx]) for every entry idx of vertex i (in place)."""
    vtx2idx_offset, idx2vtx, edge2vtx = map(as_tensor, (vtx2idx_offset, idx2vtx, edge2vtx))
    dev = _lib.require_gpu(vtx2idx_offset, idx2vtx, edge2vtx)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_edge2vtx_from_vtx2vtx(
        stream, vtx2idx_offset.shape[0] - 1, _lib.ptr(vtx2idx_offset), idx2vtx.shape[0],
        _lib.ptr(idx2vtx), _lib.ptr(edge2vtx)), "dmsh_edge2vtx_from_vtx2vtx")
