"""Mid-level contour extraction (reference: del_msh_dlpack/Edge2Vtx/__init__.py,
native function edge2vtx_contour_for_triangle_mesh at
del-msh-dlpack/src/edge2vtx.rs:19-115)."""

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
