"""Mid-level depth (reference: del_msh_dlpack/Pix2Depth/__init__.py)."""

from .. import _lib
from ..util_torch import as_tensor


def update(pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world, stream_ptr=0):
    """transform_ndc2world column-major; the inverse is taken on the device."""
    pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world = map(
        as_tensor, (pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world))
    dev = _lib.require_gpu(pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world)
    img_h, img_w = pix2depth.shape
    stream = stream_ptr or _lib.stream_of(dev)
    ws = _lib.workspace(_lib.lib().dmsh_pix2depth_workspace_bytes(), dev)
    _lib.check(
        _lib.lib().dmsh_pix2depth(
            stream, _lib.ptr(pix2depth), _lib.ptr(pix2tri), tri2vtx.shape[0], _lib.ptr(tri2vtx),
            vtx2xyz.shape[0], _lib.ptr(vtx2xyz), img_w, img_h, _lib.ptr(transform_ndc2world),
            _lib.ptr(ws), ws.numel()),
        "dmsh_pix2depth",
    )


def bwd_wrt_vtx2xyz(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world,
                    stream_ptr=0):
    """Accumulates into dldw_vtx2xyz; transform_ndc2world column-major."""
    dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world = map(
        as_tensor, (dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world))
    dev = _lib.require_gpu(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth,
                           transform_ndc2world)
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    ws = _lib.workspace(_lib.lib().dmsh_pix2depth_workspace_bytes(), dev)
    _lib.check(
        _lib.lib().dmsh_pix2depth_bwd_wrt_vtx2xyz(
            stream, _lib.ptr(dldw_vtx2xyz), _lib.ptr(pix2tri), tri2vtx.shape[0], _lib.ptr(tri2vtx),
            vtx2xyz.shape[0], _lib.ptr(vtx2xyz), _lib.ptr(dldw_pix2depth), img_w, img_h,
            _lib.ptr(transform_ndc2world), _lib.ptr(ws), ws.numel()),
        "dmsh_pix2depth_bwd_wrt_vtx2xyz",
    )
