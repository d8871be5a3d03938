"""Mid-level edge-gradient backward (reference:
del_msh_dlpack/RasterizedEdgeGradient/__init__.py:1-22)."""

from .. import _lib
from ..util_torch import as_tensor


def bwd(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val,
        stream_ptr=0):
    """Accumulates into dldw_vtx2xyz; transform_world2pix is 16 floats column-major."""
    tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val = map(
        as_tensor,
        (tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val))
    dev = _lib.require_gpu(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val,
                           dldw_pix2val)
    img_h, img_w = pix2tri.shape
    num_vdim = pix2val.shape[2]
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(
        _lib.lib().dmsh_rasterized_edge_gradient_bwd(
            stream, _lib.ptr(tri2vtx), tri2vtx.shape[0], _lib.ptr(vtx2xyz), vtx2xyz.shape[0],
            _lib.ptr(dldw_vtx2xyz), _lib.ptr(transform_world2pix), img_w, img_h,
            _lib.ptr(pix2tri), num_vdim, _lib.ptr(pix2val), _lib.ptr(dldw_pix2val)),
        "dmsh_rasterized_edge_gradient_bwd",
    )
