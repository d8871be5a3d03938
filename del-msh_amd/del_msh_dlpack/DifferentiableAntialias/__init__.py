"""Mid-level anti-aliasing (reference: del_msh_dlpack/DifferentiableAntialias/__init__.py;
native functions at del-msh-dlpack/src/differentiable_antialias.rs:16-182)."""

from .. import _lib
from ..util_torch import as_tensor


def antialias(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin, pix2vout, stream_ptr=0):
    """Writes the blended silhouette pixels into pix2vout (initialised by the
    caller); transform_world2pix is 16 floats column-major."""
    cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin, pix2vout = map(
        as_tensor, (cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin, pix2vout))
    dev = _lib.require_gpu(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin, pix2vout)
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    ws = _lib.workspace(_lib.lib().dmsh_antialias_workspace_bytes(img_w, img_h), dev)
    _lib.check(
        _lib.lib().dmsh_antialias_fwd(
            stream, cedge2vtx.shape[0], _lib.ptr(cedge2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz),
            _lib.ptr(transform_world2pix), img_w, img_h, _lib.ptr(pix2tri), _lib.ptr(pix2vin),
            _lib.ptr(pix2vout), _lib.ptr(ws), ws.numel()),
        "dmsh_antialias_fwd",
    )


def bwd_antialias(cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val, dldw_pixval,
                  pix2tri, stream_ptr=0):
    """Accumulates into dldw_vtx2xyz; transform_world2pix is 16 floats column-major."""
    cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val, dldw_pixval, pix2tri = map(
        as_tensor,
        (cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val, dldw_pixval, pix2tri))
    dev = _lib.require_gpu(cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val,
                           dldw_pixval, pix2tri)
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(
        _lib.lib().dmsh_antialias_bwd(
            stream, cedge2vtx.shape[0], _lib.ptr(cedge2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz),
            _lib.ptr(dldw_vtx2xyz), _lib.ptr(transform_world2pix), img_w, img_h,
            _lib.ptr(pix2val), _lib.ptr(dldw_pixval), _lib.ptr(pix2tri)),
        "dmsh_antialias_bwd",
    )
