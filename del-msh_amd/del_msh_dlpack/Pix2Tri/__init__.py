"""Mid-level ray-cast and interpolation (reference: native functions of del-msh-dlpack/src/pix2tri.rs:1-271)."""

from .. import _lib
from ..util_torch import as_tensor


def update_pix2tri(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world, pix2tri,
                   stream_ptr=0, pix2t=None):
    """Fills pix2tri [H, W] (and the hit coefficient pix2t if given) by BVH
    ray-casting; transform_ndc2world is column-major (16 floats)."""
    tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world, pix2tri = map(
        as_tensor, (tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world, pix2tri))
    dev = _lib.require_gpu(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world, pix2tri)
    img_h, img_w = pix2tri.shape
    num_tri = tri2vtx.shape[0]
    assert bvhnodes.shape[0] == max(2 * num_tri - 1, 0)
    ws = _lib.workspace(_lib.lib().dmsh_pix2tri_workspace_bytes(num_tri, img_w, img_h), dev)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(
        _lib.lib().dmsh_pix2tri(
            stream, _lib.ptr(pix2tri), _lib.ptr(pix2t), num_tri, _lib.ptr(tri2vtx),
            vtx2xyz.shape[0], _lib.ptr(vtx2xyz), img_w, img_h, _lib.ptr(transform_ndc2world),
            _lib.ptr(bvhnodes), _lib.ptr(bvhnode2aabb), _lib.ptr(ws), ws.numel()),
        "dmsh_pix2tri",
    )


def interpolate(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val, stream_ptr=0):
    """pix2val [H, W, num_vdim] = barycentric interpolation of vtx2val at each
    pixel's hit on its triangle's plane (0 on the background);
    transform_ndc2world is column-major (16 floats)."""
    pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val = map(
        as_tensor, (pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val))
    dev = _lib.require_gpu(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val)
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_interpolate(
        stream, _lib.ptr(pix2tri), img_w, img_h, tri2vtx.shape[0], _lib.ptr(tri2vtx),
        vtx2xyz.shape[0], _lib.ptr(vtx2xyz), vtx2val.shape[1], _lib.ptr(vtx2val),
        _lib.ptr(transform_ndc2world), _lib.ptr(pix2val)), "dmsh_interpolate")


def interpolate_bwd(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val,
                    dldw_vtx2xyz, dldw_vtx2val, stream_ptr=0):
    """Accumulates dL/d(vertex positions) and dL/d(vertex values) of interpolate."""
    args = map(as_tensor, (pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val,
                           dldw_vtx2xyz, dldw_vtx2val))
    pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val, dxyz, dval = args
    dev = _lib.require_gpu(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val,
                           dxyz, dval)
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_interpolate_bwd(
        stream, _lib.ptr(pix2tri), img_w, img_h, tri2vtx.shape[0], _lib.ptr(tri2vtx),
        vtx2xyz.shape[0], _lib.ptr(vtx2xyz), vtx2val.shape[1], _lib.ptr(vtx2val),
        _lib.ptr(transform_ndc2world), _lib.ptr(dldw_pix2val), _lib.ptr(dxyz), _lib.ptr(dval)),
        "dmsh_interpolate_bwd")
