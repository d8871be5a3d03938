"""Mid-level edge-gradient functions (reference: native functions of
del-msh-dlpack/src/rasterized_edge_gradient.rs:1-456)."""

from .. import _lib
from ..util_torch import as_tensor


def bwd(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val,
        stream_ptr=0, ordered=True):
    """Accumulates into dldw_vtx2xyz; transform_world2pix is 16 floats column-major.

    ordered=True (default): every vertex adds its terms in del-msh-cpu's loop
    order, so the result is that loop's, bit for bit, on every run.
    ordered=False: one launch with float atomics (faster, order not fixed)."""
    tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val = map(
        as_tensor,
        (tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val))
    dev = _lib.require_gpu(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val,
                           dldw_pix2val)
    _lib.check_transform(transform_world2pix, "transform_world2pix")
    img_h, img_w = pix2tri.shape
    num_vdim = pix2val.shape[2]
    stream = stream_ptr or _lib.stream_of(dev)
    L = _lib.lib()
    args = (stream, _lib.ptr(tri2vtx), tri2vtx.shape[0], _lib.ptr(vtx2xyz), vtx2xyz.shape[0],
            _lib.ptr(dldw_vtx2xyz), _lib.ptr(transform_world2pix), img_w, img_h,
            _lib.ptr(pix2tri), num_vdim, _lib.ptr(pix2val), _lib.ptr(dldw_pix2val))
    if not ordered:
        _lib.check(L.dmsh_rasterized_edge_gradient_bwd(*args), "dmsh_rasterized_edge_gradient_bwd")
        return
    ws = _lib.workspace(
        L.dmsh_rasterized_edge_gradient_bwd_workspace_bytes(img_w, img_h, vtx2xyz.shape[0]), dev)
    _lib.check(L.dmsh_rasterized_edge_gradient_bwd_ordered(*args, _lib.ptr(ws), ws.numel()),
               "dmsh_rasterized_edge_gradient_bwd_ordered")


def edge_gradient_and_type(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val,
                           hedge2type, hedge2dldr, vedge2type, vedge2dldr, stream_ptr=0):
    """Fills the pixel-edge types and gradients (in place);
    transform_world2pix is 16 floats column-major."""
    args = tuple(map(as_tensor, (tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val,
                                 dldw_pix2val, hedge2type, hedge2dldr, vedge2type, vedge2dldr)))
    dev = _lib.require_gpu(*args)
    tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val, ht, hd, vt, vd = args
    img_h, img_w = pix2tri.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_edge_gradient_and_type(
        stream, _lib.ptr(tri2vtx), tri2vtx.shape[0], _lib.ptr(vtx2xyz), vtx2xyz.shape[0],
        _lib.ptr(transform_world2pix), img_w, img_h, _lib.ptr(pix2tri), pix2val.shape[2],
        _lib.ptr(pix2val), _lib.ptr(dldw_pix2val), _lib.ptr(ht), _lib.ptr(hd), _lib.ptr(vt),
        _lib.ptr(vd)), "dmsh_edge_gradient_and_type")


def smooth_gradient(hedge2type, hedge2dldr, vedge2type, vedge2dldr, num_iter, stream_ptr=0):
    """num_iter red-black sweeps of both pixel-edge grids, in place."""
    args = tuple(map(as_tensor, (hedge2type, hedge2dldr, vedge2type, vedge2dldr)))
    dev = _lib.require_gpu(*args)
    ht, hd, vt, vd = args
    img_h, img_w = ht.shape[0] + 1, ht.shape[1]
    stream = stream_ptr or _lib.stream_of(dev)
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_smooth_gradient_workspace_bytes(img_w, img_h), dev)
    _lib.check(L.dmsh_smooth_gradient(stream, img_w, img_h, _lib.ptr(ht), _lib.ptr(vt), int(num_iter),
                                      _lib.ptr(hd), _lib.ptr(vd), _lib.ptr(ws), ws.numel()),
               "dmsh_smooth_gradient")


def interpolate(hedge2vy, vedge2vx, vtx2xy, vtx2velo, stream_ptr=0):
    """vtx2velo [num_vtx, 2] = (x velocity from the vedges, y velocity from
    the hedges) bilinear at each point."""
    args = tuple(map(as_tensor, (hedge2vy, vedge2vx, vtx2xy, vtx2velo)))
    dev = _lib.require_gpu(*args)
    hv, vv, xy, out = args
    img_h, img_w = hv.shape[0] + 1, hv.shape[1]
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_interpolate_staggered_grid(
        stream, img_w, img_h, _lib.ptr(hv), _lib.ptr(vv), xy.shape[0], _lib.ptr(xy), _lib.ptr(out)),
        "dmsh_interpolate_staggered_grid")


def smoothed_vertex_gradient(hedge2dldr, vedge2dldr, vtx2xyz, transform_world2pix_colmajor,
                             dldw_vtx2xyz, stream_ptr=0):
    """dL/dx per vertex from the smoothed pixel-edge grids, through the
    perspective projection (the tail of the smoothed backward in one launch)."""
    args = tuple(map(as_tensor, (hedge2dldr, vedge2dldr, vtx2xyz, transform_world2pix_colmajor,
                                 dldw_vtx2xyz)))
    dev = _lib.require_gpu(*args)
    hv, vv, xyz, t, out = args
    img_h, img_w = hv.shape[0] + 1, hv.shape[1]
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_smoothed_vertex_gradient(
        stream, img_w, img_h, _lib.ptr(hv), _lib.ptr(vv), xyz.shape[0], _lib.ptr(xyz), _lib.ptr(t),
        _lib.ptr(out)), "dmsh_smoothed_vertex_gradient")
