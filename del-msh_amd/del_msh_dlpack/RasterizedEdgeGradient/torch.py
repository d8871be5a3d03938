"""RasterizedEdgeGradient (reference: del_msh_dlpack/RasterizedEdgeGradient/torch.py:7-274)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def bwd(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor, transform_world2pix: torch.Tensor,
        pix2tri: torch.Tensor, pix2val: torch.Tensor, dldw_pix2val: torch.Tensor,
        ordered: bool = True) -> torch.Tensor:
    """dL/d(vertex positions) from the silhouette edges of the rasterized image.

    Args:
        tri2vtx: (num_tri, 3) uint32; vtx2xyz: (num_vtx, 3) float32
        transform_world2pix: (4, 4) float32, row-major
        pix2tri: (H, W) uint32; pix2val, dldw_pix2val: (H, W, nv) float32
        ordered: True (default) sums every vertex's terms in del-msh-cpu's loop
            order (bit-exact with it, the same bits on every run); False adds
            them with float atomics in one launch (faster, order not fixed)
    Returns:
        dldw_vtx2xyz: (num_vtx, 3) float32
    """
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w, num_vdim = pix2val.shape
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_world2pix, (4, 4), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val)
    dldw_vtx2xyz = torch.zeros_like(vtx2xyz)
    from . import bwd as native

    native(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix.T.contiguous().flatten(),
           pix2tri, pix2val, dldw_pix2val, _lib.stream_of(device), ordered=ordered)
    return dldw_vtx2xyz


class RasterizedEdgeGradientFunction(torch.autograd.Function):
    """Identity in the forward pass; the backward adds the silhouette-edge
    gradient to the vertices and passes dL/d(pixel) through unchanged."""

    @staticmethod
    def forward(ctx, tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin):
        ctx.save_for_backward(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin)
        return pix2vin

    @staticmethod
    def backward(ctx, dldw_pix2vout):
        tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin = ctx.saved_tensors
        dldw_vtx2xyz = bwd(tri2vtx, vtx2xyz.detach(), transform_world2pix, pix2tri,
                           pix2vin.detach(), dldw_pix2vout.contiguous())
        dldw_pix2vin = dldw_pix2vout.clone()
        return None, dldw_vtx2xyz, None, None, dldw_pix2vin


def edge_gradient_and_type(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                           transform_world2pix: torch.Tensor, pix2tri: torch.Tensor,
                           pix2val: torch.Tensor, dldw_pix2val: torch.Tensor):
    """Per pixel edge: its type and the loss gradient across it
    (del-msh-cpu/src/rasterized_edge_gradient.rs:77-200).

    Returns:
        hedge2type (H-1, W) uint8, hedge2dldr (H-1, W) float32,
        vedge2type (H, W-1) uint8, vedge2dldr (H, W-1) float32
    """
    vtx2xyz = vtx2xyz.detach()
    pix2val = pix2val.detach()
    assert len(tri2vtx.shape) == 2 and len(vtx2xyz.shape) == 2 and len(dldw_pix2val.shape) == 3
    img_h, img_w = pix2tri.shape
    num_tri, num_vtx, num_vdim = tri2vtx.shape[0], vtx2xyz.shape[0], dldw_pix2val.shape[2]
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_world2pix, (4, 4), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val)
    hedge2type = torch.zeros((img_h - 1, img_w), dtype=torch.uint8, device=device)
    hedge2dldr = torch.zeros((img_h - 1, img_w), dtype=torch.float32, device=device)
    vedge2type = torch.zeros((img_h, img_w - 1), dtype=torch.uint8, device=device)
    vedge2dldr = torch.zeros((img_h, img_w - 1), dtype=torch.float32, device=device)
    from . import edge_gradient_and_type as native

    native(tri2vtx, vtx2xyz, transform_world2pix.T.contiguous().flatten(), pix2tri, pix2val,
           dldw_pix2val.contiguous(), hedge2type, hedge2dldr, vedge2type, vedge2dldr,
           _lib.stream_of(device))
    return hedge2type, hedge2dldr, vedge2type, vedge2dldr


def smooth_gradient(hedge2type: torch.Tensor, vedge2type: torch.Tensor, num_iter: int,
                    hedge2dldr: torch.Tensor, vedge2dldr: torch.Tensor):
    """Smooth the pixel-edge gradients in place: num_iter red-black sweeps,
    edges of type 2 / 3 held fixed (the reference GPU path,
    rasterized_edge_gradient.cu:139-275)."""
    device = hedge2type.device
    img_h, img_w = hedge2type.shape[0] + 1, hedge2type.shape[1]
    assert_shape_dtype_device(hedge2type, (img_h - 1, img_w), torch.uint8, device)
    assert_shape_dtype_device(hedge2dldr, (img_h - 1, img_w), torch.float32, device)
    assert_shape_dtype_device(vedge2type, (img_h, img_w - 1), torch.uint8, device)
    assert_shape_dtype_device(vedge2dldr, (img_h, img_w - 1), torch.float32, device)
    _lib.require_gpu(hedge2type, hedge2dldr, vedge2type, vedge2dldr)
    from . import smooth_gradient as native

    native(hedge2type, hedge2dldr, vedge2type, vedge2dldr, num_iter, _lib.stream_of(device))


def interpolate(hedge2vy: torch.Tensor, vedge2vx: torch.Tensor, vtx2xy: torch.Tensor):
    """(x velocity, y velocity) of each point, bilinear from the pixel-edge
    grids (rasterized_edge_gradient.rs:202-248)."""
    img_h, img_w = hedge2vy.shape[0] + 1, hedge2vy.shape[1]
    num_vtx = vtx2xy.shape[0]
    device = hedge2vy.device
    assert_shape_dtype_device(hedge2vy, (img_h - 1, img_w), torch.float32, device)
    assert_shape_dtype_device(vedge2vx, (img_h, img_w - 1), torch.float32, device)
    vtx2xy = vtx2xy.detach().contiguous()
    assert_shape_dtype_device(vtx2xy, (num_vtx, 2), torch.float32, device)
    _lib.require_gpu(hedge2vy, vedge2vx, vtx2xy)
    vtx2velo = torch.empty((num_vtx, 2), dtype=torch.float32, device=device)
    from . import interpolate as native

    native(hedge2vy, vedge2vx, vtx2xy, vtx2velo, _lib.stream_of(device))
    return vtx2velo


class AutogradWithSmooth(torch.autograd.Function):
    """Identity in the forward pass; the backward spreads the pixel-edge
    gradients over the image by smoothing (100 red-black sweeps), reads them
    at the projected vertices and chains through the perspective projection
    (reference: RasterizedEdgeGradient/torch.py:237-274)."""

    @staticmethod
    def forward(ctx, tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val):
        ctx.save_for_backward(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val)
        return pix2val

    @staticmethod
    def backward(ctx, dldw_pix2val):
        tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val = ctx.saved_tensors
        hedge2type, hedge2dldr, vedge2type, vedge2dldr = edge_gradient_and_type(
            tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val.contiguous())
        smooth_gradient(hedge2type, vedge2type, 100, hedge2dldr, vedge2dldr)
        # pixel position u = (q0, q1) / q3 with q = M [x, 1] (M = world2pix,
        # possibly perspective): du/dq = [I2 / q3 | 0 | -u / q3], dq/dx = M[:, :3];
        # one kernel projects, reads dL/du from the grids and chains
        from . import smoothed_vertex_gradient as native

        xyz = vtx2xyz.detach().contiguous()
        dldx = torch.empty_like(xyz)
        native(hedge2dldr, vedge2dldr, xyz, transform_world2pix.T.contiguous(), dldx,
               _lib.stream_of(xyz.device))
        return None, dldx, None, None, dldw_pix2val
