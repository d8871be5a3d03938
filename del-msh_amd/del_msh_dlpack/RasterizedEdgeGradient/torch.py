"""RasterizedEdgeGradient (reference: del_msh_dlpack/RasterizedEdgeGradient/torch.py:7-64,214-234)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def bwd(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor, transform_world2pix: torch.Tensor,
        pix2tri: torch.Tensor, pix2val: torch.Tensor, dldw_pix2val: torch.Tensor) -> torch.Tensor:
    """dL/d(vertex positions) from the silhouette edges of the rasterized image.

    Args:
        tri2vtx: (num_tri, 3) uint32; vtx2xyz: (num_vtx, 3) float32
        transform_world2pix: (4, 4) float32, row-major
        pix2tri: (H, W) uint32; pix2val, dldw_pix2val: (H, W, nv) float32
    Returns:
        dldw_vtx2xyz: (num_vtx, 3) float32
    """
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w, num_vdim = pix2val.shape
    device = tri2vtx.device
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val)
    dldw_vtx2xyz = torch.zeros_like(vtx2xyz)
    from . import bwd as native

    native(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix.T.contiguous().flatten(),
           pix2tri, pix2val, dldw_pix2val, _lib.stream_of(device))
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
