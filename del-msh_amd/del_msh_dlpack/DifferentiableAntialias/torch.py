"""DifferentiableAntialias (reference: del_msh_dlpack/DifferentiableAntialias/torch.py:5-144)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def antialias(cedge2vtx: torch.Tensor, vtx2xyz: torch.Tensor, transform_world2pix: torch.Tensor,
              pix2tri: torch.Tensor, pix2vin: torch.Tensor) -> torch.Tensor:
    """Blends pixel values across the contour edges (fractional coverage of
    the pixels the silhouette crosses).

    Args:
        cedge2vtx: (num_contour, 2) uint32; vtx2xyz: (num_vtx, 3) float32
        transform_world2pix: (4, 4) float32, row-major
        pix2tri: (H, W) uint32 (u32::MAX on background); pix2vin: (H, W) float32
    Returns:
        pix2vout: (H, W) float32
    """
    pix2vin = pix2vin.detach()
    num_contour = cedge2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w = pix2vin.shape[0], pix2vin.shape[1]
    device = cedge2vtx.device
    assert_shape_dtype_device(cedge2vtx, (num_contour, 2), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_world2pix, (4, 4), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(pix2vin, (img_h, img_w), torch.float32, device)
    _lib.require_gpu(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin)
    pix2vout = pix2vin.clone()
    from . import antialias as native

    native(cedge2vtx, vtx2xyz, transform_world2pix.T.contiguous().flatten(), pix2tri, pix2vin,
           pix2vout, _lib.stream_of(device))
    return pix2vout


def bwd_antialias(cedge2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                  transform_world2pix: torch.Tensor, pix2val: torch.Tensor,
                  dldw_pix2val: torch.Tensor, pix2tri: torch.Tensor) -> torch.Tensor:
    """dL/d(vertex positions) of the anti-aliased image.

    Args:
        pix2val: (H, W) float32, the image fed to the forward pass
        dldw_pix2val: (H, W) float32, dL/d(anti-aliased image)
    Returns:
        dldw_vtx2xyz: (num_vtx, 3) float32
    """
    pix2val = pix2val.detach()
    vtx2xyz = vtx2xyz.detach()
    num_contour = cedge2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w = dldw_pix2val.shape[0], dldw_pix2val.shape[1]
    device = cedge2vtx.device
    assert_shape_dtype_device(cedge2vtx, (num_contour, 2), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_world2pix, (4, 4), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2val, (img_h, img_w), torch.float32, device)
    assert_shape_dtype_device(pix2val, (img_h, img_w), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    _lib.require_gpu(cedge2vtx, vtx2xyz, transform_world2pix, pix2val, dldw_pix2val, pix2tri)
    dldw_vtx2xyz = torch.zeros_like(vtx2xyz)
    from . import bwd_antialias as native

    native(cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix.T.contiguous().flatten(),
           pix2val, dldw_pix2val, pix2tri, _lib.stream_of(device))
    return dldw_vtx2xyz


class DifferentiableAntialiasFunction(torch.autograd.Function):
    """Anti-aliasing forward; the backward adds the vertex gradient and passes
    dL/d(pixel) through to the input image unchanged."""

    @staticmethod
    def forward(ctx, cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin):
        ctx.save_for_backward(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin)
        return antialias(cedge2vtx.detach(), vtx2xyz.detach(), transform_world2pix.detach(),
                         pix2tri.detach(), pix2vin)

    @staticmethod
    def backward(ctx, dldw_pix2vout):
        cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2vin = ctx.saved_tensors
        dldw_vtx2xyz = bwd_antialias(cedge2vtx, vtx2xyz, transform_world2pix, pix2vin,
                                     dldw_pix2vout.contiguous(), pix2tri)
        dldw_pix2vin = dldw_pix2vout.clone()
        return None, dldw_vtx2xyz, None, None, dldw_pix2vin
