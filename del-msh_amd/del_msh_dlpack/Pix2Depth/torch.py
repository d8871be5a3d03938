"""Pix2Depth (reference: del_msh_dlpack/Pix2Depth/torch.py:5-131)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def update(pix2depth: torch.Tensor, pix2tri: torch.Tensor, tri2vtx: torch.Tensor,
           vtx2xyz: torch.Tensor, transform_ndc2world: torch.Tensor):
    """Depth (ndc_z + 1) / 2 of the stored hit per pixel, 0 on background."""
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w = pix2depth.shape[0], pix2depth.shape[1]
    device = pix2depth.device
    assert_shape_dtype_device(pix2depth, (img_h, img_w), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_ndc2world, (4, 4), torch.float32, device)
    _lib.require_gpu(pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world)
    from . import update as native

    native(pix2depth.detach(), pix2tri, tri2vtx, vtx2xyz,
           transform_ndc2world.T.contiguous(), _lib.stream_of(device))


def bwd_wrt_vtx2xyz(dldw_vtx2xyz: torch.Tensor, pix2tri: torch.Tensor, tri2vtx: torch.Tensor,
                    vtx2xyz: torch.Tensor, dldw_pix2depth: torch.Tensor,
                    transform_ndc2world: torch.Tensor):
    """Accumulates dL/d(vertex positions) of the depth image into
    dldw_vtx2xyz (num_vtx, 3), in place."""
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    img_h, img_w = pix2tri.shape[0], pix2tri.shape[1]
    device = dldw_vtx2xyz.device
    assert_shape_dtype_device(dldw_vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2depth, (img_h, img_w), torch.float32, device)
    assert_shape_dtype_device(transform_ndc2world, (4, 4), torch.float32, device)
    _lib.require_gpu(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world)
    from . import bwd_wrt_vtx2xyz as native

    native(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz.detach(), dldw_pix2depth,
           transform_ndc2world.T.contiguous(), _lib.stream_of(device))


class AutogradFunction(torch.autograd.Function):
    """Depth image as a differentiable function of the vertex positions
    (pix2tri fixed)."""

    @staticmethod
    def forward(ctx, vtx2xyz, pix2tri, tri2vtx, transform_ndc2world):
        ctx.save_for_backward(pix2tri, tri2vtx, vtx2xyz, transform_ndc2world)
        pix2depth = torch.zeros(pix2tri.shape, dtype=torch.float32, device=vtx2xyz.device)
        update(pix2depth, pix2tri, tri2vtx, vtx2xyz.detach(), transform_ndc2world)
        return pix2depth

    @staticmethod
    def backward(ctx, dldw_pix2depth):
        pix2tri, tri2vtx, vtx2xyz, transform_ndc2world = ctx.saved_tensors
        dldw_vtx2xyz = torch.zeros_like(vtx2xyz)
        bwd_wrt_vtx2xyz(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz.detach(),
                        dldw_pix2depth.contiguous(), transform_ndc2world)
        return dldw_vtx2xyz, None, None, None


def pix2depth_using_interpolation(vtx2xyz, pix2tri, tri2vtx, transform_ndc2world):
    """Depth through attribute interpolation, differentiable in vtx2xyz
    (reference Pix2Depth/torch.py:134-146): each pixel's hit point is the
    interpolation of the vertex positions over its triangle (the HIP
    interpolation kernel, AutogradInterpolate), projected by world2ndc;
    depth = (ndc z / w + 1) / 2, and 0 on the background."""
    from ..Pix2Tri.torch import AutogradInterpolate

    pix2xyz = AutogradInterpolate.apply(pix2tri, tri2vtx, vtx2xyz, vtx2xyz, transform_ndc2world)
    transform_world2ndc = transform_ndc2world.inverse()
    pix2xyzw = torch.cat([pix2xyz, torch.ones_like(pix2xyz[:, :, :1])], dim=-1)
    background = pix2tri == torch.iinfo(torch.uint32).max
    pix2ndcw = pix2xyzw @ transform_world2ndc.T
    pix2depth = (pix2ndcw[:, :, 2] / pix2ndcw[:, :, 3] + 1.0) * 0.5
    return torch.where(background, torch.zeros_like(pix2depth), pix2depth)
