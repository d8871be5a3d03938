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
