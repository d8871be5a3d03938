"""Pix2Tri (reference: del_msh_dlpack/Pix2Tri/torch.py:8-201)."""

from typing import Tuple

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def _check_scene(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world):
    num_tri = tri2vtx.shape[0]
    num_vtx = vtx2xyz.shape[0]
    num_bvhnode = bvhnodes.shape[0]
    device = tri2vtx.device
    assert num_bvhnode == num_tri * 2 - 1
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(bvhnodes, (num_bvhnode, 3), torch.uint32, device)
    assert_shape_dtype_device(bvhnode2aabb, (num_bvhnode, 6), torch.float32, device)
    assert_shape_dtype_device(transform_ndc2world, (4, 4), torch.float32, device)
    _lib.require_gpu(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world)
    return device


def by_raycasting(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor, bvhnodes: torch.Tensor,
                  bvhnode2aabb: torch.Tensor, transform_ndc2world: torch.Tensor,
                  img_shape: Tuple[int, int]) -> torch.Tensor:
    """Hit triangle per pixel, u32::MAX on a miss.

    img_shape is (width, height); the result is (height, width) uint32.  The
    hit is the one del-msh-cpu's ordered BVH search returns
    (search_bvh3.rs:103-221).
    """
    vtx2xyz = vtx2xyz.detach()
    transform_ndc2world = transform_ndc2world.contiguous()
    device = _check_scene(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world)
    pix2tri = torch.full((img_shape[1], img_shape[0]), torch.iinfo(torch.uint32).max,
                         dtype=torch.uint32, device=device)
    from . import update_pix2tri

    update_pix2tri(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb,
                   transform_ndc2world.T.contiguous(), pix2tri, _lib.stream_of(device))
    return pix2tri


def by_raycasting_with_coeff(tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                             bvhnodes: torch.Tensor, bvhnode2aabb: torch.Tensor,
                             transform_ndc2world: torch.Tensor, img_shape: Tuple[int, int]):
    """As by_raycasting, plus the hit ray coefficient t per pixel (float32, 0 on
    a miss; the ray runs from NDC z = +1 at t = 0 to NDC z = -1 at t = 1)."""
    vtx2xyz = vtx2xyz.detach()
    transform_ndc2world = transform_ndc2world.contiguous()
    device = _check_scene(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world)
    pix2tri = torch.empty((img_shape[1], img_shape[0]), dtype=torch.uint32, device=device)
    pix2t = torch.empty((img_shape[1], img_shape[0]), dtype=torch.float32, device=device)
    from . import update_pix2tri

    update_pix2tri(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb,
                   transform_ndc2world.T.contiguous(), pix2tri, _lib.stream_of(device), pix2t)
    return pix2tri, pix2t


def interpolate_fwd(pix2tri: torch.Tensor, tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                    vtx2val: torch.Tensor, transform_ndc2world: torch.Tensor) -> torch.Tensor:
    """Per pixel, vtx2val interpolated at the pixel ray's hit on its
    triangle's plane (del-msh-cpu/src/pix2tri.rs:176-229); 0 on the
    background.  Returns pix2val (H, W, num_vdim) float32."""
    vtx2xyz = vtx2xyz.detach()
    vtx2val = vtx2val.detach()
    transform_ndc2world = transform_ndc2world.contiguous()
    num_tri, num_vtx, num_vdim = tri2vtx.shape[0], vtx2xyz.shape[0], vtx2val.shape[1]
    img_h, img_w = pix2tri.shape[0], pix2tri.shape[1]
    device = tri2vtx.device
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(vtx2val, (num_vtx, num_vdim), torch.float32, device)
    assert_shape_dtype_device(transform_ndc2world, (4, 4), torch.float32, device)
    _lib.require_gpu(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world)
    pix2val = torch.zeros((img_h, img_w, num_vdim), dtype=torch.float32, device=device)
    from . import interpolate as native

    native(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world.T.contiguous(), pix2val,
           _lib.stream_of(device))
    return pix2val


def interpolate_bwd(pix2tri: torch.Tensor, tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                    vtx2val: torch.Tensor, transform_ndc2world: torch.Tensor,
                    dldw_pix2val: torch.Tensor, dldw_vtx2xyz: torch.Tensor,
                    dldw_vtx2val: torch.Tensor):
    """Accumulates into dldw_vtx2xyz (num_vtx, 3) and dldw_vtx2val
    (num_vtx, num_vdim) (del-msh-cpu/src/pix2tri.rs:231-304)."""
    num_tri, num_vtx, num_vdim = tri2vtx.shape[0], vtx2xyz.shape[0], vtx2val.shape[1]
    img_h, img_w = pix2tri.shape[0], pix2tri.shape[1]
    device = tri2vtx.device
    assert_shape_dtype_device(pix2tri, (img_h, img_w), torch.uint32, device)
    assert_shape_dtype_device(tri2vtx, (num_tri, 3), torch.uint32, device)
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(vtx2val, (num_vtx, num_vdim), torch.float32, device)
    assert_shape_dtype_device(transform_ndc2world, (4, 4), torch.float32, device)
    assert_shape_dtype_device(dldw_pix2val, (img_h, img_w, num_vdim), torch.float32, device)
    assert_shape_dtype_device(dldw_vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(dldw_vtx2val, (num_vtx, num_vdim), torch.float32, device)
    _lib.require_gpu(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val,
                     dldw_vtx2xyz, dldw_vtx2val)
    from . import interpolate_bwd as native

    native(pix2tri, tri2vtx, vtx2xyz.detach(), vtx2val.detach(),
           transform_ndc2world.T.contiguous(), dldw_pix2val.contiguous(), dldw_vtx2xyz,
           dldw_vtx2val, _lib.stream_of(device))


class AutogradInterpolate(torch.autograd.Function):
    """pix2val = interpolate(vtx2val) with gradients to vtx2xyz and vtx2val
    (reference: Pix2Tri/torch.py:169-195)."""

    @staticmethod
    def forward(ctx, pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world):
        ctx.save_for_backward(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world)
        return interpolate_fwd(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world)

    @staticmethod
    def backward(ctx, dldw_pix2val):
        pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world = ctx.saved_tensors
        dldw_vtx2xyz = torch.zeros_like(vtx2xyz)
        dldw_vtx2val = torch.zeros_like(vtx2val)
        interpolate_bwd(pix2tri, tri2vtx, vtx2xyz.detach(), vtx2val.detach(), transform_ndc2world,
                        dldw_pix2val.contiguous(), dldw_vtx2xyz, dldw_vtx2val)
        return None, None, dldw_vtx2xyz, dldw_vtx2val, None


def interpolate(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world):
    return AutogradInterpolate.apply(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world)
