"""Pix2Tri (reference: del_msh_dlpack/Pix2Tri/torch.py:8-61)."""

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
