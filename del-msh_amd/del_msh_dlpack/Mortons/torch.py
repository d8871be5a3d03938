"""Morton codes and LBVH topology (reference: del_msh_dlpack/Mortons/torch.py:5-82)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def make_vtx2morton_from_vtx2co(vtx2co: torch.Tensor, transform_co2unit: torch.Tensor):
    """Morton code per point after mapping it into the unit cube.

    Args:
        vtx2co: (num_vtx, 2 or 3) float32
        transform_co2unit: (3x3 or 4x4) float32, row-major (applied as M @ [p, 1])
    Returns:
        vtx2morton: (num_vtx,) uint32 -- 2D: 16 bits/axis, 3D: 10 bits/axis
    """
    vtx2co = vtx2co.detach()
    num_vtx, num_dim = vtx2co.shape
    device = vtx2co.device
    assert num_dim == 2 or num_dim == 3
    assert_shape_dtype_device(vtx2co, (num_vtx, num_dim), torch.float32, device)
    assert_shape_dtype_device(transform_co2unit, (num_dim + 1, num_dim + 1), torch.float32, device)
    _lib.require_gpu(vtx2co, transform_co2unit)
    vtx2morton = torch.empty((num_vtx,), device=device, dtype=torch.uint32)
    from . import make_vtx2morton_from_vtx2co as native

    native(vtx2co, transform_co2unit.detach().T.contiguous(), vtx2morton, _lib.stream_of(device))
    return vtx2morton


def make_vtx2morton63_from_vtx2xyz(vtx2xyz: torch.Tensor, transform_co2unit: torch.Tensor):
    """63-bit Morton codes (21 bits/axis) as uint64 -- extension, 3D only."""
    vtx2xyz = vtx2xyz.detach()
    num_vtx = vtx2xyz.shape[0]
    device = vtx2xyz.device
    assert_shape_dtype_device(vtx2xyz, (num_vtx, 3), torch.float32, device)
    assert_shape_dtype_device(transform_co2unit, (4, 4), torch.float32, device)
    _lib.require_gpu(vtx2xyz, transform_co2unit)
    out = torch.empty((num_vtx,), device=device, dtype=torch.uint64)
    t = transform_co2unit.detach().T.contiguous()
    _lib.check(
        _lib.lib().dmsh_vtx2morton_u64(_lib.stream_of(device), num_vtx, _lib.ptr(vtx2xyz),
                                       _lib.ptr(t), _lib.ptr(out)),
        "dmsh_vtx2morton_u64",
    )
    return out


def make_bvhnodes_from_sorted_mortons(idx2obj: torch.Tensor, idx2morton: torch.Tensor):
    """Karras LBVH topology from sorted codes (uint32, or uint64 for 63-bit codes).

    Returns:
        bvhnodes: (2n-1, 3) uint32 rows (parent, left|obj, right|u32::MAX)
    """
    n = idx2obj.shape[0]
    device = idx2obj.device
    assert idx2obj.dtype == torch.uint32 and len(idx2obj.shape) == 1
    assert idx2morton.dtype in (torch.uint32, torch.uint64) and len(idx2morton.shape) == 1
    assert idx2obj.shape == idx2morton.shape
    _lib.require_gpu(idx2obj, idx2morton)
    bvhnodes = torch.empty((n * 2 - 1, 3), device=device, dtype=torch.uint32)
    from . import make_bvhnodes_from_sorted_mortons as native

    native(idx2obj.contiguous(), idx2morton.contiguous(), bvhnodes, _lib.stream_of(device))
    return bvhnodes
