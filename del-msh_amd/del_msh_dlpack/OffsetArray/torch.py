"""OffsetArray.aggregate (reference: del_msh_dlpack/OffsetArray/torch.py:5-52)."""

import torch

from .. import _lib


def aggregate(idx2jdx_offset: torch.Tensor, jdx2kdx: torch.Tensor, kdx2val: torch.Tensor):
    """idx2aggval[i] = sum of kdx2val[jdx2kdx[j]] over j in the i-th range,
    added in j order.

    Args:
        idx2jdx_offset: (num_idx+1,) uint32
        jdx2kdx: (num_jdx,) uint32
        kdx2val: (num_kdx, num_dim) float32
    Returns:
        idx2aggval: (num_idx, num_dim) float32
    """
    kdx2val = kdx2val.detach()
    num_idx = idx2jdx_offset.shape[0] - 1
    device = idx2jdx_offset.device
    assert idx2jdx_offset.dtype == torch.uint32 and jdx2kdx.dtype == torch.uint32
    assert kdx2val.dtype == torch.float32 and kdx2val.dim() == 2
    _lib.require_gpu(idx2jdx_offset, jdx2kdx, kdx2val)
    idx2aggval = torch.zeros((num_idx, kdx2val.shape[1]), device=device, dtype=torch.float32)
    from . import aggregate as native

    native(idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval, _lib.stream_of(device))
    return idx2aggval
