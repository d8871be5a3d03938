"""Array1D.argsort (reference: del_msh_dlpack/Array1D/torch.py:41-72).

Stable: equal keys keep their input order (the order of del-msh-cpu's
`sort_by_key`, del-msh-cpu/src/mortons.rs:143)."""

import torch

from .. import _lib


def argsort(idx2val: torch.Tensor, key_bits: int | None = None):
    """Sort a 1D uint32 (or uint64) array.

    Returns:
        jdx2idx: (n,) uint32 -- original index of the j-th smallest value
        jdx2val: (n,) same dtype -- sorted values
    """
    n = idx2val.shape[0]
    device = idx2val.device
    assert idx2val.shape == (n,)
    assert idx2val.dtype in (torch.uint32, torch.uint64)
    _lib.require_gpu(idx2val)
    jdx2idx = torch.empty(size=(n,), device=device, dtype=torch.uint32)
    jdx2val = idx2val.clone()
    from . import argsort as native

    native(jdx2val, jdx2idx, _lib.stream_of(device), key_bits)
    return jdx2idx, jdx2val
