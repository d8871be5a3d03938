"""Array1D (reference: del_msh_dlpack/Array1D/torch.py).

argsort is stable: equal keys keep their input order (the order of
del-msh-cpu's `sort_by_key`, del-msh-cpu/src/mortons.rs:143)."""

import torch

from .. import _lib
from ..util_torch import assert_shape_dtype_device


def permute(old2val: torch.Tensor, new2old: torch.Tensor) -> torch.Tensor:
    """new2val[i] = old2val[new2old[i]] for (n,) uint32 arrays (reference
    Array1D/torch.py:5-38); an index >= n raises the device index error."""
    n = old2val.shape[0]
    device = old2val.device
    assert_shape_dtype_device(old2val, (n,), torch.uint32, device)
    assert_shape_dtype_device(new2old, (n,), torch.uint32, device)
    _lib.require_gpu(old2val, new2old)
    new2val = torch.empty(size=(n,), device=device, dtype=torch.uint32)
    from . import permute as native

    native(old2val, new2old, new2val, _lib.stream_of(device))
    return new2val


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


def has_duplicate_in_sorted_array(idx2val: torch.Tensor) -> bool:
    """Whether a sorted uint32 array holds a duplicate (reference:
    del_msh_dlpack/Array1D/torch.py:77-101)."""
    n = idx2val.shape[0]
    assert idx2val.shape == (n,) and idx2val.dtype == torch.uint32
    from . import has_duplicate_sorted_array

    return has_duplicate_sorted_array(idx2val, _lib.stream_of(_lib.require_gpu(idx2val)))


def unique_for_sorted_array(idx2val: torch.Tensor):
    """Unique values of a sorted uint32 array (reference:
    del_msh_dlpack/Array1D/torch.py:104-148).

    Returns:
        idx2jdx: (num_idx,) uint32 -- group of each element
        jdx2val: (num_jdx,) uint32 -- the unique values
        jdx2idx_offset: (num_jdx+1,) uint32 -- first element of each group
    """
    num_idx = idx2val.shape[0]
    device = idx2val.device
    assert idx2val.shape == (num_idx,) and idx2val.dtype == torch.uint32
    stream = _lib.stream_of(_lib.require_gpu(idx2val))
    from . import unique_for_sorted_array as native_unique
    from . import unique_jdx2val_jdx2idx

    idx2jdx = torch.empty((num_idx,), dtype=torch.uint32, device=device)
    native_unique(idx2val, idx2jdx, stream)
    num_jdx = int(idx2jdx[-1:].view(torch.int32).item()) + 1 if num_idx else 0
    jdx2val = torch.empty((num_jdx,), dtype=torch.uint32, device=device)
    jdx2idx_offset = torch.empty((num_jdx + 1,), dtype=torch.uint32, device=device)
    unique_jdx2val_jdx2idx(idx2val, idx2jdx, jdx2val, jdx2idx_offset, stream)
    return idx2jdx, jdx2val, jdx2idx_offset
