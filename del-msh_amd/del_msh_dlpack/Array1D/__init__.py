"""Mid-level argsort (reference: del_msh_dlpack/Array1D/__init__.py)."""

from .. import _lib
from ..util_torch import as_tensor


def permute(old2val, new2old, new2val, stream_ptr=0):
    """new2val[i] = old2val[new2old[i]] for uint32 arrays (reference:
    del_msh_dlpack/Array1D/__init__.py, native array1d_permute)."""
    old2val, new2old, new2val = as_tensor(old2val), as_tensor(new2old), as_tensor(new2val)
    dev = _lib.require_gpu(old2val, new2old, new2val)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_permute_u32(stream, old2val.shape[0], _lib.ptr(old2val),
                                           _lib.ptr(new2old), _lib.ptr(new2val)),
               "dmsh_permute_u32")


def argsort(idx2val, jdx2idx, stream_ptr=0, key_bits=None):
    """Sorts idx2val in place (stable) and writes the permutation to jdx2idx."""
    idx2val, jdx2idx = as_tensor(idx2val), as_tensor(jdx2idx)
    dev = _lib.require_gpu(idx2val, jdx2idx)
    n = idx2val.shape[0]
    kb8 = idx2val.element_size() == 8
    bits = key_bits if key_bits is not None else (64 if kb8 else 32)
    ws = _lib.workspace(_lib.lib().dmsh_argsort_workspace_bytes(n, 8 if kb8 else 4), dev)
    fn = _lib.lib().dmsh_argsort_u64 if kb8 else _lib.lib().dmsh_argsort_u32
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(fn(stream, n, _lib.ptr(idx2val), _lib.ptr(jdx2idx), bits, _lib.ptr(ws), ws.numel()),
               "dmsh_argsort")


def has_duplicate_sorted_array(idx2val, stream_ptr=0) -> bool:
    """True if two neighbours of the sorted uint32 array are equal
    (reference: del-msh-dlpack/src/array1d.rs:156-193)."""
    import torch

    idx2val = as_tensor(idx2val)
    dev = _lib.require_gpu(idx2val)
    flag = torch.empty((1,), dtype=torch.uint32, device=dev)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_has_duplicate_sorted_u32(stream, idx2val.shape[0], _lib.ptr(idx2val),
                                                        _lib.ptr(flag)),
               "dmsh_has_duplicate_sorted_u32")
    return bool(flag.view(torch.int32).item())


def unique_for_sorted_array(idx2val, idx2jdx, stream_ptr=0):
    """idx2jdx[i] = number of value changes in idx2val[0..i]
    (reference: del-msh-dlpack/src/array1d.rs:195-240)."""
    idx2val, idx2jdx = as_tensor(idx2val), as_tensor(idx2jdx)
    dev = _lib.require_gpu(idx2val, idx2jdx)
    n = idx2val.shape[0]
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_unique_for_sorted_workspace_bytes(n), dev)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(L.dmsh_unique_for_sorted_u32(stream, n, _lib.ptr(idx2val), _lib.ptr(idx2jdx),
                                            _lib.ptr(ws), ws.numel()), "dmsh_unique_for_sorted_u32")


def unique_jdx2val_jdx2idx(idx2val, idx2jdx, jdx2val, jdx2idx_offset, stream_ptr=0):
    """jdx2idx_offset [num_jdx+1] and jdx2val [num_jdx] of the groups
    (reference: del-msh-dlpack/src/array1d.rs:242-300)."""
    idx2val, idx2jdx = as_tensor(idx2val), as_tensor(idx2jdx)
    jdx2val, jdx2idx_offset = as_tensor(jdx2val), as_tensor(jdx2idx_offset)
    dev = _lib.require_gpu(idx2val, idx2jdx, jdx2val, jdx2idx_offset)
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_unique_jdx2val_jdx2idx(
        stream, idx2val.shape[0], _lib.ptr(idx2val), _lib.ptr(idx2jdx), jdx2val.shape[0],
        _lib.ptr(jdx2val), _lib.ptr(jdx2idx_offset)), "dmsh_unique_jdx2val_jdx2idx")
