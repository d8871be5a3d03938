"""Mid-level argsort (reference: del_msh_dlpack/Array1D/__init__.py)."""

from .. import _lib
from ..util_torch import as_tensor


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
