"""Mid-level offset-array aggregate (reference: del-msh-dlpack/src/
offset_array.rs, offset_array_aggregate)."""

from .. import _lib
from ..util_torch import as_tensor


def aggregate(idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval, stream_ptr=0):
    """Adds, per idx, kdx2val[jdx2kdx[j]] for j in [off[idx], off[idx+1]) in
    j order onto idx2aggval."""
    idx2jdx_offset, jdx2kdx = as_tensor(idx2jdx_offset), as_tensor(jdx2kdx)
    kdx2val, idx2aggval = as_tensor(kdx2val), as_tensor(idx2aggval)
    dev = _lib.require_gpu(idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval)
    num_idx = idx2jdx_offset.shape[0] - 1
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_offset_array_aggregate(
        stream, num_idx, _lib.ptr(idx2jdx_offset), jdx2kdx.shape[0], _lib.ptr(jdx2kdx),
        kdx2val.shape[0], idx2aggval.shape[1], _lib.ptr(kdx2val), _lib.ptr(idx2aggval)),
        "dmsh_offset_array_aggregate")
