"""Mid-level Morton entry points (reference: del_msh_dlpack/Mortons/__init__.py).
Arguments may be tensors or DLPack capsules; the transform is column-major."""

from .. import _lib
from ..util_torch import as_tensor


def make_vtx2morton_from_vtx2co(vtx2co, transform_co2unit, vtx2morton, stream_ptr=0):
    vtx2co, transform_co2unit, vtx2morton = map(as_tensor, (vtx2co, transform_co2unit, vtx2morton))
    dev = _lib.require_gpu(vtx2co, transform_co2unit, vtx2morton)
    num_vtx, num_dim = vtx2co.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(
        _lib.lib().dmsh_vtx2morton_u32(stream, num_vtx, _lib.ptr(vtx2co), num_dim,
                                       _lib.ptr(transform_co2unit), _lib.ptr(vtx2morton)),
        "dmsh_vtx2morton_u32",
    )
    return vtx2morton


def make_bvhnodes_from_sorted_mortons(idx2obj, idx2morton, bvhnodes, stream_ptr=0):
    idx2obj, idx2morton, bvhnodes = map(as_tensor, (idx2obj, idx2morton, bvhnodes))
    dev = _lib.require_gpu(idx2obj, idx2morton, bvhnodes)
    n = idx2obj.shape[0]
    stream = stream_ptr or _lib.stream_of(dev)
    fn = (_lib.lib().dmsh_bvhnodes_from_sorted_mortons_u64 if idx2morton.element_size() == 8
          else _lib.lib().dmsh_bvhnodes_from_sorted_mortons_u32)
    _lib.check(fn(stream, n, _lib.ptr(bvhnodes), _lib.ptr(idx2morton), _lib.ptr(idx2obj)),
               "dmsh_bvhnodes_from_sorted_mortons")
