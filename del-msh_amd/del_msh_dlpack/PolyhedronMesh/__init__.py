"""Mid-level functions of mixed polyhedron meshes in CSR form (reference:
del-msh-dlpack/src/polyhedron_mesh.rs)."""

from .. import _lib
from ..util_torch import as_tensor


def elem2center(elem2idx_offset, idx2vtx, vtx2xyz, elem2center, stream_ptr=0):
    elem2idx_offset, idx2vtx = as_tensor(elem2idx_offset), as_tensor(idx2vtx)
    vtx2xyz, elem2center = as_tensor(vtx2xyz), as_tensor(elem2center)
    dev = _lib.require_gpu(elem2idx_offset, idx2vtx, vtx2xyz, elem2center)
    num_vtx, num_dim = vtx2xyz.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_polyhedron_elem2center(
        stream, elem2idx_offset.shape[0] - 1, _lib.ptr(elem2idx_offset), idx2vtx.shape[0],
        _lib.ptr(idx2vtx), num_vtx, num_dim, _lib.ptr(vtx2xyz), _lib.ptr(elem2center)),
        "dmsh_polyhedron_elem2center")


def bvhnode2aabb_from_bvhnodes(elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb,
                               stream_ptr=0):
    ts = [as_tensor(t) for t in (elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb)]
    elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb = ts
    dev = _lib.require_gpu(*ts)
    num_elem = elem2idx_offset.shape[0] - 1
    L = _lib.lib()
    stream = stream_ptr or _lib.stream_of(dev)
    ws = _lib.kept_workspace("dmsh_bvhnode2aabb_polyhedron",
                             L.dmsh_polyhedron_refit_workspace_bytes(num_elem), dev, stream)
    _lib.check(L.dmsh_bvhnode2aabb_polyhedron(
        stream, bvhnodes.shape[0], _lib.ptr(bvhnodes), num_elem, _lib.ptr(elem2idx_offset),
        idx2vtx.shape[0], _lib.ptr(idx2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz),
        _lib.ptr(bvhnode2aabb), _lib.ptr(ws), ws.numel()), "dmsh_bvhnode2aabb_polyhedron")


def search_elem_contain_points(bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz, wtx2xyz,
                               wtx2elem, wtx2param, stream_ptr=0):
    """Argument order of the reference's native function
    (polyhedron_mesh.rs:87-160)."""
    ts = [as_tensor(t) for t in (bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz,
                                 wtx2xyz, wtx2elem, wtx2param)]
    bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz, wtx2xyz, wtx2elem, wtx2param = ts
    dev = _lib.require_gpu(*ts)
    stream = stream_ptr or _lib.stream_of(dev)
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_polyhedron_search_workspace_bytes(wtx2xyz.shape[0]), dev)
    _lib.check(L.dmsh_polyhedron_search_elem_contain_points(
        stream, elem2idx_offset.shape[0] - 1, _lib.ptr(elem2idx_offset), idx2vtx.shape[0],
        _lib.ptr(idx2vtx), vtx2xyz.shape[0], _lib.ptr(vtx2xyz), _lib.ptr(bvhnodes),
        _lib.ptr(bvhnode2aabb), wtx2xyz.shape[0], _lib.ptr(wtx2xyz), _lib.ptr(wtx2elem),
        _lib.ptr(wtx2param), _lib.ptr(ws), ws.numel()), "dmsh_polyhedron_search_elem_contain_points")


def interpolate_values_at_points(elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param,
                                 wtx2value, stream_ptr=0):
    ts = [as_tensor(t) for t in (elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param,
                                 wtx2value)]
    elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param, wtx2value = ts
    dev = _lib.require_gpu(*ts)
    num_vtx, num_vdim = vtx2value.shape
    stream = stream_ptr or _lib.stream_of(dev)
    _lib.check(_lib.lib().dmsh_polyhedron_interpolate_at_points(
        stream, wtx2elem.shape[0], elem2idx_offset.shape[0] - 1, _lib.ptr(elem2idx_offset),
        idx2vtx.shape[0], _lib.ptr(idx2vtx), num_vtx, num_vdim, _lib.ptr(vtx2value),
        _lib.ptr(wtx2elem), _lib.ptr(wtx2param), _lib.ptr(wtx2value)),
        "dmsh_polyhedron_interpolate_at_points")
