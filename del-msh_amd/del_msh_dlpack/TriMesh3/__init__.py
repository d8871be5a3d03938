"""Mid-level TriMesh3 entry points (reference: del_msh_dlpack/TriMesh3/__init__.py)."""

from .. import _lib
from ..util_torch import as_tensor


def make_bvhnode2aabb_from_bvhnodes(tri2vtx, vtx2xyz0, vtx2xyz1, bvhnodes, bvhnode2aabb,
                                    stream_ptr=0):
    """Refit bvhnode2aabb for an LBVH over triangles.  As on the reference GPU
    path, vtx2xyz1 (continuous-collision motion) is not used."""
    tri2vtx, vtx2xyz0, bvhnodes, bvhnode2aabb = map(
        as_tensor, (tri2vtx, vtx2xyz0, bvhnodes, bvhnode2aabb))
    dev = _lib.require_gpu(tri2vtx, vtx2xyz0, bvhnodes, bvhnode2aabb)
    num_tri = tri2vtx.shape[0]
    num_node = bvhnodes.shape[0]
    stream = stream_ptr or _lib.stream_of(dev)
    ws = _lib.kept_workspace("dmsh_bvhnode2aabb_trimesh3",
                             _lib.lib().dmsh_refit_workspace_bytes(num_tri), dev, stream)
    _lib.check(
        _lib.lib().dmsh_bvhnode2aabb_trimesh3(
            stream, _lib.ptr(bvhnode2aabb), num_node, _lib.ptr(bvhnodes), num_tri,
            _lib.ptr(tri2vtx), vtx2xyz0.shape[0], _lib.ptr(vtx2xyz0), 0.0, _lib.ptr(ws),
            ws.numel()),
        "dmsh_bvhnode2aabb_trimesh3",
    )
