"""Mixed polyhedron meshes in CSR form (reference:
del_msh_dlpack/PolyhedronMesh/torch.py).  elem2volume and subdivide have no
GPU path in the reference and are not provided here."""

import torch

from .. import _lib


def _check_csr(elem2idx_offset, idx2vtx, vtx2xyz, num_dim=3):
    num_elem = elem2idx_offset.shape[0] - 1
    device = elem2idx_offset.device
    assert elem2idx_offset.dtype == torch.uint32 and elem2idx_offset.shape == (num_elem + 1,)
    assert idx2vtx.dtype == torch.uint32 and idx2vtx.dim() == 1 and idx2vtx.device == device
    assert vtx2xyz.dtype == torch.float32 and vtx2xyz.dim() == 2 and vtx2xyz.device == device
    if num_dim is not None:
        assert vtx2xyz.shape[1] == num_dim
    return num_elem, device


def make_elem2centroid(elem2idx_offset: torch.Tensor, idx2vtx: torch.Tensor,
                       vtx2xyz: torch.Tensor) -> torch.Tensor:
    """(num_elem, num_dim) float32: the mean of each element's vertices."""
    num_elem, device = _check_csr(elem2idx_offset, idx2vtx, vtx2xyz, None)
    elem2center = torch.empty((num_elem, vtx2xyz.shape[1]), dtype=torch.float32, device=device)
    from . import elem2center as native

    native(elem2idx_offset, idx2vtx, vtx2xyz, elem2center, _lib.stream_of(device))
    return elem2center


def make_bvhnodes_bvhnode2aabb(elem2idx_offset: torch.Tensor, idx2vtx: torch.Tensor,
                               vtx2xyz: torch.Tensor):
    """Element BVH: centroids fitted into the unit cube, 30-bit Morton codes,
    stable argsort, Karras topology, refit over the elements' vertices."""
    vtx2xyz = vtx2xyz.detach()
    num_elem, device = _check_csr(elem2idx_offset, idx2vtx, vtx2xyz)
    elem2centroid = make_elem2centroid(elem2idx_offset, idx2vtx, vtx2xyz)
    from ..Array1D.torch import argsort
    from ..Mat44.torch import from_fit_vtx2xyz_into_unit_cube
    from ..Mortons.torch import make_bvhnodes_from_sorted_mortons, make_vtx2morton_from_vtx2co

    transform_co2unit = from_fit_vtx2xyz_into_unit_cube(elem2centroid)
    elem2morton = make_vtx2morton_from_vtx2co(elem2centroid, transform_co2unit)
    idx2elem, idx2morton = argsort(elem2morton)
    bvhnodes = make_bvhnodes_from_sorted_mortons(idx2elem, idx2morton)
    bvhnode2aabb = torch.empty((bvhnodes.shape[0], 6), dtype=torch.float32, device=device)
    from . import bvhnode2aabb_from_bvhnodes

    bvhnode2aabb_from_bvhnodes(elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb,
                               _lib.stream_of(device))
    return bvhnodes, bvhnode2aabb


def search_elem_contain_points(elem2idx_offset: torch.Tensor, idx2vtx: torch.Tensor,
                               vtx2xyz: torch.Tensor, bvhnodes: torch.Tensor,
                               bvhnode2aabb: torch.Tensor, wtx2xyz: torch.Tensor):
    """For each query point, the first element (left-first walk of the BVH)
    that contains it and the point's parametric coordinates there.

    Returns:
        wtx2elem: (num_wtx,) uint32 -- element, 0xFFFFFFFF outside the mesh
        wtx2param: (num_wtx, 3) float32 -- tet (r0, r1, r2) with r3 implicit;
            pyramid / prism (r, s, t) in [0, 1]; hex (r, s, t) in [-1, 1]
    """
    num_wtx = wtx2xyz.shape[0]
    device = wtx2xyz.device
    assert wtx2xyz.shape == (num_wtx, 3) and wtx2xyz.dtype == torch.float32
    wtx2elem = torch.empty(num_wtx, dtype=torch.uint32, device=device)
    wtx2param = torch.empty((num_wtx, 3), dtype=torch.float32, device=device)
    from . import search_elem_contain_points as native

    native(bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz, wtx2xyz, wtx2elem,
           wtx2param, _lib.stream_of(_lib.require_gpu(wtx2xyz)))
    return wtx2elem, wtx2param


def interpolate_values_at_points(elem2idx_offset: torch.Tensor, idx2vtx: torch.Tensor,
                                 vtx2value: torch.Tensor, wtx2elem: torch.Tensor,
                                 wtx2param: torch.Tensor) -> torch.Tensor:
    """(num_wtx, num_value_dim) float32: shape-function interpolation of the
    per-vertex values (0 for points outside the mesh)."""
    num_elem, device = _check_csr(elem2idx_offset, idx2vtx, vtx2value, None)
    num_wtx = wtx2elem.shape[0]
    assert wtx2elem.dtype == torch.uint32 and wtx2param.shape == (num_wtx, 3)
    wtx2value = torch.zeros((num_wtx, vtx2value.shape[1]), dtype=torch.float32, device=device)
    from . import interpolate_values_at_points as native

    native(elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param, wtx2value,
           _lib.stream_of(_lib.require_gpu(wtx2elem)))
    return wtx2value
