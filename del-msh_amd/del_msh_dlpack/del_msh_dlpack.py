"""Native-level entry points of del_msh_dlpack on libdmsh (the names the
reference registers as its pyo3 module del_msh_dlpack.del_msh_dlpack,
del-msh-dlpack/src/lib.rs:29-57).

Every array argument is a DLPack capsule or producer (a tensor works too),
and `stream_ptr` is the HIP stream as an integer (0: the current stream of
the arrays' device), as in the reference.  Conventions kept from the
reference's del-dlpack checks:
  - sizes are read from the arrays the reference reads them from, and every
    array is checked for its shape, element type and device
    (check_{1,2,3}d_tensor + unwrap); a failed check raises PanicException,
    the BaseException pyo3 raises for a Rust panic;
  - ROCm device arrays (DLPack device kDLROCM = 10; torch's "cuda" device on
    ROCm) run the HIP kernels.  Host arrays raise NotImplementedError: this
    build has no CPU path (the reference runs del-msh-cpu there).
  - functions return None, except edge2vtx_contour_for_triangle_mesh,
    vtx2elem_from_uniform_mesh and vtx2vtx_from_uniform_mesh, which return
    DLPack capsules that own new device memory, and
    array1d_has_duplicate_sorted_array, which returns a bool.
"""

from __future__ import annotations

import torch

from . import _lib
from .util_torch import as_tensor


class PanicException(BaseException):
    """A failed argument check (the reference's unwrap() panics surface in
    Python as pyo3_runtime.PanicException, a BaseException)."""


def get_cuda_driver_version():
    """(u32::MAX, u32::MAX): the value the reference returns when it is built
    without CUDA, which this build is (del-msh-dlpack/src/lib.rs:59-74)."""
    return (0xFFFFFFFF, 0xFFFFFFFF)


def get_ptx_compiler_version():
    """(-1, -1, -1): no PTX compiler in a HIP build (lib.rs:76-88)."""
    return (-1, -1, -1)


# ---------------------------------------------------------------------------
# checks
# ---------------------------------------------------------------------------
def _get(x) -> torch.Tensor:
    return as_tensor(x)


def _device(t: torch.Tensor) -> torch.device:
    """The arrays' device; host arrays have no implementation here."""
    if t.device.type == "cuda":
        return t.device
    raise NotImplementedError(
        f"del_msh_dlpack (MI355X build) runs on ROCm device arrays only, got {t.device}")


def _chk(t: torch.Tensor, shape, dtype, device, name: str) -> None:
    if (tuple(t.shape) != tuple(shape) or t.dtype != dtype or t.device != device
            or not t.is_contiguous()):
        raise PanicException(
            f"{name}: expected a contiguous {tuple(shape)} {dtype} array on {device}, got "
            f"{tuple(t.shape)} {t.dtype} on {t.device}")


def _dim(t: torch.Tensor, axis: int, name: str) -> int:
    if t.dim() <= axis:
        raise PanicException(f"{name}: no axis {axis} in a {t.dim()}-d array")
    return int(t.shape[axis])


def _stream(stream_ptr: int, device: torch.device) -> int:
    torch.cuda.set_device(device)
    return int(stream_ptr) or _lib.stream_of(device)


u32, f32 = torch.uint32, torch.float32


# ---------------------------------------------------------------------------
# Morton codes, sort, topology, refit (mortons.rs, array1d.rs, bvhnode2aabb.rs)
# ---------------------------------------------------------------------------
def mortons_vtx2morton_from_vtx2co(vtx2co, transform_co2unit, vtx2morton, stream_ptr=0):
    """vtx2morton [n] u32 from vtx2co [n, 2|3] f32 through the column-major
    (num_dim+1)^2 transform (mortons.rs:17-77)."""
    vtx2co, transform_co2unit, vtx2morton = map(_get, (vtx2co, transform_co2unit, vtx2morton))
    dev = _device(vtx2co)
    n, num_dim = _dim(vtx2co, 0, "vtx2co"), _dim(vtx2co, 1, "vtx2co")
    _chk(vtx2co, (n, num_dim), f32, dev, "vtx2co")
    _chk(transform_co2unit, ((num_dim + 1) * (num_dim + 1),), f32, dev, "transform_co2unit")
    _chk(vtx2morton, (n,), u32, dev, "vtx2morton")
    if num_dim not in (2, 3):
        raise NotImplementedError(f"num_dim {num_dim}")
    from .Mortons import make_vtx2morton_from_vtx2co

    make_vtx2morton_from_vtx2co(vtx2co, transform_co2unit, vtx2morton, _stream(stream_ptr, dev))


def mortons_make_bvhnodes_from_sorted_mortons(idx2obj, idx2morton, bhvnodes, stream_ptr=0):
    """bhvnodes [2n-1, 3] u32 from the sorted codes (mortons.rs:79-135)."""
    idx2obj, idx2morton, bhvnodes = map(_get, (idx2obj, idx2morton, bhvnodes))
    dev = _device(idx2obj)
    n = _dim(idx2obj, 0, "idx2obj")
    _chk(idx2obj, (n,), u32, dev, "idx2obj")
    _chk(idx2morton, (n,), u32, dev, "idx2morton")
    _chk(bhvnodes, (2 * n - 1, 3), u32, dev, "bhvnodes")
    from .Mortons import make_bvhnodes_from_sorted_mortons

    make_bvhnodes_from_sorted_mortons(idx2obj, idx2morton, bhvnodes, _stream(stream_ptr, dev))


def array1d_permute(old2val, new2old, new2val, stream_ptr=0):
    """new2val[i] = old2val[new2old[i]] (array1d.rs:17-70)."""
    old2val, new2old, new2val = map(_get, (old2val, new2old, new2val))
    dev = _device(old2val)
    n = _dim(old2val, 0, "old2val")
    for t, name in ((old2val, "old2val"), (new2old, "new2old"), (new2val, "new2val")):
        _chk(t, (n,), u32, dev, name)
    _lib.check(_lib.lib().dmsh_permute_u32(_stream(stream_ptr, dev), n, _lib.ptr(old2val),
                                           _lib.ptr(new2old), _lib.ptr(new2val)),
               "dmsh_permute_u32")


def array1d_argsort(idx2val, jdx2idx, stream_ptr=0):
    """Sorts idx2val in place and writes the permutation into jdx2idx
    (array1d.rs:99-154).  Ties keep input order (the stable order of
    del-msh-cpu's mortons.rs:143)."""
    idx2val, jdx2idx = map(_get, (idx2val, jdx2idx))
    dev = _device(jdx2idx)
    n = _dim(jdx2idx, 0, "jdx2idx")
    _chk(jdx2idx, (n,), u32, dev, "jdx2idx")
    _chk(idx2val, (n,), u32, dev, "idx2val")
    from .Array1D import argsort

    argsort(idx2val, jdx2idx, _stream(stream_ptr, dev))


def array1d_has_duplicate_sorted_array(idx2val, stream_ptr=0) -> bool:
    """True if two neighbours of the sorted array are equal (array1d.rs:156-193)."""
    idx2val = _get(idx2val)
    dev = _device(idx2val)
    _chk(idx2val, (_dim(idx2val, 0, "idx2val"),), u32, dev, "idx2val")
    from .Array1D import has_duplicate_sorted_array

    return has_duplicate_sorted_array(idx2val, _stream(stream_ptr, dev))


def array1d_unique_for_sorted_array(idx2val, idx2jdx, stream_ptr=0):
    """idx2jdx[i] = number of value changes in idx2val[0..=i] (array1d.rs:195-240)."""
    idx2val, idx2jdx = map(_get, (idx2val, idx2jdx))
    dev = _device(idx2val)
    n = _dim(idx2val, 0, "idx2val")
    _chk(idx2val, (n,), u32, dev, "idx2val")
    _chk(idx2jdx, (n,), u32, dev, "idx2jdx")
    from .Array1D import unique_for_sorted_array

    unique_for_sorted_array(idx2val, idx2jdx, _stream(stream_ptr, dev))


def array1d_unique_jdx2val_jdx2idx(idx2val, idx2jdx, jdx2val, jdx2idx_offset, stream_ptr=0):
    """jdx2val [num_jdx] and jdx2idx_offset [num_jdx + 1] of the groups (array1d.rs:242-300)."""
    idx2val, idx2jdx, jdx2val, jdx2idx_offset = map(
        _get, (idx2val, idx2jdx, jdx2val, jdx2idx_offset))
    dev = _device(idx2val)
    n, num_jdx = _dim(idx2val, 0, "idx2val"), _dim(jdx2val, 0, "jdx2val")
    _chk(idx2val, (n,), u32, dev, "idx2val")
    _chk(idx2jdx, (n,), u32, dev, "idx2jdx")
    _chk(jdx2val, (num_jdx,), u32, dev, "jdx2val")
    _chk(jdx2idx_offset, (num_jdx + 1,), u32, dev, "jdx2idx_offset")
    from .Array1D import unique_jdx2val_jdx2idx

    unique_jdx2val_jdx2idx(idx2val, idx2jdx, jdx2val, jdx2idx_offset, _stream(stream_ptr, dev))


def bvhnode2aabb_update_aabb(bvhnode2aabb, i_bvhnode, bvhnodes, elem2vtx, vtx2xyz0, vtx2xyz1,
                             stream_ptr=0):
    """Refit of bvhnode2aabb [2n-1, 6] for triangle elements with the
    reference GPU path's semantics (bvhnode2aabb.rs:71-102, kernel
    from_trimesh3): the whole tree is refit from vtx2xyz0 with margin 0;
    i_bvhnode (the CPU path's subtree root) and vtx2xyz1 (the CPU path's
    swept box) are checked and not used."""
    bvhnode2aabb, bvhnodes, elem2vtx, vtx2xyz0, vtx2xyz1 = map(
        _get, (bvhnode2aabb, bvhnodes, elem2vtx, vtx2xyz0, vtx2xyz1))
    dev = _device(bvhnode2aabb)
    num_vtx0, num_dim = _dim(vtx2xyz0, 0, "vtx2xyz0"), _dim(vtx2xyz0, 1, "vtx2xyz0")
    num_vtx1 = _dim(vtx2xyz1, 0, "vtx2xyz1")
    num_bvhnode = _dim(bvhnode2aabb, 0, "bvhnode2aabb")
    num_elem, num_noel = _dim(elem2vtx, 0, "elem2vtx"), _dim(elem2vtx, 1, "elem2vtx")
    _chk(bvhnode2aabb, (num_bvhnode, num_dim * 2), f32, dev, "bvhnode2aabb")
    _chk(bvhnodes, (num_bvhnode, 3), u32, dev, "bvhnodes")
    _chk(vtx2xyz0, (num_vtx0, num_dim), f32, dev, "vtx2xyz0")
    _chk(vtx2xyz1, (num_vtx1, num_dim), f32, dev, "vtx2xyz1")
    _chk(elem2vtx, (num_elem, num_noel), u32, dev, "elem2vtx")
    if int(i_bvhnode) >= num_bvhnode:
        raise PanicException(f"i_bvhnode {i_bvhnode} >= {num_bvhnode}")
    if num_dim != 3 or num_noel != 3:
        raise NotImplementedError("triangle elements in 3D only, as the reference GPU kernel")
    from .TriMesh3 import make_bvhnode2aabb_from_bvhnodes

    make_bvhnode2aabb_from_bvhnodes(elem2vtx, vtx2xyz0, vtx2xyz1, bvhnodes, bvhnode2aabb,
                                    _stream(stream_ptr, dev))


# ---------------------------------------------------------------------------
# ray-cast, depth, interpolation (pix2tri.rs, pix2depth.rs)
# ---------------------------------------------------------------------------
def pix2tri_by_raycast(pix2tri, tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world,
                       stream_ptr=0):
    """Fills pix2tri [H, W] u32 in place (pix2tri.rs:15-94)."""
    pix2tri, tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world = map(
        _get, (pix2tri, tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world))
    dev = _device(pix2tri)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    num_bvhnode = _dim(bvhnodes, 0, "bvhnodes")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(bvhnodes, (num_bvhnode, 3), u32, dev, "bvhnodes")
    _chk(bvhnode2aabb, (num_bvhnode, 6), f32, dev, "bvhnode2aabb")
    _chk(transform_ndc2world, (16,), f32, dev, "transform_ndc2world")
    from .Pix2Tri import update_pix2tri

    update_pix2tri(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, transform_ndc2world, pix2tri,
                   _stream(stream_ptr, dev))


def pix2tri_interpolate(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val,
                        stream_ptr=0):
    """pix2val [H, W, nv] from the vertex values at each pixel's hit (pix2tri.rs:98-178)."""
    pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val = map(
        _get, (pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val))
    dev = _device(pix2tri)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    nv = _dim(vtx2val, 1, "vtx2val")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(vtx2val, (num_vtx, nv), f32, dev, "vtx2val")
    _chk(transform_ndc2world, (16,), f32, dev, "transform_ndc2world")
    _chk(pix2val, (img_h, img_w, nv), f32, dev, "pix2val")
    from .Pix2Tri import interpolate

    interpolate(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, pix2val,
                _stream(stream_ptr, dev))


def pix2tri_interpolate_bwd(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world,
                            dldw_pix2val, dldw_vtx2xyz, dldw_vtx2val, stream_ptr=0):
    """Accumulates dL/d(vertex positions) and dL/d(vertex values) (pix2tri.rs:180-271)."""
    args = tuple(map(_get, (pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world,
                            dldw_pix2val, dldw_vtx2xyz, dldw_vtx2val)))
    pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val, dxyz, dval = args
    dev = _device(pix2tri)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    nv = _dim(vtx2val, 1, "vtx2val")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(vtx2val, (num_vtx, nv), f32, dev, "vtx2val")
    _chk(transform_ndc2world, (16,), f32, dev, "transform_ndc2world")
    _chk(dldw_pix2val, (img_h, img_w, nv), f32, dev, "dldw_pix2val")
    _chk(dxyz, (num_vtx, 3), f32, dev, "dldw_vtx2xyz")
    _chk(dval, (num_vtx, nv), f32, dev, "dldw_vtx2val")
    from .Pix2Tri import interpolate_bwd

    interpolate_bwd(pix2tri, tri2vtx, vtx2xyz, vtx2val, transform_ndc2world, dldw_pix2val, dxyz,
                    dval, _stream(stream_ptr, dev))


def pix2depth_update(pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world, stream_ptr=0):
    """pix2depth [H, W] f32, (ndc_z + 1) / 2 of each hit, 0 on background
    (pix2depth.rs:14-97; the inverse transform is taken on the device)."""
    pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world = map(
        _get, (pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world))
    dev = _device(pix2depth)
    img_h, img_w = _dim(pix2depth, 0, "pix2depth"), _dim(pix2depth, 1, "pix2depth")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    _chk(pix2depth, (img_h, img_w), f32, dev, "pix2depth")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(transform_ndc2world, (16,), f32, dev, "transform_ndc2world")
    from .Pix2Depth import update

    update(pix2depth, pix2tri, tri2vtx, vtx2xyz, transform_ndc2world, _stream(stream_ptr, dev))


def pix2depth_bwd_wrt_vtx2xyz(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth,
                              transform_ndc2world, stream_ptr=0):
    """Accumulates dL/d(vertex positions) of the depth image (pix2depth.rs:101-190)."""
    dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world = map(
        _get, (dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world))
    dev = _device(dldw_vtx2xyz)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    _chk(dldw_vtx2xyz, (num_vtx, 3), f32, dev, "dldw_vtx2xyz")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(dldw_pix2depth, (img_h, img_w), f32, dev, "dldw_pix2depth")
    _chk(transform_ndc2world, (16,), f32, dev, "transform_ndc2world")
    from .Pix2Depth import bwd_wrt_vtx2xyz

    bwd_wrt_vtx2xyz(dldw_vtx2xyz, pix2tri, tri2vtx, vtx2xyz, dldw_pix2depth, transform_ndc2world,
                    _stream(stream_ptr, dev))


# ---------------------------------------------------------------------------
# silhouette gradients (rasterized_edge_gradient.rs, differentiable_antialias.rs)
# ---------------------------------------------------------------------------
def _check_edge_gradient_inputs(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val,
                                dldw_pix2val, dev):
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    nv = _dim(pix2val, 2, "pix2val")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(transform_world2pix, (16,), f32, dev, "transform_world2pix")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(pix2val, (img_h, img_w, nv), f32, dev, "pix2val")
    _chk(dldw_pix2val, (img_h, img_w, nv), f32, dev, "dldw_pix2val")
    return img_h, img_w, num_vtx


def rasterized_edge_gradient_bwd(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri,
                                 pix2val, dldw_pix2val, stream_ptr=0):
    """Adds the silhouette-edge gradient to dldw_vtx2xyz
    (rasterized_edge_gradient.rs:27-150), each vertex's terms in
    del-msh-cpu's loop order (bit-exact with its bwd)."""
    args = tuple(map(_get, (tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri,
                            pix2val, dldw_pix2val)))
    tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val = args
    dev = _device(tri2vtx)
    _, _, num_vtx = _check_edge_gradient_inputs(tri2vtx, vtx2xyz, transform_world2pix, pix2tri,
                                                pix2val, dldw_pix2val, dev)
    _chk(dldw_vtx2xyz, (num_vtx, 3), f32, dev, "dldw_vtx2xyz")
    from .RasterizedEdgeGradient import bwd

    bwd(tri2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pix2val,
        _stream(stream_ptr, dev))


def rasterized_edge_gradient_edge_gradient_and_type(tri2vtx, vtx2xyz, transform_world2pix,
                                                    pix2tri, pix2val, dldw_pixval, hedge2type,
                                                    hedge2dldr, vedge2type, vedge2dldr,
                                                    stream_ptr=0):
    """Pixel-edge types and gradients (rasterized_edge_gradient.rs:152-290)."""
    args = tuple(map(_get, (tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pixval,
                            hedge2type, hedge2dldr, vedge2type, vedge2dldr)))
    tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pixval, ht, hd, vt, vd = args
    dev = _device(tri2vtx)
    img_h, img_w, _ = _check_edge_gradient_inputs(tri2vtx, vtx2xyz, transform_world2pix, pix2tri,
                                                  pix2val, dldw_pixval, dev)
    _chk(ht, (img_h - 1, img_w), torch.uint8, dev, "hedge2type")
    _chk(hd, (img_h - 1, img_w), f32, dev, "hedge2dldr")
    _chk(vt, (img_h, img_w - 1), torch.uint8, dev, "vedge2type")
    _chk(vd, (img_h, img_w - 1), f32, dev, "vedge2dldr")
    from .RasterizedEdgeGradient import edge_gradient_and_type

    edge_gradient_and_type(tri2vtx, vtx2xyz, transform_world2pix, pix2tri, pix2val, dldw_pixval,
                           ht, hd, vt, vd, _stream(stream_ptr, dev))


def rasterized_edge_gradient_smooth_gradient(hedge2type, hedge2dldr, vedge2type, vedge2dldr,
                                             num_iter, stream_ptr=0):
    """num_iter red-black sweeps of both pixel-edge grids, in place
    (rasterized_edge_gradient.rs:292-392)."""
    ht, hd, vt, vd = map(_get, (hedge2type, hedge2dldr, vedge2type, vedge2dldr))
    dev = _device(ht)
    img_h, img_w = _dim(ht, 0, "hedge2type") + 1, _dim(ht, 1, "hedge2type")
    _chk(ht, (img_h - 1, img_w), torch.uint8, dev, "hedge2type")
    _chk(hd, (img_h - 1, img_w), f32, dev, "hedge2dldr")
    _chk(vt, (img_h, img_w - 1), torch.uint8, dev, "vedge2type")
    _chk(vd, (img_h, img_w - 1), f32, dev, "vedge2dldr")
    from .RasterizedEdgeGradient import smooth_gradient

    smooth_gradient(ht, hd, vt, vd, int(num_iter), _stream(stream_ptr, dev))


def rasterized_edge_gradient_interpolate(hedge2vy, vedge2vx, vtx2xy, vtx2velo, stream_ptr=0):
    """Bilinear read of the pixel-edge grids at points (rasterized_edge_gradient.rs:394-450)."""
    hv, vv, xy, velo = map(_get, (hedge2vy, vedge2vx, vtx2xy, vtx2velo))
    dev = _device(hv)
    img_h, img_w = _dim(hv, 0, "hedge2vy") + 1, _dim(hv, 1, "hedge2vy")
    num_vtx = _dim(xy, 0, "vtx2xy")
    _chk(hv, (img_h - 1, img_w), f32, dev, "hedge2vy")
    _chk(vv, (img_h, img_w - 1), f32, dev, "vedge2vx")
    _chk(xy, (num_vtx, 2), f32, dev, "vtx2xy")
    _chk(velo, (num_vtx, 2), f32, dev, "vtx2velo")
    from .RasterizedEdgeGradient import interpolate

    interpolate(hv, vv, xy, velo, _stream(stream_ptr, dev))


def differentiable_antialias_fwd(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pic2vin,
                                 pic2vout, stream_ptr=0):
    """Anti-aliased image along the contour edges (differentiable_antialias.rs:103-182)."""
    cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pic2vin, pic2vout = map(
        _get, (cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pic2vin, pic2vout))
    dev = _device(cedge2vtx)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_cedge, num_vtx = _dim(cedge2vtx, 0, "cedge2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    _chk(cedge2vtx, (num_cedge, 2), u32, dev, "cedge2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(transform_world2pix, (16,), f32, dev, "transform_world2pix")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    _chk(pic2vin, (img_h, img_w), f32, dev, "pic2vin")
    _chk(pic2vout, (img_h, img_w), f32, dev, "pic2vout")
    from .DifferentiableAntialias import antialias

    antialias(cedge2vtx, vtx2xyz, transform_world2pix, pix2tri, pic2vin, pic2vout,
              _stream(stream_ptr, dev))


def differentiable_antialias_bwd(cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val,
                                 dldw_pixval, pix2tri, stream_ptr=0):
    """Accumulates the antialias backward into dldw_vtx2xyz
    (differentiable_antialias.rs:16-101)."""
    args = tuple(map(_get, (cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val,
                            dldw_pixval, pix2tri)))
    cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val, dldw_pixval, pix2tri = args
    dev = _device(cedge2vtx)
    img_h, img_w = _dim(pix2tri, 0, "pix2tri"), _dim(pix2tri, 1, "pix2tri")
    num_cedge, num_vtx = _dim(cedge2vtx, 0, "cedge2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    _chk(cedge2vtx, (num_cedge, 2), u32, dev, "cedge2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(dldw_vtx2xyz, (num_vtx, 3), f32, dev, "dldw_vtx2xyz")
    _chk(transform_world2pix, (16,), f32, dev, "transform_world2pix")
    _chk(pix2val, (img_h, img_w), f32, dev, "pix2val")
    _chk(dldw_pixval, (img_h, img_w), f32, dev, "dldw_pixval")
    _chk(pix2tri, (img_h, img_w), u32, dev, "pix2tri")
    from .DifferentiableAntialias import bwd_antialias

    bwd_antialias(cedge2vtx, vtx2xyz, dldw_vtx2xyz, transform_world2pix, pix2val, dldw_pixval,
                  pix2tri, _stream(stream_ptr, dev))


# ---------------------------------------------------------------------------
# mesh topology (edge2vtx.rs, edge2elem.rs, vtx2elem.rs, vtx2vtx.rs)
# ---------------------------------------------------------------------------
def edge2vtx_contour_for_triangle_mesh(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri,
                                       stream_ptr=0):
    """Contour edges as a new (num_contour, 2) u32 device array, returned as
    a DLPack capsule that owns it (edge2vtx.rs:19-115)."""
    tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri = map(
        _get, (tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri))
    dev = _device(tri2vtx)
    num_tri, num_vtx = _dim(tri2vtx, 0, "tri2vtx"), _dim(vtx2xyz, 0, "vtx2xyz")
    num_edge = _dim(edge2vtx, 0, "edge2vtx")
    _chk(tri2vtx, (num_tri, 3), u32, dev, "tri2vtx")
    _chk(vtx2xyz, (num_vtx, 3), f32, dev, "vtx2xyz")
    _chk(transform_world2ndc, (16,), f32, dev, "transform_world2ndc")
    _chk(edge2vtx, (num_edge, 2), u32, dev, "edge2vtx")
    _chk(edge2tri, (num_edge, 2), u32, dev, "edge2tri")
    from .Edge2Vtx import contour_for_triangle_mesh

    out = contour_for_triangle_mesh(tri2vtx, vtx2xyz, transform_world2ndc, edge2vtx, edge2tri,
                                    _stream(stream_ptr, dev))
    return torch.utils.dlpack.to_dlpack(out)


def edge2vtx_from_vtx2vtx(vtx2idx_offset, idx2vtx, edge2vtx, stream_ptr=0):
    """Edges (i < j) of a vertex adjacency (edge2vtx.rs:117-174)."""
    vtx2idx_offset, idx2vtx, edge2vtx = map(_get, (vtx2idx_offset, idx2vtx, edge2vtx))
    dev = _device(vtx2idx_offset)
    _chk(vtx2idx_offset, (_dim(vtx2idx_offset, 0, "vtx2idx_offset"),), u32, dev, "vtx2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(edge2vtx, (_dim(edge2vtx, 0, "edge2vtx"), 2), u32, dev, "edge2vtx")
    from .Edge2Vtx import from_vtx2vtx

    from_vtx2vtx(vtx2idx_offset, idx2vtx, edge2vtx, _stream(stream_ptr, dev))


def edge2elem_from_edge2vtx_of_tri2vtx_with_vtx2vtx(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx,
                                                    edge2tri, stream_ptr=0):
    """The triangles on both sides of each edge (edge2elem.rs:17-83)."""
    edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri = map(
        _get, (edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri))
    dev = _device(edge2vtx)
    num_edge = _dim(edge2vtx, 0, "edge2vtx")
    _chk(edge2vtx, (num_edge, 2), u32, dev, "edge2vtx")
    _chk(tri2vtx, (_dim(tri2vtx, 0, "tri2vtx"), 3), u32, dev, "tri2vtx")
    _chk(vtx2idx_offset, (_dim(vtx2idx_offset, 0, "vtx2idx_offset"),), u32, dev, "vtx2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(edge2tri, (num_edge, 2), u32, dev, "edge2tri")
    from .Edge2Elem import from_edge2vtx_of_tri2vtx_with_vtx2vtx

    from_edge2vtx_of_tri2vtx_with_vtx2vtx(edge2vtx, tri2vtx, vtx2idx_offset, idx2vtx, edge2tri,
                                          _stream(stream_ptr, dev))


def vtx2elem_from_uniform_mesh(elem2vtx, num_vtx, stream_ptr=0):
    """(vtx2idx, idx2elem) as DLPack capsules owning new device arrays (vtx2elem.rs:14-66)."""
    elem2vtx = _get(elem2vtx)
    dev = _device(elem2vtx)
    _chk(elem2vtx, (_dim(elem2vtx, 0, "elem2vtx"), _dim(elem2vtx, 1, "elem2vtx")), u32, dev,
         "elem2vtx")
    from .Vtx2Elem import from_uniform_mesh

    a, b = from_uniform_mesh(elem2vtx, int(num_vtx), _stream(stream_ptr, dev))
    return torch.utils.dlpack.to_dlpack(a), torch.utils.dlpack.to_dlpack(b)


def vtx2vtx_from_uniform_mesh(elem2vtx, num_vtx, is_self, stream_ptr=0):
    """(vtx2idx, idx2vtx) as DLPack capsules owning new device arrays (vtx2vtx.rs:193-248)."""
    elem2vtx = _get(elem2vtx)
    dev = _device(elem2vtx)
    _chk(elem2vtx, (_dim(elem2vtx, 0, "elem2vtx"), _dim(elem2vtx, 1, "elem2vtx")), u32, dev,
         "elem2vtx")
    from .Vtx2Vtx import from_uniform_mesh

    a, b = from_uniform_mesh(elem2vtx, int(num_vtx), bool(is_self), _stream(stream_ptr, dev))
    return torch.utils.dlpack.to_dlpack(a), torch.utils.dlpack.to_dlpack(b)


def vtx2vtx_laplacian_smoothing(vtx2idx, idx2vtx, lambda_, vtx2lhs, vtx2rhs, num_iter,
                                vtx2lhstmp, stream_ptr=0):
    """num_iter Jacobi steps of [I + lambda L] lhs = rhs (vtx2vtx.rs:23-120).
    vtx2lhstmp is the reference's ping-pong buffer; checked, not needed here."""
    vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp = map(
        _get, (vtx2idx, idx2vtx, vtx2lhs, vtx2rhs, vtx2lhstmp))
    dev = _device(vtx2idx)
    num_vtx = _dim(vtx2idx, 0, "vtx2idx") - 1
    nd = _dim(vtx2lhs, 1, "vtx2lhs")
    _chk(vtx2idx, (num_vtx + 1,), u32, dev, "vtx2idx")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2lhs, (num_vtx, nd), f32, dev, "vtx2lhs")
    _chk(vtx2rhs, (num_vtx, nd), f32, dev, "vtx2rhs")
    _chk(vtx2lhstmp, (num_vtx, nd), f32, dev, "vtx2lhstmp")
    from .Vtx2Vtx import laplacian_smoothing

    laplacian_smoothing(vtx2idx, idx2vtx, float(lambda_), vtx2lhs, vtx2rhs, int(num_iter),
                        _stream(stream_ptr, dev))


def vtx2vtx_multiply_graph_laplacian(vtx2idx_offset, idx2vtx, vtx2rhs, vtx2lhs, stream_ptr=0):
    """vtx2lhs = L vtx2rhs for the graph Laplacian (vtx2vtx.rs:122-190)."""
    vtx2idx_offset, idx2vtx, vtx2rhs, vtx2lhs = map(_get, (vtx2idx_offset, idx2vtx, vtx2rhs,
                                                           vtx2lhs))
    dev = _device(vtx2idx_offset)
    num_vtx = _dim(vtx2idx_offset, 0, "vtx2idx_offset") - 1
    nd = _dim(vtx2rhs, 1, "vtx2rhs")
    _chk(vtx2idx_offset, (num_vtx + 1,), u32, dev, "vtx2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2rhs, (num_vtx, nd), f32, dev, "vtx2rhs")
    _chk(vtx2lhs, (num_vtx, nd), f32, dev, "vtx2lhs")
    from .Vtx2Vtx import multiply_graph_laplacian

    multiply_graph_laplacian(vtx2idx_offset, idx2vtx, vtx2rhs, vtx2lhs, _stream(stream_ptr, dev))


# ---------------------------------------------------------------------------
# other users of the LBVH machinery (quad_oct_tree.rs, polyhedron_mesh.rs,
# offset_array.rs)
# ---------------------------------------------------------------------------
def quad_oct_tree_bnodes_and_bnode2depth_and_bnode2onode_and_idx2bnode(
        idx2morton, num_dim, max_depth, bnodes, bnode2depth, bnode2onode, idx2bnode,
        stream_ptr=0):
    """Binary radix tree with depths and octree numbering (quad_oct_tree.rs)."""
    idx2morton, bnodes, bnode2depth, bnode2onode, idx2bnode = map(
        _get, (idx2morton, bnodes, bnode2depth, bnode2onode, idx2bnode))
    dev = _device(idx2morton)
    n = _dim(idx2morton, 0, "idx2morton")
    _chk(idx2morton, (n,), u32, dev, "idx2morton")
    _chk(bnodes, (n - 1, 3), u32, dev, "bnodes")
    _chk(bnode2depth, (n - 1,), u32, dev, "bnode2depth")
    _chk(bnode2onode, (n - 1,), u32, dev, "bnode2onode")
    _chk(idx2bnode, (n,), u32, dev, "idx2bnode")
    from .QuadOctTree import bnodes_and_bnode2depth_and_bnode2onode

    bnodes_and_bnode2depth_and_bnode2onode(idx2morton, int(num_dim), int(max_depth), bnodes,
                                           bnode2depth, bnode2onode, idx2bnode,
                                           _stream(stream_ptr, dev))


def quad_oct_tree_make_tree_from_binary_radix_tree(bnodes, bnode2onode, bnode2depth, idx2bnode,
                                                   idx2morton, num_dim, max_depth, onodes,
                                                   onode2depth, onode2center, idx2onode,
                                                   idx2center, stream_ptr=0):
    """Octree nodes, depths and centres from the radix tree (quad_oct_tree.rs)."""
    args = tuple(map(_get, (bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton, onodes,
                            onode2depth, onode2center, idx2onode, idx2center)))
    bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton = args[:5]
    onodes, onode2depth, onode2center, idx2onode, idx2center = args[5:]
    dev = _device(bnodes)
    n = _dim(idx2morton, 0, "idx2morton")
    num_onode = _dim(onodes, 0, "onodes")
    nlink = 1 + 2 ** int(num_dim)
    _chk(bnodes, (n - 1, 3), u32, dev, "bnodes")
    _chk(bnode2onode, (n - 1,), u32, dev, "bnode2onode")
    _chk(bnode2depth, (n - 1,), u32, dev, "bnode2depth")
    _chk(idx2bnode, (n,), u32, dev, "idx2bnode")
    _chk(idx2morton, (n,), u32, dev, "idx2morton")
    _chk(onodes, (num_onode, nlink), u32, dev, "onodes")
    _chk(onode2depth, (num_onode,), u32, dev, "onode2depth")
    _chk(onode2center, (num_onode, int(num_dim)), f32, dev, "onode2center")
    _chk(idx2onode, (n,), u32, dev, "idx2onode")
    _chk(idx2center, (n, int(num_dim)), f32, dev, "idx2center")
    from .QuadOctTree import make_tree_from_binary_radix_tree

    make_tree_from_binary_radix_tree(bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton,
                                     int(num_dim), int(max_depth), onodes, onode2depth,
                                     onode2center, idx2onode, idx2center,
                                     _stream(stream_ptr, dev))


def quad_oct_tree_aggregate(idx2val, idx2onode, onodes, onode2aggval, stream_ptr=0):
    """Per-node sums of the point values (quad_oct_tree.rs)."""
    idx2val, idx2onode, onodes, onode2aggval = map(_get, (idx2val, idx2onode, onodes,
                                                          onode2aggval))
    dev = _device(idx2val)
    n, nv = _dim(idx2val, 0, "idx2val"), _dim(idx2val, 1, "idx2val")
    num_onode = _dim(onodes, 0, "onodes")
    _chk(idx2val, (n, nv), f32, dev, "idx2val")
    _chk(idx2onode, (n,), u32, dev, "idx2onode")
    _chk(onodes, (num_onode, _dim(onodes, 1, "onodes")), u32, dev, "onodes")
    _chk(onode2aggval, (num_onode, nv), f32, dev, "onode2aggval")
    from .QuadOctTree import aggregate

    aggregate(idx2val, idx2onode, onodes, onode2aggval, _stream(stream_ptr, dev))


def offset_array_aggregate(idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval, stream_ptr=0):
    """idx2aggval[i] += sum of kdx2val[jdx2kdx[j]] over j in the offsets' range (offset_array.rs)."""
    idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval = map(_get, (idx2jdx_offset, jdx2kdx, kdx2val,
                                                              idx2aggval))
    dev = _device(idx2jdx_offset)
    num_idx = _dim(idx2jdx_offset, 0, "idx2jdx_offset") - 1
    nd = _dim(kdx2val, 1, "kdx2val")
    _chk(idx2jdx_offset, (num_idx + 1,), u32, dev, "idx2jdx_offset")
    _chk(jdx2kdx, (_dim(jdx2kdx, 0, "jdx2kdx"),), u32, dev, "jdx2kdx")
    _chk(kdx2val, (_dim(kdx2val, 0, "kdx2val"), nd), f32, dev, "kdx2val")
    _chk(idx2aggval, (num_idx, nd), f32, dev, "idx2aggval")
    from .OffsetArray import aggregate

    aggregate(idx2jdx_offset, jdx2kdx, kdx2val, idx2aggval, _stream(stream_ptr, dev))


def polyhedron_mesh_elem2center(elem2idx_offset, idx2vtx, vtx2xyz, elem2center, stream_ptr=0):
    """Element centroids of a CSR polyhedron mesh (polyhedron_mesh.rs)."""
    elem2idx_offset, idx2vtx, vtx2xyz, elem2center = map(_get, (elem2idx_offset, idx2vtx,
                                                                vtx2xyz, elem2center))
    dev = _device(elem2idx_offset)
    num_elem = _dim(elem2idx_offset, 0, "elem2idx_offset") - 1
    _chk(elem2idx_offset, (num_elem + 1,), u32, dev, "elem2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2xyz, (_dim(vtx2xyz, 0, "vtx2xyz"), 3), f32, dev, "vtx2xyz")
    _chk(elem2center, (num_elem, 3), f32, dev, "elem2center")
    from .PolyhedronMesh import elem2center as native

    native(elem2idx_offset, idx2vtx, vtx2xyz, elem2center, _stream(stream_ptr, dev))


def polyhedron_mesh_bvhnode2aabb_from_bvhnodes(elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes,
                                               bvhnode2aabb, stream_ptr=0):
    """Element-box refit of a polyhedron mesh's LBVH (polyhedron_mesh.rs)."""
    elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb = map(
        _get, (elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb))
    dev = _device(elem2idx_offset)
    num_elem = _dim(elem2idx_offset, 0, "elem2idx_offset") - 1
    _chk(elem2idx_offset, (num_elem + 1,), u32, dev, "elem2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2xyz, (_dim(vtx2xyz, 0, "vtx2xyz"), 3), f32, dev, "vtx2xyz")
    _chk(bvhnodes, (2 * num_elem - 1, 3), u32, dev, "bvhnodes")
    _chk(bvhnode2aabb, (2 * num_elem - 1, 6), f32, dev, "bvhnode2aabb")
    from .PolyhedronMesh import bvhnode2aabb_from_bvhnodes

    bvhnode2aabb_from_bvhnodes(elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb,
                               _stream(stream_ptr, dev))


def polyhedron_mesh_search_elem_contain_points(bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx,
                                               vtx2xyz, wtx2xyz, wtx2elem, wtx2param,
                                               stream_ptr=0):
    """The element containing each query point and its parametric
    coordinates (polyhedron_mesh.rs)."""
    args = tuple(map(_get, (bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz, wtx2xyz,
                            wtx2elem, wtx2param)))
    bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz, wtx2xyz, wtx2elem, wtx2param = args
    dev = _device(bvhnodes)
    num_elem = _dim(elem2idx_offset, 0, "elem2idx_offset") - 1
    num_wtx = _dim(wtx2xyz, 0, "wtx2xyz")
    _chk(bvhnodes, (2 * num_elem - 1, 3), u32, dev, "bvhnodes")
    _chk(bvhnode2aabb, (2 * num_elem - 1, 6), f32, dev, "bvhnode2aabb")
    _chk(elem2idx_offset, (num_elem + 1,), u32, dev, "elem2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2xyz, (_dim(vtx2xyz, 0, "vtx2xyz"), 3), f32, dev, "vtx2xyz")
    _chk(wtx2xyz, (num_wtx, 3), f32, dev, "wtx2xyz")
    _chk(wtx2elem, (num_wtx,), u32, dev, "wtx2elem")
    _chk(wtx2param, (num_wtx, 3), f32, dev, "wtx2param")
    from .PolyhedronMesh import search_elem_contain_points

    search_elem_contain_points(bvhnodes, bvhnode2aabb, elem2idx_offset, idx2vtx, vtx2xyz,
                               wtx2xyz, wtx2elem, wtx2param, _stream(stream_ptr, dev))


def polyhedron_mesh_interpolate_values_at_points(elem2idx_offset, idx2vtx, vtx2value, wtx2elem,
                                                 wtx2param, wtx2value, stream_ptr=0):
    """Shape-function interpolation of vertex values at located points (polyhedron_mesh.rs)."""
    args = tuple(map(_get, (elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param, wtx2value)))
    elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param, wtx2value = args
    dev = _device(elem2idx_offset)
    num_wtx = _dim(wtx2elem, 0, "wtx2elem")
    nv = _dim(vtx2value, 1, "vtx2value")
    _chk(elem2idx_offset, (_dim(elem2idx_offset, 0, "elem2idx_offset"),), u32, dev,
         "elem2idx_offset")
    _chk(idx2vtx, (_dim(idx2vtx, 0, "idx2vtx"),), u32, dev, "idx2vtx")
    _chk(vtx2value, (_dim(vtx2value, 0, "vtx2value"), nv), f32, dev, "vtx2value")
    _chk(wtx2elem, (num_wtx,), u32, dev, "wtx2elem")
    _chk(wtx2param, (num_wtx, 3), f32, dev, "wtx2param")
    _chk(wtx2value, (num_wtx, nv), f32, dev, "wtx2value")
    from .PolyhedronMesh import interpolate_values_at_points

    interpolate_values_at_points(elem2idx_offset, idx2vtx, vtx2value, wtx2elem, wtx2param,
                                 wtx2value, _stream(stream_ptr, dev))
