"""ctypes binding of libdmsh (the gfx950 C ABI declared in include/dmsh.h).

The library is loaded from ``del-msh_amd/lib/libdmsh.so`` (override with the
``DMSH_LIBRARY`` environment variable).  There is no fallback: if the library
is missing, or a tensor is not on a ROCm device, the call raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib

import torch

_LIB = None

c_void_p = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_f32 = ctypes.c_float
c_size = ctypes.c_size_t
c_int = ctypes.c_int

# name -> (restype, argtypes)
_PROTOTYPES = {
    "dmsh_status_string": (ctypes.c_char_p, [c_int]),
    "dmsh_abi_version": (c_int, []),
    "dmsh_get_device_errors": (c_int, [ctypes.POINTER(c_u32)]),
    "dmsh_vtx2morton_u32": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_vtx2morton_u64": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dmsh_argsort_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_argsort_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_size]),
    "dmsh_argsort_u64": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_size]),
    "dmsh_bvhnodes_from_sorted_mortons_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dmsh_bvhnodes_from_sorted_mortons_u64": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dmsh_refit_workspace_bytes": (c_size, [c_u32]),
    "dmsh_bvhnode2aabb_trimesh3": (
        c_int,
        [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_f32, c_void_p, c_size],
    ),
    "dmsh_bvhnode2aabb_points": (
        c_int,
        [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_size],
    ),
    "dmsh_build_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_build_trimesh3": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_build_trimesh3_noalloc": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_build_points": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_set_tuning": (c_int, [ctypes.c_char_p, c_int]),
    "dmsh_get_tuning": (c_int, [ctypes.c_char_p]),
    "dmsh_pix2tri_workspace_bytes": (c_size, [c_u32, c_u32, c_u32]),
    "dmsh_pix2depth_workspace_bytes": (c_size, []),
    "dmsh_pix2tri": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p,
         c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_pix2depth": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p,
         c_void_p, c_size],
    ),
    "dmsh_rasterized_edge_gradient_bwd": (
        c_int,
        [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_void_p,
         c_u32, c_void_p, c_void_p],
    ),
    "dmsh_rasterized_edge_gradient_bwd_workspace_bytes": (c_size, [c_u32, c_u32, c_u32]),
    "dmsh_rasterized_edge_gradient_bwd_ordered": (
        c_int,
        [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_void_p,
         c_u32, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_pix2depth_bwd_wrt_vtx2xyz": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_u32,
         c_void_p, c_void_p, c_size],
    ),
    "dmsh_batch_set_tuning": (c_int, [ctypes.c_char_p, c_int]),
    "dmsh_batch_get_tuning": (c_int, [ctypes.c_char_p]),
    "dmsh_batch_build_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_batch_build_trimesh3": (
        c_int,
        [c_void_p, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_batch_pix2tri_workspace_bytes": (c_size, [c_u32, c_u32, c_u32]),
    "dmsh_batch_pix2tri": (
        c_int,
        [c_void_p, c_u32, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_packed_bytes": (c_size, [c_u32]),
    "dmsh_batch_pack_workspace_bytes": (c_size, [c_u32]),
    "dmsh_batch_pack": (
        c_int,
        [c_void_p, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_batch_build_pack_trimesh3": (
        c_int,
        [c_void_p, c_u32, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_cast_requests_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_cast_requests": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_timing_enable": (c_int, [c_int]),
    "dmsh_timing_query": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_u32),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "dmsh_contour_workspace_bytes": (c_size, [c_u32]),
    "dmsh_contour_for_triangle_mesh": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_size],
    ),
    "dmsh_antialias_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_antialias_fwd": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_void_p, c_void_p,
         c_void_p, c_void_p, c_size],
    ),
    "dmsh_antialias_bwd": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_u32, c_u32, c_void_p,
         c_void_p, c_void_p],
    ),
    "dmsh_interpolate": (
        c_int, [c_void_p, c_void_p, c_u32, c_u32, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p,
                c_void_p, c_void_p]),
    "dmsh_interpolate_bwd": (
        c_int, [c_void_p, c_void_p, c_u32, c_u32, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p,
                c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmsh_edge_gradient_and_type": (
        c_int, [c_void_p, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p, c_u32,
                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmsh_smooth_gradient_workspace_bytes": (c_size, [c_u32, c_u32]),
    "dmsh_smooth_gradient": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p,
                c_size]),
    "dmsh_interpolate_staggered_grid": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_smoothed_vertex_gradient": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dmsh_exclusive_scan_workspace_bytes": (c_size, [c_u32]),
    "dmsh_exclusive_scan_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_vtx2elem_workspace_bytes": (c_size, [c_u32, c_u32, c_u32]),
    "dmsh_vtx2elem_from_uniform_mesh": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_vtx2vtx_workspace_bytes": (c_size, [c_u32, c_u32, c_u32, c_int]),
    "dmsh_vtx2vtx_count": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_u32, c_int, c_void_p, c_void_p, c_size]),
    "dmsh_vtx2vtx_fill": (
        c_int, [c_void_p, c_u32, c_u32, c_u32, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_edge2vtx_from_vtx2vtx": (c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_edge2elem_from_edge2vtx_of_tri2vtx": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_laplacian_smoothing": (
        c_int,
        [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_f32, c_u32, c_void_p, c_void_p, c_u32,
         c_void_p]),
    "dmsh_multiply_graph_laplacian": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_octree_bnodes_workspace_bytes": (c_size, [c_u32]),
    "dmsh_octree_bnodes": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_size]),
    "dmsh_octree_make_tree": (
        c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_u32, c_u32,
                c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "dmsh_octree_aggregate_workspace_bytes": (c_size, [c_u32]),
    "dmsh_octree_aggregate": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_void_p, c_u32, c_u32, c_void_p, c_void_p,
                c_void_p, c_size]),
    "dmsh_permute_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p]),
    "dmsh_has_duplicate_sorted_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_unique_for_sorted_workspace_bytes": (c_size, [c_u32]),
    "dmsh_unique_for_sorted_u32": (c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_unique_jdx2val_jdx2idx": (
        c_int, [c_void_p, c_u32, c_void_p, c_void_p, c_u32, c_void_p, c_void_p]),
    "dmsh_polyhedron_elem2center": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p, c_void_p]),
    "dmsh_polyhedron_refit_workspace_bytes": (c_size, [c_u32]),
    "dmsh_bvhnode2aabb_polyhedron": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p,
                c_void_p, c_void_p, c_size]),
    "dmsh_polyhedron_search_workspace_bytes": (c_size, [c_u32]),
    "dmsh_polyhedron_search_elem_contain_points": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_void_p, c_void_p,
                c_u32, c_void_p, c_void_p, c_void_p, c_void_p, c_size]),
    "dmsh_polyhedron_interpolate_at_points": (
        c_int, [c_void_p, c_u32, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p,
                c_void_p, c_void_p, c_void_p]),
    "dmsh_offset_array_aggregate": (
        c_int, [c_void_p, c_u32, c_void_p, c_u32, c_void_p, c_u32, c_u32, c_void_p, c_void_p]),
}


def library_path() -> pathlib.Path:
    env = os.environ.get("DMSH_LIBRARY")
    if env:
        return pathlib.Path(env)
    return pathlib.Path(__file__).resolve().parent.parent / "lib" / "libdmsh.so"


def lib():
    """Loads libdmsh once and attaches the C prototypes; raises if absent."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not path.exists():
            raise RuntimeError(
                f"libdmsh not found at {path}; build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            )
        handle = ctypes.CDLL(str(path))
        for name, (restype, argtypes) in _PROTOTYPES.items():
            fn = getattr(handle, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _LIB = handle
    return _LIB


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().dmsh_status_string(status).decode()
        raise RuntimeError(f"{what} failed: {msg} (status {status})")


def device_errors() -> int:
    """Reads and clears the sticky device error bits (synchronizes)."""
    bits = c_u32(0)
    check(lib().dmsh_get_device_errors(ctypes.byref(bits)), "dmsh_get_device_errors")
    return int(bits.value)


DEVERR_NAMES = {
    0x1: "index out of range",
    0x2: "ray traversal stack overflow",
    0x4: "radix-sort look-back gave up",
    0x8: "broken tree in refit",
    0x10: "singular transform",
    0x20: "antialias pixel walk longer than 2^26 steps",
}


def timing_enable(on: bool) -> None:
    """Start (clearing earlier records) or stop per-kernel HIP-event timing."""
    check(lib().dmsh_timing_enable(1 if on else 0), "dmsh_timing_enable")


def timing_query(label: str):
    """(launches, total_ms, max_ms) recorded for one kernel name."""
    n = c_u32(0)
    total = ctypes.c_double(0.0)
    mx = ctypes.c_double(0.0)
    check(lib().dmsh_timing_query(label.encode(), ctypes.byref(n), ctypes.byref(total),
                                  ctypes.byref(mx)), "dmsh_timing_query")
    return int(n.value), float(total.value), float(mx.value)


def raise_on_device_errors() -> None:
    bits = device_errors()
    if bits:
        names = [v for k, v in DEVERR_NAMES.items() if bits & k]
        raise RuntimeError(f"libdmsh device error bits 0x{bits:x}: {', '.join(names)}")


def require_gpu(*tensors: torch.Tensor) -> torch.device:
    """All tensors must live on one ROCm device; returns it and makes it current."""
    dev = tensors[0].device
    for t in tensors:
        if t.device.type != "cuda":
            raise NotImplementedError(
                "del_msh_dlpack (MI355X build) runs on ROCm device tensors only; "
                f"got a tensor on {t.device}"
            )
        if t.device != dev:
            raise ValueError(f"tensors on different devices: {dev} vs {t.device}")
    torch.cuda.set_device(dev)
    return dev


def check_transform(t: torch.Tensor, name: str, numel: int = 16) -> None:
    """A transform crossing the C ABI is `numel` contiguous float32 values."""
    if t.dtype != torch.float32 or t.numel() != numel or not t.is_contiguous():
        raise ValueError(f"{name} must be {numel} contiguous float32 values, got "
                         f"{tuple(t.shape)} {t.dtype}")


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    assert t.is_contiguous()
    return t.data_ptr()


def workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=device)


_KEPT = {}


def kept_workspace(entry: str, nbytes: int, device: torch.device, stream: int) -> torch.Tensor:
    """A workspace kept for `entry` on (device, stream) and handed to its later
    calls on that stream (which run in order).  The standalone refits keep the
    tree's plan in it, checked against a hash of the rows on every call."""
    key = (entry, torch.device(device).index, int(stream))
    ws = _KEPT.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = workspace(nbytes, device)
        _KEPT[key] = ws
    return ws
