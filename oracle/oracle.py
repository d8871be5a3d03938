"""numpy front end of the CPU oracle (liboracle.so, built from dmsh_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates del-msh-cpu's hot-path
functions in C; only tests/, __graft_entry__.smoke() and the CPU-baseline leg
of bench.py may use it, and only as the checker.
"""

from __future__ import annotations

import ctypes
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_L = None

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t


class TraceStats(ctypes.Structure):
    _fields_ = [("nodes_visited", ctypes.c_uint64), ("tris_tested", ctypes.c_uint64),
                ("frames_hist", ctypes.c_uint64 * 64)]


def lib():
    global _L
    if _L is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        _L = load(path)
    return _L


def load(path):
    """The oracle library at `path` (the default build, or a native build for
    the CPU baseline) with its prototypes attached."""
    L = ctypes.CDLL(str(path))
    sig = {
        "orc_expand_bits2": (ctypes.c_uint32, [ctypes.c_uint32]),
        "orc_expand_bits3": (ctypes.c_uint32, [ctypes.c_uint32]),
        "orc_expand_bits3_u64": (ctypes.c_uint64, [ctypes.c_uint32]),
        "orc_morton_code2": (ctypes.c_uint32, [ctypes.c_float] * 2),
        "orc_morton_code3": (ctypes.c_uint32, [ctypes.c_float] * 3),
        "orc_morton_code3_u64": (ctypes.c_uint64, [ctypes.c_float] * 3),
        "orc_vtx2morton": (ctypes.c_int, [_sz, _f32p, ctypes.c_int, _f32p, _u32p]),
        "orc_vtx2morton_u64": (ctypes.c_int, [_sz, _f32p, _f32p, _u64p]),
        "orc_argsort_u32": (None, [_sz, _u32p, _u32p, _u32p]),
        "orc_argsort_u64": (None, [_sz, _u64p, _u32p, _u64p]),
        "orc_split_u32": (_sz, [_u32p, _sz, _sz]),
        "orc_bvhnodes_u32": (ctypes.c_int, [_u32p, _u32p, _u32p, _sz]),
        "orc_bvhnodes_u64": (ctypes.c_int, [_u32p, _u32p, _u64p, _sz]),
        "orc_check_morton_code_range_split_u32": (ctypes.c_int, [_u32p, _sz]),
        "orc_check_binary_radix_tree_u32": (ctypes.c_int, [_u32p, _u32p, _sz]),
        "orc_check_bvh_topology": (ctypes.c_int, [_u32p, _sz]),
        "orc_bvhnode2aabb_trimesh3": (ctypes.c_int, [_f32p, _u32p, _sz, _u32p, _f32p, ctypes.c_float]),
        "orc_bvhnode2aabb_points": (ctypes.c_int, [_f32p, _u32p, _sz, _f32p]),
        "orc_transform_homogeneous": (ctypes.c_int, [_f32p, _f32p, _f32p]),
        "orc_ray_for_pixel": (None, [ctypes.c_uint32] * 4 + [_f32p, _f32p, _f32p]),
        "orc_ray_aabb": (ctypes.c_int, [_f32p, _f32p, _f32p, ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float)]),
        "orc_ray_tri": (ctypes.c_int, [_f32p, _f32p, _f32p, _f32p, _f32p,
                                       ctypes.POINTER(ctypes.c_float), _f32p]),
        "orc_line_tri": (ctypes.c_int, [_f32p, _f32p, _f32p, _f32p, _f32p,
                                        ctypes.POINTER(ctypes.c_float), _f32p]),
        "orc_mat4_inverse": (ctypes.c_int, [_f32p, _f32p]),
        "orc_tri2centroid": (None, [_sz, _u32p, _f32p, _f32p]),
        "orc_fit_into_unit_cube": (None, [_sz, _f32p, _f32p]),
        "orc_first_intersection_ray": (
            ctypes.c_int,
            [_f32p, _f32p, _u32p, _f32p, _u32p, _sz, _f32p, ctypes.POINTER(ctypes.c_float),
             ctypes.POINTER(ctypes.c_uint32), _f32p, ctypes.POINTER(TraceStats)]),
        "orc_pix2tri": (
            ctypes.c_int,
            [_u32p, ctypes.c_void_p, _sz, _u32p, _f32p, _u32p, _f32p, ctypes.c_uint32,
             ctypes.c_uint32, _f32p, ctypes.c_int, ctypes.POINTER(TraceStats)]),
        "orc_pix2tri_flat": (
            ctypes.c_int,
            [_u32p, ctypes.c_void_p, _sz, _u32p, _f32p, _u32p, _f32p, ctypes.c_uint32,
             ctypes.c_uint32, _f32p, ctypes.c_int, ctypes.POINTER(TraceStats)]),
        "orc_pix2depth": (ctypes.c_int, [_f32p, _u32p, _u32p, _f32p, ctypes.c_uint32,
                                         ctypes.c_uint32, _f32p]),
        "orc_pix2depth_bwd": (ctypes.c_int, [_f32p, _u32p, _u32p, _f32p, _f32p,
                                             ctypes.c_uint32, ctypes.c_uint32, _f32p]),
        "orc_edge_gradient_bwd": (
            None,
            [_u32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p,
             ctypes.c_uint32, _f32p, _f32p]),
        "orc_build_trimesh3_u64_mt": (ctypes.c_int, [_sz, _u32p, _f32p, ctypes.c_int, _u32p,
                                                     _f32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def build_native(out_dir) -> pathlib.Path:
    """Compiles the oracle with -O3 -march=native -ffp-contract=off for the
    host it runs on (the CPU baseline); returns the library path."""
    import subprocess

    out = pathlib.Path(out_dir) / "liboracle_native.so"
    cmd = ["gcc", "-O3", "-march=native", "-std=c11", "-fPIC", "-ffp-contract=off",
           "-fno-fast-math", "-fopenmp", "-shared", "-o", str(out),
           str(_HERE / "dmsh_oracle.c"), str(_HERE / "dmsh_cpu_mt.c"), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    return out


def build_trimesh3_mt(tri2vtx, vtx2xyz, num_threads, L=None):
    """The all-cores variant of build_trimesh3 with 63-bit codes (eps 0)."""
    t = _u32(tri2vtx)
    n = t.shape[0]
    bn = np.zeros((2 * n - 1) * 3, np.uint32)
    aabb = np.zeros((2 * n - 1) * 6, np.float32)
    rc = (L or lib()).orc_build_trimesh3_u64_mt(n, t.reshape(-1), _f32(vtx2xyz).reshape(-1),
                                                 num_threads, bn, aabb)
    assert rc == 0, rc
    return bn.reshape(-1, 3), aabb.reshape(-1, 6)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def colmajor(m_rowmajor) -> np.ndarray:
    """Row-major 4x4 (as the Python API takes it) -> 16 floats column-major."""
    return _f32(np.asarray(m_rowmajor, dtype=np.float32).T.reshape(-1))


# --- Morton / sort / topology ------------------------------------------------
def vtx2morton(vtx2co, transform_colmajor) -> np.ndarray:
    co = _f32(vtx2co)
    out = np.empty(co.shape[0], np.uint32)
    rc = lib().orc_vtx2morton(co.shape[0], co.reshape(-1), co.shape[1], _f32(transform_colmajor), out)
    assert rc == 0
    return out


def vtx2morton_u64(vtx2xyz, transform_colmajor) -> np.ndarray:
    co = _f32(vtx2xyz)
    out = np.empty(co.shape[0], np.uint64)
    lib().orc_vtx2morton_u64(co.shape[0], co.reshape(-1), _f32(transform_colmajor), out)
    return out


def argsort(keys):
    keys = np.ascontiguousarray(keys)
    n = keys.shape[0]
    idx = np.empty(n, np.uint32)
    sk = np.empty(n, keys.dtype)
    if keys.dtype == np.uint64:
        lib().orc_argsort_u64(n, keys, idx, sk)
    else:
        lib().orc_argsort_u32(n, _u32(keys), idx, sk)
    return idx, sk


def bvhnodes_from_sorted(idx2obj, idx2morton) -> np.ndarray:
    n = idx2obj.shape[0]
    bn = np.zeros((2 * n - 1) * 3, np.uint32)
    if idx2morton.dtype == np.uint64:
        rc = lib().orc_bvhnodes_u64(bn, _u32(idx2obj), np.ascontiguousarray(idx2morton), n)
    else:
        rc = lib().orc_bvhnodes_u32(bn, _u32(idx2obj), _u32(idx2morton), n)
    assert rc == 0, rc
    return bn.reshape(-1, 3)


def check_bvh_topology(bvhnodes, num_leaf) -> int:
    return lib().orc_check_bvh_topology(_u32(bvhnodes).reshape(-1), num_leaf)


def check_morton_code_range_split(sorted_codes) -> int:
    c = _u32(sorted_codes)
    return lib().orc_check_morton_code_range_split_u32(c, c.shape[0])


def check_binary_radix_tree(bvhnodes, sorted_codes) -> int:
    c = _u32(sorted_codes)
    return lib().orc_check_binary_radix_tree_u32(_u32(bvhnodes).reshape(-1), c, c.shape[0])


# --- build ---------------------------------------------------------------------
def tri2centroid(tri2vtx, vtx2xyz) -> np.ndarray:
    t = _u32(tri2vtx)
    out = np.empty((t.shape[0], 3), np.float32)
    lib().orc_tri2centroid(t.shape[0], t.reshape(-1), _f32(vtx2xyz).reshape(-1), out.reshape(-1))
    return out


def fit_into_unit_cube(pts) -> np.ndarray:
    p = _f32(pts)
    out = np.empty(16, np.float32)
    lib().orc_fit_into_unit_cube(p.shape[0], p.reshape(-1), out)
    return out


def refit_trimesh3(bvhnodes, tri2vtx, vtx2xyz, eps=0.0) -> np.ndarray:
    bn = _u32(bvhnodes).reshape(-1)
    num_node = bn.shape[0] // 3
    out = np.zeros(num_node * 6, np.float32)
    rc = lib().orc_bvhnode2aabb_trimesh3(out, bn, num_node, _u32(tri2vtx).reshape(-1),
                                         _f32(vtx2xyz).reshape(-1), eps)
    assert rc == 0, rc
    return out.reshape(-1, 6)


def refit_points(bvhnodes, vtx2xyz) -> np.ndarray:
    bn = _u32(bvhnodes).reshape(-1)
    num_node = bn.shape[0] // 3
    out = np.zeros(num_node * 6, np.float32)
    rc = lib().orc_bvhnode2aabb_points(out, bn, num_node, _f32(vtx2xyz).reshape(-1))
    assert rc == 0, rc
    return out.reshape(-1, 6)


def build_trimesh3(tri2vtx, vtx2xyz, morton_bits=30):
    """del-msh-cpu pipeline: centroid, fit, Morton, stable sort, topology, refit."""
    cntr = tri2centroid(tri2vtx, vtx2xyz)
    T = fit_into_unit_cube(cntr)
    codes = vtx2morton_u64(cntr, T) if morton_bits == 63 else vtx2morton(cntr, T)
    idx, sk = argsort(codes)
    bn = bvhnodes_from_sorted(idx, sk)
    aabb = refit_trimesh3(bn, tri2vtx, vtx2xyz)
    return bn, aabb


def build_points(pts, morton_bits=30):
    T = fit_into_unit_cube(pts)
    codes = vtx2morton_u64(pts, T) if morton_bits == 63 else vtx2morton(pts, T)
    idx, sk = argsort(codes)
    bn = bvhnodes_from_sorted(idx, sk)
    return bn, refit_points(bn, pts)


# --- ray-cast ------------------------------------------------------------------
def pix2tri(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, ndc2world_colmajor, img_w, img_h,
            num_threads=1, with_t=True):
    t = _u32(tri2vtx)
    out = np.empty(img_w * img_h, np.uint32)
    tt = np.empty(img_w * img_h, np.float32) if with_t else None
    st = TraceStats()
    lib().orc_pix2tri(out, tt.ctypes.data if with_t else None, t.shape[0], t.reshape(-1),
                      _f32(vtx2xyz).reshape(-1), _u32(bvhnodes).reshape(-1),
                      _f32(bvhnode2aabb).reshape(-1), img_w, img_h, _f32(ndc2world_colmajor),
                      num_threads, ctypes.byref(st))
    res = out.reshape(img_h, img_w)
    stats = (int(st.nodes_visited), int(st.tris_tested), np.array(st.frames_hist[:], np.int64))
    if with_t:
        return res, tt.reshape(img_h, img_w), stats
    return res, None, stats


def pix2tri_flat(tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, ndc2world_colmajor, img_w, img_h,
                 num_threads=1, with_t=True):
    """The image by the flat closest-hit search the GPU kernels run; returns
    (pix2tri, pix2t or None, (internal nodes expanded, triangles tested,
    histogram of the deepest stack per ray))."""
    t = _u32(tri2vtx)
    out = np.empty(img_w * img_h, np.uint32)
    tt = np.empty(img_w * img_h, np.float32) if with_t else None
    st = TraceStats()
    rc = lib().orc_pix2tri_flat(out, tt.ctypes.data if with_t else None, t.shape[0], t.reshape(-1),
                                _f32(vtx2xyz).reshape(-1), _u32(bvhnodes).reshape(-1),
                                _f32(bvhnode2aabb).reshape(-1), img_w, img_h,
                                _f32(ndc2world_colmajor), num_threads, ctypes.byref(st))
    if rc != 0:
        raise RuntimeError("flat search stack overflow")
    stats = (int(st.nodes_visited), int(st.tris_tested), np.array(st.frames_hist[:], np.int64))
    return (out.reshape(img_h, img_w), tt.reshape(img_h, img_w) if with_t else None, stats)


def first_intersection_ray(org, dirv, tri2vtx, vtx2xyz, bvhnodes, bvhnode2aabb):
    t = ctypes.c_float(0)
    tri = ctypes.c_uint32(0)
    bc = np.zeros(3, np.float32)
    st = TraceStats()
    bn = _u32(bvhnodes).reshape(-1)
    hit = lib().orc_first_intersection_ray(_f32(org), _f32(dirv), _u32(tri2vtx).reshape(-1),
                                           _f32(vtx2xyz).reshape(-1), bn, bn.shape[0] // 3,
                                           _f32(bvhnode2aabb).reshape(-1), ctypes.byref(t),
                                           ctypes.byref(tri), bc, ctypes.byref(st))
    return (float(t.value), int(tri.value), bc) if hit else None


def ray_for_pixel(iw, ih, img_w, img_h, ndc2world_colmajor):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_ray_for_pixel(iw, ih, img_w, img_h, _f32(ndc2world_colmajor), o, d)
    return o, d


def ray_tri(p0, p1, p2, org, dirv):
    t = ctypes.c_float(0)
    bc = np.zeros(3, np.float32)
    hit = lib().orc_ray_tri(_f32(p0), _f32(p1), _f32(p2), _f32(org), _f32(dirv), ctypes.byref(t), bc)
    return (float(t.value), bc) if hit else None


def ray_aabb(aabb, org, dirv):
    a = ctypes.c_float(0)
    b = ctypes.c_float(0)
    hit = lib().orc_ray_aabb(_f32(aabb), _f32(org), _f32(dirv), ctypes.byref(a), ctypes.byref(b))
    return (float(a.value), float(b.value)) if hit else None


def mat4_inverse(m_colmajor):
    out = np.zeros(16, np.float32)
    ok = lib().orc_mat4_inverse(_f32(m_colmajor), out)
    return out if ok else None


def pix2depth(pix2tri_, tri2vtx, vtx2xyz, ndc2world_colmajor, img_w, img_h):
    out = np.zeros(img_w * img_h, np.float32)
    rc = lib().orc_pix2depth(out, _u32(pix2tri_).reshape(-1), _u32(tri2vtx).reshape(-1),
                             _f32(vtx2xyz).reshape(-1), img_w, img_h, _f32(ndc2world_colmajor))
    assert rc == 0
    return out.reshape(img_h, img_w)


def pix2depth_bwd(pix2tri_, tri2vtx, vtx2xyz, dldw_pix2depth, ndc2world_colmajor):
    h, w = pix2tri_.shape
    out = np.zeros(_f32(vtx2xyz).size, np.float32)
    rc = lib().orc_pix2depth_bwd(out, _u32(pix2tri_).reshape(-1), _u32(tri2vtx).reshape(-1),
                                 _f32(vtx2xyz).reshape(-1), _f32(dldw_pix2depth).reshape(-1),
                                 w, h, _f32(ndc2world_colmajor))
    assert rc == 0
    return out.reshape(-1, 3)


def edge_gradient_bwd(tri2vtx, vtx2xyz, world2pix_colmajor, pix2tri_, pix2val, dldw_pix2val,
                      dldw_init=None):
    """dL/d(vertex) added term after term to dldw_init (zeros if None)."""
    h, w, nv = pix2val.shape
    out = (np.zeros(_f32(vtx2xyz).size, np.float32) if dldw_init is None
           else _f32(dldw_init).reshape(-1).copy())
    lib().orc_edge_gradient_bwd(_u32(tri2vtx).reshape(-1), _f32(vtx2xyz).reshape(-1), out,
                                _f32(world2pix_colmajor), w, h, _u32(pix2tri_).reshape(-1), nv,
                                _f32(pix2val).reshape(-1), _f32(dldw_pix2val).reshape(-1))
    return out.reshape(-1, 3)


# --- contour / anti-aliasing -----------------------------------------------------
def _sig_aa():
    L = lib()
    if getattr(L, "_aa_ready", False):
        return L
    L.orc_contour_for_triangle_mesh.restype = ctypes.c_long
    L.orc_contour_for_triangle_mesh.argtypes = [_u32p, _f32p, _f32p, _sz, _u32p, _u32p, _u32p]
    L.orc_intersection_edge2.restype = ctypes.c_int
    L.orc_intersection_edge2.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]
    L.orc_dldw_intersection_edge2.restype = None
    L.orc_dldw_intersection_edge2.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_float, _f32p, _f32p]
    L.orc_antialias.restype = ctypes.c_long
    L.orc_antialias.argtypes = [_sz, _u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                _f32p, _f32p]
    L.orc_bwd_antialias.restype = ctypes.c_long
    L.orc_bwd_antialias.argtypes = [_sz, _u32p, _f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                    _f32p, _f32p, _u32p]
    L._aa_ready = True
    return L


def contour(tri2vtx, vtx2xyz, world2ndc_colmajor, edge2vtx, edge2tri):
    L = _sig_aa()
    e2v = _u32(edge2vtx)
    out = np.zeros(e2v.size, np.uint32)
    cnt = L.orc_contour_for_triangle_mesh(_u32(tri2vtx).reshape(-1), _f32(vtx2xyz).reshape(-1),
                                          _f32(world2ndc_colmajor), e2v.shape[0], e2v.reshape(-1),
                                          _u32(edge2tri).reshape(-1), out)
    assert cnt >= 0
    return out[: 2 * cnt].reshape(-1, 2)


def intersection_edge2(c0, c1, s0, e1):
    L = _sig_aa()
    rc = ctypes.c_float(0)
    re = ctypes.c_float(0)
    ok = L.orc_intersection_edge2(_f32(c0), _f32(c1), _f32(s0), _f32(e1), ctypes.byref(rc), ctypes.byref(re))
    return (float(rc.value), float(re.value)) if ok else None


def dldw_intersection_edge2(c0, c1, s0, e1, dldrc):
    L = _sig_aa()
    a = np.zeros(2, np.float32)
    b = np.zeros(2, np.float32)
    L.orc_dldw_intersection_edge2(_f32(c0), _f32(c1), _f32(s0), _f32(e1), dldrc, a, b)
    return a, b


def antialias(cedge2vtx, vtx2xyz, world2pix_colmajor, pix2tri_, pix2vin):
    L = _sig_aa()
    h, w = pix2vin.shape
    out = _f32(pix2vin).copy()
    c = _u32(cedge2vtx)
    n = L.orc_antialias(c.shape[0], c.reshape(-1), _f32(vtx2xyz).reshape(-1),
                        _f32(world2pix_colmajor), w, h, _u32(pix2tri_).reshape(-1),
                        _f32(pix2vin).reshape(-1), out.reshape(-1))
    if n:
        raise RuntimeError(f"{n} antialias pixel walks exceeded 2^26 steps")
    return out


def bwd_antialias(cedge2vtx, vtx2xyz, world2pix_colmajor, pix2val, dldw_pix2val, pix2tri_):
    L = _sig_aa()
    h, w = pix2val.shape
    g = np.zeros(_f32(vtx2xyz).size, np.float32)
    c = _u32(cedge2vtx)
    n = L.orc_bwd_antialias(c.shape[0], c.reshape(-1), _f32(vtx2xyz).reshape(-1), g,
                            _f32(world2pix_colmajor), w, h, _f32(pix2val).reshape(-1),
                            _f32(dldw_pix2val).reshape(-1), _u32(pix2tri_).reshape(-1))
    if n:
        raise RuntimeError(f"{n} antialias pixel walks exceeded 2^26 steps")
    return g.reshape(-1, 3)


# ---------------------------------------------------------------------------
# mesh topology
# ---------------------------------------------------------------------------
def _sig_topo():
    L = lib()
    if getattr(L, "_topo_ready", False):
        return L
    _u32p_or_null = ctypes.c_void_p
    L.orc_vtx2elem.restype = ctypes.c_int
    L.orc_vtx2elem.argtypes = [_sz, _sz, _u32p, _sz, _u32p, _u32p]
    L.orc_vtx2vtx.restype = _sz
    L.orc_vtx2vtx.argtypes = [_sz, _sz, _u32p, _sz, _u32p, _u32p, ctypes.c_int, _u32p,
                              _u32p_or_null]
    L.orc_edge2vtx_from_vtx2vtx.restype = None
    L.orc_edge2vtx_from_vtx2vtx.argtypes = [_sz, _u32p, _u32p, _u32p]
    L.orc_edge2elem.restype = ctypes.c_int
    L.orc_edge2elem.argtypes = [_sz, _u32p, _u32p, _u32p, _u32p, _u32p]
    L.orc_laplacian_smoothing.restype = None
    L.orc_laplacian_smoothing.argtypes = [_sz, _u32p, _u32p, ctypes.c_float, _sz, _f32p, _f32p, _sz,
                                          _f32p]
    L.orc_multiply_graph_laplacian.restype = None
    L.orc_multiply_graph_laplacian.argtypes = [_sz, _u32p, _u32p, _sz, _f32p, _f32p]
    L._topo_ready = True
    return L


def vtx2elem(elem2vtx, num_vtx):
    """(vtx2idx [num_vtx+1], idx2elem [num_elem*num_node]) of vtx2elem.rs:58-96."""
    L = _sig_topo()
    e = _u32(elem2vtx)
    vtx2idx = np.zeros(num_vtx + 1, np.uint32)
    idx2elem = np.zeros(e.size, np.uint32)
    if L.orc_vtx2elem(e.shape[0], e.shape[1], e.reshape(-1), num_vtx, vtx2idx, idx2elem) != 0:
        raise ValueError("vertex index out of range")
    return vtx2idx, idx2elem


def vtx2vtx(elem2vtx, num_vtx, is_self):
    """(vtx2idx, idx2vtx) of vtx2vtx.rs:78-102 from_uniform_mesh."""
    L = _sig_topo()
    e = _u32(elem2vtx)
    vtx2jdx, jdx2elem = vtx2elem(e, num_vtx)
    vtx2idx = np.zeros(num_vtx + 1, np.uint32)
    n = L.orc_vtx2vtx(e.shape[0], e.shape[1], e.reshape(-1), num_vtx, vtx2jdx, jdx2elem,
                      int(bool(is_self)), vtx2idx, None)
    idx2vtx = np.zeros(max(n, 1), np.uint32)
    L.orc_vtx2vtx(e.shape[0], e.shape[1], e.reshape(-1), num_vtx, vtx2jdx, jdx2elem,
                  int(bool(is_self)), vtx2idx, idx2vtx.ctypes.data)
    return vtx2idx, idx2vtx[:n]


def edge2vtx_from_vtx2vtx(vtx2idx, idx2vtx):
    L = _sig_topo()
    out = np.zeros((idx2vtx.shape[0], 2), np.uint32)
    L.orc_edge2vtx_from_vtx2vtx(vtx2idx.shape[0] - 1, _u32(vtx2idx), _u32(idx2vtx), out.reshape(-1))
    return out


def edge2elem(edge2vtx, tri2vtx, vtx2idx, idx2tri, fill=0xFFFFFFFF):
    """edge2elem.rs:20-70; entries not found keep `fill`."""
    L = _sig_topo()
    e2v = _u32(edge2vtx)
    out = np.full((e2v.shape[0], 2), fill, np.uint32)
    rc = L.orc_edge2elem(e2v.shape[0], e2v.reshape(-1), _u32(tri2vtx).reshape(-1), _u32(vtx2idx),
                         _u32(idx2tri), out.reshape(-1))
    if rc != 0:
        raise ValueError("a triangle holds an edge's first vertex more than once")
    return out


def laplacian_smoothing(vtx2idx, idx2vtx, lam, vtx2lhs, vtx2rhs, num_iter):
    L = _sig_topo()
    lhs = _f32(vtx2lhs).copy()
    tmp = np.zeros_like(lhs)
    L.orc_laplacian_smoothing(vtx2idx.shape[0] - 1, _u32(vtx2idx), _u32(idx2vtx), lam, lhs.shape[1],
                              lhs.reshape(-1), _f32(vtx2rhs).reshape(-1), num_iter, tmp.reshape(-1))
    return lhs


def multiply_graph_laplacian(vtx2idx, idx2vtx, vtx2rhs):
    L = _sig_topo()
    rhs = _f32(vtx2rhs)
    out = np.zeros_like(rhs)
    L.orc_multiply_graph_laplacian(vtx2idx.shape[0] - 1, _u32(vtx2idx), _u32(idx2vtx), rhs.shape[1],
                                   rhs.reshape(-1), out.reshape(-1))
    return out


# ---------------------------------------------------------------------------
# attribute interpolation
# ---------------------------------------------------------------------------
def _sig_interp():
    L = lib()
    if getattr(L, "_interp_ready", False):
        return L
    L.orc_interpolate.restype = ctypes.c_int
    L.orc_interpolate.argtypes = [_f32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p, _sz, _f32p,
                                  _sz, ctypes.c_uint32, _f32p, _f32p]
    L.orc_interpolate_bwd.restype = ctypes.c_int
    L.orc_interpolate_bwd.argtypes = [_f32p, _f32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                      _sz, _f32p, _sz, ctypes.c_uint32, _f32p, _f32p, _f32p]
    L._interp_ready = True
    return L


def interpolate(pix2tri_, tri2vtx, vtx2xyz, vtx2val, ndc2world_colmajor):
    """pix2tri.rs:176-229 -> pix2val [H, W, num_vdim]."""
    L = _sig_interp()
    h, w = pix2tri_.shape
    val = _f32(vtx2val)
    nv = val.shape[1]
    out = np.zeros((h, w, nv), np.float32)
    t = _u32(tri2vtx)
    x = _f32(vtx2xyz)
    if L.orc_interpolate(out.reshape(-1), _u32(pix2tri_).reshape(-1), w, h, t.reshape(-1), t.shape[0],
                         x.reshape(-1), x.shape[0], nv, val.reshape(-1), _f32(ndc2world_colmajor)):
        raise ValueError("index out of range")
    return out


def interpolate_bwd(pix2tri_, tri2vtx, vtx2xyz, vtx2val, ndc2world_colmajor, dldw_pix2val):
    """pix2tri.rs:231-304 -> (dL/dvtx2xyz, dL/dvtx2val)."""
    L = _sig_interp()
    h, w = pix2tri_.shape
    val = _f32(vtx2val)
    nv = val.shape[1]
    t = _u32(tri2vtx)
    x = _f32(vtx2xyz)
    gx = np.zeros(x.shape, np.float32)
    gv = np.zeros(val.shape, np.float32)
    if L.orc_interpolate_bwd(gx.reshape(-1), gv.reshape(-1), _u32(pix2tri_).reshape(-1), w, h,
                             t.reshape(-1), t.shape[0], x.reshape(-1), x.shape[0], nv, val.reshape(-1),
                             _f32(ndc2world_colmajor), _f32(dldw_pix2val).reshape(-1)):
        raise ValueError("index out of range")
    return gx, gv


# ---------------------------------------------------------------------------
# smoothed edge gradient
# ---------------------------------------------------------------------------
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def _sig_smooth():
    L = lib()
    if getattr(L, "_smooth_ready", False):
        return L
    L.orc_edge_gradient_and_type.restype = None
    L.orc_edge_gradient_and_type.argtypes = [_u32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_uint32,
                                             _u32p, ctypes.c_uint32, _f32p, _f32p, _u8p, _f32p,
                                             _u8p, _f32p]
    for name in ("orc_smooth_gradient_serial", "orc_smooth_gradient_red_black"):
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u8p, _u8p, ctypes.c_uint32, _f32p, _f32p]
    L.orc_interpolate_staggered_grid.restype = None
    L.orc_interpolate_staggered_grid.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p, _sz,
                                                 _f32p, _f32p]
    L._smooth_ready = True
    return L


def edge_gradient_and_type(tri2vtx, vtx2xyz, world2pix_colmajor, pix2tri_, pix2val, dldw_pix2val):
    """-> hedge2type [H-1,W] u8, hedge2dldr [H-1,W], vedge2type [H,W-1] u8, vedge2dldr [H,W-1]."""
    L = _sig_smooth()
    h, w = pix2tri_.shape
    pv = _f32(pix2val)
    nv = pv.shape[2]
    ht = np.zeros((h - 1, w), np.uint8)
    hd = np.zeros((h - 1, w), np.float32)
    vt = np.zeros((h, w - 1), np.uint8)
    vd = np.zeros((h, w - 1), np.float32)
    L.orc_edge_gradient_and_type(_u32(tri2vtx).reshape(-1), _f32(vtx2xyz).reshape(-1),
                                 _f32(world2pix_colmajor), w, h, _u32(pix2tri_).reshape(-1), nv,
                                 pv.reshape(-1), _f32(dldw_pix2val).reshape(-1), ht.reshape(-1),
                                 hd.reshape(-1), vt.reshape(-1), vd.reshape(-1))
    return ht, hd, vt, vd


def smooth_gradient(hedge2type, vedge2type, num_iter, hedge2dldr, vedge2dldr, red_black=True):
    """Returns smoothed copies (red-black as the reference GPU path, or the
    serial sweeps of del-msh-cpu)."""
    L = _sig_smooth()
    h = hedge2type.shape[0] + 1
    w = hedge2type.shape[1]
    hd = _f32(hedge2dldr).copy()
    vd = _f32(vedge2dldr).copy()
    fn = L.orc_smooth_gradient_red_black if red_black else L.orc_smooth_gradient_serial
    fn(w, h, np.ascontiguousarray(hedge2type, np.uint8).reshape(-1),
       np.ascontiguousarray(vedge2type, np.uint8).reshape(-1), num_iter, hd.reshape(-1), vd.reshape(-1))
    return hd, vd


def interpolate_staggered_grid(hedge2vy, vedge2vx, vtx2xy):
    L = _sig_smooth()
    h = hedge2vy.shape[0] + 1
    w = hedge2vy.shape[1]
    xy = _f32(vtx2xy)
    out = np.zeros_like(xy)
    L.orc_interpolate_staggered_grid(w, h, _f32(hedge2vy).reshape(-1), _f32(vedge2vx).reshape(-1),
                                     xy.shape[0], xy.reshape(-1), out.reshape(-1))
    return out


# ---------------------------------------------------------------------------
# quad/oct-tree and sorted-array helpers
# ---------------------------------------------------------------------------
def _sig_octree():
    L = lib()
    if getattr(L, "_octree_ready", False):
        return L
    u32 = ctypes.c_uint32
    L.orc_octree_bnodes.restype = ctypes.c_int
    L.orc_octree_bnodes.argtypes = [_sz, _u32p, u32, u32, _u32p, _u32p, _u32p, _u32p]
    L.orc_morton2center.restype = None
    L.orc_morton2center.argtypes = [u32, u32, u32, u32, _f32p]
    L.orc_octree_make_tree.restype = ctypes.c_int
    L.orc_octree_make_tree.argtypes = [_sz, _u32p, _u32p, _u32p, _u32p, _u32p, _sz, u32, u32,
                                       _u32p, _u32p, _f32p, _u32p, _f32p]
    L.orc_octree_aggregate.restype = None
    L.orc_octree_aggregate.argtypes = [_sz, u32, _f32p, _u32p, u32, _sz, _u32p, _f32p]
    L.orc_octree_aggregate_tree.restype = None
    L.orc_octree_aggregate_tree.argtypes = [_sz, u32, _u32p, u32, _f32p, _f32p]
    L.orc_unique_for_sorted.restype = None
    L.orc_unique_for_sorted.argtypes = [_sz, _u32p, _u32p]
    L.orc_unique_jdx2val_jdx2idx.restype = None
    L.orc_unique_jdx2val_jdx2idx.argtypes = [_sz, _u32p, _u32p, _sz, _u32p, _u32p]
    L.orc_offset_array_aggregate.restype = None
    L.orc_offset_array_aggregate.argtypes = [_sz, _u32p, _u32p, u32, _f32p, _f32p]
    L._octree_ready = True
    return L


def octree_bnodes(idx2morton, num_dim, max_depth=None):
    """(bnodes [n-1,3], bnode2depth, bnode2onode, idx2bnode) of
    quad_oct_tree.rs:1-109."""
    L = _sig_octree()
    if max_depth is None:
        max_depth = 32 // num_dim
    m = _u32(idx2morton)
    n = m.shape[0]
    bnodes = np.zeros((n - 1, 3), np.uint32)
    depth = np.zeros(n - 1, np.uint32)
    b2o = np.zeros(n - 1, np.uint32)
    i2b = np.full(n, 0xFFFFFFFF, np.uint32)
    if L.orc_octree_bnodes(n, m, num_dim, max_depth, bnodes.reshape(-1), depth, b2o, i2b) != 0:
        raise ValueError("binary radix tree failed")
    return bnodes, depth, b2o, i2b


def morton2center(morton, num_dim, depth, max_depth):
    c = np.zeros(3, np.float32)
    _sig_octree().orc_morton2center(int(morton), num_dim, depth, max_depth, c)
    return c[:num_dim]


def octree_make_tree(bnodes, bnode2onode, bnode2depth, idx2bnode, idx2morton, num_dim,
                     max_depth=None):
    """(onodes, onode2depth, onode2center, idx2onode, idx2center) of
    quad_oct_tree.rs:111-227."""
    L = _sig_octree()
    if max_depth is None:
        max_depth = 32 // num_dim
    n = _u32(idx2morton).shape[0]
    num_onode = int(bnode2onode[-1]) + 1
    nlink = 1 + 2 ** num_dim
    onodes = np.full((num_onode, nlink), 0xFFFFFFFF, np.uint32)
    o2d = np.zeros(num_onode, np.uint32)
    o2c = np.zeros((num_onode, num_dim), np.float32)
    i2o = np.zeros(n, np.uint32)
    i2c = np.zeros((n, num_dim), np.float32)
    rc = L.orc_octree_make_tree(n, _u32(bnodes).reshape(-1), _u32(bnode2onode), _u32(bnode2depth),
                                _u32(idx2bnode), _u32(idx2morton), num_onode, max_depth, num_dim,
                                onodes.reshape(-1), o2d, o2c.reshape(-1), i2o, i2c.reshape(-1))
    if rc != 0:
        raise ValueError("octree construction failed (duplicated child)")
    return onodes, o2d, o2c, i2o, i2c


def octree_aggregate(idx2val, idx2onode, onodes):
    L = _sig_octree()
    v = _f32(idx2val)
    on = _u32(onodes)
    out = np.zeros((on.shape[0], v.shape[1]), np.float32)
    L.orc_octree_aggregate(v.shape[0], v.shape[1], v.reshape(-1), _u32(idx2onode), on.shape[1],
                           on.shape[0], on.reshape(-1), out.reshape(-1))
    return out


def octree_aggregate_tree(idx2val, onodes):
    """The kernels' summation order: each node adds its children in slot order."""
    L = _sig_octree()
    v = _f32(idx2val)
    on = _u32(onodes)
    out = np.zeros((on.shape[0], v.shape[1]), np.float32)
    L.orc_octree_aggregate_tree(on.shape[0], on.shape[1], on.reshape(-1), v.shape[1], v.reshape(-1),
                                out.reshape(-1))
    return out


def unique_for_sorted(idx2val):
    """(idx2jdx, jdx2val, jdx2idx_offset) of array1d.rs:1-15 + map_idx.rs:83-115."""
    L = _sig_octree()
    v = _u32(idx2val)
    n = v.shape[0]
    idx2jdx = np.zeros(n, np.uint32)
    L.orc_unique_for_sorted(n, v, idx2jdx)
    num_jdx = int(idx2jdx[-1]) + 1 if n else 0
    jdx2val = np.zeros(num_jdx, np.uint32)
    off = np.zeros(num_jdx + 1, np.uint32)
    L.orc_unique_jdx2val_jdx2idx(n, v, idx2jdx, num_jdx, jdx2val, off)
    return idx2jdx, jdx2val, off


def offset_array_aggregate(idx2jdx_offset, jdx2kdx, kdx2val):
    L = _sig_octree()
    off = _u32(idx2jdx_offset)
    v = _f32(kdx2val)
    out = np.zeros((off.shape[0] - 1, v.shape[1]), np.float32)
    L.orc_offset_array_aggregate(off.shape[0] - 1, off, _u32(jdx2kdx), v.shape[1], v.reshape(-1),
                                 out.reshape(-1))
    return out


# ---------------------------------------------------------------------------
# mixed polyhedron meshes (CSR)
# ---------------------------------------------------------------------------
def _sig_poly():
    L = lib()
    if getattr(L, "_poly_ready", False):
        return L
    u32 = ctypes.c_uint32
    L.orc_polyhedron_elem2center.restype = None
    L.orc_polyhedron_elem2center.argtypes = [_sz, _u32p, _u32p, u32, _f32p, _f32p]
    L.orc_polyhedron_refit.restype = None
    L.orc_polyhedron_refit.argtypes = [_f32p, _u32p, _sz, _u32p, _u32p, _f32p]
    L.orc_shape.restype = None
    L.orc_shape.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p,
                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.orc_param_coord.restype = ctypes.c_int
    L.orc_param_coord.argtypes = [_f32p, _u32p, u32, _f32p, _f32p]
    L.orc_polyhedron_search.restype = None
    L.orc_polyhedron_search.argtypes = [_sz, _u32p, _u32p, _f32p, _u32p, _f32p, _sz, _f32p, _u32p,
                                        _f32p]
    L.orc_polyhedron_interpolate.restype = None
    L.orc_polyhedron_interpolate.argtypes = [_sz, _u32p, _u32p, u32, _f32p, _u32p, _f32p, _f32p]
    L._poly_ready = True
    return L


def polyhedron_elem2center(elem2idx_offset, idx2vtx, vtx2xyz):
    L = _sig_poly()
    off, iv, x = _u32(elem2idx_offset), _u32(idx2vtx), _f32(vtx2xyz)
    out = np.zeros((off.shape[0] - 1, x.shape[1]), np.float32)
    L.orc_polyhedron_elem2center(off.shape[0] - 1, off, iv, x.shape[1], x.reshape(-1),
                                 out.reshape(-1))
    return out


def polyhedron_refit(bvhnodes, elem2idx_offset, idx2vtx, vtx2xyz):
    L = _sig_poly()
    bn = _u32(bvhnodes)
    out = np.zeros((bn.shape[0], 6), np.float32)
    off = _u32(elem2idx_offset)
    L.orc_polyhedron_refit(out.reshape(-1), bn.reshape(-1), off.shape[0] - 1, off, _u32(idx2vtx),
                           _f32(vtx2xyz).reshape(-1))
    return out


def shape_functions(num_node, r, s, t):
    N = np.zeros(8, np.float32)
    _sig_poly().orc_shape(num_node, r, s, t, N, None, None, None)
    return N[:num_node]


def param_coord(q, node2vtx, vtx2xyz):
    p = np.zeros(3, np.float32)
    ok = _sig_poly().orc_param_coord(_f32(q), _u32(node2vtx), len(node2vtx),
                                     _f32(vtx2xyz).reshape(-1), p)
    return p if ok else None


def polyhedron_search(elem2idx_offset, idx2vtx, vtx2xyz, bvhnodes, bvhnode2aabb, wtx2xyz):
    L = _sig_poly()
    off = _u32(elem2idx_offset)
    w = _f32(wtx2xyz)
    elem = np.zeros(w.shape[0], np.uint32)
    par = np.zeros((w.shape[0], 3), np.float32)
    L.orc_polyhedron_search(off.shape[0] - 1, off, _u32(idx2vtx), _f32(vtx2xyz).reshape(-1),
                            _u32(bvhnodes).reshape(-1), _f32(bvhnode2aabb).reshape(-1),
                            w.shape[0], w.reshape(-1), elem, par.reshape(-1))
    return elem, par


def polyhedron_interpolate(elem2idx_offset, idx2vtx, vtx2val, wtx2elem, wtx2param):
    L = _sig_poly()
    v = _f32(vtx2val)
    e = _u32(wtx2elem)
    out = np.zeros((e.shape[0], v.shape[1]), np.float32)
    L.orc_polyhedron_interpolate(e.shape[0], _u32(elem2idx_offset), _u32(idx2vtx), v.shape[1],
                                 v.reshape(-1), e, _f32(wtx2param).reshape(-1), out.reshape(-1))
    return out
