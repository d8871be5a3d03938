"""GPU: the native-level entry points (del_msh_dlpack.del_msh_dlpack, the
reference's pyo3 names of del-msh-dlpack/src/lib.rs:29-57) called with DLPack
capsules of device arrays and a stream pointer, against the oracle; the
reference's argument checks raise PanicException."""

import numpy as np
import pytest
import torch
from torch.utils.dlpack import from_dlpack, to_dlpack

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.del_msh_dlpack as N
import del_msh_dlpack.Pix2Depth.torch as Pix2Depth
import del_msh_dlpack.Array1D.torch as Array1D
import del_msh_dlpack.Mat44.torch as Mat44

pytestmark = pytest.mark.gpu

U32MAX = 0xFFFFFFFF


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def C(t):
    return to_dlpack(t)


def bits_equal(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def no_device_errors():
    bits = _lib.device_errors()
    assert bits == 0, f"device error bits 0x{bits:x}"


@pytest.fixture(scope="module")
def scene(gpu, oracle):
    tri, vtx = primitives.torus_zup(1.3, 0.4, 64, 32)
    c, s = np.cos(1.15), np.sin(1.15)
    vtx = (vtx.astype(np.float64) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).T
           + [0, 0.3, 0]).astype(np.float32)
    w2n = np.diag([0.5, 0.5, 0.5, 1.0])
    n2w = np.linalg.inv(w2n).astype(np.float32)
    W, H_ = 128, 96
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * H_, 0, 0.5 * H_], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]])
    w2p = (n2p @ w2n).astype(np.float32)
    return dict(tri=tri, vtx=vtx, w2n=w2n.astype(np.float32), n2w=n2w, w2p=w2p, W=W, H=H_,
                t=G(tri, gpu), v=G(vtx, gpu))


def test_native_morton_sort_topology_refit(gpu, oracle, scene):
    """mortons_vtx2morton_from_vtx2co -> array1d_argsort ->
    mortons_make_bvhnodes_from_sorted_mortons -> bvhnode2aabb_update_aabb,
    each through capsules, equal to the oracle's del-msh-cpu pipeline."""
    tri, vtx = scene["tri"], scene["vtx"]
    stream = torch.cuda.current_stream(gpu).cuda_stream
    cntr = oracle.tri2centroid(tri, vtx)
    T = oracle.fit_into_unit_cube(cntr)
    d_cntr = G(cntr, gpu)
    codes = torch.empty(tri.shape[0], dtype=torch.uint32, device=gpu)
    N.mortons_vtx2morton_from_vtx2co(C(d_cntr), C(G(T, gpu)), C(codes), stream)
    want_codes = oracle.vtx2morton(cntr, T)
    assert np.array_equal(H(codes), want_codes)
    idx = torch.empty_like(codes)
    N.array1d_argsort(C(codes), C(idx), stream)
    widx, wsorted = oracle.argsort(want_codes)
    assert np.array_equal(H(idx), widx) and np.array_equal(H(codes), wsorted)
    n = tri.shape[0]
    bn = torch.empty((2 * n - 1, 3), dtype=torch.uint32, device=gpu)
    N.mortons_make_bvhnodes_from_sorted_mortons(C(idx), C(codes), C(bn), stream)
    wbn = oracle.bvhnodes_from_sorted(widx, wsorted)
    assert np.array_equal(H(bn), wbn)
    aabb = torch.zeros((2 * n - 1, 6), dtype=torch.float32, device=gpu)
    N.bvhnode2aabb_update_aabb(C(aabb), 0, C(bn), C(scene["t"]), C(scene["v"]), C(scene["v"]),
                               stream)
    assert bits_equal(H(aabb), oracle.refit_trimesh3(wbn, tri, vtx))
    no_device_errors()


def test_native_raycast_depth_and_gradients(gpu, oracle, scene):
    """pix2tri_by_raycast, pix2depth_update, pix2depth_bwd_wrt_vtx2xyz,
    rasterized_edge_gradient_bwd, edge2vtx_contour_for_triangle_mesh and
    differentiable_antialias_fwd / _bwd through capsules."""
    tri, vtx, W, H_ = scene["tri"], scene["vtx"], scene["W"], scene["H"]
    stream = torch.cuda.current_stream(gpu).cuda_stream
    wbn, waabb = oracle.build_trimesh3(tri, vtx)
    n2w_c = oracle.colmajor(scene["n2w"])
    pix = torch.full((H_, W), U32MAX, dtype=torch.uint32, device=gpu)
    N.pix2tri_by_raycast(C(pix), C(scene["t"]), C(scene["v"]), C(G(wbn, gpu)), C(G(waabb, gpu)),
                         C(G(n2w_c, gpu)), stream)
    wpix, _, _ = oracle.pix2tri(tri, vtx, wbn, waabb, n2w_c, W, H_, num_threads=8)
    assert np.array_equal(H(pix), wpix) and (wpix != U32MAX).any()
    depth = torch.zeros((H_, W), dtype=torch.float32, device=gpu)
    N.pix2depth_update(C(depth), C(pix), C(scene["t"]), C(scene["v"]), C(G(n2w_c, gpu)), stream)
    assert bits_equal(H(depth), oracle.pix2depth(wpix, tri, vtx, n2w_c, W, H_))
    rng = np.random.default_rng(7)
    dldd = rng.normal(size=(H_, W)).astype(np.float32)
    gd = torch.zeros_like(scene["v"])
    N.pix2depth_bwd_wrt_vtx2xyz(C(gd), C(pix), C(scene["t"]), C(scene["v"]), C(G(dldd, gpu)),
                                C(G(n2w_c, gpu)), stream)
    want = oracle.pix2depth_bwd(wpix, tri, vtx, dldd, n2w_c)
    assert np.abs(H(gd) - want).max() <= 1e-5 * np.abs(want).max()
    # edge gradient: bit-exact (the serial loop's order)
    occ = (wpix != U32MAX).astype(np.float32)[..., None]
    dldo = rng.random((H_, W, 1), dtype=np.float32)
    w2p_c = oracle.colmajor(scene["w2p"])
    ge = torch.zeros_like(scene["v"])
    N.rasterized_edge_gradient_bwd(C(scene["t"]), C(scene["v"]), C(ge), C(G(w2p_c, gpu)), C(pix),
                                   C(G(occ, gpu)), C(G(dldo, gpu)), stream)
    assert bits_equal(H(ge), oracle.edge_gradient_bwd(tri, vtx, w2p_c, wpix, occ, dldo))
    # contour: a capsule owning a new device array
    e2v, e2t = primitives.edge2vtx_edge2tri(tri)
    cap = N.edge2vtx_contour_for_triangle_mesh(C(scene["t"]), C(scene["v"]),
                                               C(G(oracle.colmajor(scene["w2n"]), gpu)),
                                               C(G(e2v, gpu)), C(G(e2t, gpu)), stream)
    cedge = from_dlpack(cap)
    assert cedge.device == gpu and cedge.dtype == torch.uint32
    wce = oracle.contour(tri, vtx, oracle.colmajor(scene["w2n"]), e2v, e2t)
    assert np.array_equal(H(cedge), wce)
    pin = G(occ[..., 0], gpu)
    pout = pin.clone()  # the kernel writes the pixels next to contour edges only
    N.differentiable_antialias_fwd(C(cedge), C(scene["v"]), C(G(w2p_c, gpu)), C(pix), C(pin),
                                   C(pout), stream)
    assert bits_equal(H(pout), oracle.antialias(wce, vtx, w2p_c, wpix, occ[..., 0]))
    ga = torch.zeros_like(scene["v"])
    dlda = rng.normal(size=(H_, W)).astype(np.float32)
    N.differentiable_antialias_bwd(C(cedge), C(scene["v"]), C(ga), C(G(w2p_c, gpu)), C(pin),
                                   C(G(dlda, gpu)), C(pix), stream)
    want = oracle.bwd_antialias(wce, vtx, w2p_c, occ[..., 0], dlda, wpix)
    assert np.abs(want).max() > 0
    assert np.abs(H(ga) - want).max() <= 2e-5 * np.abs(want).max()
    no_device_errors()


def test_native_permute_and_checks(gpu):
    """array1d_permute gathers; a wrong shape or element type panics as the
    reference's unwrap() does; an index out of range sets the device error."""
    rng = np.random.default_rng(3)
    old = rng.integers(0, 1 << 32, 1000, dtype=np.uint64).astype(np.uint32)
    perm = rng.permutation(1000).astype(np.uint32)
    out = torch.empty(1000, dtype=torch.uint32, device=gpu)
    N.array1d_permute(C(G(old, gpu)), C(G(perm, gpu)), C(out), 0)
    assert np.array_equal(H(out), old[perm])
    assert np.array_equal(H(Array1D.permute(G(old, gpu), G(perm, gpu))), old[perm])
    with pytest.raises(N.PanicException):
        N.array1d_permute(C(G(old, gpu)), C(G(perm[:999], gpu)), C(out), 0)
    with pytest.raises(N.PanicException):
        N.array1d_argsort(C(G(old.astype(np.float32), gpu)), C(out), 0)
    no_device_errors()
    bad = perm.copy()
    bad[5] = 1000
    N.array1d_permute(C(G(old, gpu)), C(G(bad, gpu)), C(out), 0)
    assert _lib.device_errors() & 0x1


def test_pix2depth_using_interpolation(gpu, oracle, scene):
    """Depth from interpolated hit positions (reference Pix2Depth/torch.py:134-146)
    agrees with the direct depth to f32 rounding and is differentiable."""
    tri, vtx, W, H_ = scene["tri"], scene["vtx"], scene["W"], scene["H"]
    wbn, waabb = oracle.build_trimesh3(tri, vtx)
    n2w = scene["n2w"]
    wpix, _, _ = oracle.pix2tri(tri, vtx, wbn, waabb, oracle.colmajor(n2w), W, H_, num_threads=8)
    v = scene["v"].clone().requires_grad_(True)
    d = Pix2Depth.pix2depth_using_interpolation(v, G(wpix, gpu), scene["t"], G(n2w, gpu))
    want = oracle.pix2depth(wpix, tri, vtx, oracle.colmajor(n2w), W, H_)
    assert np.abs(H(d) - want).max() <= 1e-5
    assert np.all(H(d)[wpix == U32MAX] == 0)
    d.sum().backward()
    g = H(v.grad)
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    # the depth is affine in z for this orthographic camera: dd/dz = 0.5 * 0.5
    # per unit of barycentric weight, so the z gradients sum to 0.25 per pixel hit
    assert abs(g[:, 2].sum() - 0.25 * (wpix != U32MAX).sum()) < 1e-3 * (wpix != U32MAX).sum()
    no_device_errors()


def test_from_transfrom_world2unit(gpu):
    v = torch.tensor([[0.0, 0.0, 0.0], [2.0, 1.0, 1.0], [1.0, 3.0, 0.5]], device=gpu)
    T = Mat44.from_transfrom_world2unit(v, gpu)
    q = (torch.cat([v, torch.ones(3, 1, device=gpu)], 1) @ T.T)[:, :3]
    lo, hi = q.min(0).values, q.max(0).values
    assert torch.allclose((lo + hi) * 0.5, torch.full((3,), 0.5, device=gpu))
    assert abs((hi - lo).max().item() - 1.0) < 1e-6
