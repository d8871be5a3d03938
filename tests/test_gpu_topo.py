"""GPU parity of the mesh-topology builders, the graph Laplacian and the
exclusive scan against the oracle (bit-exact: integer work, and f32 sums in
the CPU's order), through the reference's Python entry points."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Edge2Elem.torch as Edge2Elem
import del_msh_dlpack.Edge2Vtx.torch as Edge2Vtx
import del_msh_dlpack.TriMesh3.torch as TriMesh3
import del_msh_dlpack.Vtx2Elem.torch as Vtx2Elem
import del_msh_dlpack.Vtx2Vtx.torch as Vtx2Vtx

pytestmark = pytest.mark.gpu


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def no_device_errors():
    bits = _lib.device_errors()
    assert bits == 0, f"device error bits 0x{bits:x}"


def _case(name):
    if name == "sphere":
        tri, vtx = primitives.sphere_yup(1.0, 32, 64)
        return tri, vtx.shape[0]
    if name == "torus":
        tri, vtx = primitives.torus_zup(1.3, 0.4, 64, 32)
        return tri, vtx.shape[0]
    if name == "octa7":
        tri, vtx = primitives.octahedral_sphere(7, noise_sigma=0.01, seed=2)
        return tri, vtx.shape[0]
    if name == "ragged":  # isolated vertices, a repeated vertex in one element
        return np.array([[1, 2, 3], [3, 2, 4], [4, 6, 6], [7, 8, 1], [2, 8, 7]], np.uint32), 10
    if name == "quads":
        rng = np.random.default_rng(7)
        return rng.integers(0, 5000, (20000, 4)).astype(np.uint32), 5003
    if name == "fan":  # hub with 3000 neighbours: lists longer than a wave's LDS share
        n = 3000
        k = np.arange(n, dtype=np.uint32)
        return np.stack([np.zeros(n, np.uint32), 1 + k, 1 + (k + 1) % n], axis=1), n + 1
    if name == "hubs":  # a few vertices in hundreds of random triangles, with repeats
        rng = np.random.default_rng(11)
        tri = rng.integers(0, 400, (4000, 3)).astype(np.uint32)
        tri[::3, 0] = rng.integers(0, 3, tri[::3].shape[0])
        return tri, 401
    raise KeyError(name)


@pytest.mark.parametrize("name", ["sphere", "torus", "octa7", "ragged", "quads", "fan", "hubs"])
def test_vtx2elem(gpu, oracle, name):
    elem, nv = _case(name)
    offs, idx2elem = Vtx2Elem.from_uniform_mesh(G(elem, gpu), nv)
    w_offs, w_idx = oracle.vtx2elem(elem, nv)
    assert np.array_equal(H(offs), w_offs)
    assert np.array_equal(H(idx2elem), w_idx)
    no_device_errors()


@pytest.mark.parametrize("is_self", [False, True])
@pytest.mark.parametrize("name", ["sphere", "torus", "octa7", "ragged", "quads", "fan", "hubs"])
def test_vtx2vtx(gpu, oracle, name, is_self):
    elem, nv = _case(name)
    offs, idx2vtx = Vtx2Vtx.from_uniform_mesh(G(elem, gpu), nv, is_self)
    w_offs, w_idx = oracle.vtx2vtx(elem, nv, is_self)
    assert np.array_equal(H(offs), w_offs)
    assert np.array_equal(H(idx2vtx), w_idx)
    no_device_errors()


@pytest.mark.parametrize("name", ["sphere", "torus", "octa7"])
def test_edges_and_edge2tri(gpu, oracle, name):
    """TriMesh3.make_edge2vtx / make_edge2tri (the reference's composition)
    against the oracle's edge2vtx_from_vtx2vtx and edge2elem."""
    tri, nv = _case(name)
    t = G(tri, gpu)
    e2v = TriMesh3.make_edge2vtx(t, nv)
    offs, idx2vtx = oracle.vtx2vtx(tri, nv, False)
    w_e2v = oracle.edge2vtx_from_vtx2vtx(offs, idx2vtx)
    assert np.array_equal(H(e2v), w_e2v)
    e2t = TriMesh3.make_edge2tri(t, nv, e2v)
    v2e_off, v2e = oracle.vtx2elem(tri, nv)
    assert np.array_equal(H(e2t), oracle.edge2elem(w_e2v, tri, v2e_off, v2e))
    no_device_errors()


def test_edge2tri_open_edges_keep_their_values(gpu, oracle):
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    e2v = np.array([[0, 2], [0, 1], [2, 3]], np.uint32)
    v2e_off, v2e = Vtx2Elem.from_uniform_mesh(G(tri, gpu), 4)
    out = torch.full((3, 2), 77, dtype=torch.uint32, device=gpu)
    Edge2Elem.from_edge2vtx_of_tri2vtx_with_vtx2vtx(G(e2v, gpu), G(tri, gpu), v2e_off, v2e, out)
    assert H(out).tolist() == [[0, 1], [0, 77], [1, 77]]
    no_device_errors()


def test_contour_through_gpu_edges_matches_oracle(gpu, oracle):
    """The reference's contour pipeline (test_differentiable_antialias.py:
    126-196) with the edge lists built on the GPU."""
    tri, vtx = primitives.torus_zup(1.3, 0.4, 64, 32)
    c, s = np.cos(1.15), np.sin(1.15)
    rot = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    vtx = (vtx.astype(np.float64) @ rot.T + [0, 0.6, 0]).astype(np.float32)
    w2n = np.diag([0.5, 0.5, 0.5, 1.0]).astype(np.float32)
    t, v = G(tri, gpu), G(vtx, gpu)
    e2v = TriMesh3.make_edge2vtx(t, vtx.shape[0])
    e2t = TriMesh3.make_edge2tri(t, vtx.shape[0], e2v)
    cedge = Edge2Vtx.contour_for_triangle_mesh(t, v, G(w2n, gpu), e2v, e2t)
    want = oracle.contour(tri, vtx, oracle.colmajor(w2n), H(e2v), H(e2t))
    assert np.array_equal(H(cedge), want)
    # each silhouette edge appears once from each end
    fwd = want[want[:, 0] < want[:, 1]]
    assert 2 * fwd.shape[0] == want.shape[0] and fwd.shape[0] > 0
    no_device_errors()


@pytest.mark.parametrize("num_dim", [1, 3, 5])
def test_graph_laplacian_and_smoothing(gpu, oracle, num_dim):
    tri, nv = _case("torus")
    offs, idx2vtx = oracle.vtx2vtx(tri, nv, False)
    rng = np.random.default_rng(num_dim)
    x = rng.standard_normal((nv, num_dim)).astype(np.float32)
    b = rng.standard_normal((nv, num_dim)).astype(np.float32)
    d_offs, d_idx = G(offs, gpu), G(idx2vtx, gpu)
    lx = Vtx2Vtx.multiply_graph_laplacian(d_offs, d_idx, G(x, gpu))
    assert np.array_equal(H(lx), oracle.multiply_graph_laplacian(offs, idx2vtx, x))
    lhs = G(x, gpu)
    Vtx2Vtx.laplacian_smoothing(d_offs, d_idx, 0.7, lhs, G(b, gpu), 5, None)
    assert np.array_equal(H(lhs), oracle.laplacian_smoothing(offs, idx2vtx, 0.7, x, b, 5))
    no_device_errors()


def test_graph_laplacian_autograd(gpu, oracle):
    tri, nv = _case("sphere")
    offs, idx2vtx = oracle.vtx2vtx(tri, nv, False)
    rng = np.random.default_rng(3)
    x = G(rng.standard_normal((nv, 3)).astype(np.float32), gpu).requires_grad_(True)
    w = rng.standard_normal((nv, 3)).astype(np.float32)
    y = Vtx2Vtx.GraphLaplacian.apply(G(offs, gpu), G(idx2vtx, gpu), x)
    (y * G(w, gpu)).sum().backward()
    assert np.array_equal(H(x.grad), oracle.multiply_graph_laplacian(offs, idx2vtx, w))


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 300001])
def test_exclusive_scan(gpu, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 1000, n).astype(np.uint32)
    d_in = G(a, gpu) if n else torch.empty((0,), dtype=torch.uint32, device=gpu)
    out = torch.full((n + 1,), 0xDEADBEEF, dtype=torch.uint32, device=gpu)
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_exclusive_scan_workspace_bytes(n), gpu)
    _lib.check(L.dmsh_exclusive_scan_u32(_lib.stream_of(gpu), n, _lib.ptr(d_in), _lib.ptr(out),
                                         _lib.ptr(ws), ws.numel()), "scan")
    want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)]).astype(np.uint32)
    assert np.array_equal(H(out), want)


def test_out_of_range_vertex_raises(gpu):
    """A vertex index >= num_vtx is clamped on the device and reported."""
    _lib.device_errors()
    Vtx2Elem.from_uniform_mesh(G(np.array([[0, 1, 9]], np.uint32), gpu), 5)
    assert _lib.device_errors() & 0x1


def test_full_size_topology(gpu, oracle):
    """2^21-triangle octahedral sphere: vertex adjacency and edges bit-exact."""
    tri, vtx = primitives.octahedral_sphere(9, noise_sigma=0.01, seed=1)
    nv = vtx.shape[0]
    t = G(tri, gpu)
    e2v = TriMesh3.make_edge2vtx(t, nv)
    e2t = TriMesh3.make_edge2tri(t, nv, e2v)
    offs, idx2vtx = oracle.vtx2vtx(tri, nv, False)
    w_e2v = oracle.edge2vtx_from_vtx2vtx(offs, idx2vtx)
    assert np.array_equal(H(e2v), w_e2v)
    v2e_off, v2e = oracle.vtx2elem(tri, nv)
    assert np.array_equal(H(e2t), oracle.edge2elem(w_e2v, tri, v2e_off, v2e))
    no_device_errors()
