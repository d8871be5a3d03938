"""GPU parity of the mixed polyhedron mesh functions against the oracle
(bit-exact: centroids, element BVH, point location with parameters, and
interpolation share the oracle's operations), through the reference's
Python entry points: mixed grids with and without jitter, the reference's
CFD mesh (asset/cfd_mesh.txt, kept in tests/golden), and a BVH deeper than a
64-entry stack."""

import pathlib

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.MixMesh3.torch as MixMesh3
import del_msh_dlpack.PolyhedronMesh.torch as PolyhedronMesh

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).parent / "golden"


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def bits_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                                                 np.ascontiguousarray(b).view(np.uint32))


def _check_mesh(gpu, oracle, off, iv, x, q):
    d_off, d_iv, d_x = G(off, gpu), G(iv, gpu), G(x, gpu)
    cog = PolyhedronMesh.make_elem2centroid(d_off, d_iv, d_x)
    w_cog = oracle.polyhedron_elem2center(off, iv, x)
    assert bits_equal(H(cog), w_cog)
    bn, aabb = PolyhedronMesh.make_bvhnodes_bvhnode2aabb(d_off, d_iv, d_x)
    w_bn, _ = oracle.build_points(w_cog)
    assert np.array_equal(H(bn), w_bn)
    w_aabb = oracle.polyhedron_refit(w_bn, off, iv, x)
    assert bits_equal(H(aabb), w_aabb)
    elem, par = PolyhedronMesh.search_elem_contain_points(d_off, d_iv, d_x, bn, aabb, G(q, gpu))
    w_elem, w_par = oracle.polyhedron_search(off, iv, x, w_bn, w_aabb, q)
    assert np.array_equal(H(elem), w_elem)
    assert bits_equal(H(par), w_par)
    vals = np.random.default_rng(9).standard_normal((x.shape[0], 2)).astype(np.float32)
    got = PolyhedronMesh.interpolate_values_at_points(d_off, d_iv, G(vals, gpu), elem, par)
    assert bits_equal(H(got), oracle.polyhedron_interpolate(off, iv, vals, w_elem, w_par))
    assert _lib.device_errors() == 0
    return w_elem


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_mixed_grid(gpu, oracle, jitter):
    off, iv, x = primitives.mixed_hex_grid(12, 10, 8, seed=3, jitter=jitter)
    rng = np.random.default_rng(4)
    q = rng.uniform([-0.5, -0.5, -0.5], [12.5, 10.5, 8.5], (20000, 3)).astype(np.float32)
    elem = _check_mesh(gpu, oracle, off, iv, x, q)
    assert (elem != 0xFFFFFFFF).mean() > 0.7


def test_reference_cfd_mesh(gpu, oracle):
    from test_oracle_polymesh import load_cfd_mesh

    off, iv, x = load_cfd_mesh(GOLDEN / "cfd_mesh.txt")
    q = (np.random.default_rng(0).random((2000, 3)) * [1.0, 1.0, 3.0] + [0.0, 0.0, -2.0])
    elem = _check_mesh(gpu, oracle, off, iv, x, q.astype(np.float32))
    assert set(np.unique(elem[elem != 0xFFFFFFFF])) == {0, 1, 2, 3}


def test_deep_bvh(gpu, oracle):
    """300 copies of one tet (one centroid, so equal Morton codes chain into
    a tree ~300 deep) among a few others: the stackless walk still visits the
    recursion's nodes in its order and returns the first copy it reaches."""
    tet = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    x = np.concatenate([tet, tet + 5.0, tet * 0.5 + [2.0, 2.0, 2.0]]).astype(np.float32)
    elems = [[0, 1, 2, 3]] * 300 + [[4, 5, 6, 7], [8, 9, 10, 11]]
    off = np.arange(0, 4 * len(elems) + 1, 4, dtype=np.uint32)
    iv = np.asarray(elems, np.uint32).reshape(-1)
    q = np.array([[0.1, 0.1, 0.1], [0.2, 0.3, 0.1], [5.2, 5.2, 5.2], [2.1, 2.1, 2.1],
                  [3.0, 3.0, 3.0]], np.float32)
    bn, _ = oracle.build_points(oracle.polyhedron_elem2center(off, iv, x))
    depth, node = 0, None
    parent = {int(c): i for i in range(len(elems) - 1) for c in bn[i, 1:]}
    for leaf in range(len(elems) - 1, 2 * len(elems) - 1):
        d, n = 0, leaf
        while n in parent:
            n, d = parent[n], d + 1
        depth = max(depth, d)
    assert depth > 64
    elem = _check_mesh(gpu, oracle, off, iv, x, q)
    assert elem[0] < 300 and elem[2] == 300 and elem[3] == 301 and elem[4] == 0xFFFFFFFF


def test_to_polyhedral_mesh(gpu):
    rng = np.random.default_rng(2)
    arrs = [rng.integers(0, 100, (k, n)).astype(np.uint32) for k, n in ((3, 4), (0, 5), (2, 6),
                                                                        (4, 8))]
    off, iv = MixMesh3.to_polyhedral_mesh(*[G(a, gpu) for a in arrs])
    want_iv = np.concatenate([a.reshape(-1) for a in arrs])
    want_off = np.concatenate([[0], np.cumsum([4] * 3 + [6] * 2 + [8] * 4)]).astype(np.uint32)
    assert np.array_equal(H(off), want_off) and np.array_equal(H(iv), want_iv)


@pytest.mark.parametrize("layout", ["lbvh", "relabelled"])
def test_refit_stored_tree(gpu, oracle, layout):
    """The element-box refit of a stored tree: the LBVH built from the
    element centres (treelet path), and the same tree with its internal ids
    permuted (generic fallback); both bit-exact against the oracle."""
    from test_gpu_parity import relabel_internal
    import del_msh_dlpack.PolyhedronMesh as PM

    off, iv, x = primitives.mixed_hex_grid(20, 16, 12, seed=5, jitter=0.2)
    w_bn, _ = oracle.build_points(oracle.polyhedron_elem2center(off, iv, x))
    bn = w_bn if layout == "lbvh" else relabel_internal(w_bn, 8)
    aabb = torch.empty((bn.shape[0], 6), dtype=torch.float32, device=gpu)
    PM.bvhnode2aabb_from_bvhnodes(G(off, gpu), G(iv, gpu), G(x, gpu), G(bn, gpu), aabb)
    assert bits_equal(H(aabb), oracle.polyhedron_refit(bn, off, iv, x))
    assert _lib.device_errors() == 0


def test_search_in_morton_order(gpu, oracle):
    """100 000 queries (>= 65536): the search runs in the Morton order of the
    query points; every result still lands at its own query, bit-exact."""
    off, iv, x = primitives.mixed_hex_grid(16, 12, 10, seed=7, jitter=0.2)
    rng = np.random.default_rng(8)
    q = rng.uniform([-0.5, -0.5, -0.5], [16.5, 12.5, 10.5], (100_000, 3)).astype(np.float32)
    elem = _check_mesh(gpu, oracle, off, iv, x, q)
    assert (elem != 0xFFFFFFFF).mean() > 0.7
