"""Pins of the oracle's mixed polyhedron mesh functions (del-msh-cpu/src/
polyhedron_mesh.rs, elem2center.rs, bvhnode2aabb3.rs) and of the written
parametric-coordinate rule the kernels share: closed forms on affine
elements, reproduction of the query point from the returned parameters,
the tiling of a mixed grid (every point in exactly one element), search
through the BVH against brute force, and the reference test's linear-field
interpolation on its own CFD mesh (del-msh-dlpack/tests/test_mix_mesh.py:
62-93, asset/cfd_mesh.txt, kept in tests/golden)."""

import pathlib

import numpy as np
import pytest

from del_msh_dlpack import primitives

GOLDEN = pathlib.Path(__file__).parent / "golden"


def load_cfd_mesh(path):
    """The text format of asset/cfd_mesh.txt: '# ...' headers, node lines
    'id x y z', then element lines 'id v0 v1 ...' (1-based) grouped as tet,
    pyramid, prism, hex."""
    vtx, groups, cur = [], {}, None
    for line in path.read_text().splitlines():
        if line.startswith("#"):
            for key, n in (("tetra", 4), ("pyramid", 5), ("prism", 6), ("hexahedra", 8)):
                if key in line:
                    cur = n
                    groups.setdefault(n, [])
            if "nodes" in line:
                cur = 0
            continue
        tok = line.split()
        if not tok:
            continue
        if cur == 0:
            vtx.append([float(v) for v in tok[1:4]])
        elif cur:
            groups[cur].append([int(v) - 1 for v in tok[1:]])
    elems = [e for n in (4, 5, 6, 8) for e in groups.get(n, [])]
    off = np.zeros(len(elems) + 1, np.uint32)
    off[1:] = np.cumsum([len(e) for e in elems])
    return off, np.concatenate([np.asarray(e, np.uint32) for e in elems]), np.asarray(vtx, np.float32)


def _bvh(oracle, off, iv, x):
    cog = oracle.polyhedron_elem2center(off, iv, x)
    bn, _ = oracle.build_points(cog)
    return bn, oracle.polyhedron_refit(bn, off, iv, x)


def _recon(oracle, off, iv, x, e, p):
    n = int(off[e + 1] - off[e])
    nodes = iv[off[e]:off[e + 1]]
    if n == 4:
        sf = np.array([p[0], p[1], p[2], 1 - p[0] - p[1] - p[2]], np.float64)
    else:
        sf = oracle.shape_functions(n, float(p[0]), float(p[1]), float(p[2])).astype(np.float64)
    return sf @ x[nodes].astype(np.float64)


def test_elem2center_and_refit(oracle):
    off, iv, x = primitives.mixed_hex_grid(3, 3, 2, seed=2, jitter=0.1)
    cog = oracle.polyhedron_elem2center(off, iv, x)
    for e in range(off.size - 1):
        acc = np.zeros(3, np.float32)
        for v in iv[off[e]:off[e + 1]]:
            acc = (acc + x[v]).astype(np.float32)
        assert np.array_equal(cog[e], (acc * np.float32(1.0 / (off[e + 1] - off[e]))).astype(np.float32))
    bn, aabb = _bvh(oracle, off, iv, x)
    n = off.size - 1
    for k in range(n):
        e = bn[n - 1 + k, 1]
        pts = x[iv[off[e]:off[e + 1]]]
        assert np.array_equal(aabb[n - 1 + k], np.concatenate([pts.min(0), pts.max(0)]))
    used = x[np.unique(iv)]
    assert np.array_equal(aabb[0], np.concatenate([used.min(0), used.max(0)]))


def test_param_closed_forms(oracle):
    # unit hex at [2,3]x[1,2]x[0,1]: r = 2 (q - lo) - 1 per axis
    cube = np.array([[2, 1, 0], [3, 1, 0], [3, 2, 0], [2, 2, 0], [2, 1, 1], [3, 1, 1], [3, 2, 1],
                     [2, 2, 1]], np.float32)
    rng = np.random.default_rng(0)
    for q in rng.uniform([2, 1, 0], [3, 2, 1], (50, 3)).astype(np.float32):
        p = oracle.param_coord(q, np.arange(8), cube)
        assert p is not None
        assert np.allclose(p, 2 * (q - [2, 1, 0]) - 1, atol=2e-6)
    assert oracle.param_coord(np.array([3.5, 1.5, 0.5], np.float32), np.arange(8), cube) is None
    # tet: barycentrics against a float64 solve
    tet = rng.standard_normal((4, 3)).astype(np.float32)
    for _ in range(50):
        w = rng.dirichlet(np.ones(4))
        q = (w @ tet.astype(np.float64)).astype(np.float32)
        p = oracle.param_coord(q, np.arange(4), tet)
        A = np.vstack([tet.T.astype(np.float64), np.ones(4)])
        want = np.linalg.solve(A, np.append(q.astype(np.float64), 1.0))
        assert p is not None and np.allclose(p, want[:3], atol=1e-5)
    # shape functions are a partition of unity
    for nn in (5, 6, 8):
        for r, s, t in rng.uniform(0, 0.5, (10, 3)):
            assert abs(float(oracle.shape_functions(nn, r, s, t).astype(np.float64).sum()) - 1) < 1e-6


@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_search_tiles_and_reconstructs(oracle, jitter):
    off, iv, x = primitives.mixed_hex_grid(4, 3, 3, seed=5, jitter=jitter)
    bn, aabb = _bvh(oracle, off, iv, x)
    rng = np.random.default_rng(1)
    q = rng.uniform([0.3, 0.3, 0.3], [3.7, 2.7, 2.7], (400, 3)).astype(np.float32)
    elem, par = oracle.polyhedron_search(off, iv, x, bn, aabb, q)
    n = off.size - 1
    counts = np.zeros(len(q), int)
    single = {}
    for w in range(len(q)):
        hits = [e for e in range(n) if oracle.param_coord(q[w], iv[off[e]:off[e + 1]], x) is not None]
        counts[w] = len(hits)
        if len(hits) == 1:
            single[w] = hits[0]
    if jitter == 0.0:
        assert np.all(counts >= 1) and np.mean(counts == 1) > 0.97
    for w, e in single.items():
        assert elem[w] == e
    found = np.nonzero(elem != 0xFFFFFFFF)[0]
    assert found.size > 0.9 * len(q)
    for w in found:
        assert np.allclose(_recon(oracle, off, iv, x, int(elem[w]), par[w]), q[w], atol=2e-5)
    assert np.all(par[elem == 0xFFFFFFFF] == 0)


def test_linear_field_on_reference_cfd_mesh(oracle):
    """test_mix_mesh.py:62-93 on the reference's own mesh file: a linear field
    is reproduced at points inside the mesh (error < 1e-4)."""
    off, iv, x = load_cfd_mesh(GOLDEN / "cfd_mesh.txt")
    assert np.array_equal(np.diff(off), [4, 5, 6, 8])
    A = np.array([[1.0, 2.0, 3.0], [0.0, 1.0, -1.0], [2.0, -1.0, 0.0]], np.float32)
    b = np.array([1.0, -1.0, 0.5], np.float32)
    vals = x @ A.T + b
    rng = np.random.default_rng(0)
    q = (rng.random((400, 3)) * [1.0, 1.0, 3.0] + [0.0, 0.0, -2.0]).astype(np.float32)
    bn, aabb = _bvh(oracle, off, iv, x)
    elem, par = oracle.polyhedron_search(off, iv, x, bn, aabb, q)
    inside = elem != 0xFFFFFFFF
    assert inside.sum() > 100 and set(np.unique(elem[inside])) == {0, 1, 2, 3}
    got = oracle.polyhedron_interpolate(off, iv, vals, elem, par)
    assert np.abs(got[inside] - (q[inside] @ A.T + b)).max() < 1e-4
    assert np.all(got[~inside] == 0)
