"""CPU: pins of the oracle's triangle centroid, unit-cube fit and the
per-vertex distribution of the edge-gradient backward against what the
mathematics fixes (closed forms, error bounds of f32 arithmetic, the area
gradient of a triangle and the dilation identity), independently of the
oracle's own formulas.

References: centroid del-msh-cpu/src/elem2center.rs:8-37; unit-cube fit
del-msh-dlpack/del_msh_dlpack/Mat44/torch.py:120-140; edge-gradient
backward del-msh-cpu/src/rasterized_edge_gradient.rs:434-603 (the
barycentric distribution at :503,518,581,596)."""

import numpy as np
import pytest

from del_msh_dlpack import primitives

U32MAX = 0xFFFFFFFF
F32_EPS = float(np.finfo(np.float32).eps)


# ---------------------------------------------------------------------------
# triangle centroid
# ---------------------------------------------------------------------------
def test_centroid_hand_checked(oracle):
    """Three vertices that all differ on every axis; the sums are exact and
    the products with RN(1/3) round back to the integers."""
    vtx = np.array([[1, 10, 100], [2, 20, 200], [6, 30, 300],
                    [3, 0, 0], [0, 6, 0], [0, 0, 9]], np.float32)
    tri = np.array([[0, 1, 2], [3, 4, 5], [2, 0, 1]], np.uint32)
    c = oracle.tri2centroid(tri, vtx)
    assert np.array_equal(c, np.array([[3, 20, 200], [1, 2, 3], [3, 20, 200]], np.float32))


def test_centroid_within_ulps_of_f64(oracle):
    """Same-sign coordinates: two f32 additions, a product with RN(1/3) and
    its rounding stay within 2 ulp of the exact centroid of the f32 inputs."""
    rng = np.random.default_rng(11)
    vtx = rng.uniform(1.0, 2.0, size=(3000, 3)).astype(np.float32)
    vtx *= np.float32(2.0) ** rng.integers(-6, 7, size=(3000, 1)).astype(np.float32)
    tri = rng.integers(0, 3000, size=(5000, 3)).astype(np.uint32)
    c = oracle.tri2centroid(tri, vtx).astype(np.float64)
    exact = vtx.astype(np.float64)[tri].sum(axis=1) / 3.0
    ulp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(c - exact) <= 2.0 * ulp)


def test_centroid_mixed_signs_error_bound(oracle):
    """Any signs: |c - exact| <= 3 eps (|p0| + |p1| + |p2|) / 3 (two sums
    and two products, each within eps/2 of its operands' magnitude)."""
    rng = np.random.default_rng(12)
    vtx = rng.normal(size=(2000, 3)).astype(np.float32)
    tri = rng.integers(0, 2000, size=(4000, 3)).astype(np.uint32)
    c = oracle.tri2centroid(tri, vtx).astype(np.float64)
    p = vtx.astype(np.float64)[tri]
    exact = p.sum(axis=1) / 3.0
    bound = 3.0 * F32_EPS * np.abs(p).sum(axis=1) / 3.0
    assert np.all(np.abs(c - exact) <= bound)


def test_centroid_translation_covariance(oracle):
    """Moving every vertex by d moves every centroid by d (to the rounding
    bound of the f32 sums)."""
    rng = np.random.default_rng(13)
    vtx = rng.uniform(-1, 1, size=(500, 3)).astype(np.float32)
    tri = rng.integers(0, 500, size=(800, 3)).astype(np.uint32)
    d = np.array([3.25, -7.5, 0.125], np.float32)
    c0 = oracle.tri2centroid(tri, vtx).astype(np.float64)
    c1 = oracle.tri2centroid(tri, vtx + d).astype(np.float64)
    mag = np.abs(vtx.astype(np.float64) + d)[tri].sum(axis=1) / 3.0 + np.abs(vtx)[tri].sum(axis=1) / 3.0
    assert np.all(np.abs((c1 - c0) - d) <= 3.0 * F32_EPS * mag)


# ---------------------------------------------------------------------------
# unit-cube fit
# ---------------------------------------------------------------------------
def _apply(T_colmajor, p):
    M = T_colmajor.reshape(4, 4).T.astype(np.float64)
    return p.astype(np.float64) @ M[:3, :3].T + M[:3, 3]


def test_fit_two_points_closed_form(oracle):
    """Box (0,0,0)-(2,1,1): s = 1/2, centre (1, .5, .5), translation
    -c s + 1/2 = (0, .25, .25); every value is exact in f32."""
    T = oracle.fit_into_unit_cube(np.array([[0, 0, 0], [2, 1, 1]], np.float32))
    want = np.array([[0.5, 0, 0, 0], [0, 0.5, 0, 0.25], [0, 0, 0.5, 0.25], [0, 0, 0, 1]], np.float32)
    assert np.array_equal(T, want.T.reshape(-1))


def test_fit_longest_axis_to_unit_interval_and_centred(oracle):
    """The longest axis of the box maps to [0, 1]; every axis is centred on
    1/2; the scale is uniform; all points land in the unit cube."""
    rng = np.random.default_rng(14)
    for it in range(20):
        lo = rng.uniform(-5, 5, size=3)
        ext = rng.uniform(0.1, 4, size=3)
        p = (lo + ext * rng.uniform(0, 1, size=(777, 3))).astype(np.float32)
        T = oracle.fit_into_unit_cube(p)
        M = T.reshape(4, 4).T
        assert M[0, 0] == M[1, 1] == M[2, 2] and M[3, 3] == 1
        assert np.count_nonzero(M[:3, :3] - np.diag(np.diag(M[:3, :3]))) == 0
        mn, mx = p.min(0), p.max(0)
        k = int(np.argmax(mx.astype(np.float64) - mn))
        a, b = _apply(T, mn[None])[0], _apply(T, mx[None])[0]
        tol = 4 * F32_EPS * (1.0 + np.abs(M[:3, 3]).max() + np.abs(mn * M[0, 0]).max())
        assert abs(a[k] - 0.0) <= tol and abs(b[k] - 1.0) <= tol
        assert np.all(np.abs(0.5 * (a + b) - 0.5) <= tol)
        q = _apply(T, p)
        assert q.min() >= -tol and q.max() <= 1 + tol


# ---------------------------------------------------------------------------
# edge-gradient backward: per-vertex distribution
# ---------------------------------------------------------------------------
IMG = 512


def _ortho():
    """World xy in [-1, 1] -> pixels, row 0 at the top; rays along -z."""
    n2p = np.array([[0.5 * IMG, 0, 0, 0.5 * IMG], [0, -0.5 * IMG, 0, 0.5 * IMG], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]], np.float64)
    return np.eye(4, dtype=np.float32), n2p.astype(np.float32)


def _coverage_gradient(oracle, tri, vtx):
    n2w, w2p = _ortho()
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), IMG, IMG, num_threads=4)
    occ = (pix != U32MAX).astype(np.float32)[..., None]
    g = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(w2p), pix, occ, np.ones_like(occ))
    return g.astype(np.float64), occ.sum(), w2p.astype(np.float64)


def test_edge_gradient_single_triangle_area_gradient(oracle):
    """L = number of covered pixels ~ the triangle's area in pixels.  The
    area's gradient at vertex k is half the opposite edge turned outward:
    dA/dq_k = 1/2 (y_{k+1} - y_{k+2}, x_{k+2} - x_{k+1}) for a counter-
    clockwise triangle.  Each vertex's gradient from the edge pass matches
    it within 2 % (a scalene triangle, so a vertex given another vertex's
    barycentric weight misses by far more)."""
    vtx = np.array([[-0.7, -0.6, 0.0], [0.75, -0.35, 0.0], [-0.2, 0.72, 0.0]], np.float32)
    tri = np.array([[0, 1, 2]], np.uint32)
    g, area_pix, w2p = _coverage_gradient(oracle, tri, vtx)
    q = vtx.astype(np.float64) @ w2p[:2, :3].T + w2p[:2, 3]
    signed = 0.5 * ((q[1, 0] - q[0, 0]) * (q[2, 1] - q[0, 1]) - (q[2, 0] - q[0, 0]) * (q[1, 1] - q[0, 1]))
    assert abs(abs(signed) - area_pix) < 0.01 * abs(signed)
    for k in range(3):
        a, b = q[(k + 1) % 3], q[(k + 2) % 3]
        dq = np.sign(signed) * 0.5 * np.array([a[1] - b[1], b[0] - a[0]])
        want = np.array([dq[0] * w2p[0, 0], dq[1] * w2p[1, 1], 0.0])  # chain rule: J^T dq
        assert np.linalg.norm(g[k] - want) <= 0.02 * np.linalg.norm(want), (k, g[k], want)
    # translation moves the covered area rigidly: the vertex gradients sum to ~0
    assert np.linalg.norm(g.sum(0)) <= 0.01 * np.abs(g).sum()


@pytest.mark.parametrize("order", [(0, 1, 2), (1, 2, 0), (2, 1, 0)])
def test_edge_gradient_single_triangle_vertex_order(oracle, order):
    """Relabelling the vertices (and flipping the winding) permutes the
    per-vertex gradients and changes nothing else."""
    vtx = np.array([[-0.55, -0.66, 0.0], [0.8, -0.2, 0.0], [-0.4, 0.6, 0.0]], np.float32)
    g0, _, _ = _coverage_gradient(oracle, np.array([[0, 1, 2]], np.uint32), vtx)
    g1, _, _ = _coverage_gradient(oracle, np.array([order], np.uint32), vtx)
    assert np.allclose(g0, g1, rtol=1e-5, atol=1e-6 * np.abs(g0).max())


def test_edge_gradient_dilation_identity_on_sphere(oracle):
    """Scaling world x and y by (1 + e) about the camera axis scales the
    silhouette's pixel area by (1 + e)^2, so sum_v (x_v dL/dx_v + y_v dL/dy_v)
    = 2 A for L = covered pixel count (Euler's identity for a function
    homogeneous of degree 2)."""
    tri, vtx = primitives.sphere_yup(0.7, 48, 96)
    g, area_pix, _ = _coverage_gradient(oracle, tri, vtx)
    p = vtx.astype(np.float64)
    moment = (p[:, 0] * g[:, 0] + p[:, 1] * g[:, 1]).sum()
    assert abs(moment - 2.0 * area_pix) <= 0.02 * 2.0 * area_pix
    assert np.abs(g[:, 2]).max() == 0.0  # orthographic along z: depth does not move pixels
