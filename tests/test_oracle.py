"""Pins for the CPU oracle (runs without a GPU).

Each check comes from outside the oracle: the reference's literal known-answer
tests, its own validators, closed forms, or brute force on small inputs.
"""

import json
import math
import pathlib

import numpy as np
import pytest

from del_msh_dlpack import primitives

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------------------
# Morton codes
# ---------------------------------------------------------------------------
def test_morton_reference_kats(oracle):
    kats = json.loads((GOLDEN / "morton_kats.json").read_text())
    L = oracle.lib()
    for x, want in kats["expand_bits2"]:
        assert L.orc_expand_bits2(x) == want
    for (x, y), want in kats["morton_code2"]:
        assert L.orc_morton_code2(x, y) == want
    for x, want in kats["expand_bits3"]:
        assert L.orc_expand_bits3(x) == want
    for (x, y, z), want in kats["morton_code3"]:
        assert L.orc_morton_code3(x, y, z) == want


def _interleave(bits_per_axis, ix, iy, iz=None):
    """Brute-force bit interleaver: x takes the highest bit of each group."""
    code = 0
    dims = [ix, iy] if iz is None else [ix, iy, iz]
    nd = len(dims)
    for b in range(bits_per_axis):
        for k, v in enumerate(dims):
            if (v >> b) & 1:
                code |= 1 << (nd * b + (nd - 1 - k))
    return code


def test_expand_bits3_exhaustive(oracle):
    """All 1024 inputs: the oracle's shift form equals a brute-force spread and
    the multiply form used by the reference CUDA kernel (mortons.h:14-22)."""
    L = oracle.lib()
    for v in range(1024):
        got = L.orc_expand_bits3(v)
        assert got == _interleave(10, v, 0, 0) >> 2
        m = v
        m = (m * 0x00010001) & 0xFF0000FF
        m = (m * 0x00000101) & 0x0F00F00F
        m = (m * 0x00000011) & 0xC30C30C3
        m = (m * 0x00000005) & 0x49249249
        assert got == m


def test_expand_bits2_exhaustive(oracle):
    L = oracle.lib()
    for v in range(0, 65536, 7):
        assert L.orc_expand_bits2(v) == _interleave(16, v, 0) >> 1


def test_morton63_against_interleave(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(0)
    vals = [0, 1, (1 << 21) - 1, 1 << 20] + [int(v) for v in rng.integers(0, 1 << 21, 200)]
    for v in vals:
        assert L.orc_expand_bits3_u64(v) == _interleave(21, v, 0, 0) >> 2
    for _ in range(200):
        x, y, z = rng.random(3).astype(np.float32)
        ix, iy, iz = (min(int(np.float32(c) * np.float32(2097152.0)), 2097151) for c in (x, y, z))
        assert L.orc_morton_code3_u64(float(x), float(y), float(z)) == _interleave(21, ix, iy, iz)
    # corners
    assert L.orc_morton_code3_u64(1.0, 1.0, 1.0) == (1 << 63) - 1
    assert L.orc_morton_code3_u64(0.0, 0.0, 0.0) == 0
    assert L.orc_morton_code3_u64(1.0, 0.0, 0.0) == _interleave(21, (1 << 21) - 1, 0, 0)


def test_morton30_quantization_against_interleave(oracle):
    """clamp(x * 1024, 0, 1023) as u32 with NaN -> 0 (Rust clamp + saturating cast)."""
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for _ in range(500):
        c = rng.uniform(-0.2, 1.2, 3).astype(np.float32)
        q = []
        for v in c:
            s = np.float32(v) * np.float32(1024.0)
            q.append(int(min(max(s, 0.0), 1023.0)))
        assert L.orc_morton_code3(*map(float, c)) == _interleave(10, *q)
    assert L.orc_morton_code3(float("nan"), -1.0, 2.0) == _interleave(10, 0, 0, 1023)


def test_vtx2morton_transform_equivalence(oracle):
    """test_mortons.py:68-98: the transform applied inside the kernel equals
    pre-transformed points with the identity."""
    rng = np.random.default_rng(0)
    vtx = (rng.random((1000, 3), dtype=np.float32) * 3.0 - 1.0).astype(np.float32)
    m = np.diag([1 / 3, 1 / 3, 1 / 3, 1.0]) @ np.array(
        [[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [0, 0, 0, 1.0]])
    m = m.astype(np.float32)
    c0 = oracle.vtx2morton(vtx, oracle.colmajor(m))
    vtx_unit = (np.concatenate([vtx, np.ones((1000, 1), np.float32)], 1) @ m.T)[:, :3]
    c1 = oracle.vtx2morton(vtx_unit.astype(np.float32), oracle.colmajor(np.eye(4)))
    assert np.array_equal(c0, c1)


# ---------------------------------------------------------------------------
# sort and topology
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_argsort_is_stable_sort(oracle, dtype):
    rng = np.random.default_rng(2)
    for n in [1, 2, 7, 1000, 5000]:
        keys = rng.integers(0, 50, n).astype(dtype)  # many duplicates
        idx, sk = oracle.argsort(keys)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(idx, ref.astype(np.uint32))
        assert np.array_equal(sk, keys[ref])


def _topdown(codes, nbits):
    """Independent top-down radix tree: split a range at the last index that
    shares more than the range's common prefix; an all-equal range splits at
    its first element (mortons.rs:303-305).  Node ids follow Karras: a left
    child internal node is numbered by its last index, a right child by its
    first; leaves are n-1+k."""
    n = len(codes)
    bn = np.zeros((2 * n - 1, 3), np.uint32)
    bn[0, 0] = 0xFFFFFFFF

    def clz(v):
        return nbits if v == 0 else nbits - int(v).bit_length()

    stack = [(0, n - 1, 0)]
    while stack:
        a, b, node = stack.pop()
        p = clz(int(codes[a]) ^ int(codes[b]))
        if p == nbits:
            s = a
        else:
            s = a
            for k in range(a, b):
                if clz(int(codes[a]) ^ int(codes[k])) > p:
                    s = k
        left = (n - 1 + s) if s == a else s
        right = (n - 1 + s + 1) if s + 1 == b else s + 1
        bn[node, 1], bn[node, 2] = left, right
        bn[left, 0] = node
        bn[right, 0] = node
        if s == a:
            bn[left, 2] = 0xFFFFFFFF
        else:
            stack.append((a, s, left))
        if s + 1 == b:
            bn[right, 2] = 0xFFFFFFFF
        else:
            stack.append((s + 1, b, right))
    return bn


@pytest.mark.parametrize("dups", [False, True])
def test_topology_matches_topdown_construction(oracle, dups):
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(2, 120))
        hi = 40 if dups else (1 << 30)
        codes = np.sort(rng.integers(0, hi, n)).astype(np.uint32)
        if not dups:
            codes = np.unique(codes)
            n = len(codes)
            if n < 2:
                continue
        idx2obj = rng.permutation(n).astype(np.uint32)
        bn = oracle.bvhnodes_from_sorted(idx2obj, codes)
        td = _topdown(codes, 32)
        td[n - 1:, 1] = idx2obj  # leaf rows hold the object index
        assert np.array_equal(bn[:, 1:], td[:, 1:]), trial
        assert np.array_equal(bn[:, 0], td[:, 0]), trial


def test_topology_u64_matches_topdown(oracle):
    rng = np.random.default_rng(4)
    for trial in range(30):
        n = int(rng.integers(2, 100))
        codes = np.sort(rng.integers(0, 1 << 63, n, dtype=np.uint64))
        if trial % 3 == 0:
            codes = np.sort(np.repeat(codes[: max(1, n // 4)], 4))[:n]
            n = len(codes)
        idx2obj = np.arange(n, dtype=np.uint32)
        bn = oracle.bvhnodes_from_sorted(idx2obj, codes)
        td = _topdown(codes, 64)
        td[n - 1:, 1] = idx2obj
        assert np.array_equal(bn, td), trial


def test_topology_reference_validators(oracle):
    """bvhnodes_morton.rs:52-75: 100k random points pass the reference's own
    check_morton_code_range_split / check_binary_radix_tree / check_bvh_topology."""
    rng = np.random.default_rng(5)
    pts = rng.random((100_000, 3), dtype=np.float32)
    codes = oracle.vtx2morton(pts, oracle.colmajor(np.eye(4)))
    idx, sk = oracle.argsort(codes)
    assert np.all(sk[:-1] <= sk[1:])
    assert oracle.check_morton_code_range_split(sk) == 0
    bn = oracle.bvhnodes_from_sorted(idx, sk)
    assert oracle.check_binary_radix_tree(bn, sk) == 0
    assert oracle.check_bvh_topology(bn, len(sk)) == 0


def test_topology_validator_catches_planted_error(oracle):
    rng = np.random.default_rng(6)
    codes = np.sort(rng.integers(0, 1 << 30, 200)).astype(np.uint32)
    bn = oracle.bvhnodes_from_sorted(np.arange(200, dtype=np.uint32), codes)
    bad = bn.copy()
    bad[0, 1], bad[0, 2] = bad[0, 2], bad[0, 1]
    assert oracle.check_binary_radix_tree(bad, codes) != 0


# ---------------------------------------------------------------------------
# refit
# ---------------------------------------------------------------------------
def _leaf_sets(bn, n):
    out = {}

    def rec(node):
        if bn[node, 2] == 0xFFFFFFFF:
            out[node] = [int(bn[node, 1])]
            return out[node]
        s = rec(int(bn[node, 1])) + rec(int(bn[node, 2]))
        out[node] = s
        return s

    rec(0)
    return out


def test_refit_equals_subtree_minmax(oracle):
    tri, vtx = primitives.torus_zup(1.0, 0.3, 12, 8)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    sets = _leaf_sets(bn, tri.shape[0])
    for node, leaves in sets.items():
        pts = vtx[tri[leaves].reshape(-1)]
        assert np.array_equal(aabb[node, :3], pts.min(0)), node
        assert np.array_equal(aabb[node, 3:], pts.max(0)), node


def test_refit_sphere_root_box(oracle):
    """del-msh-numpy/tests/test_bvh.py:10-20: root box of a unit sphere is [-1,1]^3."""
    tri, vtx = primitives.sphere_yup(1.0, 32, 32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    assert np.allclose(aabb[0], [-1, -1, -1, 1, 1, 1], atol=1e-5)
    assert oracle.check_bvh_topology(bn, tri.shape[0]) == 0


def test_refit_points(oracle):
    rng = np.random.default_rng(7)
    pts = rng.random((3000, 3), dtype=np.float32)
    bn, aabb = oracle.build_points(pts)
    assert oracle.check_bvh_topology(bn, 3000) == 0
    assert np.array_equal(aabb[0, :3], pts.min(0)) and np.array_equal(aabb[0, 3:], pts.max(0))
    leaves = bn[2999:, 1]
    assert np.array_equal(aabb[2999:, :3], pts[leaves])


# ---------------------------------------------------------------------------
# ray-cast
# ---------------------------------------------------------------------------
def _brute_first_hit(oracle, tri, vtx, org, dirv):
    best = None
    for it in range(tri.shape[0]):
        r = oracle.ray_tri(vtx[tri[it, 0]], vtx[tri[it, 1]], vtx[tri[it, 2]], org, dirv)
        if r is not None and (best is None or r[0] < best[0]):
            best = (r[0], it)
    return best


def test_first_hit_equals_brute_force(oracle):
    """The ordered BVH search returns the nearest hit of the brute-force loop
    (the reference's commented-out brute force, pix2tri.cu:29-41)."""
    tri, vtx = primitives.torus_zup(1.0, 0.35, 24, 12)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    rng = np.random.default_rng(8)
    mismatches = 0
    for _ in range(300):
        org = rng.uniform(-2, 2, 3).astype(np.float32)
        org[2] = 3.0
        tgt = rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        dirv = (tgt - org).astype(np.float32)
        got = oracle.first_intersection_ray(org, dirv, tri, vtx, bn, aabb)
        want = _brute_first_hit(oracle, tri, vtx, org, dirv)
        if want is None:
            assert got is None
            continue
        assert got is not None
        assert got[0] == want[0]
        if got[1] != want[1]:
            mismatches += 1  # a tie in t (shared edge): both triangles have the same t
            r = oracle.ray_tri(vtx[tri[got[1], 0]], vtx[tri[got[1], 1]], vtx[tri[got[1], 2]], org, dirv)
            assert r is not None and r[0] == want[0]
    assert mismatches <= 3


def test_sphere_hits_inside_and_twice(oracle):
    """search_bvh3.rs:248-304: rays along -z over [-1,1]^2 hit a 128x128 unit
    sphere inside the unit ball, and the line crosses exactly two triangles."""
    tri, vtx = primitives.sphere_yup(1.0, 128, 128)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    rng = np.random.default_rng(0)
    p0, p1, p2 = vtx[tri[:, 0]], vtx[tri[:, 1]], vtx[tri[:, 2]]
    lo = np.minimum(np.minimum(p0, p1), p2)
    hi = np.maximum(np.maximum(p0, p1), p2)
    for _ in range(1000):
        x, y = rng.random(2) * 2.0 - 1.0
        org = np.array([x, y, 1.0], np.float32)
        dirv = np.array([0, 0, -1.0], np.float32)
        res = oracle.first_intersection_ray(org, dirv, tri, vtx, bn, aabb)
        # all hits by brute force over the triangles whose xy box contains (x, y)
        org_far = np.array([x, y, 3.0], np.float32)
        cand = np.nonzero((lo[:, 0] <= x) & (x <= hi[:, 0]) & (lo[:, 1] <= y) & (y <= hi[:, 1]))[0]
        n_hits = sum(oracle.ray_tri(p0[k], p1[k], p2[k], org_far, dirv) is not None for k in cand)
        if res is not None:
            t = res[0]
            p = org + t * dirv
            assert np.linalg.norm(p) < 1.0
            assert n_hits == 2


def test_raycast_coverage_fraction(oracle):
    """test_trimesh.py:54-86: a radius-0.8 sphere under the identity ndc2world
    covers pi 0.8^2 / 4 of a 300x300 image within 1.3e-4."""
    tri, vtx = primitives.sphere_yup(0.8, 64, 32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(np.eye(4)), 300, 300, num_threads=4)
    ratio = np.count_nonzero(pix != 0xFFFFFFFF) / 300.0 / 300.0
    assert abs(ratio - 0.8 * 0.8 * math.pi * 0.25) < 1.3e-4


def test_raycast_t_closed_form_sphere(oracle):
    """Orthographic rays from z = +2 along -4 z hit the unit sphere at
    z = sqrt(1 - x^2 - y^2): t = (2 - z) / 4 up to the tessellation gap."""
    tri, vtx = primitives.sphere_yup(1.0, 64, 128)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    T = oracle.colmajor(np.diag([1.2, 1.2, 2.0, 1.0]))
    pix, t, _ = oracle.pix2tri(tri, vtx, bn, aabb, T, 64, 64, num_threads=4)
    ih, iw = np.nonzero(pix != 0xFFFFFFFF)
    x = (2.0 * (iw + 0.5) / 64 - 1.0) * 1.2
    y = (1.0 - 2.0 * (ih + 0.5) / 64) * 1.2
    r2 = x * x + y * y
    inner = r2 < 0.8
    z = np.sqrt(np.clip(1.0 - r2, 0.0, None))
    want = (2.0 - z) / 4.0
    got = t[ih, iw]
    assert np.max(np.abs(got[inner] - want[inner])) < 2e-3


def test_ray_aabb_and_tri_closed_forms(oracle):
    box = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    assert oracle.ray_aabb(box, [0, 0, 5], [0, 0, -1]) == (4.0, 6.0)
    assert oracle.ray_aabb(box, [2, 0, 5], [0, 0, -1]) is None      # parallel axis outside
    assert oracle.ray_aabb(box, [0, 0, 5], [0, 0, 1]) is None       # box behind the ray
    assert oracle.ray_aabb(box, [0, 0, 0], [1, 0, 0]) == (-1.0, 1.0)  # origin inside
    r = oracle.ray_tri([0, 0, 0], [1, 0, 0], [0, 1, 0], [0.25, 0.25, 1], [0, 0, -1])
    assert r is not None and r[0] == 1.0 and np.allclose(r[1], [0.5, 0.25, 0.25])
    assert oracle.ray_tri([0, 0, 0], [1, 0, 0], [0, 1, 0], [0.25, 0.25, -1], [0, 0, -1]) is None


def test_mat4_inverse_against_numpy(oracle):
    rng = np.random.default_rng(9)
    for _ in range(50):
        m = rng.normal(size=(4, 4)).astype(np.float32)
        inv = oracle.mat4_inverse(oracle.colmajor(m))
        want = np.linalg.inv(m.astype(np.float64))
        got = inv.reshape(4, 4).T
        assert np.allclose(got, want, rtol=1e-3, atol=1e-3 * np.abs(want).max())
    assert oracle.mat4_inverse(np.zeros(16, np.float32)) is None


def test_pix2depth_plane_closed_form(oracle):
    """A square at z = 0.25 under ndc2world = diag(1,1,1,1): ndc z = world z,
    depth = (0.25 + 1) / 2 on every covered pixel."""
    vtx = np.array([[-0.5, -0.5, 0.25], [0.5, -0.5, 0.25], [0.5, 0.5, 0.25], [-0.5, 0.5, 0.25]], np.float32)
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    T = oracle.colmajor(np.eye(4))
    pix, t, _ = oracle.pix2tri(tri, vtx, bn, aabb, T, 32, 32)
    d = oracle.pix2depth(pix, tri, vtx, T, 32, 32)
    cov = pix != 0xFFFFFFFF
    assert cov.sum() == 16 * 16
    assert np.allclose(d[cov], 0.625, atol=1e-6)
    assert np.all(d[~cov] == 0)
    assert np.allclose(t[cov], (1.0 - 0.25) / 2.0, atol=1e-7)


# ---------------------------------------------------------------------------
# edge-gradient backward
# ---------------------------------------------------------------------------
def _scene_sphere(oracle, img=96, radius=0.6, n_long=24, n_lat=48):
    tri, vtx = primitives.sphere_yup(radius, n_long, n_lat)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    w2n = np.diag([1.0, 1.0, 0.5, 1.0]).astype(np.float32)
    n2w = np.linalg.inv(w2n).astype(np.float32)
    n2p = np.array([[0.5 * img, 0, 0, 0.5 * img], [0, -0.5 * img, 0, 0.5 * img], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]], np.float32)
    w2p = (n2p @ w2n).astype(np.float32)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), img, img, num_threads=4)
    return tri, vtx, pix, w2p


def test_edge_gradient_translation_closed_form(oracle):
    """With L = sum_pix w(pix) occ(pix) and w = pixel x (resp. y), moving the
    whole mesh by d in pixel x changes L by the covered area (divergence
    theorem), so sum_v dL/dx_v = (pixels per world unit) * area."""
    img = 96
    tri, vtx, pix, w2p = _scene_sphere(oracle, img)
    occ = (pix != 0xFFFFFFFF).astype(np.float32)[..., None]
    area = occ.sum()
    iw = np.arange(img, dtype=np.float32)[None, :, None] + 0.5
    ih = np.arange(img, dtype=np.float32)[:, None, None] + 0.5
    gx = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(w2p), pix, occ, np.broadcast_to(iw, occ.shape).copy())
    gy = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(w2p), pix, occ, np.broadcast_to(ih, occ.shape).copy())
    sx = w2p[0, 0]
    sy = w2p[1, 1]  # negative: pixel y runs down
    assert abs(gx[:, 0].sum() / sx - area) < 0.05 * area
    assert abs(gy[:, 1].sum() / sy - area) < 0.05 * area
    # a uniform weight gives no net translation gradient
    g1 = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(w2p), pix, occ, np.ones_like(occ))
    assert abs(g1[:, 0].sum()) < 0.02 * abs(gx[:, 0]).sum()


def test_edge_gradient_nonzero_only_on_silhouette_triangles(oracle):
    tri, vtx, pix, w2p = _scene_sphere(oracle, 64)
    occ = (pix != 0xFFFFFFFF).astype(np.float32)[..., None]
    rng = np.random.default_rng(10)
    g = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(w2p), pix, occ,
                                 rng.normal(size=occ.shape).astype(np.float32))
    # only vertices of triangles adjacent to a foreground/background pixel pair move
    fg = pix != 0xFFFFFFFF
    edge = np.zeros_like(fg)
    edge[:-1, :] |= fg[:-1, :] != fg[1:, :]
    edge[1:, :] |= fg[:-1, :] != fg[1:, :]
    edge[:, :-1] |= fg[:, :-1] != fg[:, 1:]
    edge[:, 1:] |= fg[:, :-1] != fg[:, 1:]
    tris = np.unique(pix[edge & fg])
    allowed = np.zeros(vtx.shape[0], bool)
    allowed[tri[tris].reshape(-1)] = True
    moved = np.abs(g).sum(1) > 0
    assert moved.any()
    assert not np.any(moved & ~allowed)


# ---------------------------------------------------------------------------
# generators
# ---------------------------------------------------------------------------
def test_generator_counts():
    tri, vtx = primitives.sphere_yup(1.0, 64, 128)
    assert tri.shape == (16128, 3) and vtx.shape == (8066, 3)
    tri, vtx = primitives.sphere_yup(1.0, 126, 200)
    assert tri.shape == (50000, 3) and vtx.shape == (25002, 3)
    tri, vtx = primitives.torus_zup(1.0, 0.3, 16, 12)
    assert tri.shape == (2 * 16 * 12, 3)
    tri, vtx = primitives.cylinder_closed_end_yup(1.0, 2.0, 16, 5)
    assert tri.shape == (2 * 16 * 6, 3) and vtx.shape == (16 * 6 + 2, 3)
    tri, vtx = primitives.octahedral_sphere(3)
    assert tri.shape == (8 * 4 ** 3, 3) and vtx.shape == (4 * 4 ** 3 + 2, 3)
    assert tri.max() < vtx.shape[0]


@pytest.mark.parametrize("maker", [
    lambda: primitives.sphere_yup(0.8, 64, 32),
    lambda: primitives.octahedral_sphere(4, radius=0.8),
])
def test_sphere_area_and_closedness(maker):
    tri, vtx = maker()
    p = vtx.astype(np.float64)
    a = 0.5 * np.linalg.norm(np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1).sum()
    assert abs(a - 4 * math.pi * 0.8 ** 2) / (4 * math.pi * 0.8 ** 2) < 0.01
    # closed 2-manifold: every undirected edge is shared by exactly two triangles
    e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)
