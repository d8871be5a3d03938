"""Pins of the oracle's quad/oct-tree (del-msh-cpu/src/quad_oct_tree.rs) and
sorted-array helpers: the reference's morton2center cases
(quad_oct_tree.rs:65-75), its check_octree invariants (:229-336, restated),
a brute-force octree from recursive Morton-digit partitioning, and closed
forms for the aggregates."""

import numpy as np
import pytest


def _codes(oracle, n, num_dim, seed):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, num_dim)).astype(np.float32)
    m = oracle.vtx2morton(pts, oracle.colmajor(np.eye(num_dim + 1, dtype=np.float32)))
    return np.unique(np.sort(m)), pts


def test_morton2center_reference_cases(oracle):
    for p, num_dim, max_depth in (([0.321, 0.123], 2, 16), ([0.321, 0.123, 0.87321], 3, 10)):
        m = oracle.vtx2morton(np.array([p], np.float32),
                              oracle.colmajor(np.eye(num_dim + 1, dtype=np.float32)))[0]
        c = oracle.morton2center(m, num_dim, max_depth, max_depth)
        h = 1.0 / (1 << max_depth)
        assert np.all(np.abs(c - np.array(p, np.float32)) < h)
    # depth 0 is the unit cell's centre; depth 1 a quadrant's
    assert np.array_equal(oracle.morton2center(0, 3, 0, 10), [0.5, 0.5, 0.5])
    assert np.array_equal(oracle.morton2center(0b111 << 27, 3, 1, 10), [0.75, 0.75, 0.75])


def _brute_octree(codes, num_dim, max_depth):
    """Branching cells by recursive partition: a cell is a node when its points
    fall in at least two child cells; returns {(depth, prefix)} of the non-root
    nodes and, per point, the (depth, prefix) of its node."""
    nchild = 1 << num_dim
    nodes = set()
    parent = {}

    def digit(c, d):
        return (int(c) >> ((max_depth - 1 - d) * num_dim)) & (nchild - 1)

    def rec(lo, hi, d, owner, is_root):
        while d < max_depth and digit(codes[lo], d) == digit(codes[hi - 1], d):
            d += 1  # a single occupied child: no node here
        key = (d, int(codes[lo]) >> ((max_depth - d) * num_dim))
        if not is_root:
            nodes.add(key)
        me = key if not is_root else "root"
        s = lo
        while s < hi:
            e = s
            while e < hi and digit(codes[e], d) == digit(codes[s], d):
                e += 1
            if e - s == 1:
                parent[s] = me
            else:
                rec(s, e, d + 1, me, False)
            s = e

    rec(0, len(codes), 0, None, True)
    return nodes, parent


def _prefix_of_center(center, depth, num_dim, max_depth):
    # cell index per axis at this depth, interleaved back into a prefix
    idx = [int(np.floor(float(c) * (1 << depth))) for c in center]
    p = 0
    for b in range(depth - 1, -1, -1):
        for a in range(num_dim):
            p = (p << 1) | ((idx[a] >> b) & 1)
    return p


@pytest.mark.parametrize("num_dim", [2, 3])
def test_octree_matches_brute_force_and_invariants(oracle, num_dim):
    max_depth = 32 // num_dim
    codes, _ = _codes(oracle, 3000, num_dim, 1 + num_dim)
    bn, dep, b2o, i2b = oracle.octree_bnodes(codes, num_dim)
    on, o2d, o2c, i2o, i2c = oracle.octree_make_tree(bn, b2o, dep, i2b, codes, num_dim)
    num_onode = on.shape[0]
    nodes, parent = _brute_octree(codes, num_dim, max_depth)
    got = {(int(o2d[o]), _prefix_of_center(o2c[o], int(o2d[o]), num_dim, max_depth))
           for o in range(1, num_onode)}
    assert got == nodes and num_onode == len(nodes) + 1
    for i in range(len(codes)):
        o = int(i2o[i])
        want = parent[i]
        if want == "root":
            assert o == 0
        else:
            assert (int(o2d[o]), _prefix_of_center(o2c[o], int(o2d[o]), num_dim, max_depth)) == want
    # check_octree: every point reached once from the root; children nest
    nchild = 1 << num_dim
    seen = np.zeros(len(codes), np.int64)
    stack = [0]
    while stack:
        o = stack.pop()
        for c in on[o, 1:]:
            if c == 0xFFFFFFFF:
                continue
            if c >= num_onode:
                seen[c - num_onode] += 1
            else:
                assert on[c, 0] == o
                stack.append(int(c))
    assert np.all(seen == 1)
    for o in range(1, num_onode):
        p = on[o, 0]
        h, hp = 0.5 / (1 << int(o2d[o])), 0.5 / (1 << int(o2d[p]))
        assert o2d[o] > o2d[p]
        assert np.all(np.abs(o2c[o] - o2c[p]) + h < hp * (1 + 1e-7))
    assert on.shape[1] == nchild + 1


@pytest.mark.parametrize("num_dim", [2, 3])
def test_octree_aggregate_closed_forms(oracle, num_dim):
    codes, _ = _codes(oracle, 2000, num_dim, 7)
    bn, dep, b2o, i2b = oracle.octree_bnodes(codes, num_dim)
    on, o2d, o2c, i2o, i2c = oracle.octree_make_tree(bn, b2o, dep, i2b, codes, num_dim)
    n = len(codes)
    agg = oracle.octree_aggregate(np.ones((n, 2), np.float32), i2o, on)
    assert np.array_equal(agg[0], [n, n])
    # every node's count equals the points under it (brute force over prefixes)
    max_depth = 32 // num_dim
    for o in range(1, on.shape[0]):
        d = int(o2d[o])
        pref = _prefix_of_center(o2c[o], d, num_dim, max_depth)
        cnt = int(np.sum((codes.astype(np.int64) >> ((max_depth - d) * num_dim)) == pref))
        assert agg[o, 0] == cnt
    # values: the sum in ascending point order, checked against float64
    rng = np.random.default_rng(3)
    v = rng.standard_normal((n, 3)).astype(np.float32)
    agg = oracle.octree_aggregate(v, i2o, on)
    assert np.allclose(agg[0], v.astype(np.float64).sum(0), rtol=1e-4, atol=1e-3)
    # the kernels' slot order: the same sums as the point-by-point order when
    # every partial sum is exact, and within rounding otherwise
    tree = oracle.octree_aggregate_tree(v, on)
    assert np.allclose(tree, agg, rtol=1e-4, atol=1e-4)
    ints = rng.integers(-8, 9, (n, 2)).astype(np.float32)
    assert np.array_equal(oracle.octree_aggregate_tree(ints, on),
                          oracle.octree_aggregate(ints, i2o, on))


def test_unique_and_offset_aggregate(oracle):
    rng = np.random.default_rng(4)
    v = np.sort(rng.integers(0, 500, 5000).astype(np.uint32))
    idx2jdx, jdx2val, off = oracle.unique_for_sorted(v)
    u, first, inv = np.unique(v, return_index=True, return_inverse=True)
    assert np.array_equal(jdx2val, u)
    assert np.array_equal(off[:-1], first) and off[-1] == v.size
    assert np.array_equal(idx2jdx, inv)
    perm = rng.permutation(v.size).astype(np.uint32)
    vals = rng.standard_normal((v.size, 2)).astype(np.float32)
    agg = oracle.offset_array_aggregate(off, perm, vals)
    want = np.zeros_like(agg)
    for i in range(u.size):
        for j in range(off[i], off[i + 1]):
            want[i] = (want[i] + vals[perm[j]]).astype(np.float32)
    assert np.array_equal(agg, want)
