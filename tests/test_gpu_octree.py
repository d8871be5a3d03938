"""GPU parity of the quad/oct-tree and the sorted-array helpers against the
oracle (bit-exact: integer work, dyadic centres, the offset aggregate in the
CPU's order and the node aggregate in the kernels' slot order; the node
aggregate also against the CPU's point-by-point order, exactly for exact
partial sums as in the reference test, else to rounding), through the
reference's Python entry points, at the reference test's size (1M points,
del-msh-dlpack/tests/test_quad_oct_tree.py)."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib
import del_msh_dlpack.Array1D.torch as Array1D
import del_msh_dlpack.Mortons.torch as Mortons
import del_msh_dlpack.OffsetArray.torch as OffsetArray
import del_msh_dlpack.QuadOctTree.torch as QuadOctTree

pytestmark = pytest.mark.gpu


def H(t):
    return t.detach().cpu().numpy()


def _sorted_codes(gpu, n, num_dim, seed):
    torch.manual_seed(seed)
    vtx2co = torch.rand((n, num_dim)).to(gpu)
    vtx2morton = Mortons.make_vtx2morton_from_vtx2co(vtx2co, torch.eye(num_dim + 1).to(gpu))
    idx2vtx, idx2morton = Array1D.argsort(vtx2morton)
    return vtx2co, idx2vtx, idx2morton


@pytest.mark.parametrize("num_dim", [2, 3])
def test_tree_construction(gpu, oracle, num_dim):
    _, _, idx2morton = _sorted_codes(gpu, 1_000_000, num_dim, 0)
    idx2jdx, jdx2morton, jdx2idx_offset = Array1D.unique_for_sorted_array(idx2morton)
    w_i2j, w_j2m, w_off = oracle.unique_for_sorted(H(idx2morton))
    assert np.array_equal(H(idx2jdx), w_i2j)
    assert np.array_equal(H(jdx2morton), w_j2m)
    assert np.array_equal(H(jdx2idx_offset), w_off)
    assert Array1D.has_duplicate_in_sorted_array(idx2morton) == bool(
        np.any(np.diff(H(idx2morton).astype(np.int64)) == 0))
    assert not Array1D.has_duplicate_in_sorted_array(jdx2morton)
    tree = QuadOctTree.QuadOctTree()
    tree.construct_from_idx2morton(jdx2morton, num_dim, True)
    codes = H(jdx2morton)
    bn, dep, b2o, i2b = oracle.octree_bnodes(codes, num_dim)
    assert np.array_equal(H(tree.bnodes), bn)
    assert np.array_equal(H(tree.bnode2depth), dep)
    assert np.array_equal(H(tree.bnode2onode), b2o)
    assert np.array_equal(H(tree.idx2bnode), i2b)
    on, o2d, o2c, i2o, i2c = oracle.octree_make_tree(bn, b2o, dep, i2b, codes, num_dim)
    assert np.array_equal(H(tree.onodes), on)
    assert np.array_equal(H(tree.onode2depth), o2d)
    assert np.array_equal(H(tree.onode2center), o2c)
    assert np.array_equal(H(tree.idx2onode), i2o)
    assert np.array_equal(H(tree.idx2center), i2c)
    assert _lib.device_errors() == 0


def test_tree_aggregation(gpu, oracle):
    """The reference test's pipeline (all-ones values, 1000 points) and random
    values at 1M points, both bit-exact with the oracle's serial sums."""
    for n, ones in ((1000, True), (1_000_000, False)):
        vtx2co, jdx2vtx, jdx2morton = _sorted_codes(gpu, n, 3, 0)
        jdx2idx, idx2morton, idx2jdx_offset = Array1D.unique_for_sorted_array(jdx2morton)
        rng = np.random.default_rng(1)
        vals = (np.ones((n, 3), np.float32) if ones
                else rng.standard_normal((n, 3)).astype(np.float32))
        vtx2val = torch.from_numpy(vals).to(gpu)
        idx2aggval = OffsetArray.aggregate(idx2jdx_offset, jdx2vtx, vtx2val)
        w_agg = oracle.offset_array_aggregate(H(idx2jdx_offset), H(jdx2vtx), vals)
        assert np.array_equal(H(idx2aggval), w_agg)
        tree = QuadOctTree.QuadOctTree()
        tree.construct_from_idx2morton(idx2morton, 3, False)
        onode2aggval = tree.aggregate(idx2aggval)
        assert np.array_equal(H(onode2aggval), oracle.octree_aggregate_tree(w_agg, H(tree.onodes)))
        w_cpu = oracle.octree_aggregate(w_agg, H(tree.idx2onode), H(tree.onodes))
        if ones:  # exact partial sums: the CPU's point-by-point order agrees bit for bit
            assert np.array_equal(H(onode2aggval), w_cpu)
            assert np.array_equal(H(onode2aggval)[0], [n, n, n])
        else:
            assert np.allclose(H(onode2aggval), w_cpu, rtol=1e-3, atol=1e-2)
    assert _lib.device_errors() == 0


def test_small_and_clustered_inputs(gpu, oracle):
    """Two points, points sharing a long prefix (the root's own depth is
    deeper than 0), and a 2D set on one quadrant."""
    cases = [np.array([1, 2], np.uint32),
             np.array([0x3FFFFFF0, 0x3FFFFFF1, 0x3FFFFFF7, 0x3FFFFFFF], np.uint32),
             np.unique(np.random.default_rng(2).integers(0, 1 << 28, 500).astype(np.uint32))]
    for codes in cases:
        for num_dim in (2, 3):
            if num_dim == 3 and codes.max() >= (1 << 30):
                continue
            tree = QuadOctTree.QuadOctTree()
            tree.construct_from_idx2morton(torch.from_numpy(codes).to(gpu), num_dim, True)
            bn, dep, b2o, i2b = oracle.octree_bnodes(codes, num_dim)
            on, o2d, o2c, i2o, i2c = oracle.octree_make_tree(bn, b2o, dep, i2b, codes, num_dim)
            assert np.array_equal(H(tree.onodes), on)
            assert np.array_equal(H(tree.onode2center), o2c)
            assert np.array_equal(H(tree.idx2center), i2c)
    assert _lib.device_errors() == 0
