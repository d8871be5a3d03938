"""GPU parity: libdmsh's HIP kernels (called through the C ABI via the
del_msh_dlpack drop-in) against the CPU oracle on the same inputs.

Integer and min/max outputs must be bit-exact; the edge-gradient backward
(float atomics, order-dependent) is compared with a stated tolerance.
"""

import math

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Mortons.torch as Mortons
import del_msh_dlpack.Array1D.torch as Array1D
import del_msh_dlpack.TriMesh3.torch as TriMesh3
import del_msh_dlpack.Pix2Tri.torch as Pix2Tri
import del_msh_dlpack.Pix2Depth.torch as Pix2Depth
import del_msh_dlpack.RasterizedEdgeGradient.torch as RasterizedEdgeGradient

pytestmark = pytest.mark.gpu

U32MAX = 0xFFFFFFFF


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def no_device_errors():
    bits = _lib.device_errors()
    assert bits == 0, f"device error bits 0x{bits:x}"


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------------------
# Morton codes
# ---------------------------------------------------------------------------
def _morton_inputs(n, dim, seed):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-0.3, 1.3, (n, dim)).astype(np.float32)
    pts[:8] = np.array([[np.nan] * dim, [0.0] * dim, [1.0] * dim, [-0.0] * dim,
                        [np.inf] * dim, [-np.inf] * dim, [0.99999994] * dim, [1e-30] * dim],
                       np.float32)[:, :dim]
    return pts


@pytest.mark.parametrize("dim", [2, 3])
def test_vtx2morton_u32(gpu, oracle, dim):
    pts = _morton_inputs(100_003, dim, 11)
    m = np.eye(dim + 1, dtype=np.float32)
    m[:dim, :dim] *= np.float32(0.7)
    m[:dim, dim] = np.float32(0.1)
    got = H(Mortons.make_vtx2morton_from_vtx2co(G(pts, gpu), G(m, gpu)))
    want = oracle.vtx2morton(pts, m.T.reshape(-1))
    assert np.array_equal(got, want)
    no_device_errors()


def test_vtx2morton_u64(gpu, oracle):
    pts = _morton_inputs(100_003, 3, 12)
    m = np.array([[0.8, 0.1, 0, 0.05], [0, 0.9, 0, 0.02], [0.05, 0, 0.7, 0.1], [0, 0, 0, 1]], np.float32)
    got = H(Mortons.make_vtx2morton63_from_vtx2xyz(G(pts, gpu), G(m, gpu)))
    want = oracle.vtx2morton_u64(pts, m.T.reshape(-1))
    assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# argsort
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 255, 2048, 2049, 100_000, 1 << 20])
@pytest.mark.parametrize("spread", ["dups", "wide"])
def test_argsort_u32(gpu, oracle, n, spread):
    rng = np.random.default_rng(n)
    hi = 1000 if spread == "dups" else 1 << 32
    keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    idx, sk = Array1D.argsort(G(keys, gpu))
    widx, wsk = oracle.argsort(keys)
    assert np.array_equal(H(sk), wsk)
    assert np.array_equal(H(idx), widx)
    no_device_errors()


@pytest.mark.parametrize("key_bits", [40, 48, 56, 63, 64])
@pytest.mark.parametrize("groups", [None, 4000, 300, 1])
def test_argsort_u64_grouped(gpu, oracle, key_bits, groups):
    """Keys of five or more digits: top-digit passes, then the in-LDS finish
    per group of equal top bits.  groups=None: random keys (tiny groups);
    4000: ~250 keys per group with repeated low parts; 300: ~3300 per group,
    beyond the finish's window, so k_sort_groups sorts each of them; 1: one
    group of every key (one workgroup sorts all 10^6)."""
    n = 1_000_003
    rng = np.random.default_rng(key_bits * 7 + (groups or 1))
    mask = np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    if groups:
        shift = np.uint64(8 * (-(-key_bits // 8) - 3))
        top = rng.integers(0, groups, n, dtype=np.uint64) << shift
        low = rng.integers(0, 50, n, dtype=np.uint64)  # repeats inside a group
        keys = top | low
    keys &= mask
    idx, sk = Array1D.argsort(G(keys, gpu), key_bits=key_bits)
    widx, wsk = oracle.argsort(keys)
    assert np.array_equal(H(sk), wsk)
    assert np.array_equal(H(idx), widx)
    no_device_errors()


@pytest.mark.parametrize("n", [3, 4096, 300_001])
@pytest.mark.parametrize("key_bits", [63, 64, 17])
def test_argsort_u64(gpu, oracle, n, key_bits):
    rng = np.random.default_rng(n + key_bits)
    keys = rng.integers(0, 1 << 62, n, dtype=np.uint64) * np.uint64(2)
    keys[: n // 3] = keys[n // 2]  # a long duplicate run
    keys &= np.uint64((1 << key_bits) - 1) if key_bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    idx, sk = Array1D.argsort(G(keys, gpu), key_bits=key_bits)
    widx, wsk = oracle.argsort(keys)
    assert np.array_equal(H(sk), wsk)
    assert np.array_equal(H(idx), widx)


# ---------------------------------------------------------------------------
# topology, refit, whole build
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["u32", "u32dups", "u64", "u64dups", "u32runs", "u64runs",
                                  "u64equal"])
def test_topology(gpu, oracle, kind):
    """Random codes, short duplicate runs, runs of 1 to 3000 equal codes (the
    searches that the wave finishes: ranges and runs past 64 codes), and all
    codes equal."""
    rng = np.random.default_rng(21)
    n = 50_000
    dt = np.uint32 if kind.startswith("u32") else np.uint64
    if kind.endswith("runs"):
        lens = rng.integers(1, 3000, 100)
        vals = np.sort(rng.choice(1 << 30, lens.size, replace=False)).astype(dt)
        if dt == np.uint64:
            vals = vals << np.uint64(30)
        codes = np.repeat(vals, lens)[:n]
        n = codes.size
    elif kind == "u64equal":
        codes = np.full(n, 12345, dtype=np.uint64)
    elif dt == np.uint32:
        hi = 5000 if kind.endswith("dups") else 1 << 30
        codes = np.sort(rng.integers(0, hi, n)).astype(np.uint32)
    else:
        hi = 5000 if kind.endswith("dups") else 1 << 63
        codes = np.sort(rng.integers(0, hi, n, dtype=np.uint64))
    idx2obj = rng.permutation(n).astype(np.uint32)
    got = H(Mortons.make_bvhnodes_from_sorted_mortons(G(idx2obj, gpu), G(codes, gpu)))
    want = oracle.bvhnodes_from_sorted(idx2obj, codes)
    assert np.array_equal(got, want)


def test_topology_single_and_pair(gpu, oracle):
    for n in (1, 2):
        codes = np.arange(n, dtype=np.uint32)
        idx = np.arange(n, dtype=np.uint32)[::-1].copy()
        got = H(Mortons.make_bvhnodes_from_sorted_mortons(G(idx, gpu), G(codes, gpu)))
        assert np.array_equal(got, oracle.bvhnodes_from_sorted(idx, codes))


@pytest.mark.parametrize("n", [3, 64, 65, 127, 128, 129, 130, 257, 4097])
def test_topology_sizes_near_the_serial_cut(gpu, oracle, n):
    """Sizes around the point (128 codes) past which the root's split and
    long ranges go to the wave-wide search, with and without equal codes."""
    rng = np.random.default_rng(n)
    for hi in (1 << 30, 7, 1):
        codes = np.sort(rng.integers(0, hi, n)).astype(np.uint32)
        idx = rng.permutation(n).astype(np.uint32)
        got = H(Mortons.make_bvhnodes_from_sorted_mortons(G(idx, gpu), G(codes, gpu)))
        assert np.array_equal(got, oracle.bvhnodes_from_sorted(idx, codes)), hi


def _clustered():
    """A 32k-triangle sphere shrunk to 1e-6 beside one far triangle: every
    clustered centroid shares its top 24 code bits, so the sort's finish
    leaves one 32k-element group to k_sort_groups."""
    tri, vtx = primitives.octahedral_sphere(6, noise_sigma=0.01, seed=2)
    vtx = (vtx.astype(np.float64) * 1e-6 + 0.3).astype(np.float32)
    far = np.array([[-1.0, -1.0, -1.0], [-1.0, -0.9, -1.0], [-0.9, -1.0, -1.0]], np.float32)
    nv = vtx.shape[0]
    tri = np.concatenate([tri, np.array([[nv, nv + 1, nv + 2]], tri.dtype)])
    return tri, np.concatenate([vtx, far])


MESHES = {
    "sphere_c1": lambda: primitives.sphere_yup(1.0, 64, 128),
    "clustered": _clustered,
    "torus": lambda: primitives.torus_zup(1.3, 0.4, 64, 32),
    "octa6_noise": lambda: primitives.octahedral_sphere(6, noise_sigma=0.01, seed=1),
    "cylinder": lambda: primitives.cylinder_closed_end_yup(0.5, 2.0, 40, 30),
}


@pytest.mark.parametrize("mesh", list(MESHES))
@pytest.mark.parametrize("bits", [30, 63])
def test_build_trimesh3(gpu, oracle, mesh, bits):
    tri, vtx = MESHES[mesh]()
    bn, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb(G(tri, gpu), G(vtx, gpu), morton_bits=bits)
    wbn, waabb = oracle.build_trimesh3(tri, vtx, morton_bits=bits)
    assert np.array_equal(H(bn), wbn)
    assert bits_equal(H(aabb), waabb)
    no_device_errors()


def test_refit_standalone_matches_oracle(gpu, oracle):
    tri, vtx = MESHES["torus"]()
    wbn, waabb = oracle.build_trimesh3(tri, vtx)
    aabb = TriMesh3.make_bvhnode2aabb_from_bvhnodes(G(tri, gpu), G(vtx, gpu), G(wbn, gpu))
    assert bits_equal(H(aabb), waabb)


def relabel_internal(bn, seed):
    """The same tree with its internal ids (root 0 kept) randomly permuted:
    leaf rows stay at n-1.., but internal ids no longer lie in their own leaf
    ranges, so the standalone refit must take its generic fallback."""
    nb = bn.shape[0] // 2
    perm = np.concatenate([[0], 1 + np.random.default_rng(seed).permutation(nb - 1)])
    idmap = np.arange(bn.shape[0], dtype=np.int64)
    idmap[:nb] = perm

    def f(x):
        x = x.astype(np.int64)
        return np.where(x < idmap.size, idmap[np.minimum(x, idmap.size - 1)], x).astype(np.uint32)

    out = bn.copy()
    out[perm] = np.stack([f(bn[:nb, c]) for c in range(3)], 1)
    out[nb:, 0] = f(bn[nb:, 0])
    return out


def swap_leaves(bn, k1, k2):
    """The tree with leaves k1 and k2 exchanged (rows and parent links):
    subtree leaf sets are no longer contiguous ranges."""
    bn = bn.copy()
    nb = bn.shape[0] // 2
    r1, r2 = nb + k1, nb + k2
    p1, p2 = int(bn[r1, 0]), int(bn[r2, 0])
    assert p1 != p2
    bn[p1, 1:] = np.where(bn[p1, 1:] == r1, r2, bn[p1, 1:])
    bn[p2, 1:] = np.where(bn[p2, 1:] == r2, r1, bn[p2, 1:])
    bn[[r1, r2]] = bn[[r2, r1]]
    return bn


def refit_tri_abi(gpu, tri, vtx, bn, eps):
    L = _lib.lib()
    t_tri, t_vtx, t_bn = G(tri, gpu), G(vtx, gpu), G(bn, gpu)
    aabb = torch.empty((bn.shape[0], 6), dtype=torch.float32, device=gpu)
    ws = _lib.workspace(L.dmsh_refit_workspace_bytes(tri.shape[0]), gpu)
    _lib.check(L.dmsh_bvhnode2aabb_trimesh3(
        _lib.stream_of(gpu), _lib.ptr(aabb), bn.shape[0], _lib.ptr(t_bn), tri.shape[0],
        _lib.ptr(t_tri), vtx.shape[0], _lib.ptr(t_vtx), eps, _lib.ptr(ws), ws.numel()),
        "dmsh_bvhnode2aabb_trimesh3")
    return H(aabb)


def refit_points_abi(gpu, pts, bn):
    L = _lib.lib()
    t_pts, t_bn = G(pts, gpu), G(bn, gpu)
    aabb = torch.empty((bn.shape[0], 6), dtype=torch.float32, device=gpu)
    ws = _lib.workspace(L.dmsh_refit_workspace_bytes(pts.shape[0]), gpu)
    _lib.check(L.dmsh_bvhnode2aabb_points(
        _lib.stream_of(gpu), _lib.ptr(aabb), bn.shape[0], _lib.ptr(t_bn), pts.shape[0],
        _lib.ptr(t_pts), _lib.ptr(ws), ws.numel()), "dmsh_bvhnode2aabb_points")
    return H(aabb)


@pytest.mark.parametrize("eps", [0.0, 0.01])
def test_refit_standalone_large(gpu, oracle, eps):
    """131,072 triangles (128 treelets and the top) through the treelet path
    of the standalone refit, margin eps included."""
    tri, vtx = primitives.octahedral_sphere(7, noise_sigma=0.01, seed=5)
    wbn, _ = oracle.build_trimesh3(tri, vtx, morton_bits=63)
    got = refit_tri_abi(gpu, tri, vtx, wbn, eps)
    assert bits_equal(got, oracle.refit_trimesh3(wbn, tri, vtx, eps))
    no_device_errors()


@pytest.mark.parametrize("layout", ["relabelled", "swapped"])
@pytest.mark.parametrize("mesh", ["torus", "octa6_noise"])
def test_refit_standalone_non_lbvh(gpu, oracle, layout, mesh):
    """Trees whose rows fail the LBVH-layout check take the generic climb
    and still match the oracle's recursion bit for bit, with no device error."""
    tri, vtx = MESHES[mesh]()
    wbn, _ = oracle.build_trimesh3(tri, vtx)
    n = tri.shape[0]
    bn = relabel_internal(wbn, 3) if layout == "relabelled" else swap_leaves(wbn, 3, n - 5)
    got = refit_tri_abi(gpu, tri, vtx, bn, 0.0)
    assert bits_equal(got, oracle.refit_trimesh3(bn, tri, vtx, 0.0))
    no_device_errors()


def _plan_gate_word(ws, n):
    """counters[4] of the standalone refit workspace (1: the last call built
    the plan, 0: it kept the one already there); the layout is four arrays of
    n - 1 words, each 256-byte aligned, then the counters."""
    off = 0
    for _ in range(4):
        off = (off + 255) // 256 * 256 + 4 * max(n - 1, 1)
    off = (off + 255) // 256 * 256
    return int(ws[off + 16: off + 20].cpu().numpy().view(np.uint32)[0])


def test_refit_standalone_kept_plan(gpu, oracle):
    """One workspace through a sequence of refits: the plan is kept while the
    rows stay the same (moved vertices only), rebuilt when they change (a
    non-LBVH tree of the same size, another tree of the same size) and when
    keeping is switched off; every result equals the oracle bit for bit."""
    L = _lib.lib()
    tri, vtx = primitives.octahedral_sphere(7, noise_sigma=0.01, seed=5)
    n = tri.shape[0]
    bn_a, _ = oracle.build_trimesh3(tri, vtx, morton_bits=63)
    tri_b, vtx_b = primitives.octahedral_sphere(7, noise_sigma=0.02, seed=6)
    bn_b, _ = oracle.build_trimesh3(tri_b, vtx_b, morton_bits=63)
    assert not np.array_equal(bn_a, bn_b)
    vtx_moved = (vtx * np.float32(1.01) + np.float32(0.003)).astype(np.float32)
    ws = _lib.workspace(L.dmsh_refit_workspace_bytes(n), gpu)
    ws.zero_()  # not a block that held an earlier test's plan
    aabb = torch.empty((2 * n - 1, 6), dtype=torch.float32, device=gpu)

    def refit(t, v, bn):
        t_tri, t_vtx, t_bn = G(t, gpu), G(v, gpu), G(bn, gpu)
        _lib.check(L.dmsh_bvhnode2aabb_trimesh3(
            _lib.stream_of(gpu), _lib.ptr(aabb), bn.shape[0], _lib.ptr(t_bn), n, _lib.ptr(t_tri),
            v.shape[0], _lib.ptr(t_vtx), 0.0, _lib.ptr(ws), ws.numel()), "refit")
        assert bits_equal(H(aabb), oracle.refit_trimesh3(bn, t, v, 0.0))
        return _plan_gate_word(ws, n)

    assert refit(tri, vtx, bn_a) == 1          # a new workspace builds the plan
    assert refit(tri, vtx_moved, bn_a) == 0    # same rows: kept
    assert refit(tri, vtx, bn_a) == 0
    swapped = swap_leaves(bn_a, 3, n - 5)
    assert refit(tri, vtx, swapped) == 1       # rows changed (and fail the layout check)
    assert refit(tri, vtx, swapped) == 1       # a failed check is never kept
    assert refit(tri, vtx_moved, bn_a) == 1    # back to the first tree: rebuilt
    assert refit(tri_b, vtx_b, bn_b) == 1      # another tree of the same size
    assert refit(tri_b, vtx_b, bn_b) == 0
    assert L.dmsh_set_tuning(b"refit.plan_cache", 0) == 0
    try:
        assert refit(tri_b, vtx_b, bn_b) == 1  # keeping switched off
    finally:
        assert L.dmsh_set_tuning(b"refit.plan_cache", 1) == 0
    no_device_errors()


@pytest.mark.parametrize("layout", ["lbvh", "relabelled"])
def test_refit_standalone_points(gpu, oracle, layout):
    pts = np.random.default_rng(33).random((70_000, 3), dtype=np.float32)
    wbn, _ = oracle.build_points(pts, morton_bits=30)
    bn = wbn if layout == "lbvh" else relabel_internal(wbn, 4)
    assert bits_equal(refit_points_abi(gpu, pts, bn), oracle.refit_points(bn, pts))
    no_device_errors()


def test_refit_standalone_tiny(gpu, oracle):
    """One, two and three leaves (no treelet above the root)."""
    for n in (1, 2, 3):
        pts = np.random.default_rng(n).random((n, 3), dtype=np.float32)
        wbn, _ = oracle.build_points(pts, morton_bits=30)
        assert bits_equal(refit_points_abi(gpu, pts, wbn), oracle.refit_points(wbn, pts))
    no_device_errors()


@pytest.mark.parametrize("bits", [30, 63])
def test_build_points(gpu, oracle, bits):
    rng = np.random.default_rng(31)
    pts = rng.random((200_000, 3), dtype=np.float32)
    bn, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb_for_points(G(pts, gpu), morton_bits=bits)
    wbn, waabb = oracle.build_points(pts, morton_bits=bits)
    assert np.array_equal(H(bn), wbn)
    assert bits_equal(H(aabb), waabb)


# ---------------------------------------------------------------------------
# ray-cast and depth
# ---------------------------------------------------------------------------
def _ortho_c1():
    return np.diag([1.2, 1.2, 2.0, 1.0]).astype(np.float32)


def _perspective(eye=(0.0, 0.0, 3.0), fovy=45.0, near=1.0, far=5.0):
    import del_msh_dlpack.Mat44.torch as Mat44

    w2c = Mat44.look_at(eye, (0, 0, 0), (0, 1, 0), dtype=torch.float64).numpy()
    proj = Mat44.perspective(fovy, 1.0, near, far, dtype=torch.float64).numpy()
    w2n = proj @ w2c
    return np.linalg.inv(w2n).astype(np.float32), w2n.astype(np.float32)


@pytest.mark.parametrize("mesh", ["sphere_c1", "torus", "octa6_noise"])
@pytest.mark.parametrize("camera", ["ortho", "persp"])
def test_pix2tri(gpu, oracle, mesh, camera):
    tri, vtx = MESHES[mesh]()
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    n2w = _ortho_c1() if camera == "ortho" else _perspective()[0]
    W, H_ = 256, 192
    pix, t = Pix2Tri.by_raycasting_with_coeff(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu),
                                              G(n2w, gpu), (W, H_))
    wpix, wt, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), W, H_, num_threads=8)
    assert np.array_equal(H(pix), wpix)
    assert bits_equal(H(t), wt)
    pix2 = Pix2Tri.by_raycasting(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu), G(n2w, gpu), (W, H_))
    assert np.array_equal(H(pix2), wpix)
    no_device_errors()


def test_pix2tri_flags_rows_out_of_layout(gpu, oracle):
    """The cast packs each row by itself: a leaf row with a second child, or
    a child id past the last node, becomes a never-hit triangle / node and
    raises DMSH_DEVERR_INDEX; every other ray ends as before."""
    tri, vtx = MESHES["torus"]()
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    n = tri.shape[0]
    n2w = _ortho_c1()
    W, H_ = 128, 96
    wpix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), W, H_, num_threads=8)
    no_device_errors()
    for row, col, value in ((n - 1 + 5, 2, 0), (3, 1, 2 * n + 7)):
        bad = bn.reshape(-1, 3).copy()
        bad[row, col] = value
        pix = Pix2Tri.by_raycasting(G(tri, gpu), G(vtx, gpu), G(bad.reshape(bn.shape), gpu),
                                    G(aabb, gpu), G(n2w, gpu), (W, H_))
        assert _lib.device_errors() & 0x1
        got = H(pix)
        assert got.shape == wpix.shape
        assert np.all((got == wpix) | (got == 0xFFFFFFFF) | (got < n))


def test_c1_config_at_size(gpu, oracle):
    """The plain-path workload at its stated size: UV sphere (r = 1, 64 x 128
    segments, 16,128 triangles), 30-bit codes, LBVH build + refit on the
    device, orthographic ray-cast of exactly 256 x 256 rays from z = +2 over
    xy in [-1.2, 1.2]^2 looking down -z; hit triangle, hit coefficient and
    depth bit-exact with the oracle, and the coverage of a unit disc in a
    2.4-wide square: pi / 2.4^2 (to the tessellation and pixel error)."""
    tri, vtx = primitives.sphere_yup(1.0, 64, 128)
    assert tri.shape == (16128, 3) and vtx.shape == (8066, 3)
    t_tri, t_vtx = G(tri, gpu), G(vtx, gpu)
    bn, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb(t_tri, t_vtx)
    wbn, waabb = oracle.build_trimesh3(tri, vtx)
    assert np.array_equal(H(bn), wbn)
    assert bits_equal(H(aabb), waabb)
    n2w = _ortho_c1()
    W = H_ = 256
    pix, t = Pix2Tri.by_raycasting_with_coeff(t_tri, t_vtx, bn, aabb, G(n2w, gpu), (W, H_))
    T = oracle.colmajor(n2w)
    wpix, wt, _ = oracle.pix2tri(tri, vtx, wbn, waabb, T, W, H_, num_threads=8)
    assert np.array_equal(H(pix), wpix)
    assert bits_equal(H(t), wt)
    d = torch.zeros((H_, W), dtype=torch.float32, device=gpu)
    Pix2Depth.update(d, pix, t_tri, t_vtx, G(n2w, gpu))
    assert bits_equal(H(d), oracle.pix2depth(wpix, tri, vtx, T, W, H_))
    frac = np.count_nonzero(wpix != U32MAX) / (W * H_)
    assert abs(frac - math.pi / 2.4 ** 2) < 0.01
    no_device_errors()


def _pixel_grid(n):
    """z = 0 grid with its vertices under the pixel centres of an n x n
    orthographic image: rays through vertices and edges tie between the
    triangles around them."""
    x = ((2.0 * (np.arange(n) + 0.5) / n) - 1.0).astype(np.float32)
    xx, yy = np.meshgrid(x, x)
    vtx = np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n, np.float32)], 1).astype(np.float32)
    tri = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i, (j + 1) * n + i + 1
            tri += [[a, b, d], [a, d, c]]
    return np.array(tri, np.uint32), vtx


@pytest.mark.parametrize("scene", ["sphere_twice", "grid_twice", "coplanar_soup"])
def test_pix2tri_exact_ties(gpu, oracle, scene):
    """Scenes where hits tie exactly (every triangle duplicated, rays through
    grid vertices, overlapping coplanar triangles): the winner of each tie is
    fixed by the reference's recursion, and the flat search must pick it."""
    img = 64
    if scene == "sphere_twice":
        tri, vtx = MESHES["sphere_c1"]()
        tri = np.concatenate([tri, tri])
        n2w = _perspective()[0]
    elif scene == "grid_twice":
        tri, vtx = _pixel_grid(img)
        tri = np.concatenate([tri[::-1], tri])
        n2w = np.eye(4, dtype=np.float32)
    else:
        rng = np.random.default_rng(5)
        soup = rng.uniform(-1, 1, (800, 3, 3)).astype(np.float32)
        soup[:, :, 2] = np.round(soup[:, :, 2] * 2) / 4
        vtx = soup.reshape(-1, 3)
        tri = np.arange(vtx.shape[0], dtype=np.uint32).reshape(-1, 3)
        n2w = np.eye(4, dtype=np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, t = Pix2Tri.by_raycasting_with_coeff(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu),
                                              G(n2w, gpu), (img, img))
    wpix, wt, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), img, img, num_threads=8)
    assert (wpix != U32MAX).sum() > img
    assert np.array_equal(H(pix), wpix)
    assert bits_equal(H(t), wt)
    no_device_errors()


@pytest.mark.parametrize("camera", ["ortho", "persp"])
def test_pix2tri_deep_stack_path(gpu, oracle, camera):
    """A run of equal Morton codes whose triangles are stacked along the view
    ray, listed far to near, makes the near child the rest of the run at every
    level: the traversal needs ~K frames, more than the LDS stack holds, so the
    rays go through the global-stack kernel.  The orthographic camera has
    parallel axes (general slab test), the perspective one none."""
    K = 80
    quad = []
    vtx = []
    for k in range(K):  # far (k=0) to near; centroids differ by < 1 Morton cell
        z = -0.001 + k * 2.0e-5
        base = len(vtx)
        vtx += [[-0.5, -0.5, z], [0.6, -0.45, z], [-0.45, 0.6, z]]
        quad.append([base, base + 1, base + 2])
    # two anchors far away set the unit-cube fit
    for s in (-1.0, 1.0):
        base = len(vtx)
        vtx += [[s * 4, s * 4, 0], [s * 4 + 0.01, s * 4, 0], [s * 4, s * 4 + 0.01, 0]]
        quad.append([base, base + 1, base + 2])
    tri = np.array(quad, np.uint32)
    vtx = np.array(vtx, np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    if camera == "ortho":
        n2w = np.diag([1.0, 1.0, 1.0, 1.0]).astype(np.float32)
    else:
        n2w = _perspective(eye=(0.0, 0.0, 2.0), fovy=20.0)[0]
    pix, t = Pix2Tri.by_raycasting_with_coeff(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu),
                                              G(n2w, gpu), (32, 32))
    wpix, wt, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), 32, 32)
    assert np.array_equal(H(pix), wpix)
    assert bits_equal(H(t), wt)
    assert (wpix == K - 1).sum() > 0  # the nearest sheet is what the rays see
    no_device_errors()


def test_pix2tri_nonfinite_vertices(gpu, oracle):
    """NaN / infinite vertices give non-finite boxes, which switch the
    traversal to the general slab test; results still match bit for bit."""
    tri, vtx = primitives.torus_zup(0.6, 0.2, 32, 16)
    vtx = vtx.copy()
    vtx[5] = np.nan
    vtx[17] = [np.inf, 0.1, 0.0]
    vtx[40, 2] = -np.inf
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    assert not np.isfinite(aabb).all()
    n2w = _perspective(eye=(0.2, 0.3, 2.5))[0]
    W, H_ = 96, 64
    pix, t = Pix2Tri.by_raycasting_with_coeff(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu),
                                              G(n2w, gpu), (W, H_))
    wpix, wt, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), W, H_, num_threads=8)
    assert np.array_equal(H(pix), wpix)
    assert bits_equal(H(t), wt)
    no_device_errors()


def test_pix2tri_empty_and_miss(gpu, oracle):
    tri, vtx = MESHES["torus"]()
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    n2w = np.diag([1.0, 1.0, 1.0, 1.0]).astype(np.float32)
    n2w[0, 3] = 50.0  # camera looks far away from the mesh
    pix = Pix2Tri.by_raycasting(G(tri, gpu), G(vtx, gpu), G(bn, gpu), G(aabb, gpu), G(n2w, gpu), (64, 64))
    assert np.all(H(pix) == U32MAX)


@pytest.mark.parametrize("camera", ["ortho", "persp"])
def test_pix2depth(gpu, oracle, camera):
    tri, vtx = MESHES["torus"]()
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    n2w = _ortho_c1() if camera == "ortho" else _perspective()[0]
    W, H_ = 128, 128
    wpix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), W, H_, num_threads=8)
    depth = torch.zeros((H_, W), dtype=torch.float32, device=gpu)
    Pix2Depth.update(depth, G(wpix, gpu), G(tri, gpu), G(vtx, gpu), G(n2w, gpu))
    want = oracle.pix2depth(wpix, tri, vtx, n2w.T.reshape(-1), W, H_)
    assert bits_equal(H(depth), want)
    no_device_errors()


# ---------------------------------------------------------------------------
# edge-gradient backward
# ---------------------------------------------------------------------------
def _edge_gradient_scene(oracle, camera, W, H_, nv, seed):
    tri, vtx = primitives.sphere_yup(0.8, 48, 96)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    if camera == "ortho":
        w2n = np.diag([1.0, 1.0, 0.5, 1.0]).astype(np.float32)
        n2w = np.linalg.inv(w2n).astype(np.float32)
    else:
        n2w, w2n = _perspective(eye=(0.3, 0.2, 3.0))
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * H_, 0, 0.5 * H_], [0, 0, 0.5, 0.5], [0, 0, 0, 1]],
                   np.float32)
    w2p = (n2p @ w2n).astype(np.float32)
    wpix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, n2w.T.reshape(-1), W, H_, num_threads=8)
    rng = np.random.default_rng(seed)
    occ = (wpix != U32MAX).astype(np.float32)[..., None]
    pix2val = np.concatenate([occ, rng.random((H_, W, nv - 1), dtype=np.float32)], axis=2)
    dldw = rng.normal(size=(H_, W, nv)).astype(np.float32)
    return tri, vtx, w2p, wpix, pix2val, dldw


@pytest.mark.parametrize("camera", ["ortho", "persp"])
@pytest.mark.parametrize("nv", [1, 3])
@pytest.mark.parametrize("ordered", [True, False])
def test_edge_gradient_bwd(gpu, oracle, camera, nv, ordered):
    """Ordered: bit-exact with the oracle's serial loop.  Atomic: float
    atomics reorder the per-vertex sums, so to summation rounding."""
    W, H_ = 160, 128
    tri, vtx, w2p, wpix, pix2val, dldw = _edge_gradient_scene(oracle, camera, W, H_, nv, 40 + nv)
    got = H(RasterizedEdgeGradient.bwd(G(tri, gpu), G(vtx, gpu), G(w2p, gpu), G(wpix, gpu),
                                       G(pix2val, gpu), G(dldw, gpu), ordered=ordered))
    want = oracle.edge_gradient_bwd(tri, vtx, w2p.T.reshape(-1), wpix, pix2val, dldw)
    scale = np.abs(want).max()
    assert scale > 0
    if ordered:
        assert bits_equal(got, want)
    else:
        assert np.abs(got - want).max() <= 1e-5 * scale + 5e-6
    no_device_errors()


@pytest.mark.parametrize("shape", [(1, 90), (90, 1), (2, 2), (97, 61)])
def test_edge_gradient_bwd_ordered_shapes_and_accumulation(gpu, oracle, shape):
    """Ragged images (one row, one column, 2 x 2, odd sizes) and a nonzero
    output to accumulate into: the ordered backward adds to what dldw_vtx2xyz
    holds, term after term, as the serial loop does."""
    from del_msh_dlpack.RasterizedEdgeGradient import bwd as native

    W, H_ = shape
    tri, vtx, w2p, wpix, pix2val, dldw = _edge_gradient_scene(oracle, "persp", W, H_, 2, 7)
    init = np.random.default_rng(8).normal(size=vtx.shape).astype(np.float32)
    out = G(init, gpu)
    native(G(tri, gpu), G(vtx, gpu), out, G(oracle.colmajor(w2p), gpu), G(wpix, gpu),
           G(pix2val, gpu), G(dldw, gpu))
    want = oracle.edge_gradient_bwd(tri, vtx, w2p.T.reshape(-1), wpix, pix2val, dldw, dldw_init=init)
    assert bits_equal(H(out), want)
    no_device_errors()


def test_edge_gradient_bwd_ordered_repeats_bit_for_bit(gpu, oracle):
    """Three runs of the ordered backward give the same bits."""
    W, H_ = 200, 150
    tri, vtx, w2p, wpix, pix2val, dldw = _edge_gradient_scene(oracle, "persp", W, H_, 1, 3)
    runs = [H(RasterizedEdgeGradient.bwd(G(tri, gpu), G(vtx, gpu), G(w2p, gpu), G(wpix, gpu),
                                         G(pix2val, gpu), G(dldw, gpu))) for _ in range(3)]
    assert bits_equal(runs[0], runs[1]) and bits_equal(runs[0], runs[2])
    no_device_errors()


# ---------------------------------------------------------------------------
# full-size properties (2^21 triangles, 2048^2 rays, 2^22 points)
# ---------------------------------------------------------------------------
def test_full_size_build_and_cast(gpu, oracle):
    tri, vtx = primitives.octahedral_sphere(9, noise_sigma=0.01, seed=1)
    assert tri.shape[0] == 1 << 21
    t_tri, t_vtx = G(tri, gpu), G(vtx, gpu)
    bn, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb(t_tri, t_vtx, morton_bits=63)
    wbn, waabb = oracle.build_trimesh3(tri, vtx, morton_bits=63)
    assert np.array_equal(H(bn), wbn)
    assert bits_equal(H(aabb), waabb)
    assert oracle.check_bvh_topology(wbn, tri.shape[0]) == 0
    n2w, _ = _perspective()
    W = H_ = 2048
    pix, t = Pix2Tri.by_raycasting_with_coeff(t_tri, t_vtx, bn, aabb, G(n2w, gpu), (W, H_))
    pix, t = H(pix), H(t)
    no_device_errors()
    # every one of the 4,194,304 rays against the oracle's ordered recursion
    T = n2w.T.reshape(-1)
    wpix, wt, _ = oracle.pix2tri(tri, vtx, wbn, waabb, T, W, H_, num_threads=16)
    assert np.array_equal(pix, wpix)
    assert bits_equal(t, wt)
    # depth of every hit, bit-exact with the oracle
    d = torch.zeros((H_, W), dtype=torch.float32, device=gpu)
    Pix2Depth.update(d, G(pix, gpu), t_tri, t_vtx, G(n2w, gpu))
    wd = oracle.pix2depth(wpix, tri, vtx, T, W, H_)
    assert bits_equal(H(d), wd)
    # coverage of a sphere seen from distance 3 with fovy 45: the silhouette of
    # radius r covers (tan(asin(r/3)) / tan(22.5 deg))^2 pi / 4 of the image;
    # the radial noise (sigma 0.01) keeps the silhouette radius within 1 +- 0.05
    frac = np.count_nonzero(pix != U32MAX) / (W * H_)

    def cover(r):
        return (math.tan(math.asin(r / 3.0)) / math.tan(math.radians(22.5))) ** 2 * math.pi / 4

    assert cover(0.95) < frac < cover(1.05)
    assert np.all((t[pix != U32MAX] > 0.0) & (t[pix != U32MAX] < 1.0))


def test_full_size_points_build(gpu, oracle):
    rng = np.random.default_rng(2)
    pts = rng.random((1 << 22, 3), dtype=np.float32)
    bn, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb_for_points(G(pts, gpu), morton_bits=63)
    wbn, waabb = oracle.build_points(pts, morton_bits=63)
    assert np.array_equal(H(bn), wbn)
    assert bits_equal(H(aabb), waabb)
