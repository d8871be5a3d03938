"""GPU parity for the batched build and ray-cast: every mesh of a batch gets
exactly the tree and image the single-mesh path (and the oracle) gives it."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Batched.torch as Batched

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


def _edge_batch():
    """Generated meshes plus the edge cases: an empty mesh, one and two
    triangles, and one just over the 1024-leaf treelet size."""
    b = primitives.batch_meshes(14, seed=11, tri_range=(6, 6000), pix_range=(16 * 16, 96 * 96))
    tri_list, vtx_list, w_list, cam_list = [], [], [], []
    for m in range(14):
        t0, t1 = b["mesh2tri"][m], b["mesh2tri"][m + 1]
        v0, v1 = b["mesh2vtx"][m], b["mesh2vtx"][m + 1]
        tri_list.append(b["tri2vtx"][t0:t1])
        vtx_list.append(b["vtx2xyz"][v0:v1])
        w_list.append(int(b["mesh2width"][m]))
        cam_list.append(b["ndc2world"][m])
    one = (np.array([[0, 1, 2]], np.uint32), np.array([[-.5, -.5, 0], [.6, -.4, .1], [0, .7, -.1]], np.float32))
    two = (np.array([[0, 1, 2], [0, 2, 3]], np.uint32),
           np.array([[-.5, -.5, 0], [.5, -.5, .1], [.5, .5, 0], [-.5, .5, -.1]], np.float32))
    tri_k, vtx_k = primitives.torus_zup(0.7, 0.3, 25, 21)  # 1050 triangles
    empty = (np.zeros((0, 3), np.uint32), np.zeros((0, 3), np.float32))
    for k, (t, v) in enumerate([empty, one, two, (tri_k, vtx_k), empty]):
        pos = [3, 0, 7, 9, 16][k]
        tri_list.insert(pos, t)
        vtx_list.insert(pos, v)
        w_list.insert(pos, 24 + 8 * k)
        cam_list.insert(pos, np.diag([1.2, 1.2, 2.0, 1.0]).astype(np.float32))
    csr = lambda c: np.concatenate([[0], np.cumsum(c)]).astype(np.uint32)
    return {
        "tris": tri_list, "vtxs": vtx_list,
        "mesh2tri": csr([len(t) for t in tri_list]),
        "mesh2vtx": csr([len(v) for v in vtx_list]),
        "mesh2pix": csr([w * w for w in w_list]),
        "tri2vtx": np.concatenate(tri_list), "vtx2xyz": np.concatenate(vtx_list),
        "mesh2width": np.array(w_list, np.uint32), "ndc2world": np.stack(cam_list),
    }


def _pipeline(b, gpu, with_coeff=True):
    return Batched.Pipeline(b["mesh2tri"], b["mesh2vtx"], G(b["tri2vtx"], gpu),
                            G(b["vtx2xyz"], gpu), b["mesh2pix"], b["mesh2width"],
                            G(b["ndc2world"], gpu), with_coeff)


def _check_against_oracle(oracle, p, b, meshes):
    for m in meshes:
        tri, vtx = b["tris"][m], b["vtxs"][m]
        bn, ab = p.mesh_tree(m)
        pix, t = p.mesh_image(m)
        w = int(b["mesh2width"][m])
        if len(tri) == 0:
            assert bn.shape[0] == 0
            assert np.all(H(pix) == U32MAX)
            continue
        wbn, wab = oracle.build_trimesh3(tri, vtx)
        assert np.array_equal(H(bn), wbn), f"mesh {m} topology"
        assert bits_equal(H(ab), wab), f"mesh {m} boxes"
        wpix, wt, _ = oracle.pix2tri(tri, vtx, wbn, wab, oracle.colmajor(b["ndc2world"][m]), w, w,
                                     num_threads=8)
        assert np.array_equal(H(pix), wpix), f"mesh {m} pix2tri"
        assert bits_equal(H(t), wt), f"mesh {m} t"


def test_batch_build_and_cast_match_single_meshes(gpu, oracle):
    b = _edge_batch()
    p = _pipeline(b, gpu)
    p.run()
    torch.cuda.synchronize()
    assert np.array_equal(H(p.mesh2node), p.mesh2node_h)
    _check_against_oracle(oracle, p, b, range(len(b["tris"])))
    no_device_errors()


def _duplicate_batch():
    """Meshes whose codes repeat: 3000 copies of one triangle beside 600
    spread ones (one group of equal top code bits longer than the finish's
    1024-element margin), a mesh of copies only, and a plain torus."""
    rng = np.random.default_rng(21)
    base = np.array([[-.3, -.2, .1], [.4, -.1, 0], [0, .5, -.2]], np.float32)
    spread_v = rng.uniform(-1, 1, (1800, 3)).astype(np.float32)
    tri_a = np.concatenate([np.tile(np.array([[0, 1, 2]], np.uint32), (3000, 1)),
                            3 + np.arange(1800, dtype=np.uint32).reshape(600, 3)])
    tri_a = tri_a[rng.permutation(len(tri_a))]
    vtx_a = np.concatenate([base, spread_v])
    tri_b = np.tile(np.array([[0, 1, 2]], np.uint32), (1500, 1))
    tri_c, vtx_c = primitives.torus_zup(0.7, 0.3, 30, 20)
    tris, vtxs = [tri_a, tri_b, tri_c], [vtx_a, base.copy(), vtx_c]
    csr = lambda c: np.concatenate([[0], np.cumsum(c)]).astype(np.uint32)
    cam = np.diag([1.2, 1.2, 2.0, 1.0]).astype(np.float32)
    return {
        "tris": tris, "vtxs": vtxs,
        "mesh2tri": csr([len(t) for t in tris]), "mesh2vtx": csr([len(v) for v in vtxs]),
        "mesh2pix": csr([48 * 48] * 3),
        "tri2vtx": np.concatenate(tris), "vtx2xyz": np.concatenate(vtxs),
        "mesh2width": np.array([48] * 3, np.uint32), "ndc2world": np.stack([cam] * 3),
    }


@pytest.mark.parametrize("seg", [1, 0])
def test_batch_duplicate_codes(gpu, oracle, seg):
    """Long runs of equal codes: in the segmented sort they are sorted by
    k_sort_groups on the low code bits; both sorts give the oracle's trees."""
    L = _lib.lib()
    old = L.dmsh_batch_get_tuning(b"batch.sort_seg")
    try:
        assert L.dmsh_batch_set_tuning(b"batch.sort_seg", seg) == 0
        b = _duplicate_batch()
        p = _pipeline(b, gpu)
        p.run()
        torch.cuda.synchronize()
        _check_against_oracle(oracle, p, b, range(3))
    finally:
        L.dmsh_batch_set_tuning(b"batch.sort_seg", old)
    no_device_errors()


def _tiny_batch(num_mesh: int, seed: int):
    """Triangle soups of 0..40 triangles each: many meshes share a wave of
    the centroid kernel, a round of the key kernel and a 64-mesh chunk of the
    sort offsets."""
    rng = np.random.default_rng(seed)
    tris, vtxs = [], []
    for _ in range(num_mesh):
        k = int(rng.integers(0, 41))
        nv = max(3, k)
        vtxs.append(rng.uniform(-1, 1, (nv, 3)).astype(np.float32))
        tris.append(rng.integers(0, nv, (k, 3)).astype(np.uint32))
    csr = lambda c: np.concatenate([[0], np.cumsum(c)]).astype(np.uint32)
    cam = np.diag([1.2, 1.2, 2.0, 1.0]).astype(np.float32)
    return {
        "tris": tris, "vtxs": vtxs,
        "mesh2tri": csr([len(t) for t in tris]), "mesh2vtx": csr([len(v) for v in vtxs]),
        "mesh2pix": csr([16 * 16] * num_mesh),
        "tri2vtx": np.concatenate(tris), "vtx2xyz": np.concatenate(vtxs),
        "mesh2width": np.array([16] * num_mesh, np.uint32),
        "ndc2world": np.stack([cam] * num_mesh),
    }


@pytest.mark.parametrize("seg", [1, 0])
def test_batch_many_tiny_meshes(gpu, seg):
    """3000 meshes of 0..40 triangles: the batched build and cast equal the
    single-mesh entry points mesh by mesh, with either sort."""
    L = _lib.lib()
    old = L.dmsh_batch_get_tuning(b"batch.sort_seg")
    try:
        assert L.dmsh_batch_set_tuning(b"batch.sort_seg", seg) == 0
        b = _tiny_batch(3000, seed=8)
        p = _pipeline(b, gpu)
        p.run()
        loop = Batched.PerMeshLoop(p)
        loop.build()
        loop.cast()
        torch.cuda.synchronize()
        assert loop.equals_pipeline()
    finally:
        L.dmsh_batch_set_tuning(b"batch.sort_seg", old)
    no_device_errors()


def test_batch_results_do_not_depend_on_tuning(gpu):
    b = _edge_batch()
    p = _pipeline(b, gpu)
    p.run()
    ref = (H(p.bvhnodes).copy(), H(p.bvhnode2aabb).copy(), H(p.pix2tri).copy(), H(p.pix2t).copy())
    L = _lib.lib()
    climb = L.dmsh_get_tuning(b"refit.climb")
    saved = {k: L.dmsh_batch_get_tuning(k) for k in (b"batch.stack", b"batch.wgs_per_cu",
                                                      b"batch.treelet_grid", b"batch.sort_items",
                                                      b"batch.cast_waves", b"batch.sort_seg")}
    try:
        # 19 meshes: 35-bit composite keys, sorted by three segmented passes
        # and the finish, or by five digit passes, each with its own status
        # region at every tile size
        for stack, wgs, tg, items, waves, seg in ((16, 1, 3, 8, 2, 0), (32, 64, 4096, 12, 4, 0),
                                                  (24, 3, 1024, 16, 1, 0), (16, 2, 7, 8, 2, 1),
                                                  (20, 4, 512, 12, 1, 1)):
            assert L.dmsh_batch_set_tuning(b"batch.stack", stack) == 0
            assert L.dmsh_batch_set_tuning(b"batch.wgs_per_cu", wgs) == 0
            assert L.dmsh_batch_set_tuning(b"batch.treelet_grid", tg) == 0
            assert L.dmsh_batch_set_tuning(b"batch.sort_items", items) == 0
            assert L.dmsh_batch_set_tuning(b"batch.cast_waves", waves) == 0
            assert L.dmsh_batch_set_tuning(b"batch.sort_seg", seg) == 0
            # the treelet climb: leaf-pair starts (2) or every leaf climbing (0)
            assert L.dmsh_set_tuning(b"refit.climb", 0 if items == 12 else 2) == 0
            p.run()
            got = (H(p.bvhnodes), H(p.bvhnode2aabb), H(p.pix2tri), H(p.pix2t))
            assert np.array_equal(got[0], ref[0]) and bits_equal(got[1], ref[1])
            assert np.array_equal(got[2], ref[2]) and bits_equal(got[3], ref[3])
    finally:
        for k, v in saved.items():
            L.dmsh_batch_set_tuning(k, v)
        L.dmsh_set_tuning(b"refit.climb", climb)
    no_device_errors()


def test_packed_pipeline_images_equal(gpu):
    """A packed pipeline (build into packed buffers, cast as requests) gives
    the images of the default one, eagerly and replayed from a graph; meshes
    of 0, 1, 2 and 1050 triangles included."""
    b = _edge_batch()
    p = _pipeline(b, gpu)
    p.run()
    torch.cuda.synchronize()
    ref = (H(p.pix2tri).copy(), H(p.pix2t).copy())
    q = Batched.Pipeline(b["mesh2tri"], b["mesh2vtx"], G(b["tri2vtx"], gpu),
                         G(b["vtx2xyz"], gpu), b["mesh2pix"], b["mesh2width"],
                         G(b["ndc2world"], gpu), True, packed=True)
    q.run()
    torch.cuda.synchronize()
    assert np.array_equal(H(q.pix2tri), ref[0]) and bits_equal(H(q.pix2t), ref[1])
    assert np.array_equal(H(q.mesh2node), H(p.mesh2node))
    q.capture()
    for _ in range(2):
        q.pix2tri.fill_(0)
        torch.cuda.synchronize()
        q.replay()
        torch.cuda.synchronize()
        assert np.array_equal(H(q.pix2tri), ref[0]) and bits_equal(H(q.pix2t), ref[1])
    with pytest.raises(ValueError):
        q.mesh_tree(0)
    no_device_errors()


def test_batch_graph_replays_match_eager(gpu):
    b = _edge_batch()
    p = _pipeline(b, gpu)
    p.run()
    torch.cuda.synchronize()
    ref = (H(p.bvhnodes).copy(), H(p.bvhnode2aabb).copy(), H(p.pix2tri).copy(), H(p.pix2t).copy())
    p.capture()
    for k in range(3):
        # eager work and a sync between replays
        p.bvhnodes.fill_(0)
        p.pix2tri.fill_(7)
        torch.cuda.synchronize()
        p.replay()
        torch.cuda.synchronize()
        assert np.array_equal(H(p.bvhnodes), ref[0]), f"replay {k}"
        assert bits_equal(H(p.bvhnode2aabb), ref[1])
        assert np.array_equal(H(p.pix2tri), ref[2])
        assert bits_equal(H(p.pix2t), ref[3])
    no_device_errors()


def test_batch_generated_sample(gpu, oracle):
    """300 generated meshes (256..65536 triangles); 16 of them checked
    against the oracle, all of them for device errors."""
    b = primitives.batch_meshes(300, seed=3, tri_range=(256, 65536),
                                pix_range=(64 * 64, 160 * 160))
    b["tris"] = [b["tri2vtx"][b["mesh2tri"][m]:b["mesh2tri"][m + 1]] for m in range(300)]
    b["vtxs"] = [b["vtx2xyz"][b["mesh2vtx"][m]:b["mesh2vtx"][m + 1]] for m in range(300)]
    p = _pipeline(b, gpu)
    p.run()
    torch.cuda.synchronize()
    no_device_errors()
    rng = np.random.default_rng(0)
    sizes = np.diff(b["mesh2tri"].astype(np.int64))
    pick = set(rng.choice(300, size=12, replace=False).tolist())
    pick |= {int(np.argmax(sizes)), int(np.argmin(sizes)), 0, 299}
    _check_against_oracle(oracle, p, b, sorted(pick))


def test_c3_full_batch_at_size(gpu, oracle):
    """The batched workload at its stated size: 4096 meshes (seed 3,
    256..262144 triangles log-uniform, 64^2..512^2 images), autotuned.
    Composite keys mesh << 30 | code take 42 bits here: six digit passes, or
    three segmented passes over the codes and the finish, as tuned.
    Every mesh's tree and image from the batched launches equal the ones the
    single-mesh entry points give it, one call per mesh; the largest, the
    smallest and 16 random meshes equal the oracle's."""
    b = primitives.batch_meshes(4096, seed=3)
    sizes = np.diff(b["mesh2tri"].astype(np.int64))
    assert len(sizes) == 4096 and sizes.max() > 200_000 and sizes.min() < 400
    assert int(b["mesh2tri"][-1]) > 1.4e8
    assert 30 + int(np.ceil(np.log2(4096))) == 42
    p = _pipeline(b, gpu)
    tuned = Batched.autotune(p, reps=1)
    print("autotuned:", {k: v for k, v in tuned.items() if k != "trials"})
    p.run()
    torch.cuda.synchronize()
    no_device_errors()
    loop = Batched.PerMeshLoop(p)
    loop.build()
    loop.cast()
    torch.cuda.synchronize()
    assert loop.equals_pipeline()
    no_device_errors()
    b["tris"] = {}
    b["vtxs"] = {}
    rng = np.random.default_rng(3)
    pick = set(rng.choice(4096, size=16, replace=False).tolist())
    pick |= {int(np.argmax(sizes)), int(np.argmin(sizes))}
    for m in pick:
        b["tris"][m] = b["tri2vtx"][b["mesh2tri"][m]:b["mesh2tri"][m + 1]]
        b["vtxs"][m] = b["vtx2xyz"][b["mesh2vtx"][m]:b["mesh2vtx"][m + 1]]
    _check_against_oracle(oracle, p, b, sorted(pick))
