"""GPU: the serving loop's images equal what each request gets alone from the
single-mesh entry points, with trees evicted and rebuilt along the way."""

import copy

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib
import del_msh_dlpack.Serve.torch as Serve

pytestmark = pytest.mark.gpu


def H(t):
    return t.detach().cpu().numpy()


def test_served_images_equal_each_request_alone(gpu):
    cat = Serve.MeshCatalog(tri_range=(200, 20000), seed=2)
    reqs = Serve.request_stream(80, rate=5000.0, num_ids=40, side_range=(24, 96), seed=5)
    alone = copy.deepcopy(reqs)
    # room for a handful of trees only: the mix evicts and rebuilds
    srv = Serve.Server(cat, budget_bytes=3 << 20, device=gpu, max_step_requests=16,
                       max_step_pixels=1 << 15)
    steps = srv.run(reqs, keep_results=True)
    torch.cuda.synchronize()
    assert steps > 1
    assert srv.cache.evicted > 0 and srv.cache.hits > 0
    assert all(r.done >= r.arrival for r in reqs)
    Serve.PlainServer(cat, budget_bytes=1 << 30, device=gpu).run(alone, keep_results=True)
    for a, b in zip(reqs, alone):
        pa, ta = a.result
        pb, tb = b.result
        assert np.array_equal(H(pa), H(pb)), a.rid
        assert np.array_equal(H(ta).view(np.uint32), H(tb).view(np.uint32)), a.rid
        assert (H(pa) != 0xFFFFFFFF).any(), a.rid  # the camera sees the mesh
    bits = _lib.device_errors()
    assert bits == 0, hex(bits)


@pytest.mark.parametrize("climb", [2, 0])
def test_fused_pack_equals_build_then_pack(gpu, climb):
    """dmsh_batch_build_pack_trimesh3 writes the packed buffers byte for byte
    as dmsh_batch_build_trimesh3 followed by dmsh_batch_pack: meshes of one
    triangle, of a single treelet and of many, an empty mesh, runs of equal
    codes and a non-finite vertex; with either treelet climb."""
    from del_msh_dlpack import primitives
    L = _lib.lib()
    old_climb = L.dmsh_get_tuning(b"refit.climb")
    assert L.dmsh_set_tuning(b"refit.climb", climb) == 0
    try:
        _fused_pack_case(gpu, L, primitives)
    finally:
        L.dmsh_set_tuning(b"refit.climb", old_climb)


def _fused_pack_case(gpu, L, primitives):
    rng = np.random.default_rng(4)
    meshes = [primitives.torus_zup(0.7, 0.3, 40, 30), primitives.torus_zup(0.6, 0.2, 12, 9),
              (np.array([[0, 1, 2]], np.uint32),
               np.array([[0, 0, 0], [1, 0, 0], [0, 1, .5]], np.float32)),
              (np.zeros((0, 3), np.uint32), np.zeros((0, 3), np.float32)),
              (np.tile(np.array([[0, 1, 2]], np.uint32), (2000, 1)),
               rng.uniform(-1, 1, (3, 3)).astype(np.float32)),
              primitives.octahedral_sphere(6, noise_sigma=0.01, seed=3)]
    # a mesh with an infinite vertex: its boxes are not finite
    t_inf, v_inf = primitives.torus_zup(0.5, 0.2, 10, 8)
    v_inf = v_inf.copy()
    v_inf[3, 1] = np.inf
    meshes.append((t_inf, v_inf))
    ntri = np.array([len(t) for t, _ in meshes], np.int64)
    nvtx = np.array([len(v) for _, v in meshes], np.int64)
    m2t = np.concatenate([[0], np.cumsum(ntri)]).astype(np.uint32)
    m2v = np.concatenate([[0], np.cumsum(nvtx)]).astype(np.uint32)
    M, T, V = len(meshes), int(m2t[-1]), int(m2v[-1])
    tri = torch.from_numpy(np.concatenate([t for t, _ in meshes])).to(gpu)
    vtx = torch.from_numpy(np.concatenate([v for _, v in meshes])).to(gpu)
    d_m2t, d_m2v = torch.from_numpy(m2t).to(gpu), torch.from_numpy(m2v).to(gpu)
    rows = int((2 * ntri[ntri > 0] - 1).sum())
    sp = torch.cuda.current_stream().cuda_stream
    out = []
    for fused in (False, True):
        m2n = torch.empty(M + 1, dtype=torch.uint32, device=gpu)
        nodes = torch.empty((rows, 3), dtype=torch.uint32, device=gpu)
        aabb = torch.empty((rows, 6), dtype=torch.float32, device=gpu)
        bufs = [torch.full((int(L.dmsh_packed_bytes(int(n))),), 0xA5, dtype=torch.uint8,
                           device=gpu) for n in ntri]
        addrs = torch.tensor([b.data_ptr() for b in bufs], dtype=torch.uint64, device=gpu)
        nonfinite = torch.full((M,), 7, dtype=torch.uint32, device=gpu)
        ws = _lib.workspace(L.dmsh_batch_build_workspace_bytes(M, T), gpu)
        args = (sp, M, T, V, d_m2t.data_ptr(), d_m2v.data_ptr(), tri.data_ptr(), vtx.data_ptr(),
                m2n.data_ptr(), nodes.data_ptr(), aabb.data_ptr())
        if fused:
            _lib.check(L.dmsh_batch_build_pack_trimesh3(*args, addrs.data_ptr(),
                                                        nonfinite.data_ptr(), ws.data_ptr(),
                                                        ws.numel()), "build + pack")
        else:
            _lib.check(L.dmsh_batch_build_trimesh3(*args, ws.data_ptr(), ws.numel()), "build")
            wsp = _lib.workspace(L.dmsh_batch_pack_workspace_bytes(M), gpu)
            _lib.check(L.dmsh_batch_pack(sp, M, T, V, d_m2t.data_ptr(), d_m2v.data_ptr(),
                                         tri.data_ptr(), vtx.data_ptr(), m2n.data_ptr(),
                                         nodes.data_ptr(), aabb.data_ptr(), addrs.data_ptr(),
                                         nonfinite.data_ptr(), wsp.data_ptr(), wsp.numel()),
                       "pack")
        torch.cuda.synchronize()
        out.append(([H(b).copy() for b in bufs], H(nonfinite).copy(), H(nodes).copy(),
                    H(aabb).view(np.uint32).copy()))
    (b_a, nf_a, n_a, a_a), (b_b, nf_b, n_b, a_b) = out
    # the fused build keeps only the boxes its top merge needs; the roots'
    # are among them
    root_rows = np.concatenate([[0], np.cumsum(2 * ntri[ntri > 0] - 1)[:-1]])
    assert np.array_equal(n_a, n_b) and np.array_equal(a_a[root_rows], a_b[root_rows])
    assert np.array_equal(nf_a, nf_b) and nf_b[-1] != 0
    for m in range(M):
        assert np.array_equal(b_a[m], b_b[m]), f"mesh {m}"
    bits = _lib.device_errors()
    assert bits == 0, hex(bits)


def test_cast_requests_rejects_bad_offsets(gpu):
    """A request whose pixel offset is not the running sum is rejected on the
    device: nothing is written and DMSH_DEVERR_INDEX is raised."""
    cat = Serve.MeshCatalog(tri_range=(200, 400), seed=2)
    srv = Serve.Server(cat, budget_bytes=1 << 30, device=gpu)
    reqs = Serve.request_stream(2, rate=1e6, num_ids=3, side_range=(16, 16), seed=1)
    srv.run(reqs)  # builds the trees
    _lib.device_errors()
    recs = np.zeros(2, Serve.REQ_DTYPE)
    for k, r in enumerate(reqs):
        buf, n, nf = srv.cache.entries[r.mesh_id].value
        recs[k] = (buf.data_ptr(), n, 0, 16, 16, 0, 0, r.ndc2world.T.reshape(-1))  # both start at 0
    d_req = torch.from_numpy(recs.view(np.uint8)).to(gpu)
    pix = torch.full((512,), 7, dtype=torch.uint32, device=gpu)
    L = _lib.lib()
    ws = _lib.workspace(L.dmsh_cast_requests_workspace_bytes(2, 512), gpu)
    _lib.check(L.dmsh_cast_requests(torch.cuda.current_stream().cuda_stream, 2, d_req.data_ptr(),
                                    512, pix.data_ptr(), None, ws.data_ptr(), ws.numel()), "cast")
    assert _lib.device_errors() & 0x1
    assert np.all(H(pix) == 7)


def test_c5_served_at_config_sizes(gpu, oracle):
    """The served workload at its stated sizes: 100 Poisson requests (seed 7)
    at 20,000 req/s, mesh ids Zipf(1.1) over 10,000 ids, 1k..1M triangles
    per mesh with radial noise, perspective viewports of 256^2..1024^2, a
    1 GiB tree cache.  Every served image (hit triangle and hit coefficient)
    equals what the request gets alone from the single-mesh entry points, and
    four requests (the largest and the smallest mesh, two at random) equal the
    oracle's build and ordered ray-cast."""
    cat = Serve.MeshCatalog(tri_range=(1000, 1_000_000), seed=7)
    reqs = Serve.request_stream(100, rate=20000.0, num_ids=10_000, side_range=(256, 1024), seed=7)
    sides = [r.width for r in reqs]
    ntri = [cat.num_tri(r.mesh_id) for r in reqs]
    print(f"sides {min(sides)}..{max(sides)}, triangles {min(ntri)}..{max(ntri)}, "
          f"{len(set(r.mesh_id for r in reqs))} meshes")
    assert min(sides) >= 256 and max(sides) <= 1024 and max(sides) > 700
    assert min(ntri) >= 500 and max(ntri) > 300_000
    alone = copy.deepcopy(reqs)
    srv = Serve.Server(cat, budget_bytes=1 << 30, device=gpu)
    steps = srv.run(reqs, keep_results=True)
    torch.cuda.synchronize()
    assert steps >= 1
    Serve.PlainServer(cat, budget_bytes=1 << 32, device=gpu).run(alone, keep_results=True)
    for a, b in zip(reqs, alone):
        pa, ta = a.result
        pb, tb = b.result
        assert torch.equal(pa, pb), a.rid
        assert torch.equal(ta.view(torch.int32), tb.view(torch.int32)), a.rid
    bits = _lib.device_errors()
    assert bits == 0, hex(bits)
    rng = np.random.default_rng(7)
    pick = {int(np.argmax(ntri)), int(np.argmin(ntri))} | set(rng.choice(100, 2, replace=False).tolist())
    for k in sorted(pick):
        r = reqs[k]
        tri, vtx = cat.get(r.mesh_id)
        bn, ab = oracle.build_trimesh3(tri, vtx)
        wpix, wt, _ = oracle.pix2tri(tri, vtx, bn, ab, oracle.colmajor(r.ndc2world), r.width,
                                     r.height, num_threads=16)
        pix, t = r.result
        assert np.array_equal(H(pix), wpix), r.rid
        assert np.array_equal(H(t).view(np.uint32), wt.view(np.uint32)), r.rid
        assert (wpix != 0xFFFFFFFF).any(), r.rid
