"""GPU: differentiation workload.  A 50,000-triangle sphere rendered at
512 x 512, upstream dL/dpixel ~ N(0, 1) (seed 4): the anti-aliased
silhouette's vertex gradient through DifferentiableAntialiasFunction against
the oracle and against central differences (h = 1e-3) on 64 contour
vertices (seed 5); the edge-gradient backward against the oracle; and a
batch of 64 uneven meshes (5k-200k triangles, seed 6) whose hit buffers are
recomputed in the backward pass."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Batched.torch as Batched
import del_msh_dlpack.DifferentiableAntialias.torch as DifferentiableAntialias
import del_msh_dlpack.Edge2Vtx.torch as Edge2Vtx
import del_msh_dlpack.Mat44.torch as Mat44
import del_msh_dlpack.Pix2Tri.torch as Pix2Tri
import del_msh_dlpack.RasterizedEdgeGradient.torch as RasterizedEdgeGradient
import del_msh_dlpack.TriMesh3.torch as TriMesh3

pytestmark = pytest.mark.gpu

U32MAX = 0xFFFFFFFF
W = H_ = 512


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def no_device_errors():
    bits = _lib.device_errors()
    assert bits == 0, f"device error bits 0x{bits:x}"


def _camera():
    w2c = Mat44.look_at((0.0, 0.0, 3.0), (0, 0, 0), (0, 1, 0), dtype=torch.float64).numpy()
    proj = Mat44.perspective(45.0, 1.0, 1.0, 5.0, dtype=torch.float64).numpy()
    w2n = proj @ w2c
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * H_, 0, 0.5 * H_], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]])
    return (w2n.astype(np.float32), np.linalg.inv(w2n).astype(np.float32),
            (n2p @ w2n).astype(np.float32))


@pytest.fixture(scope="module")
def scene(gpu):
    tri, vtx = primitives.sphere_yup(1.0, 126, 200)
    assert tri.shape[0] == 50000 and vtx.shape[0] == 25002
    w2n, n2w, w2p = _camera()
    t, v = G(tri, gpu), G(vtx, gpu)
    bvhnodes, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v)
    pix = Pix2Tri.by_raycasting(t, v, bvhnodes, aabb, G(n2w, gpu), (W, H_))
    e2v = TriMesh3.make_edge2vtx(t, vtx.shape[0])
    e2t = TriMesh3.make_edge2tri(t, vtx.shape[0], e2v)
    one = e2v[:, 0].to(torch.int64) < e2v[:, 1].to(torch.int64)
    # (torch indexes int32, not uint32, rows)
    keep = lambda a: a.view(torch.int32)[one].view(torch.uint32).contiguous()
    cedge = Edge2Vtx.contour_for_triangle_mesh(t, v, G(w2n, gpu), keep(e2v), keep(e2t))
    cedge_both = Edge2Vtx.contour_for_triangle_mesh(t, v, G(w2n, gpu), e2v, e2t)
    occ = (pix != U32MAX).to(torch.float32)
    dldw = np.random.default_rng(4).standard_normal((H_, W)).astype(np.float32)
    return dict(tri=tri, vtx=vtx, w2n=w2n, n2w=n2w, w2p=w2p, t=t, v=v, pix=pix, cedge=cedge,
                cedge_both=cedge_both, occ=occ, dldw=dldw)


def test_c4_hit_buffers_match_oracle(gpu, oracle, scene):
    s = scene
    bn, aabb = oracle.build_trimesh3(s["tri"], s["vtx"])
    wpix, _, _ = oracle.pix2tri(s["tri"], s["vtx"], bn, aabb, oracle.colmajor(s["n2w"]), W, H_,
                                num_threads=8)
    assert np.array_equal(H(s["pix"]), wpix)
    cover = (wpix != U32MAX).mean()
    assert 0.4 < cover < 0.8  # the sphere fills about 60% of the image
    e2v, e2t = primitives.edge2vtx_edge2tri(s["tri"])
    want = oracle.contour(s["tri"], s["vtx"], oracle.colmajor(s["w2n"]), e2v, e2t)
    got = H(s["cedge"])
    key = lambda a: np.sort(a[:, 0].astype(np.int64) * 1_000_003 + a[:, 1])
    assert np.array_equal(key(got), key(want))
    no_device_errors()


def _autograd_grad(s, cedge, gpu):
    v = s["v"].clone().requires_grad_(True)
    out = DifferentiableAntialias.DifferentiableAntialiasFunction.apply(
        cedge, v, G(s["w2p"], gpu), s["pix"], s["occ"])
    (G(s["dldw"], gpu) * out).sum().backward()
    return H(v.grad), out


def test_c4_antialias_gradient_vs_oracle(gpu, oracle, scene):
    s = scene
    g, out = _autograd_grad(s, s["cedge"], gpu)
    occ = H(s["occ"])
    want_out = oracle.antialias(H(s["cedge"]), s["vtx"], oracle.colmajor(s["w2p"]), H(s["pix"]), occ)
    assert np.array_equal(H(out).view(np.uint32), want_out.view(np.uint32))
    want = oracle.bwd_antialias(H(s["cedge"]), s["vtx"], oracle.colmajor(s["w2p"]), occ, s["dldw"],
                                H(s["pix"]))
    scale = np.abs(want).max()
    assert scale > 0
    # per-vertex sums of float atomics in another order than the serial loop
    assert np.abs(g - want).max() <= 2e-5 * scale
    no_device_errors()


def test_c4_edges_from_both_ends_double_the_gradient(gpu, scene):
    """The reference's make_edge2vtx lists every edge from both ends, so its
    contour holds each silhouette edge twice: the image is the same up to the
    rounding of the reversed edge's crossing parameters, and the antialias
    backward adds each edge's contribution twice."""
    s = scene
    g1, out1 = _autograd_grad(s, s["cedge"], gpu)
    g2, out2 = _autograd_grad(s, s["cedge_both"], gpu)
    assert s["cedge_both"].shape[0] == 2 * s["cedge"].shape[0]
    d_out = (out1 - out2).abs().max().item()
    scale = np.abs(g1).max()
    d_g = np.abs(g2 - 2.0 * g1).max()
    print(f"image difference {d_out:.3e}, gradient difference {d_g / scale:.3e} of max")
    assert d_out <= 1e-5
    assert d_g <= 1e-5 * scale


def test_c4_antialias_gradient_vs_central_differences(gpu, scene):
    """L(v) = sum dldw * AA(occ; v) with pix2tri and the contour held fixed
    (as the autograd function holds them), against central differences with
    h = 1e-3 on each coordinate of 64 contour vertices (seed 5); 64 vertices
    off the contour have zero gradient and leave the image unchanged.

    The anti-aliased image is piecewise smooth in the vertex positions: a
    walk that gains or loses a pixel pair, or a crossing that leaves its edge
    segment, makes a jump, and at 512^2 a 1e-3 move (0.2 px) crosses one in
    most samples.  So each pixel is also sampled at +-h/2: a side whose three
    values are not on one line holds a jump, and the pixel's difference is
    taken on its smooth side (central when both are).  A sample with a pixel
    that jumps on both sides is set aside.  Where the move is 1e-5, jumps
    are rare and plain central differences are checked as well."""
    s = scene
    g, _ = _autograd_grad(s, s["cedge"], gpu)
    dldw = G(s["dldw"], gpu).to(torch.float64)
    w2p = G(s["w2p"], gpu)

    def image(vv):
        return DifferentiableAntialias.antialias(s["cedge"], vv, w2p, s["pix"], s["occ"]).to(torch.float64)

    def moved(iv, k, d):
        vv = s["v"].clone()
        vv[iv, k] += d
        return vv, float(vv[iv, k]) - float(s["v"][iv, k])  # the step f32 actually took

    rng = np.random.default_rng(5)
    on = np.unique(H(s["cedge"]).reshape(-1))
    sel = rng.choice(on, size=64, replace=False)
    base = image(s["v"])
    scale = np.abs(g).max()
    h = 1e-3
    fds, ana, skipped = [], [], 0
    small_fd, small_ana = [], []
    for iv in sel:
        for k in range(3):
            (vp1, hp), (vp2, _), (vm2, _), (vm1, hm) = (moved(iv, k, d) for d in (h, h / 2, -h / 2, -h))
            p1, p2, m2, m1 = image(vp1), image(vp2), image(vm2), image(vm1)
            right = (p1 - 2.0 * p2 + base).abs() <= 1e-4
            left = (base - 2.0 * m2 + m1).abs() <= 1e-4
            slope = torch.where(right & left, (p1 - m1) / (hp - hm),
                                torch.where(right, (p1 - base) / hp, (base - m1) / -hm))
            if bool(((~right) & (~left) & (dldw != 0)).any()):
                skipped += 1
            else:
                fds.append(float((slope * dldw).sum()))
                ana.append(float(g[iv, k]))
            (vs, hs), (vt, ht) = moved(iv, k, 1e-5), moved(iv, k, -1e-5)
            small_fd.append(float(((image(vs) - image(vt)) * dldw).sum()) / (hs - ht))
            small_ana.append(float(g[iv, k]))
    fds, ana = np.array(fds), np.array(ana)
    err = np.abs(fds - ana)
    small_err = np.abs(np.array(small_fd) - np.array(small_ana))
    print(f"h=1e-3: {len(fds)} samples, {skipped} set aside; |g| max {scale:.3e}; err median "
          f"{np.median(err):.3e} max {err.max():.3e}; h=1e-5: err median {np.median(small_err):.3e} "
          f"p90 {np.quantile(small_err, 0.9):.3e}")
    assert len(fds) >= 64  # of 192
    assert np.abs(fds).max() > 0.1 * scale  # the silhouette does move
    rel = err / (np.maximum(np.abs(fds), np.abs(ana)) + 1e-3 * scale)
    # a jump below the 1e-4 detection threshold can still tilt a sample
    assert np.median(rel) < 1e-3
    assert np.mean(rel <= 0.02) >= 0.95
    assert np.mean(small_err <= 0.02 * np.abs(np.array(small_ana)) + 1e-3 * scale) >= 0.9
    off = np.setdiff1d(np.arange(s["vtx"].shape[0]), on)
    for iv in rng.choice(off, size=64, replace=False):
        assert np.all(g[iv] == 0.0)
        assert torch.equal(image(moved(iv, 0, h)[0]), base)


def test_c4_edge_gradient_vs_oracle(gpu, oracle, scene):
    s = scene
    occ = H(s["occ"])[..., None]
    dldw = s["dldw"][..., None]
    got = H(RasterizedEdgeGradient.bwd(s["t"], s["v"], G(s["w2p"], gpu), s["pix"],
                                       G(occ, gpu), G(dldw, gpu)))
    want = oracle.edge_gradient_bwd(s["tri"], s["vtx"], oracle.colmajor(s["w2p"]), H(s["pix"]), occ,
                                    dldw)
    assert np.abs(want).max() > 0
    # the ordered backward adds each vertex's terms in the serial loop's order
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the one-launch atomic variant agrees to summation rounding
    fast = H(RasterizedEdgeGradient.bwd(s["t"], s["v"], G(s["w2p"], gpu), s["pix"],
                                        G(occ, gpu), G(dldw, gpu), ordered=False))
    assert np.abs(fast - want).max() <= 1e-5 * np.abs(want).max() + 5e-6
    no_device_errors()


@pytest.fixture(scope="module")
def batch64(gpu):
    b = primitives.batch_meshes(64, seed=6, tri_range=(5000, 200000),
                                pix_range=(512 * 512, 512 * 512))
    n = np.diff(b["mesh2tri"].astype(np.int64))
    assert n.min() >= 2500 and n.max() <= 400000
    sb = Batched.SilhouetteBatch(b["mesh2tri"], b["mesh2vtx"], G(b["tri2vtx"], gpu),
                                 G(b["vtx2xyz"], gpu), b["mesh2pix"], b["mesh2width"],
                                 torch.from_numpy(b["ndc2world"]))
    return b, sb


def test_batch_recompute_equals_stored(gpu, batch64):
    b, sb = batch64
    dldw = G(np.random.default_rng(4).standard_normal(int(b["mesh2pix"][-1])).astype(np.float32), gpu)
    grads, imgs = [], []
    for recompute in (True, False):
        v = G(b["vtx2xyz"], gpu).requires_grad_(True)
        img = Batched.AntialiasedSilhouette.apply(v, sb, recompute)
        (img * dldw).sum().backward()
        grads.append(H(v.grad))
        imgs.append(img)
    assert torch.equal(imgs[0], imgs[1])
    scale = np.abs(grads[1]).max()
    assert scale > 0
    assert np.abs(grads[0] - grads[1]).max() <= 1e-5 * scale
    # the recomputed hit buffers are the forward's, bit for bit
    h1, h2 = sb.hits(G(b["vtx2xyz"], gpu)), sb.hits(G(b["vtx2xyz"], gpu))
    assert torch.equal(h1["pix2tri"], h2["pix2tri"])
    assert all(torch.equal(x, y) for x, y in zip(h1["cedges"], h2["cedges"]))
    no_device_errors()


def test_batch_meshes_equal_single_mesh_path(gpu, oracle, batch64):
    """Three meshes of the batch against the single-mesh entry points (image
    bit-exact, gradient to atomics rounding) and the smallest against the
    oracle's antialias backward."""
    b, sb = batch64
    v = G(b["vtx2xyz"], gpu)
    dldw_all = np.random.default_rng(4).standard_normal(int(b["mesh2pix"][-1])).astype(np.float32)
    hits = sb.hits(v)
    img = sb.images(v, hits)
    grad = H(sb.vertex_gradient(v, hits, G(dldw_all, gpu)))
    n = np.diff(b["mesh2tri"].astype(np.int64))
    for m in sorted({0, 31, int(np.argmin(n))}):
        t0, t1 = int(b["mesh2tri"][m]), int(b["mesh2tri"][m + 1])
        v0, v1 = int(b["mesh2vtx"][m]), int(b["mesh2vtx"][m + 1])
        p0, p1 = int(b["mesh2pix"][m]), int(b["mesh2pix"][m + 1])
        tri, vtx = b["tri2vtx"][t0:t1], b["vtx2xyz"][v0:v1]
        n2w = b["ndc2world"][m]
        t, vv = G(tri, gpu), G(vtx, gpu)
        bn, ab = TriMesh3.make_bvhnodes_bvhnode2aabb(t, vv, morton_bits=30)
        pix = Pix2Tri.by_raycasting(t, vv, bn, ab, G(n2w, gpu), (W, H_))
        assert torch.equal(pix.view(-1), hits["pix2tri"][p0:p1])
        occ = (pix != U32MAX).to(torch.float32)
        w2p = sb.world2pix[m]
        c = hits["cedges"][m]
        out = DifferentiableAntialias.antialias(c, vv, w2p, pix, occ)
        assert torch.equal(out.view(-1), img[p0:p1])
        dl = dldw_all[p0:p1].reshape(H_, W)
        g1 = H(DifferentiableAntialias.bwd_antialias(c, vv, w2p, occ, G(dl, gpu), pix))
        scale = np.abs(g1).max()
        assert scale > 0
        assert np.abs(grad[v0:v1] - g1).max() <= 1e-5 * scale
        if m == int(np.argmin(n)):
            want = oracle.bwd_antialias(H(c), vtx, oracle.colmajor(H(w2p)), H(occ), dl, H(pix))
            assert np.abs(g1 - want).max() <= 2e-5 * np.abs(want).max()
    no_device_errors()


def test_c4_rasterized_edge_gradient_function_autograd(gpu, oracle, scene):
    """RasterizedEdgeGradientFunction through torch.autograd on the C4 scene:
    the forward is the identity on pix2vin, the vertex gradient equals the
    oracle's edge-gradient backward bit for bit, and dL/dpix2vin passes
    through unchanged (reference RasterizedEdgeGradient/torch.py:214-234)."""
    s = scene
    v = s["v"].clone().requires_grad_(True)
    pin = s["occ"][..., None].clone().requires_grad_(True)
    w2p = G(s["w2p"], gpu)
    out = RasterizedEdgeGradient.RasterizedEdgeGradientFunction.apply(s["t"], v, w2p, s["pix"], pin)
    assert torch.equal(out, pin)
    dldw = G(s["dldw"][..., None], gpu)
    (dldw * out).sum().backward()
    assert torch.equal(pin.grad, dldw)
    want = oracle.edge_gradient_bwd(s["tri"], s["vtx"], oracle.colmajor(s["w2p"]), H(s["pix"]),
                                    H(s["occ"])[..., None], s["dldw"][..., None])
    got = H(v.grad)
    scale = np.abs(want).max()
    print(f"edge-gradient scale max|dL/dx| = {scale:.4g}, max |got - want| = "
          f"{np.abs(got - want).max():.3g}")
    assert scale > 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    no_device_errors()


def _reference_example1(gpu):
    """The scene of del-msh-dlpack/tests/test_differentiable_antialias.py:29-37:
    torus (1.3, 0.4, 64, 32) rotated about x by 1.15 and moved up 0.6,
    world2ndc = scale(0.5), 128 x 128."""
    import del_msh_dlpack.Vtx2Xyz.torch as Vtx2Xyz

    tri2vtx, vtx2xyz = TriMesh3.torus(1.3, 0.4, 64, 32)
    transform = Mat44.from_translation(0.0, 0.6, 0.0) @ Mat44.from_x_rotation(1.15)
    vtx2xyz = Vtx2Xyz.transform_affine(vtx2xyz, transform)
    world2ndc = Mat44.from_scale(0.5, 0.5, 0.5)
    return tri2vtx, vtx2xyz, world2ndc, (128, 128)


def test_reference_match_cpu_gpu_microedge_bwd(gpu, oracle):
    """The reference's own CPU<->GPU check of the edge-gradient backward
    (del-msh-dlpack/tests/test_rasterized_edge_gradient.py:58-93): the same
    scene, dL/docc ~ U[0, 1) from torch's generator seeded 0, and its bound
    max |gpu - cpu| < 5e-6, here against the oracle."""
    tri2vtx, vtx2xyz, world2ndc, img_shape = _reference_example1(gpu)
    ndc2world = world2ndc.inverse()
    world2pix = Mat44.from_transform_ndc2pix(img_shape) @ world2ndc
    t, v = tri2vtx.to(gpu), vtx2xyz.to(gpu)
    bvhnodes, aabb = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v)
    pix2tri = Pix2Tri.by_raycasting(t, v, bvhnodes, aabb, ndc2world.to(gpu), img_shape)
    pix2occ = torch.where(pix2tri == U32MAX, 0.0, 1.0).to(torch.float32).unsqueeze(-1)
    torch.random.manual_seed(0)
    dldw_pix2occ = torch.rand_like(pix2occ.cpu())
    got = RasterizedEdgeGradient.bwd(t, v, world2pix.to(gpu), pix2tri, pix2occ, dldw_pix2occ.to(gpu))
    tri, vtx = tri2vtx.numpy(), vtx2xyz.numpy()
    bn, ab = oracle.build_trimesh3(tri, vtx)
    wpix, _, _ = oracle.pix2tri(tri, vtx, bn, ab, oracle.colmajor(ndc2world.numpy()), 128, 128)
    assert np.array_equal(H(pix2tri), wpix)
    want = oracle.edge_gradient_bwd(tri, vtx, oracle.colmajor(world2pix.numpy()), wpix,
                                    H(pix2occ), dldw_pix2occ.numpy())
    err = np.abs(H(got) - want).max()
    print(f"example1: max|dL/dx| = {np.abs(want).max():.4g}, max |gpu - oracle| = {err:.3g}")
    assert np.abs(want).max() > 0
    assert err < 5.0e-6
    # the largest gradient (about 69) has an ulp of 7.6e-6, above that bound:
    # only the serial summation order meets it, and the ordered backward has it
    assert np.array_equal(H(got).view(np.uint32), want.view(np.uint32))
    no_device_errors()
