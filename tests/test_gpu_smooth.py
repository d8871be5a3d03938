"""GPU parity of the smoothed edge gradient: pixel-edge types and gradients,
red-black smoothing and staggered-grid interpolation are bit-exact against
the oracle (same binary32 operations); AutogradWithSmooth's vertex gradient
against the oracle pipeline followed by the projection's chain rule in
float64."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Pix2Tri.torch as Pix2Tri
import del_msh_dlpack.RasterizedEdgeGradient.torch as REG
import del_msh_dlpack.TriMesh3.torch as TriMesh3

pytestmark = pytest.mark.gpu
U32MAX = 0xFFFFFFFF


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def _scene(camera, W=160, Hh=128):
    tri, vtx = primitives.torus_zup(1.0, 0.35, 48, 24)
    c, s = np.cos(0.8), np.sin(0.8)
    vtx = (vtx.astype(np.float64) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).T).astype(np.float32)
    if camera == "ortho":
        w2n = np.diag([0.7, 0.7, 0.35, 1.0])
    else:
        f = 1.0 / np.tan(np.radians(22.5))
        proj = np.array([[f, 0, 0, 0], [0, f, 0, 0], [0, 0, 1.5, 2.5], [0, 0, -1, 0]])
        view = np.eye(4)
        view[:3, 3] = [0.1, -0.1, -3.2]
        w2n = proj @ view
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * Hh, 0, 0.5 * Hh], [0, 0, 0.5, 0.5], [0, 0, 0, 1]])
    return tri, vtx, np.linalg.inv(w2n).astype(np.float32), (n2p @ w2n).astype(np.float32), (W, Hh)


def _render(gpu, tri, vtx, n2w, shape):
    t, v = G(tri, gpu), G(vtx, gpu)
    bn, ab = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v)
    return Pix2Tri.by_raycasting(t, v, bn, ab, G(n2w, gpu), shape)


@pytest.mark.parametrize("camera", ["ortho", "persp"])
def test_edge_gradient_and_type(gpu, oracle, camera):
    tri, vtx, n2w, w2p, shape = _scene(camera)
    pix = _render(gpu, tri, vtx, n2w, shape)
    rng = np.random.default_rng(5)
    occ = (H(pix) != U32MAX).astype(np.float32)[..., None]
    val = np.concatenate([occ, rng.random(occ.shape, dtype=np.float32)], 2)
    dl = rng.standard_normal(val.shape).astype(np.float32)
    got = REG.edge_gradient_and_type(G(tri, gpu), G(vtx, gpu), G(w2p, gpu), pix, G(val, gpu), G(dl, gpu))
    want = oracle.edge_gradient_and_type(tri, vtx, oracle.colmajor(w2p), H(pix), val, dl)
    for g, w in zip(got, want):
        assert np.array_equal(H(g), w)
    assert (want[0] == 2).any() and (want[0] == 3).any() and (want[0] == 1).any()
    assert _lib.device_errors() == 0


@pytest.mark.parametrize("num_iter", [1, 7, 100])
def test_smooth_gradient_red_black(gpu, oracle, num_iter):
    tri, vtx, n2w, w2p, shape = _scene("persp")
    pix = _render(gpu, tri, vtx, n2w, shape)
    occ = (H(pix) != U32MAX).astype(np.float32)[..., None]
    dl = (occ - np.roll(occ, 4, axis=0)).astype(np.float32)
    ht, hd, vt, vd = REG.edge_gradient_and_type(G(tri, gpu), G(vtx, gpu), G(w2p, gpu), pix, G(occ, gpu),
                                                G(dl, gpu))
    want_h, want_v = oracle.smooth_gradient(H(ht), H(vt), num_iter, H(hd), H(vd))
    REG.smooth_gradient(ht, vt, num_iter, hd, vd)
    assert np.array_equal(H(hd), want_h)
    assert np.array_equal(H(vd), want_v)


@pytest.mark.parametrize("num_iter", [8, 16, 23])
@pytest.mark.parametrize("shape", [(2, 2), (77, 301), (515, 33)])
def test_smooth_gradient_random_grids(gpu, oracle, shape, num_iter):
    """Random types (fixed cells anywhere), -0 values, image sizes that are
    not tile multiples, and launch counts of both parities; compared bit for
    bit, so a -0 turned into +0 fails."""
    W, Hh = shape
    rng = np.random.default_rng(W * 1000 + Hh + num_iter)
    ht = rng.integers(0, 5, (Hh - 1, W)).astype(np.uint8)
    vt = rng.integers(0, 5, (Hh, W - 1)).astype(np.uint8)
    hd = rng.standard_normal((Hh - 1, W)).astype(np.float32)
    vd = rng.standard_normal((Hh, W - 1)).astype(np.float32)
    hd[rng.random(hd.shape) < 0.1] = -0.0
    vd[rng.random(vd.shape) < 0.1] = -0.0
    want_h, want_v = oracle.smooth_gradient(ht, vt, num_iter, hd, vd)
    gh, gv = G(hd, gpu), G(vd, gpu)
    REG.smooth_gradient(G(ht, gpu), G(vt, gpu), num_iter, gh, gv)
    assert np.array_equal(H(gh).view(np.uint32), want_h.view(np.uint32))
    assert np.array_equal(H(gv).view(np.uint32), want_v.view(np.uint32))


def test_interpolate_staggered(gpu, oracle):
    rng = np.random.default_rng(4)
    Hh, W = 37, 53
    hd = rng.standard_normal((Hh - 1, W)).astype(np.float32)
    vd = rng.standard_normal((Hh, W - 1)).astype(np.float32)
    xy = (rng.random((5000, 2)) * [W + 4, Hh + 4] - 2).astype(np.float32)  # also outside the image
    got = REG.interpolate(G(hd, gpu), G(vd, gpu), G(xy, gpu))
    assert np.array_equal(H(got), oracle.interpolate_staggered_grid(hd, vd, xy))


def test_autograd_with_smooth(gpu, oracle):
    tri, vtx, n2w, w2p, shape = _scene("persp")
    pix = _render(gpu, tri, vtx, n2w, shape)
    occ = (H(pix) != U32MAX).astype(np.float32)[..., None]
    trg = np.roll(occ, 3, axis=1)
    v = G(vtx, gpu).requires_grad_(True)
    out = REG.AutogradWithSmooth.apply(G(tri, gpu), v, G(w2p, gpu), pix, G(occ, gpu))
    ((out - G(trg, gpu)) ** 2).sum().backward()
    dl = 2.0 * (occ - trg)
    ht, hd, vt, vd = oracle.edge_gradient_and_type(tri, vtx, oracle.colmajor(w2p), H(pix), occ, dl)
    hs, vs = oracle.smooth_gradient(ht, vt, 100, hd, vd)
    m = w2p.astype(np.float64)
    q = np.concatenate([vtx.astype(np.float64), np.ones((vtx.shape[0], 1))], 1) @ m.T
    u = (q[:, :2] / q[:, 3:]).astype(np.float32)
    g = oracle.interpolate_staggered_grid(hs, vs, u).astype(np.float64)
    w = q[:, 3:]
    dq = np.concatenate([g / w, np.zeros_like(w), -(g * (q[:, :2] / w)).sum(1, keepdims=True) / w], 1)
    want = dq @ m[:, :3]
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(H(v.grad) - want).max() <= 1e-4 * scale
