"""GPU parity of attribute interpolation over pix2tri (forward bit-exact:
the same binary32 plane barycentrics and weighted sum; backward to the
rounding of float atomics) and of Pix2Tri.AutogradInterpolate."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.Pix2Tri.torch as Pix2Tri
import del_msh_dlpack.TriMesh3.torch as TriMesh3

pytestmark = pytest.mark.gpu
U32MAX = 0xFFFFFFFF


def G(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def H(t):
    return t.detach().cpu().numpy()


def _scene(camera):
    tri, vtx = primitives.sphere_yup(0.8, 48, 96)
    if camera == "ortho":
        w2n = np.diag([1.0, 1.0, 0.5, 1.0])
    else:
        f = 1.0 / np.tan(np.radians(22.5))
        proj = np.array([[f, 0, 0, 0], [0, f, 0, 0], [0, 0, 1.5, 2.5], [0, 0, -1, 0]])
        view = np.eye(4)
        view[:3, 3] = [-0.2, 0.1, -3.0]
        w2n = proj @ view
    return tri, vtx, np.linalg.inv(w2n).astype(np.float32)


@pytest.mark.parametrize("camera", ["ortho", "persp"])
@pytest.mark.parametrize("nv", [1, 3])
def test_interpolate_fwd_bwd(gpu, oracle, camera, nv):
    tri, vtx, n2w = _scene(camera)
    W, Hh = 200, 160
    t, v = G(tri, gpu), G(vtx, gpu)
    bn, ab = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v)
    pix = Pix2Tri.by_raycasting(t, v, bn, ab, G(n2w, gpu), (W, Hh))
    rng = np.random.default_rng(7 + nv)
    val = rng.random((vtx.shape[0], nv)).astype(np.float32)
    out = Pix2Tri.interpolate_fwd(pix, t, v, G(val, gpu), G(n2w, gpu))
    want = oracle.interpolate(H(pix), tri, vtx, val, oracle.colmajor(n2w))
    assert np.array_equal(H(out).view(np.uint32), want.view(np.uint32))
    dl = rng.standard_normal((Hh, W, nv)).astype(np.float32)
    gx = torch.zeros_like(v)
    gv = torch.zeros((vtx.shape[0], nv), dtype=torch.float32, device=gpu)
    Pix2Tri.interpolate_bwd(pix, t, v, G(val, gpu), G(n2w, gpu), G(dl, gpu), gx, gv)
    wx, wv = oracle.interpolate_bwd(H(pix), tri, vtx, val, oracle.colmajor(n2w), dl)
    for got, ref in ((H(gx), wx), (H(gv), wv)):
        scale = np.abs(ref).max()
        assert scale > 0
        assert np.abs(got - ref).max() <= 1e-5 * scale
    bits = _lib.device_errors()
    assert bits == 0, hex(bits)


def test_autograd_interpolate(gpu, oracle):
    tri, vtx, n2w = _scene("persp")
    W, Hh = 128, 128
    t = G(tri, gpu)
    v = G(vtx, gpu).requires_grad_(True)
    bn, ab = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v.detach())
    pix = Pix2Tri.by_raycasting(t, v.detach(), bn, ab, G(n2w, gpu), (W, Hh))
    rng = np.random.default_rng(3)
    val = G(rng.random((vtx.shape[0], 3)).astype(np.float32), gpu).requires_grad_(True)
    out = Pix2Tri.interpolate(pix, t, v, val, G(n2w, gpu))
    w = rng.standard_normal((Hh, W, 3)).astype(np.float32)
    (out * G(w, gpu)).sum().backward()
    wx, wv = oracle.interpolate_bwd(H(pix), tri, vtx, H(val), oracle.colmajor(n2w), w)
    assert np.abs(H(v.grad) - wx).max() <= 1e-5 * np.abs(wx).max()
    assert np.abs(H(val.grad) - wv).max() <= 1e-5 * np.abs(wv).max()
