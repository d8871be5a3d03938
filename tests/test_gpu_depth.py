"""GPU parity for the depth backward, Pix2Depth.AutogradFunction and the
depth -> anti-aliasing autograd chain, against the CPU oracle.

The depth forward is bit-exact; backward passes add with float atomics and
are compared with a stated tolerance."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.DifferentiableAntialias.torch as DifferentiableAntialias
import del_msh_dlpack.Mat44.torch as Mat44
import del_msh_dlpack.Pix2Depth.torch as Pix2Depth

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


def _scene(oracle, kind):
    """(tri, vtx, world2ndc f64, ndc2world f32, (W, H), pix2tri)."""
    if kind == "reference":  # test_pix2depth.py:16-27 / test_differentiable_antialias.py
        tri, vtx = primitives.torus_zup(1.3, 0.4, 64, 32)
        c, s = np.cos(1.15), np.sin(1.15)
        vtx = (vtx.astype(np.float64) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).T
               + [0, 0.3, -4]).astype(np.float32)
        w2n = Mat44.camera_perspective_blender(1.0, 30.0, 2.0, 6.0, True,
                                               dtype=torch.float64).numpy()
        W = H_ = 128
    else:
        tri, vtx = primitives.octahedral_sphere(5, 0.8, noise_sigma=0.02, seed=4)
        w2n = np.diag([1.0, 1.0, 0.4, 1.0])
        W, H_ = 160, 120
    n2w = np.linalg.inv(w2n).astype(np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), W, H_, num_threads=8)
    return tri, vtx, w2n, n2w, (W, H_), pix


@pytest.mark.parametrize("kind", ["reference", "ortho"])
def test_pix2depth_bwd(gpu, oracle, kind):
    tri, vtx, w2n, n2w, (W, H_), pix = _scene(oracle, kind)
    rng = np.random.default_rng(31)
    dldw = rng.normal(size=pix.shape).astype(np.float32)
    got = torch.zeros((vtx.shape[0], 3), dtype=torch.float32, device=gpu)
    Pix2Depth.bwd_wrt_vtx2xyz(got, G(pix, gpu), G(tri, gpu), G(vtx, gpu), G(dldw, gpu), G(n2w, gpu))
    want = oracle.pix2depth_bwd(pix, tri, vtx, dldw, oracle.colmajor(n2w))
    scale = np.abs(want).max()
    assert scale > 0
    # per-vertex sums of a few pixels each, reordered by the atomics
    assert np.abs(H(got) - want).max() <= 1e-5 * scale
    no_device_errors()


def test_pix2depth_autograd(gpu, oracle):
    tri, vtx, w2n, n2w, (W, H_), pix = _scene(oracle, "reference")
    v = G(vtx, gpu).requires_grad_(True)
    depth = Pix2Depth.AutogradFunction.apply(v, G(pix, gpu), G(tri, gpu), G(n2w, gpu))
    assert bits_equal(H(depth), oracle.pix2depth(pix, tri, vtx, oracle.colmajor(n2w), W, H_))
    rng = np.random.default_rng(32)
    trgt = rng.random(pix.shape, dtype=np.float32)
    (depth * G(trgt, gpu)).sum().backward()
    want = oracle.pix2depth_bwd(pix, tri, vtx, trgt, oracle.colmajor(n2w))
    assert np.abs(H(v.grad) - want).max() <= 1e-5 * np.abs(want).max()
    no_device_errors()


def test_depth_then_antialias_autograd(gpu, oracle):
    """The chain of test_differentiable_antialias.py:239-299: depth image ->
    anti-aliased depth -> dot with a target.  The oracle side applies the
    chain rule by hand: the antialias backward for the vertices plus the
    depth backward fed with the pass-through image gradient."""
    tri, vtx, w2n, n2w, (W, H_), pix = _scene(oracle, "reference")
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * H_, 0, 0.5 * H_], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]])
    w2p = (n2p @ w2n).astype(np.float32)
    e2v, e2t = primitives.edge2vtx_edge2tri(tri)
    cedge = oracle.contour(tri, vtx, oracle.colmajor(w2n.astype(np.float32)), e2v, e2t)
    rng = np.random.default_rng(0)
    trgt = rng.random((H_, W), dtype=np.float32)

    v = G(vtx, gpu).requires_grad_(True)
    depth = Pix2Depth.AutogradFunction.apply(v, G(pix, gpu), G(tri, gpu), G(n2w, gpu))
    deptha = DifferentiableAntialias.DifferentiableAntialiasFunction.apply(
        G(cedge, gpu), v, G(w2p, gpu), G(pix, gpu), depth)
    loss = torch.dot(G(trgt, gpu).flatten(), deptha.flatten())
    loss.backward()

    d0 = oracle.pix2depth(pix, tri, vtx, oracle.colmajor(n2w), W, H_)
    da = oracle.antialias(cedge, vtx, oracle.colmajor(w2p), pix, d0)
    assert bits_equal(H(deptha), da)
    want = (oracle.bwd_antialias(cedge, vtx, oracle.colmajor(w2p), d0, trgt, pix)
            + oracle.pix2depth_bwd(pix, tri, vtx, trgt, oracle.colmajor(n2w)))
    scale = np.abs(want).max()
    assert np.abs(H(v.grad) - want).max() <= 2e-5 * scale
    no_device_errors()
