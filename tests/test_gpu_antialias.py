"""GPU parity for the contour extraction and the anti-aliased silhouette
(forward, backward, autograd wrapper) against the CPU oracle.

Contour sets and the anti-aliased forward image are bit-exact (the forward
keeps the serial last-writer order on the device); the backward sums with
float atomics, so it is compared with a stated tolerance.
"""

import numpy as np
import pytest
import torch

from del_msh_dlpack import _lib, primitives
import del_msh_dlpack.DifferentiableAntialias.torch as DifferentiableAntialias
import del_msh_dlpack.Edge2Vtx.torch as Edge2Vtx

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


def _ndc2pix(w, h):
    return np.array([[0.5 * w, 0, 0, 0.5 * w], [0, -0.5 * h, 0, 0.5 * h], [0, 0, 0.5, 0.5],
                     [0, 0, 0, 1]], np.float64)


def _perspective_w2n(eye, fovy=45.0, near=1.0, far=5.0):
    import del_msh_dlpack.Mat44.torch as Mat44

    w2c = Mat44.look_at(eye, (0, 0, 0), (0, 1, 0), dtype=torch.float64).numpy()
    proj = Mat44.perspective(fovy, 1.0, near, far, dtype=torch.float64).numpy()
    return proj @ w2c


def _scene(name):
    """(tri, vtx, world2ndc f32, (W, H)).  "torus_ref" is the scene of the
    reference's antialias test (test_differentiable_antialias.py:28-36)."""
    if name == "torus_ref":
        tri, vtx = primitives.torus_zup(1.3, 0.4, 64, 32)
        c, s = np.cos(1.15), np.sin(1.15)
        rot = np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]])
        tr = np.eye(4)
        tr[1, 3] = 0.6
        m = tr @ rot
        vtx = (vtx.astype(np.float64) @ m[:3, :3].T + m[:3, 3]).astype(np.float32)
        return tri, vtx, np.diag([0.5, 0.5, 0.5, 1.0]).astype(np.float32), (128, 128)
    if name == "sphere_persp":
        tri, vtx = primitives.sphere_yup(0.8, 32, 64)
        return tri, vtx, _perspective_w2n((0.3, 0.2, 3.0)).astype(np.float32), (256, 192)
    if name == "sphere_overflow":
        # the silhouette (radius 1.3 in NDC) leaves the image on every side
        tri, vtx = primitives.sphere_yup(1.0, 24, 48)
        return tri, vtx, np.diag([1.3, 1.3, 0.5, 1.0]).astype(np.float32), (96, 80)
    if name == "octa_noise":
        tri, vtx = primitives.octahedral_sphere(5, 0.7, noise_sigma=0.02, seed=9)
        return tri, vtx, _perspective_w2n((0.5, -0.4, 2.8)).astype(np.float32), (200, 160)
    raise KeyError(name)


SCENES = ["torus_ref", "sphere_persp", "sphere_overflow", "octa_noise"]


def _setup(oracle, name):
    tri, vtx, w2n, (W, H_) = _scene(name)
    n2w = np.linalg.inv(w2n.astype(np.float64)).astype(np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), W, H_, num_threads=8)
    e2v, e2t = primitives.edge2vtx_edge2tri(tri)
    cedge = oracle.contour(tri, vtx, oracle.colmajor(w2n), e2v, e2t)
    w2p = (_ndc2pix(W, H_) @ w2n.astype(np.float64)).astype(np.float32)
    return dict(tri=tri, vtx=vtx, w2n=w2n, w2p=w2p, pix=pix, e2v=e2v, e2t=e2t, cedge=cedge)


@pytest.mark.parametrize("name", SCENES)
def test_contour(gpu, oracle, name):
    s = _setup(oracle, name)
    got = Edge2Vtx.contour_for_triangle_mesh(G(s["tri"], gpu), G(s["vtx"], gpu), G(s["w2n"], gpu),
                                             G(s["e2v"], gpu), G(s["e2t"], gpu))
    assert got.dtype == torch.uint32 and got.shape[1] == 2
    assert len(s["cedge"]) > 0
    assert np.array_equal(H(got), s["cedge"])
    no_device_errors()


def test_contour_empty(gpu):
    tri, vtx = primitives.sphere_yup(1.0, 8, 12)
    e = np.zeros((0, 2), np.uint32)
    got = Edge2Vtx.contour_for_triangle_mesh(G(tri, gpu), G(vtx, gpu),
                                             G(np.eye(4, dtype=np.float32), gpu), G(e, gpu), G(e, gpu))
    assert tuple(got.shape) == (0, 2)
    no_device_errors()


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("image", ["occupancy", "random"])
def test_antialias_fwd(gpu, oracle, name, image):
    s = _setup(oracle, name)
    occ = (s["pix"] != U32MAX).astype(np.float32)
    if image == "random":
        rng = np.random.default_rng(21)
        vin = (occ * rng.random(occ.shape, dtype=np.float32)).astype(np.float32)
    else:
        vin = occ
    got = DifferentiableAntialias.antialias(G(s["cedge"], gpu), G(s["vtx"], gpu), G(s["w2p"], gpu),
                                            G(s["pix"], gpu), G(vin, gpu))
    want = oracle.antialias(s["cedge"], s["vtx"], oracle.colmajor(s["w2p"]), s["pix"], vin)
    assert not np.array_equal(want, vin)  # the silhouette was blended somewhere
    assert bits_equal(H(got), want)
    no_device_errors()


def _bwd_inputs(oracle, name):
    s = _setup(oracle, name)
    occ = (s["pix"] != U32MAX).astype(np.float32)
    rng = np.random.default_rng(22)
    dldw = rng.normal(size=occ.shape).astype(np.float32)
    return s, occ, dldw


@pytest.mark.parametrize("name", SCENES)
def test_antialias_bwd(gpu, oracle, name):
    s, occ, dldw = _bwd_inputs(oracle, name)
    got = H(DifferentiableAntialias.bwd_antialias(G(s["cedge"], gpu), G(s["vtx"], gpu),
                                                  G(s["w2p"], gpu), G(occ, gpu), G(dldw, gpu),
                                                  G(s["pix"], gpu)))
    want = oracle.bwd_antialias(s["cedge"], s["vtx"], oracle.colmajor(s["w2p"]), occ, dldw, s["pix"])
    # each vertex row is a sum over its contour edges' walks; the device adds
    # one partial sum per edge with float atomics, the oracle adds per pair
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(got - want).max() <= 2e-5 * scale
    no_device_errors()


def test_antialias_autograd(gpu, oracle):
    s, occ, _ = _bwd_inputs(oracle, "torus_ref")
    rng = np.random.default_rng(23)
    trgt = rng.random(occ.shape, dtype=np.float32)
    vtx = G(s["vtx"], gpu).requires_grad_(True)
    vin = G(occ, gpu).requires_grad_(True)
    out = DifferentiableAntialias.DifferentiableAntialiasFunction.apply(
        G(s["cedge"], gpu), vtx, G(s["w2p"], gpu), G(s["pix"], gpu), vin)
    loss = torch.dot(G(trgt, gpu).flatten(), out.flatten())
    loss.backward()
    want_out = oracle.antialias(s["cedge"], s["vtx"], oracle.colmajor(s["w2p"]), s["pix"], occ)
    assert bits_equal(H(out), want_out)
    want_g = oracle.bwd_antialias(s["cedge"], s["vtx"], oracle.colmajor(s["w2p"]), occ, trgt, s["pix"])
    scale = np.abs(want_g).max()
    assert np.abs(H(vtx.grad) - want_g).max() <= 2e-5 * scale
    assert bits_equal(H(vin.grad), trgt)
    no_device_errors()
