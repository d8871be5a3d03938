"""Pins of the oracle's attribute interpolation (pix2tri.rs:176-304):
closed forms of the forward (interpolating the positions gives the ray's
hit point, constants stay constant) and the backward against central
differences of an independent float64 forward that solves the line-plane
system with numpy.linalg.solve."""

import numpy as np
import pytest

from del_msh_dlpack import primitives

U32MAX = 0xFFFFFFFF


def _scene(oracle, persp):
    tri, vtx = primitives.torus_zup(0.7, 0.3, 24, 12)
    c, s = np.cos(0.9), np.sin(0.9)
    vtx = (vtx.astype(np.float64) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).T).astype(np.float32)
    if persp:
        f = 1.0 / np.tan(np.radians(22.5))
        n, fa = 1.0, 5.0
        proj = np.array([[f, 0, 0, 0], [0, f, 0, 0], [0, 0, (fa + n) / (fa - n), 2 * fa * n / (fa - n)],
                         [0, 0, -1, 0]])
        view = np.eye(4)
        view[2, 3] = -3.0
        w2n = proj @ view
    else:
        w2n = np.diag([1.0, 1.0, 0.5, 1.0])
    n2w = np.linalg.inv(w2n).astype(np.float32)
    W, H = 48, 40
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), W, H, num_threads=4)
    return tri, vtx, n2w, pix


def _rays(n2w, W, H):
    """float64 pixel rays: NDC (x, y, +1) -> (x, y, -1)."""
    ih, iw = np.mgrid[0:H, 0:W]
    x = 2.0 * (iw + 0.5) / W - 1.0
    y = 1.0 - 2.0 * (ih + 0.5) / H
    m = n2w.astype(np.float64)

    def tr(z):
        p = np.stack([x, y, np.full_like(x, z), np.ones_like(x)], -1) @ m.T
        return p[..., :3] / p[..., 3:]

    o = tr(1.0)
    return o, tr(-1.0) - o


def _forward64(pix, tri, vtx, val, n2w):
    H, W = pix.shape
    o, d = _rays(n2w, W, H)
    out = np.zeros((H, W, val.shape[1]))
    for ih in range(H):
        for iw in range(W):
            it = pix[ih, iw]
            if it == U32MAX:
                continue
            p = vtx[tri[it]].astype(np.float64)
            A = np.stack([p[1] - p[0], p[2] - p[0], -d[ih, iw]], 1)
            u, v, _ = np.linalg.solve(A, o[ih, iw] - p[0])
            out[ih, iw] = (1 - u - v) * val[tri[it][0]] + u * val[tri[it][1]] + v * val[tri[it][2]]
    return out


@pytest.mark.parametrize("persp", [False, True])
def test_interpolate_closed_forms(oracle, persp):
    tri, vtx, n2w, pix = _scene(oracle, persp)
    H, W = pix.shape
    out = oracle.interpolate(pix, tri, vtx, vtx, oracle.colmajor(n2w))
    o, d = _rays(n2w, W, H)
    cov = pix != U32MAX
    assert cov.sum() > 200
    # the interpolated position is on the pixel's ray
    r = out[cov].astype(np.float64) - o[cov]
    dn = d[cov] / np.linalg.norm(d[cov], axis=1, keepdims=True)
    off = r - (r * dn).sum(1, keepdims=True) * dn
    assert np.abs(off).max() < 1e-5
    assert np.all(out[~cov] == 0.0)
    const = np.full((vtx.shape[0], 2), 0.75, np.float32)
    c = oracle.interpolate(pix, tri, vtx, const, oracle.colmajor(n2w))
    assert np.abs(c[cov] - 0.75).max() < 1e-6


@pytest.mark.parametrize("persp", [False, True])
def test_interpolate_bwd_central_differences(oracle, persp):
    tri, vtx, n2w, pix = _scene(oracle, persp)
    rng = np.random.default_rng(11)
    val = rng.random((vtx.shape[0], 2)).astype(np.float32)
    w = rng.standard_normal(pix.shape + (2,)).astype(np.float32)
    gx, gv = oracle.interpolate_bwd(pix, tri, vtx, val, oracle.colmajor(n2w), w)
    used = np.unique(tri[pix[pix != U32MAX]].reshape(-1))
    sel = rng.choice(used, size=6, replace=False)
    h = 1e-5
    base_v = vtx.astype(np.float64)

    def loss(vv, va):
        return float((_forward64(pix, tri, vv, va, n2w) * w).sum())

    for iv in sel:
        for k in range(3):
            vp, vm = base_v.copy(), base_v.copy()
            vp[iv, k] += h
            vm[iv, k] -= h
            fd = (loss(vp, val) - loss(vm, val)) / (2 * h)
            assert abs(fd - gx[iv, k]) <= 2e-3 * max(abs(fd), 1.0), (iv, k, fd, gx[iv, k])
        for k in range(2):
            ap, am = val.astype(np.float64), val.astype(np.float64)
            ap[iv, k] += h
            am[iv, k] -= h
            fd = (loss(base_v, ap) - loss(base_v, am)) / (2 * h)
            assert abs(fd - gv[iv, k]) <= 1e-4 * max(abs(fd), 1.0), (iv, k, fd, gv[iv, k])
