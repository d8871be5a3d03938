"""Pins for the oracle's depth backward; runs without a GPU.

The depth of a covered pixel is a smooth function of its triangle's three
vertices (ray-plane intersection followed by the projection), so central
differences of an independent float64 evaluation pin the analytic gradient."""

import numpy as np
import pytest
import torch

from del_msh_dlpack import primitives
import del_msh_dlpack.Mat44.torch as Mat44


def _ray64(n2w, iw, ih, w, h):
    x = 2.0 * (iw + 0.5) / w - 1.0
    y = 1.0 - 2.0 * (ih + 0.5) / h
    a = n2w @ np.array([x, y, 1.0, 1.0])
    b = n2w @ np.array([x, y, -1.0, 1.0])
    o = a[:3] / a[3]
    return o, b[:3] / b[3] - o


def _loss64(vtx, tri, pix, dldw, n2w, w2n):
    """sum over covered pixels of dldw * depth, depth from the ray-plane hit"""
    h, w = pix.shape
    total = 0.0
    for ip in np.nonzero(pix.reshape(-1) != 0xFFFFFFFF)[0]:
        ih, iw = divmod(int(ip), w)
        o, d = _ray64(n2w, iw, ih, w, h)
        p0, p1, p2 = (vtx[k] for k in tri[pix.reshape(-1)[ip]])
        n = np.cross(p1 - p0, p2 - p0)
        t = n @ (p0 - o) / (n @ d)
        q = w2n @ np.append(o + t * d, 1.0)
        total += dldw.reshape(-1)[ip] * (q[2] / q[3] + 1.0) * 0.5
    return total


def _scene(oracle, kind):
    if kind == "reference":  # test_pix2depth.py:16-27
        tri, vtx = primitives.torus_zup(1.3, 0.4, 32, 16)
        c, s = np.cos(1.15), np.sin(1.15)
        vtx = vtx.astype(np.float64) @ np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).T + [0, 0.3, -4]
        w2n = Mat44.camera_perspective_blender(1.0, 30.0, 2.0, 6.0, True, dtype=torch.float64).numpy()
        W = H = 48
    else:
        tri, vtx = primitives.sphere_yup(0.8, 12, 24)
        vtx = vtx.astype(np.float64)
        w2n = np.diag([1.0, 1.0, 0.4, 1.0])
        W, H = 40, 32
    vtx = vtx.astype(np.float32)
    n2w = np.linalg.inv(w2n)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w.astype(np.float32)), W, H)
    return tri, vtx, pix, n2w, w2n


@pytest.mark.parametrize("kind", ["reference", "ortho"])
def test_depth_bwd_matches_central_differences(oracle, kind):
    tri, vtx, pix, n2w, w2n = _scene(oracle, kind)
    assert (pix != 0xFFFFFFFF).sum() > 100
    rng = np.random.default_rng(7)
    dldw = rng.normal(size=pix.shape).astype(np.float32)
    g = oracle.pix2depth_bwd(pix, tri, vtx, dldw, oracle.colmajor(n2w.astype(np.float32)))
    used = np.unique(tri[pix[pix != 0xFFFFFFFF]].reshape(-1))
    sel = rng.choice(used, size=min(12, len(used)), replace=False)
    v64 = vtx.astype(np.float64)
    h = 1e-5
    scale = np.abs(g).max()
    for iv in sel:
        for k in range(3):
            vp, vm = v64.copy(), v64.copy()
            vp[iv, k] += h
            vm[iv, k] -= h
            fd = (_loss64(vp, tri, pix, dldw, n2w, w2n) - _loss64(vm, tri, pix, dldw, n2w, w2n)) / (2 * h)
            assert abs(fd - g[iv, k]) < 2e-3 * scale + 1e-4 * abs(fd), (iv, k, fd, g[iv, k])


def test_depth_bwd_translation_closed_form(oracle):
    """Moving the whole mesh by d_z along an orthographic view axis z moves
    every hit point by exactly d_z, so every covered pixel's depth moves by
    s_z d_z / 2 (s_z = world2ndc[2][2]): sum_v dL/dz_v = s_z / 2 sum_pix g."""
    tri, vtx, pix, n2w, w2n = _scene(oracle, "ortho")
    rng = np.random.default_rng(8)
    dldw = rng.normal(size=pix.shape).astype(np.float32)
    g = oracle.pix2depth_bwd(pix, tri, vtx, dldw, oracle.colmajor(n2w.astype(np.float32)))
    covered = pix != 0xFFFFFFFF
    want_z = float(dldw[covered].astype(np.float64).sum()) * w2n[2, 2] / 2.0
    assert abs(g[:, 2].sum() - want_z) < 1e-5 * np.abs(dldw).sum()
