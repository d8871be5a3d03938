"""Pins for the oracle's contour extraction and anti-aliased silhouette
(forward and backward); runs without a GPU."""

import numpy as np
import pytest

from del_msh_dlpack import primitives


def test_intersection_edge2_closed_form(oracle):
    # c: (0,0)->(2,0); s: (1,-1)->(1,1) cross at rc = 0.5, re = 0.5
    assert oracle.intersection_edge2([0, 0], [2, 0], [1, -1], [1, 1]) == (0.5, 0.5)
    # touching at an end point counts, parallel and disjoint do not
    assert oracle.intersection_edge2([0, 0], [2, 0], [2, -1], [2, 1]) == (1.0, 0.5)
    assert oracle.intersection_edge2([0, 0], [2, 0], [0, 1], [2, 1]) is None
    assert oracle.intersection_edge2([0, 0], [2, 0], [3, -1], [3, 1]) is None


def test_dldw_intersection_edge2_finite_difference(oracle):
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(200):
        c0, c1 = rng.uniform(-1, 1, 2), rng.uniform(-1, 1, 2)
        s0, e1 = rng.uniform(-1, 1, 2), rng.uniform(-1, 1, 2)
        base = oracle.intersection_edge2(c0, c1, s0, e1)
        if base is None or not (0.05 < base[0] < 0.95 and 0.05 < base[1] < 0.95):
            continue
        gs, ge = oracle.dldw_intersection_edge2(c0, c1, s0, e1, 1.0)
        h = 1e-3
        for which, g in ((0, gs), (1, ge)):
            for k in range(2):
                pts = [np.array(s0, np.float64), np.array(e1, np.float64)]
                pts[which][k] += h
                rp = oracle.intersection_edge2(c0, c1, pts[0], pts[1])
                pts[which][k] -= 2 * h
                rm = oracle.intersection_edge2(c0, c1, pts[0], pts[1])
                if rp is None or rm is None:
                    continue
                fd = (rp[0] - rm[0]) / (2 * h)
                assert abs(fd - g[k]) < 2e-3 * (1.0 + abs(fd)), (fd, g[k])
                checked += 1
    assert checked > 100


def _edges(tri):
    return primitives.edge2vtx_edge2tri(tri)


def test_edge_topology_helper():
    tri, vtx = primitives.sphere_yup(1.0, 8, 12)
    e2v, e2t = _edges(tri)
    assert e2v.shape[0] == 3 * tri.shape[0] // 2  # closed mesh: E = 3F/2
    for (a, b), (t0, t1) in zip(e2v[:50], e2t[:50]):
        assert a < b and t0 != t1
        assert {int(a), int(b)} <= set(tri[t0].tolist()) and {int(a), int(b)} <= set(tri[t1].tolist())


def test_contour_is_the_silhouette_loop(oracle):
    """View along -z (orthographic): the contour of a y-up UV sphere is the
    ring of meridian edges at x = +-r (z ~ 0), forming closed loops."""
    tri, vtx = primitives.sphere_yup(1.0, 16, 32)
    e2v, e2t = _edges(tri)
    w2n = np.diag([1.0, 1.0, 0.5, 1.0]).astype(np.float32)
    c = oracle.contour(tri, vtx, oracle.colmajor(w2n), e2v, e2t)
    assert len(c) > 0
    mid = 0.5 * (vtx[c[:, 0]] + vtx[c[:, 1]])
    spacing = 2 * np.pi / 32
    assert np.all(np.abs(mid[:, 2]) < spacing)
    deg = np.bincount(c.reshape(-1), minlength=vtx.shape[0])
    assert np.all(deg % 2 == 0)
    # flipping the view direction gives the same contour
    c2 = oracle.contour(tri, vtx, oracle.colmajor(np.diag([1.0, 1.0, -0.5, 1.0]).astype(np.float32)), e2v, e2t)
    assert np.array_equal(c, c2)


def _aa_scene(oracle, img=64, n_long=24, n_lat=48, persp=False):
    tri, vtx = primitives.sphere_yup(0.6, n_long, n_lat)
    rot = np.array([[1, 0, 0], [0, np.cos(0.4), -np.sin(0.4)], [0, np.sin(0.4), np.cos(0.4)]])
    vtx = (vtx.astype(np.float64) @ rot.T + [0.05, -0.03, 0.0]).astype(np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    w2n = np.diag([1.0, 1.0, 0.5, 1.0]).astype(np.float32)
    n2w = np.linalg.inv(w2n).astype(np.float32)
    n2p = np.array([[0.5 * img, 0, 0, 0.5 * img], [0, -0.5 * img, 0, 0.5 * img], [0, 0, 0.5, 0.5],
                    [0, 0, 0, 1]], np.float32)
    w2p = (n2p @ w2n).astype(np.float32)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), img, img, num_threads=4)
    e2v, e2t = _edges(tri)
    cedge = oracle.contour(tri, vtx, oracle.colmajor(w2n), e2v, e2t)
    return tri, vtx, pix, w2p, cedge


def test_antialias_forward_properties(oracle):
    tri, vtx, pix, w2p, cedge = _aa_scene(oracle)
    occ = (pix != 0xFFFFFFFF).astype(np.float32)
    out = oracle.antialias(cedge, vtx, oracle.colmajor(w2p), pix, occ)
    changed = out != occ
    assert changed.any()
    assert np.all((out >= 0) & (out <= 1))
    # only pixels next to the silhouette change
    fg = pix != 0xFFFFFFFF
    edge = np.zeros_like(fg)
    edge[:-1] |= fg[:-1] != fg[1:]
    edge[1:] |= fg[:-1] != fg[1:]
    edge[:, :-1] |= fg[:, :-1] != fg[:, 1:]
    edge[:, 1:] |= fg[:, :-1] != fg[:, 1:]
    assert not np.any(changed & ~edge)
    # the blurred silhouette area stays close to the analytic disk area
    area = np.pi * (0.6 * 32) ** 2
    assert abs(out.sum() - area) < abs(occ.sum() - area) + 0.02 * area


def test_antialias_backward_finite_difference(oracle):
    """L = sum w * AA(occ) with pix2tri and the contour fixed: the analytic
    backward matches central differences (h = 1e-3) on the contour vertices."""
    tri, vtx, pix, w2p, cedge = _aa_scene(oracle)
    occ = (pix != 0xFFFFFFFF).astype(np.float32)
    rng = np.random.default_rng(3)
    w = rng.random(occ.shape).astype(np.float32)
    T = oracle.colmajor(w2p)
    g = oracle.bwd_antialias(cedge, vtx, T, occ, w, pix)

    def loss(v):
        return float((w.astype(np.float64) * oracle.antialias(cedge, v, T, pix, occ)).sum())

    verts = np.unique(cedge.reshape(-1))
    sel = rng.choice(verts, size=min(24, len(verts)), replace=False)
    h = 1e-3
    errs, scale = [], np.abs(g).max()
    for iv in sel:
        for k in range(2):  # x, y move the projected edge; z does not (orthographic)
            vp = vtx.copy()
            vp[iv, k] += h
            vm = vtx.copy()
            vm[iv, k] -= h
            fd = (loss(vp) - loss(vm)) / (2 * h)
            errs.append(abs(fd - g[iv, k]))
    errs = np.array(errs)
    # away from pixel-walk discontinuities the two agree; allow a few outliers
    assert np.median(errs) < 1e-3 * scale + 1e-3
    assert np.mean(errs < 0.05 * scale + 1e-2) > 0.8
    assert np.allclose(g[:, 2], 0.0, atol=1e-5 * scale)
