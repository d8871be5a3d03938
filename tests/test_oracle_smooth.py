"""Pins of the oracle's smoothed edge gradient (rasterized_edge_gradient.rs:
77-432): pixel-edge types against what the ray-cast already decided
(background next to a surface), the across-edge gradient against a float64
restatement of its sum, the red-black and serial smoothings against each
other at convergence (same fixed point) and the staggered-grid
interpolation against linear fields, which bilinear interpolation
reproduces."""

import numpy as np
import pytest

from del_msh_dlpack import primitives

U32MAX = 0xFFFFFFFF


def _scene(oracle):
    tri, vtx = primitives.sphere_yup(0.7, 24, 48)
    w2n = np.diag([1.0, 1.0, 0.5, 1.0])
    W, H = 48, 40
    n2p = np.array([[0.5 * W, 0, 0, 0.5 * W], [0, -0.5 * H, 0, 0.5 * H], [0, 0, 0.5, 0.5], [0, 0, 0, 1]])
    n2w = np.linalg.inv(w2n).astype(np.float32)
    w2p = (n2p @ w2n).astype(np.float32)
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    pix, _, _ = oracle.pix2tri(tri, vtx, bn, aabb, oracle.colmajor(n2w), W, H, num_threads=4)
    return tri, vtx, w2p, pix


def test_edge_types_and_gradient(oracle):
    tri, vtx, w2p, pix = _scene(oracle)
    rng = np.random.default_rng(2)
    occ = (pix != U32MAX).astype(np.float32)[..., None]
    val = np.concatenate([occ, rng.random(pix.shape + (1,), dtype=np.float32)], 2)
    dl = rng.standard_normal(pix.shape + (2,)).astype(np.float32)
    ht, hd, vt, vd = oracle.edge_gradient_and_type(tri, vtx, oracle.colmajor(w2p), pix, val, dl)
    assert np.array_equal(ht == 0, pix[:-1, :] == pix[1:, :])
    assert np.array_equal(vt == 0, pix[:, :-1] == pix[:, 1:])
    # a background pixel's centre lies in no projected triangle: its surface
    # neighbour is "in front" (type 3 when the background pixel comes first)
    bg = pix == U32MAX
    for t, a, b in ((ht, bg[:-1, :], bg[1:, :]), (vt, bg[:, :-1], bg[:, 1:])):
        first_bg = a & ~b
        second_bg = ~a & b
        assert first_bg.sum() > 20 and second_bg.sum() > 20
        assert np.mean(t[first_bg] == 3) > 0.99
        assert np.mean(t[second_bg] == 2) > 0.99
    assert np.any(ht == 1) and np.any(vt == 1)  # surface-surface edges inside the disc
    v64, g64 = val.astype(np.float64), dl.astype(np.float64)
    want_h = ((g64[:-1] + g64[1:]) * 0.5 * (v64[:-1] - v64[1:])).sum(2)
    want_v = ((g64[:, :-1] + g64[:, 1:]) * 0.5 * (v64[:, :-1] - v64[:, 1:])).sum(2)
    assert np.abs(hd - want_h).max() < 1e-5
    assert np.abs(vd - want_v).max() < 1e-5


def test_smoothings_share_a_fixed_point(oracle):
    tri, vtx, w2p, pix = _scene(oracle)
    occ = (pix != U32MAX).astype(np.float32)[..., None]
    # target: the silhouette shifted by 3 pixels, as the reference's test does with its own target
    trg = np.roll(occ, 3, axis=1)
    dl = (occ - trg).astype(np.float32)
    ht, hd, vt, vd = oracle.edge_gradient_and_type(tri, vtx, oracle.colmajor(w2p), pix, occ, dl)
    rb_h, rb_v = oracle.smooth_gradient(ht, vt, 4000, hd, vd, red_black=True)
    gs_h, gs_v = oracle.smooth_gradient(ht, vt, 4000, hd, vd, red_black=False)
    scale = max(np.abs(hd).max(), np.abs(vd).max())
    assert scale > 0
    assert np.abs(rb_h - gs_h).max() < 1e-4 * scale
    assert np.abs(rb_v - gs_v).max() < 1e-4 * scale
    fixed_h, fixed_v = (ht == 2) | (ht == 3), (vt == 2) | (vt == 3)
    assert fixed_h.any() and np.array_equal(rb_h[fixed_h], hd[fixed_h])
    assert np.array_equal(rb_v[fixed_v], vd[fixed_v])
    # the fixed point is discrete-harmonic: one more sweep changes nothing
    again_h, _ = oracle.smooth_gradient(ht, vt, 1, rb_h, rb_v, red_black=True)
    assert np.abs(again_h - rb_h).max() < 1e-5 * scale


def test_red_black_half_step_by_hand(oracle):
    H, W = 6, 7
    ht = np.zeros((H - 1, W), np.uint8)
    vt = np.zeros((H, W - 1), np.uint8)
    ht[2, 4] = 2  # a fixed neighbour that blocks the north link of (3, 4)
    rng = np.random.default_rng(9)
    hd = rng.random((H - 1, W)).astype(np.float32)
    vd = rng.random((H, W - 1)).astype(np.float32)
    got_h, _ = oracle.smooth_gradient(ht, vt, 1, hd, vd)
    # (ih, iw) = (3, 4) has colour 1: updated in the second half-step from
    # its colour-0 neighbours, which the first half-step already updated
    first, _ = oracle.smooth_gradient(ht, vt, 0, hd, vd)
    h1 = hd.copy()
    for ih in range(H - 1):
        for iw in range(W):
            if (ih + iw) % 2 == 0 and ht[ih, iw] not in (2, 3):
                nb = []
                if ih > 0 and ht[ih - 1, iw] != 2:
                    nb.append(hd[ih - 1, iw])
                if ih < H - 2 and ht[ih + 1, iw] != 3:
                    nb.append(hd[ih + 1, iw])
                if iw > 0:
                    nb.append(hd[ih, iw - 1])
                if iw < W - 1:
                    nb.append(hd[ih, iw + 1])
                s = np.float32(0)
                for x in nb:
                    s = np.float32(s + x)
                h1[ih, iw] = np.float32(s / np.float32(len(nb)))
    s = np.float32(0)
    for x in (h1[4, 4], h1[3, 3], h1[3, 5]):  # south, west, east (north is fixed type 2)
        s = np.float32(s + x)
    assert got_h[3, 4] == np.float32(s / np.float32(3))
    assert np.array_equal(first, hd)


@pytest.mark.parametrize("shape", [(9, 11), (2, 2), (5, 2)])
def test_staggered_interpolation_reproduces_linear_fields(oracle, shape):
    H, W = shape
    a, b, c = 0.3, -0.7, 1.25
    ih, iw = np.mgrid[0:H, 0:W - 1]
    vd = (a * (iw + 1.0) + b * (ih + 0.5) + c).astype(np.float32)  # vedges at (iw+1, ih+0.5)
    ih, iw = np.mgrid[0:H - 1, 0:W]
    hd = (a * (iw + 0.5) + b * (ih + 1.0) + c).astype(np.float32)  # hedges at (iw+0.5, ih+1)
    rng = np.random.default_rng(1)
    lo = np.array([1.0, 1.0])
    hi = np.array([max(W - 1.0, 1.0), max(H - 1.0, 1.0)])
    xy = (lo + rng.random((200, 2)) * (hi - lo)).astype(np.float32)
    out = oracle.interpolate_staggered_grid(hd, vd, xy)
    want = a * xy[:, 0] + b * xy[:, 1] + c
    inside_x = (xy[:, 0] <= W - 1) & (xy[:, 1] >= 0.5) & (xy[:, 1] <= H - 0.5)
    inside_y = (xy[:, 0] >= 0.5) & (xy[:, 0] <= W - 0.5) & (xy[:, 1] <= H - 1)
    if W > 2:
        assert np.abs(out[inside_x, 0] - want[inside_x]).max() < 1e-5
    if H > 2:
        assert np.abs(out[inside_y, 1] - want[inside_y]).max() < 1e-5
    assert np.all(np.isfinite(out))
