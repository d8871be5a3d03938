"""The ray-cast kernels search the BVH with one best hit and a stack of far
children (ties to the later hit in near-first order) instead of replaying
del-msh-cpu's recursion with per-subtree results (search_bvh3.rs:103-221).
These tests pin that the two searches give the same image, bit for bit, on
scenes built to produce exact ties: every triangle duplicated, rays through
grid vertices and edges, coplanar overlapping soups, and ordinary meshes
under orthographic and perspective cameras."""

import numpy as np
import pytest

from del_msh_dlpack import primitives


def _ortho(scale=1.0):
    return np.diag([scale, scale, scale, 1.0])


def _persp(eye=(0.3, 0.4, 3.0)):
    eye = np.asarray(eye, np.float64)
    f = -eye / np.linalg.norm(eye)
    up = np.array([0.0, 1.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    near, far, g = 1.0, 6.0, 1.0 / np.tan(np.radians(22.5))
    proj = np.array([[g, 0, 0, 0], [0, g, 0, 0],
                     [0, 0, -(far + near) / (near - far), -2 * far * near / (near - far)],
                     [0, 0, -1, 0]])
    return np.linalg.inv(proj @ view)


def _pixel_grid(n):
    """A z = 0 grid whose vertices sit exactly under the pixel centres of an
    n x n orthographic image, so rays pass through vertices and edges."""
    x = ((2.0 * (np.arange(n) + 0.5) / n) - 1.0).astype(np.float32)
    xx, yy = np.meshgrid(x, x)
    vtx = np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n, np.float32)], 1).astype(np.float32)
    tri = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i, (j + 1) * n + i + 1
            tri += [[a, b, d], [a, d, c]]
    return np.array(tri, np.uint32), vtx


def _scenes():
    tri, vtx = primitives.sphere_yup(0.8, 24, 48)
    yield "sphere_ortho", tri, vtx, _ortho(), 64
    yield "sphere_persp", tri, vtx, _persp(), 64
    yield "sphere_duplicated", np.concatenate([tri, tri]), vtx, _persp(), 64
    tri, vtx = primitives.torus_zup(0.6, 0.25, 40, 20)
    yield "torus_duplicated_reversed", np.concatenate([tri, tri[::-1]]), vtx, _ortho(), 48
    tri, vtx = _pixel_grid(33)
    yield "pixel_grid", tri, vtx, _ortho(), 33
    yield "pixel_grid_duplicated", np.concatenate([tri[::-1], tri]), vtx, _ortho(), 33
    rng = np.random.default_rng(5)
    soup = rng.uniform(-1, 1, (600, 3, 3)).astype(np.float32)
    soup[:, :, 2] = np.round(soup[:, :, 2] * 2) / 4  # five planes: coplanar overlaps
    vtx = soup.reshape(-1, 3)
    tri = np.arange(vtx.shape[0], dtype=np.uint32).reshape(-1, 3)
    yield "coplanar_soup", tri, vtx, _ortho(), 48
    tri, vtx = primitives.octahedral_sphere(4, noise_sigma=0.01, seed=3)
    yield "octa4_persp", tri, vtx, _persp((0.0, 0.0, 3.0)), 80


@pytest.mark.parametrize("scene", list(_scenes()), ids=lambda s: s[0])
def test_flat_search_equals_recursion(oracle, scene):
    name, tri, vtx, n2w, img = scene
    bn, aabb = oracle.build_trimesh3(tri, vtx)
    T = oracle.colmajor(np.asarray(n2w, np.float32))
    want, want_t, (nv_rec, nt_rec, _) = oracle.pix2tri(tri, vtx, bn, aabb, T, img, img)
    got, got_t, (nv_flat, nt_flat, _) = oracle.pix2tri_flat(tri, vtx, bn, aabb, T, img, img,
                                                            num_threads=4)
    assert (want != 0xFFFFFFFF).any()
    assert np.array_equal(got, want)
    assert np.array_equal(got_t.view(np.uint32), want_t.view(np.uint32))
    # the flat search prunes against the best hit so far, never less
    assert nv_flat <= nv_rec and nt_flat <= nt_rec


def test_duplicated_triangles_do_tie(oracle):
    """In the duplicated scenes every hit is an exact tie: the twin of the
    winning triangle gives the same t.  The two copies are adjacent leaves
    with equal boxes, so near-first order visits the second copy first and
    the recursion keeps the far (first) copy; a search that kept the earlier
    of two equal hits would return the second copy and fail the comparison
    above."""
    tri, vtx = primitives.sphere_yup(0.8, 24, 48)
    n = tri.shape[0]
    both = np.concatenate([tri, tri])
    bn, aabb = oracle.build_trimesh3(both, vtx)
    T = oracle.colmajor(np.asarray(_persp(), np.float32))
    pix, pt, _ = oracle.pix2tri(both, vtx, bn, aabb, T, 64, 64)
    ys, xs = np.nonzero(pix != 0xFFFFFFFF)
    assert ys.size > 100
    assert (pix[ys, xs] < n).all()
    for y, x in list(zip(ys, xs))[::37]:
        org, d = oracle.ray_for_pixel(int(x), int(y), 64, 64, T)
        twin = both[pix[y, x] + n]
        hit = oracle.ray_tri(vtx[twin[0]], vtx[twin[1]], vtx[twin[2]], org, d)
        assert hit is not None and np.float32(hit[0]) == pt[y, x]
