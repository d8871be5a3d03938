"""Host-side triangle-mesh generators (inputs for the hot path).

Topologies follow del-msh-cpu/src/trimesh3_primitive.rs; coordinates are
computed in float64 and rounded once to float32, so every consumer (device
kernels and the CPU oracle) receives identical bytes.
"""

from __future__ import annotations

import numpy as np


def cylinder_like_topology(ndiv_side: int, ndiv_circumference: int) -> np.ndarray:
    """trimesh3_primitive.rs:117-156: bottom fan, side quads, top fan."""
    nc = ndiv_circumference
    nlong = ndiv_side + 2
    ic = np.arange(nc)
    bottom = np.stack([np.zeros(nc, np.int64), ic + 1, (ic + 1) % nc + 1], axis=1)
    ih = np.arange(nlong - 2)[:, None]
    i1 = ih * nc + 1 + ic % nc
    i2 = ih * nc + 1 + (ic + 1) % nc
    i3 = (ih + 1) * nc + 1 + (ic + 1) % nc
    i4 = (ih + 1) * nc + 1 + ic % nc
    side = np.stack([i3, i2, i1, i4, i3, i1], axis=-1).reshape(-1, 3)
    i0 = np.full(nc, nc * (nlong - 1) + 1)
    j1 = (nlong - 2) * nc + 1 + (ic + 1) % nc
    j2 = (nlong - 2) * nc + 1 + ic % nc
    top = np.stack([i0, j1, j2], axis=1)
    return np.concatenate([bottom, side, top]).astype(np.uint32)


def sphere_yup(radius: float, n_longitude: int, n_latitude: int):
    """trimesh3_primitive.rs:377-420 (pole on +y).
    2 n_latitude (n_longitude - 1) triangles, 2 + (n_longitude - 1) n_latitude vertices."""
    if n_longitude <= 1 or n_latitude <= 2:
        return np.zeros((0, 3), np.uint32), np.zeros((0, 3), np.float32)
    dl = np.pi / n_longitude
    dr = 2.0 * np.pi / n_latitude
    pts = [[0.0, radius, 0.0]]
    ila = np.arange(1, n_longitude)[:, None]
    ilo = np.arange(n_latitude)[None, :]
    y0 = np.cos(dl * ila) * np.ones_like(ilo)
    r0 = np.sin(dl * ila)
    x0 = r0 * np.sin(dr * ilo)
    z0 = r0 * np.cos(dr * ilo)
    mid = radius * np.stack([x0, y0, z0], axis=-1).reshape(-1, 3)
    # rings 0 and n_longitude keep only their first vertex (ilo = 0)
    ys, rs = np.cos(dl * n_longitude), np.sin(dl * n_longitude)
    south = radius * np.array([[rs * np.sin(0.0), ys, rs * np.cos(0.0)]])
    vtx = np.concatenate([np.array(pts), mid, south]).astype(np.float32)
    tri = cylinder_like_topology(n_longitude - 2, n_latitude)
    return tri, vtx


def torus_zup(major_radius: float, minor_radius: float, ndiv_major: int, ndiv_minor: int):
    """trimesh3_primitive.rs:326-366; 2 ndiv_major ndiv_minor triangles."""
    rlg = 2.0 * np.pi / ndiv_major
    rlt = 2.0 * np.pi / ndiv_minor
    ilg = np.arange(ndiv_major)[:, None]
    ilt = np.arange(ndiv_minor)[None, :]
    lt = ilt * rlt
    lg = ilg * rlg
    r0 = major_radius + minor_radius * np.cos(lt)
    x = r0 * np.sin(lg)
    y = r0 * np.cos(lg)
    z = minor_radius * np.sin(lt) * np.ones_like(lg)
    vtx = np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)
    iug = (ilg + 1) % ndiv_major
    iut = (ilt + 1) % ndiv_minor
    a = ilg * ndiv_minor + ilt
    b = iug * ndiv_minor + ilt
    c = iug * ndiv_minor + iut
    d = ilg * ndiv_minor + iut * np.ones_like(ilg)
    tri = np.stack([a, c, b, a, d, c], axis=-1).reshape(-1, 3).astype(np.uint32)
    return tri, vtx


def cylinder_closed_end_yup(radius: float, length: float, ndiv_circumference: int,
                            ndiv_length: int, is_center: bool = True):
    """trimesh3_primitive.rs:159-203; 2 ndiv_circumference (ndiv_length + 1) triangles."""
    nc = ndiv_circumference
    dl = length / ndiv_length
    dr = 2.0 * np.pi / nc
    y_min = -0.5 * length if is_center else 0.0
    il = np.arange(ndiv_length + 1)[:, None]
    ilo = np.arange(nc)[None, :]
    x = radius * np.cos(dr * ilo) * np.ones_like(il)
    y = (y_min + dl * il) * np.ones_like(ilo)
    z = radius * np.sin(dr * ilo) * np.ones_like(il)
    ring = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    vtx = np.concatenate([[[0.0, y_min, 0.0]], ring, [[0.0, y_min + length, 0.0]]])
    tri = cylinder_like_topology(ndiv_length, nc)
    return tri, vtx.astype(np.float32)


def octahedral_sphere(level: int, radius: float = 1.0, noise_sigma: float = 0.0,
                      seed: int = 0):
    """Closed sphere from an octahedron subdivided `level` times:
    8 * 4^level triangles, 4 * 4^level + 2 vertices (level 9 gives exactly 2^21
    triangles, which no icosphere level does).  Built on the octahedral
    unfolding of the square [-M, M]^2, M = 2^level: the grid cells are split
    along the octahedron's edges and the folded square border is welded.
    Optional radial noise: r = radius * (1 + N(0, noise_sigma^2)) per vertex."""
    M = 1 << level
    g = np.arange(-M, M + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    ax, ay = np.abs(X), np.abs(Y)
    upper = ax + ay <= M
    Z = np.where(upper, M - ax - ay, -(ax + ay - M))
    Xf = np.where(upper, X, np.sign(X) * (M - ay))
    Yf = np.where(upper, Y, np.sign(Y) * (M - ax))
    key = np.stack([Xf, Yf, Z], axis=-1).reshape(-1, 3)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    n = 2 * M + 1
    gid = inv.reshape(n, n)
    i = np.arange(2 * M)[:, None]
    j = np.arange(2 * M)[None, :]
    v00 = gid[i, j]
    v10 = gid[i + 1, j]
    v01 = gid[i, j + 1]
    v11 = gid[i + 1, j + 1]
    cx = i + 0.5 - M
    cy = j + 0.5 - M
    anti = (cx * cy) > 0  # quadrants where the face edge runs along x + y = const
    t0 = np.where(anti[..., None], np.stack([v00, v10, v01], -1), np.stack([v00, v10, v11], -1))
    t1 = np.where(anti[..., None], np.stack([v10, v11, v01], -1), np.stack([v00, v11, v01], -1))
    tri = np.concatenate([t0.reshape(-1, 3), t1.reshape(-1, 3)]).astype(np.int64)
    p = uniq.astype(np.float64)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    if noise_sigma > 0.0:
        rng = np.random.default_rng(seed)
        p *= (1.0 + rng.normal(0.0, noise_sigma, size=(p.shape[0], 1)))
    p *= radius
    # orient every triangle outward
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    flip = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    return tri.astype(np.uint32), p.astype(np.float32)


def edge2vtx_edge2tri(tri2vtx: np.ndarray):
    """Undirected edges (i < j, sorted) of a closed triangle mesh and the two
    triangles sharing each edge (inputs of the contour extraction)."""
    t = tri2vtx.astype(np.int64)
    ntri = t.shape[0]
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    owner = np.concatenate([np.arange(ntri)] * 3)
    e.sort(axis=1)
    key = e[:, 0] * (int(t.max()) + 1) + e[:, 1]
    order = np.argsort(key, kind="stable")
    key_s = key[order]
    first = np.ones(len(key_s), bool)
    first[1:] = key_s[1:] != key_s[:-1]
    starts = np.nonzero(first)[0]
    counts = np.diff(np.append(starts, len(key_s)))
    if not np.all(counts == 2):
        raise ValueError("mesh is not a closed 2-manifold")
    edge2vtx = e[order[starts]].astype(np.uint32)
    edge2tri = np.stack([owner[order[starts]], owner[order[starts + 1]]], 1).astype(np.uint32)
    return edge2vtx, edge2tri


def _divisions_for(kind: int, target: float):
    """Generator divisions whose triangle count is near `target`:
    sphere 2 n_lat (n_long - 1) with n_lat = 2 (n_long - 1); torus 2 a b with
    a = 2 b; cylinder 2 c (l + 1) with c ~ sqrt(target)."""
    if kind == 0:
        k = max(1, int(round(np.sqrt(target / 4.0))))
        return (k + 1, 2 * k) if k >= 2 else (2, 3)
    if kind == 1:
        b = max(3, int(round(np.sqrt(target / 4.0))))
        return (2 * b, b)
    c = max(3, int(round(np.sqrt(target))))
    ln = max(1, int(round(target / (2.0 * c))) - 1)
    return (c, ln)


def _quaternion_matrix(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def batch_meshes(num_mesh: int, seed: int = 3, tri_range=(256, 262144),
                 pix_range=(64 * 64, 512 * 512)):
    """A batch of sphere / torus / closed-cylinder meshes, concatenated in CSR
    form, each with its own orthographic image.

    Per mesh (seeded): the generator type uniform over the three; a triangle
    target log-uniform in tri_range mapped to the nearest divisions; a uniform
    scale in [0.5, 2] and a rotation from a normalised 4D Gaussian quaternion
    (applied in f64, rounded once to f32); a square image whose pixel count is
    log-uniform in pix_range, looking down -z through a box 1.1x the mesh's
    bounding radius R (ndc2world = diag(1.1 R, 1.1 R, 2.2 R, 1)).

    Returns a dict: mesh2tri, mesh2vtx, mesh2pix (u32 [M+1]), tri2vtx (u32
    [T,3], vertex indices local to each mesh), vtx2xyz (f32 [V,3]),
    mesh2width (u32 [M]), ndc2world (f32 [M,4,4], row-major), kinds (u8 [M])."""
    rng = np.random.default_rng(seed)
    tris, vtxs, widths, cams, kinds = [], [], [], [], []
    lo, hi = np.log(tri_range[0]), np.log(tri_range[1])
    plo, phi = np.log(pix_range[0]), np.log(pix_range[1])
    for _ in range(num_mesh):
        kind = int(rng.integers(3))
        target = float(np.exp(rng.uniform(lo, hi)))
        scale = float(rng.uniform(0.5, 2.0))
        rot = _quaternion_matrix(rng.normal(size=4))
        side = int(round(np.sqrt(np.exp(rng.uniform(plo, phi)))))
        a, b = _divisions_for(kind, target)
        if kind == 0:
            tri, vtx = sphere_yup(1.0, a, b)
        elif kind == 1:
            tri, vtx = torus_zup(0.7, 0.3, a, b)
        else:
            tri, vtx = cylinder_closed_end_yup(0.5, 1.4, a, b)
        v = (vtx.astype(np.float64) @ rot.T) * scale
        r = 1.1 * float(np.sqrt((v * v).sum(axis=1)).max())
        tris.append(tri.astype(np.uint32))
        vtxs.append(v.astype(np.float32))
        widths.append(side)
        cams.append(np.diag([r, r, 2.0 * r, 1.0]).astype(np.float32))
        kinds.append(kind)
    ntri = np.array([t.shape[0] for t in tris], np.int64)
    nvtx = np.array([v.shape[0] for v in vtxs], np.int64)
    npix = np.array(widths, np.int64) ** 2
    csr = lambda c: np.concatenate([[0], np.cumsum(c)]).astype(np.uint32)
    return {
        "mesh2tri": csr(ntri),
        "mesh2vtx": csr(nvtx),
        "mesh2pix": csr(npix),
        "tri2vtx": np.concatenate(tris) if tris else np.zeros((0, 3), np.uint32),
        "vtx2xyz": np.concatenate(vtxs) if vtxs else np.zeros((0, 3), np.float32),
        "mesh2width": np.array(widths, np.uint32),
        "ndc2world": np.stack(cams) if cams else np.zeros((0, 4, 4), np.float32),
        "kinds": np.array(kinds, np.uint8),
    }


def mixed_hex_grid(nx: int, ny: int, nz: int, seed: int = 0, jitter: float = 0.0,
                   kinds=(0, 1, 2, 3)):
    """A mixed tet / pyramid / prism / hex mesh of the box [0, nx] x [0, ny] x
    [0, nz] in CSR form: each unit cube, with a kind drawn from `kinds`
    (seeded), becomes one hex (0), two prisms (1), six pyramids around the
    cube's centre (2) or six tets around its main diagonal (3).  With jitter
    > 0 the grid vertices move by U(-jitter, jitter) per axis (hexes then
    become trilinear, and cubes of different kinds no longer tile exactly).
    Node orders are those of the shape functions of the reference's
    polyhedron_mesh.cu.  Returns (elem2idx_offset u32, idx2vtx u32,
    vtx2xyz f32 [V, 3])."""
    rng = np.random.default_rng(seed)
    gx, gy, gz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    vtx = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.float64)
    if jitter > 0:
        vtx = vtx + rng.uniform(-jitter, jitter, vtx.shape)

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    extra = []
    elems = []
    cube_kind = rng.choice(np.asarray(kinds), size=(nx, ny, nz))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = {(a, b, d): vid(i + a, j + b, k + d) for a in (0, 1) for b in (0, 1)
                     for d in (0, 1)}
                kind = int(cube_kind[i, j, k])
                if kind == 0:
                    elems.append([c[0, 0, 0], c[1, 0, 0], c[1, 1, 0], c[0, 1, 0],
                                  c[0, 0, 1], c[1, 0, 1], c[1, 1, 1], c[0, 1, 1]])
                elif kind == 1:
                    elems.append([c[0, 0, 0], c[1, 0, 0], c[0, 1, 0], c[0, 0, 1], c[1, 0, 1],
                                  c[0, 1, 1]])
                    elems.append([c[1, 1, 0], c[0, 1, 0], c[1, 0, 0], c[1, 1, 1], c[0, 1, 1],
                                  c[1, 0, 1]])
                elif kind == 2:
                    ctr = len(vtx) + len(extra)
                    extra.append(np.mean([vtx[v] for v in c.values()], axis=0))
                    for face in ((c[0, 0, 0], c[1, 0, 0], c[1, 1, 0], c[0, 1, 0]),
                                 (c[0, 0, 1], c[1, 0, 1], c[1, 1, 1], c[0, 1, 1]),
                                 (c[0, 0, 0], c[1, 0, 0], c[1, 0, 1], c[0, 0, 1]),
                                 (c[0, 1, 0], c[1, 1, 0], c[1, 1, 1], c[0, 1, 1]),
                                 (c[0, 0, 0], c[0, 1, 0], c[0, 1, 1], c[0, 0, 1]),
                                 (c[1, 0, 0], c[1, 1, 0], c[1, 1, 1], c[1, 0, 1])):
                        elems.append(list(face) + [ctr])
                else:
                    a, h = c[0, 0, 0], c[1, 1, 1]
                    ring = (c[1, 0, 0], c[1, 1, 0], c[0, 1, 0], c[0, 1, 1], c[0, 0, 1], c[1, 0, 1])
                    for m in range(6):
                        elems.append([a, ring[m], ring[(m + 1) % 6], h])
    if extra:
        vtx = np.concatenate([vtx, np.array(extra)], 0)
    off = np.zeros(len(elems) + 1, np.uint32)
    off[1:] = np.cumsum([len(e) for e in elems])
    idx2vtx = np.concatenate([np.asarray(e, np.uint32) for e in elems])
    return off, idx2vtx, vtx.astype(np.float32)
