"""Batched build / ray-cast over a CSR batch of triangle meshes.

A batch is given by host-side offsets (numpy or CPU tensors, num_mesh + 1
entries, first entry 0): triangles [mesh2tri[m], mesh2tri[m+1]) of tri2vtx
with vertex indices local to vertices [mesh2vtx[m], mesh2vtx[m+1]) of vtx2xyz,
and for the ray-cast pixels [mesh2pix[m], mesh2pix[m+1]) of an image mesh2width[m]
wide.  Bulk arrays live on the GPU.  Mesh m's rows of bvhnodes / bvhnode2aabb
start at mesh2node[m] = sum_{j<m} (2 N_j - 1).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib


def _host_u32(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def node_offsets(mesh2tri) -> np.ndarray:
    n = np.diff(_host_u32(mesh2tri).astype(np.int64))
    return np.concatenate([[0], np.cumsum(np.where(n > 0, 2 * n - 1, 0))]).astype(np.uint32)


class Pipeline:
    """Build and ray-cast of one batch with its workspaces allocated once, so
    that build(), cast() and run() allocate nothing; capture() records run()
    into a HIP graph that replay() launches.

    packed=True (with images): build() writes every mesh's packed traversal
    buffer as it refits (dmsh_batch_build_pack_trimesh3) and cast() casts
    each mesh's image from it as one request of dmsh_cast_requests, so no
    cast packs the trees again; the images are the same, and bvhnode2aabb
    is then scratch (mesh_tree is not available)."""

    def __init__(self, mesh2tri, mesh2vtx, tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                 mesh2pix=None, mesh2width=None, ndc2world: torch.Tensor | None = None,
                 with_coeff: bool = True, packed: bool = False):
        dev = tri2vtx.device
        self.device = dev
        _lib.require_gpu(tri2vtx, vtx2xyz)
        self.mesh2tri_h = _host_u32(mesh2tri)
        self.mesh2vtx_h = _host_u32(mesh2vtx)
        self.num_mesh = len(self.mesh2tri_h) - 1
        if self.num_mesh < 1 or len(self.mesh2vtx_h) != self.num_mesh + 1:
            raise ValueError("mesh2tri / mesh2vtx need num_mesh + 1 >= 2 entries")
        self.total_tri = int(self.mesh2tri_h[-1])
        self.total_vtx = int(self.mesh2vtx_h[-1])
        if tri2vtx.shape != (self.total_tri, 3) or tri2vtx.dtype != torch.uint32:
            raise ValueError("tri2vtx must be (mesh2tri[-1], 3) uint32")
        if vtx2xyz.shape != (self.total_vtx, 3) or vtx2xyz.dtype != torch.float32:
            raise ValueError("vtx2xyz must be (mesh2vtx[-1], 3) float32")
        self.tri2vtx = tri2vtx.contiguous()
        self.vtx2xyz = vtx2xyz.contiguous()
        self.mesh2tri = torch.from_numpy(self.mesh2tri_h).to(dev)
        self.mesh2vtx = torch.from_numpy(self.mesh2vtx_h).to(dev)
        self.mesh2node_h = node_offsets(self.mesh2tri_h)
        self.num_rows = int(self.mesh2node_h[-1])
        self.mesh2node = torch.empty(self.num_mesh + 1, dtype=torch.uint32, device=dev)
        self.bvhnodes = torch.empty((max(self.num_rows, 1), 3), dtype=torch.uint32, device=dev)
        self.bvhnode2aabb = torch.empty((max(self.num_rows, 1), 6), dtype=torch.float32, device=dev)
        L = _lib.lib()
        self.ws_build = _lib.workspace(
            L.dmsh_batch_build_workspace_bytes(self.num_mesh, self.total_tri), dev)
        self.has_cast = mesh2pix is not None
        if self.has_cast:
            self.mesh2pix_h = _host_u32(mesh2pix)
            self.mesh2width_h = _host_u32(mesh2width)
            if len(self.mesh2pix_h) != self.num_mesh + 1 or len(self.mesh2width_h) != self.num_mesh:
                raise ValueError("mesh2pix needs num_mesh + 1 entries, mesh2width num_mesh")
            self.total_pix = int(self.mesh2pix_h[-1])
            self.mesh2pix = torch.from_numpy(self.mesh2pix_h).to(dev)
            self.mesh2width = torch.from_numpy(self.mesh2width_h).to(dev)
            if ndc2world is None or tuple(ndc2world.shape) != (self.num_mesh, 4, 4):
                raise ValueError("ndc2world must be (num_mesh, 4, 4)")
            # column-major per mesh, as the single-mesh entry points take it
            self.transforms = ndc2world.to(dev, torch.float32).transpose(1, 2).contiguous()
            self.pix2tri = torch.empty(max(self.total_pix, 1), dtype=torch.uint32, device=dev)
            self.pix2t = (torch.empty(max(self.total_pix, 1), dtype=torch.float32, device=dev)
                          if with_coeff else None)
            self.ws_cast = _lib.workspace(
                L.dmsh_batch_pix2tri_workspace_bytes(self.num_mesh, self.total_tri,
                                                     self.total_pix), dev)
        self.packed = bool(packed)
        if self.packed:
            if not self.has_cast:
                raise ValueError("packed=True needs the images")
            self._init_packed(L, dev)
        self.graph = None

    def _init_packed(self, L, dev):
        """One allocation of every mesh's packed buffer (256-byte aligned) and
        the request table that names them, written once."""
        from ..Serve.torch import REQ_DTYPE
        ntri = np.diff(self.mesh2tri_h.astype(np.int64))
        size = np.array([int(L.dmsh_packed_bytes(int(n))) for n in ntri], np.int64)
        off = np.concatenate([[0], np.cumsum((size + 255) // 256 * 256)])
        self.packed_buf = torch.empty(max(int(off[-1]), 1), dtype=torch.uint8, device=dev)
        addrs = self.packed_buf.data_ptr() + off[:-1]
        self.packed_addr = torch.from_numpy(addrs.astype(np.uint64)).to(dev)
        self.nonfinite = torch.empty(self.num_mesh, dtype=torch.uint32, device=dev)
        pix = np.diff(self.mesh2pix_h.astype(np.int64))
        w = self.mesh2width_h.astype(np.int64)
        if np.any(w == 0) or np.any(pix % np.maximum(w, 1)) or np.any(pix == 0):
            raise ValueError("packed=True needs a non-empty image of whole rows per mesh")
        recs = np.zeros(self.num_mesh, REQ_DTYPE)
        recs["packed"] = addrs.astype(np.uint64)
        recs["num_tri"] = ntri
        recs["img_w"] = w
        recs["img_h"] = pix // w
        recs["pix_start"] = self.mesh2pix_h[:-1]
        recs["ndc2world"] = self.transforms.reshape(self.num_mesh, 16).cpu().numpy()
        self.requests = torch.from_numpy(recs.view(np.uint8)).to(dev)
        self.ws_req = _lib.workspace(
            L.dmsh_cast_requests_workspace_bytes(self.num_mesh, self.total_pix), dev)

    def build(self):
        L = _lib.lib()
        if self.packed:
            _lib.check(L.dmsh_batch_build_pack_trimesh3(
                _lib.stream_of(self.device), self.num_mesh, self.total_tri, self.total_vtx,
                _lib.ptr(self.mesh2tri), _lib.ptr(self.mesh2vtx), _lib.ptr(self.tri2vtx),
                _lib.ptr(self.vtx2xyz), _lib.ptr(self.mesh2node), _lib.ptr(self.bvhnodes),
                _lib.ptr(self.bvhnode2aabb), _lib.ptr(self.packed_addr), _lib.ptr(self.nonfinite),
                _lib.ptr(self.ws_build), self.ws_build.numel()),
                "dmsh_batch_build_pack_trimesh3")
            return
        _lib.check(L.dmsh_batch_build_trimesh3(
            _lib.stream_of(self.device), self.num_mesh, self.total_tri, self.total_vtx,
            _lib.ptr(self.mesh2tri), _lib.ptr(self.mesh2vtx), _lib.ptr(self.tri2vtx),
            _lib.ptr(self.vtx2xyz), _lib.ptr(self.mesh2node), _lib.ptr(self.bvhnodes),
            _lib.ptr(self.bvhnode2aabb), _lib.ptr(self.ws_build), self.ws_build.numel()),
            "dmsh_batch_build_trimesh3")

    def cast(self):
        if not self.has_cast:
            raise ValueError("no images were given")
        L = _lib.lib()
        if self.packed:
            _lib.check(L.dmsh_cast_requests(
                _lib.stream_of(self.device), self.num_mesh, _lib.ptr(self.requests),
                self.total_pix, _lib.ptr(self.pix2tri),
                _lib.ptr(self.pix2t) if self.pix2t is not None else None,
                _lib.ptr(self.ws_req), self.ws_req.numel()), "dmsh_cast_requests")
            return
        _lib.check(L.dmsh_batch_pix2tri(
            _lib.stream_of(self.device), self.num_mesh, self.total_tri, self.total_vtx,
            self.total_pix,
            _lib.ptr(self.mesh2tri), _lib.ptr(self.mesh2vtx), _lib.ptr(self.tri2vtx),
            _lib.ptr(self.vtx2xyz), _lib.ptr(self.mesh2node), _lib.ptr(self.bvhnodes),
            _lib.ptr(self.bvhnode2aabb), _lib.ptr(self.mesh2pix), _lib.ptr(self.mesh2width),
            _lib.ptr(self.transforms), _lib.ptr(self.pix2tri),
            _lib.ptr(self.pix2t) if self.pix2t is not None else None,
            _lib.ptr(self.ws_cast), self.ws_cast.numel()),
            "dmsh_batch_pix2tri")

    def run(self):
        self.build()
        if self.has_cast:
            self.cast()

    def set_vertices(self, vtx2xyz: torch.Tensor):
        """New vertex positions (same shape) for the next build; a graph
        captured earlier keeps reading the old tensor."""
        if vtx2xyz.shape != self.vtx2xyz.shape or vtx2xyz.dtype != torch.float32:
            raise ValueError("vtx2xyz must keep its shape and dtype")
        _lib.require_gpu(vtx2xyz)
        self.vtx2xyz = vtx2xyz.detach().contiguous()

    def capture(self):
        """Record run() into a HIP graph (warm it up eagerly first)."""
        self.run()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    # views -------------------------------------------------------------
    def mesh_tree(self, m: int):
        if self.packed:
            raise ValueError("a packed pipeline keeps no boxes in bvhnode2aabb")
        a, b = int(self.mesh2node_h[m]), int(self.mesh2node_h[m + 1])
        return self.bvhnodes[a:b], self.bvhnode2aabb[a:b]

    def mesh_image(self, m: int):
        a, b = int(self.mesh2pix_h[m]), int(self.mesh2pix_h[m + 1])
        w = int(self.mesh2width_h[m])
        h = (b - a) // w if w else 0
        pix = self.pix2tri[a:b].view(h, w)
        t = self.pix2t[a:b].view(h, w) if self.pix2t is not None else None
        return pix, t


class PerMeshLoop:
    """The same batch through the single-mesh entry points, one mesh per
    call (dmsh_build_trimesh3 and dmsh_pix2tri, as a caller of the reference's
    TriMesh3.make_bvhnodes_bvhnode2aabb / Pix2Tri.by_raycasting would run it),
    into buffers of its own laid out like the pipeline's, so the two results
    compare byte for byte."""

    def __init__(self, p: Pipeline, morton_bits: int = 30):
        self.p = p
        self.morton_bits = morton_bits
        dev = p.device
        L = _lib.lib()
        self.ntri = np.diff(p.mesh2tri_h.astype(np.int64))
        nmax = int(self.ntri.max()) if self.num_mesh else 0
        self.bvhnodes = torch.empty_like(p.bvhnodes)
        self.bvhnode2aabb = torch.empty_like(p.bvhnode2aabb)
        self.ws_build = _lib.workspace(L.dmsh_build_workspace_bytes(max(nmax, 1), morton_bits), dev)
        if p.has_cast:
            wmax = int(p.mesh2width_h.max())
            hmax = max(int(p.mesh2pix_h[m + 1] - p.mesh2pix_h[m]) // max(int(p.mesh2width_h[m]), 1)
                       for m in range(self.num_mesh))
            self.pix2tri = torch.empty_like(p.pix2tri)
            self.pix2t = torch.empty_like(p.pix2t) if p.pix2t is not None else None
            self.ws_cast = _lib.workspace(
                L.dmsh_pix2tri_workspace_bytes(max(nmax, 1), max(wmax, 1), max(hmax, 1)), dev)

    @property
    def num_mesh(self) -> int:
        return self.p.num_mesh

    def build(self):
        p, L = self.p, _lib.lib()
        stream = _lib.stream_of(p.device)
        m2t, m2v, m2n = p.mesh2tri_h, p.mesh2vtx_h, p.mesh2node_h
        tri_ptr, xyz_ptr = p.tri2vtx.data_ptr(), p.vtx2xyz.data_ptr()
        for m in range(self.num_mesh):
            n = int(self.ntri[m])
            if n == 0:
                continue
            _lib.check(L.dmsh_build_trimesh3(
                stream, n, tri_ptr + int(m2t[m]) * 12, int(m2v[m + 1] - m2v[m]),
                xyz_ptr + int(m2v[m]) * 12, self.morton_bits,
                self.bvhnodes.data_ptr() + int(m2n[m]) * 12,
                self.bvhnode2aabb.data_ptr() + int(m2n[m]) * 24,
                self.ws_build.data_ptr(), self.ws_build.numel()), "dmsh_build_trimesh3")

    def cast(self):
        """Casts over this loop's own trees (build() first)."""
        p, L = self.p, _lib.lib()
        stream = _lib.stream_of(p.device)
        m2t, m2v, m2n, m2p = p.mesh2tri_h, p.mesh2vtx_h, p.mesh2node_h, p.mesh2pix_h
        tri_ptr, xyz_ptr = p.tri2vtx.data_ptr(), p.vtx2xyz.data_ptr()
        for m in range(self.num_mesh):
            w = int(p.mesh2width_h[m])
            if w == 0:
                continue
            _lib.check(L.dmsh_pix2tri(
                stream, self.pix2tri.data_ptr() + int(m2p[m]) * 4,
                self.pix2t.data_ptr() + int(m2p[m]) * 4 if self.pix2t is not None else None,
                int(self.ntri[m]), tri_ptr + int(m2t[m]) * 12, int(m2v[m + 1] - m2v[m]),
                xyz_ptr + int(m2v[m]) * 12, w, int(m2p[m + 1] - m2p[m]) // w,
                p.transforms.data_ptr() + m * 64, self.bvhnodes.data_ptr() + int(m2n[m]) * 12,
                self.bvhnode2aabb.data_ptr() + int(m2n[m]) * 24, self.ws_cast.data_ptr(),
                self.ws_cast.numel()), "dmsh_pix2tri")

    def equals_pipeline(self) -> bool:
        """Every tree row, box bit, hit index and hit coefficient bit equal."""
        p, r = self.p, self.p.num_rows
        same = (torch.equal(self.bvhnodes[:r], p.bvhnodes[:r])
                and torch.equal(self.bvhnode2aabb[:r].view(torch.int32),
                                p.bvhnode2aabb[:r].view(torch.int32)))
        if p.has_cast:
            same = same and torch.equal(self.pix2tri, p.pix2tri)
            if self.pix2t is not None:
                same = same and torch.equal(self.pix2t.view(torch.int32), p.pix2t.view(torch.int32))
        return bool(same)


def make_bvhnodes_bvhnode2aabb(mesh2tri, mesh2vtx, tri2vtx: torch.Tensor,
                               vtx2xyz: torch.Tensor):
    """(mesh2node [host u32, M+1], bvhnodes [R,3] u32, bvhnode2aabb [R,6] f32)."""
    p = Pipeline(mesh2tri, mesh2vtx, tri2vtx, vtx2xyz)
    p.build()
    return p.mesh2node_h, p.bvhnodes[:p.num_rows], p.bvhnode2aabb[:p.num_rows]


def by_raycasting(mesh2tri, mesh2vtx, tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                  mesh2pix, mesh2width, ndc2world: torch.Tensor, with_coeff: bool = True):
    """Build and cast: (pix2tri [P] u32, pix2t [P] f32 or None), flat over all
    images in mesh order."""
    p = Pipeline(mesh2tri, mesh2vtx, tri2vtx, vtx2xyz, mesh2pix, mesh2width, ndc2world,
                 with_coeff)
    p.run()
    return p.pix2tri[:p.total_pix], (p.pix2t[:p.total_pix] if p.pix2t is not None else None)


def _time_ms(fn, reps: int) -> float:
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps


def autotune(p: Pipeline, reps: int = 2, stacks=(16, 20, 24, 32), wgs_per_cu=(4, 8, 16, 32),
             cast_waves=(1, 2, 4), treelet_grids=(512, 1024, 2048), sort_items=(8, 12, 16),
             sort_seg=(1, 0)):
    """Times the launch parameters of the batched kernels on this batch and
    keeps the fastest (results do not depend on them).  The ray-cast is tuned
    over (LDS stack frames, ray-cast waves per CU), then over the workgroup
    size (waves per workgroup) at that pair; the build over (refit grid,
    elements per thread of the radix-sort digit passes, the sort: three
    segmented passes over the codes and the LDS finish, or one pass per 8-bit
    digit of the whole composite key).  Returns {"cast": {...}, "build":
    {...}, "trials": [...]}."""
    L = _lib.lib()
    trials = []

    def set_(key, value):
        _lib.check(L.dmsh_batch_set_tuning(key.encode(), value), "tuning")

    best_cast = None
    if p.has_cast:
        p.build()
        set_("batch.cast_waves", 1)
        for s in stacks:
            for w in wgs_per_cu:
                set_("batch.stack", s)
                set_("batch.wgs_per_cu", w)
                ms = _time_ms(p.cast, reps)
                trials.append({"stage": "cast", "batch.stack": s, "batch.wgs_per_cu": w,
                               "batch.cast_waves": 1, "ms": round(ms, 4)})
                if best_cast is None or ms < best_cast[0]:
                    best_cast = (ms, s, w, 1)
        set_("batch.stack", best_cast[1])
        set_("batch.wgs_per_cu", best_cast[2])
        for cw in cast_waves:
            if cw == 1:
                continue
            set_("batch.cast_waves", cw)
            ms = _time_ms(p.cast, reps)
            trials.append({"stage": "cast", "batch.stack": best_cast[1],
                           "batch.wgs_per_cu": best_cast[2], "batch.cast_waves": cw,
                           "ms": round(ms, 4)})
            if ms < best_cast[0]:
                best_cast = (ms, best_cast[1], best_cast[2], cw)
        set_("batch.cast_waves", best_cast[3])
    best_build = None
    for sg in sort_seg:
        for g in treelet_grids:
            for it in sort_items:
                set_("batch.sort_seg", sg)
                set_("batch.treelet_grid", g)
                set_("batch.sort_items", it)
                ms = _time_ms(p.build, reps)
                trials.append({"stage": "build", "batch.sort_seg": sg, "batch.treelet_grid": g,
                               "batch.sort_items": it, "ms": round(ms, 4)})
                if best_build is None or ms < best_build[0]:
                    best_build = (ms, g, it, sg)
    set_("batch.treelet_grid", best_build[1])
    set_("batch.sort_items", best_build[2])
    set_("batch.sort_seg", best_build[3])
    out = {"build": {"batch.treelet_grid": best_build[1], "batch.sort_items": best_build[2],
                     "batch.sort_seg": best_build[3], "ms": round(best_build[0], 4)},
           "trials": trials}
    if best_cast:
        out["cast"] = {"batch.stack": best_cast[1], "batch.wgs_per_cu": best_cast[2],
                       "batch.cast_waves": best_cast[3], "ms": round(best_cast[0], 4)}
    return out


# ---------------------------------------------------------------------------
# differentiable anti-aliased silhouettes of a batch
# ---------------------------------------------------------------------------
class SilhouetteBatch:
    """Anti-aliased silhouette images of a CSR mesh batch, one image per mesh
    (occupancy of the ray-cast, blended across the contour edges as
    DifferentiableAntialias does), and dL/d(vertex positions) of them.

    The "hit buffers" -- the batch's BVHs, pix2tri, the occupancy and the
    contour edges -- come from hits(): one batched build and one batched
    ray-cast for all meshes, then one contour extraction per mesh.  The mesh
    edges are built once on the GPU (TriMesh3.make_edge2vtx / make_edge2tri)
    and kept one direction per edge, so each silhouette edge is walked once
    and the gradient is that of the forward image.

    No reference counterpart (the reference renders and differentiates one
    mesh per call)."""

    def __init__(self, mesh2tri, mesh2vtx, tri2vtx: torch.Tensor, vtx2xyz: torch.Tensor,
                 mesh2pix, mesh2width, ndc2world: torch.Tensor):
        from ..Mat44.torch import from_transform_ndc2pix
        from ..TriMesh3.torch import make_edge2tri, make_edge2vtx

        self.pipe = Pipeline(mesh2tri, mesh2vtx, tri2vtx, vtx2xyz, mesh2pix, mesh2width, ndc2world,
                             with_coeff=False)
        p = self.pipe
        dev = p.device
        self.num_vtx = p.total_vtx
        n2w = ndc2world.detach().cpu().to(torch.float64)
        self.world2ndc, self.world2pix, self.edges, self.shapes = [], [], [], []
        for m in range(p.num_mesh):
            a, b = int(p.mesh2pix_h[m]), int(p.mesh2pix_h[m + 1])
            w = int(p.mesh2width_h[m])
            h = (b - a) // w if w else 0
            self.shapes.append((h, w))
            w2n = torch.linalg.inv(n2w[m])
            self.world2ndc.append(w2n.to(torch.float32).contiguous().to(dev))
            self.world2pix.append((from_transform_ndc2pix((w, h), dtype=torch.float64) @ w2n)
                                  .to(torch.float32).contiguous().to(dev))
            t0, t1 = int(p.mesh2tri_h[m]), int(p.mesh2tri_h[m + 1])
            v0, v1 = int(p.mesh2vtx_h[m]), int(p.mesh2vtx_h[m + 1])
            tri = p.tri2vtx[t0:t1]
            if t1 > t0:
                e2v = make_edge2vtx(tri, v1 - v0)
                e2t = make_edge2tri(tri, v1 - v0, e2v)
                keep = e2v[:, 0].to(torch.int64) < e2v[:, 1].to(torch.int64)
                rows = lambda a: a.view(torch.int32)[keep].view(torch.uint32).contiguous()
                self.edges.append((rows(e2v), rows(e2t)))
            else:
                z = torch.zeros((0, 2), dtype=torch.uint32, device=dev)
                self.edges.append((z, z))

    def _slices(self, m: int):
        p = self.pipe
        return ((int(p.mesh2tri_h[m]), int(p.mesh2tri_h[m + 1])),
                (int(p.mesh2vtx_h[m]), int(p.mesh2vtx_h[m + 1])),
                (int(p.mesh2pix_h[m]), int(p.mesh2pix_h[m + 1])))

    def hits(self, vtx2xyz: torch.Tensor) -> dict:
        """BVHs + pix2tri (batched), occupancy, contour edges per mesh.  The
        meshes' contour extractions are enqueued back to back into slices of
        one buffer, sharing one workspace, and their counts are read with one
        synchronization (the same kernels and results as one call per mesh)."""
        from .. import _lib
        from ..Edge2Vtx import contour_for_triangle_mesh_into

        p = self.pipe
        p.set_vertices(vtx2xyz)
        p.run()
        pix2tri = p.pix2tri[: p.total_pix].clone()
        occ = (pix2tri != 0xFFFFFFFF).to(torch.float32)
        v = p.vtx2xyz
        dev = pix2tri.device
        sizes = [int(self.edges[m][0].shape[0]) for m in range(p.num_mesh)]
        offs = [0]
        for n in sizes:
            offs.append(offs[-1] + n)
        cbuf = torch.empty((max(offs[-1], 1), 2), dtype=torch.uint32, device=dev)
        counts = torch.zeros((max(p.num_mesh, 1),), dtype=torch.uint32, device=dev)
        ws = _lib.workspace(_lib.lib().dmsh_contour_workspace_bytes(max(sizes + [1])), dev)
        live = []
        for m in range(p.num_mesh):
            (t0, t1), (v0, v1), _ = self._slices(m)
            e2v, e2t = self.edges[m]
            if t1 == t0 or e2v.shape[0] == 0:
                continue
            contour_for_triangle_mesh_into(p.tri2vtx[t0:t1], v[v0:v1],
                                           self.world2ndc[m].T.contiguous().flatten(), e2v, e2t,
                                           cbuf[offs[m]:offs[m + 1]], counts[m:m + 1], ws)
            live.append(m)
        counts_h = counts.view(torch.int32).cpu().tolist()
        cedges = []
        for m in range(p.num_mesh):
            if m in live:
                cedges.append(cbuf[offs[m]:offs[m] + counts_h[m]])
            else:
                cedges.append(self.edges[m][0])
        return {"pix2tri": pix2tri, "occ": occ, "cedges": cedges}

    @staticmethod
    def hit_bytes(hits: dict) -> int:
        return (hits["pix2tri"].numel() * 4 + hits["occ"].numel() * 4 +
                sum(c.numel() * 4 for c in hits["cedges"]))

    def images(self, vtx2xyz: torch.Tensor, hits: dict) -> torch.Tensor:
        from ..DifferentiableAntialias.torch import antialias

        p = self.pipe
        out = hits["occ"].clone()
        v = vtx2xyz.detach()
        for m in range(p.num_mesh):
            _, (v0, v1), (a, b) = self._slices(m)
            c = hits["cedges"][m]
            if c.shape[0] == 0 or b == a:
                continue
            h, w = self.shapes[m]
            out[a:b] = antialias(c, v[v0:v1], self.world2pix[m], hits["pix2tri"][a:b].view(h, w),
                                 hits["occ"][a:b].view(h, w)).view(-1)
        return out

    def vertex_gradient(self, vtx2xyz: torch.Tensor, hits: dict,
                        dldw: torch.Tensor) -> torch.Tensor:
        from ..DifferentiableAntialias.torch import bwd_antialias

        p = self.pipe
        v = vtx2xyz.detach()
        grad = torch.zeros_like(v)
        for m in range(p.num_mesh):
            _, (v0, v1), (a, b) = self._slices(m)
            c = hits["cedges"][m]
            if c.shape[0] == 0 or b == a:
                continue
            h, w = self.shapes[m]
            grad[v0:v1] = bwd_antialias(c, v[v0:v1], self.world2pix[m],
                                        hits["occ"][a:b].view(h, w), dldw[a:b].view(h, w),
                                        hits["pix2tri"][a:b].view(h, w))
        return grad


class AntialiasedSilhouette(torch.autograd.Function):
    """images = AntialiasedSilhouette.apply(vtx2xyz, batch, recompute):
    the concatenated silhouette images of a SilhouetteBatch.  With
    recompute=True nothing but the vertices is kept between forward and
    backward: the backward pass rebuilds the BVHs, re-casts pix2tri and
    re-extracts the contours (bit-identical, the build and the cast are
    deterministic) before the antialias backward."""

    @staticmethod
    def forward(ctx, vtx2xyz, batch: SilhouetteBatch, recompute: bool = True):
        hits = batch.hits(vtx2xyz)
        out = batch.images(vtx2xyz, hits)
        ctx.batch = batch
        ctx.hits = None if recompute else hits
        ctx.save_for_backward(vtx2xyz)
        return out

    @staticmethod
    def backward(ctx, dldw):
        (vtx2xyz,) = ctx.saved_tensors
        hits = ctx.hits if ctx.hits is not None else ctx.batch.hits(vtx2xyz)
        ctx.hits = None
        return ctx.batch.vertex_gradient(vtx2xyz, hits, dldw.contiguous()), None, None
