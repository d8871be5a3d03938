"""Continuous-batching ray-cast server over device-resident trees.

A request is (mesh id, camera, viewport).  Trees are built the first time a
mesh id is seen and kept in HBM as packed traversal buffers in a byte-budget
LRU cache.  Each step admits the queued requests that fit (first come first
served, pinning their trees), builds the missing trees of the step with one
batched build and one batched pack, and ray-casts every admitted request with
one launch.  The next step is prepared and enqueued while the previous one
runs; completion times come from GPU events.  A request's image equals what
the single-mesh entry points give for that mesh and camera.
"""

from __future__ import annotations

import time
from collections import deque

import numpy as np
import torch

from .. import _lib, primitives
from .cache import ByteLRU, Request, admit

REQ_DTYPE = np.dtype([("packed", "<u8"), ("num_tri", "<u4"), ("nonfinite", "<u4"),
                      ("img_w", "<u4"), ("img_h", "<u4"), ("pix_start", "<u4"),
                      ("reserved", "<u4"), ("ndc2world", "<f4", (16,))])
assert REQ_DTYPE.itemsize == 96


class MeshCatalog:
    """Mesh id -> (tri2vtx u32 [N,3], vtx2xyz f32 [V,3]) on the host, fixed by
    the id: generator type id % 3 (sphere, torus, closed cylinder), triangle
    target log-uniform in tri_range and radial noise N(0, noise^2), both from
    a generator seeded with (seed, id)."""

    def __init__(self, tri_range=(1000, 1_000_000), seed: int = 7, noise: float = 0.01):
        self.tri_range = tri_range
        self.seed = seed
        self.noise = noise
        self._cache: dict[int, tuple] = {}

    def get(self, mid: int):
        hit = self._cache.get(mid)
        if hit is not None:
            return hit
        rng = np.random.default_rng([self.seed, mid])
        kind = mid % 3
        lo, hi = np.log(self.tri_range[0]), np.log(self.tri_range[1])
        a, b = primitives._divisions_for(kind, float(np.exp(rng.uniform(lo, hi))))
        if kind == 0:
            tri, vtx = primitives.sphere_yup(1.0, a, b)
        elif kind == 1:
            tri, vtx = primitives.torus_zup(0.7, 0.3, a, b)
        else:
            tri, vtx = primitives.cylinder_closed_end_yup(0.5, 1.4, a, b)
        v = vtx.astype(np.float64) * (1.0 + rng.normal(0.0, self.noise, size=(vtx.shape[0], 1)))
        out = (tri.astype(np.uint32), v.astype(np.float32))
        self._cache[mid] = out
        return out

    def pinned(self, mid: int):
        """The mesh as page-locked host tensors, so uploads are asynchronous."""
        key = ("pinned", mid)
        hit = self._cache.get(key)
        if hit is None:
            tri, vtx = self.get(mid)
            hit = (torch.from_numpy(tri).pin_memory(), torch.from_numpy(vtx).pin_memory())
            self._cache[key] = hit
        return hit

    def num_tri(self, mid: int) -> int:
        return int(self.get(mid)[0].shape[0])


def perspective_ndc2world(eye, fovy=45.0, near=1.0, far=5.0) -> np.ndarray:
    """Camera at eye looking at the origin (y up); near plane -> NDC z = +1."""
    eye = np.asarray(eye, np.float64)
    f = -eye / np.linalg.norm(eye)
    up = np.array([0.0, 1.0, 0.0]) if abs(f[1]) < 0.99 else np.array([1.0, 0.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    g = 1.0 / np.tan(np.radians(fovy) * 0.5)
    proj = np.array([[g, 0, 0, 0], [0, g, 0, 0],
                     [0, 0, -(far + near) / (near - far), -2 * far * near / (near - far)],
                     [0, 0, -1, 0]])
    return np.linalg.inv(proj @ view).astype(np.float32)


def request_stream(num_req: int, rate: float, num_ids: int = 10_000, zipf_s: float = 1.1,
                   side_range=(256, 1024), seed: int = 7):
    """Poisson arrivals at `rate` per second; mesh id Zipf(zipf_s) over
    num_ids ids; eye uniform on the sphere of radius 3; square viewport with a
    pixel count log-uniform in side_range^2."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=num_req))
    ranks = np.arange(1, num_ids + 1, dtype=np.float64)
    p = ranks ** -zipf_s
    p /= p.sum()
    ids = rng.choice(num_ids, size=num_req, p=p)
    reqs = []
    lo, hi = 2 * np.log(side_range[0]), 2 * np.log(side_range[1])
    for k in range(num_req):
        d = rng.normal(size=3)
        eye = 3.0 * d / np.linalg.norm(d)
        side = int(round(np.sqrt(np.exp(rng.uniform(lo, hi)))))
        reqs.append(Request(k, float(t[k]), int(ids[k]), side, side, perspective_ndc2world(eye)))
    return reqs


def _upload(a: np.ndarray, dev) -> torch.Tensor:
    """Host array -> device without a stream sync (through pinned memory)."""
    return torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)


class PinnedRing:
    """Page-locked staging slots for the per-step uploads (request tables,
    mesh offsets, buffer addresses), allocated once and reused: a slot is
    handed out again only after the copy that read it has run (an event
    recorded behind the copy on the stream)."""

    def __init__(self, slots: int = 8, nbytes: int = 1 << 16):
        self.bufs = [torch.empty(nbytes, dtype=torch.uint8).pin_memory() for _ in range(slots)]
        self.events = [None] * slots
        self.k = 0

    def upload(self, a: np.ndarray, dev, stream) -> torch.Tensor:
        raw = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
        k = self.k
        self.k = (k + 1) % len(self.bufs)
        if self.events[k] is not None:
            self.events[k].synchronize()
        if self.bufs[k].numel() < raw.size:
            self.bufs[k] = torch.empty(max(raw.size, 2 * self.bufs[k].numel()),
                                       dtype=torch.uint8).pin_memory()
        host = self.bufs[k][:raw.size]
        host.numpy()[:] = raw
        out = torch.empty(raw.size, dtype=torch.uint8, device=dev)
        with torch.cuda.stream(stream):
            out.copy_(host, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.events[k] = ev
        return out


class Server:
    def __init__(self, catalog: MeshCatalog, budget_bytes: int, device=None,
                 max_step_requests: int = 256, max_step_pixels: int = 1 << 27,
                 fused_pack: bool = True):
        self.catalog = catalog
        # True: one dmsh_batch_build_pack_trimesh3 per step (the refit writes
        # the packed buffers); False: dmsh_batch_build_trimesh3, then
        # dmsh_batch_pack (the build alone is then timed too)
        self.fused_pack = fused_pack
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.L = _lib.lib()
        self.cache = ByteLRU(budget_bytes)
        self.max_step_requests = max_step_requests
        self.max_step_pixels = max_step_pixels
        self.stream = torch.cuda.current_stream(self.device)
        # uploads of new meshes run here, beside the previous step's casts
        self.copy_stream = torch.cuda.Stream(device=self.device)
        self.ring = PinnedRing()
        # per step: host seconds spent preparing and enqueueing it; GPU build
        # (+ pack) events and the triangles they built
        self.host_s: list = []
        self.builds: list = []       # build + pack into the traversal layout
        self.builds_only: list = []  # the batched build alone (unfused pack only)
        self.step_gpu: list = []     # (event before the step's work, event after its casts)

    def tree_bytes(self, mid: int) -> int:
        return int(self.L.dmsh_packed_bytes(self.catalog.num_tri(mid)))

    # one step -------------------------------------------------------------
    def _build_trees(self, new_ids, planned):
        """Batched build + pack of the step's new meshes into cache entries."""
        if not new_ids:
            return []
        L, dev = self.L, self.device
        meshes = [self.catalog.pinned(mid) for mid in new_ids]
        ntri = np.array([t.shape[0] for t, _ in meshes], np.int64)
        nvtx = np.array([v.shape[0] for _, v in meshes], np.int64)
        m2t = np.concatenate([[0], np.cumsum(ntri)]).astype(np.uint32)
        m2v = np.concatenate([[0], np.cumsum(nvtx)]).astype(np.uint32)
        rows = np.concatenate([[0], np.cumsum(2 * ntri - 1)]).astype(np.uint32)
        # the step's meshes, concatenated on the device straight from pinned
        # memory on the copy stream, so the transfer overlaps the casts already
        # queued on the compute stream; the compute stream waits for it, and
        # record_stream keeps the staging memory from reuse until the build and
        # pack that read it have run
        with torch.cuda.stream(self.copy_stream):
            tri = torch.empty((int(m2t[-1]), 3), dtype=torch.uint32, device=dev)
            vtx = torch.empty((int(m2v[-1]), 3), dtype=torch.float32, device=dev)
            for k, (t, v) in enumerate(meshes):
                tri[int(m2t[k]):int(m2t[k + 1])].copy_(t, non_blocking=True)
                vtx[int(m2v[k]):int(m2v[k + 1])].copy_(v, non_blocking=True)
            d_m2t = self.ring.upload(m2t, dev, self.copy_stream).view(torch.uint32)
            d_m2v = self.ring.upload(m2v, dev, self.copy_stream).view(torch.uint32)
            uploaded = torch.cuda.Event()
            uploaded.record(self.copy_stream)
        self.stream.wait_event(uploaded)
        for t in (tri, vtx, d_m2t, d_m2v):
            t.record_stream(self.stream)
        m2n = torch.empty(len(new_ids) + 1, dtype=torch.uint32, device=dev)
        nodes = torch.empty((int(rows[-1]), 3), dtype=torch.uint32, device=dev)
        aabb = torch.empty((int(rows[-1]), 6), dtype=torch.float32, device=dev)
        M, T, V = len(new_ids), int(m2t[-1]), int(m2v[-1])
        ws = _lib.workspace(L.dmsh_batch_build_workspace_bytes(M, T), dev)
        sp = self.stream.cuda_stream
        bufs = [torch.empty(int(L.dmsh_packed_bytes(int(n))), dtype=torch.uint8, device=dev)
                for n in ntri]
        addrs = self.ring.upload(np.array([b.data_ptr() for b in bufs], np.uint64), dev,
                                 self.stream)
        nonfinite = torch.empty(M, dtype=torch.uint32, device=dev)
        b0 = torch.cuda.Event(enable_timing=True)
        b0.record(self.stream)
        if self.fused_pack:
            _lib.check(L.dmsh_batch_build_pack_trimesh3(
                sp, M, T, V, _lib.ptr(d_m2t), _lib.ptr(d_m2v), _lib.ptr(tri), _lib.ptr(vtx),
                _lib.ptr(m2n), _lib.ptr(nodes), _lib.ptr(aabb), _lib.ptr(addrs),
                _lib.ptr(nonfinite), _lib.ptr(ws), ws.numel()), "batch build + pack")
        else:
            _lib.check(L.dmsh_batch_build_trimesh3(sp, M, T, V, _lib.ptr(d_m2t), _lib.ptr(d_m2v),
                                                   _lib.ptr(tri), _lib.ptr(vtx), _lib.ptr(m2n),
                                                   _lib.ptr(nodes), _lib.ptr(aabb), _lib.ptr(ws),
                                                   ws.numel()), "batch build")
            bm = torch.cuda.Event(enable_timing=True)
            bm.record(self.stream)
            wsp = _lib.workspace(L.dmsh_batch_pack_workspace_bytes(M), dev)
            _lib.check(L.dmsh_batch_pack(sp, M, T, V, _lib.ptr(d_m2t), _lib.ptr(d_m2v),
                                         _lib.ptr(tri), _lib.ptr(vtx), _lib.ptr(m2n),
                                         _lib.ptr(nodes), _lib.ptr(aabb), _lib.ptr(addrs),
                                         _lib.ptr(nonfinite), _lib.ptr(wsp), wsp.numel()),
                       "batch pack")
            self.builds_only.append((b0, bm, T))
        b1 = torch.cuda.Event(enable_timing=True)
        b1.record(self.stream)
        self.builds.append((b0, b1, T))
        # the tree's own non-finite flag is in its buffer; staging tensors are
        # freed here and reused only by later work on the same stream
        for k, mid in enumerate(new_ids):
            self.cache.insert(mid, self.tree_bytes(mid), (bufs[k], int(ntri[k]), 0),
                              pins=planned[mid])
        return bufs

    def launch_step(self, batch):
        """Enqueue the casts of `batch` (admitted requests); returns the step
        record (outputs, event)."""
        L, dev = self.L, self.device
        recs = np.zeros(len(batch), REQ_DTYPE)
        vals = [self.cache.entries[r.mesh_id].value for r in batch]
        px = np.array([r.width * r.height for r in batch], np.int64)
        recs["packed"] = [v[0].data_ptr() for v in vals]
        recs["num_tri"] = [v[1] for v in vals]
        recs["nonfinite"] = [v[2] for v in vals]
        recs["img_w"] = [r.width for r in batch]
        recs["img_h"] = [r.height for r in batch]
        recs["pix_start"] = np.concatenate([[0], np.cumsum(px)[:-1]])
        recs["ndc2world"] = np.stack([np.asarray(r.ndc2world, np.float32).T.reshape(-1)
                                      for r in batch])
        start = int(px.sum())
        d_req = self.ring.upload(recs, dev, self.stream)
        pix = torch.empty(max(start, 1), dtype=torch.uint32, device=dev)
        pt = torch.empty(max(start, 1), dtype=torch.float32, device=dev)
        ws = _lib.workspace(L.dmsh_cast_requests_workspace_bytes(len(batch), start), dev)
        _lib.check(L.dmsh_cast_requests(self.stream.cuda_stream, len(batch), _lib.ptr(d_req), start,
                                        _lib.ptr(pix), _lib.ptr(pt), _lib.ptr(ws), ws.numel()),
                   "cast requests")
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        return {"batch": batch, "pix": pix, "t": pt, "event": ev, "keep": (d_req, ws)}

    def retire(self, step, t_done: float, keep_results: bool):
        off = 0
        for r in step["batch"]:
            px = r.width * r.height
            r.done = t_done
            if keep_results:
                r.result = (step["pix"][off:off + px].view(r.height, r.width),
                            step["t"][off:off + px].view(r.height, r.width))
            off += px
            self.cache.unpin(r.mesh_id)

    def run(self, requests, keep_results: bool = False, max_inflight: int = 3):
        """Serves `requests` (sorted by arrival, seconds from now) in real
        time; fills r.done (same clock) and, if asked, r.result.  Up to
        max_inflight steps are queued on the GPU while the host prepares the
        next (three kept the GPU busier than two at saturation, with lower
        p50 / p99 latency too)."""
        pending = deque(requests)
        queue: list = []
        inflight: deque = deque()
        steps = 0
        torch.cuda.synchronize(self.device)
        ref = torch.cuda.Event(enable_timing=True)
        ref.record(self.stream)
        ref.synchronize()
        t0 = time.perf_counter()

        def done_time(ev):
            return ref.elapsed_time(ev) * 1e-3  # seconds after ref (= t0)

        while pending or queue or inflight:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while inflight and inflight[0]["event"].query():
                s = inflight.popleft()
                self.retire(s, done_time(s["event"]), keep_results)
            if queue and len(inflight) < max_inflight:
                h0 = time.perf_counter()
                batch, new_ids, planned = admit(queue, self.cache, self.tree_bytes,
                                                self.max_step_requests, self.max_step_pixels)
                if batch:
                    s0 = torch.cuda.Event(enable_timing=True)
                    s0.record(self.stream)
                    self._build_trees(new_ids, planned)
                    step = self.launch_step(batch)
                    self.step_gpu.append((s0, step["event"]))
                    inflight.append(step)
                    self.host_s.append(time.perf_counter() - h0)
                    steps += 1
                    continue
                if not inflight:
                    if pending:
                        time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                        continue
                    raise RuntimeError("a queued request can never be admitted "
                                       "(its tree is larger than the cache budget)")
            if len(inflight) >= max_inflight:
                # the GPU has max_inflight steps queued: the next launch waits anyway
                s = inflight.popleft()
                s["event"].synchronize()
                self.retire(s, done_time(s["event"]), keep_results)
            elif inflight or pending:
                # poll for arrivals and completions
                time.sleep(0.0002)
        return steps


class PlainServer:
    """The per-request baseline: one request at a time with the single-mesh
    entry points (build on first sight, trees in a byte-budget LRU cache)."""

    def __init__(self, catalog: MeshCatalog, budget_bytes: int, device=None):
        self.catalog = catalog
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.L = _lib.lib()
        self.cache = ByteLRU(budget_bytes)
        self.stream = torch.cuda.current_stream(self.device)
        self.builds: list = []  # (start event, end event, triangles) per build

    def _tree(self, mid: int):
        if mid in self.cache:
            self.cache.hits += 1
            return self.cache.touch(mid).value
        self.cache.misses += 1
        tri, vtx = self.catalog.pinned(mid)
        n = tri.shape[0]
        nbytes = (2 * n - 1) * 36 + n * 12 + vtx.shape[0] * 12
        if not self.cache.reserve(nbytes):
            raise RuntimeError("tree larger than the cache budget")
        dev, L = self.device, self.L
        t_tri = tri.to(dev, non_blocking=True)
        t_vtx = vtx.to(dev, non_blocking=True)
        nodes = torch.empty((2 * n - 1, 3), dtype=torch.uint32, device=dev)
        aabb = torch.empty((2 * n - 1, 6), dtype=torch.float32, device=dev)
        ws = _lib.workspace(L.dmsh_build_workspace_bytes(n, 30), dev)
        b0 = torch.cuda.Event(enable_timing=True)
        b0.record(self.stream)
        _lib.check(L.dmsh_build_trimesh3(self.stream.cuda_stream, n, _lib.ptr(t_tri), vtx.shape[0],
                                         _lib.ptr(t_vtx), 30, _lib.ptr(nodes), _lib.ptr(aabb),
                                         _lib.ptr(ws), ws.numel()), "build")
        b1 = torch.cuda.Event(enable_timing=True)
        b1.record(self.stream)
        self.builds.append((b0, b1, n))
        val = (t_tri, t_vtx, nodes, aabb)
        self.cache.insert(mid, nbytes, val)
        return val

    def run(self, requests, keep_results: bool = False):
        torch.cuda.synchronize(self.device)
        ref = torch.cuda.Event(enable_timing=True)
        ref.record(self.stream)
        ref.synchronize()
        t0 = time.perf_counter()
        L = self.L
        events = []
        for r in requests:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            t_tri, t_vtx, nodes, aabb = self._tree(r.mesh_id)
            n = t_tri.shape[0]
            pix = torch.empty((r.height, r.width), dtype=torch.uint32, device=self.device)
            pt = torch.empty((r.height, r.width), dtype=torch.float32, device=self.device)
            T = _upload(np.asarray(r.ndc2world, np.float32).T, self.device)
            ws = _lib.workspace(L.dmsh_pix2tri_workspace_bytes(n, r.width, r.height), self.device)
            _lib.check(L.dmsh_pix2tri(self.stream.cuda_stream, _lib.ptr(pix), _lib.ptr(pt), n,
                                      _lib.ptr(t_tri), t_vtx.shape[0], _lib.ptr(t_vtx), r.width,
                                      r.height, _lib.ptr(T), _lib.ptr(nodes), _lib.ptr(aabb),
                                      _lib.ptr(ws), ws.numel()), "pix2tri")
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(self.stream)
            events.append(ev)
            if keep_results:
                r.result = (pix, pt)
        torch.cuda.synchronize(self.device)
        for r, ev in zip(requests, events):
            r.done = ref.elapsed_time(ev) * 1e-3


def build_stats(builds) -> dict:
    """Triangles built per second of GPU build time (events around each build,
    read after the run; the served figure includes the pack into the
    traversal layout, the plain one does not need it)."""
    if not builds:
        return {"builds": 0}
    ms = sum(a.elapsed_time(b) for a, b, _ in builds)
    tris = sum(t for _, _, t in builds)
    return {"builds": len(builds), "tris": int(tris), "gpu_ms": round(ms, 3),
            "mtri_per_s": round(tris / (ms * 1e-3) / 1e6, 1)}


def gpu_busy_s(step_gpu) -> float:
    """Seconds the stream spent on the steps: each step from the later of its
    start event and the previous step's end to its own end."""
    busy, prev = 0.0, None
    for a, b in step_gpu:
        start = a if prev is None or prev.elapsed_time(a) > 0 else prev
        busy += start.elapsed_time(b) * 1e-3
        prev = b
    return busy


def latency_stats(requests) -> dict:
    lat = np.array([r.done - r.arrival for r in requests]) * 1e3
    rays = sum(r.width * r.height for r in requests)
    span = max(r.done for r in requests) - min(r.arrival for r in requests)
    return {"p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3),
            "mrays_per_s": round(rays / span / 1e6, 1),
            "requests": len(requests), "rays": int(rays), "span_s": round(span, 3)}
