"""Benchmark of the del-msh LBVH / ray-cast hot path on one MI355X.

Workload (single large mesh, plain kernels): an octahedron-subdivided sphere
with exactly 2^21 triangles (level 9), vertices displaced radially by seeded
N(0, 0.01^2) noise; LBVH build with 63-bit Morton codes (centroid, unit-cube
fit, Morton, stable radix argsort, Karras topology, refit); perspective
ray-cast of 2048 x 2048 rays (eye (0,0,3), target origin, fovy 45, near 1,
far 5; the near plane maps to NDC z = +1 where pixel rays start) producing
the hit triangle and hit coefficient per pixel.

One step = one build + one ray-cast.  Inputs are resident in HBM before the
timed region.  `value` is the build throughput in Mtri/s; the ray-cast
throughput and the separate 2^22-point build / sort-alone figures are
reported beside it.  Per-kernel launch times come from the library's HIP
event timing on the launch stream.  Prints one JSON line.

Other modes: --pmc-workload (the short run a rocprofv3 counter pass
profiles) and --summarize-pmc (turns two counter passes into the traffic
summary under profiles/, which the default run reads back).
"""

from __future__ import annotations

import argparse
import csv
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "del-msh_amd"))

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--img", type=int, default=2048)
    ap.add_argument("--level", type=int, default=9)
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run only what a counter pass needs: 3 builds, 3 ray-casts and 3 "
                         "standalone refits of the bench workload and 3 device-to-device copies "
                         "of 512 MiB (the calibration kernel)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="DIR",
                    help="turn rocprofv3 --pmc output folders (one per counter pass) into the "
                         "per-kernel counter summary text")
    ap.add_argument("--batch-meshes", type=int, default=4096,
                    help="meshes of the batched workload (0 skips it)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests per arrival rate of the served workload (0 skips it)")
    ap.add_argument("--serve-rates", default="1000,4000,10000,20000",
                    help="arrival rates, requests/s")
    ap.add_argument("--serve-budget-mib", type=int, default=4096)
    ap.add_argument("--diff-reps", type=int, default=5,
                    help="repetitions of the differentiation workload (0 skips it)")
    ap.add_argument("--tree-reps", type=int, default=5,
                    help="repetitions of the octree and point-location workloads (0 skips them)")
    return ap.parse_args()


# kernels whose launch time the bench reads back from the library's event timing
TIMED_KERNELS = ["k_pix2tri", "k_pix2tri_tier2", "k_pix2tri_tier3", "k_pack_tree",
                 "k_build_init", "k_centroid_aabb", "k_morton_hist", "k_hist_all", "k_onesweep",
                 "k_sort_finish", "k_sort_groups", "k_topology",
                 "k_refit_treelet", "k_refit_top", "k_fill_u32"]
COUNTER_SUMMARY = ROOT / "profiles" / "pmc_counters_octa9_2048.txt"
CALIB_BYTES = 512 << 20
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def perspective_ndc2world(eye=(0.0, 0.0, 3.0), fovy=45.0, near=1.0, far=5.0):
    import numpy as np

    eye = np.asarray(eye, np.float64)
    f = -eye / np.linalg.norm(eye)
    up = np.array([0.0, 1.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    g = 1.0 / math.tan(math.radians(fovy) * 0.5)
    # near plane -> NDC z = +1 (where pixel rays start), far plane -> -1
    proj = np.array([[g, 0, 0, 0], [0, g, 0, 0],
                     [0, 0, -(far + near) / (near - far), -2 * far * near / (near - far)],
                     [0, 0, -1, 0]])
    return np.linalg.inv(proj @ view).astype(np.float32)


# ---------------------------------------------------------------------------
# counter passes
# ---------------------------------------------------------------------------
# Each pass is its own rocprofv3 run (the hardware collects one counter set
# per pass); the summary joins them per kernel and per phase of the counter
# workload (build, cast, refit: a kernel such as k_refit_treelet runs in the
# build and in the standalone refit with different inputs).  Bytes at the
# L2-to-fabric interface come from the size-split request counts summed over
# every TCC instance (all channels of all XCDs): reads 32 / 64 / 128 B
# requests, writes 64 B (WRREQ_64B) or 32 B (the rest); the bytes that go on
# to DRAM from the 32-byte sector counts of DRAM-bound requests.  The 512 MiB
# device copy checks both against bytes known to move.
COUNTER_PASSES = {
    "reads": ["TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum",
              "TCC_EA0_RDREQ_DRAM_32B_sum"],
    "writes": ["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum",
               "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum"],
    "hits": ["TCC_HIT_sum", "TCC_MISS_sum", "TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_DRAM_sum"],
    "issue": ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM", "SQ_INSTS_SALU",
              "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "SQ_BUSY_CU_CYCLES", "GRBM_GUI_ACTIVE"],
    "waits": ["SQ_WAIT_INST_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_WAIT_ANY", "SQ_ACTIVE_INST_LDS",
              "SQ_WAIT_INST_ANY", "SQ_ACTIVE_INST_ANY", "SQ_INSTS_SMEM", "SQ_BUSY_CYCLES"],
}
# (phase, kernel-name fragment) rows of the summary
PROFILED_KERNELS = [
    ("cast", "k_pix2tri<"), ("cast", "k_pack_tree"),
    ("build", "k_centroid_aabb<true>"), ("build", "k_morton_hist"), ("build", "k_onesweep"),
    ("build", "k_sort_finish"), ("build", "k_topology"), ("build", "k_refit_treelet<1,"),
    ("build", "k_refit_top_lds"),
    ("refit", "k_plan_hash"), ("refit", "k_leaf_ranges"), ("refit", "k_list_treelets"),
    ("refit", "k_refit_treelet<1,"),
    ("refit", "k_refit_top_lds"),
]
NUM_CU = 256
NUM_XCD = 8  # GRBM_GUI_ACTIVE comes back summed over the XCDs


def _read_counters(folders):
    """{(phase, kernel name): {counter: [value per dispatch]}} and the
    dispatch durations (ns), from rocprofv3 counter_collection CSV files.  The
    counter workload runs 3 builds, 3 casts, 3 refits and 3 copies in that
    order: per file, dispatches before the first k_pix2tri are the build's,
    those after the last one the refit's (copies keep their own name)."""
    per, dur = {}, {}
    for folder in folders:
        files = sorted(pathlib.Path(folder).rglob("*counter_collection.csv"))
        if not files:
            raise SystemExit(f"no counter_collection.csv under {folder}")
        for f in files:
            with open(f, newline="") as fh:
                rows = list(csv.DictReader(fh))
            cast_ids = [int(r["Dispatch_Id"]) for r in rows if "k_pix2tri<" in r["Kernel_Name"]]
            lo, hi = (min(cast_ids), max(cast_ids)) if cast_ids else (0, 0)
            for r in rows:
                d = int(r["Dispatch_Id"])
                phase = "build" if d < lo else ("refit" if d > hi else "cast")
                name = r["Kernel_Name"]
                if COPY_KERNEL in name:
                    phase = f"copy{r.get('Grid_Size', '')}"
                key = (phase, name)
                per.setdefault(key, {}).setdefault(r["Counter_Name"], []).append(
                    float(r["Counter_Value"]))
                dur.setdefault(key, {})[(folder, d)] = (float(r["End_Timestamp"])
                                                        - float(r["Start_Timestamp"]))
    return per, {k: list(v.values()) for k, v in dur.items()}


def _avg(vals):
    return sum(vals) / len(vals) if vals else float("nan")


def summarize_counters(folders, title: str) -> str:
    per, dur = _read_counters(folders)

    def pick(phase, short):
        for key in per:
            if key[0] == phase and short in key[1]:
                return key
        return None

    def val(key, c):
        return _avg(per.get(key, {}).get(c, []))

    def l2_read(key):
        return (32 * val(key, "TCC_EA0_RDREQ_32B_sum") + 64 * val(key, "TCC_EA0_RDREQ_64B_sum")
                + 128 * val(key, "TCC_EA0_RDREQ_128B_sum"))

    def l2_write(key):
        wr, w64 = val(key, "TCC_EA0_WRREQ_sum"), val(key, "TCC_EA0_WRREQ_64B_sum")
        return (wr - w64) * 32 + w64 * 64

    def dram_read(key):
        return 32 * val(key, "TCC_EA0_RDREQ_DRAM_32B_sum")

    def dram_write(key):
        return 32 * (val(key, "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")
                     + val(key, "TCC_EA0_WRREQ_ATOMIC_DRAM_32B_sum"))

    copies = sorted((k for k in per if k[0].startswith("copy")),
                    key=lambda k: -int(k[0][4:] or 0))
    lines = [title, "Commands (MI355X, ROCm 7.2), one rocprofv3 pass per counter set:",
             "  cd /tmp && export TMPDIR=/tmp"]
    for key, cs in COUNTER_PASSES.items():
        lines.append(f"  timeout -s KILL 150 rocprofv3 --pmc {' '.join(cs)} --output-format csv "
                     f"-d <{key}> -- python3 bench.py --pmc-workload")
    lines.append("  python3 bench.py --summarize-counters <reads> <writes> <hits> <issue> <waits>")
    if copies:
        c = copies[0]
        lines.append(f"check: {COPY_KERNEL} of {CALIB_BYTES} B (read once, written once): L2<->fabric "
                     f"read {l2_read(c) / CALIB_BYTES:.3f} x, write {l2_write(c) / CALIB_BYTES:.3f} x; "
                     f"to / from DRAM read {dram_read(c) / CALIB_BYTES:.3f} x, write "
                     f"{dram_write(c) / CALIB_BYTES:.3f} x")
    diffs = [abs(l2_read(k) + l2_write(k) - dram_read(k) - dram_write(k))
             / max(l2_read(k) + l2_write(k), 1.0)
             for k in per if not k[0].startswith("copy") and per[k].get("TCC_EA0_RDREQ_32B_sum")]
    lines.append("the DRAM-bound request counts (RDREQ_DRAM_32B, WRREQ_*_DRAM_32B) match the "
                 f"L2-to-fabric totals for every kernel here (largest relative difference "
                 f"{max(diffs) if diffs else float('nan'):.1e}): every L2 miss is a memory request, "
                 "and whether the MALL (infinity cache) or HBM serves it is not visible to the L2 "
                 "counters, so the miss bytes below bound the HBM bytes from above")
    lines.append("per dispatch, averaged over its dispatches in that phase: us (kernel time in the "
                 "counter passes); MB read / written by L2 misses (to MALL / HBM); L2 hit %; VALU "
                 "busy % (SQ_ACTIVE_INST_VALU / (256 CUs x GRBM_GUI_ACTIVE / 8 XCDs)); VALU / LDS / "
                 "VMEM instructions per wave; LDS-wait share of wave cycles %; LDS bank-conflict "
                 "cycles per LDS instruction; waves")
    lines.append(f"{'phase':<6}{'kernel':<24}{'us':>8}{'missRdMB':>9}{'missWrMB':>9}"
                 f"{'L2hit%':>7}{'VALU%':>6}{'valu/w':>8}{'lds/w':>7}{'vmem/w':>7}"
                 f"{'ldsW%':>6}{'bankc':>6}{'waves':>8}")
    for phase, short in PROFILED_KERNELS:
        key = pick(phase, short)
        if not key:
            continue
        waves = val(key, "SQ_WAVES")
        hit, miss = val(key, "TCC_HIT_sum"), val(key, "TCC_MISS_sum")
        gui = val(key, "GRBM_GUI_ACTIVE")
        lds = val(key, "SQ_INSTS_LDS")
        us = _avg(dur.get(key, [])) * 1e-3
        label = short.split('<')[0]
        lines.append(
            f"{phase:<6}{label:<24}{us:>8.1f}{l2_read(key) / 1e6:>9.2f}{l2_write(key) / 1e6:>9.2f}"
            f"{100 * hit / (hit + miss):>7.1f}"
            f"{100 * val(key, 'SQ_ACTIVE_INST_VALU') / NUM_CU / (gui / NUM_XCD):>6.1f}"
            f"{val(key, 'SQ_INSTS_VALU') / waves:>8.1f}{lds / waves:>7.1f}"
            f"{val(key, 'SQ_INSTS_VMEM') / waves:>7.1f}"
            f"{100 * val(key, 'SQ_WAIT_INST_LDS') / val(key, 'SQ_WAVE_CYCLES'):>6.1f}"
            f"{(val(key, 'SQ_LDS_BANK_CONFLICT') / lds) if lds else 0.0:>6.2f}{waves:>8.0f}")
    lines.append("per launch, read back by bench.py: L2-miss bytes to memory (read + write; the "
                 "roofline `traffic`), L2 hit %, VALU busy %, kernel us:")
    for phase, short in PROFILED_KERNELS:
        key = pick(phase, short)
        if key:
            hit, miss = val(key, "TCC_HIT_sum"), val(key, "TCC_MISS_sum")
            gui = val(key, "GRBM_GUI_ACTIVE")
            lines.append(f"per_launch {phase}:{short.split('<')[0]} "
                         f"mem_bytes={l2_read(key) + l2_write(key):.6e} "
                         f"l2_hit_pct={100 * hit / (hit + miss):.2f} "
                         f"valu_busy_pct={100 * val(key, 'SQ_ACTIVE_INST_VALU') / NUM_CU / (gui / NUM_XCD):.2f} "
                         f"us={_avg(dur.get(key, [])) * 1e-3:.2f}")
    return "\n".join(lines) + "\n"


def fp64_error(tri, vtx, n2w_rowmajor, pix2tri_h, pix2t_h, depth_h):
    """Error of the f32 hit coefficient and depth against a float64
    evaluation of the same pixel ray, the same (stored) triangle and the same
    projection: Moller-Trumbore and the homogeneous transforms in f64 on the
    f32 inputs.  Returns max / mean absolute error, max relative error and
    the error in units of the f32 spacing at the f64 value."""
    import numpy as np

    h, w = pix2tri_h.shape
    hit = pix2tri_h.reshape(-1) != 0xFFFFFFFF
    ip = np.nonzero(hit)[0]
    iw, ih = (ip % w).astype(np.float64), (ip // w).astype(np.float64)
    T = np.asarray(n2w_rowmajor, np.float32).astype(np.float64)  # ndc -> world, row-major

    def xform(m, p):
        q = p @ m[:3, :3].T + m[:3, 3]
        wq = p @ m[3, :3] + m[3, 3]
        return q / wq[:, None]

    x0 = 2.0 * (iw + 0.5) / w - 1.0
    y0 = 1.0 - 2.0 * (ih + 0.5) / h
    one = np.ones_like(x0)
    org = xform(T, np.stack([x0, y0, one], 1))
    far = xform(T, np.stack([x0, y0, -one], 1))
    d = far - org
    t3 = tri[pix2tri_h.reshape(-1)[ip]].astype(np.int64)
    v = vtx.astype(np.float64)
    p0, p1, p2 = v[t3[:, 0]], v[t3[:, 1]], v[t3[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    pv = np.cross(d, e2)
    det = np.einsum("ij,ij->i", e1, pv)
    qv = np.cross(org - p0, e1)
    t64 = np.einsum("ij,ij->i", e2, qv) / det
    depth64 = (xform(np.linalg.inv(T), org + t64[:, None] * d)[:, 2] + 1.0) * 0.5

    def stats(got, want):
        got = got.astype(np.float64)
        err = np.abs(got - want)
        ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        return {"max_abs": float(err.max()), "mean_abs": float(err.mean()),
                "max_rel": float((err / np.abs(want)).max()),
                "max_f32_ulps": round(float((err / ulp).max()), 2),
                "mean_f32_ulps": round(float((err / ulp).mean()), 3)}

    return {"hit_pixels": int(ip.size),
            "pix2t": stats(pix2t_h.reshape(-1)[ip], t64),
            "depth": stats(depth_h.reshape(-1)[ip], depth64),
            "reference": "float64 ray, Moller-Trumbore and projection on the f32 inputs and the "
                         "stored triangle"}


def read_counters(kernel: str = "cast:k_pix2tri"):
    """Per-launch figures of `kernel` (phase:name) from the committed counter
    summary: {mem_bytes, l2_hit_pct, valu_busy_pct, us}, or
    None when the summary does not list it."""
    try:
        for line in COUNTER_SUMMARY.read_text().splitlines():
            parts = line.split()
            if parts[:2] == ["per_launch", kernel]:
                return {k: float(v) for k, v in (p.split("=") for p in parts[2:])}
    except OSError:
        return None
    return None


# ---------------------------------------------------------------------------
# batched workload: many meshes per launch vs one mesh per launch
# ---------------------------------------------------------------------------
def run_batched(num_mesh: int, reps: int):
    """Seeded batch of sphere/torus/cylinder meshes (log-uniform 256..262144
    triangles, scale U[0.5,2], random rotation, one orthographic image per
    mesh with log-uniform 64^2..512^2 pixels, seed 3).  Times the batched
    build and ray-cast (autotuned, eager and replayed from a HIP graph)
    against the plain single-mesh entry points launched once per mesh on the
    same data, and checks that both give identical trees and images."""
    import numpy as np
    import torch

    from del_msh_dlpack import _lib, primitives
    import del_msh_dlpack.Batched.torch as Batched

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    b = primitives.batch_meshes(num_mesh, seed=3)
    gen_s = time.perf_counter() - t0
    p = Batched.Pipeline(b["mesh2tri"], b["mesh2vtx"], torch.from_numpy(b["tri2vtx"]).to(dev),
                         torch.from_numpy(b["vtx2xyz"]).to(dev), b["mesh2pix"], b["mesh2width"],
                         torch.from_numpy(b["ndc2world"]).to(dev), with_coeff=True)
    tuned = Batched.autotune(p, reps=1)
    build_ms = Batched._time_ms(p.build, reps)
    cast_ms = Batched._time_ms(p.cast, reps)
    p.capture()
    graph_ms = Batched._time_ms(p.replay, reps)
    errs = _lib.device_errors()
    if errs:
        raise RuntimeError(f"device error bits 0x{errs:x} in the batched run")

    # plain: the single-mesh entry points, one mesh per call, same data
    plain = Batched.PerMeshLoop(p)
    plain_build_ms = Batched._time_ms(plain.build, 1)
    plain_cast_ms = Batched._time_ms(plain.cast, 1)
    same = plain.equals_pipeline()
    errs = _lib.device_errors()
    if errs:
        raise RuntimeError(f"device error bits 0x{errs:x} in the plain loop")
    del plain

    # packed: the build writes every mesh's traversal buffer as it refits and
    # the cast reads them as one request per mesh, so no cast packs again
    ref_pix, ref_t = p.pix2tri.clone(), p.pix2t.clone()
    q = Batched.Pipeline(b["mesh2tri"], b["mesh2vtx"], p.tri2vtx, p.vtx2xyz, b["mesh2pix"],
                         b["mesh2width"], torch.from_numpy(b["ndc2world"]).to(dev),
                         with_coeff=True, packed=True)
    q.run()
    packed_build_ms = Batched._time_ms(q.build, reps)
    packed_cast_ms = Batched._time_ms(q.cast, reps)
    q.capture()
    packed_graph_ms = Batched._time_ms(q.replay, reps)
    torch.cuda.synchronize()
    packed_same = bool(torch.equal(q.pix2tri, ref_pix)
                       and torch.equal(q.pix2t.view(torch.int32), ref_t.view(torch.int32)))
    errs = _lib.device_errors()
    if errs:
        raise RuntimeError(f"device error bits 0x{errs:x} in the packed pipeline")
    tri, px = p.total_tri, p.total_pix
    packed = {
        "build_ms": round(packed_build_ms, 3),
        "build_mtri_per_s": round(tri / packed_build_ms / 1e3, 1),
        "cast_ms": round(packed_cast_ms, 3),
        "cast_mrays_per_s": round(px / packed_cast_ms / 1e3, 1),
        "graph_build_cast_ms": round(packed_graph_ms, 3),
        "speedup_build": round(plain_build_ms / packed_build_ms, 2),
        "speedup_cast": round(plain_cast_ms / packed_cast_ms, 2),
        "images_equal": packed_same,
        "how": "dmsh_batch_build_pack_trimesh3 + dmsh_cast_requests (one request per mesh)",
    }
    return {
        "workload": (f"{num_mesh} meshes (sphere/torus/cylinder, 256..262144 tris log-uniform, "
                     f"scale U[0.5,2], random rotation, seed 3), 30-bit codes, per-mesh "
                     f"orthographic images 64^2..512^2 log-uniform"),
        "num_mesh": num_mesh, "total_tri": tri, "total_rays": px,
        "generate_s": round(gen_s, 2),
        "build_ms": round(build_ms, 3), "build_mtri_per_s": round(tri / build_ms / 1e3, 1),
        "cast_ms": round(cast_ms, 3), "cast_mrays_per_s": round(px / cast_ms / 1e3, 1),
        "graph_build_cast_ms": round(graph_ms, 3),
        "plain_build_ms": round(plain_build_ms, 3),
        "plain_build_mtri_per_s": round(tri / plain_build_ms / 1e3, 1),
        "plain_cast_ms": round(plain_cast_ms, 3),
        "plain_cast_mrays_per_s": round(px / plain_cast_ms / 1e3, 1),
        "speedup_build": round(plain_build_ms / build_ms, 2),
        "speedup_cast": round(plain_cast_ms / cast_ms, 2),
        "plain_equals_batched": bool(same),
        "autotuned": {k: v for k, v in tuned.items() if k != "trials"},
        "packed": packed,
    }


# ---------------------------------------------------------------------------
# served workload: continuous batching vs one request at a time
# ---------------------------------------------------------------------------
def run_trees(reps: int):
    """Other users of the Morton / LBVH machinery: a 3D octree over 2^22
    uniform points (codes, sort, unique, tree, aggregate of 3 values per
    point), and point location + interpolation of 2^22 query points in a
    mixed tet / pyramid / prism / hex grid of 48^3 jittered cubes (element
    BVH build included).  Medians over `reps` runs, HIP events."""
    import numpy as np
    import torch

    import del_msh_dlpack.Array1D.torch as Array1D
    import del_msh_dlpack.Mortons.torch as Mortons
    import del_msh_dlpack.OffsetArray.torch as OffsetArray
    import del_msh_dlpack.PolyhedronMesh.torch as PolyhedronMesh
    import del_msh_dlpack.QuadOctTree.torch as QuadOctTree
    from del_msh_dlpack import primitives

    dev = torch.device("cuda", 0)

    def timed(fn, n):
        ev = []
        for _ in range(n):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            ev.append((a, b))
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    n = 1 << 22
    g = torch.Generator(device="cpu").manual_seed(0)
    vtx2co = torch.rand((n, 3), generator=g).to(dev)
    eye = torch.eye(4, device=dev)
    vals = torch.ones((n, 3), dtype=torch.float32, device=dev)
    state = {}

    def codes():
        m = Mortons.make_vtx2morton_from_vtx2co(vtx2co, eye)
        state["jdx2vtx"], state["sorted"] = Array1D.argsort(m)
        _, state["uniq"], state["off"] = Array1D.unique_for_sorted_array(state["sorted"])

    def tree():
        t = QuadOctTree.QuadOctTree()
        t.construct_from_idx2morton(state["uniq"], 3)
        state["tree"] = t

    def agg():
        idx2agg = OffsetArray.aggregate(state["off"], state["jdx2vtx"], vals)
        state["agg"] = state["tree"].aggregate(idx2agg)

    for f in (codes, tree, agg):
        f()
    torch.cuda.synchronize()
    codes_ms, tree_ms, agg_ms = timed(codes, reps), timed(tree, reps), timed(agg, reps)
    octree = {
        "workload": "2^22 points uniform in [0,1]^3 (seed 0), 30-bit codes, max depth 10",
        "points": n, "unique_codes": int(state["uniq"].shape[0]),
        "onodes": int(state["tree"].onodes.shape[0]),
        "codes_sort_unique_ms": round(codes_ms, 4), "tree_ms": round(tree_ms, 4),
        "aggregate_ms": round(agg_ms, 4),
        "mpts_per_s": round(n / ((codes_ms + tree_ms) * 1e-3) / 1e6, 1),
        "root_sum_ok": bool(state["agg"][0, 0].item() == n),
    }

    off, iv, x = primitives.mixed_hex_grid(48, 48, 48, seed=1, jitter=0.15)
    d_off, d_iv, d_x = (torch.from_numpy(a).to(dev) for a in (off, iv, x))
    q = torch.from_numpy(np.random.default_rng(2).uniform(0.0, 48.0, (n, 3)).astype(np.float32)).to(dev)
    v = torch.from_numpy(x[:, :2].copy()).to(dev)

    def build():
        state["bvh"] = PolyhedronMesh.make_bvhnodes_bvhnode2aabb(d_off, d_iv, d_x)

    def search():
        bn, ab = state["bvh"]
        state["found"] = PolyhedronMesh.search_elem_contain_points(d_off, d_iv, d_x, bn, ab, q)

    def interp():
        e, p = state["found"]
        state["val"] = PolyhedronMesh.interpolate_values_at_points(d_off, d_iv, v, e, p)

    for f in (build, search, interp):
        f()
    torch.cuda.synchronize()
    build_ms, search_ms, interp_ms = timed(build, reps), timed(search, reps), timed(interp, reps)
    e = state["found"][0]
    located = {
        "workload": (f"mixed grid 48^3 cubes (hex / 2 prisms / 6 pyramids / 6 tets per cube, "
                     f"seed 1, jitter 0.15): {off.size - 1} elements; {n} query points U[0,48]^3"),
        "elements": int(off.size - 1), "queries": n,
        "found_fraction": round(float((e != 0xFFFFFFFF).float().mean().item()), 4),
        "bvh_build_ms": round(build_ms, 4), "search_ms": round(search_ms, 4),
        "interpolate_ms": round(interp_ms, 4),
        "mqueries_per_s": round(n / (search_ms * 1e-3) / 1e6, 1),
    }
    return {"octree": octree, "point_location": located}


def run_diff(reps: int):
    """Differentiation workload.  (1) A 50,000-triangle sphere (sphere_yup(1,
    126, 200)) seen from (0, 0, 3) at 512^2, dL/dpixel ~ N(0, 1) (seed 4):
    forward = LBVH build + ray-cast + contour + anti-aliased silhouette,
    backward = antialias backward; the edge-gradient backward of the same
    image timed beside it.  (2) A batch of 64 meshes (5k..200k triangles
    log-uniform, seed 6, 512^2 each) through Batched.AntialiasedSilhouette,
    with the hit buffers kept from forward or recomputed in backward."""
    import numpy as np
    import torch

    from del_msh_dlpack import _lib, primitives
    import del_msh_dlpack.Batched.torch as Batched
    import del_msh_dlpack.DifferentiableAntialias.torch as DA
    import del_msh_dlpack.Edge2Vtx.torch as Edge2Vtx
    import del_msh_dlpack.Mat44.torch as Mat44
    import del_msh_dlpack.Pix2Tri.torch as Pix2Tri
    import del_msh_dlpack.RasterizedEdgeGradient.torch as REG
    import del_msh_dlpack.TriMesh3.torch as TriMesh3

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    W = 512
    tri, vtx = primitives.sphere_yup(1.0, 126, 200)
    w2c = Mat44.look_at((0.0, 0.0, 3.0), (0, 0, 0), (0, 1, 0), dtype=torch.float64)
    w2n = Mat44.perspective(45.0, 1.0, 1.0, 5.0, dtype=torch.float64) @ w2c
    w2p = (Mat44.from_transform_ndc2pix((W, W), dtype=torch.float64) @ w2n).float().to(dev)
    n2w = torch.linalg.inv(w2n).float().contiguous().to(dev)
    w2n = w2n.float().contiguous().to(dev)
    t = torch.from_numpy(tri).to(dev)
    v0 = torch.from_numpy(vtx).to(dev)
    e2v = TriMesh3.make_edge2vtx(t, vtx.shape[0])
    e2t = TriMesh3.make_edge2tri(t, vtx.shape[0], e2v)
    one = e2v[:, 0].to(torch.int64) < e2v[:, 1].to(torch.int64)
    e2v = e2v.view(torch.int32)[one].view(torch.uint32).contiguous()
    e2t = e2t.view(torch.int32)[one].view(torch.uint32).contiguous()
    dldw = torch.from_numpy(np.random.default_rng(4).standard_normal((W, W)).astype(np.float32)).to(dev)

    def forward(v):
        bn, ab = TriMesh3.make_bvhnodes_bvhnode2aabb(t, v)
        pix = Pix2Tri.by_raycasting(t, v, bn, ab, n2w, (W, W))
        cedge = Edge2Vtx.contour_for_triangle_mesh(t, v, w2n, e2v, e2t)
        occ = (pix != 0xFFFFFFFF).to(torch.float32)
        return DA.DifferentiableAntialiasFunction.apply(cedge, v, w2p, pix, occ), pix, occ, cedge

    def timed(fn, n):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(n)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    state = {}

    def fwd():
        v = v0.clone().requires_grad_(True)
        out = forward(v)
        state["v"], state["out"] = v, out

    def bwd():
        (state["out"][0] * dldw).sum().backward()

    for _ in range(2):
        fwd()
        bwd()
    f_ms, b_ms = [], []
    for _ in range(reps):
        f_ms.append(timed(fwd, 1))
        b_ms.append(timed(bwd, 1))
    _, pix, occ, cedge = state["out"]
    occ3 = occ[..., None].contiguous()
    dl3 = dldw[..., None].contiguous()
    eg_ms = timed(lambda: REG.bwd(t, v0, w2p, pix, occ3, dl3), max(reps, 5))

    def smooth_bwd():
        v = v0.clone().requires_grad_(True)
        out = REG.AutogradWithSmooth.apply(t, v, w2p, pix, occ3)
        (out * dl3).sum().backward()

    smooth_bwd()
    smooth_ms = timed(smooth_bwd, max(reps, 5))
    _lib.timing_enable(True)
    for _ in range(3):
        REG.bwd(t, v0, w2p, pix, occ3, dl3)
        DA.bwd_antialias(cedge, v0, w2p, occ, dldw, pix)
    smooth_bwd()
    torch.cuda.synchronize()
    _lib.timing_enable(False)
    kern = {}
    for name in ("k_edge_grad_bwd", "k_aa_bwd", "k_edge_type_grad", "k_smooth_links",
                 "k_smooth_tiled", "k_interpolate_staggered"):
        n, total, _ = _lib.timing_query(name)
        if n:
            kern[name] = {"avg_ms": round(total / n, 5), "launches": n}
    ph = pix.cpu().numpy()
    active = int((ph[:, 1:] != ph[:, :-1]).sum() + (ph[1:, :] != ph[:-1, :]).sum())
    # edge-gradient basis: 12 B per pixel, 12 B per vertex, 84 B per active pair
    eg_bytes = W * W * 12 + vtx.shape[0] * 12 + active * 84
    single = {
        "workload": "sphere_yup(1,126,200): 50000 tris, perspective eye (0,0,3) fovy 45, 512^2, "
                    "dL/dpixel N(0,1) seed 4",
        "forward_ms": round(float(np.median(f_ms)), 4),
        "backward_ms": round(float(np.median(b_ms)), 4),
        "contour_edges": int(cedge.shape[0]),
        "edge_grad_bwd_ms": round(eg_ms, 4),
        "smoothed_edge_grad_bwd_ms": round(smooth_ms, 4),
        "kernels_ms": kern,
        "edge_grad_active_pairs": active,
    }
    if "k_edge_grad_bwd" in kern:
        single["edge_grad_bwd_algorithmic_gbs"] = round(
            eg_bytes / (kern["k_edge_grad_bwd"]["avg_ms"] * 1e-3) / 1e9, 1)

    b = primitives.batch_meshes(64, seed=6, tri_range=(5000, 200000), pix_range=(W * W, W * W))
    sb = Batched.SilhouetteBatch(b["mesh2tri"], b["mesh2vtx"], torch.from_numpy(b["tri2vtx"]).to(dev),
                                 torch.from_numpy(b["vtx2xyz"]).to(dev), b["mesh2pix"], b["mesh2width"],
                                 torch.from_numpy(b["ndc2world"]))
    bdl = torch.from_numpy(np.random.default_rng(4).standard_normal(int(b["mesh2pix"][-1]))
                           .astype(np.float32)).to(dev)
    bv0 = torch.from_numpy(b["vtx2xyz"]).to(dev)
    batch = {"workload": "64 meshes (sphere/torus/cylinder, 5k..200k tris log-uniform, seed 6), 512^2 each",
             "total_tri": int(b["mesh2tri"][-1])}
    for recompute in (False, True):
        def step():
            v = bv0.clone().requires_grad_(True)
            img = Batched.AntialiasedSilhouette.apply(v, sb, recompute)
            (img * bdl).sum().backward()

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        batch["recompute" if recompute else "stored"] = {
            "fwd_bwd_ms": round((time.perf_counter() - t0) * 1e3 / reps, 3)}
    batch["hit_buffer_bytes_not_kept"] = Batched.SilhouetteBatch.hit_bytes(sb.hits(bv0))
    return {"single": single, "batch64": batch}


def run_served(num_req: int, rates, budget_bytes: int):
    """Poisson request streams (mesh id Zipf(1.1) over 10k ids, 1k..1M
    triangles with radial noise, eye on the sphere of radius 3, viewport
    256^2..1024^2 log-uniform, seed 7) served by the continuous-batching
    server and by the per-request baseline (single-mesh entry points), both
    building trees on first sight into a cache of `budget_bytes`."""
    import copy

    import numpy as np

    import del_msh_dlpack.Serve.torch as Serve

    cat = Serve.MeshCatalog(tri_range=(1000, 1_000_000), seed=7)
    out = {"workload": (f"{num_req} requests per rate, mesh id Zipf(1.1) over 10000 ids, "
                        f"1k..1M tris, perspective 256^2..1024^2, cache {budget_bytes >> 20} MiB"),
           "rates": {}}
    # warm-up, not timed, as a server starts: the allocator's pool is
    # reserved (the cache budget and 8 GiB for a step's buffers; otherwise
    # the pool grows inside the first timed steps while the GPU waits on the
    # host), and a kernel's code object loads on its first launch, so both
    # servers first run a short stream of other mesh ids
    import torch
    pool = torch.empty(budget_bytes + (8 << 30), dtype=torch.uint8, device="cuda:0")
    del pool
    warm = Serve.request_stream(24, rate=1e5, num_ids=12, seed=11)
    for r in warm:
        r.mesh_id += 10_000
    Serve.Server(cat, budget_bytes).run(copy.deepcopy(warm))
    Serve.PlainServer(cat, budget_bytes).run(copy.deepcopy(warm))
    for rate in rates:
        reqs = Serve.request_stream(num_req, rate=rate, num_ids=10_000, seed=7)
        t0 = time.perf_counter()
        for mid in sorted({r.mesh_id for r in reqs}):
            cat.pinned(mid)  # host-side meshes exist (pinned) before the stream starts
        gen_s = time.perf_counter() - t0
        served = copy.deepcopy(reqs)
        srv = Serve.Server(cat, budget_bytes)
        steps = srv.run(served)
        s_stats = Serve.latency_stats(served)
        host = np.array(srv.host_s) * 1e3
        s_stats.update({"steps": steps, "cache_hits": srv.cache.hits,
                        "cache_misses": srv.cache.misses, "evictions": srv.cache.evicted,
                        # build with the pack into the traversal layout fused
                        # into the refit (dmsh_batch_build_pack_trimesh3)
                        "build": Serve.build_stats(srv.builds),
                        # share of the span the GPU stream was working on steps
                        # (near 1: the GPU, not the host loop, sets the pace)
                        "gpu_busy_share": round(Serve.gpu_busy_s(srv.step_gpu)
                                                / s_stats["span_s"], 4),
                        # host time to admit, build-enqueue and cast-enqueue a step
                        "host_ms_per_step": {"mean": round(float(host.mean()), 3),
                                             "p99": round(float(np.percentile(host, 99)), 3),
                                             "share_of_span": round(float(host.sum() * 1e-3
                                                                          / s_stats["span_s"]), 4)}})
        plain = copy.deepcopy(reqs)
        psrv = Serve.PlainServer(cat, budget_bytes)
        psrv.run(plain)
        p_stats = Serve.latency_stats(plain)
        p_stats.update({"cache_hits": psrv.cache.hits, "cache_misses": psrv.cache.misses,
                        "build": Serve.build_stats(psrv.builds)})
        ratio = {"mrays_per_s": round(s_stats["mrays_per_s"] / p_stats["mrays_per_s"], 2)}
        if s_stats["build"].get("mtri_per_s") and p_stats["build"].get("mtri_per_s"):
            ratio["build_mtri_per_s"] = round(s_stats["build"]["mtri_per_s"]
                                              / p_stats["build"]["mtri_per_s"], 2)
        progress(f"served rate {rate:g}: {ratio}")
        out["rates"][str(rate)] = {"served": s_stats, "plain": p_stats, "served_over_plain": ratio,
                                   "distinct_meshes": len({r.mesh_id for r in reqs}),
                                   "host_generate_s": round(gen_s, 2)}
    return out


def progress(msg: str) -> None:
    """A line on stderr as each stage ends (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _median_s(fn, warmup=3, reps=10):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2])


def run_cpu_baseline(tri, vtx, n2w, W, H, nodes_h, aabb_h, pix_gpu):
    """The oracle (a C restatement of del-msh-cpu) built for this host with
    -O3 -march=native -ffp-contract=off and timed on the same workload: the
    serial build (del-msh-cpu's single thread), the all-cores build
    (oracle/dmsh_cpu_mt.c, the same tree bit for bit) and the all-cores
    ray-cast; 3 warm-ups, then the median of 10.  Also counts the traversal
    work of the flat closest-hit search and checks the GPU image against it."""
    import ctypes
    import tempfile

    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    tmp = tempfile.mkdtemp(prefix="dmsh_cpu_")
    Lnat = oracle.load(oracle.build_native(tmp))
    num_tri = tri.shape[0]
    num_pix = W * H
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = max(1, min(16, avail))
    cm = n2w.T.reshape(-1)
    t3 = oracle._u32(tri)

    def serial_build():
        cntr = np.empty((num_tri, 3), np.float32)
        Lnat.orc_tri2centroid(num_tri, t3.reshape(-1), oracle._f32(vtx).reshape(-1), cntr.reshape(-1))
        T = np.empty(16, np.float32)
        Lnat.orc_fit_into_unit_cube(num_tri, cntr.reshape(-1), T)
        codes = np.empty(num_tri, np.uint64)
        Lnat.orc_vtx2morton_u64(num_tri, cntr.reshape(-1), T, codes)
        idx = np.empty(num_tri, np.uint32)
        sk = np.empty(num_tri, np.uint64)
        Lnat.orc_argsort_u64(num_tri, codes, idx, sk)
        bn = np.zeros((2 * num_tri - 1) * 3, np.uint32)
        assert Lnat.orc_bvhnodes_u64(bn, idx, sk, num_tri) == 0
        ab = np.zeros((2 * num_tri - 1) * 6, np.float32)
        assert Lnat.orc_bvhnode2aabb_trimesh3(ab, bn, 2 * num_tri - 1, t3.reshape(-1),
                                              oracle._f32(vtx).reshape(-1), 0.0) == 0
        return bn, ab

    bn_s, ab_s = serial_build()
    serial_s = _median_s(serial_build)
    bn_m, ab_m = oracle.build_trimesh3_mt(tri, vtx, threads, L=Lnat)
    mt_s = _median_s(lambda: oracle.build_trimesh3_mt(tri, vtx, threads, L=Lnat))
    same_tree = bool(np.array_equal(bn_s.reshape(-1, 3), nodes_h)
                     and np.array_equal(bn_m, nodes_h)
                     and np.array_equal(ab_s.view(np.uint32), aabb_h.reshape(-1).view(np.uint32))
                     and np.array_equal(ab_m.reshape(-1).view(np.uint32),
                                        aabb_h.reshape(-1).view(np.uint32)))

    def cast():
        out = np.empty(num_pix, np.uint32)
        st = oracle.TraceStats()
        Lnat.orc_pix2tri(out, None, num_tri, t3.reshape(-1), oracle._f32(vtx).reshape(-1),
                         nodes_h.reshape(-1), aabb_h.reshape(-1), W, H, cm, threads,
                         ctypes.byref(st))
        return out

    wpix = cast()
    cast_s = _median_s(cast)
    mism = int(np.count_nonzero(wpix.reshape(H, W) != pix_gpu))
    # the work of the search the kernel runs (one best hit, far children on a
    # stack), which equals del-msh-cpu's recursion in result
    fpix, _, (n_exp, n_leaf, fhist) = oracle.pix2tri_flat(tri, vtx, nodes_h, aabb_h, cm, W, H,
                                                          num_threads=threads, with_t=False)
    mism += int(np.count_nonzero(fpix != wpix.reshape(H, W)))
    cast_counts = (n_exp, n_leaf, int(np.nonzero(fhist)[0].max()))
    cpu = {
        "value": round(num_tri / serial_s / 1e6, 3),
        "unit": "Mtri/s",
        "cores": 1,
        "threads": 1,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "nproc": nproc,
        "cores_available": avail,
        "build_flags": "gcc -O3 -march=native -ffp-contract=off -fno-fast-math -fopenmp",
        "timing": "3 warm-ups, median of 10",
        "sample": (f"the full {num_tri}-triangle LBVH build with 63-bit codes (centroid, fit, "
                   f"Morton, stable argsort, topology, refit) on one thread, as del-msh-cpu runs "
                   f"it; trees and boxes equal the GPU build's: {same_tree}"),
        "all_cores_build": {"value": round(num_tri / mt_s / 1e6, 3), "unit": "Mtri/s",
                            "threads": threads, "label": "all cores (oracle/dmsh_cpu_mt.c: "
                            "parallel stages, radix argsort, atomic-counter refit)"},
        "all_cores_raycast": {"value": round(num_pix / cast_s / 1e6, 3), "unit": "Mrays/s",
                              "threads": threads, "label": f"all {W}x{H} rays, del-msh-cpu's "
                              "recursion per ray, rays spread over threads"},
        "gpu_oracle_pix2tri_mismatches": mism,
    }
    return cpu, cast_counts


# ---------------------------------------------------------------------------
# main
# ---------------------------------------------------------------------------
def main():
    args = parse()
    if args.summarize_counters:
        sys.stdout.write(summarize_counters(
            args.summarize_counters,
            "Counters per kernel on the bench workload (octahedral sphere level 9 = 2^21 tris, "
            "radial noise\nN(0,0.01^2) seed 1, 63-bit LBVH build, standalone refit of the built tree, "
            "perspective 2048x2048 cast)."))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the batched/served paths shard meshes per GPU, which this build does not run)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from del_msh_dlpack import _lib, primitives

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    L = _lib.lib()

    # ---------------- inputs (resident before timing) ----------------
    tri, vtx = primitives.octahedral_sphere(args.level, noise_sigma=0.01, seed=1)
    num_tri, num_vtx = tri.shape[0], vtx.shape[0]
    W = H = args.img
    num_pix = W * H
    n2w = perspective_ndc2world()
    t_tri = torch.from_numpy(tri).to(dev)
    t_vtx = torch.from_numpy(vtx).to(dev)
    t_n2w = torch.from_numpy(np.ascontiguousarray(n2w.T)).to(dev)  # column-major
    num_node = 2 * num_tri - 1
    bvhnodes = torch.empty((num_node, 3), dtype=torch.uint32, device=dev)
    aabb = torch.empty((num_node, 6), dtype=torch.float32, device=dev)
    ws_build = _lib.workspace(L.dmsh_build_workspace_bytes(num_tri, 63), dev)
    pix2tri = torch.empty((H, W), dtype=torch.uint32, device=dev)
    pix2t = torch.empty((H, W), dtype=torch.float32, device=dev)
    ws_cast = _lib.workspace(L.dmsh_pix2tri_workspace_bytes(num_tri, W, H), dev)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def build():
        _lib.check(L.dmsh_build_trimesh3(sp, num_tri, t_tri.data_ptr(), num_vtx, t_vtx.data_ptr(), 63,
                                         bvhnodes.data_ptr(), aabb.data_ptr(), ws_build.data_ptr(),
                                         ws_build.numel()), "build")

    def cast():
        _lib.check(L.dmsh_pix2tri(sp, pix2tri.data_ptr(), pix2t.data_ptr(), num_tri, t_tri.data_ptr(),
                                  num_vtx, t_vtx.data_ptr(), W, H, t_n2w.data_ptr(),
                                  bvhnodes.data_ptr(), aabb.data_ptr(), ws_cast.data_ptr(),
                                  ws_cast.numel()), "pix2tri")

    if args.pmc_workload:
        r_ws = _lib.workspace(L.dmsh_refit_workspace_bytes(num_tri), dev)
        r_box = torch.empty_like(aabb)
        # every profiled refit builds its plan (the kept-plan call runs the
        # same kernels minus the two plan kernels)
        assert L.dmsh_set_tuning(b"refit.plan_cache", 0) == 0
        for _ in range(3):
            build()
        for _ in range(3):
            cast()
        for _ in range(3):
            _lib.check(L.dmsh_bvhnode2aabb_trimesh3(sp, r_box.data_ptr(), num_node,
                                                    bvhnodes.data_ptr(), num_tri, t_tri.data_ptr(),
                                                    num_vtx, t_vtx.data_ptr(), 0.0, r_ws.data_ptr(),
                                                    r_ws.numel()), "refit")
        src = torch.empty(CALIB_BYTES, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_workload": "done", "builds": 3, "casts": 3, "refits": 3,
                          "copies": 3}))
        return

    for _ in range(args.warmup):
        build()
        cast()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        build()
        e1.record(stream)
        cast()
        e2.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the same steps again with the library's per-kernel events on (the event
    # records add gaps between kernels, so this pass is kept out of `value`)
    _lib.timing_enable(True)
    for _ in range(args.steps):
        build()
        cast()
    torch.cuda.synchronize()
    _lib.timing_enable(False)
    kernels = {}
    for name in TIMED_KERNELS:
        n, total, mx = _lib.timing_query(name)
        if n:
            kernels[name] = {"launches_per_step": round(n / args.steps, 3),
                             "avg_ms": round(total / n, 5), "max_ms": round(mx, 5)}
    build_ms = [a.elapsed_time(b) for a, b, _ in ev]
    cast_ms = [b.elapsed_time(c) for _, b, c in ev]
    errs = _lib.device_errors()
    if errs:
        raise RuntimeError(f"device error bits 0x{errs:x}")
    ms_step = wall * 1e3 / args.steps
    build_avg = float(np.mean(build_ms))
    cast_avg = float(np.mean(cast_ms))
    build_mtri = num_tri / (build_avg * 1e-3) / 1e6
    cast_mrays = num_pix / (cast_avg * 1e-3) / 1e6

    progress(f"build + cast timed: build {build_avg:.4f} ms, cast {cast_avg:.4f} ms")
    # ---------------- standalone refit of the built tree ----------------
    # dmsh_bvhnode2aabb_trimesh3 on the tree just built (the reference's
    # bvhnode2aabb call, as an optimisation loop makes it after moving the
    # vertices): leaf ranges from the rows, treelet list, treelets, top.
    r_aabb = torch.empty_like(aabb)
    ws_refit = _lib.workspace(L.dmsh_refit_workspace_bytes(num_tri), dev)

    def refit():
        _lib.check(L.dmsh_bvhnode2aabb_trimesh3(sp, r_aabb.data_ptr(), num_node, bvhnodes.data_ptr(),
                                                num_tri, t_tri.data_ptr(), num_vtx,
                                                t_vtx.data_ptr(), 0.0, ws_refit.data_ptr(),
                                                ws_refit.numel()), "refit")

    def refit_median(reps=20):
        for _ in range(3):
            refit()
        r_ev = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            refit()
            b.record(stream)
            r_ev.append((a, b))
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in r_ev]))

    # every call builds the plan (leaf ranges, treelet and top lists) ...
    assert L.dmsh_set_tuning(b"refit.plan_cache", 0) == 0
    refit_plan_ms = refit_median()
    assert L.dmsh_set_tuning(b"refit.plan_cache", 1) == 0
    # ... or keeps it from the previous call on the same rows (hash-checked):
    # the optimisation-loop case, where only the vertices move
    refit_ms = refit_median()
    _lib.timing_enable(True)
    for _ in range(10):
        refit()
    torch.cuda.synchronize()
    _lib.timing_enable(False)
    refit_kernels = {}
    for name in ("k_plan_hash", "k_plan_gate", "k_leaf_ranges", "k_list_treelets",
                 "k_refit_treelet", "k_refit_top", "k_refit"):
        n, total, _ = _lib.timing_query(name)
        if n:
            refit_kernels[name] = round(total / n, 5)
    refit_same = bool(torch.equal(r_aabb.view(torch.int32), aabb.view(torch.int32)))
    # algorithmic bytes per triangle (refit basis): read tri2vtx 12 + vertices
    # 12 V/N ~ 6 + node rows 12 (2N-1)/N ~ 24; write boxes 24 (2N-1)/N ~ 48;
    # arrival counters 8 => 98 B
    refit_bytes = 98 * num_tri
    refit_line = {
        "entry": "dmsh_bvhnode2aabb_trimesh3 (stored tree, treelet path)",
        "ms": round(refit_ms, 4),
        "plan": "kept in the workspace from the previous call on the same rows (hash-checked)",
        "ms_plan_rebuilt_every_call": round(refit_plan_ms, 4),
        "frac_of_hbm_peak_plan_rebuilt": round(refit_bytes / (refit_plan_ms * 1e-3) / 1e9
                                               / PEAK_HBM_GBS, 4),
        "mtri_per_s": round(num_tri / (refit_ms * 1e-3) / 1e6, 1),
        "algorithmic_gbs": round(refit_bytes / (refit_ms * 1e-3) / 1e9, 1),
        "frac_of_hbm_peak": round(refit_bytes / (refit_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "kernels_avg_ms": refit_kernels,
        "equals_build_boxes": refit_same,
    }
    # per-kernel DRAM traffic of the standalone refit from the counter summary
    rk = {k: read_counters(f"refit:{k}") for k in ("k_plan_hash", "k_leaf_ranges",
                                                   "k_list_treelets", "k_refit_treelet",
                                                   "k_refit_top_lds")}
    if all(rk.values()):
        kept = ("k_plan_hash", "k_refit_treelet", "k_refit_top_lds")
        dram = sum(v["mem_bytes"] for k, v in rk.items() if k in kept)
        refit_line["roofline"] = {
            "bound": "hbm",
            "achieved": refit_line["algorithmic_gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": refit_line["frac_of_hbm_peak"],
            "traffic": dram,
            "traffic_plan_rebuilt": sum(v["mem_bytes"] for v in rk.values()),
            "traffic_meaning": "L2-miss bytes per call (served by MALL or HBM)",
            "limiter": "dependent-load and LDS-climb latency per treelet (see per kernel)",
            "per_kernel": {k: {"miss_bytes": v["mem_bytes"], "us": v["us"],
                               "miss_gbs": round(v["mem_bytes"] / (v["us"] * 1e-6) / 1e9, 1),
                               "l2_hit_pct": v["l2_hit_pct"], "valu_busy_pct": v["valu_busy_pct"]}
                           for k, v in rk.items()},
        }

    progress("standalone refit timed")
    # ---------------- f32 error of the hit coefficient and depth ----------------
    depth = torch.empty((H, W), dtype=torch.float32, device=dev)
    ws_depth = _lib.workspace(max(1, L.dmsh_pix2depth_workspace_bytes()), dev)
    _lib.check(L.dmsh_pix2depth(sp, depth.data_ptr(), pix2tri.data_ptr(), num_tri, t_tri.data_ptr(),
                                num_vtx, t_vtx.data_ptr(), W, H, t_n2w.data_ptr(),
                                ws_depth.data_ptr(), ws_depth.numel()), "pix2depth")
    torch.cuda.synchronize()
    f64err = fp64_error(tri, vtx, n2w, pix2tri.cpu().numpy(), pix2t.cpu().numpy(),
                        depth.cpu().numpy())
    del depth, ws_depth

    progress("f32 error of t and depth")
    # ---------------- secondary: 2^22-point build and sort alone ----------------
    rng = np.random.default_rng(2)
    pts = torch.from_numpy(rng.random((1 << 22, 3), dtype=np.float32)).to(dev)
    npnt = pts.shape[0]
    p_nodes = torch.empty((2 * npnt - 1, 3), dtype=torch.uint32, device=dev)
    p_aabb = torch.empty((2 * npnt - 1, 6), dtype=torch.float32, device=dev)
    ws_p = _lib.workspace(L.dmsh_build_workspace_bytes(npnt, 63), dev)
    keys0 = torch.from_numpy(rng.integers(0, 1 << 63, npnt, dtype=np.uint64)).to(dev)
    keys = torch.empty_like(keys0)
    idx = torch.empty((npnt,), dtype=torch.uint32, device=dev)
    ws_s = _lib.workspace(L.dmsh_argsort_workspace_bytes(npnt, 8), dev)

    def pbuild():
        _lib.check(L.dmsh_build_points(sp, npnt, pts.data_ptr(), 63, p_nodes.data_ptr(),
                                       p_aabb.data_ptr(), ws_p.data_ptr(), ws_p.numel()), "points")

    def psort():
        _lib.check(L.dmsh_argsort_u64(sp, npnt, keys.data_ptr(), idx.data_ptr(), 63, ws_s.data_ptr(),
                                      ws_s.numel()), "sort")

    def timed(fn, reps, prep=None):
        out = []
        for _ in range(reps):
            if prep:
                prep()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            out.append((a, b))
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in out]))

    timed(pbuild, 2)
    pbuild_ms = timed(pbuild, 10)
    timed(psort, 2, lambda: keys.copy_(keys0))
    psort_ms = timed(psort, 10, lambda: keys.copy_(keys0))

    progress("points build and sort alone timed")
    # ---------------- CPU baseline and traversal counts ----------------
    cast_counts = None
    cpu = None
    if not args.no_cpu_baseline:
        cpu, cast_counts = run_cpu_baseline(tri, vtx, n2w, W, H, bvhnodes.cpu().numpy(),
                                            aabb.cpu().numpy(), pix2tri.cpu().numpy())

    # ---------------- batched workload ----------------
    progress("cpu baseline done")
    batched = run_batched(args.batch_meshes, args.batch_reps) if args.batch_meshes > 0 else None
    progress("batched done")
    served = (run_served(args.serve_requests, [float(x) for x in args.serve_rates.split(",")],
                         args.serve_budget_mib << 20) if args.serve_requests > 0 else None)
    progress("served done")
    diff = run_diff(args.diff_reps) if args.diff_reps > 0 else None
    progress("differentiation done")
    trees = run_trees(args.tree_reps) if args.tree_reps > 0 else None
    progress("trees done")

    # ---------------- roofline of the dominant kernel ----------------
    # build: compulsory bytes per triangle with u64 codes
    build_bytes = num_tri * 12 + num_vtx * 12 + num_tri * 12 + num_node * 12 + num_node * 24
    stack = L.dmsh_get_tuning(b"pix2tri.stack")
    k1 = kernels.get("k_pix2tri")
    if cast_counts and k1:
        # ray-cast, in the reference's layout: per ray the root box (24 B) and
        # 8 B of output; per expanded internal node its row (12 B) and its two
        # child boxes (48 B); per tested leaf the triangle's indices (12 B) and
        # vertices (36 B)
        cast_bytes = num_pix * (24 + 8) + 60 * cast_counts[0] + 48 * cast_counts[1]
        complete = cast_counts[2] <= stack  # every ray finished in k_pix2tri
        k_ms = k1["avg_ms"]
        name = "k_pix2tri"
        if not complete:
            k_ms = sum(kernels[k]["avg_ms"] * kernels[k]["launches_per_step"]
                       for k in ("k_pix2tri", "k_pix2tri_tier2", "k_pix2tri_tier3") if k in kernels)
            name = "k_pix2tri + restart tiers"
        cc = read_counters("cast:k_pix2tri") if complete else None
        roof = {
            "bound": "hbm",
            "kernel": name,
            "achieved": round(cast_bytes / (k_ms * 1e-3) / 1e9, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(cast_bytes / (k_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "traffic": cc["mem_bytes"] if cc else None,
            "algorithmic_bytes": cast_bytes,
            "kernel_ms": k_ms,
        }
        if cc:
            # The algorithmic bytes are what the reference's layout would
            # read from memory; the tree and triangles stay in L2 / MALL, so
            # memory moves a small share of them and the kernel is limited by
            # issue and dependent-load latency, not by HBM bandwidth.
            roof["limiter"] = "issue / dependent-load latency (working set L2- and MALL-resident)"
            roof["frac_meaning"] = "HBM-equivalent: algorithmic bytes / time / HBM peak"
            roof["traffic_meaning"] = "L2-miss bytes per launch (served by MALL or HBM)"
            roof["miss_gbs"] = round(cc["mem_bytes"] / (k_ms * 1e-3) / 1e9, 1)
            roof["miss_frac_of_hbm_peak"] = round(cc["mem_bytes"] / (k_ms * 1e-3) / 1e9
                                                  / PEAK_HBM_GBS, 4)
            roof["miss_share_of_algorithmic"] = round(cc["mem_bytes"] / cast_bytes, 4)
            roof["l2_hit_pct"] = cc["l2_hit_pct"]
            roof["valu_busy_pct"] = cc["valu_busy_pct"]
            roof["counter_us"] = cc["us"]
    else:
        roof = {
            "bound": "hbm",
            "kernel": "dmsh_build_trimesh3 (all build kernels)",
            "achieved": round(build_bytes / (build_avg * 1e-3) / 1e9, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(build_bytes / (build_avg * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "traffic": None,
            "algorithmic_bytes": build_bytes,
            "kernel_ms": build_avg,
        }
    line = {
        "metric": "LBVH build Mtri/s + Mrays/s, batched/served vs plain kernel; p50/p99 latency",
        "value": round(build_mtri, 2),
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"single large mesh, plain kernels: octahedral sphere level {args.level} "
                        f"({num_tri} tris, radial noise N(0,0.01^2), seed 1), 63-bit Morton LBVH build "
                        f"+ refit, perspective ray-cast {W}x{H}",
            "num_tri": num_tri,
            "rays": num_pix,
        },
        "build_ms": round(build_avg, 4),
        "build_algorithmic_gbs": round(build_bytes / (build_avg * 1e-3) / 1e9, 1),
        "raycast_ms": round(cast_avg, 4),
        "raycast_mrays_per_s": round(cast_mrays, 2),
        "refit_standalone": refit_line,
        "points_build_2p22_ms": round(pbuild_ms, 4),
        "points_build_mpts_per_s": round(npnt / (pbuild_ms * 1e-3) / 1e6, 2),
        "sort_alone_u64_2p22_ms": round(psort_ms, 4),
        "sort_alone_melem_per_s": round(npnt / (psort_ms * 1e-3) / 1e6, 2),
        "fp64_error": f64err,
        "raycast_traversal": ({"internal_expansions": cast_counts[0], "leaf_tests": cast_counts[1],
                               "max_stack_frames": cast_counts[2], "lds_stack_frames": stack,
                               "waves_per_block": L.dmsh_get_tuning(b"pix2tri.waves")}
                              if cast_counts else None),
        "kernels": kernels,
        "batched": batched,
        "served": served,
        "differentiation": diff,
        "trees": trees,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
