#!/usr/bin/env python3
"""Benchmark of the MI355X sparse-voxel-DAG hot path (one GPU).

Headline (`value`): full-depth DAG build of the 1024^3 grid (heightfield with
4 octaves, amplitude 256, 2000 carved caves, depth-band materials) in
Gvoxels/s - input PyrVoxel records resident in HBM, one step = one complete
pyrhip_build (encode, radix sort, unique, bricks, leaf and node dedup,
compaction, final read-back of the sizes).

Beside it: traversal of the 4096^3 skin DAG (1920x1080 primary rays, 30
seeded poses, plus one shadow ray per hit), the batched build of 1024 uneven
chunks against one build call per chunk, the roofline of the dominant kernel
and the CPU baseline (the oracle's restatement of the reference's CPU builder
and traversal, timed on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "DAG build Gvoxels/s + traversal Mrays/s, % HBM peak; p99 chunk-build latency"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="smaller traversal/batch sections")
    ap.add_argument("--only", default="", help="comma list of sections: build,rtprims,trace,scenes,batch,serve,cpu,probe")
    ap.add_argument("--radix-bits", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="build with the built-in defaults instead of the setting the autotuner picks for the "
                         "workload class (tuning runs before the timed region)")
    ap.add_argument("--log2", type=int, default=10, help="edge of the headline grid (10 = 1024^3)")
    ap.add_argument("--serve-eager-only", action="store_true",
                    help="serving section: only the run that builds every batch eagerly (no graph replays)")
    return ap.parse_args()


def measured_traffic(kernel: str, voxels: int):
    """HBM bytes per launch of `kernel` on this workload from the committed
    rocprofv3 counter summary (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE
    scaled by the copy-probe calibration), or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    for e in data.get("kernels", []):
        if e.get("kernel") == kernel and e.get("voxels") == voxels:
            return e.get("bytes_per_launch")
    return None


def build_section(args, out):
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import Events, synchronize

    w = workloads.heightfield_grid(log2_size=args.log2)
    log(f"generating {w.name} on the device")
    vox, offs, n = w.generate_device()
    log(f"{n} voxels resident ({n * 16 / 1e9:.2f} GB)")
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = gpu.autotune(args.log2, vox, count=n, repeats=2)
        tune["seconds"] = time.perf_counter() - t0
        log(f"autotune: {tune['best']} {tune['best_ms']:.2f} ms (defaults {tune['default_ms']:.2f} ms), "
            f"{tune['candidates']} settings in {tune['seconds']:.1f} s")
    L = lib()
    t = abi.PyrhipBuildTuning()
    t.radix_bits = args.radix_bits
    built = abi.PyrhipBuilt()

    def one():
        abi.check(L, L.pyrhip_build(args.log2, 0, 0, 0, vox.ptr, n, C.byref(t), None, C.byref(built)), "build")

    for i in range(args.warmup):
        one()
        info = (built.voxel_count, built.node_words, built.leaf_count)
        L.pyrhip_built_free(C.byref(built), None)
    synchronize()
    gpu.set_timing(True)
    ev = Events()
    ev.start()
    sort_ms = []
    phases = []
    for i in range(args.steps):
        one()
        tm = gpu.last_build_times()
        phases.append(tm)
        sort_ms.append(tm["sort_ms"])
        info = (built.voxel_count, built.node_words, built.leaf_count)
        L.pyrhip_built_free(C.byref(built), None)
    ms = ev.stop() / args.steps
    synchronize()
    gpu.set_timing(False)
    passes = max(1, phases[-1]["sort_passes"])  # passes that moved records
    rb = phases[-1]["radix_bits"]
    rec = phases[-1]["record_bytes"]  # 8: packed key << 32 | attribute, 12: key + attribute arrays
    items = phases[-1]["sort_items"]
    pass_ms = float(np.mean(sort_ms)) / passes
    pass_bytes = 2.0 * rec * n  # every record read once and written once per pass
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9
    U, W, Lc = info
    out["build"] = {
        "workload": w.name, "voxels": n, "unique": U, "node_words": W, "leaves": Lc, "ms": ms,
        "gvoxels_per_s": n / (ms * 1e-3) / 1e9, "sort_passes": passes, "radix_bits": rb,
        "record_bytes": rec, "sort_items": items, "finish": phases[-1]["finish"],
        "phase_ms": {k: float(np.mean([p[k] for p in phases])) for k in
                     ("encode_ms", "sort_ms", "unique_ms", "leaves_ms", "levels_ms", "total_ms")},
        "sort_pass_ms": pass_ms,
        "autotune": tune,
        # every buffer the build requests (scratch and outputs), all held at
        # once: the build's peak device memory
        "device_bytes": phases[-1]["device_bytes"], "device_bytes_per_voxel": phases[-1]["device_bytes"] / n,
    }
    # algorithmic bytes of the whole build: encode 16N read + rec*N write, P
    # moving passes x 2*rec*N, unique (bucket tables or scan) rec*N read + 12U
    # write, outputs 4W + 8L (brick keys and masks, hash tables and level
    # tuples are small beside these)
    alg = (16.0 + rec) * n + passes * 2.0 * rec * n + rec * n + 12.0 * U + 4.0 * W + 8.0 * Lc
    out["build"]["algorithmic_bytes"] = alg
    out["build"]["algorithmic_GBps"] = alg / (ms * 1e-3) / 1e9
    narrow = phases[-1].get("status_bytes", 8) == 4
    kname = "k_sort_pass<%d,%d,%s,%s>" % (rb, items, "true" if rec == 8 else "false", "true" if narrow else "false")
    out["roofline"] = {"kernel": kname, "bound": "hbm", "achieved": achieved,
                       "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                       "traffic": measured_traffic(kname, n), "bytes_per_launch": pass_bytes,
                       "avg_launch_ms": pass_ms}
    vox.free()
    return n, ms


def rt_prims_section(args, out):
    """Box primitives of the built 1024^3 grid (rt_prims.extract as its own
    operator): wall time of the counting call plus the writing call, each of
    which ends synchronised, at the reference's default edge 2^7 and at 2^3."""
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import Buffer

    L = lib()
    vox, _, n = workloads.heightfield_grid(log2_size=args.log2).generate_device()
    built = abi.PyrhipBuilt()
    abi.check(L, L.pyrhip_build(args.log2, 0, 0, 0, vox.ptr, n, None, None, C.byref(built)), "build")
    b = gpu.Built(built)
    res = {}
    for req in (7, 3):
        count, k = C.c_uint32(), C.c_uint32()
        d = b.dag()
        abi.check(L, L.pyrhip_rt_prims(C.byref(d), b.node_words, b.leaf_count, req, 0, 0, 0, None, C.byref(count),
                                       C.byref(k)), "rt prims")
        cap = count.value
        pb, ab = Buffer(max(cap, 1) * 16), Buffer(max(cap, 1) * 24)
        times = []
        for i in range(6):
            t0 = time.perf_counter()
            abi.check(L, L.pyrhip_rt_prims(C.byref(d), b.node_words, b.leaf_count, req, 0, 0, 0, None,
                                           C.byref(count), C.byref(k)), "rt prims")
            abi.check(L, L.pyrhip_rt_prims(C.byref(d), b.node_words, b.leaf_count, req, pb.ptr, ab.ptr, cap, None,
                                           C.byref(count), C.byref(k)), "rt prims")
            times.append((time.perf_counter() - t0) * 1e3)
        res[f"k{req}"] = {"primitives": count.value, "ms": float(np.median(times[1:]))}
        pb.free()
        ab.free()
    b.free()
    vox.free()
    out["rt_prims"] = res
    log(f"rt prims: {res}")


def trace_section(args, out):
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import Buffer, Events, synchronize

    log2 = 12 if not args.quick else 10
    w = workloads.skin_grid(log2_size=log2, caves=20000 if log2 == 12 else 2000)
    log(f"generating {w.name}")
    vox, offs, n = w.generate_device()
    L = lib()
    built = abi.PyrhipBuilt()
    t0 = time.time()
    abi.check(L, L.pyrhip_build(log2, 0, 0, 0, vox.ptr, n, None, None, C.byref(built)), "build skin DAG")
    log(f"built {n} skin voxels in {time.time() - t0:.2f}s: {built.node_words} words, {built.leaf_count} leaves")
    dag = abi.PyrhipDag()
    dag.nodes, dag.leaves, dag.attrs = built.nodes, built.leaves, built.attrs
    dag.root, dag.log2_size, dag.default_attribute = built.root, log2, 1
    W, H = 1920, 1080
    poses = 30 if not args.quick else 6
    cams = workloads.camera_poses(w, poses, W, H)
    hits = Buffer(W * H * 16)
    rays = Buffer(W * H * 32)
    occ = Buffer(W * H * 4)
    sun = workloads.normalize(workloads.SUN)
    sunc = (C.c_float * 3)(*[float(v) for v in sun])
    for c in cams[:2]:  # warm-up
        abi.check(L, L.pyrhip_trace_camera(C.byref(dag), C.byref(c), hits.ptr, 0, None), "trace")
    synchronize()
    ev = Events()
    ev.start()
    for c in cams:
        abi.check(L, L.pyrhip_trace_camera(C.byref(dag), C.byref(c), hits.ptr, 0, None), "trace")
    prim_ms = ev.stop()
    # primary + shadow: rays materialised, primary from them, shadow per hit
    hit_total = 0
    ev.start()
    for c in cams:
        abi.check(L, L.pyrhip_camera_rays(C.byref(c), rays.ptr, None), "rays")
        abi.check(L, L.pyrhip_trace_dag(C.byref(dag), rays.ptr, hits.ptr, W * H, 0, None), "trace")
        abi.check(L, L.pyrhip_trace_shadow(C.byref(dag), rays.ptr, hits.ptr, sunc, occ.ptr, W * H, None),
                  "shadow")
    both_ms = ev.stop()
    h = hits.download(abi.HIT, W * H)
    hit_total = int((h["instance"] != 0xFFFFFFFF).sum())
    nrays = W * H * poses
    out["traversal"] = {
        "workload": w.name, "dag_voxels": built.voxel_count, "dag_node_words": built.node_words,
        "dag_leaves": built.leaf_count, "resolution": [W, H], "poses": poses,
        "primary_ms_per_frame": prim_ms / poses, "primary_grays_per_s": nrays / (prim_ms * 1e-3) / 1e9,
        "primary_plus_shadow_ms_per_frame": both_ms / poses,
        "primary_plus_shadow_mrays_per_s": (nrays + hit_total * poses) / (both_ms * 1e-3) / 1e6,
        "last_frame_hit_fraction": hit_total / (W * H),
    }
    out["traversal"]["render"] = render_leg(args, log2, vox, n, cams, W, H, hits)
    vox.free()
    if args.cpu_trace:
        cb = cpu_trace_baseline(args, built, cams[0], W, H)
        out["traversal"]["parity_mismatches"] = cb["parity"]["mismatches"]
        out["traversal"]["parity"] = cb.pop("parity")
        out["traversal"]["cpu_baseline"] = cb
        # traversal roofline: 32 B ray + 16 B hit + 4 B per distinct node word,
        # 8 B per distinct leaf and 4 B per attribute a ray reads, the distinct
        # counts from the instrumented CPU trace of the same DAG (pose 0, half
        # resolution); the primary kernel reads no ray (camera rays are made
        # in registers), so its bytes are the hit plus the DAG reads
        tc = cb["touch"]
        per_ray = 16.0 + (4.0 * tc["node_words"] + 8.0 * tc["leaves"] + 4.0 * tc["attributes"]) / tc["rays"]
        achieved = out["traversal"]["primary_grays_per_s"] * per_ray
        out["traversal"]["roofline"] = {
            "kernel": "k_trace_camera", "bound": "hbm", "bytes_per_ray": per_ray, "achieved": achieved,
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
            "note": "distinct DAG bytes per ray; almost all of them hit the per-CU L1, and the kernel is bound "
                    "by vector instruction issue (VALU busy most of its cycles), not by these bytes"}
        log(f"traversal: {per_ray:.0f} algorithmic bytes per primary ray, {achieved:.0f} GB/s")
    L.pyrhip_built_free(C.byref(built), None)
    for b in (hits, rays, occ):
        b.free()
    return dag, w, cams


def render_leg(args, log2, vox, n, cams, W, H, cam_hits):
    """The same DAG through the reference's C ABI: pyr_geometry_build from the
    resident records, one instance, pyr_commit, then pyr_render (hits + depth)
    per pose. pyr_render returns synchronised, so the time is wall clock per
    call. The last pose's hits must equal the camera kernel's bit for bit."""
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer

    L = lib()
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    info.node_pool_bytes = info.leaf_pool_bytes = 256 << 20
    info.attribute_pool_bytes = max(256 << 20, 4 * n + (64 << 20))
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    try:
        geo, inst, view = C.c_void_p(), C.c_void_p(), C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, log2, vox.ptr, n, 0, C.byref(geo)), "geometry build")
        abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        abi.check(L, L.pyr_view_create(ctx, C.byref(view)), "view")
        hb, db = Buffer(W * H * 16), Buffer(W * H * 4)
        t = abi.PyrTargets()
        t.hits, t.depth = hb.ptr, db.ptr
        for c in cams[:2]:
            abi.check(L, L.pyr_render(ctx, view, C.byref(c), C.byref(t)), "render")
        t0 = time.perf_counter()
        for c in cams:
            abi.check(L, L.pyr_render(ctx, view, C.byref(c), C.byref(t)), "render")
        ms = (time.perf_counter() - t0) * 1e3 / len(cams)
        same = bool(np.array_equal(hb.download(np.uint32, W * H * 4), cam_hits.download(np.uint32, W * H * 4)))
        hb.free()
        db.free()
    finally:
        L.pyr_destroy(ctx)
    log(f"pyr_render {ms:.3f} ms/frame (hits + depth), hits equal to the camera kernel: {same}")
    return {"ms_per_frame": ms, "grays_per_s": W * H / (ms * 1e-3) / 1e9, "hits_equal_camera_kernel": same,
            "timing": "wall clock per synchronous pyr_render call (hits + depth targets)"}


# published frame times of the reference's CUDA traversal for the same scenes
# (RTX 5070 Laptop, 1920x1080 primary rays + depth + motion vectors,
# pyrit-plan.md:37-41): instances -> ms per frame
PUBLISHED_INSTANCE_MS = {3: 0.83, 64: 4.14, 1000: 31.6}


def scenes_section(args, out):
    """The reference's instance benchmark (tests/gpu_test.zig:394-470): the
    128^3 test scene as one geometry, 3, 64 and 1000 instances on a grid,
    pyr_commit + pyr_render (hits + depth) per frame at 1920x1080; GPU time
    by HIP events on the context's stream over 40 frames after 5."""
    from types import SimpleNamespace

    from pyrit_amd import abi, lib, workloads
    from pyrit_amd.device import Buffer, Events

    L = lib()
    grid = workloads.scene_grid(128)
    res = {}
    for count in (3, 64, 1000):
        info = abi.PyrCreateInfo()
        info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
        info.max_instances, info.max_geometries = 4096, 16
        info.node_pool_bytes = info.leaf_pool_bytes = info.attribute_pool_bytes = 64 << 20
        ctx = C.c_void_p()
        abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
        try:
            dag, geo, view = C.c_void_p(), C.c_void_p(), C.c_void_p()
            abi.check(L, L.pyr_dag_build_dense(7, grid.ctypes.data_as(C.c_void_p), 0, C.byref(dag)), "dag")
            abi.check(L, L.pyr_geometry_create(ctx, dag, C.byref(geo)), "geometry")
            L.pyr_dag_destroy(dag)
            xf, eye, target = workloads.instance_grid(count)
            for m in xf:
                inst = C.c_void_p()
                abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
                abi.check(L, L.pyr_instance_set_transform(ctx, inst, (C.c_float * 12)(*m)), "transform")
            abi.check(L, L.pyr_view_create(ctx, C.byref(view)), "view")
            cam = abi.PyrCamera()
            L.pyr_camera_look_at(C.byref(cam), (C.c_float * 3)(*eye), (C.c_float * 3)(*target),
                                 (C.c_float * 3)(0.0, 1.0, 0.0))
            L.pyr_camera_perspective(C.byref(cam), 0.9, 1920, 1080, 0.1)
            hb, db = Buffer(1920 * 1080 * 16), Buffer(1920 * 1080 * 4)
            t = abi.PyrTargets()
            t.hits, t.depth = hb.ptr, db.ptr
            abi.check(L, L.pyr_commit(ctx, None), "commit")
            ev = Events(SimpleNamespace(handle=L.pyr_cuda_stream(ctx)))
            frames = 40
            for f in range(frames + 5):
                if f == 5:
                    ev.start()
                abi.check(L, L.pyr_commit(ctx, None), "commit")
                abi.check(L, L.pyr_render(ctx, view, C.byref(cam), C.byref(t)), "render")
            ms = ev.stop() / frames
            hits = hb.download(abi.HIT, 1920 * 1080)
            ev.destroy()
            hb.free()
            db.free()
        finally:
            L.pyr_destroy(ctx)
        res[str(count)] = {"ms_per_frame": ms, "grays_per_s": 1920 * 1080 / (ms * 1e-3) / 1e9,
                           "hit_fraction": float((hits["instance"] != 0xFFFFFFFF).mean()),
                           "published_ms_rtx5070_laptop": PUBLISHED_INSTANCE_MS[count],
                           "speedup_vs_published": PUBLISHED_INSTANCE_MS[count] / ms}
        log(f"scene of {count} instances: {ms:.3f} ms/frame ({res[str(count)]['hit_fraction']:.2f} hit), "
            f"published {PUBLISHED_INSTANCE_MS[count]} ms")
    out["scenes"] = {"workload": "128^3 test scene, instance grids of tests/gpu_test.zig:394-470, 1920x1080",
                     "timing": "HIP events over 40 frames of pyr_commit + pyr_render (hits + depth)",
                     "published": "RTX 5070 Laptop, CUDA software traversal, primary rays + depth + motion vectors",
                     "instances": res}


def _dag_of(built):
    from pyrit_amd import abi

    dag = abi.PyrhipDag()
    dag.nodes, dag.leaves, dag.attrs = built.nodes, built.leaves, built.attrs
    dag.root, dag.log2_size, dag.default_attribute = built.root, built.log2_size, 1
    return dag


def cpu_trace_baseline(args, built, cam, W, H):
    """The oracle's restatement of the reference traversal (walk + traceScene)
    on the same DAG (downloaded from the GPU build) and the first pose's primary
    rays at half resolution, on 1 host core and on every host core (rays are
    independent, as in the reference's CPU tests)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    from pyrit_amd.device import download

    nodes = download(built.nodes, np.uint32, built.node_words)
    leaves = download(built.leaves, np.uint64, built.leaf_count)
    attrs = download(built.attrs, np.uint32, built.voxel_count)
    dag = O.Dag(built.log2_size, built.root, nodes, leaves, attrs, built.voxel_count)
    scene = O.single_scene(dag)
    oc = O.OrCamera.from_buffer_copy(bytes(cam))
    oc.width, oc.height = W // 2, H // 2  # same aspect: the projection scale is unchanged
    rays = O.camera_rays(oc)
    res = {"sample": f"pose 0 primary rays at {W // 2}x{H // 2} ({len(rays)} rays), same DAG", "unit": "Mrays/s",
           "kind": "port"}
    for label, threads in (("one_core", 1), ("all_cores", os.cpu_count() or 1)):
        t0 = time.perf_counter()
        want = O.trace_rays(scene, rays, threads=threads)
        dt = time.perf_counter() - t0
        res[label] = {"value": len(rays) / dt / 1e6, "cores": threads, "seconds": dt}
    # parity of the timed kernel on the same rays: the camera kernel traces
    # the half-resolution pose and its hits are compared with the oracle's
    # (instance, attribute, face/inside bits and fp32 t, bit for bit)
    from pyrit_amd import abi, gpu

    hc = abi.PyrCamera.from_buffer_copy(bytes(oc))
    got = gpu.trace_camera(_dag_of(built), hc)
    gw, ww = got.view(np.uint32).reshape(-1, 4), want.view(np.uint32).reshape(-1, 4)
    res["parity"] = {"rays": int(len(rays)), "mismatches": int(np.count_nonzero((gw != ww).any(axis=1))),
                     "hits": int(np.count_nonzero(want["instance"] != 0xFFFFFFFF)),
                     "check": "k_trace_camera hits vs the oracle's traceScene on the same DAG, bit for bit"}
    # the same rays once more, instrumented: distinct words each ray reads
    res["touch"] = O.trace_touch(scene, rays)
    log(f"cpu traversal {res['one_core']['value']:.2f} Mrays/s (1 core), "
        f"{res['all_cores']['value']:.1f} Mrays/s ({res['all_cores']['cores']} cores)")
    return res


def batch_section(args, out):
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import Buffer, Events, synchronize, to_device

    count = 1024 if not args.quick else 128
    w = workloads.uneven_chunks(count=count)
    log(f"generating {w.name}")
    vox, offs, n = w.generate_device()
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    L = lib()
    d_offs = to_device(offs.astype(np.uint32))
    d_sizes = to_device(sizes)
    batch = abi.PyrhipChunkBatch()
    batch.count, batch.capacity, batch.voxels, batch.offsets = count, 0, vox.ptr, d_offs.ptr
    batch.total, batch.chunk_log2 = n, d_sizes.ptr
    roots = np.zeros(count, np.uint32)
    first = np.zeros(count, np.uint32)
    built = abi.PyrhipBuilt()
    top = int(sizes.max())
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tune = gpu.autotune(top, vox, offsets=offs.astype(np.uint32), chunk_log2=sizes, repeats=2)
        tune["seconds"] = time.perf_counter() - t0
        log(f"batch autotune: {tune['best']} {tune['best_ms']:.2f} ms (defaults {tune['default_ms']:.2f} ms)")

    def batched():
        abi.check(L, L.pyrhip_build_chunks(top, C.byref(batch), None, None, C.byref(built),
                                           roots.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           first.ctypes.data_as(C.POINTER(C.c_uint32))), "batch")
        L.pyrhip_built_free(C.byref(built), None)

    batched()
    synchronize()
    ev = Events()
    reps = 3
    ev.start()
    for _ in range(reps):
        batched()
    b_ms = ev.stop() / reps
    # palette packing of the batch's attributes, as a world installs a batch
    # (16 entries + 4-bit indices per chunk with at most 16 distinct values);
    # this build also reports the phases for the batch's algorithmic bytes
    gpu.set_timing(True)
    abi.check(L, L.pyrhip_build_chunks(top, C.byref(batch), None, None, C.byref(built),
                                       roots.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       first.ctypes.data_as(C.POINTER(C.c_uint32))), "batch")
    bt = gpu.last_build_times()
    gpu.set_timing(False)
    # algorithmic bytes as for the single build: encode 16N read + rN write,
    # P moving passes x 2rN, unique rN read + 12U write, outputs 4W + 8L
    rec, passes = bt["record_bytes"], max(1, bt["sort_passes"])
    U, W, Lc = built.voxel_count, built.node_words, built.leaf_count
    b_alg = (16.0 + rec) * n + passes * 2.0 * rec * n + rec * n + 12.0 * U + 4.0 * W + 8.0 * Lc
    nonempty = roots != 0xFFFFFFFF
    pfirst = np.ascontiguousarray(np.append(first[nonempty], built.voxel_count), np.uint32)
    k = len(pfirst) - 1
    woff = np.zeros(k + 1, np.uint64)
    pcount = np.zeros(k, np.uint32)
    P32, P64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)

    def pack(out_ptr, words):
        abi.check(L, L.pyrhip_palette_pack(built.attrs, pfirst.ctypes.data_as(P32), k, out_ptr, words,
                                           woff.ctypes.data_as(P64), pcount.ctypes.data_as(P32), None), "palette")

    pack(0, 0)
    words = int(woff[-1])
    pal_out = Buffer(max(words, 1) * 4)
    pack(pal_out.ptr, words)
    synchronize()
    ev.start()
    for _ in range(reps):
        pack(pal_out.ptr, words)
    pal_ms = ev.stop() / reps
    U = int(built.voxel_count)
    L.pyrhip_built_free(C.byref(built), None)
    pal_out.free()
    palette_info = {"chunks": k, "palette_chunks": int((pcount != 0xFFFFFFFF).sum()), "attributes": U,
                    "words": words, "bytes_saved_fraction": 1.0 - words / max(U, 1),
                    "ms": pal_ms, "note": "scan + pack incl. the count read-back; algorithmic bytes 2 x 4U read + "
                                          "4 x words write",
                    "GBps": (8.0 * U + 4.0 * words) / (pal_ms * 1e-3) / 1e9}
    graph_info = graph_measure(top, vox, offs, sizes)
    # one build call per chunk on the same GPU (same inputs)
    lat = []
    ev.start()
    for c in range(count):
        t0 = time.perf_counter()
        cnt = int(offs[c + 1] - offs[c])
        abi.check(L, L.pyrhip_build(int(sizes[c]), 0, 0, 0, vox.ptr + int(offs[c]) * 16, cnt, None, None,
                                    C.byref(built)), "chunk")
        L.pyrhip_built_free(C.byref(built), None)
        lat.append((time.perf_counter() - t0) * 1e3)
    p_ms = ev.stop()
    out["batched"] = {
        "workload": w.name, "chunks": count, "voxels": n, "batched_ms": b_ms,
        "batched_gvoxels_per_s": n / (b_ms * 1e-3) / 1e9, "per_chunk_ms": p_ms,
        "per_chunk_gvoxels_per_s": n / (p_ms * 1e-3) / 1e9, "speedup_vs_per_chunk": p_ms / b_ms,
        "sort_passes": passes, "record_bytes": rec, "finish": bt["finish"], "unique": U, "algorithmic_bytes": b_alg,
        "device_bytes": bt["device_bytes"], "device_bytes_per_voxel": bt["device_bytes"] / max(1, n),
        "algorithmic_GBps": b_alg / (b_ms * 1e-3) / 1e9, "hbm_frac": b_alg / (b_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
        "phase_ms": {k: bt[k] for k in ("encode_ms", "sort_ms", "unique_ms", "leaves_ms", "levels_ms", "total_ms")},
        "per_chunk_latency_ms_p50": float(np.percentile(lat, 50)),
        "per_chunk_latency_ms_p99": float(np.percentile(lat, 99)),
        "autotune": tune,
        "palette": palette_info,
        "graph": graph_info,
    }
    for b in (vox, d_offs, d_sizes):
        b.free()


def graph_measure(top, vox, offs, sizes):
    """The batch build replayed from a HIP graph (captured once, buffers
    allocated once, inputs already in the graph's input buffer) against the
    eager batch build with its inputs on the device: the whole batch, its
    first 16 chunks, and 16 edit-sized chunks where launch overhead counts."""
    from pyrit_amd import abi, gpu, lib
    from pyrit_amd.device import hip, synchronize, to_device

    L = lib()
    # 16 chunks of 32^3 holding 4096 scattered voxels each
    rng = np.random.default_rng(12)
    small = np.zeros(16 * 4096, abi.VOXEL)
    for f in ("x", "y", "z"):
        small[f] = rng.integers(0, 32, len(small))
    small["attribute"] = rng.integers(1, 200, len(small))
    small_dev = to_device(small)
    small_offs = (np.arange(17) * 4096).astype(np.uint32)
    res = {}
    cases = (("full", top, vox, offs, sizes, len(sizes)), ("16_chunks", top, vox, offs, sizes, 16),
             ("16_chunks_of_4096_voxels", 5, small_dev, small_offs, None, 16))
    for label, l2, src, src_offs, src_sizes, k in cases:
        o = np.ascontiguousarray(src_offs[:k + 1], np.uint32)
        total = int(o[-1])
        sz = np.ascontiguousarray(src_sizes[:k], np.uint8) if src_sizes is not None else None
        d_o = to_device(o)
        d_sz = to_device(sz) if sz is not None else None
        batch = abi.PyrhipChunkBatch()
        batch.count, batch.capacity, batch.voxels, batch.offsets = k, 0, src.ptr, d_o.ptr
        batch.total, batch.chunk_log2 = total, d_sz.ptr if d_sz else 0
        roots_e = np.zeros(k, np.uint32)
        first_e = np.zeros(k, np.uint32)
        built = abi.PyrhipBuilt()

        def eager():
            abi.check(L, L.pyrhip_build_chunks(l2, C.byref(batch), None, None, C.byref(built),
                                               roots_e.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               first_e.ctypes.data_as(C.POINTER(C.c_uint32))), "batch")

        g = gpu.BuildGraph(l2, k, total, mixed_sizes=sz is not None)
        try:
            # inputs resident in the graph's buffer, as a loop producing them in place
            if hip().hipMemcpy(C.c_void_p(g.voxels_ptr()), C.c_void_p(src.ptr), C.c_size_t(total * 16), 3) != 0:
                raise RuntimeError("hipMemcpy")
            g.launch(None, o, sz)
            reps = 3 if k > 16 else 50
            synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                out_g, roots_g, first_g = g.launch(None, o, sz)
            g_ms = (time.perf_counter() - t0) * 1e3 / reps
            eager()
            same = (built.root == out_g.root and built.node_words == out_g.node_words
                    and built.leaf_count == out_g.leaf_count and np.array_equal(roots_e, roots_g)
                    and np.array_equal(first_e, first_g))
            L.pyrhip_built_free(C.byref(built), None)
            synchronize()
            # both end in the counter read-back the host waits for
            t0 = time.perf_counter()
            for _ in range(reps):
                eager()
                L.pyrhip_built_free(C.byref(built), None)
            synchronize()
            e_ms = (time.perf_counter() - t0) * 1e3 / reps
        finally:
            g.destroy()
            d_o.free()
            if d_sz:
                d_sz.free()
        res[label] = {"chunks": k, "voxels": total, "graph_ms": g_ms, "eager_ms": e_ms,
                      "graph_gvoxels_per_s": total / (g_ms * 1e-3) / 1e9, "speedup": e_ms / g_ms,
                      "same_output": bool(same)}
        log(f"graph {label}: {g_ms:.3f} ms vs eager {e_ms:.3f} ms ({k} chunks, {total} voxels), same={same}")
    small_dev.free()
    return res


def serve_section(args, out):
    """Poisson arrivals (mean 2000 requests/s) of chunk builds with the uneven
    sizes of the batched workload, mixed 80/20 with edits of 1..512 voxels of a
    resident chunk; 64 GB device budget; latency = submit -> DAG resident.
    Run with the server's graph replays of small batches (its default) and
    again with every batch built eagerly."""
    from pyrit_amd import serving, workloads

    count = 1024 if not args.quick else 128
    w = workloads.uneven_chunks(count=count, seed=31)
    log(f"serving: generating {w.name} inputs")
    vox, offs, n = w.generate_device()
    inputs = [(w.chunks[c].log2_size, vox.ptr + int(offs[c]) * 16, int(offs[c + 1] - offs[c])) for c in range(count)]
    duration = 3.0 if not args.quick else 1.0

    def run(graphs, rate=2000.0, seconds=duration):
        sv = serving.Server(budget=64 << 30, workers=2, graphs=graphs)
        # warm the pool and the kernels (and the graphs) with a few batches
        for c in range(8):
            sv.build(c, *inputs[c])
        sv.drain()
        for c in range(8):
            sv.build(c, *inputs[c])
            sv.drain()
        sv.poll()
        res, elapsed, n_req = serving.poisson_run(sv, inputs, rate=rate, duration=seconds, seed=7)
        st = sv.stats()
        sv.close()
        lat = np.array([(c.done_s - c.submit_s) * 1e3 for c in res if c.result == 0])
        builds = [c for c in res if c.user == 0 and c.result == 0]
        edits = [c for c in res if c.user == 1 and c.result == 0]
        bl = np.array([(c.done_s - c.submit_s) * 1e3 for c in builds]) if builds else np.zeros(1)
        el = np.array([(c.done_s - c.submit_s) * 1e3 for c in edits]) if edits else np.zeros(1)
        sizes = np.array([c.batch_chunks for c in res if c.result == 0])
        return {
            "requests": n_req, "completed": len(res), "rejected": int(sum(1 for c in res if c.result != 0)),
            "elapsed_s": elapsed, "builds_per_s": len(builds) / elapsed, "requests_per_s": len(res) / elapsed,
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "build_latency_ms_p50": float(np.percentile(bl, 50)),
            "build_latency_ms_p99": float(np.percentile(bl, 99)),
            "edit_latency_ms_p50": float(np.percentile(el, 50)), "edit_latency_ms_p99": float(np.percentile(el, 99)),
            "batches": st["batches"], "mean_batch_chunks": float(sizes.mean()) if len(sizes) else 0.0,
            "device_busy_fraction": st["busy_s"] / elapsed, "voxels_built": st["batched_voxels"],
            "resident_chunks": st["resident_chunks"], "resident_GB": st["resident_bytes"] / 1e9,
            "offered_rate": rate,
        }

    if args.serve_eager_only:
        eager = run(False)
        eager["workload"] = (f"Poisson 2000 req/s, 80% builds ({w.name} sizes) / 20% edits of 1..512 voxels, "
                             "64 GB budget; every batch built eagerly")
        out["serving"] = eager
        vox.free()
        return
    res = run(True)
    eager = run(False)
    # the same stream at higher arrival rates: batches grow as requests wait
    # for a free worker, which is what keeps up with the load
    sweep = []
    for rate in (4000.0, 8000.0, 16000.0):
        r = run(True, rate=rate, seconds=min(duration, 2.0))
        sweep.append({k: r[k] for k in ("offered_rate", "requests", "rejected", "builds_per_s", "requests_per_s",
                                         "mean_batch_chunks", "latency_ms_p50", "latency_ms_p99",
                                         "device_busy_fraction", "voxels_built", "elapsed_s")})
        log(f"serving at {rate:.0f}/s: {r['requests_per_s']:.0f} req/s, mean batch {r['mean_batch_chunks']:.2f}, "
            f"p50 {r['latency_ms_p50']:.2f} ms p99 {r['latency_ms_p99']:.2f} ms, busy {r['device_busy_fraction']:.2f}")
    res["load_sweep"] = [{k: res[k] for k in sweep[0]}] + sweep
    res["workload"] = (f"Poisson 2000 req/s, 80% builds ({w.name} sizes) / 20% edits of 1..512 voxels, 64 GB budget; "
                       "small batches replay HIP graphs")
    res["eager"] = eager
    out["serving"] = res
    log(f"serving eager: p50 {eager['latency_ms_p50']:.2f} ms p99 {eager['latency_ms_p99']:.2f} ms, "
        f"{eager['builds_per_s']:.0f} builds/s")
    vox.free()

def cpu_section(args, out):
    """The oracle's restatement of dag_builder.buildPoints (single-threaded by
    construction in the reference): one build of the benchmark grid itself
    (the 1024^3 records, made on the device and copied to the host), or with
    --quick a few builds of its 256^3 instance."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    from pyrit_amd import workloads

    if args.quick:
        log2 = 8
        vox, _ = workloads.heightfield_grid(log2_size=8, caves=2000 // 64).generate_host()
        what = "the 256^3 instance of the 1024^3 scheme"
    else:
        log2 = args.log2
        dv, _, n = workloads.heightfield_grid(log2_size=log2).generate_device()
        vox = dv.download(O.VOXEL, n)
        dv.free()
        what = f"the {1 << log2}^3 benchmark grid itself"
    reps, t_total, nvox = 0, 0.0, 0
    t_end = time.time() + 3.0
    while reps == 0 or time.time() < t_end:
        t0 = time.perf_counter()
        O.build_points(log2, vox)
        t_total += time.perf_counter() - t0
        nvox += len(vox)
        reps += 1
    del vox
    cores = 1
    out["cpu_baseline"] = {"value": nvox / t_total / 1e9, "unit": "Gvoxels/s", "cores": cores, "kind": "port",
                           "sample": f"oracle buildPoints restatement on {what} ({nvox // reps} voxels) x{reps}, "
                                     f"{t_total:.1f}s",
                           "host_nproc": os.cpu_count()}
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
        out["cpu_baseline"]["host_cpu"] = model
    except Exception:
        pass


def probe_section(args, out):
    """Copy-kernel calibration of the HBM counters (also a bandwidth probe)."""
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, Events, synchronize

    L = lib()
    nbytes = 4 << 30
    a, b = Buffer(nbytes), Buffer(nbytes)
    abi.check(L, L.pyrhip_copy_probe(a.ptr, b.ptr, nbytes, None), "probe")
    synchronize()
    ev = Events()
    ev.start()
    for _ in range(5):
        abi.check(L, L.pyrhip_copy_probe(a.ptr, b.ptr, nbytes, None), "probe")
    ms = ev.stop() / 5
    out["copy_probe"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms,
                         "GBps": 2 * nbytes / (ms * 1e-3) / 1e9}
    a.free()
    b.free()


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only (the build and traversal kernels of one device); run with --gpus 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2
    rank = int(os.environ.get("RANK", "0"))
    from pyrit_amd import device

    if not device.gpu_available():
        print(json.dumps({"error": "no HIP device visible"}), flush=True)
        return 1
    only = set(args.only.split(",")) if args.only else {"build", "rtprims", "trace", "scenes", "batch", "serve", "cpu",
                                                         "probe"}
    args.cpu_trace = "cpu" in only
    out = {}
    n, ms = 0, float("nan")
    if "probe" in only:
        probe_section(args, out)
        log(f"copy probe {out['copy_probe']['GBps']:.0f} GB/s")
    if "build" in only:
        n, ms = build_section(args, out)
        log(f"build {out['build']['gvoxels_per_s']:.2f} Gvox/s, sort pass {out['roofline']['achieved']:.0f} GB/s, "
            f"finish {out['build']['finish']}, phases {out['build']['phase_ms']}")
    if "rtprims" in only:
        rt_prims_section(args, out)
    if "batch" in only:
        batch_section(args, out)
        log(f"batched {out['batched']['batched_gvoxels_per_s']:.2f} Gvox/s, "
            f"x{out['batched']['speedup_vs_per_chunk']:.1f} vs per chunk")
        pi = out["batched"]["palette"]
        log(f"palette: {pi['palette_chunks']}/{pi['chunks']} chunks packed, {100 * pi['bytes_saved_fraction']:.0f} % "
            f"of attribute bytes saved, {pi['ms']:.2f} ms ({pi['GBps']:.0f} GB/s)")
    if "serve" in only:
        serve_section(args, out)
        log(f"serving p50 {out['serving']['latency_ms_p50']:.2f} ms p99 {out['serving']['latency_ms_p99']:.2f} ms, "
            f"{out['serving']['builds_per_s']:.0f} builds/s")
    if "trace" in only:
        trace_section(args, out)
        log(f"primary {out['traversal']['primary_grays_per_s']:.2f} Grays/s")
    if "scenes" in only:
        scenes_section(args, out)
    if "cpu" in only and rank == 0:
        cpu_section(args, out)
    res = {
        "metric": METRIC, "value": out.get("build", {}).get("gvoxels_per_s"), "unit": "Gvoxels/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "1024^3 heightfield grid (4 octaves, amplitude 256, 2000 caves), full-depth build",
                   "voxels": n, "log2_size": args.log2, "parallelism": "single GPU"},
    }
    for k in ("roofline", "cpu_baseline", "build", "rt_prims", "traversal", "scenes", "batched", "serving",
              "copy_probe"):
        if k in out:
            res[k] = out[k]
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
