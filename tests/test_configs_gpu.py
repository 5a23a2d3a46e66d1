"""Parity at the benchmark workloads' own shapes: the 1024^3 heightfield grid
(and its 256^3 instance at default tuning), batches of uneven chunks with the
real edges {32, 64, 128, 256}, the skin terrain traversal at 2048^3 and
4096^3, and the Poisson 80/20 request stream against the build server.

Builds: GPU-native arrays bit-exact against the serial restatement of the
reference's GPU build (src/gpu_build.zig:135-360) where the oracle finishes
in seconds, and canon(GPU) == the reference CPU builder (src/dag_builder.zig)
at every size. Traversal: identical hit instance, attribute and face bits,
fp32 t within 1e-5 relative (the kernels reach bit equality, which is what
is asserted)."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import final_points

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _same_build(b, ref):
    assert b.voxel_count == ref["voxel_count"] and b.root == ref["root"]
    assert np.array_equal(b.nodes(), ref["nodes"])
    assert np.array_equal(b.leaves(), ref["leaves"])
    assert np.array_equal(b.attrs(), ref["attrs"])
    assert np.array_equal(b.keys(), ref["keys"])


def _same_as_builder(b, cpu, log2):
    assert O.canon(b.nodes(), b.leaves(), b.root, log2).same(cpu)
    assert np.array_equal(b.attrs(), cpu.attrs)


def _bits(h):
    return h.view(np.uint32).reshape(len(h), 4)


def test_heightfield_256_default_tuning(hip_device):
    """The 1024^3 scheme at 256^3: ~8M records, 1024 buckets of ~7.8k records,
    so the default build takes the bucket finish with 32-bit look-back words,
    as the full grid does."""
    from pyrit_amd import gpu, workloads

    _say("case: heightfield 256^3 at default tuning")
    w = workloads.heightfield_grid(log2_size=8)
    vox, _ = w.generate_host()
    ref = O.gb_build(8, vox)
    b = gpu.build(8, vox)
    _same_build(b, ref)
    _same_as_builder(b, O.build_points(8, vox), 8)
    gpu.set_timing(True)
    try:
        b2 = gpu.build(8, vox)
        t = gpu.last_build_times()
    finally:
        gpu.set_timing(False)
    assert t["finish"] == 2 and t["status_bytes"] == 4, t
    _same_build(b2, ref)


def test_heightfield_1024_full_size(hip_device):
    """The benchmark grid itself (1024^3, 2000 caves, ~5.1e8 records) built
    from device-generated records, against the CPU builder on the same records."""
    from pyrit_amd import gpu, workloads

    _say("case: heightfield 1024^3 full size")
    w = workloads.heightfield_grid(log2_size=10)
    vox, _, n = w.generate_device()
    b = gpu.build(10, vox, count=n)
    host = vox.download(O.VOXEL, n)
    vox.free()
    _say(f"  {n} records, GPU {b.node_words} words {b.leaf_count} leaves; CPU builder")
    cpu = O.build_points(10, host)
    del host
    assert b.voxel_count == cpu.voxel_count
    _same_as_builder(b, cpu, 10)


def _uneven(count, seed):
    from pyrit_amd import workloads

    w = workloads.uneven_chunks(seed=seed, count=count)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    return w, sizes


def test_uneven_chunks_real_edges_bit_exact(hip_device):
    """64 chunks of edges {32, 64, 128, 256} with p = {.4, .3, .2, .1}, one
    batch at default tuning: arrays, roots and first voxels bit-exact against
    the serial batch build; every chunk equal to its own CPU build (batch ==
    single, cpu_test.zig:1001-1060)."""
    from pyrit_amd import gpu

    _say("case: 64 uneven chunks, edges 32..256")
    w, sizes = _uneven(64, 3)
    assert set(sizes.tolist()) == {5, 6, 7, 8}
    vox, offs64 = w.generate_host()
    offs = offs64.astype(np.uint32)
    ref, rroots, rfirst = O.gb_build_chunks(8, vox, offs, chunk_log2=sizes)
    b, roots, first = gpu.build_chunks(8, vox, offs, chunk_log2=sizes)
    _same_build(b, ref)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    nodes, leaves, attrs = b.nodes(), b.leaves(), b.attrs()
    for c in range(len(sizes)):
        seg = vox[offs[c]:offs[c + 1]]
        cpu = O.build_points(int(sizes[c]), seg)
        if cpu.voxel_count == 0:
            assert roots[c] == 0xFFFFFFFF
            continue
        assert O.canon(nodes, leaves, int(roots[c]), int(sizes[c])).same(cpu), c
        assert np.array_equal(attrs[first[c]:first[c] + cpu.voxel_count], cpu.attrs), c


def test_uneven_chunks_1024_full_batch(hip_device):
    """The benchmark batch itself (1024 chunks, ~9.5e8 records) in one launch
    sequence: every chunk equal to a single GPU build of its records (batch ==
    single), and every 16th chunk equal to the CPU builder."""
    from pyrit_amd import gpu

    _say("case: 1024 uneven chunks, full batch")
    w, sizes = _uneven(1024, 3)
    vox, offs64, n = w.generate_device()
    offs = offs64.astype(np.uint32)
    b, roots, first = gpu.build_chunks(8, vox, offs, chunk_log2=sizes)
    nodes, leaves, attrs = b.nodes(), b.leaves(), b.attrs()
    b.free()
    _say(f"  {n} records, {len(nodes)} words, {len(leaves)} leaves")
    for c in range(len(sizes)):
        l2, lo, hi = int(sizes[c]), int(offs[c]), int(offs[c + 1])
        single = gpu.build(l2, vox.ptr + lo * 16, count=hi - lo)
        if single.voxel_count == 0:
            assert roots[c] == 0xFFFFFFFF
            continue
        got = O.canon(nodes, leaves, int(roots[c]), l2)
        want = O.canon(single.nodes(), single.leaves(), single.root, l2)
        assert got.same(want), c
        ga = attrs[first[c]:first[c] + single.voxel_count]
        assert np.array_equal(ga, single.attrs()), c
        if c % 16 == 0:
            cpu = O.build_points(l2, O_download(vox, lo, hi))
            assert got.same(cpu) and np.array_equal(ga, cpu.attrs), c
        single.free()
    vox.free()


def O_download(buf, lo, hi):
    from pyrit_amd.device import download

    return download(buf.ptr + lo * 16, O.VOXEL, hi - lo)


def _skin(log2, caves):
    from pyrit_amd import gpu, workloads

    w = workloads.skin_grid(log2_size=log2, caves=caves)
    vox, _, n = w.generate_device()
    b = gpu.build(log2, vox, count=n)
    host = vox.download(O.VOXEL, n)
    vox.free()
    return w, b, host


def test_skin_2048_build_and_traversal(hip_device):
    """Skin terrain at 2048^3 (the 4096^3 scheme at half scale): the build
    bit-exact against the serial GPU build and the CPU builder; a seeded pose
    at 960x540 traced by the camera kernel, by pyr_trace and pyr_render on a
    one-instance scene, and one shadow ray per hit, all identical to the
    oracle's traversal of the CPU builder's DAG."""
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import Buffer, to_device

    _say("case: skin 2048^3 build")
    w, b, host = _skin(11, 5000)
    ref = O.gb_build(11, host)
    _same_build(b, ref)
    cpu = O.build_points(11, host)
    _same_as_builder(b, cpu, 11)
    scene = O.single_scene(cpu)
    cam = workloads.camera_poses(w, 3, 960, 540)[1]
    oc = O.OrCamera.from_buffer_copy(bytes(cam))
    rays = O.camera_rays(oc)
    want = O.trace_rays(scene, rays, threads=16)
    assert (want["instance"] == 0).sum() > len(rays) // 4

    _say("case: skin 2048^3 camera kernel")
    got = gpu.trace_camera(b.dag(), cam)
    assert np.array_equal(_bits(got), _bits(want))

    _say("case: skin 2048^3 shadow rays")
    sun = workloads.normalize(workloads.SUN)
    occ = gpu.trace_shadow(b.dag(), rays, got, sun)
    srays, valid = O.shadow_rays(rays, want, sun)
    sref = O.trace_rays(scene, srays, flags=3, threads=16)
    socc = np.where(valid, (sref["instance"] != 0xFFFFFFFF).astype(np.uint32), np.uint32(0xFFFFFFFF))
    assert np.array_equal(occ, socc)
    assert 0 < (socc == 1).sum() < valid.sum()

    _say("case: skin 2048^3 pyr_trace / pyr_render")
    L = lib()
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    info.node_pool_bytes = max(64 << 20, 2 * b.node_words * 4)
    info.leaf_pool_bytes = max(64 << 20, 2 * b.leaf_count * 8)
    info.attribute_pool_bytes = max(64 << 20, 2 * b.voxel_count * 4)
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    try:
        geo, inst = C.c_void_p(), C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 11, host.ctypes.data_as(C.c_void_p), len(host),
                                          abi.PYR_BUILD_HOST_INPUT, C.byref(geo)), "build")
        abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        rb, hb = to_device(rays), Buffer(len(rays) * 16)
        abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), 0, 0), "trace")
        abi.check(L, L.pyr_synchronize(ctx), "sync")
        assert np.array_equal(_bits(hb.download(O.HIT, len(rays))), _bits(want))
        view = C.c_void_p()
        abi.check(L, L.pyr_view_create(ctx, C.byref(view)), "view")
        t = abi.PyrTargets()
        hb2, db = Buffer(len(rays) * 16), Buffer(len(rays) * 4)
        t.hits, t.depth = hb2.ptr, db.ptr
        abi.check(L, L.pyr_render(ctx, view, C.byref(cam), C.byref(t)), "render")
        rh = hb2.download(O.HIT, len(rays))
        dep = db.download(np.float32, len(rays))
        rhits = np.zeros(len(rays), O.HIT)
        rdep = np.zeros(len(rays), np.float32)
        O.lib().or_render_primary(C.byref(scene), C.byref(oc), rhits.ctypes.data_as(C.c_void_p),
                                  rdep.ctypes.data_as(C.c_void_p), 16)
        assert np.array_equal(_bits(rh), _bits(rhits))
        assert np.array_equal(_bits(rh), _bits(want))
        assert np.array_equal(dep.view(np.uint32), rdep.view(np.uint32))
    finally:
        L.pyr_destroy(ctx)


def test_skin_4096_full_size_traversal(hip_device):
    """The traversal DAG itself (4096^3 skin, 20000 caves): canon(GPU build)
    == the CPU builder on the same records, and the first benchmark pose at
    960x540 hits identically through the camera kernel."""
    from pyrit_amd import gpu, workloads

    _say("case: skin 4096^3 full size")
    w, b, host = _skin(12, 20000)
    _say(f"  {len(host)} records, GPU {b.node_words} words {b.leaf_count} leaves; CPU builder")
    cpu = O.build_points(12, host)
    del host
    _same_as_builder(b, cpu, 12)
    cam = workloads.camera_poses(w, 1, 960, 540)[0]
    rays = O.camera_rays(O.OrCamera.from_buffer_copy(bytes(cam)))
    want = O.trace_rays(O.single_scene(cpu), rays, threads=16)
    got = gpu.trace_camera(b.dag(), cam)
    assert np.array_equal(_bits(got), _bits(want))
    assert (want["instance"] == 0).sum() > len(rays) // 4


def _final_sets(inputs, sched):
    state = {}
    for _, kind, j, e in sched:
        l2, pts = inputs[j]
        if kind == "build":
            state[j] = final_points(pts, l2)
        else:
            state[j] = final_points(np.concatenate([state.get(j, np.zeros(0, O.VOXEL)), e]), l2)
    return state


def test_poisson_serving_80_20_same_alone_and_mixed(hip_device):
    """The serving workload's request stream (Poisson arrivals at 2000/s,
    80 % builds of uneven chunks with edges 32..256, 20 % edits of 1..512
    voxels of a chunk built earlier): each chunk's final DAG is the same when
    every request runs alone, when the stream arrives in real time, and when
    it is submitted at once (largest batches), and equals the CPU builder on
    the chunk's final voxel set."""
    from pyrit_amd.device import to_device
    from pyrit_amd.serving import Server, poisson_run, poisson_schedule

    _say("case: Poisson 80/20 serving, uneven chunks 32..256")
    w, sizes = _uneven(24, 17)
    vox, offs = w.generate_host()
    inputs = [(int(sizes[c]), vox[offs[c]:offs[c + 1]]) for c in range(w.count)]
    bufs = [to_device(p) if len(p) else None for _, p in inputs]
    dev = [(l2, bufs[c].ptr if bufs[c] else 0, len(p)) for c, (l2, p) in enumerate(inputs)]
    rate, duration, seed = 2000.0, 0.08, 21
    sched = poisson_schedule(dev, rate, duration, seed)
    assert sum(k == "edit" for _, k, _, _ in sched) > 0
    expect = _final_sets(inputs, sched)

    def arrays(server):
        out = {}
        for c in expect:
            nodes, leaves, attrs, root, l2, n = server.chunk_arrays(c)
            out[c] = None if n == 0 else (O.canon(nodes, leaves, root, l2), attrs.copy())
        return out

    alone = Server(workers=1)
    for _, kind, j, e in sched:
        l2, ptr, cnt = dev[j]
        if kind == "build":
            alone.build(j, l2, ptr, cnt)
        else:
            alone.edit(j, e, log2_size=l2)
        alone.drain()
    ra = alone.poll()
    got_alone = arrays(alone)
    err_alone = alone.error()
    alone.close()

    live = Server(workers=2)
    rl, _, nreq = poisson_run(live, dev, rate, duration, seed)
    got_live = arrays(live)
    err_live = live.error()
    live.close()

    burst = Server(workers=2)
    for _, kind, j, e in sched:
        l2, ptr, cnt = dev[j]
        if kind == "build":
            burst.build(j, l2, ptr, cnt)
        else:
            burst.edit(j, e, log2_size=l2)
    burst.drain()
    rb = burst.poll()
    got_burst = arrays(burst)
    err_burst = burst.error()
    burst.close()

    for name, res, err in (("alone", ra, err_alone), ("live", rl, err_live), ("burst", rb, err_burst)):
        bad = [(c.chunk, c.result, c.batch_chunks) for c in res if c.result != 0]
        assert len(res) == len(sched) == nreq and not bad, (name, bad[:8], err)
        assert err == "", (name, err)
    assert max(c.batch_chunks for c in rb) > 1, "the burst was never batched"
    for c, pts in expect.items():
        cpu = O.build_points(inputs[c][0], pts)
        for got in (got_alone, got_live, got_burst):
            if cpu.voxel_count == 0:
                assert got[c] is None
                continue
            g, a = got[c]
            assert g.same(cpu) and np.array_equal(a, cpu.attrs), c
