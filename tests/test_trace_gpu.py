"""Parity of the HIP traversal against the oracle's restatement of
src/device/dag.zig walk / src/device/trace.zig traceScene: identical hit
instance, attribute, face/inside bits and bit-identical fp32 t (the
tolerance stated for t is 1e-5 relative; the kernels reach 0)."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import attr_of, grid_to_points, scene_grid

pytestmark = pytest.mark.gpu


def same_hits(a, b, rtol=1e-5):
    assert np.array_equal(a["instance"], b["instance"])
    assert np.array_equal(a["attribute"], b["attribute"])
    assert np.array_equal(a["meta"], b["meta"])
    ta, tb = a["t"].astype(np.float64), b["t"].astype(np.float64)
    assert np.all(np.abs(ta - tb) <= rtol * np.maximum(1.0, np.abs(tb)))
    return int(np.count_nonzero(a["t"].view(np.uint32) != b["t"].view(np.uint32)))


def product_camera(eye, target, up, fov, w, h):
    from pyrit_amd import abi, lib

    c = abi.PyrCamera()
    L = lib()
    L.pyr_camera_look_at(C.byref(c), (C.c_float * 3)(*eye), (C.c_float * 3)(*target), (C.c_float * 3)(*up))
    L.pyr_camera_perspective(C.byref(c), fov, w, h, 0.0)
    return c


def test_terrain_chunk_64_primary_rays(hip_device):
    print("case: terrain chunk 64^3, 256x256 primary rays", flush=True)
    from pyrit_amd import gpu, workloads

    vox, _ = workloads.terrain_chunk64().generate_host()
    b = gpu.build(6, vox)
    eye, tgt, up = (-32, 96, -32), (32, 32, 32), (0, 1, 0)
    pc = product_camera(eye, tgt, up, np.pi / 3, 256, 256)
    oc = O.camera(eye, tgt, up, np.pi / 3, 256, 256)
    assert bytes(pc) == bytes(oc)
    hits = gpu.trace_camera(b.dag(), pc)
    cpu = O.build_points(6, vox)
    rays = O.camera_rays(oc)
    ref = O.trace_dag_rays(cpu, rays)
    assert same_hits(hits, ref) == 0
    assert (hits["instance"] == 0).sum() > 1000
    # device-generated rays equal the oracle's
    drays = gpu.camera_rays(pc)
    assert np.array_equal(drays.view(np.uint32), rays.view(np.uint32))


@pytest.mark.parametrize("log2,count", [(6, 100000), (8, 50000)])
def test_random_rays_against_oracle(hip_device, log2, count):
    # cpu_test.zig:42-103 ray distribution, traced on the GPU and by the oracle
    print(f"case: random rays log2={log2} count={count}", flush=True)
    from pyrit_amd import gpu

    n = 1 << log2
    grid = scene_grid(n)
    pts = grid_to_points(grid)
    b = gpu.build(log2, pts)
    rng = np.random.default_rng(7)
    o = n * 0.5 + (rng.random((count, 3)) * 2 - 1) * n * 1.5
    d = n * 0.5 + (rng.random((count, 3)) * 2 - 1) * n * 0.45 - o
    d[::8, 0] = 0
    d[1::16, :2] = 0
    rays = np.zeros(count, O.RAY)
    rays["origin"], rays["direction"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 0, np.float32(3.402823466e38)
    hits = gpu.trace_dag(b.dag(), rays)
    cpu = O.Dag(log2, b.root, b.nodes(), b.leaves(), b.attrs(), b.voxel_count)
    ref = O.trace_rays(O.single_scene(cpu), rays)
    # identity instance: scene trace and bare walk agree
    assert same_hits(hits, ref) == 0
    hit = hits["instance"] == 0
    assert hit.sum() > count // 4
    # attrOf always sets bit 0 (tests/common.zig:18)
    assert np.all(hits["attribute"][hit] & 1)


def test_no_attribute_and_any_hit_flags(hip_device):
    print("case: trace flags", flush=True)
    from pyrit_amd import gpu

    grid = scene_grid(64)
    b = gpu.build(6, grid_to_points(grid))
    c = O.camera((-20, 80, -20), (32, 20, 32), (0, 1, 0), 1.0, 64, 64)
    rays = O.camera_rays(c)
    cpu = O.Dag(6, b.root, b.nodes(), b.leaves(), b.attrs(), b.voxel_count)
    for flags in (2, 1, 3):
        same_hits(gpu.trace_dag(b.dag(), rays, flags=flags), O.trace_dag_rays(cpu, rays, flags=flags))


def test_shadow_rays(hip_device):
    print("case: shadow rays", flush=True)
    from pyrit_amd import gpu, workloads

    vox, _ = workloads.terrain_chunk64().generate_host()
    b = gpu.build(6, vox)
    c = O.camera((-32, 96, -32), (32, 32, 32), (0, 1, 0), np.pi / 3, 128, 128)
    rays = O.camera_rays(c)
    hits = gpu.trace_dag(b.dag(), rays)
    sun = workloads.normalize(workloads.SUN)
    occ = gpu.trace_shadow(b.dag(), rays, hits, sun)
    srays, valid = O.shadow_rays(rays, hits, sun)
    cpu = O.build_points(6, vox)
    ref = O.trace_dag_rays(cpu, srays, flags=3)
    want = np.where(valid, (ref["instance"] != 0xFFFFFFFF).astype(np.uint32), np.uint32(0xFFFFFFFF))
    assert np.array_equal(occ, want)
    assert 0 < (want == 1).sum() < valid.sum()


def test_scene_api_instances_and_render(hip_device):
    # pyr_geometry_build / instances / commit / pyr_trace / pyr_render against
    # the oracle's traceScene with the same instance data
    print("case: pyr_* scene API", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    L = lib()
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    try:
        grid = scene_grid(64)
        pts = grid_to_points(grid)
        geo = C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 6, pts.ctypes.data_as(C.c_void_p), len(pts),
                                          abi.PYR_BUILD_HOST_INPUT | abi.PYR_BUILD_EDITABLE, C.byref(geo)), "build")
        insts = []
        xf = [[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0],
              [0.5, 0, 0, 80, 0, 0.5, 0, 3, 0, 0, 0.5, -5],
              [0, 0, 1.5, -90, 0, 1.5, 0, 10, -1.5, 0, 0, 40]]
        for m in xf:
            inst = C.c_void_p()
            abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
            abi.check(L, L.pyr_instance_set_transform(ctx, inst, (C.c_float * 12)(*m)), "xform")
            insts.append(inst)
        abi.check(L, L.pyr_instance_set_mask(ctx, insts[2], 0x2), "mask")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        # oracle scene with the same instance data (inverse and bounds recomputed by the oracle)
        cpu = O.build_points(6, pts)
        OL = O.lib()
        oi = (O.OrInstance * 3)()
        for k, m in enumerate(xf):
            oi[k].object_to_world[:] = [float(v) for v in m]
            OL.or_inverse(oi[k].object_to_world, oi[k].world_to_object)
            OL.or_transformed_box(oi[k].object_to_world, 64.0, oi[k].bounds_min, oi[k].bounds_max)
            oi[k].mask, oi[k].flags = (0x2 if k == 2 else 0xFF), 1
        geo_o = (O.OrGeometry * 1)()
        geo_o[0].root, geo_o[0].log2_size, geo_o[0].flags, geo_o[0].default_attribute = cpu.root, 6, 1, 1
        s = O.OrScene()
        s.nodes = cpu.nodes.ctypes.data_as(C.POINTER(C.c_uint32))
        s.leaves = cpu.leaves.ctypes.data_as(C.POINTER(C.c_uint64))
        s.attrs = cpu.attrs.ctypes.data_as(C.POINTER(C.c_uint32))
        s.geometries, s.instances, s.instance_count = geo_o, oi, 3
        # canonical node layout differs from the GPU layout: compare hits
        # (instance, attribute, face, t), which do not depend on the layout
        cam = O.camera((-60, 90, -70), (30, 10, 30), (0, 1, 0), 1.1, 160, 96)
        rays = O.camera_rays(cam)
        rb = to_device(rays)
        hb = Buffer(len(rays) * 16)
        for mask in (0, 0x2, 0x1):
            abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), mask, 0), "trace")
            abi.check(L, L.pyr_synchronize(ctx), "sync")
            got = hb.download(O.HIT, len(rays))
            ref = O.trace_rays(s, rays, ray_mask=mask)
            assert same_hits(got, ref) == 0
        assert len(set(got["instance"].tolist())) >= 2
        # render: hits equal trace of the camera rays, depth equals the oracle's
        view = C.c_void_p()
        abi.check(L, L.pyr_view_create(ctx, C.byref(view)), "view")
        pc = abi.PyrCamera.from_buffer_copy(bytes(cam))
        t = abi.PyrTargets()
        hb2 = Buffer(len(rays) * 16)
        db = Buffer(len(rays) * 4)
        t.hits, t.depth = hb2.ptr, db.ptr
        abi.check(L, L.pyr_render(ctx, view, C.byref(pc), C.byref(t)), "render")
        rh = hb2.download(O.HIT, len(rays))
        dep = db.download(np.float32, len(rays))
        rhits = np.zeros(len(rays), O.HIT)
        rdep = np.zeros(len(rays), np.float32)
        OL.or_render_primary(C.byref(s), C.byref(cam), rhits.ctypes.data_as(C.c_void_p),
                             rdep.ctypes.data_as(C.c_void_p), 4)
        assert same_hits(rh, rhits) == 0
        assert np.array_equal(dep.view(np.uint32), rdep.view(np.uint32))
        # edit: remove the floor, add a column; hits follow the rebuilt set
        ed = np.array([(x, 0, z, 0) for x in range(64) for z in range(64)] + [(30, y, 30, 7) for y in range(64)],
                      O.VOXEL)
        abi.check(L, L.pyr_geometry_edit(ctx, geo, ed.ctypes.data_as(C.c_void_p), len(ed),
                                         abi.PYR_BUILD_HOST_INPUT), "edit")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        g2 = grid.copy()
        g2[:, 0, :] = 0
        for y in range(64):
            g2[30, y, 30] = 7
        cpu2 = O.build_dense(6, g2)
        s.nodes = cpu2.nodes.ctypes.data_as(C.POINTER(C.c_uint32))
        s.leaves = cpu2.leaves.ctypes.data_as(C.POINTER(C.c_uint64))
        s.attrs = cpu2.attrs.ctypes.data_as(C.POINTER(C.c_uint32))
        geo_o[0].root = cpu2.root
        abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), 0, 0), "trace")
        abi.check(L, L.pyr_synchronize(ctx), "sync")
        assert same_hits(hb.download(O.HIT, len(rays)), O.trace_rays(s, rays)) == 0
        # destroy rules: in use while instanced
        assert L.pyr_geometry_destroy(ctx, geo) == -5
        for inst in insts:
            abi.check(L, L.pyr_instance_destroy(ctx, inst), "destroy instance")
        abi.check(L, L.pyr_geometry_destroy(ctx, geo), "destroy geometry")
        assert L.pyr_geometry_destroy(ctx, geo) == -2
    finally:
        L.pyr_destroy(ctx)
