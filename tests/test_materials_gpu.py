"""Material sets of the traversal through the pyr_* API: PYR_TRACE_SKIP_TRANSPARENT
(dag.zig:113-115, 236-289 traceSkipping; trace.zig:93-98) and cutout holes
(dag.zig:133-151, trace.zig:80, 161-173), set by pyr_material_set
(context.zig:1948-1964) and taking effect at pyr_commit. Hits of pyr_trace
and pyr_render against the oracle's traceScene on a scene with the same
instances and material sets, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import grid_to_points, scene_grid

pytestmark = pytest.mark.gpu

SKIP, NO_ATTR, ANY = 0x8, 0x2, 0x1
TRANS = [m for m in range(256) if m % 5 == 0]
CUT = [m for m in range(256) if m % 3 == 0 and m % 5 != 0]
# identity; half size (voxel edge 0.5: cutouts apply); 1.5 x rotated (voxel
# edge 1.5, squared row norm of world_to_object 0.44 <= 0.5: no cutouts)
XF = [[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0],
      [0.5, 0, 0, 80, 0, 0.5, 0, 3, 0, 0, 0.5, -5],
      [0, 0, 1.5, -90, 0, 1.5, 0, 10, -1.5, 0, 0, 40]]


def _bits(h):
    return h.view(np.uint32)


def _set(materials):
    m = [0, 0, 0, 0]
    for i in materials:
        m[i >> 6] |= 1 << (i & 63)
    return m


def _oracle_scene(cpu, transparent, cutout):
    OL = O.lib()
    oi = (O.OrInstance * len(XF))()
    for k, m in enumerate(XF):
        oi[k].object_to_world[:] = [float(v) for v in m]
        OL.or_inverse(oi[k].object_to_world, oi[k].world_to_object)
        OL.or_transformed_box(oi[k].object_to_world, 64.0, oi[k].bounds_min, oi[k].bounds_max)
        oi[k].mask, oi[k].flags = 0xFF, 1
    geo = (O.OrGeometry * 1)()
    geo[0].root, geo[0].log2_size, geo[0].flags, geo[0].default_attribute = cpu.root, 6, 1, 1
    s = O.OrScene()
    s.nodes = cpu.nodes.ctypes.data_as(C.POINTER(C.c_uint32))
    s.leaves = cpu.leaves.ctypes.data_as(C.POINTER(C.c_uint64))
    s.attrs = cpu.attrs.ctypes.data_as(C.POINTER(C.c_uint32))
    s.geometries, s.instances, s.instance_count = geo, oi, len(XF)
    s.transparent[:] = _set(transparent)
    s.cutout[:] = _set(cutout)
    s._keep = (geo, oi, cpu)
    return s


def test_skip_transparent_and_cutouts(hip_device):
    print("case: material sets through pyr_trace / pyr_render", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    L = lib()
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    try:
        pts = grid_to_points(scene_grid(64))
        geo = C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 6, pts.ctypes.data_as(C.c_void_p), len(pts),
                                          abi.PYR_BUILD_HOST_INPUT, C.byref(geo)), "build")
        for m in XF:
            inst = C.c_void_p()
            abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
            abi.check(L, L.pyr_instance_set_transform(ctx, inst, (C.c_float * 12)(*m)), "xform")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        cpu = O.build_points(6, pts)
        cam = O.camera((-60, 90, -70), (30, 10, 30), (0, 1, 0), 1.1, 192, 128)
        rays = O.camera_rays(cam)
        rng = np.random.default_rng(11)
        extra = np.zeros(20000, O.RAY)
        extra["origin"] = rng.uniform(-120, 140, (len(extra), 3))
        extra["direction"] = rng.uniform(-1, 1, (len(extra), 3))
        extra["tmax"] = np.float32(3.0e38)
        rays = np.concatenate([rays, extra])
        rb, hb = to_device(rays), Buffer(len(rays) * 16)

        def trace(flags):
            abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), 0, flags), "trace")
            abi.check(L, L.pyr_synchronize(ctx), "sync")
            return hb.download(O.HIT, len(rays))

        # no material flags yet: the skip flag changes nothing
        plain = O.trace_rays(_oracle_scene(cpu, [], []), rays)
        assert np.array_equal(_bits(trace(SKIP)), _bits(plain))

        # materials are set now, but take effect at the commit
        m = abi.PyrMaterial()
        L.pyr_material_default(C.byref(m))
        assert m.flags == abi.PYR_MATERIAL_VOXEL_COLOR and abs(m.ior - 1.5) < 1e-7
        for i in TRANS:
            m.flags = abi.PYR_MATERIAL_TRANSPARENT
            abi.check(L, L.pyr_material_set(ctx, i, C.byref(m)), "material")
        for i in CUT:
            m.flags = abi.PYR_MATERIAL_CUTOUT | abi.PYR_MATERIAL_VOXEL_COLOR
            abi.check(L, L.pyr_material_set(ctx, i, C.byref(m)), "material")
        assert L.pyr_material_set(ctx, 256, C.byref(m)) == abi.PYR_ERROR_INVALID_ARGUMENT
        assert np.array_equal(_bits(trace(SKIP)), _bits(plain))
        abi.check(L, L.pyr_commit(ctx, None), "commit")

        s = _oracle_scene(cpu, TRANS, CUT)
        for flags in (0, SKIP, SKIP | NO_ATTR, SKIP | ANY, NO_ATTR):
            print(f"  flags {flags:#x}", flush=True)
            want = O.trace_rays(s, rays, flags=flags)
            assert np.array_equal(_bits(trace(flags)), _bits(want)), flags
        # the sets change hits: cutouts without the flag, transparency with it
        cut_only = O.trace_rays(s, rays)
        both = O.trace_rays(s, rays, flags=SKIP)
        assert not np.array_equal(_bits(cut_only), _bits(plain))
        assert not np.array_equal(_bits(both), _bits(cut_only))
        assert len(set(both["instance"].tolist())) == 4  # three instances and misses

        # render: primary hits and depth with the cutouts
        view = C.c_void_p()
        abi.check(L, L.pyr_view_create(ctx, C.byref(view)), "view")
        pc = abi.PyrCamera.from_buffer_copy(bytes(cam))
        t = abi.PyrTargets()
        n = cam.width * cam.height
        hb2, db = Buffer(n * 16), Buffer(n * 4)
        t.hits, t.depth = hb2.ptr, db.ptr
        abi.check(L, L.pyr_render(ctx, view, C.byref(pc), C.byref(t)), "render")
        rhits = np.zeros(n, O.HIT)
        rdep = np.zeros(n, np.float32)
        O.lib().or_render_primary(C.byref(s), C.byref(cam), rhits.ctypes.data_as(C.c_void_p),
                                  rdep.ctypes.data_as(C.c_void_p), 8)
        assert np.array_equal(_bits(hb2.download(O.HIT, n)), _bits(rhits))
        assert np.array_equal(db.download(np.uint32, n), rdep.view(np.uint32))
        assert not np.array_equal(_bits(rhits), _bits(plain[:n]))
    finally:
        L.pyr_destroy(ctx)
