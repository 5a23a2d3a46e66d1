"""pyr_trace with PYR_TRACE_EXTENDED (gpu_kernels.zig:74-85, trace.zig:125-157):
every PyrHitEx (hit, voxel, position, normal) against the oracle's
extendedHit, bit for bit, on three instances (identity, half size, rotated
1.5x) with and without material sets, and the plain hits unchanged."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import grid_to_points, scene_grid
from test_materials_gpu import CUT, TRANS, XF, _oracle_scene

pytestmark = pytest.mark.gpu

SKIP, EXT, NO_ATTR, ANY = 0x8, 0x4, 0x2, 0x1


def test_extended_hits(hip_device):
    print("case: PYR_TRACE_EXTENDED through pyr_trace", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    assert abi.PYR_TRACE_EXTENDED == EXT and abi.HIT_EX.itemsize == 64
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
        rays = O.camera_rays(O.camera((-60, 90, -70), (30, 10, 30), (0, 1, 0), 1.1, 160, 120))
        rng = np.random.default_rng(17)
        extra = np.zeros(30000, O.RAY)
        extra["origin"] = rng.uniform(-120, 140, (len(extra), 3))
        extra["direction"] = rng.uniform(-1, 1, (len(extra), 3))
        extra["tmax"] = np.float32(3.0e38)
        # rays starting inside filled voxels of the identity instance
        z, y, x = np.nonzero(scene_grid(64))
        pick = rng.integers(0, len(x), 4000)
        extra["origin"][:4000] = np.stack([x[pick], y[pick], z[pick]], 1) + 0.25
        rays = np.concatenate([rays, extra])
        rb, hb = to_device(rays), Buffer(len(rays) * 64)

        def trace(flags, dtype):
            abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), 0, flags), "trace")
            abi.check(L, L.pyr_synchronize(ctx), "sync")
            return hb.download(dtype, len(rays))

        def words(h, k):
            return h.view(np.uint32).reshape(-1, k)

        s = _oracle_scene(cpu, [], [])
        for flags in (0, NO_ATTR, ANY):
            print(f"  flags {flags:#x}", flush=True)
            want = O.trace_rays_ex(s, rays, flags=flags)
            got = trace(flags | EXT, abi.HIT_EX)
            assert np.array_equal(words(got, 16), words(want, 16)), flags
            # the plain launch writes the same hits in 16 bytes
            assert np.array_equal(words(trace(flags, abi.HIT), 4), words(want, 16)[:, :4])
        hit = want["hit"]["instance"] != 0xFFFFFFFF
        assert set(want["hit"]["instance"][hit].tolist()) == {0, 1, 2}
        assert (want["hit"]["meta"] & 0x20).sum() >= 3000

        m = abi.PyrMaterial()
        L.pyr_material_default(C.byref(m))
        for i in TRANS:
            m.flags = abi.PYR_MATERIAL_TRANSPARENT
            abi.check(L, L.pyr_material_set(ctx, i, C.byref(m)), "material")
        for i in CUT:
            m.flags = abi.PYR_MATERIAL_CUTOUT | abi.PYR_MATERIAL_VOXEL_COLOR
            abi.check(L, L.pyr_material_set(ctx, i, C.byref(m)), "material")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        s = _oracle_scene(cpu, TRANS, CUT)
        for flags in (0, SKIP, SKIP | ANY):
            print(f"  materials, flags {flags:#x}", flush=True)
            want = O.trace_rays_ex(s, rays, flags=flags)
            assert np.array_equal(words(trace(flags | EXT, abi.HIT_EX), 16), words(want, 16)), flags
        # zero rays are a no-op with the flag too
        abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, 0, 0, EXT), "trace 0")
    finally:
        L.pyr_destroy(ctx)
