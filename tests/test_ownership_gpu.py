"""Ownership of geometry ranges across pyr_commit (docs/API.md:91,
src/context.zig:972-1005, 1446-1545): an edited geometry's previous version
and a destroyed geometry stay traceable until the next pyr_commit, even when
other builds are installed meanwhile; the commit releases them."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import final_points, grid_to_points, random_points, scene_grid

pytestmark = pytest.mark.gpu


def _ctx(L, abi, pool_bytes):
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    info.node_pool_bytes = info.leaf_pool_bytes = info.attribute_pool_bytes = pool_bytes
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    return ctx


def _build(L, abi, ctx, log2, pts, flags=0):
    geo = C.c_void_p()
    abi.check(L, L.pyr_geometry_build(ctx, log2, pts.ctypes.data_as(C.c_void_p), len(pts),
                                      abi.PYR_BUILD_HOST_INPUT | flags, C.byref(geo)), "build")
    return geo


def _hits(L, abi, ctx, rb, hb, n):
    abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, n, 0, 0), "trace")
    abi.check(L, L.pyr_synchronize(ctx), "sync")
    return hb.download(O.HIT, n)


def _same(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_edit_and_destroy_stay_traceable_until_commit(hip_device):
    print("case: deferred release of edited and destroyed geometry", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    L = lib()
    # small pools, so a build after the edit would land in the old version's
    # ranges if they were released at once (first fit from offset 0)
    ctx = _ctx(L, abi, 8 << 20)
    try:
        grid = scene_grid(64)
        pts = grid_to_points(grid)
        geo = _build(L, abi, ctx, 6, pts, abi.PYR_BUILD_EDITABLE)
        inst = C.c_void_p()
        abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        cam = O.camera((-40, 90, -50), (32, 16, 32), (0, 1, 0), 1.0, 128, 96)
        rays = O.camera_rays(cam)
        rb, hb = to_device(rays), Buffer(len(rays) * 16)
        want_a = O.trace_rays(O.single_scene(O.build_points(6, pts)), rays)
        assert (want_a["instance"] == 0).sum() > 1000
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want_a)

        # edit: remove the floor and the sphere's upper half, add a column
        ed = [(x, y, z, 0) for (z, y, x) in zip(*np.nonzero(grid)) if y < 3 or y > 32]
        ed += [(20, y, 20, 9) for y in range(64)]
        ed = np.array(ed, O.VOXEL)
        abi.check(L, L.pyr_geometry_edit(ctx, geo, ed.ctypes.data_as(C.c_void_p), len(ed),
                                         abi.PYR_BUILD_HOST_INPUT), "edit")
        # other builds fill the pools meanwhile
        others = [_build(L, abi, ctx, 6, random_points(20000, 6, seed)) for seed in (5, 6)]
        # not committed yet: the scene still shows the old version, bit for bit
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want_a)
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        g2 = grid.copy()
        g2[:, :3, :] = 0
        g2[:, 33:, :] = 0
        g2[20, :, 20] = 9
        want_b = O.trace_rays(O.single_scene(O.build_dense(6, g2)), rays)
        assert not _same(want_a, want_b)
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want_b)

        # destroy the instance and the geometry, build into the pools again:
        # the committed instance still traces the destroyed geometry
        index = L.pyr_geometry_index(geo)
        abi.check(L, L.pyr_instance_destroy(ctx, inst), "destroy instance")
        abi.check(L, L.pyr_geometry_destroy(ctx, geo), "destroy geometry")
        more = [_build(L, abi, ctx, 6, random_points(30000, 6, seed)) for seed in (7, 8)]
        # the destroyed geometry's slot is not handed out before the commit
        assert index not in [L.pyr_geometry_index(g) for g in others + more]
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want_b)
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        h = _hits(L, abi, ctx, rb, hb, len(rays))
        assert np.all(h["instance"] == 0xFFFFFFFF)
        # after the commit the slot and the ranges are free again: a build of
        # the first scene reuses the slot and traces like the oracle
        geo3 = _build(L, abi, ctx, 6, pts)
        assert L.pyr_geometry_index(geo3) == index
        inst3 = C.c_void_p()
        abi.check(L, L.pyr_instance_create(ctx, geo3, C.byref(inst3)), "instance")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want_a)
    finally:
        L.pyr_destroy(ctx)


def test_repeated_edits_without_commit_release_every_version(hip_device):
    # versions pile up between commits; one commit releases all of them, so
    # the pools do not run out over many edit/commit cycles
    print("case: many edits between commits", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    L = lib()
    ctx = _ctx(L, abi, 4 << 20)
    try:
        pts = random_points(20000, 6, 11)
        geo = _build(L, abi, ctx, 6, pts, abi.PYR_BUILD_EDITABLE)
        inst = C.c_void_p()
        abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
        abi.check(L, L.pyr_commit(ctx, None), "commit")
        cur = pts
        for cycle in range(12):
            for k in range(3):
                ed = random_points(500, 6, 100 * cycle + k, dup_frac=0.0)
                ed["attribute"][::3] = 0
                abi.check(L, L.pyr_geometry_edit(ctx, geo, ed.ctypes.data_as(C.c_void_p), len(ed),
                                                 abi.PYR_BUILD_HOST_INPUT), "edit")
                cur = np.concatenate([cur, ed])
            abi.check(L, L.pyr_commit(ctx, None), "commit")
        rays = O.camera_rays(O.camera((-20, 80, -30), (32, 32, 32), (0, 1, 0), 1.0, 64, 64))
        rb, hb = to_device(rays), Buffer(len(rays) * 16)
        want = O.trace_rays(O.single_scene(O.build_points(6, final_points(cur, 6))), rays)
        assert _same(_hits(L, abi, ctx, rb, hb, len(rays)), want)
    finally:
        L.pyr_destroy(ctx)
