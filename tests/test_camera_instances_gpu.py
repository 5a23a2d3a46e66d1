"""Cameras and instance changes through the pyr_* API against the oracle:
device camera rays (render.zig:16-32) for orthographic cameras, shifted and
jittered perspective cameras and near/far planes, bit for bit, with pyr_render
hits and depth on a three-instance scene; pyr_instance_set_geometry and
pyr_instance_set_user (context.zig:1619-1646), which reach the device at the
next pyr_commit."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import grid_to_points, scene_grid, sphere_noise_grid

pytestmark = pytest.mark.gpu

XF = [[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0],
      [0.5, 0, 0, 80, 0, 0.5, 0, 3, 0, 0, 0.5, -5],
      [0, 0, 1.5, -90, 0, 1.5, 0, 10, -1.5, 0, 0, 40]]


def _bits(h):
    return h.view(np.uint32)


class Ctx:
    def __init__(self):
        from pyrit_amd import abi, lib

        self.abi, self.L = abi, lib()
        info = abi.PyrCreateInfo()
        info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
        self.ctx = C.c_void_p()
        abi.check(self.L, self.L.pyr_create(C.byref(info), C.byref(self.ctx)), "create")

    def check(self, rc, what):
        self.abi.check(self.L, rc, what)

    def geometry(self, log2, pts):
        g = C.c_void_p()
        self.check(self.L.pyr_geometry_build(self.ctx, log2, pts.ctypes.data_as(C.c_void_p), len(pts),
                                             self.abi.PYR_BUILD_HOST_INPUT, C.byref(g)), "build")
        return g

    def instance(self, geo, m):
        i = C.c_void_p()
        self.check(self.L.pyr_instance_create(self.ctx, geo, C.byref(i)), "instance")
        self.check(self.L.pyr_instance_set_transform(self.ctx, i, (C.c_float * 12)(*m)), "xform")
        return i

    def destroy(self):
        self.L.pyr_destroy(self.ctx)


def test_camera_variants(hip_device):
    print("case: orthographic, shifted, jittered and clipped cameras", flush=True)
    from pyrit_amd import abi, gpu
    from pyrit_amd.device import Buffer

    c = Ctx()
    try:
        pts = grid_to_points(scene_grid(64))
        geo = c.geometry(6, pts)
        for m in XF:
            c.instance(geo, m)
        c.check(c.L.pyr_commit(c.ctx, None), "commit")
        s = O.multi_scene([O.build_points(6, pts)], [(m, 0, 0xFF) for m in XF])
        view = C.c_void_p()
        c.check(c.L.pyr_view_create(c.ctx, C.byref(view)), "view")
        eye, tgt, up = (-60.0, 90.0, -70.0), (30.0, 10.0, 30.0), (0.0, 1.0, 0.0)
        cams = []
        for kind in ("ortho", "ortho planes", "shifted", "clipped"):
            pc = abi.PyrCamera()
            c.L.pyr_camera_look_at(C.byref(pc), (C.c_float * 3)(*eye), (C.c_float * 3)(*tgt), (C.c_float * 3)(*up))
            if kind.startswith("ortho"):
                c.L.pyr_camera_orthographic(C.byref(pc), 70.0, 200, 120)
                if kind == "ortho planes":
                    pc.near_plane, pc.far_plane = 40.0, 130.0
            else:
                c.L.pyr_camera_perspective(C.byref(pc), 1.1, 192, 128, 0.0)
                pc.shift[:] = [0.15, -0.1]
                pc.jitter[:] = [0.3, -0.45]
                if kind == "clipped":
                    pc.near_plane, pc.far_plane = 60.0, 140.0
            cams.append((kind, pc))
        misses = {}
        for kind, pc in cams:
            print(f"  camera {kind}", flush=True)
            oc = O.OrCamera.from_buffer_copy(bytes(pc))
            rays = O.camera_rays(oc)
            assert np.array_equal(_bits(gpu.camera_rays(pc)), _bits(rays)), kind
            n = pc.width * pc.height
            t = abi.PyrTargets()
            hb, db = Buffer(n * 16), Buffer(n * 4)
            t.hits, t.depth = hb.ptr, db.ptr
            c.check(c.L.pyr_render(c.ctx, view, C.byref(pc), C.byref(t)), "render")
            want_h = np.zeros(n, O.HIT)
            want_d = np.zeros(n, np.float32)
            O.lib().or_render_primary(C.byref(s), C.byref(oc), want_h.ctypes.data_as(C.c_void_p),
                                      want_d.ctypes.data_as(C.c_void_p), 8)
            got_h = hb.download(O.HIT, n)
            assert np.array_equal(_bits(got_h), _bits(want_h)), kind
            assert np.array_equal(db.download(np.uint32, n), want_d.view(np.uint32)), kind
            hit = want_h["instance"] != 0xFFFFFFFF
            assert hit.sum() > n // 10, kind
            misses[kind] = int((~hit).sum())
        # the planes cut rays off: more misses than without them
        assert misses["ortho planes"] > misses["ortho"]
        assert misses["clipped"] > misses["shifted"]
    finally:
        c.destroy()


def _instance_table(c, count):
    from pyrit_amd import abi
    from pyrit_amd.device import download

    dev = c.L.pyr_scene_device(c.ctx)
    sc = abi.PyrScene.from_buffer_copy(download(dev, np.uint8, C.sizeof(abi.PyrScene)).tobytes())
    raw = download(sc.instances, np.uint8, count * C.sizeof(abi.PyrInstanceData)).tobytes()
    return [abi.PyrInstanceData.from_buffer_copy(raw, k * C.sizeof(abi.PyrInstanceData)) for k in range(count)]


def _stats(c):
    st = c.abi.PyrStats()
    c.check(c.L.pyr_get_stats(c.ctx, C.byref(st)), "stats")
    return st


def _words(c, geo):
    """node words and leaves of a geometry, from its download"""
    d = C.c_void_p()
    c.check(c.L.pyr_geometry_download(c.ctx, geo, C.byref(d)), "download")
    info = c.abi.PyrDagInfo()
    c.L.pyr_dag_get_info(d, C.byref(info))
    r = (info.node_words, info.leaf_count)
    c.L.pyr_dag_destroy(d)
    return r


def test_instance_set_geometry_and_user(hip_device):
    """Also pyr_get_stats (context.zig:2999-3017): pool use in bytes follows
    the geometries' node words and leaves, and a destroyed geometry's ranges
    stay counted until the next commit releases them."""
    print("case: pyr_instance_set_geometry / set_user / get_stats", flush=True)
    from pyrit_amd import abi
    from pyrit_amd.device import Buffer, to_device

    c = Ctx()
    try:
        pa = grid_to_points(scene_grid(64))
        pb = grid_to_points(sphere_noise_grid(32, seed=77))
        st0 = _stats(c)
        assert (st0.node_pool_used, st0.leaf_pool_used, st0.geometries, st0.instances, st0.features) == (0, 0, 0, 0, 0)
        assert st0.node_pool_capacity > 0 and st0.leaf_pool_capacity > 0 and st0.attribute_pool_capacity > 0
        ga, gb = c.geometry(6, pa), c.geometry(5, pb)
        da, db = O.build_points(6, pa), O.build_points(5, pb)
        (wa, la), (wb, lb) = _words(c, ga), _words(c, gb)
        st1 = _stats(c)
        assert (st1.node_pool_used, st1.leaf_pool_used) == (4 * (wa + wb), 8 * (la + lb))
        assert st1.attribute_pool_used > 0 and st1.geometries == 2
        insts = [c.instance(ga, m) for m in XF]
        for k, i in enumerate(insts):
            c.check(c.L.pyr_instance_set_user(c.ctx, i, 1000 + k), "user")
        c.check(c.L.pyr_commit(c.ctx, None), "commit")
        assert [d.user for d in _instance_table(c, 3)] == [1000, 1001, 1002]
        st2 = _stats(c)
        assert (st2.instances, st2.frame) == (3, st0.frame + 1)

        rays = O.camera_rays(O.camera((-60, 90, -70), (30, 10, 30), (0, 1, 0), 1.1, 160, 120))
        rng = np.random.default_rng(5)
        extra = np.zeros(20000, O.RAY)
        extra["origin"] = rng.uniform(-120, 140, (len(extra), 3))
        extra["direction"] = rng.uniform(-1, 1, (len(extra), 3))
        extra["tmax"] = np.float32(3.0e38)
        rays = np.concatenate([rays, extra])
        rb, hb = to_device(rays), Buffer(len(rays) * 16)

        def trace():
            c.check(c.L.pyr_trace(c.ctx, rb.ptr, hb.ptr, len(rays), 0, 0), "trace")
            c.check(c.L.pyr_synchronize(c.ctx), "sync")
            return hb.download(O.HIT, len(rays))

        all_a = O.trace_rays(O.multi_scene([da, db], [(m, 0, 0xFF) for m in XF]), rays)
        assert np.array_equal(_bits(trace()), _bits(all_a))

        # instance 1 takes the 32^3 geometry: nothing changes before the commit
        c.check(c.L.pyr_instance_set_geometry(c.ctx, insts[1], gb), "set geometry")
        assert np.array_equal(_bits(trace()), _bits(all_a))
        c.check(c.L.pyr_commit(c.ctx, None), "commit")
        mixed = O.trace_rays(O.multi_scene([da, db], [(XF[0], 0, 0xFF), (XF[1], 1, 0xFF), (XF[2], 0, 0xFF)]), rays)
        got = trace()
        assert np.array_equal(_bits(got), _bits(mixed))
        assert not np.array_equal(_bits(mixed), _bits(all_a))
        table = _instance_table(c, 3)
        assert [d.geometry for d in table] == [0, 1, 0] and [d.user for d in table] == [1000, 1001, 1002]
        assert table[1].history > table[0].history  # a new shape starts a new history
        c.check(c.L.pyr_instance_reset_history(c.ctx, insts[2]), "reset history")
        c.check(c.L.pyr_commit(c.ctx, None), "commit")
        t2 = _instance_table(c, 3)
        assert t2[2].history > table[1].history and t2[0].history == table[0].history
        assert np.array_equal(_bits(trace()), _bits(mixed))
        view = C.c_void_p()
        c.check(c.L.pyr_view_create(c.ctx, C.byref(view)), "view")
        c.check(c.L.pyr_view_reset_history(c.ctx, view), "view reset")
        assert c.L.pyr_view_reset_history(c.ctx, None) == abi.PYR_ERROR_INVALID_HANDLE

        # the first geometry stays in use until no instance refers to it
        assert c.L.pyr_geometry_destroy(c.ctx, ga) == abi.PYR_ERROR_IN_USE
        for i in (insts[0], insts[2]):
            c.check(c.L.pyr_instance_set_geometry(c.ctx, i, gb), "set geometry")
        c.check(c.L.pyr_geometry_destroy(c.ctx, ga), "destroy")
        st3 = _stats(c)
        assert st3.geometries == 1 and st3.node_pool_used == st1.node_pool_used  # released by the commit
        c.check(c.L.pyr_commit(c.ctx, None), "commit")
        st4 = _stats(c)
        assert (st4.node_pool_used, st4.leaf_pool_used) == (4 * wb, 8 * lb)
        assert st4.attribute_pool_used < st1.attribute_pool_used
        all_b = O.trace_rays(O.multi_scene([db], [(m, 0, 0xFF) for m in XF]), rays)
        assert np.array_equal(_bits(trace()), _bits(all_b))
        assert len(set(all_b["instance"].tolist())) == 4  # three instances and misses
        # stale handles are refused
        assert c.L.pyr_instance_set_geometry(c.ctx, insts[0], ga) == abi.PYR_ERROR_INVALID_HANDLE
        assert c.L.pyr_instance_set_user(c.ctx, None, 1) == abi.PYR_ERROR_INVALID_HANDLE
    finally:
        c.destroy()
