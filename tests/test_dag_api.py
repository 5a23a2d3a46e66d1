"""Host DAGs and their way to and from the GPU: pyr_dag_* (src/root.zig:162-260),
the PYRD format (src/io.zig:86-160), pyr_geometry_create / download /
downsample (src/context.zig:799-816, 1009-1050).

The CPU tests hold the PYRD layout and the argument checks; the GPU tests
compare GPU-built host DAGs with the oracle's CPU builder after
canonicalisation, trace uploaded DAGs against the oracle, and check the LOD
rule of downsampling (per coarse cell the source voxel with the largest key,
as the reference's own test states it, tests/cpu_test.zig:960-1000)."""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle_py as O
from scenes import grid_to_points, random_points, scene_grid, sphere_noise_grid

INVALID = -1


def _pyrd(dag: O.Dag, voxel_count=None, version=1, log2=None, root=None, flags=None, attrs="dag") -> bytes:
    """PYRD bytes laid out field by field as io.zig:86-93 describes them."""
    a = dag.attrs if attrs == "dag" else attrs
    head = b"PYRD" + struct.pack("<4I", version, dag.log2_size if log2 is None else log2,
                                 dag.root if root is None else root,
                                 (1 if a is not None else 0) if flags is None else flags)
    na = 0 if a is None else len(a)
    head += struct.pack("<4Q", dag.voxel_count if voxel_count is None else voxel_count, len(dag.nodes),
                        len(dag.leaves), na)
    body = dag.nodes.astype("<u4").tobytes() + dag.leaves.astype("<u8").tobytes()
    if a is not None:
        body += np.asarray(a).astype("<u4").tobytes()
    return head + body


def _host():
    from pyrit_amd import hostdag

    return hostdag.HostDag


def _load_rc(data: bytes) -> int:
    from pyrit_amd import lib

    out = C.c_void_p()
    buf = C.create_string_buffer(data, len(data))
    rc = lib().pyr_dag_load(buf, len(data), C.byref(out))
    if rc == 0:
        lib().pyr_dag_destroy(out)
    return rc


# ---- CPU: format and argument checks ------------------------------------------

@pytest.mark.parametrize("store_attrs", [True, False])
def test_pyrd_load_save_round_trip(store_attrs):
    grid = sphere_noise_grid(32)
    dag = O.build_dense(5, grid, store_attrs)
    data = _pyrd(dag)
    assert len(data) == 52 + 4 * len(dag.nodes) + 8 * len(dag.leaves) + (4 * len(dag.attrs) if store_attrs else 0)
    with _host().load(data) as h:
        a = h.arrays()
        assert (a["log2_size"], a["root"], a["voxel_count"]) == (5, dag.root, dag.voxel_count)
        assert np.array_equal(a["nodes"], dag.nodes) and np.array_equal(a["leaves"], dag.leaves)
        if store_attrs:
            assert np.array_equal(a["attrs"], dag.attrs)
        else:
            assert a["attrs"] is None
        assert h.save() == data
    # a loaded DAG still answers lookups like the grid
    loaded = O.Dag(a["log2_size"], a["root"], a["nodes"], a["leaves"], a["attrs"], a["voxel_count"])
    for (x, y, z) in [(15, 15, 15), (0, 0, 0), (31, 31, 31), (3, 20, 9)]:
        want = grid[z, y, x] if store_attrs else (1 if grid[z, y, x] else 0)
        assert loaded.lookup(x, y, z) == want


def test_pyrd_rejects_bad_data():
    dag = O.build_dense(5, sphere_noise_grid(32))
    good = _pyrd(dag)
    assert _load_rc(good) == 0
    cases = {
        "magic": b"PYRX" + good[4:],
        "version": _pyrd(dag, version=2),
        "edge too small": _pyrd(dag, log2=2),
        "edge too large": _pyrd(dag, log2=21),
        "truncated": good[:-1],
        "header only": good[:51],
        "root out of range": _pyrd(dag, root=len(dag.nodes)),
        "attribute count": _pyrd(dag, voxel_count=dag.voxel_count + 1),
    }
    for name, data in cases.items():
        assert _load_rc(data) == INVALID, name
    # attributes present in the file but the flag clear: geometry only
    with _host().load(_pyrd(dag, flags=0)) as h:
        assert h.arrays()["attrs"] is None


def test_pyrd_save_size_query_and_small_buffer():
    from pyrit_amd import lib

    dag = O.build_dense(5, sphere_noise_grid(32))
    with _host().load(_pyrd(dag)) as h:
        n = C.c_size_t(0)
        assert lib().pyr_dag_save(h.handle, None, C.byref(n)) == 0
        small = C.c_size_t(n.value - 1)
        buf = C.create_string_buffer(n.value)
        assert lib().pyr_dag_save(h.handle, buf, C.byref(small)) == INVALID


def test_dag_build_argument_checks():
    """buildPoints rejects points outside the cube and empty attributes
    (dag_builder.zig:333-335) before any device work."""
    from pyrit_amd import lib

    L = lib()
    out = C.c_void_p()
    for v in [(8, 0, 0, 1), (-1, 0, 0, 1), (0, 0, 0, 0)]:
        pts = np.array([(1, 1, 1, 5), v], O.VOXEL)
        assert L.pyr_dag_build_points(3, pts.ctypes.data_as(C.c_void_p), 2, 0, C.byref(out)) == INVALID
    pts = np.array([(1, 1, 1, 5)], O.VOXEL)
    assert L.pyr_dag_build_points(2, pts.ctypes.data_as(C.c_void_p), 1, 0, C.byref(out)) == INVALID
    assert L.pyr_dag_build_points(21, pts.ctypes.data_as(C.c_void_p), 1, 0, C.byref(out)) == INVALID
    grid = np.zeros(8 ** 3, np.uint32)
    assert L.pyr_dag_build_dense(11, grid.ctypes.data_as(C.c_void_p), 0, C.byref(out)) == INVALID
    assert L.pyr_dag_build_dense(3, None, 0, C.byref(out)) == INVALID


# ---- GPU ---------------------------------------------------------------------------

def _same_as_cpu(h, cpu: O.Dag, store_attrs=True):
    a = h.arrays()
    assert a["voxel_count"] == cpu.voxel_count
    assert O.canon(a["nodes"], a["leaves"], a["root"], a["log2_size"]).same(cpu)
    if store_attrs:
        assert np.array_equal(a["attrs"], cpu.attrs)
    else:
        assert a["attrs"] is None
    return a


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [0, 1])
def test_dag_build_points_and_dense(hip_device, flags):
    print(f"case: pyr_dag_build_points/dense flags={flags}", flush=True)
    HD = _host()
    grid = scene_grid(64)
    cpu = O.build_dense(6, grid, flags == 0)
    with HD.build_dense(6, grid, flags) as h:
        _same_as_cpu(h, cpu, flags == 0)
    pts = grid_to_points(grid)
    pts = pts[np.random.default_rng(3).permutation(len(pts))]
    with HD.build_points(6, pts, flags) as h:
        _same_as_cpu(h, cpu, flags == 0)
    # duplicates: the later point wins, as in buildPoints
    rp = random_points(5000, 8, 21, dup_frac=0.3)
    with HD.build_points(8, rp, flags) as h:
        _same_as_cpu(h, O.build_points(8, rp, flags == 0), flags == 0)
    # empty input: an empty DAG
    with HD.build_points(5, np.zeros(0, O.VOXEL), flags) as h:
        _same_as_cpu(h, O.build_points(5, np.zeros(0, O.VOXEL), flags == 0), flags == 0)


@pytest.mark.gpu
def test_dag_build_fn_with_region_skips(hip_device):
    print("case: pyr_dag_build_fn", flush=True)
    HD = _host()
    grid = scene_grid(32)
    calls = {"voxel": 0, "region": 0}

    def voxel(x, y, z):
        calls["voxel"] += 1
        return grid[z, y, x]

    occupied = np.zeros((2, 2, 2), bool)
    for z, y, x in zip(*np.nonzero(grid)):
        occupied[z // 16, y // 16, x // 16] = True

    def region_empty(x, y, z, size):
        calls["region"] += 1
        return size == 16 and not occupied[z // 16, y // 16, x // 16]

    cpu = O.build_dense(5, grid)
    with HD.build_fn(5, voxel) as h:
        _same_as_cpu(h, cpu)
    assert calls["voxel"] == 32 ** 3
    calls["voxel"] = 0
    with HD.build_fn(5, voxel, region_empty) as h:
        _same_as_cpu(h, cpu)
    skipped = int(np.count_nonzero(~occupied))
    assert calls["voxel"] == 32 ** 3 - skipped * 16 ** 3


def _ctx():
    from pyrit_amd import abi, lib

    L = lib()
    info = abi.PyrCreateInfo()
    info.struct_size, info.version = C.sizeof(abi.PyrCreateInfo), abi.PYR_VERSION
    ctx = C.c_void_p()
    abi.check(L, L.pyr_create(C.byref(info), C.byref(ctx)), "create")
    return L, ctx


def _trace(L, ctx, geo, rays):
    from pyrit_amd import abi
    from pyrit_amd.device import Buffer, to_device

    inst = C.c_void_p()
    abi.check(L, L.pyr_instance_create(ctx, geo, C.byref(inst)), "instance")
    abi.check(L, L.pyr_commit(ctx, None), "commit")
    rb = to_device(rays)
    hb = Buffer(len(rays) * 16)
    abi.check(L, L.pyr_trace(ctx, rb.ptr, hb.ptr, len(rays), 0, 0), "trace")
    abi.check(L, L.pyr_synchronize(ctx), "sync")
    hits = hb.download(O.HIT, len(rays))
    abi.check(L, L.pyr_instance_destroy(ctx, inst), "instance destroy")
    return hits


def _same_hits(a, b):
    assert np.array_equal(a["instance"], b["instance"])
    assert np.array_equal(a["attribute"], b["attribute"])
    assert np.array_equal(a["meta"], b["meta"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("store_attrs", [True, False])
def test_geometry_create_download_round_trip(hip_device, store_attrs):
    """A CPU-built DAG (canonical layout) loaded from PYRD bytes and uploaded
    with pyr_geometry_create traces like the oracle; a downloaded GPU build
    saves and reloads unchanged and traces like the geometry it came from."""
    print(f"case: geometry create/download attrs={store_attrs}", flush=True)
    from pyrit_amd import abi

    HD = _host()
    grid = scene_grid(64)
    cpu = O.build_dense(6, grid, store_attrs)
    cam = O.camera((-60, 90, -70), (30, 10, 30), (0, 1, 0), 1.1, 128, 80)
    rays = O.camera_rays(cam)
    ref = O.trace_rays(O.single_scene(cpu, 6), rays)
    L, ctx = _ctx()
    try:
        with HD.load(_pyrd(cpu)) as h:
            geo = h.create_geometry(ctx)
        _same_hits(_trace(L, ctx, geo, rays), ref)
        with HD.download(ctx, geo) as d:
            a = d.arrays()
            assert np.array_equal(a["nodes"], cpu.nodes) and np.array_equal(a["leaves"], cpu.leaves)
            assert a["voxel_count"] == cpu.voxel_count and a["root"] == cpu.root
            assert (a["attrs"] is None) == (not store_attrs)
            assert d.save() == _pyrd(cpu)
        abi.check(L, L.pyr_geometry_destroy(ctx, geo), "destroy")
        # GPU build -> download -> PYRD -> load -> create
        pts = grid_to_points(grid)
        g2 = C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 6, pts.ctypes.data_as(C.c_void_p), len(pts), abi.PYR_BUILD_HOST_INPUT,
                                          C.byref(g2)), "build")
        h2 = _trace(L, ctx, g2, rays)
        with HD.download(ctx, g2) as d:
            _same_as_cpu(d, O.build_dense(6, grid))
            data = d.save()
        with HD.load(data) as back:
            g3 = back.create_geometry(ctx)
        _same_hits(_trace(L, ctx, g3, rays), h2)
        for g in (g2, g3):
            abi.check(L, L.pyr_geometry_destroy(ctx, g), "destroy")
    finally:
        L.pyr_destroy(ctx)


def _lod_reference(pts: np.ndarray, shift: int) -> np.ndarray:
    """Per coarse cell the point with the largest key morton(x>>2, y>>2,
    z>>2) << 6 | (x&3) + 4 (y&3) + 16 (z&3), coordinates >> shift."""
    OL = O.lib()
    x, y, z = (pts[c].astype(np.int64) for c in ("x", "y", "z"))
    key = np.array([(OL.or_morton(int(a) >> 2, int(b) >> 2, int(c) >> 2) << 6) | ((a & 3) + 4 * (b & 3) + 16 * (c & 3))
                    for a, b, c in zip(x, y, z)], np.uint64)
    cx, cy, cz = x >> shift, y >> shift, z >> shift
    best = {}
    for i in range(len(pts)):
        c = (int(cx[i]), int(cy[i]), int(cz[i]))
        if c not in best or key[i] > key[best[c]]:
            best[c] = i
    out = np.zeros(len(best), O.VOXEL)
    for k, (c, i) in enumerate(best.items()):
        out[k] = (c[0], c[1], c[2], pts["attribute"][i])
    return out


@pytest.mark.gpu
def test_geometry_downsample(hip_device):
    print("case: pyr_geometry_downsample", flush=True)
    from pyrit_amd import abi

    HD = _host()
    grid = scene_grid(64)
    pts = grid_to_points(grid)
    L, ctx = _ctx()
    try:
        geo = C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 6, pts.ctypes.data_as(C.c_void_p), len(pts),
                                          abi.PYR_BUILD_HOST_INPUT | abi.PYR_BUILD_EDITABLE, C.byref(geo)), "build")
        for shift in (1, 2, 3):
            lod = C.c_void_p()
            abi.check(L, L.pyr_geometry_downsample(ctx, geo, shift, abi.PYR_BUILD_EDITABLE, C.byref(lod)), "lod")
            with HD.download(ctx, lod) as d:
                a = _same_as_cpu(d, O.build_points(6 - shift, _lod_reference(pts, shift)))
                assert a["log2_size"] == 6 - shift
            if shift == 1:
                # an editable LOD downsamples again; the rule applies to the
                # LOD's own keys (two halvings need not pick the voxel one
                # quartering picks: keys order bricks by Morton code, voxels
                # inside a brick z-major)
                lod2 = C.c_void_p()
                abi.check(L, L.pyr_geometry_downsample(ctx, lod, 1, 0, C.byref(lod2)), "lod of lod")
                with HD.download(ctx, lod2) as d:
                    _same_as_cpu(d, O.build_points(4, _lod_reference(_lod_reference(pts, 1), 1)))
                abi.check(L, L.pyr_geometry_destroy(ctx, lod2), "destroy")
            abi.check(L, L.pyr_geometry_destroy(ctx, lod), "destroy")
        bad = C.c_void_p()
        assert L.pyr_geometry_downsample(ctx, geo, 0, 0, C.byref(bad)) == INVALID
        assert L.pyr_geometry_downsample(ctx, geo, 4, 0, C.byref(bad)) == INVALID
        plain = C.c_void_p()
        abi.check(L, L.pyr_geometry_build(ctx, 6, pts.ctypes.data_as(C.c_void_p), len(pts), abi.PYR_BUILD_HOST_INPUT,
                                          C.byref(plain)), "build")
        assert L.pyr_geometry_downsample(ctx, plain, 1, 0, C.byref(bad)) == INVALID
        for g in (geo, plain):
            abi.check(L, L.pyr_geometry_destroy(ctx, g), "destroy")
    finally:
        L.pyr_destroy(ctx)
