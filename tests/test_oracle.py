"""Pins the CPU oracle against the reference's own known answers, equalities
and brute force (no GPU). References are to files of the reference repo."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import attr_of, final_points, grid_to_points, random_points, scene_grid, sphere_noise_grid

FLT_MAX = np.float32(3.402823466e38)


def test_morton_known_answers():
    # dag_builder.zig:420-426
    L = O.lib()
    assert L.or_morton(1, 0, 0) == 1
    assert L.or_morton(0, 1, 0) == 2
    assert L.or_morton(0, 0, 1) == 4
    out = (C.c_uint32 * 3)()
    for v in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (5, 9, 1023), (0x1FFFFF, 3, 0x12345)]:
        L.or_morton_decode(L.or_morton(*v), out)
        assert tuple(out) == v


def test_empty_geometry():
    # dag_builder.zig:428-433
    d = O.build_points(3, np.zeros(0, O.VOXEL))
    assert d.nodes[d.root] & 0xFF == 0
    assert d.lookup(1, 2, 3) == 0


def test_single_voxel():
    # dag_builder.zig:435-442
    pts = np.array([(5, 6, 7, 42)], O.VOXEL)
    d = O.build_points(4, pts)
    assert d.voxel_count == 1
    assert d.lookup(5, 6, 7) == 42
    assert d.lookup(5, 6, 6) == 0


def test_dense_equals_points_and_lookup_brute_force():
    # dag_builder.zig:444-498: dense == points (shuffled) bit-identical, every
    # voxel's lookup equals the dense input
    grid = sphere_noise_grid(32)
    pts = grid_to_points(grid)
    np.random.default_rng(5).shuffle(pts)
    a = O.build_dense(5, grid)
    b = O.build_points(5, pts)
    assert np.array_equal(a.nodes, b.nodes)
    assert np.array_equal(a.leaves, b.leaves)
    assert np.array_equal(a.attrs, b.attrs)
    assert a.root == b.root and a.voxel_count == len(pts)
    z, y, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    got = np.array([a.lookup(int(xx), int(yy), int(zz)) for xx, yy, zz in zip(x.ravel(), y.ravel(), z.ravel())])
    assert np.array_equal(got, grid.ravel())


def test_periodic_pattern_dedups():
    # dag_builder.zig:500-520: period-8 pattern -> at most 4 x 10 node words
    n = 64
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    grid = ((x % 8) + (y % 8) < (z % 8) + 2).astype(np.uint32)
    d = O.build_dense(6, grid, store_attrs=False)
    assert len(d.nodes) <= 40
    sample = np.random.default_rng(0).integers(0, n, (2000, 3))
    for xx, yy, zz in sample:
        assert d.lookup(int(xx), int(yy), int(zz)) == grid[zz, yy, xx]


def test_no_duplicate_node_tuples():
    # minimality: a DAG from the builder never stores two equal tuples
    d = O.build_dense(6, scene_grid(64))
    seen = set()
    i = 0
    while i < len(d.nodes):
        ln = 2 + bin(int(d.nodes[i]) & 0xFF).count("1")
        t = tuple(d.nodes[i:i + ln].tolist())
        assert t not in seen
        seen.add(t)
        i += ln
    assert len(set(d.leaves.tolist())) == len(d.leaves)


def test_canon_is_identity_on_builder_output():
    d = O.build_dense(6, scene_grid(64))
    c = O.canon(d.nodes, d.leaves, d.root, 6)
    assert np.array_equal(c.nodes, d.nodes) and np.array_equal(c.leaves, d.leaves) and c.root == d.root


def test_canon_merges_repeated_leaves_and_tuples():
    """A DAG that repeats a leaf and a node tuple in fresh slots (what path
    edits leave behind) canonicalises to the minimal DAG of the same voxels."""
    d = O.build_dense(6, scene_grid(64))
    nodes, leaves = d.nodes.copy(), d.leaves.copy()
    # find a level-2 node (children are leaves): walk down the first children
    root = d.root
    node, s = root, 5
    path = []
    while s > 2:
        path.append(node)
        node = int(nodes[node + 2])
        s -= 1
    # copy its first leaf to a new slot and the node to new words pointing at the copy
    leaf_copy = len(leaves)
    leaves = np.append(leaves, leaves[nodes[node + 2]])
    n = bin(int(nodes[node]) & 0xFF).count("1")
    copy = len(nodes)
    words = nodes[node:node + 2 + n].copy()
    words[2] = leaf_copy
    nodes = np.concatenate([nodes, words])
    # the parent refers to the copy; equal voxels, one more leaf and tuple
    parent = path[-1]
    nodes[parent + 2] = copy
    assert len(leaves) == len(d.leaves) + 1
    c = O.canon(nodes, leaves, root, 6)
    assert c.same(d) and np.array_equal(c.leaves, d.leaves)


@pytest.mark.parametrize("log2,seed", [(3, 1), (5, 2), (6, 3), (7, 4)])
def test_serial_gpu_build_matches_cpu_builder(log2, seed):
    # cpu_test.zig:758-807: gpu build (serial) == cpu builder by canonical
    # form, voxel count and attribute array; duplicates, deletes, out of range
    pts = random_points(3000 if log2 > 4 else 200, log2, seed, dup_frac=0.2, delete_frac=0.05, out_frac=0.02)
    fin = final_points(pts, log2)
    ref = O.build_points(log2, fin)
    g = O.gb_build(log2, pts)
    assert g["voxel_count"] == ref.voxel_count == len(fin)
    assert np.array_equal(g["attrs"], ref.attrs)
    c = O.canon(g["nodes"], g["leaves"], g["root"], log2)
    assert np.array_equal(c.nodes, ref.nodes) and np.array_equal(c.leaves, ref.leaves) and c.root == ref.root
    # leaves: both builders emit them in first-appearance Morton order
    assert np.array_equal(g["leaves"], ref.leaves)


def test_serial_gpu_build_scene_with_duplicates():
    # cpu_test.zig:758-807 with the scene of tests/common.zig
    grid = scene_grid(64)
    pts = grid_to_points(grid)
    rng = np.random.default_rng(21)
    rng.shuffle(pts)
    dup = pts[:200].copy()
    dup["attribute"] = attr_of(dup["x"], dup["y"], dup["z"])
    extra = np.array([(-1, 0, 0, 5), (64, 0, 0, 5)], O.VOXEL)
    allp = np.concatenate([pts, dup, extra])
    g = O.gb_build(6, allp)
    ref = O.build_dense(6, grid)
    assert g["voxel_count"] == ref.voxel_count
    assert np.array_equal(g["attrs"], ref.attrs)
    c = O.canon(g["nodes"], g["leaves"], g["root"], 6)
    assert c.same(ref)


def test_serial_gpu_edit_equals_rebuild():
    # cpu_test.zig:809-850
    base = np.array([(x, 0, z, 0x10) for z in range(32) for x in range(32)], O.VOXEL)
    b1 = O.gb_build(5, base)
    edits = [(10, y, 10, 0x20) for y in range(1, 20)] + [(x, 0, 3, 0) for x in range(8)] + [(20, 0, 20, 0x30)]
    b2 = O.gb_build(5, np.array(edits, O.VOXEL), b1["keys"], b1["attrs"])
    final = []
    for v in base:
        if v["z"] == 3 and v["x"] < 8:
            continue
        a = 0x30 if (v["x"] == 20 and v["z"] == 20) else int(v["attribute"])
        final.append((int(v["x"]), int(v["y"]), int(v["z"]), a))
    final += [(10, y, 10, 0x20) for y in range(1, 20)]
    ref = O.build_points(5, np.array(final, O.VOXEL))
    c = O.canon(b2["nodes"], b2["leaves"], b2["root"], 5)
    assert c.same(ref) and b2["voxel_count"] == len(final)
    assert np.array_equal(b2["attrs"], ref.attrs)
    # empty geometry
    b3 = O.gb_build(4, np.zeros(0, O.VOXEL))
    assert b3["voxel_count"] == 0
    assert b3["nodes"].tolist() == [0, 0] and b3["root"] == 0


def _shell_batch(k=5, cap=8000, empty=3):
    # cpu_test.zig:1001-1060: chunk c holds a spherical shell of radius 6 + 2c,
    # chunk `empty` stays empty
    vox = np.zeros(k * cap, O.VOXEL)
    cnt = np.zeros(k, np.int64)
    z, y, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    d2 = (x - 15.5) ** 2 + (y - 15.5) ** 2 + (z - 15.5) ** 2
    for c in range(k):
        if c == empty:
            continue
        r = 6 + 2 * c
        sel = (d2 < r * r) & (d2 > (r - 2) ** 2)
        xs, ys, zs = x[sel], y[sel], z[sel]
        m = min(len(xs), cap)
        seg = vox[c * cap:c * cap + m]
        seg["x"], seg["y"], seg["z"] = xs[:m], ys[:m], zs[:m]
        seg["attribute"] = attr_of(xs[:m] + c, ys[:m], zs[:m])
        cnt[c] = m
    offs = np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32)
    return vox, cnt, offs, cap


def test_serial_batch_equals_single_builds():
    vox, cnt, offs, cap = _shell_batch()
    b, roots, first = O.gb_build_chunks(5, vox, offs, capacity=cap)
    assert roots[3] == 0xFFFFFFFF
    for c in range(5):
        if c == 3:
            continue
        single = O.gb_build(5, vox[c * cap:c * cap + cnt[c]])
        ref = O.canon(single["nodes"], single["leaves"], single["root"], 5)
        got = O.canon(b["nodes"], b["leaves"], int(roots[c]), 5)
        assert got.same(ref)
        n = int(b["nodes"][roots[c] + 1])
        assert np.array_equal(b["attrs"][first[c]:first[c] + n], single["attrs"])


def test_serial_mixed_size_batch_equals_single_builds():
    # this repository's extension: chunks of different edges in one batch
    rng = np.random.default_rng(9)
    sizes = [3, 5, 4, 5, 3, 4]
    chunks = []
    for l2 in sizes:
        pts = random_points(int(rng.integers(20, 400)), l2, int(rng.integers(1 << 30)), dup_frac=0.1)
        chunks.append(pts)
    chunks[2] = np.zeros(0, O.VOXEL)
    offs = np.concatenate([[0], np.cumsum([len(p) for p in chunks])]).astype(np.uint32)
    vox = np.concatenate(chunks)
    b, roots, first = O.gb_build_chunks(5, vox, offs, chunk_log2=np.array(sizes, np.uint8))
    assert roots[2] == 0xFFFFFFFF
    for c, l2 in enumerate(sizes):
        if c == 2:
            continue
        ref = O.build_points(l2, final_points(chunks[c], l2))
        got = O.canon(b["nodes"], b["leaves"], int(roots[c]), l2)
        assert got.same(ref), c
        assert np.array_equal(b["attrs"][first[c]:first[c] + ref.voxel_count], ref.attrs)


def test_axis_rays_known_answers():
    # cpu_test.zig:15-40: voxel (5,6,7) attr 42 in an 8^3 DAG
    d = O.build_points(3, np.array([(5, 6, 7, 42)], O.VOXEL))
    cases = [((-10, 6.5, 7.5), (1, 0, 0), 15, 1), ((20, 6.5, 7.5), (-1, 0, 0), 14, 0),
             ((5.5, -3, 7.5), (0, 1, 0), 9, 3), ((5.5, 6.5, 30), (0, 0, -2), 11, 4)]
    for o, dv, t, face in cases:
        h, _ = O.walk(d, o, dv)
        assert h is not None
        assert abs(h.t - t) <= 1e-5 and h.face == face
        assert tuple(h.voxel) == (5, 6, 7) and h.attribute == 42 and not h.inside
    assert O.walk(d, (-10, 6.5, 6.5), (1, 0, 0))[0] is None
    h, _ = O.walk(d, (5.5, 6.5, 7.5), (1, 0, 0))
    assert h.inside


def test_trace_touch_counts():
    """Distinct words read per ray (traversal algorithmic bytes), pinned by
    hand on the one-voxel 8^3 DAG: the root is the only node (mask word +
    one child ref), the child is the one leaf."""
    d = O.build_points(3, np.array([(5, 6, 7, 42)], O.VOXEL))
    scene = O.single_scene(d)

    def rays(*rs):
        out = np.zeros(len(rs), O.RAY)
        for i, (o, dv) in enumerate(rs):
            out[i]["origin"], out[i]["direction"], out[i]["tmin"], out[i]["tmax"] = o, dv, 0.0, 1e30
        return out

    hit = ((-10, 6.5, 7.5), (1, 0, 0))   # enters an empty octant, then the voxel's: hit
    miss = ((-10, 6.5, 6.5), (1, 0, 0))  # same octants and leaf, misses the voxel
    away = ((-10, 20.0, 7.5), (1, 0, 0))  # outside the instance bounds: reads nothing
    t = O.trace_touch(scene, rays(hit))
    assert (t["node_words"], t["leaves"], t["attributes"], t["hits"]) == (2, 1, 1, 1)
    t = O.trace_touch(scene, rays(miss))
    assert (t["node_words"], t["leaves"], t["attributes"], t["hits"]) == (2, 1, 0, 0)
    t = O.trace_touch(scene, rays(away))
    assert (t["node_words"], t["leaves"], t["attributes"], t["hits"]) == (0, 0, 0, 0)
    # distinct per ray: two equal rays count twice; hits agree with the tracer
    t = O.trace_touch(scene, rays(hit, hit, miss))
    assert (t["node_words"], t["leaves"], t["attributes"], t["hits"]) == (6, 3, 2, 2)
    hits = O.trace_rays(scene, rays(hit, hit, miss), threads=1)
    assert int((hits["instance"] != 0xFFFFFFFF).sum()) == 2


@pytest.mark.parametrize("log2,rays", [(6, 20000), (8, 4000)])
def test_walk_against_f64_dda(log2, rays):
    # cpu_test.zig:42-103: random rays vs an f64 Amanatides-Woo DDA on the
    # dense grid; mismatches <= 1e-4 of rays, attributes exact
    n = 1 << log2
    grid = scene_grid(n)
    d = O.build_dense(log2, grid)
    rng = np.random.default_rng(7)
    mism = face_mism = hits = 0
    for i in range(rays):
        o = n * 0.5 + (rng.random(3) * 2 - 1) * n * 1.5
        dv = n * 0.5 + (rng.random(3) * 2 - 1) * n * 0.45 - o
        if i % 8 == 0:
            dv[i % 3] = 0
        if i % 16 == 1:
            dv[0] = dv[1] = 0
        if not dv.any():
            continue
        r = O.ref_dda(grid, n, o.tolist(), dv.tolist())
        h, _ = O.walk(d, np.float32(o), np.float32(dv))
        if (r is None) != (h is None) or (r is not None and tuple(h.voxel) != r[1]):
            mism += 1
            continue
        if r is None:
            continue
        hits += 1
        ln = float(np.sqrt((dv * dv).sum()))
        assert abs(h.t - r[0]) * ln < 1e-3 * n
        if r[2] >= 0 and h.face != r[2]:
            face_mism += 1
        assert h.attribute == attr_of(*h.voxel)
    assert mism * 10000 <= rays  # at most 1e-4 of the rays (f32 vs f64 edge grazes)
    assert face_mism * 10000 <= rays
    assert hits > rays // 4


def test_camera_closed_forms():
    c = O.camera((-32, 96, -32), (32, 32, 32), (0, 1, 0), np.pi / 3, 257, 257)
    rays = O.camera_rays(c)
    d = rays["direction"]
    assert np.allclose((d * d).sum(1), 1, atol=1e-6)
    centre = rays[128 * 257 + 128]
    want = np.array([64, -64, 64], np.float64) / np.sqrt(3 * 64 * 64)
    assert np.allclose(centre["direction"], want, atol=1e-6)
    assert np.allclose(rays["origin"], (-32, 96, -32))


def test_trace_scene_equals_walk_for_identity_instance():
    d = O.build_dense(6, scene_grid(64))
    c = O.camera((-32, 96, -32), (32, 32, 32), (0, 1, 0), np.pi / 3, 64, 64)
    rays = O.camera_rays(c)
    s = O.single_scene(d)
    a = O.trace_rays(s, rays)
    b = O.trace_dag_rays(d, rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_range_alloc_sequence():
    # range_alloc.zig:61-78
    L = O.lib()
    r = O.OrRangeAlloc()
    L.or_range_init(C.byref(r), 100)
    a = L.or_range_alloc(C.byref(r), 30)
    b = L.or_range_alloc(C.byref(r), 30)
    c = L.or_range_alloc(C.byref(r), 30)
    assert L.or_range_alloc(C.byref(r), 20) == -1
    L.or_range_release(C.byref(r), b, 30)
    L.or_range_release(C.byref(r), a, 30)
    assert r.n == 2
    assert L.or_range_alloc(C.byref(r), 60) == 0
    L.or_range_release(C.byref(r), 0, 60)
    L.or_range_release(C.byref(r), c, 30)
    assert r.n == 1 and r.used == 0
    L.or_range_free(C.byref(r))
