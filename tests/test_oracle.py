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


def test_camera_shift_planes_and_orthographic_closed_forms():
    """cameraRay (render.zig:16-32) in closed form, in f64: a perspective ray
    through pixel (x, y) points along right vx + up vy - back with
    vx = (nx + shift x) scale x, nx from the jittered pixel centre, and its
    tmin/tmax are the near/far planes times |(vx, vy, -1)|; an orthographic
    ray starts at eye + right vx + up vy, points along -back, and its planes
    are not scaled."""
    eye, tgt = np.array([-32.0, 96.0, -32.0]), np.array([32.0, 32.0, 32.0])
    back = (eye - tgt) / np.linalg.norm(eye - tgt)
    right = np.cross([0.0, 1.0, 0.0], back)
    right /= np.linalg.norm(right)
    up = np.cross(back, right)
    w, h = 40, 30
    for ortho in (False, True):
        c = O.camera(tuple(eye), tuple(tgt), (0, 1, 0), 0.9, w, h, 3.0)
        if ortho:
            c.projection, c.scale[0], c.scale[1] = 1, 20.0 * w / h, 20.0
        c.shift[:] = [0.1, -0.2]
        c.jitter[:] = [0.25, -0.375]
        c.far_plane = 150.0
        rays = O.camera_rays(c)
        y, x = np.divmod(np.arange(w * h), w)
        nx = 2 * (x + 0.5 + 0.25) / w - 1
        ny = 1 - 2 * (y + 0.5 - 0.375) / h
        vx, vy = (nx + 0.1) * c.scale[0], (ny - 0.2) * c.scale[1]
        if ortho:
            o = eye + vx[:, None] * right + vy[:, None] * up
            d = np.broadcast_to(-back, o.shape)
            ln = np.ones(w * h)
        else:
            o = np.broadcast_to(eye, (w * h, 3))
            dv = vx[:, None] * right + vy[:, None] * up - back
            ln = np.linalg.norm(dv, axis=1)
            d = dv / ln[:, None]
        assert np.abs(rays["origin"] - o).max() < 1e-4
        assert np.abs(rays["direction"] - d).max() < 1e-5
        assert np.allclose(rays["tmin"], 3.0 * ln, rtol=1e-5)
        assert np.allclose(rays["tmax"], 150.0 * ln, rtol=1e-5)


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


def make_transform(axis, a, s, t):
    """Rotation about `axis` by a, uniform scale s, translation t, in f64 and
    rounded to f32 (tests/common.zig:169-191 makeTransform)"""
    ax = np.asarray(axis, np.float64)
    ax = ax / np.sqrt((ax * ax).sum())
    c, sn, k = np.cos(a), np.sin(a), 1 - np.cos(a)
    x, y, z = ax
    r = np.array([[c + x * x * k, x * y * k - z * sn, x * z * k + y * sn],
                  [y * x * k + z * sn, c + y * y * k, y * z * k - x * sn],
                  [z * x * k - y * sn, z * y * k + x * sn, c + z * z * k]])
    m = np.zeros((3, 4))
    m[:, :3] = r * s
    m[:, 3] = t
    return m.astype(np.float32).ravel()


def apply_f64(m, p, w):
    """tests/common.zig:163-167 applyF64: f32 matrix on an f64 point (w = 1) or vector (w = 0)"""
    m = np.asarray(m, np.float64).reshape(3, 4)
    return m[:, :3] @ np.asarray(p, np.float64) + m[:, 3] * w


def transformed_scene(dag, n, xforms, masks):
    OL = O.lib()
    inst = (O.OrInstance * len(xforms))()
    for k, m in enumerate(xforms):
        inst[k].object_to_world[:] = [float(v) for v in m]
        assert OL.or_inverse(inst[k].object_to_world, inst[k].world_to_object) == 1
        OL.or_transformed_box(inst[k].object_to_world, float(n), inst[k].bounds_min, inst[k].bounds_max)
        inst[k].geometry, inst[k].mask, inst[k].flags = 0, masks[k], 1
    s = O.single_scene(dag)
    s.instances, s.instance_count = inst, len(xforms)
    s._keep_inst = inst
    return s, inst


def test_transformed_instances_against_f64_dda():
    # cpu_test.zig:105-171: three transformed instances of the 64^3 test scene
    # (one masked to 0x2), random world rays (odd rays mask 0xFF, even 0x1);
    # reference: the f64 DDA in each instance's object space, nearest hit.
    # Mismatches <= rays / 5000, attributes exact, the object-space hit
    # point of a hit from outside lies on the voxel face within 2e-3.
    n = 64
    grid = scene_grid(n)
    dag = O.build_dense(6, grid)
    xf = [make_transform((0, 1, 0), 0.0, 1.0, (-80, 0, 0)),
          make_transform((1, 2, 3), 0.7, 0.5, (20, 10, -5)),
          make_transform((-1, 0.3, 0.2), 2.1, 1.7, (0, -40, 60))]
    s, inst = transformed_scene(dag, n, xf, [0xFF, 0xFF, 0x2])
    rays_n = 50000
    rng = np.random.default_rng(11)
    o = (rng.random((rays_n, 3)) * 2 - 1) * 200
    d = (rng.random((rays_n, 3)) * 2 - 1) * 70 - o
    rays = np.zeros(rays_n, O.RAY)
    rays["origin"], rays["direction"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 0.0, np.float32(3.402823466e38)
    masks = np.where(np.arange(rays_n) & 1, 0xFF, 0x1)
    got = np.concatenate([O.trace_rays(s, rays[masks == m], ray_mask=int(m)) for m in (0xFF, 0x1)])
    order = np.concatenate([np.nonzero(masks == m)[0] for m in (0xFF, 0x1)])
    hits = np.zeros(rays_n, O.HIT)
    hits[order] = got
    w2o = [np.array(inst[k].world_to_object[:], np.float64) for k in range(3)]
    mism = nhits = inside = 0
    for i in range(rays_n):
        of, df = o[i].astype(np.float32).astype(np.float64), d[i].astype(np.float32).astype(np.float64)
        best, best_k = None, 0xFFFFFFFF
        for k in range(3):
            if inst[k].mask & int(masks[i]) == 0:
                continue
            r = O.ref_dda(grid, n, apply_f64(w2o[k], of, 1), apply_f64(w2o[k], df, 0), 0.0,
                          best[0] if best else 1e30)
            if r is not None and (best is None or r[0] < best[0]):
                best, best_k = r, k
        h = hits[i]
        if (best is None) != (h["instance"] == 0xFFFFFFFF) or (
                best is not None and (h["instance"] != best_k or h["attribute"] != attr_of(*best[1]))):
            mism += 1
            continue
        if best is None:
            continue
        nhits += 1
        if h["meta"] & 0x20:
            inside += 1
            continue
        face = int(h["meta"] & 7)
        ax = face >> 1
        p = apply_f64(w2o[best_k], of, 1) + float(h["t"]) * apply_f64(w2o[best_k], df, 0)
        assert abs(p[ax] - (best[1][ax] + (0 if face & 1 else 1))) < 2e-3
    assert mism * 5000 <= rays_n
    assert nhits > rays_n // 20
    # every instance is hit, and the masked one only by 0xFF rays
    assert set(hits["instance"][hits["instance"] != 0xFFFFFFFF].tolist()) == {0, 1, 2}
    assert not np.any(hits["instance"][masks == 0x1] == 2)


def test_inverse_known_answer_and_against_lu():
    # xform.zig:60-68: m * inverse(m) applied to a point is the point; a
    # singular matrix has no inverse
    L = O.lib()
    m = (C.c_float * 12)(2, 0, 0, 1, 0, 0, -3, 2, 0, 1, 0, 3)
    inv = (C.c_float * 12)()
    assert L.or_inverse(m, inv) == 1
    p = np.array([0.5, -1.25, 7], np.float32)
    q = apply_f64(np.array(m[:]), apply_f64(np.array(inv[:]), p, 1), 1)
    assert np.allclose(q, p, atol=1e-5)
    assert L.or_inverse((C.c_float * 12)(1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0), inv) == 0
    # random well-conditioned transforms: equal to the f64 LU inverse rounded to f32
    rng = np.random.default_rng(3)
    for _ in range(200):
        m = make_transform(rng.normal(size=3), rng.uniform(0, 6), rng.uniform(0.2, 4), rng.uniform(-100, 100, 3))
        assert L.or_inverse((C.c_float * 12)(*m.tolist()), inv) == 1
        a = np.eye(4)
        a[:3, :] = m.reshape(3, 4)
        want = np.linalg.inv(a)[:3, :]
        assert np.allclose(np.array(inv[:]).reshape(3, 4), want, rtol=1e-6, atol=1e-6 * (1 + np.abs(want).max()))


def test_transformed_box_encloses_corners_brute_force():
    # xform.zig:37-58: world AABB of the cube [0, size]^3 under m (Arvo's
    # method) plus a margin of 1e-5 * max(1, |lo|, |hi|): equal to the box of
    # the eight transformed corners, widened by that margin
    L = O.lib()
    rng = np.random.default_rng(4)
    lo, hi = (C.c_float * 3)(), (C.c_float * 3)()
    for _ in range(300):
        m = rng.uniform(-3, 3, 12).astype(np.float32)
        size = float(np.float32(rng.choice([1.0, 16.0, 64.0, 4096.0])))
        L.or_transformed_box((C.c_float * 12)(*m.tolist()), size, lo, hi)
        corners = np.array([[x, y, z] for x in (0, size) for y in (0, size) for z in (0, size)])
        w = np.array([apply_f64(m, c, 1) for c in corners])
        cmin, cmax = w.min(0), w.max(0)
        eps = 1e-5 * np.maximum(1.0, np.maximum(np.abs(cmin), np.abs(cmax)))
        tol = 1e-6 * (np.abs(cmin) + np.abs(cmax) + size * 3 * np.abs(m).max() + 1)
        assert np.all(np.abs(np.array(lo[:]) - (cmin - eps)) <= tol)
        assert np.all(np.abs(np.array(hi[:]) - (cmax + eps)) <= tol)
        assert np.all(np.array(lo[:]) < cmin) and np.all(np.array(hi[:]) > cmax)


def test_render_depth_closed_form():
    # render.zig:143-166: depth = -z of the hit point in view space. For a
    # pinhole camera the primary ray of pixel (px, py) has view direction
    # (vx, vy, -1) / |(vx, vy, -1)|, so depth = t / sqrt(vx^2 + vy^2 + 1)
    # with vx = (2 px / w - 1) scale_x, vy = (1 - 2 py / h) scale_y
    d = O.build_dense(6, scene_grid(64))
    s = O.single_scene(d)
    w, h = 96, 64
    cam = O.camera((-40, 90, -50), (32, 20, 32), (0, 1, 0), 1.0, w, h)
    hits = np.zeros(w * h, O.HIT)
    depth = np.zeros(w * h, np.float32)
    O.lib().or_render_primary(C.byref(s), C.byref(cam), hits.ctypes.data_as(C.c_void_p),
                              depth.ctypes.data_as(C.c_void_p), 2)
    py, px = np.divmod(np.arange(w * h), w)
    vx = (2 * (px + 0.5) / w - 1) * cam.scale[0]
    vy = (1 - 2 * (py + 0.5) / h) * cam.scale[1]
    hit = hits["instance"] != 0xFFFFFFFF
    assert hit.sum() > w * h // 8 and (~hit).sum() > 0
    want = hits["t"].astype(np.float64) / np.sqrt(vx * vx + vy * vy + 1)
    assert np.allclose(depth[hit], want[hit], rtol=2e-5)
    assert np.all(depth[~hit] == np.float32(3.402823466e38))
    # hits equal the plain trace of the camera rays
    assert np.array_equal(hits.view(np.uint32), O.trace_rays(s, O.camera_rays(cam)).view(np.uint32))


def _morton(x, y, z):
    return O.lib().or_morton(int(x), int(y), int(z))


def _brute_prims(pts: np.ndarray, log2_size: int, k: int):
    """occupied cubes of edge 2^k in Morton order: (cell, lo, hi, attr_base,
    voxel count), from the voxel set alone"""
    cells = {}
    for p in pts:
        x, y, z = int(p["x"]), int(p["y"]), int(p["z"])
        c = (x >> k, y >> k, z >> k)
        lo, hi, n = cells.get(c, ((1 << 30,) * 3, (0, 0, 0), 0))
        cells[c] = (tuple(min(a, b) for a, b in zip(lo, (x, y, z))),
                    tuple(max(a, b + 1) for a, b in zip(hi, (x, y, z))), n + 1)
    out, rank = [], 0
    for c in sorted(cells, key=lambda c: _morton(*c)):
        lo, hi, n = cells[c]
        out.append((c, lo, hi, rank, n))
        rank += n
    return out


def _check_prims_brute_force(dag, pts, requested):
    k, prims, aabbs = O.rt_prims(dag, requested)
    assert k == min(max(requested, 3), dag.log2_size)
    want = _brute_prims(pts, dag.log2_size, k)
    assert len(prims) == len(want)
    eps = np.float32(1.0 / 256.0)
    by_cell = {}
    for p in pts:
        by_cell.setdefault((int(p["x"]) >> k, int(p["y"]) >> k, int(p["z"]) >> k), []).append(p)
    L = O.lib()
    for p, box, (c, lo, hi, rank, n) in zip(prims, aabbs, want):
        assert int(p["cell"]) == c[0] | c[1] << 21 | c[2] << 42
        assert int(p["attr_base"]) == rank
        assert np.array_equal(box[:3], np.float32(lo) - eps) and np.array_equal(box[3:], np.float32(hi) + eps)
        # the primitive's node is the subtree of exactly this cell's voxels:
        # its count word and the lookups relative to attr_base
        node = int(p["node"])
        if k > 2:
            assert int(dag.nodes[node + 1]) == n
        sub = dag.attrs[rank:] if dag.attrs is not None else None
        base = [v << k for v in c]
        for v in by_cell[c]:
            r = [int(v["x"]) - base[0], int(v["y"]) - base[1], int(v["z"]) - base[2]]
            got = L.or_lookup(O._p(dag.nodes), O._p(dag.leaves), O._p(sub), node, k, *r)
            assert got == int(v["attribute"])
    return prims


def test_rt_prims_known_answer():
    # rt_prims.zig:139-160
    pts = np.array([(1, 2, 3, 7), (30, 31, 17, 8), (31, 31, 17, 9)], O.VOXEL)
    d = O.build_points(5, pts)
    k, prims, aabbs = O.rt_prims(d, 4)
    assert k == 4 and len(prims) == 2
    assert abs(aabbs[0][0] - 1) < 0.01 and abs(aabbs[0][5] - 4) < 0.01
    assert prims[0]["attr_base"] == 0 and prims[0]["cell"] == 0
    assert prims[1]["cell"] == 1 | 1 << 21 | 1 << 42 and prims[1]["attr_base"] == 1
    assert abs(aabbs[1][0] - 30) < 0.01 and abs(aabbs[1][3] - 32) < 0.01


@pytest.mark.parametrize("log2,seed", [(5, 1), (6, 2), (7, 3)])
def test_rt_prims_brute_force_random(log2, seed):
    pts = final_points(random_points(3000, log2, seed, delete_frac=0.1), log2)
    d = O.build_points(log2, pts)
    for req in (0, 3, 4, log2 - 1, log2, log2 + 3):
        _check_prims_brute_force(d, pts, req)


def test_rt_prims_brute_force_dense_and_shared():
    # a dense scene and a periodic pattern whose nodes are shared by many
    # cells (the box memo per (level, node) of rt_prims.zig:61, 75-76)
    grid = scene_grid(64)
    d = O.build_dense(6, grid)
    for req in (3, 4, 5):
        _check_prims_brute_force(d, grid_to_points(grid), req)
    n = 64
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    per = ((x % 8) + (y % 8) < (z % 8) + 2).astype(np.uint32)
    d = O.build_dense(6, per)
    prims = _check_prims_brute_force(d, grid_to_points(per), 3)
    assert len(prims) == 512 and len(set(prims["node"].tolist())) == 1


def test_rt_prims_empty_and_single():
    d = O.build_points(5, np.zeros(0, O.VOXEL))
    for req in (3, 5):
        k, prims, _ = O.rt_prims(d, req)
        assert len(prims) == 0
    pts = np.array([(31, 0, 17, 5)], O.VOXEL)
    d = O.build_points(5, pts)
    for req in (3, 4, 5):
        _check_prims_brute_force(d, pts, req)


SKIP = 0x8


def _material_set(materials):
    m = [0, 0, 0, 0]
    for i in materials:
        m[i >> 6] |= 1 << (i & 63)
    return m


def _probe_rays(n, count, seed):
    """random rays through the cube plus axis-aligned ones (the DDA test's mix)"""
    rng = np.random.default_rng(seed)
    rays = np.zeros(count, O.RAY)
    for i in range(count):
        o = n * 0.5 + (rng.random(3) * 2 - 1) * n * 1.5
        dv = n * 0.5 + (rng.random(3) * 2 - 1) * n * 0.45 - o
        if i % 8 == 0:
            dv[i % 3] = 0
        rays[i]["origin"], rays[i]["direction"] = o, dv
    rays["tmin"], rays["tmax"] = 0.0, np.float32(3.0e38)
    return rays


def test_skip_transparent_equals_removing_the_voxels():
    """traceSkipping (dag.zig:236-289) with transparent materials: every hit
    (t, instance, face, inside, attribute) equals the plain walk of the DAG
    built without those voxels; an empty set equals the plain walk."""
    n = 64
    grid = scene_grid(n)
    trans = {m for m in range(256) if m % 5 == 0}
    mat = grid & 0xFF
    removed = np.where(np.isin(mat, list(trans)), 0, grid).astype(np.uint32)
    assert (grid != 0).sum() > (removed != 0).sum() > 0
    full, rem = O.build_dense(6, grid), O.build_dense(6, removed)
    s_full, s_rem = O.single_scene(full), O.single_scene(rem)
    s_full.transparent[:] = _material_set(trans)
    rays = np.concatenate([_probe_rays(n, 6000, 3),
                           O.camera_rays(O.camera((-32, 96, -32), (32, 32, 32), (0, 1, 0), 1.0, 64, 64))])
    # rays starting inside voxels too: origins on voxel centres
    inside = _probe_rays(n, 2000, 4)
    z, y, x = np.nonzero(grid)
    pick = np.random.default_rng(5).integers(0, len(x), len(inside))
    inside["origin"] = np.stack([x[pick], y[pick], z[pick]], 1) + 0.5
    rays = np.concatenate([rays, inside])
    a = O.trace_rays(s_full, rays, flags=SKIP)
    b = O.trace_rays(s_rem, rays)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (a["instance"] == 0).sum() > len(rays) // 4
    plain = O.trace_rays(s_full, rays)
    assert not np.array_equal(plain.view(np.uint32), a.view(np.uint32))
    # without the flag the set is ignored; with an empty set the flag changes nothing
    s_full.transparent[:] = [0, 0, 0, 0]
    assert np.array_equal(O.trace_rays(s_full, rays, flags=SKIP).view(np.uint32), plain.view(np.uint32))


def test_cutout_holes():
    """Cutout materials (dag.zig:133-151, trace.zig:161-173): rays pass about
    30 % (77/256) of the entry points on cutout voxels, a ray passing a hole
    goes on to the same hit as with that voxel removed or further, a ray
    starting inside a cutout voxel passes it, and an instance whose voxel edge
    exceeds the limit traces as if there were no cutouts."""
    n = 64
    grid = scene_grid(n)
    cut = {m for m in range(256) if m % 3 == 0}
    d = O.build_dense(6, grid)
    s = O.single_scene(d)
    rays = np.concatenate([_probe_rays(n, 8000, 6),
                           O.camera_rays(O.camera((-32, 96, -32), (32, 32, 32), (0, 1, 0), 1.0, 96, 96))])
    plain = O.trace_rays(s, rays)
    s.cutout[:] = _material_set(cut)
    got = O.trace_rays(s, rays)
    on_cut = (plain["instance"] == 0) & np.isin(plain["attribute"] & 0xFF, list(cut))
    passed = on_cut & (got.view(np.uint32).reshape(-1, 4) != plain.view(np.uint32).reshape(-1, 4)).any(1)
    frac = passed.sum() / on_cut.sum()
    assert on_cut.sum() > 2000 and 0.25 < frac < 0.36, (on_cut.sum(), frac)
    # rays that did not first meet a cutout voxel are unchanged
    same = ~on_cut
    assert np.array_equal(got[same].view(np.uint32), plain[same].view(np.uint32))
    # a passed ray hits later (or misses)
    later = passed & (got["instance"] == 0)
    assert np.all(got["t"][later] >= plain["t"][later])
    # starting inside a cutout voxel: the plain walk reports an inside hit at
    # tmin, the cutout walk never stops in that voxel
    z, y, x = np.nonzero(np.isin(grid & 0xFF, list(cut)) & (grid != 0))
    start = _probe_rays(n, 500, 7)
    start["origin"] = np.stack([x[:500], y[:500], z[:500]], 1) + 0.5
    p0 = O.trace_rays(O.single_scene(d), start)
    assert np.all(p0["meta"] & 0x20)
    g0 = O.trace_rays(s, start)
    assert not np.any(g0["meta"] & 0x20)
    # voxel edge 2 world units: w2o row 0 = (0.5, 0, 0), squared norm 0.25 <= 0.5
    s2 = O.single_scene(d)
    m = [2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0]
    w = [0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0]
    s2._keep[1][0].object_to_world[:] = m
    s2._keep[1][0].world_to_object[:] = w
    O.lib().or_transformed_box(s2._keep[1][0].object_to_world, float(n), s2._keep[1][0].bounds_min,
                               s2._keep[1][0].bounds_max)
    big = rays.copy()
    big["origin"] *= 2
    a = O.trace_rays(s2, big)
    s2.cutout[:] = _material_set(cut)
    assert np.array_equal(O.trace_rays(s2, big).view(np.uint32), a.view(np.uint32))


def test_extended_hits_against_f64_dda():
    """extendedHit (trace.zig:125-157) on three transformed instances of the
    64^3 scene: the hit part equals the plain trace, the voxel equals the f64
    DDA's voxel in the hit instance's object space (the start voxel for rays
    starting inside), the position lies on the f64 ray at the DDA's t, the
    normal equals object_to_world times the face normal for these similarity
    transforms and faces the ray, and misses are zero but for the hit."""
    n = 64
    grid = scene_grid(n)
    dag = O.build_dense(6, grid)
    xf = [make_transform((0, 1, 0), 0.0, 1.0, (-80, 0, 0)),
          make_transform((1, 2, 3), 0.7, 0.5, (20, 10, -5)),
          make_transform((-1, 0.3, 0.2), 2.1, 1.7, (0, -40, 60))]
    s, inst = transformed_scene(dag, n, xf, [0xFF, 0xFF, 0xFF])
    rays_n = 12000
    rng = np.random.default_rng(21)
    o = (rng.random((rays_n, 3)) * 2 - 1) * 200
    d = (rng.random((rays_n, 3)) * 2 - 1) * 70 - o
    # a tenth of the rays start inside a filled voxel of instance 0
    z, y, x = np.nonzero(grid)
    pick = rng.integers(0, len(x), rays_n // 10)
    o[: rays_n // 10] = np.stack([x[pick], y[pick], z[pick]], 1) + 0.5 + np.array([-80, 0, 0])
    rays = np.zeros(rays_n, O.RAY)
    rays["origin"], rays["direction"] = o.astype(np.float32), d.astype(np.float32)
    rays["tmin"], rays["tmax"] = 0.0, np.float32(3.402823466e38)
    ex = O.trace_rays_ex(s, rays)
    plain = O.trace_rays(s, rays)
    assert np.array_equal(ex.view(np.uint32).reshape(-1, 16)[:, :4], plain.view(np.uint32).reshape(-1, 4))
    miss = plain["instance"] == 0xFFFFFFFF
    rest = ex[miss].view(np.uint32).reshape(-1, 16)[:, 4:]
    assert not rest.any()
    w2o = [np.array(inst[k].world_to_object[:], np.float64) for k in range(3)]
    o2w = [np.array(inst[k].object_to_world[:], np.float64).reshape(3, 4)[:, :3] for k in range(3)]
    off = nhits = inside = 0
    for i in np.nonzero(~miss)[0]:
        h, e = plain[i], ex[i]
        k = int(h["instance"])
        of, df = rays[i]["origin"].astype(np.float64), rays[i]["direction"].astype(np.float64)
        oo, dd = apply_f64(w2o[k], of, 1), apply_f64(w2o[k], df, 0)
        face, ax = int(h["meta"] & 7), int(h["meta"] & 7) >> 1
        nobj = np.zeros(3)
        nobj[ax] = -1.0 if face & 1 else 1.0
        nw = o2w[k] @ nobj
        assert np.abs(e["normal"] - nw / np.sqrt(nw @ nw)).max() < 1e-6
        pw = of + df * float(h["t"])
        assert np.abs(e["position"] - pw).max() <= 1e-5 * max(1.0, np.abs(pw).max())
        if h["meta"] & 0x20:
            inside += 1
            assert tuple(e["voxel"]) == tuple(np.floor(oo).astype(int))
            continue
        nhits += 1
        assert float(e["normal"] @ df) < 0
        r = O.ref_dda(grid, n, oo, dd, 0.0, 1e30)
        if r is None or tuple(e["voxel"]) != r[1]:
            off += 1
    assert nhits > rays_n // 20 and inside >= rays_n // 10 - 5
    assert off * 2000 <= nhits, (off, nhits)
