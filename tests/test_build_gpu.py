"""Parity of the HIP build against the oracle: GPU-native arrays bit-exact
against the serial restatement of the reference's GPU build, and canonical
form equal to the reference's CPU builder (src/dag_builder.zig)."""
import numpy as np
import pytest

import oracle_py as O
from scenes import final_points, grid_to_points, random_points, scene_grid

pytestmark = pytest.mark.gpu


def _gpu():
    from pyrit_amd import gpu

    return gpu


def assert_same_build(b, ref, log2):
    assert b.voxel_count == ref["voxel_count"]
    assert b.root == ref["root"]
    assert b.node_words == len(ref["nodes"])
    assert b.leaf_count == len(ref["leaves"])
    assert np.array_equal(b.nodes(), ref["nodes"])
    assert np.array_equal(b.leaves(), ref["leaves"])
    assert np.array_equal(b.attrs(), ref["attrs"])
    assert np.array_equal(b.keys(), ref["keys"])


def test_terrain_chunk_64_bit_exact(hip_device):
    from pyrit_amd import workloads

    print("case: terrain chunk 64^3 build", flush=True)
    vox, _ = workloads.terrain_chunk64().generate_host()
    ref = O.gb_build(6, vox)
    b = _gpu().build(6, vox)
    assert_same_build(b, ref, 6)
    c = O.canon(b.nodes(), b.leaves(), b.root, 6)
    cpu = O.build_points(6, vox)
    assert c.same(cpu)
    assert np.array_equal(b.attrs(), cpu.attrs)


@pytest.mark.parametrize("log2,n,seed", [(3, 50, 1), (4, 300, 2), (5, 3000, 3), (6, 20000, 4), (7, 60000, 5),
                                         (8, 200000, 6), (10, 100000, 7), (12, 50000, 8), (20, 30000, 9)])
def test_random_points_bit_exact(hip_device, log2, n, seed):
    print(f"case: random points log2={log2} n={n} seed={seed}", flush=True)
    pts = random_points(n, log2, seed, dup_frac=0.2, delete_frac=0.05, out_frac=0.02)
    ref = O.gb_build(log2, pts)
    b = _gpu().build(log2, pts)
    assert_same_build(b, ref, log2)
    cpu = O.build_points(log2, final_points(pts, log2))
    assert O.canon(b.nodes(), b.leaves(), b.root, log2).same(cpu)


@pytest.mark.parametrize("case", ["low corner", "two corners", "no invalid 256", "with invalid 256"])
def test_passes_that_move_nothing(hip_device, case):
    """Radix passes whose digit is the same for every record are skipped on
    the device; the result must not depend on which passes ran."""
    print(f"case: skipped passes, {case}", flush=True)
    rng = np.random.default_rng(11)
    if case == "low corner":  # 1024^3 grid, coordinates < 8: only the lowest digits vary
        log2, pts = 10, random_points(4000, 3, 12, dup_frac=0.3, delete_frac=0.1)
    elif case == "two corners":  # middle digits constant, the top one varies
        log2 = 10
        a = random_points(3000, 2, 13, dup_frac=0.3, delete_frac=0.1)
        b = random_points(3000, 2, 14, dup_frac=0.3, delete_frac=0.1)
        for c in ("x", "y", "z"):
            b[c] += 512
        pts = np.concatenate([a, b])
        pts = pts[rng.permutation(len(pts))]
    elif case == "no invalid 256":  # 25-bit keys: the top pass only holds the invalid bit
        log2, pts = 8, random_points(100000, 8, 15, dup_frac=0.2, delete_frac=0.05)
    else:
        log2, pts = 8, random_points(100000, 8, 16, dup_frac=0.2, delete_frac=0.05, out_frac=0.01)
    ref = O.gb_build(log2, pts)
    b = _gpu().build(log2, pts)
    assert_same_build(b, ref, log2)
    # a batch whose chunks all skip the same passes
    offs = np.array([0, len(pts) // 3, len(pts)], np.uint32)
    rb_, rr, rf = O.gb_build_chunks(log2, pts, offs)
    bb, roots, first = _gpu().build_chunks(log2, pts, offs)
    assert_same_build(bb, rb_, log2)
    assert np.array_equal(roots, rr) and np.array_equal(first, rf)


@pytest.mark.parametrize("radix_bits", [8, 9, 10, 11])
def test_radix_widths_identical(hip_device, radix_bits):
    print(f"case: radix bits {radix_bits}", flush=True)
    pts = random_points(120000, 9, 77, dup_frac=0.3, delete_frac=0.1)
    ref = O.gb_build(9, pts)
    b = _gpu().build(9, pts, radix_bits=radix_bits)
    assert_same_build(b, ref, 9)


@pytest.mark.parametrize("sort_items", [8, 12, 16, 20, 24])
@pytest.mark.parametrize("pair_layout", [0, 1, 2])
def test_tile_sizes_and_layouts_identical(hip_device, sort_items, pair_layout):
    # log2 9: 27-bit keys, so layout 0 sorts packed key << 32 | attribute words
    print(f"case: sort_items {sort_items} pair_layout {pair_layout}", flush=True)
    pts = random_points(90000, 9, 78, dup_frac=0.3, delete_frac=0.1, out_frac=0.01)
    ref = O.gb_build(9, pts)
    b = _gpu().build(9, pts, sort_items=sort_items, pair_layout=pair_layout)
    assert_same_build(b, ref, 9)
    # edits through the same tuning
    ed = random_points(7000, 9, 79, dup_frac=0.2, delete_frac=0.3)
    b2 = _gpu().build(9, ed, old=b, sort_items=sort_items, pair_layout=pair_layout)
    assert_same_build(b2, O.gb_build(9, ed, ref["keys"], ref["attrs"]), 9)


@pytest.mark.parametrize("scan_items", [8, 12, 16, 20])
@pytest.mark.parametrize("sort_items", [12, 20])
def test_scan_tile_sizes_identical(hip_device, scan_items, sort_items):
    # unique / brick scan tiles of 256 * scan_items records, cut from sort tiles
    # of 256 * sort_items (pieces of unequal length), single and batched
    print(f"case: scan_items {scan_items} sort_items {sort_items}", flush=True)
    pts = random_points(150000, 8, 80 + scan_items, dup_frac=0.3, delete_frac=0.1, out_frac=0.01)
    ref = O.gb_build(8, pts)
    b = _gpu().build(8, pts, sort_items=sort_items, scan_items=scan_items)
    assert_same_build(b, ref, 8)


@pytest.mark.parametrize("counted_pass", [1, 2])
def test_counted_and_lookback_first_pass_identical(hip_device, counted_pass):
    """The first radix pass takes its tile offsets from the encode step's
    per-tile counts (1) or from the decoupled look-back (2): single builds
    with out-of-range records, edits, both finishes, batches (tiles of many
    chunks in one encode run) and mixed chunk sizes, against the oracle."""
    from pyrit_amd import workloads

    g = _gpu()
    print(f"case: counted pass {counted_pass}, single and edit", flush=True)
    pts = random_points(400000, 9, 91, dup_frac=0.3, delete_frac=0.1, out_frac=0.01)
    ref = O.gb_build(9, pts)
    for finish in (1, 2):
        b = g.build(9, pts, counted_pass=counted_pass, finish=finish)
        assert_same_build(b, ref, 9)
    ed = random_points(30000, 9, 92, dup_frac=0.2, delete_frac=0.3)
    b2 = g.build(9, ed, old=b, counted_pass=counted_pass)
    assert_same_build(b2, O.gb_build(9, ed, ref["keys"], ref["attrs"]), 9)
    print(f"case: counted pass {counted_pass}, terrain then scattered records", flush=True)
    # coherent terrain records followed by random ones
    tv, _ = workloads.heightfield_grid(log2_size=8, caves=60).generate_host()
    mix = np.concatenate([tv, random_points(300000, 8, 95, dup_frac=0.2, delete_frac=0.05)])
    assert_same_build(g.build(8, tv, counted_pass=counted_pass), O.gb_build(8, tv), 8)
    assert_same_build(g.build(8, mix, counted_pass=counted_pass), O.gb_build(8, mix), 8)
    print(f"case: counted pass {counted_pass}, packed batch of 700 chunks", flush=True)
    rng = np.random.default_rng(93)
    chunks = [random_points(int(rng.integers(0, 9000)), 6, int(rng.integers(1 << 30)), dup_frac=0.1,
                            delete_frac=0.05, out_frac=0.01) for _ in range(700)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.uint32)
    vox = np.concatenate(chunks)
    rb_, rr, rf = O.gb_build_chunks(6, vox, offs)
    for finish in (1, 2):
        bb, roots, first = g.build_chunks(6, vox, offs, counted_pass=counted_pass, finish=finish)
        assert_same_build(bb, rb_, 6)
        assert np.array_equal(roots, rr) and np.array_equal(first, rf)
    print(f"case: counted pass {counted_pass}, mixed sizes", flush=True)
    w = workloads.uneven_chunks(seed=94, count=48, edges=(8, 16, 32, 64, 128))
    vox, offs64 = w.generate_host()
    offs = offs64.astype(np.uint32)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    rb_, rr, rf = O.gb_build_chunks(7, vox, offs, chunk_log2=sizes)
    bb, roots, first = g.build_chunks(7, vox, offs, chunk_log2=sizes, counted_pass=counted_pass)
    assert_same_build(bb, rb_, 7)
    assert np.array_equal(roots, rr) and np.array_equal(first, rf)


@pytest.mark.parametrize("levels", [1, 2])
def test_level_paths_identical(hip_device, levels):
    """Node levels built by four launches per level over the GPU (1) or by one
    workgroup in one launch (2): single builds from 2^3 to 2^12 with
    duplicates, deletes and out-of-range records, an edit, an empty build,
    a batch whose chunks end at different levels (roots recorded on the way)
    and a dense solid grid (hot tuples), against the oracle."""
    from pyrit_amd import workloads

    g = _gpu()
    print(f"case: levels {levels}, single builds", flush=True)
    for log2, n, seed in [(3, 40, 1), (5, 3000, 2), (7, 40000, 3), (9, 120000, 4), (12, 60000, 5)]:
        pts = random_points(n, log2, seed, dup_frac=0.2, delete_frac=0.05, out_frac=0.02)
        ref = O.gb_build(log2, pts)
        b = g.build(log2, pts, levels=levels)
        assert_same_build(b, ref, log2)
    ed = random_points(5000, 12, 6, dup_frac=0.2, delete_frac=0.3)
    assert_same_build(g.build(12, ed, old=b, levels=levels), O.gb_build(12, ed, ref["keys"], ref["attrs"]), 12)
    empty = np.zeros(0, O.VOXEL)
    assert_same_build(g.build(6, empty, levels=levels), O.gb_build(6, empty), 6)
    solid = grid_to_points(np.ones((32, 32, 32), np.uint32) * 3)
    assert_same_build(g.build(5, solid, levels=levels), O.gb_build(5, solid), 5)
    print(f"case: levels {levels}, batch of mixed sizes", flush=True)
    w = workloads.uneven_chunks(seed=97, count=40, edges=(8, 16, 32, 64))
    vox, offs64 = w.generate_host()
    offs = offs64.astype(np.uint32)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    rb_, rr, rf = O.gb_build_chunks(6, vox, offs, chunk_log2=sizes)
    bb, roots, first = g.build_chunks(6, vox, offs, chunk_log2=sizes, levels=levels)
    assert_same_build(bb, rb_, 6)
    assert np.array_equal(roots, rr) and np.array_equal(first, rf)


def test_build_memory_per_voxel(hip_device):
    """The reference's bound on the build's device memory (tests/cpu_test.zig:
    1559-1629): every buffer the build requests, scratch and outputs, held
    at once, below 130 bytes per voxel for the surface of a sphere in a 256^3
    grid (y = 127.5 + sqrt(110^2 - dx^2 - dz^2) per column, f32)."""
    print("case: build memory per voxel, sphere surface 256^3", flush=True)
    g = _gpu()
    c = np.arange(256, dtype=np.float32) - np.float32(127.5)
    dz, dx = np.meshgrid(c, c, indexing="ij")
    r2 = np.float32(110.0 * 110.0) - dx * dx - dz * dz
    z, x = np.nonzero(r2 > 0)
    y = (np.float32(127.5) + np.sqrt(r2[z, x])).astype(np.uint32)
    pts = np.zeros(len(x), O.VOXEL)
    pts["x"], pts["y"], pts["z"], pts["attribute"] = x, y, z, 0x00FF00FF
    g.set_timing(True)
    try:
        b = g.build(8, pts)
        t = g.last_build_times()
    finally:
        g.set_timing(False)
    per_voxel = t["device_bytes"] / len(pts)
    print(f"  {len(pts)} voxels, {t['device_bytes'] / 2**20:.1f} MiB, {per_voxel:.0f} bytes per voxel", flush=True)
    assert b.voxel_count == len(pts)
    assert per_voxel < 130
    assert_same_build(b, O.gb_build(8, pts), 8)


@pytest.mark.parametrize("load", [25, 50, 95])
def test_hash_load_identical(hip_device, load):
    # the hash capacity changes probe lengths only: owners are the smallest
    # index with a mask / tuple, so the layout stays the same
    print(f"case: hash load {load}%", flush=True)
    pts = random_points(200000, 9, 31, dup_frac=0.2, delete_frac=0.05)
    ref = O.gb_build(9, pts)
    assert_same_build(_gpu().build(9, pts, hash_load_pct=load), ref, 9)
    offs = np.array([0, 70000, 70000, len(pts)], np.uint32)
    rb_, rr, rf = O.gb_build_chunks(9, pts, offs)
    bb, roots, first = _gpu().build_chunks(9, pts, offs, hash_load_pct=load)
    assert_same_build(bb, rb_, 9)
    assert np.array_equal(roots, rr) and np.array_equal(first, rf)


@pytest.mark.parametrize("status_words", [1, 2])
def test_lookback_word_widths_identical(hip_device, status_words):
    # 64-bit epoch-tagged and 32-bit cleared look-back words, single and batched
    print(f"case: look-back words {status_words}", flush=True)
    pts = random_points(250000, 9, 33, dup_frac=0.2, delete_frac=0.05, out_frac=0.01)
    assert_same_build(_gpu().build(9, pts, status_words=status_words), O.gb_build(9, pts), 9)
    offs = np.array([0, 1000, 90000, 90000, len(pts)], np.uint32)
    rb_, rr, rf = O.gb_build_chunks(9, pts, offs)
    bb, roots, first = _gpu().build_chunks(9, pts, offs, status_words=status_words)
    assert_same_build(bb, rb_, 9)
    assert np.array_equal(roots, rr) and np.array_equal(first, rf)


@pytest.mark.parametrize("log2,n,seed,layout", [(3, 50, 1, 0), (4, 300, 2, 0), (5, 3000, 3, 0), (6, 20000, 4, 0),
                                                (8, 200000, 6, 0), (9, 90000, 7, 1), (10, 100000, 8, 0),
                                                (11, 40000, 9, 0)])
def test_bucket_finish_bit_exact(hip_device, log2, n, seed, layout):
    """Partial sort (key bits from 14 up) + one key table per bucket: the
    same arrays as the full sort and unique scan, for packed and separate
    records (log2 9 with separate arrays, log2 11 with 34-bit keys)."""
    print(f"case: bucket finish log2={log2} n={n} seed={seed} layout={layout}", flush=True)
    pts = random_points(n, log2, seed, dup_frac=0.2, delete_frac=0.05, out_frac=0.02)
    ref = O.gb_build(log2, pts)
    g = _gpu()
    for finish in (1, 2):
        b = g.build(log2, pts, finish=finish, pair_layout=layout)
        assert_same_build(b, ref, log2)
    g.set_timing(True)
    g.build(log2, pts, finish=2, pair_layout=layout)
    assert g.last_build_times()["finish"] == 2
    g.set_timing(False)
    # an edit build through the bucket finish
    ed = random_points(max(n // 10, 5), log2, seed + 100, dup_frac=0.2, delete_frac=0.3)
    b2 = g.build(log2, ed, old=b, finish=2, pair_layout=layout)
    assert_same_build(b2, O.gb_build(log2, ed, ref["keys"], ref["attrs"]), log2)


@pytest.mark.parametrize("case", ["solid 64", "one key repeated", "low corner 1024", "terrain chunk 64"])
def test_bucket_finish_edge_cases(hip_device, case):
    """Buckets whose 16384 slots are all set, a bucket of far more records
    than slots (one key 100000 times, the last wins), every record in bucket
    0 of a 1024^3 grid with every pass skipped, and a real terrain chunk."""
    print(f"case: bucket finish, {case}", flush=True)
    rng = np.random.default_rng(21)
    if case == "solid 64":
        log2 = 6
        pts = grid_to_points(np.ones((64, 64, 64), np.uint32) * 7)
        pts = pts[rng.permutation(len(pts))]
    elif case == "one key repeated":
        log2 = 7
        pts = np.zeros(100000, O.VOXEL)
        pts["x"], pts["y"], pts["z"] = 100, 3, 77
        pts["attribute"] = rng.integers(1, 1000, len(pts))
        pts = np.concatenate([pts, random_points(5000, 7, 22, dup_frac=0.2, delete_frac=0.1)])
    elif case == "low corner 1024":
        log2, pts = 10, random_points(4000, 3, 12, dup_frac=0.3, delete_frac=0.1)
    else:
        from pyrit_amd import workloads

        log2 = 6
        pts, _ = workloads.terrain_chunk64().generate_host()
    ref = O.gb_build(log2, pts)
    assert_same_build(_gpu().build(log2, pts, finish=2), ref, log2)


@pytest.mark.parametrize("case", ["all out of range", "one voxel", "deletes only", "last bucket", "empty chunks"])
def test_bucket_finish_degenerate_inputs(hip_device, case):
    """Inputs that leave buckets empty or drop every record: the bucket
    finish must give the oracle's arrays (an empty DAG where nothing stays)."""
    print(f"case: bucket finish, {case}", flush=True)
    g = _gpu()
    rng = np.random.default_rng(33)
    if case in ("all out of range", "one voxel", "deletes only", "last bucket"):
        log2 = 10
        if case == "all out of range":
            pts = random_points(3000, 10, 34)
            pts["x"] = np.where(rng.random(len(pts)) < 0.5, -1, 1024)
        elif case == "one voxel":
            pts = np.array([(1023, 0, 511, 77)], O.VOXEL)
        elif case == "deletes only":
            pts = random_points(3000, 10, 35)
            pts["attribute"] = 0
        else:  # the top corner: the last bucket of the grid, then deleted again
            pts = np.array([(1023, 1023, 1023, 5), (1022, 1023, 1023, 6), (1023, 1023, 1023, 0)], O.VOXEL)
        ref = O.gb_build(log2, pts)
        assert_same_build(g.build(log2, pts, finish=2), ref, log2)
        return
    # a batch whose chunks are empty, hold only out-of-range records, or one voxel
    parts = [np.zeros(0, O.VOXEL), random_points(500, 6, 36), np.zeros(0, O.VOXEL),
             np.array([(70, 0, 0, 3), (-1, 5, 5, 4)], O.VOXEL), np.array([(63, 63, 63, 9)], O.VOXEL)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in parts])]).astype(np.uint32)
    vox = np.concatenate(parts)
    ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs)
    b, roots, first = g.build_chunks(6, vox, offs, finish=2)
    assert_same_build(b, ref, 6)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    assert roots[0] == roots[2] == roots[3] == 0xFFFFFFFF


def test_tuning_argument_checks(hip_device):
    g = _gpu()
    pts = random_points(1000, 5, 3)
    for kw in ({"hash_load_pct": 10}, {"hash_load_pct": 99}, {"pair_layout": 3}, {"radix_bits": 12},
               {"sort_items": 10}, {"scan_items": 24}, {"status_words": 3}, {"finish": 3}, {"counted_pass": 3}):
        with pytest.raises(RuntimeError, match="INVALID_ARGUMENT"):
            g.build(5, pts, **kw)
    # the bucket finish: single builds of at most 2^20 buckets (log2 <= 11)
    with pytest.raises(RuntimeError, match="INVALID_ARGUMENT"):
        g.build(12, random_points(1000, 12, 3), finish=2)


def test_autotune_installs_an_identical_setting(hip_device):
    print("case: autotune single build and batch", flush=True)
    g = _gpu()
    default = {"radix_bits": 8, "sort_items": 20, "pair_layout": 2, "scan_items": 12, "hash_load_pct": 80,
               "status_words": 2, "finish": 0, "counted_pass": 1, "levels": 0}
    g.tuning_clear()
    try:
        pts = random_points(300000, 10, 41, dup_frac=0.1, delete_frac=0.02)
        rep = g.autotune(10, pts, repeats=1)
        print(rep, flush=True)
        assert rep["rejected"] == 0 and rep["candidates"] >= 10
        assert rep["best_ms"] <= rep["default_ms"]
        assert g.tuning_get(**rep["class"]) == rep["best"]
        # same class (bit length of the count), other class untouched
        assert g.tuning_get(10, 300000 - 5000, 0) == rep["best"]
        assert g.tuning_get(10, 3000, 0) == default
        # builds that leave the fields 0 use the installed setting: same output
        assert_same_build(g.build(10, pts), O.gb_build(10, pts), 10)
        rng = np.random.default_rng(7)
        chunks = [random_points(int(rng.integers(0, 3000)), 6, int(rng.integers(1 << 30)), dup_frac=0.1)
                  for _ in range(200)]
        offs = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.uint32)
        vox = np.concatenate(chunks)
        rep2 = g.autotune(6, vox, offsets=offs, repeats=1)
        print(rep2, flush=True)
        assert rep2["rejected"] == 0 and rep2["class"]["chunk_count"] == 200
        assert g.tuning_get(**rep2["class"]) == rep2["best"]
        ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs)
        b, roots, first = g.build_chunks(6, vox, offs)
        assert_same_build(b, ref, 6)
        assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
        # a saved setting installs again; None removes it
        g.tuning_set(10, 300000, 0, None)
        assert g.tuning_get(10, 300000, 0) == default
        g.tuning_set(10, 300000, 0, rep["best"])
        assert g.tuning_get(10, 300000, 0) == rep["best"]
    finally:
        g.tuning_clear()


def test_scene_dense_and_empty_and_single(hip_device):
    print("case: scene 64^3 / empty / single voxel", flush=True)
    g = _gpu()
    grid = scene_grid(64)
    pts = grid_to_points(grid)
    np.random.default_rng(3).shuffle(pts)
    b = g.build(6, pts)
    assert_same_build(b, O.gb_build(6, pts), 6)
    e = g.build(4, np.zeros(0, O.VOXEL))
    assert e.voxel_count == 0 and e.nodes().tolist() == [0, 0] and e.root == 0
    s = g.build(4, np.array([(5, 6, 7, 42)], O.VOXEL))
    assert_same_build(s, O.gb_build(4, np.array([(5, 6, 7, 42)], O.VOXEL)), 4)
    # all records deleted or out of range
    z = g.build(5, np.array([(1, 1, 1, 0), (99, 0, 0, 3), (-1, 2, 2, 4)], O.VOXEL))
    assert z.voxel_count == 0 and z.nodes().tolist() == [0, 0]


def test_wave_leader_mask_high_bits(hip_device):
    # regression: one wave of bricks whose first brick has mask 0x80000000 and
    # whose others have 0xFFFFFFFF80000000 - distinct leaves that a
    # sign-extended lane broadcast would merge
    print("case: leaf masks differing only in the high dword", flush=True)
    pts = []
    for m in range(64):
        bx, by, bz = _decode(m)
        bits = [31] if m == 0 else [31] + list(range(32, 64))
        for bit in bits:
            pts.append((bx * 4 + (bit & 3), by * 4 + ((bit >> 2) & 3), bz * 4 + (bit >> 4), 5 + m))
    pts = np.array(pts, O.VOXEL)
    ref = O.gb_build(5, pts)
    b = _gpu().build(5, pts)
    assert_same_build(b, ref, 5)
    assert len(ref["leaves"]) == 2


def _decode(m):
    import ctypes as C

    out = (C.c_uint32 * 3)()
    O.lib().or_morton_decode(m, out)
    return tuple(out)


def test_edit_equals_rebuild(hip_device):
    # cpu_test.zig:809-850 on the GPU path
    print("case: edit = rebuild", flush=True)
    g = _gpu()
    base = np.array([(x, 0, z, 0x10) for z in range(32) for x in range(32)], O.VOXEL)
    b1 = g.build(5, base)
    edits = np.array([(10, y, 10, 0x20) for y in range(1, 20)] + [(x, 0, 3, 0) for x in range(8)]
                     + [(20, 0, 20, 0x30)], O.VOXEL)
    b2 = g.build(5, edits, old=b1)
    r1 = O.gb_build(5, base)
    r2 = O.gb_build(5, edits, r1["keys"], r1["attrs"])
    assert_same_build(b2, r2, 5)
    # large random edit on a large base
    pts = random_points(50000, 8, 11)
    b3 = g.build(8, pts)
    ed = random_points(5000, 8, 12, dup_frac=0.1, delete_frac=0.3)
    b4 = g.build(8, ed, old=b3)
    r3 = O.gb_build(8, pts)
    assert_same_build(b4, O.gb_build(8, ed, r3["keys"], r3["attrs"]), 8)


def _shell_batch(k=5, cap=8000, empty=3):
    from test_oracle import _shell_batch as sb

    return sb(k, cap, empty)


def test_batch_reference_layout_bit_exact(hip_device):
    # cpu_test.zig:1001-1060 layout (segments of `capacity`)
    print("case: chunk batch, capacity layout", flush=True)
    vox, cnt, offs, cap = _shell_batch()
    ref, rroots, rfirst = O.gb_build_chunks(5, vox, offs, capacity=cap)
    b, roots, first = _gpu().build_chunks(5, vox, offs, capacity=cap)
    assert_same_build(b, ref, 5)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    assert roots[3] == 0xFFFFFFFF


def test_batch_packed_many_chunks(hip_device):
    print("case: packed batch of 300 chunks", flush=True)
    rng = np.random.default_rng(4)
    chunks = [random_points(int(rng.integers(0, 800)), 5, int(rng.integers(1 << 30)), dup_frac=0.1,
                            delete_frac=0.05) for _ in range(300)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.uint32)
    vox = np.concatenate(chunks)
    ref, rroots, rfirst = O.gb_build_chunks(5, vox, offs)
    b, roots, first = _gpu().build_chunks(5, vox, offs)
    assert_same_build(b, ref, 5)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)


def test_mixed_size_batch(hip_device):
    print("case: mixed-size batch", flush=True)
    from pyrit_amd import workloads

    w = workloads.uneven_chunks(seed=5, count=24, edges=(8, 16, 32, 64))
    vox, offs64 = w.generate_host()
    offs = offs64.astype(np.uint32)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs, chunk_log2=sizes)
    b, roots, first = _gpu().build_chunks(6, vox, offs, chunk_log2=sizes)
    assert_same_build(b, ref, 6)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    nodes, leaves = b.nodes(), b.leaves()
    attrs = b.attrs()
    for c in range(len(sizes)):
        seg = vox[offs[c]:offs[c + 1]]
        if len(seg) == 0:
            assert roots[c] == 0xFFFFFFFF
            continue
        cpu = O.build_points(int(sizes[c]), seg)
        got = O.canon(nodes, leaves, int(roots[c]), int(sizes[c]))
        assert got.same(cpu), c
        assert np.array_equal(attrs[first[c]:first[c] + cpu.voxel_count], cpu.attrs)


@pytest.mark.parametrize("finish", [1, 2])
def test_batch_bucket_finish_bit_exact(hip_device, finish):
    """Chunk batches through the full sort + unique scan and through the
    bucket finish (buckets per chunk from its own size): packed and capacity
    layouts, mixed sizes with chunks below and above 2^14 keys, empty chunks,
    out-of-range records, duplicates and deletes."""
    from pyrit_amd import workloads

    g = _gpu()
    print(f"case: batch finish {finish}, capacity layout", flush=True)
    vox, cnt, offs, cap = _shell_batch()
    ref, rroots, rfirst = O.gb_build_chunks(5, vox, offs, capacity=cap)
    b, roots, first = g.build_chunks(5, vox, offs, capacity=cap, finish=finish)
    assert_same_build(b, ref, 5)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    print(f"case: batch finish {finish}, packed 120 chunks of 64^3", flush=True)
    rng = np.random.default_rng(14)
    chunks = [random_points(int(rng.integers(0, 6000)), 6, int(rng.integers(1 << 30)), dup_frac=0.2,
                            delete_frac=0.05, out_frac=0.01) for _ in range(120)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.uint32)
    vox = np.concatenate(chunks)
    ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs)
    b, roots, first = g.build_chunks(6, vox, offs, finish=finish)
    assert_same_build(b, ref, 6)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)
    print(f"case: batch finish {finish}, mixed sizes 8..64", flush=True)
    w = workloads.uneven_chunks(seed=9, count=40, edges=(8, 16, 32, 64))
    vox, offs64 = w.generate_host()
    offs = offs64.astype(np.uint32)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs, chunk_log2=sizes)
    b, roots, first = g.build_chunks(6, vox, offs, chunk_log2=sizes, finish=finish)
    assert_same_build(b, ref, 6)
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)


def test_device_generator_matches_host(hip_device):
    print("case: terrain generator host == device", flush=True)
    from pyrit_amd import workloads

    for w in (workloads.terrain_chunk64(), workloads.uneven_chunks(seed=8, count=16, edges=(8, 16, 32, 64)),
              workloads.heightfield_grid(log2_size=6, caves=40), workloads.skin_grid(log2_size=7, caves=30)):
        hv, ho = w.generate_host()
        dv, do, total = w.generate_device()
        assert total == len(hv) and np.array_equal(do, ho)
        got = dv.download(O.VOXEL, total)
        assert np.array_equal(got.view(np.uint32), hv.view(np.uint32)), w.name
