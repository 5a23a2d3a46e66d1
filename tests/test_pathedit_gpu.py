"""Path edits (the capability of src/device/pathedit.zig): voxels set,
inserted and removed in a finished DAG by writing new copies of the touched
bricks and nodes into spare capacity. The
edited DAG must describe exactly the edited voxel set: canonical form equal
to the CPU builder's DAG of that set, attributes equal in DFS rank order."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import grid_to_points, scene_grid

pytestmark = pytest.mark.gpu


def _apply(pts: np.ndarray, edits: np.ndarray) -> np.ndarray:
    """the voxel set after the edits, in order (attribute 0 removes)"""
    d = {(int(p["x"]), int(p["y"]), int(p["z"])): int(p["attribute"]) for p in pts}
    for x, y, z, a in edits.tolist():
        if a:
            d[(x, y, z)] = a
        else:
            d.pop((x, y, z), None)
    out = np.zeros(len(d), O.VOXEL)
    for i, ((x, y, z), a) in enumerate(sorted(d.items())):
        out[i] = (x, y, z, a)
    return out


def _edit(log2, pts, edits, node_extra=None, leaf_extra=None):
    """GPU build of pts, then pyrhip_path_edit with the edits; returns
    (flags, root, count, nodes, leaves, attrs) downloaded"""
    from pyrit_amd import abi, gpu, lib
    from pyrit_amd.device import Buffer, to_device

    L = lib()
    b = gpu.build(log2, pts)
    nodes, leaves, attrs = b.nodes(), b.leaves(), b.attrs()
    ne = len(edits)
    levels = log2 - 2
    ncap = len(nodes) + (node_extra if node_extra is not None else (ne + 1) * levels * 10)
    lcap = len(leaves) + (leaf_extra if leaf_extra is not None else ne)
    nb = to_device(np.concatenate([nodes, np.zeros(ncap - len(nodes), np.uint32)]))
    lb = to_device(np.concatenate([leaves, np.zeros(max(lcap - len(leaves), 1), np.uint64)]))
    ab = to_device(attrs if len(attrs) else np.zeros(1, np.uint32))
    out = Buffer((len(attrs) + ne + 1) * 4)
    d = abi.PyrhipDag()
    d.nodes, d.leaves, d.attrs = nb.ptr, lb.ptr, ab.ptr
    d.root, d.log2_size, d.default_attribute = b.root, log2, 1
    e = np.ascontiguousarray(np.asarray(edits, np.uint32).reshape(-1, 4))
    root, count, flags = C.c_uint32(), C.c_uint32(), C.c_uint32()
    abi.check(L, L.pyrhip_path_edit(C.byref(d), len(attrs), len(nodes), ncap, len(leaves), lcap,
                                    e.ctypes.data_as(C.POINTER(C.c_uint32)), ne, out.ptr, None, C.byref(root),
                                    C.byref(count), C.byref(flags)), "path edit")
    return (flags.value, root.value, count.value, nb.download(np.uint32, ncap), lb.download(np.uint64, lcap),
            out.download(np.uint32, count.value))


def _check(log2, pts, edits):
    flags, root, count, nodes, leaves, attrs = _edit(log2, pts, edits)
    assert flags == 0
    final = _apply(pts, edits)
    ref = O.build_points(log2, final)
    assert count == ref.voxel_count == len(final)
    assert O.canon(nodes, leaves, root, log2).same(ref)
    assert np.array_equal(attrs, ref.attrs)


def test_path_edits_scene(hip_device):
    print("case: path edits on the 32^3 scene", flush=True)
    pts = grid_to_points(scene_grid(32))
    rng = np.random.default_rng(5)
    # sets of existing voxels, removals, inserts next to existing voxels and
    # in empty space (missing nodes), repeated positions
    ex = pts[rng.choice(len(pts), 60, replace=False)]
    edits = [(int(p["x"]), int(p["y"]), int(p["z"]), int(rng.integers(1, 1000))) for p in ex[:20]]
    edits += [(int(p["x"]), int(p["y"]), int(p["z"]), 0) for p in ex[20:40]]
    edits += [(int(x), int(y), int(z), int(rng.integers(1, 1000))) for x, y, z in rng.integers(0, 32, (40, 3))]
    edits += [(31, 31, 31, 7), (31, 31, 31, 0), (31, 31, 31, 9), (0, 31, 0, 3), (1, 31, 0, 4), (0, 31, 0, 0)]
    order = rng.permutation(len(edits))
    _check(5, pts, np.array([edits[i] for i in order], np.uint32))


@pytest.mark.parametrize("log2,seed", [(4, 1), (6, 2), (8, 3)])
def test_path_edits_random(hip_device, log2, seed):
    print(f"case: path edits log2={log2} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    n = 1 << log2
    pts = _apply(np.zeros(0, O.VOXEL), np.concatenate(
        [rng.integers(0, n, (3000, 3)), rng.integers(1, 50, (3000, 1))], axis=1).astype(np.uint32))
    k = 300
    pos = np.concatenate([rng.integers(0, n, (k, 3)),
                          np.stack([pts["x"], pts["y"], pts["z"]], 1)[rng.choice(len(pts), k)]]).astype(np.uint32)
    att = np.where(rng.random(2 * k) < 0.35, 0, rng.integers(1, 50, 2 * k)).astype(np.uint32)
    _check(log2, pts, np.concatenate([pos, att[:, None]], axis=1))


@pytest.mark.parametrize("log2,points,edits,seed", [(8, 40000, 4000, 4), (12, 5000, 1500, 5), (5, 2000, 3000, 6)])
def test_path_edits_many_bricks(hip_device, log2, points, edits, seed):
    """Thousands of edits over many bricks (many rounds of 64 bricks per wave),
    deep trees whose edits create whole new paths (4096^3), and a dense 32^3
    grid where most voxels are touched several times."""
    print(f"case: path edits log2={log2} edits={edits} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    n = 1 << log2
    pts = _apply(np.zeros(0, O.VOXEL), np.concatenate(
        [rng.integers(0, n, (points, 3)), rng.integers(1, 1 << 20, (points, 1))], axis=1).astype(np.uint32))
    half = edits // 2
    pos = np.concatenate([rng.integers(0, n, (half, 3)),
                          np.stack([pts["x"], pts["y"], pts["z"]], 1)[rng.choice(len(pts), edits - half)]])
    att = np.where(rng.random(edits) < 0.4, 0, rng.integers(1, 1 << 20, edits))
    e = np.concatenate([pos, att[:, None]], axis=1).astype(np.uint32)
    _check(log2, pts, e[rng.permutation(edits)])


def test_path_edit_empty_and_overflow(hip_device):
    print("case: path edit to empty, capacity overflow", flush=True)
    pts = np.array([(1, 2, 3, 5), (9, 9, 9, 6)], O.VOXEL)
    flags, *_ = _edit(4, pts, np.array([(1, 2, 3, 0), (9, 9, 9, 0)], np.uint32))
    assert flags == 2
    flags, *_ = _edit(4, pts, np.array([(15, 0, 15, 4)], np.uint32), node_extra=3)
    assert flags == 4
    flags, *_ = _edit(4, pts, np.array([(15, 0, 15, 4)], np.uint32), leaf_extra=0)
    assert flags == 4
    # nothing changes: sets only
    _check(4, pts, np.array([(1, 2, 3, 8), (9, 9, 9, 6), (1, 2, 3, 2)], np.uint32))
