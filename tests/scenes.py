"""Test scenes restated from the reference's tests (tests/common.zig:14-35,
dag_builder.zig:444-520) and small seeded point sets."""
from __future__ import annotations

import numpy as np

from oracle_py import VOXEL


from pyrit_amd.workloads import attr_of, scene_grid  # noqa: E402,F401  (sceneVoxel, tests/common.zig:14-35)


def grid_to_points(grid: np.ndarray) -> np.ndarray:
    z, y, x = np.nonzero(grid)
    pts = np.zeros(len(x), VOXEL)
    pts["x"], pts["y"], pts["z"], pts["attribute"] = x, y, z, grid[z, y, x]
    return pts


def sphere_noise_grid(n: int = 32, seed: int = 1234) -> np.ndarray:
    """dag_builder.zig:444-466 style: sphere of radius 12 plus 1/50 noise voxels,
    attributes 1..1000 (own seeded RNG)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    d = (x - 15.5) ** 2 + (y - 15.5) ** 2 + (z - 15.5) ** 2
    inside = (d < 144.0) | (rng.integers(0, 50, size=x.shape) == 0)
    return np.where(inside, rng.integers(1, 1001, size=x.shape), 0).astype(np.uint32)


def random_points(n: int, log2_size: int, seed: int, dup_frac=0.1, delete_frac=0.0, out_frac=0.0) -> np.ndarray:
    """Shuffled points with duplicates (later entries win), deletions
    (attribute 0) and out-of-range entries, as cpu_test.zig:758-807 uses."""
    rng = np.random.default_rng(seed)
    size = 1 << log2_size
    pts = np.zeros(n, VOXEL)
    pts["x"] = rng.integers(0, size, n)
    pts["y"] = rng.integers(0, size, n)
    pts["z"] = rng.integers(0, size, n)
    pts["attribute"] = rng.integers(1, 1 << 30, n)
    k = int(n * dup_frac)
    if k:
        dup = pts[rng.integers(0, n, k)].copy()
        dup["attribute"] = rng.integers(1, 1 << 30, k)
        pts = np.concatenate([pts, dup])
    k = int(n * delete_frac)
    if k:
        dele = pts[rng.integers(0, len(pts), k)].copy()
        dele["attribute"] = 0
        pts = np.concatenate([pts, dele])
    k = int(n * out_frac)
    if k:
        out = pts[rng.integers(0, len(pts), k)].copy()
        out["x"] = np.where(rng.integers(0, 2, k) == 0, -1 - rng.integers(0, 5, k), size + rng.integers(0, 5, k))
        pts = np.concatenate([pts, out])
    rng.shuffle(pts)
    return pts


def final_points(pts: np.ndarray, log2_size: int) -> np.ndarray:
    """The voxel set a build of `pts` must contain: in range, last entry per
    coordinate wins, attribute 0 deletes."""
    size = 1 << log2_size
    ok = (pts["x"] >= 0) & (pts["y"] >= 0) & (pts["z"] >= 0) & (pts["x"] < size) & (pts["y"] < size) & (
        pts["z"] < size)
    p = pts[ok]
    key = (p["x"].astype(np.int64) * size + p["y"]) * size + p["z"]
    # last occurrence of each key
    rev = key[::-1]
    _, idx = np.unique(rev, return_index=True)
    last = len(key) - 1 - idx
    q = p[np.sort(last)]
    return q[q["attribute"] != 0]
