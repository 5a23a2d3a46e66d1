"""Writes tests/golden/config1.npz: the 64^3 terrain chunk of the first
benchmark configuration (seeded 2-octave heightfield, mean 32, amplitude 16,
categorical materials), its DAG from the CPU builder and from the serial
GPU-layout build (both from the oracle), and the hits of the 256 x 256
primary rays of a 60-degree pinhole camera at (-32, 96, -32) looking at the
chunk centre, traced by the oracle.

    python tests/golden/make_config1.py

The fixture is data: inputs (the voxel records) and expected outputs. The
tests regenerate the records and check them against the fixture, so a
change of the generator shows up as a fixture mismatch."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_py as O  # noqa: E402

OUT = os.path.join(HERE, "config1.npz")
EYE, TARGET, UP = (-32.0, 96.0, -32.0), (32.0, 32.0, 32.0), (0.0, 1.0, 0.0)
FOV, SIZE = float(np.pi / 3), 256


def records() -> np.ndarray:
    from pyrit_amd import workloads

    vox, _ = workloads.terrain_chunk64(seed=1).generate_host()
    return np.ascontiguousarray(vox).view(O.VOXEL)


def expected(vox: np.ndarray) -> dict:
    cpu = O.build_points(6, vox)
    gb = O.gb_build(6, vox)
    cam = O.camera(EYE, TARGET, UP, FOV, SIZE, SIZE)
    rays = O.camera_rays(cam)
    hits = O.trace_rays(O.single_scene(cpu), rays, threads=8)
    return {
        "cpu_root": np.uint32(cpu.root), "cpu_nodes": cpu.nodes, "cpu_leaves": cpu.leaves, "cpu_attrs": cpu.attrs,
        "gb_root": np.uint32(gb["root"]), "gb_nodes": gb["nodes"], "gb_leaves": gb["leaves"], "gb_keys": gb["keys"],
        "gb_attrs": gb["attrs"], "hits": hits.view(np.uint32).reshape(-1, 4),
    }


def main() -> None:
    vox = records()
    np.savez_compressed(OUT, voxels=vox.view(np.uint32).reshape(-1, 4), **expected(vox))
    print(f"wrote {OUT}: {len(vox)} records, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
