"""The committed config-1 fixture (tests/golden/config1.npz, written by
tests/golden/make_config1.py): the 64^3 terrain chunk's records, its DAG in
both layouts and the hits of 65,536 primary rays. On the CPU: the generator
still makes these records and the oracle still gives these arrays and hits.
With a GPU: the HIP build and the camera kernel give them too."""
import os

import numpy as np
import pytest

import oracle_py as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1.npz")


def _load():
    f = np.load(FIXTURE)  # plain arrays, no pickles
    return {k: f[k] for k in f.files}


def _voxels(fx):
    return np.ascontiguousarray(fx["voxels"]).view(O.VOXEL).reshape(-1)


def test_generator_reproduces_the_fixture_records():
    from golden.make_config1 import records

    fx = _load()
    assert np.array_equal(records().view(np.uint32).reshape(-1, 4), fx["voxels"])
    assert len(fx["voxels"]) == 141558


def test_oracle_reproduces_the_fixture_dag_and_hits():
    from golden.make_config1 import expected

    fx = _load()
    want = expected(_voxels(fx))
    for k, v in want.items():
        assert np.array_equal(np.asarray(v), fx[k]), k
    hits = fx["hits"]
    assert hits.shape == (65536, 4)
    assert (hits[:, 1] == 0).sum() > 16000  # a quarter of the view is the chunk


@pytest.mark.gpu
def test_hip_build_and_camera_kernel_match_the_fixture(hip_device):
    print("case: config-1 fixture on the GPU", flush=True)
    from golden.make_config1 import EYE, FOV, SIZE, TARGET, UP
    from pyrit_amd import abi, gpu, lib
    import ctypes as C

    fx = _load()
    b = gpu.build(6, _voxels(fx))
    assert b.root == int(fx["gb_root"])
    assert np.array_equal(b.nodes(), fx["gb_nodes"])
    assert np.array_equal(b.leaves(), fx["gb_leaves"])
    assert np.array_equal(b.keys(), fx["gb_keys"])
    assert np.array_equal(b.attrs(), fx["gb_attrs"])
    cam = abi.PyrCamera()
    L = lib()
    L.pyr_camera_look_at(C.byref(cam), (C.c_float * 3)(*EYE), (C.c_float * 3)(*TARGET), (C.c_float * 3)(*UP))
    L.pyr_camera_perspective(C.byref(cam), FOV, SIZE, SIZE, 0.0)
    hits = gpu.trace_camera(b.dag(), cam)
    assert np.array_equal(hits.view(np.uint32).reshape(-1, 4), fx["hits"])
