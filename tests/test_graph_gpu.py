"""Chunk-batch builds captured into a HIP graph (pyrhip_graph_*): every
replay must be bit-exact to the eager batch build of the same input and to
the oracle, whatever ran on the GPU between replays (eager builds, other
batches, synchronisations), for smaller batches than the graph was sized
for, the capacity layout and chunks of mixed sizes."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import random_points

pytestmark = pytest.mark.gpu


def _pad(offs, count):
    out = np.full(count + 1, offs[-1], np.uint32)
    out[:len(offs)] = offs
    return out


def _check(out, roots, first, ref, rroots, rfirst):
    assert out.voxel_count == ref["voxel_count"]
    assert out.root == ref["root"]
    assert np.array_equal(out.nodes(), ref["nodes"])
    assert np.array_equal(out.leaves(), ref["leaves"])
    assert np.array_equal(out.attrs(), ref["attrs"])
    assert np.array_equal(out.keys(), ref["keys"])
    assert np.array_equal(roots, rroots) and np.array_equal(first, rfirst)


def _random_batch(seed, chunks, log2=5, most=800):
    rng = np.random.default_rng(seed)
    parts = [random_points(int(rng.integers(0, most)), log2, int(rng.integers(1 << 30)), dup_frac=0.1,
                           delete_frac=0.05) for _ in range(chunks)]
    offs = np.concatenate([[0], np.cumsum([len(c) for c in parts])]).astype(np.uint32)
    return np.concatenate(parts), offs


def test_graph_replays_packed(hip_device):
    from pyrit_amd import gpu
    from pyrit_amd.device import synchronize

    count = 64
    batches = [_random_batch(1, 64), _random_batch(2, 40), _random_batch(3, 64, most=300),
               (np.zeros(0, O.VOXEL), np.zeros(65, np.uint32)), _random_batch(1, 64)]
    total_max = max(int(o[-1]) for _, o in batches)
    g = gpu.BuildGraph(5, count, total_max)
    try:
        for i, (vox, offs) in enumerate(batches):
            print(f"case: packed graph replay {i}, {len(offs) - 1} chunks, {int(offs[-1])} voxels", flush=True)
            po = _pad(offs, count)
            ref, rroots, rfirst = O.gb_build_chunks(5, vox, po)
            out, roots, first = g.launch(vox, offs)
            _check(out, roots, first, ref, rroots, rfirst)
            # eager work between replays: a batch build of another input
            ev, eo = _random_batch(100 + i, 30)
            b, er, ef = gpu.build_chunks(5, ev, eo)
            eref, err, erf = O.gb_build_chunks(5, ev, eo)
            assert np.array_equal(b.nodes(), eref["nodes"]) and np.array_equal(er, err)
            b.free()
            synchronize()
            # the same input again through the eager path: identical arrays
            b, er, ef = gpu.build_chunks(5, vox, po)
            _check(b, er, ef, ref, rroots, rfirst)
            b.free()
    finally:
        g.destroy()


def test_graph_replays_capacity_layout(hip_device):
    from pyrit_amd import gpu
    from test_oracle import _shell_batch

    vox, cnt, offs, cap = _shell_batch()
    k = len(offs) - 1
    g = gpu.BuildGraph(5, k, int(offs[-1]), capacity=cap)
    try:
        for i in range(3):
            print(f"case: capacity graph replay {i}", flush=True)
            # replay i keeps the first voxels of each chunk (fewer every time)
            keep = np.diff(offs.astype(np.int64)) >> i
            o = np.concatenate([[0], np.cumsum(keep)]).astype(np.uint32)
            ref, rroots, rfirst = O.gb_build_chunks(5, vox, o, capacity=cap)
            out, roots, first = g.launch(vox, o)
            _check(out, roots, first, ref, rroots, rfirst)
    finally:
        g.destroy()


@pytest.mark.parametrize("finish", [1, 2])
def test_graph_replays_mixed_sizes(hip_device, finish):
    from pyrit_amd import gpu, workloads

    ws = [workloads.uneven_chunks(seed=s, count=24, edges=(8, 16, 32, 64)) for s in (5, 6, 7)]
    data = []
    for w in ws:
        vox, offs64 = w.generate_host()
        data.append((vox, offs64.astype(np.uint32), np.array([c.log2_size for c in w.chunks], np.uint8)))
    g = gpu.BuildGraph(6, 24, max(int(o[-1]) for _, o, _ in data), mixed_sizes=True, finish=finish)
    try:
        for rep in range(2):
            for i, (vox, offs, sizes) in enumerate(data):
                print(f"case: mixed-size graph replay {rep}/{i}, finish {finish}", flush=True)
                ref, rroots, rfirst = O.gb_build_chunks(6, vox, offs, chunk_log2=sizes)
                out, roots, first = g.launch(vox, offs, chunk_log2=sizes)
                _check(out, roots, first, ref, rroots, rfirst)
    finally:
        g.destroy()


def test_graph_input_in_place_and_argument_checks(hip_device):
    from pyrit_amd import abi, gpu, lib
    from pyrit_amd.device import hip

    vox, offs = _random_batch(9, 16)
    g = gpu.BuildGraph(5, 16, int(offs[-1]))
    try:
        print("case: graph with voxels already in its input buffer", flush=True)
        assert g.voxels_ptr() != 0
        H = hip()
        assert H.hipMemcpy(C.c_void_p(g.voxels_ptr()), vox.ctypes.data_as(C.c_void_p), C.c_size_t(vox.nbytes),
                           1) == 0
        ref, rroots, rfirst = O.gb_build_chunks(5, vox, offs)
        out, roots, first = g.launch(None, offs)
        _check(out, roots, first, ref, rroots, rfirst)
        print("case: graph argument checks", flush=True)
        big = offs.copy()
        big[-1] += 1
        with pytest.raises(RuntimeError):
            g.launch(vox, big)  # more voxels than the graph holds
        bad = offs.copy()
        bad[1] = offs[2] + 1  # offsets decrease after chunk 1
        with pytest.raises(RuntimeError):
            g.launch(vox, bad)
        with pytest.raises(RuntimeError):
            g.launch(vox, offs, chunk_log2=np.full(16, 5, np.uint8))  # not a mixed-size graph
        # still usable after rejected launches
        out, roots, first = g.launch(vox, offs)
        _check(out, roots, first, ref, rroots, rfirst)
    finally:
        g.destroy()
    L = lib()
    h = C.c_void_p()
    assert L.pyrhip_graph_create(5, 0, 0, 10, 0, None, C.byref(h)) == abi.PYR_ERROR_INVALID_ARGUMENT
    assert L.pyrhip_graph_create(5, 4, 10, 41, 0, None, C.byref(h)) == abi.PYR_ERROR_INVALID_ARGUMENT
    assert not h.value
