"""Box primitives of a DAG (src/rt_prims.zig:127-137 extract, emitted by the
reference's GPU build at rt_log2, src/device/gbuild.zig:574-605): the GPU
operator against the oracle's restatement, bit for bit (prims and float
boxes), on DAGs of the CPU builder and of the GPU build, including the
benchmark grid at full size and the chunks of a batch."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import final_points, grid_to_points, random_points, scene_grid

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _device_dag(nodes, leaves, root, log2):
    from pyrit_amd import abi
    from pyrit_amd.device import to_device

    nb = to_device(np.ascontiguousarray(nodes, np.uint32))
    lb = to_device(leaves if len(leaves) else np.zeros(1, np.uint64))
    d = abi.PyrhipDag()
    d.nodes, d.leaves, d.attrs = nb.ptr, lb.ptr, 0
    d.root, d.log2_size, d.default_attribute = root, log2, 1
    d._keep = (nb, lb)
    return d


def _same(gpu_res, cpu_res):
    k, prims, aabbs = gpu_res
    kc, pc, ac = cpu_res
    assert k == kc
    assert len(prims) == len(pc)
    assert np.array_equal(prims.view(np.uint8), pc.view(np.uint8))
    assert np.array_equal(aabbs.view(np.uint32), ac.view(np.uint32))


def _check_host_dag(dag: O.Dag, requests):
    from pyrit_amd import gpu

    d = _device_dag(dag.nodes, dag.leaves, dag.root, dag.log2_size)
    for req in requests:
        _same(gpu.rt_prims(d, len(dag.nodes), len(dag.leaves), req), O.rt_prims(dag, req))


def test_known_answer(hip_device):
    # rt_prims.zig:139-160
    _say("case: rt prims known answer")
    from pyrit_amd import gpu

    pts = np.array([(1, 2, 3, 7), (30, 31, 17, 8), (31, 31, 17, 9)], O.VOXEL)
    dag = O.build_points(5, pts)
    d = _device_dag(dag.nodes, dag.leaves, dag.root, 5)
    k, prims, aabbs = gpu.rt_prims(d, len(dag.nodes), len(dag.leaves), 4)
    assert k == 4 and len(prims) == 2
    assert prims[0]["attr_base"] == 0 and prims[0]["cell"] == 0
    assert prims[1]["attr_base"] == 1 and prims[1]["cell"] == 1 | 1 << 21 | 1 << 42
    assert abs(aabbs[0][0] - 1) < 0.01 and abs(aabbs[0][5] - 4) < 0.01
    assert abs(aabbs[1][0] - 30) < 0.01 and abs(aabbs[1][3] - 32) < 0.01


@pytest.mark.parametrize("log2,seed", [(5, 1), (7, 2), (9, 3)])
def test_random_points_every_level(hip_device, log2, seed):
    _say(f"case: rt prims random points log2 {log2} seed {seed}")
    pts = final_points(random_points(20000, log2, seed, delete_frac=0.1), log2)
    _check_host_dag(O.build_points(log2, pts), [0, 3, 4, 5, log2 - 1, log2, log2 + 4])


def test_scene_and_shared_nodes(hip_device):
    _say("case: rt prims dense scene and periodic pattern")
    _check_host_dag(O.build_dense(6, scene_grid(64)), [3, 4, 5, 6])
    n = 64
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    per = ((x % 8) + (y % 8) < (z % 8) + 2).astype(np.uint32)
    _check_host_dag(O.build_dense(6, per), [3, 4, 6])


def test_empty_and_single_voxel(hip_device):
    _say("case: rt prims empty DAG and single voxel")
    from pyrit_amd import gpu

    dag = O.build_points(5, np.zeros(0, O.VOXEL))
    d = _device_dag(dag.nodes, dag.leaves, dag.root, 5)
    for req in (3, 5):
        k, prims, _ = gpu.rt_prims(d, len(dag.nodes), max(len(dag.leaves), 1), req)
        assert len(prims) == 0
    _check_host_dag(O.build_points(6, np.array([(63, 0, 40, 5)], O.VOXEL)), [3, 5, 6])


def test_capacity_and_bad_references(hip_device):
    _say("case: rt prims capacity and reference checks")
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer

    L = lib()
    pts = final_points(random_points(5000, 6, 9), 6)
    dag = O.build_points(6, pts)
    d = _device_dag(dag.nodes, dag.leaves, dag.root, 6)
    want = O.rt_prims(dag, 3)[1]
    count, k = C.c_uint32(), C.c_uint32()
    pb, ab = Buffer(16 * len(want)), Buffer(24 * len(want))
    rc = L.pyrhip_rt_prims(C.byref(d), len(dag.nodes), len(dag.leaves), 3, pb.ptr, ab.ptr, len(want) - 1, None,
                           C.byref(count), C.byref(k))
    assert rc == abi.PYR_ERROR_CAPACITY and count.value == len(want)
    # word and leaf counts that leave references outside are reported, not read
    rc = L.pyrhip_rt_prims(C.byref(d), len(dag.nodes), len(dag.leaves) // 2, 3, pb.ptr, ab.ptr, len(want), None,
                           C.byref(count), C.byref(k))
    assert rc == abi.PYR_ERROR_INVALID_ARGUMENT
    rc = L.pyrhip_rt_prims(C.byref(d), dag.root + 1, len(dag.leaves), 3, pb.ptr, ab.ptr, len(want), None,
                           C.byref(count), C.byref(k))
    assert rc == abi.PYR_ERROR_INVALID_ARGUMENT  # the root's own words run past the count
    rc = L.pyrhip_rt_prims(C.byref(d), len(dag.nodes), len(dag.leaves), 3, pb.ptr, 0, len(want), None,
                           C.byref(count), C.byref(k))
    assert rc == abi.PYR_ERROR_INVALID_ARGUMENT


def test_gpu_built_benchmark_grid(hip_device):
    """The 1024^3 heightfield grid built on the GPU: primitives at the
    reference's default edge 2^7, at 2^3 (the most) and the whole grid."""
    _say("case: rt prims of the GPU-built 1024^3 heightfield")
    from pyrit_amd import gpu, workloads

    w = workloads.heightfield_grid(log2_size=10)
    vox, _, n = w.generate_device()
    b = gpu.build(10, vox, count=n)
    vox.free()
    dag = O.Dag(10, b.root, b.nodes(), b.leaves(), None, b.voxel_count)
    for req in (7, 3, 10):
        g = gpu.rt_prims(b.dag(), b.node_words, b.leaf_count, req)
        c = O.rt_prims(dag, req)
        _say(f"  k={req}: {len(c[1])} primitives")
        _same(g, c)
    # the primitives partition the voxels: attribute ranks step by the count words
    k, prims, _ = gpu.rt_prims(b.dag(), b.node_words, b.leaf_count, 5)
    counts = dag.nodes[prims["node"].astype(np.int64) + 1].astype(np.int64)
    assert np.array_equal(np.cumsum(counts)[:-1], prims["attr_base"][1:].astype(np.int64))
    assert counts.sum() == b.voxel_count


def test_chunk_batch_roots(hip_device):
    """64 chunks of edges 32..256 in one batch share the node array: each
    chunk root's primitives equal the oracle's for that root."""
    _say("case: rt prims per chunk of an uneven batch")
    from pyrit_amd import gpu, workloads

    w = workloads.uneven_chunks(seed=3, count=64)
    sizes = np.array([c.log2_size for c in w.chunks], np.uint8)
    vox, offs64 = w.generate_host()
    b, roots, first = gpu.build_chunks(8, vox, offs64.astype(np.uint32), chunk_log2=sizes)
    nodes, leaves = b.nodes(), b.leaves()
    total = 0
    for c, r in enumerate(roots):
        if r == 0xFFFFFFFF:
            continue
        log2 = int(sizes[c])
        dag = O.Dag(log2, int(r), nodes, leaves, None, 0)
        d = b.dag()
        d.root, d.log2_size = int(r), log2
        res = gpu.rt_prims(d, b.node_words, b.leaf_count, log2 - 2)
        _same(res, O.rt_prims(dag, log2 - 2))
        total += len(res[1])
    assert total > 0
