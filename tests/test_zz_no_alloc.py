"""pyrhip_build_static: the build without device allocation (scratch from a
static device array) gives the same arrays as pyrhip_build, call after call
on one device, and rejects builds or outputs that do not fit."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from scenes import random_points

pytestmark = pytest.mark.gpu


def _static(L, abi, log2, vox_ptr, n, caps, old=None):
    from pyrit_amd.device import Buffer

    nodes, leaves = Buffer(max(caps[0], 1) * 4), Buffer(max(caps[1], 1) * 8)
    keys, attrs = Buffer(max(caps[2], 1) * 8), Buffer(max(caps[2], 1) * 4)
    dst = abi.PyrhipBuildOutput(nodes.ptr, caps[0], leaves.ptr, caps[1], keys.ptr, attrs.ptr, caps[2])
    out = abi.PyrhipBuilt()
    ok, ov, on = (old.raw.keys, old.raw.attrs, old.voxel_count) if old is not None else (0, 0, 0)
    code = L.pyrhip_build_static(log2, ok, ov, on, vox_ptr, n, None, None, C.byref(dst), C.byref(out))
    return code, out, (nodes, leaves, keys, attrs)


def _same(out, bufs, ref):
    nodes, leaves, keys, attrs = bufs
    assert out.root == ref.root and out.voxel_count == ref.voxel_count
    assert out.node_words == ref.node_words and out.leaf_count == ref.leaf_count
    assert out.nodes == nodes.ptr and out.leaves == leaves.ptr and out.keys == keys.ptr and out.attrs == attrs.ptr
    assert np.array_equal(nodes.download(np.uint32, out.node_words), ref.nodes())
    assert np.array_equal(leaves.download(np.uint64, out.leaf_count), ref.leaves())
    assert np.array_equal(keys.download(np.uint64, out.voxel_count), ref.keys())
    assert np.array_equal(attrs.download(np.uint32, out.voxel_count), ref.attrs())


def test_static_build_matches_pyrhip_build(hip_device):
    from pyrit_amd import abi, gpu, lib, workloads
    from pyrit_amd.device import to_device

    L = lib()
    assert L.pyrhip_build_static_limit() == 64 << 20
    print("case: static build, terrain chunk 64^3", flush=True)
    vox, _ = workloads.terrain_chunk64().generate_host()
    vb = to_device(vox)
    ref = gpu.build(6, vb, count=len(vox))
    caps = (ref.node_words + 64, ref.leaf_count + 64, len(vox))
    code, out, bufs = _static(L, abi, 6, vb.ptr, len(vox), caps)
    assert code == 0, L.pyr_error_message()
    _same(out, bufs, ref)
    assert O.canon(bufs[0].download(np.uint32, out.node_words), bufs[1].download(np.uint64, out.leaf_count),
                   out.root, 6).same(O.build_points(6, vox))

    print("case: static build, smaller second call (random points 32^3)", flush=True)
    pts = random_points(3000, 5, 21, dup_frac=0.2, delete_frac=0.05, out_frac=0.02)
    pb = to_device(pts)
    ref2 = gpu.build(5, pb, count=len(pts))
    code, out2, bufs2 = _static(L, abi, 5, pb.ptr, len(pts), (ref2.node_words, ref2.leaf_count, len(pts)))
    assert code == 0, L.pyr_error_message()
    _same(out2, bufs2, ref2)

    print("case: static build, edit of an earlier build", flush=True)
    ed = random_points(500, 6, 22, dup_frac=0.0, delete_frac=0.3)
    eb = to_device(ed)
    old = gpu.build(6, vb, count=len(vox))
    ref3 = gpu.build(6, eb, count=len(ed), old=old)
    code, out3, bufs3 = _static(L, abi, 6, eb.ptr, len(ed), (ref3.node_words, ref3.leaf_count, len(vox) + len(ed)),
                                old=old)
    assert code == 0, L.pyr_error_message()
    _same(out3, bufs3, ref3)


def test_static_build_rejects_what_does_not_fit(hip_device):
    from pyrit_amd import abi, lib, workloads
    from pyrit_amd.device import to_device

    L = lib()
    print("case: static build, scratch too small", flush=True)
    big = random_points(4_000_000, 10, 23, dup_frac=0.0)
    bb = to_device(big)
    code, out, _ = _static(L, abi, 10, bb.ptr, len(big), (1 << 20, 1 << 20, len(big)))
    assert code == abi.PYR_ERROR_CAPACITY
    assert b"static scratch" in L.pyr_error_message()
    print("case: static build, output buffers too small", flush=True)
    vox, _ = workloads.terrain_chunk64().generate_host()
    vb = to_device(vox)
    code, out, _ = _static(L, abi, 6, vb.ptr, len(vox), (8, 8, len(vox)))
    assert code == abi.PYR_ERROR_CAPACITY
    assert b"exceeds the buffers" in L.pyr_error_message()
