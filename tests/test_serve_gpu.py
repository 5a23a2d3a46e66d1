"""The continuous-batching build server: every chunk's result is the same when
its requests run alone and when they run mixed into batches with others, and
equal to the reference CPU builder on the chunk's final voxel set; admission
against the memory budget rejects what can never fit and lets the rest run."""
import numpy as np
import pytest

import oracle_py as O
from scenes import final_points, random_points

pytestmark = pytest.mark.gpu


def _requests(seed=3, chunks=12, steps=40):
    """A seeded request script: builds of chunks with edges 8..64 and edits of
    built chunks; returns (inputs, script)."""
    rng = np.random.default_rng(seed)
    inputs = []
    for c in range(chunks):
        l2 = int(rng.integers(3, 7))
        pts = random_points(int(rng.integers(0, 3000)), l2, int(rng.integers(1 << 30)), dup_frac=0.1,
                            delete_frac=0.05)
        inputs.append((l2, pts))
    script, built = [], set()
    for s in range(steps):
        if built and rng.random() < 0.4:
            c = int(rng.choice(sorted(built)))
            l2 = inputs[c][0]
            script.append(("edit", c, random_points(int(rng.integers(1, 200)), l2, int(rng.integers(1 << 30)),
                                                     dup_frac=0.1, delete_frac=0.4)))
        else:
            c = int(rng.integers(0, chunks))
            script.append(("build", c, None))
            built.add(c)
    return inputs, script


def _final_sets(inputs, script):
    """Expected voxel set per chunk after the script (oracle semantics:
    builds replace, edits append records that win, attribute 0 deletes)."""
    state = {}
    for kind, c, e in script:
        l2 = inputs[c][0]
        if kind == "build":
            state[c] = final_points(inputs[c][1], l2)
        else:
            state[c] = final_points(np.concatenate([state.get(c, np.zeros(0, O.VOXEL)), e]), l2)
    return state


def _run(server, dev_inputs, inputs, script, alone):
    for kind, c, e in script:
        if kind == "build":
            buf, n = dev_inputs[c]
            server.build(c, inputs[c][0], buf.ptr if buf else 0, n)
        else:
            server.edit(c, e, log2_size=inputs[c][0])
        if alone:
            server.drain()
    server.drain()
    return server.poll()


def _chunk_canon(server, c):
    nodes, leaves, attrs, root, l2, n = server.chunk_arrays(c)
    if n == 0:
        return None, attrs
    return O.canon(nodes, leaves, root, l2), attrs


@pytest.mark.parametrize("graphs", [True, False])
def test_same_alone_and_in_the_mix(hip_device, graphs):
    """graphs: small batches replay HIP graphs of the batch build (the
    default), or every batch runs eagerly"""
    print(f"case: server alone vs mixed, graphs {graphs}", flush=True)
    from pyrit_amd.device import to_device
    from pyrit_amd.serving import Server

    inputs, script = _requests()
    dev = [(to_device(p) if len(p) else None, len(p)) for _, p in inputs]
    expect = _final_sets(inputs, script)

    alone = Server(workers=1, graphs=graphs)
    ra = _run(alone, dev, inputs, script, alone=True)
    mixed = Server(workers=2, graphs=graphs)
    rm = _run(mixed, dev, inputs, script, alone=False)
    assert all(c.result == 0 for c in ra) and all(c.result == 0 for c in rm)
    assert len(ra) == len(rm) == len(script)
    assert max(c.batch_chunks for c in rm) > 1, "requests were never batched together"
    for c, pts in expect.items():
        ref = O.build_points(inputs[c][0], pts)
        ga, aa = _chunk_canon(alone, c)
        gm, am = _chunk_canon(mixed, c)
        if ref.voxel_count == 0:
            assert ga is None and gm is None
            continue
        assert ga.same(ref) and gm.same(ref), c
        assert np.array_equal(aa, ref.attrs) and np.array_equal(am, ref.attrs), c
    st = mixed.stats()
    assert st["completed"] == len(script) and st["rejected"] == 0
    assert alone.error() == "" and mixed.error() == ""
    alone.close()
    mixed.close()


def test_resident_chunk_traces_like_the_oracle(hip_device):
    print("case: server chunk traversal", flush=True)
    from pyrit_amd import gpu, workloads
    from pyrit_amd.device import to_device
    from pyrit_amd.serving import Server

    w = workloads.uneven_chunks(seed=12, count=6, edges=(16, 32, 64), probs=(0.4, 0.3, 0.3))
    vox, offs = w.generate_host()
    sv = Server(workers=2)
    bufs = []
    for c in range(w.count):
        seg = vox[offs[c]:offs[c + 1]]
        b = to_device(seg)
        bufs.append(b)
        sv.build(c, w.chunks[c].log2_size, b.ptr, len(seg))
    sv.drain()
    for c in range(w.count):
        seg = vox[offs[c]:offs[c + 1]]
        l2 = w.chunks[c].log2_size
        d, n, _, _ = sv.chunk_dag(c)
        cpu = O.build_points(l2, seg)
        n3 = 1 << l2
        cam = O.camera((-n3 * 0.5, n3 * 1.5, -n3 * 0.5), (n3 / 2, n3 / 3, n3 / 2), (0, 1, 0), 1.0, 48, 32)
        rays = O.camera_rays(cam)
        hits = gpu.trace_dag(d, rays)
        ref = O.trace_dag_rays(cpu, rays)
        assert np.array_equal(hits.view(np.uint32), ref.view(np.uint32)), c
    sv.close()


def test_admission_rejects_what_cannot_fit(hip_device):
    print("case: server admission", flush=True)
    from pyrit_amd.device import to_device
    from pyrit_amd.serving import Server

    pts = random_points(20000, 6, 5)
    small = random_points(500, 5, 6)
    b1, b2 = to_device(pts), to_device(small)
    # budget below what a 20000-record request needs, above a 500-record one
    sv = Server(budget=(1 << 20) + 500 * 72 + 200000, workers=2)
    t_big = sv.build(1, 6, b1.ptr, len(pts))
    t_small = sv.build(2, 5, b2.ptr, len(small))
    sv.drain()
    res = {c.ticket: c for c in sv.poll()}
    assert res[t_big].result == -3  # PYR_ERROR_OUT_OF_MEMORY
    assert res[t_small].result == 0 and res[t_small].voxel_count == len(final_points(small, 5))
    # the rejection's message stays readable on the server (the worker thread failed it)
    assert sv.error() == f"request of {len(pts)} records on chunk 1 exceeds the batch limit or budget"
    sv.close()
