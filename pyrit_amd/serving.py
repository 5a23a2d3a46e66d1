"""Python side of the continuous-batching build server (libpyrit.so
pyrhip_server_*), and the Poisson request driver of the serving benchmark:
Poisson arrivals at a mean rate, build requests with the uneven chunk sizes of
the batched workload mixed 80/20 with edit requests that set or remove 1..512
seeded voxels of a resident chunk."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import abi, lib
from .device import download


class Server:
    def __init__(self, budget: int = 64 << 30, workers: int = 2, max_batch_chunks: int = 1024,
                 max_batch_voxels: int = 1 << 28, radix_bits: int = 0, sort_items: int = 0, graphs: bool = True):
        """graphs=False: every batch eagerly (PYRHIP_SERVER_NO_GRAPHS), else
        small batches replay a HIP graph of the batch build"""
        self.L = lib()
        info = abi.PyrhipServerInfo()
        info.memory_budget, info.workers = budget, workers
        info.flags = 0 if graphs else abi.PYRHIP_SERVER_NO_GRAPHS
        info.max_batch_chunks, info.max_batch_voxels = max_batch_chunks, max_batch_voxels
        info.tuning.radix_bits, info.tuning.sort_items = radix_bits, sort_items
        self.h = C.c_void_p()
        abi.check(self.L, self.L.pyrhip_server_create(C.byref(info), C.byref(self.h)), "server create")

    def build(self, chunk: int, log2_size: int, dev_voxels: int, count: int, user: int = 0) -> int:
        q = abi.PyrhipRequest()
        q.chunk, q.kind, q.log2_size, q.voxels, q.count, q.user = chunk, abi.PYRHIP_REQUEST_BUILD, log2_size, \
            dev_voxels, count, user
        t = C.c_uint64()
        abi.check(self.L, self.L.pyrhip_server_submit(self.h, C.byref(q), C.byref(t)), "submit build")
        return t.value

    def edit(self, chunk: int, edits: np.ndarray, log2_size: int = 0, user: int = 0) -> int:
        edits = np.ascontiguousarray(edits, dtype=abi.VOXEL)
        q = abi.PyrhipRequest()
        q.chunk, q.kind, q.log2_size, q.count, q.user = chunk, abi.PYRHIP_REQUEST_EDIT, log2_size, len(edits), user
        q.voxels = edits.ctypes.data if len(edits) else 0
        t = C.c_uint64()
        abi.check(self.L, self.L.pyrhip_server_submit(self.h, C.byref(q), C.byref(t)), "submit edit")
        return t.value

    def poll(self, max_n: int = 4096) -> list:
        buf = (abi.PyrhipCompletion * max_n)()
        n = self.L.pyrhip_server_poll(self.h, buf, max_n)
        return [buf[i] for i in range(n)]

    def drain(self) -> None:
        abi.check(self.L, self.L.pyrhip_server_drain(self.h), "drain")

    def chunk_dag(self, chunk: int):
        """(PyrhipDag, voxel_count, node_words, leaf_count) of a resident chunk"""
        d = abi.PyrhipDag()
        n, w, lc = C.c_uint32(), C.c_uint32(), C.c_uint32()
        abi.check(self.L, self.L.pyrhip_server_chunk(self.h, chunk, C.byref(d), C.byref(n), C.byref(w),
                                                     C.byref(lc)), "chunk")
        return d, n.value, w.value, lc.value

    def chunk_arrays(self, chunk: int):
        """host copies (nodes, leaves, attrs, root, log2_size, voxel_count) of a
        resident chunk: the node/leaf arrays of its batch and its attributes"""
        d, n, w, lc = self.chunk_dag(chunk)
        if not d.nodes:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint64), np.zeros(0, np.uint32), 0, d.log2_size, 0
        nodes = download(d.nodes, np.uint32, w)
        leaves = download(d.leaves, np.uint64, lc)
        attrs = download(d.attrs, np.uint32, n)
        return nodes, leaves, attrs, d.root, d.log2_size, n

    def stats(self) -> dict:
        s = abi.PyrhipServerStats()
        self.L.pyrhip_server_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in abi.PyrhipServerStats._fields_}

    def error(self) -> str:
        """message of the last failed batch or rejected request ("" if none)"""
        buf = C.create_string_buffer(512)
        self.L.pyrhip_server_error(self.h, buf, len(buf))
        return buf.value.decode(errors="replace")

    def close(self) -> None:
        if self.h:
            self.L.pyrhip_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clock() -> float:
    return lib().pyrhip_server_clock()


def random_edits(rng: np.random.Generator, log2_size: int, count: int) -> np.ndarray:
    """`count` seeded voxel edits in a chunk: 2/3 set (attribute 1..8), 1/3 remove."""
    n = 1 << log2_size
    e = np.zeros(count, abi.VOXEL)
    xyz = rng.integers(0, n, size=(count, 3))
    e["x"], e["y"], e["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    keep = rng.integers(0, 3, size=count) != 0
    e["attribute"] = np.where(keep, rng.integers(1, 9, size=count), 0)
    return e


def poisson_schedule(inputs, rate: float, duration: float, seed: int = 5, edit_frac: float = 0.2):
    """Arrival offsets (s) and requests of a Poisson stream, generated before
    the run so that the load generator only submits: ("build", j) rebuilds chunk j
    (uniform over inputs) from inputs[j]; ("edit", j, records) edits a chunk
    an earlier request of the stream built."""
    rng = np.random.default_rng(seed)
    t, out, built = 0.0, [], []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        if built and rng.random() < edit_frac:
            j = built[int(rng.integers(len(built)))]
            out.append((t, "edit", j, random_edits(rng, inputs[j][0], int(rng.integers(1, 513)))))
        else:
            j = int(rng.integers(len(inputs)))
            built.append(j)
            out.append((t, "build", j, None))


def poisson_run(server: Server, inputs, rate: float, duration: float, seed: int = 5, edit_frac: float = 0.2):
    """Drive the server with Poisson arrivals for `duration` seconds.

    inputs: list of (log2_size, device pointer, count) chunk inputs. Requests
    are submitted at their arrival times (open loop: a late driver submits at
    once, it never waits for completions). Build completions carry user 0,
    edit completions user 1. Returns (completions, elapsed s, requests)."""
    sched = poisson_schedule(inputs, rate, duration, seed, edit_frac)
    results = []
    t0 = clock()
    for t, kind, j, e in sched:
        while True:
            now = clock() - t0
            if now >= t:
                break
            if t - now > 0.0005:
                time.sleep(min(t - now - 0.0003, 0.002))
                results.extend(server.poll())
        l2, ptr, cnt = inputs[j]
        if kind == "edit":
            server.edit(j, e, log2_size=l2, user=1)
        else:
            server.build(j, l2, ptr, cnt, user=0)
    server.drain()
    results.extend(server.poll())
    return results, clock() - t0, len(sched)
