"""Palette packing of chunk attributes (src/device/palette.zig, layout of
src/context.zig:1311-1370, read back by attributeAt, src/device/dag.zig:64-68).

CPU: the oracle's packing is pinned by decoding every rank back (brute force),
the 16-value boundary and the word counts 16 + ceil(n / 8) or n. GPU:
pyrhip_palette_pack is bit-exact against the oracle (layout, counts, palette
order, index words), including chunks whose tiles each see at most 16 values
while the chunk holds more."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

OVER = 0xFFFFFFFF


def decode(counts, offs, packed, first):
    """attributeAt over every rank of every chunk"""
    out = []
    for j in range(len(counts)):
        n = int(first[j + 1] - first[j])
        w = packed[int(offs[j]):int(offs[j + 1])]
        if counts[j] == OVER:
            out.append(w[:n])
            continue
        pal, idx = w[:16], w[16:]
        r = np.arange(n)
        out.append(pal[(idx[r >> 3] >> ((r & 7) * 4).astype(np.uint32)) & 15])
    return np.concatenate(out) if out else np.zeros(0, np.uint32)


def batch(seed: int):
    """chunks with 1, 2, 16 and 17 distinct values, 0xFFFFFFFF, long runs, 16
    values spread so that each 4096-attribute tile sees only a few, 17 the same
    way, a single voxel and lengths around multiples of 8"""
    rng = np.random.default_rng(seed)
    chunks = [
        np.full(5, 7, np.uint32),
        rng.choice(np.array([3, 9], np.uint32), 1001),
        rng.choice(np.arange(1, 17, dtype=np.uint32) * 1000003, 20000),
        rng.choice(np.arange(1, 18, dtype=np.uint32) * 77, 20000),
        np.array([1, 2, OVER, 2, 1], np.uint32),
        np.repeat(np.array([5, 6, 5], np.uint32), 30000),
        np.repeat(np.arange(1, 17, dtype=np.uint32), 4096),
        np.repeat(np.arange(1, 18, dtype=np.uint32), 4096),
        np.array([42], np.uint32),
        rng.integers(1, 1 << 32, 24, dtype=np.uint64).astype(np.uint32),
        rng.choice(np.array([1, 2, 3], np.uint32), 4097),
    ]
    chunks += [rng.choice(rng.integers(1, 1 << 31, int(rng.integers(1, 20)), dtype=np.uint64).astype(np.uint32),
                          int(rng.integers(1, 9000))) for _ in range(40)]
    attrs = np.concatenate(chunks)
    attrs[attrs == 0] = 1  # attributes of built voxels are never 0
    first = np.concatenate([[0], np.cumsum([len(c) for c in chunks])]).astype(np.uint32)
    return attrs, first


def test_oracle_palette_decodes_every_rank():
    attrs, first = batch(1)
    counts, offs, packed = O.palette_pack(attrs, first)
    assert np.array_equal(decode(counts, offs, packed, first), attrs)
    for j in range(len(counts)):
        a = attrs[first[j]:first[j + 1]]
        d = np.unique(a)
        pal = len(d) <= 16 and OVER not in d
        assert counts[j] == (len(d) if pal else OVER), j
        n = len(a)
        assert offs[j + 1] - offs[j] == (16 + (n + 7) // 8 if pal else n)
        if pal:
            p = packed[int(offs[j]):int(offs[j]) + 16]
            assert np.array_equal(p[:len(d)], d) and np.all(p[len(d):] == d[0])
    assert counts[2] == 16 and counts[3] == OVER and counts[4] == OVER and counts[6] == 16 and counts[7] == OVER


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_palette_pack_bit_exact(hip_device, seed):
    print(f"case: palette pack seed {seed}", flush=True)
    from pyrit_amd import abi, lib
    from pyrit_amd.device import Buffer, to_device

    attrs, first = batch(seed)
    counts, offs, packed = O.palette_pack(attrs, first)
    L = lib()
    k = len(first) - 1
    da = to_device(attrs)
    f = np.ascontiguousarray(first, np.uint32)
    got_offs = np.zeros(k + 1, np.uint64)
    got_counts = np.zeros(k, np.uint32)
    P32, P64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
    args = (f.ctypes.data_as(P32), k)
    abi.check(L, L.pyrhip_palette_pack(da.ptr, *args, 0, 0, got_offs.ctypes.data_as(P64),
                                       got_counts.ctypes.data_as(P32), None), "layout")
    assert np.array_equal(got_counts, counts) and np.array_equal(got_offs, offs)
    out = Buffer(int(offs[-1]) * 4)
    # too small an output is refused
    assert L.pyrhip_palette_pack(da.ptr, *args, out.ptr, int(offs[-1]) - 1, got_offs.ctypes.data_as(P64),
                                 got_counts.ctypes.data_as(P32), None) == -1
    abi.check(L, L.pyrhip_palette_pack(da.ptr, *args, out.ptr, int(offs[-1]), got_offs.ctypes.data_as(P64),
                                       got_counts.ctypes.data_as(P32), None), "pack")
    got = out.download(np.uint32, int(offs[-1]))
    assert np.array_equal(got, packed)
    assert np.array_equal(decode(got_counts, got_offs, got, first), attrs)
