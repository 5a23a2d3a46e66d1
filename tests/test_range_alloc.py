"""The product's pool allocator (pyrit_amd/csrc/range_alloc.h, used by every
pyr_geometry_* call) on the reference's own sequence (range_alloc.zig:61-78)
and on random sequences against the oracle's restatement
(range_alloc.zig:10-59), compiled for the host with g++ (no GPU)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("ra") / "range_alloc_driver")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/pyrit_amd/csrc",
                    os.path.join(ROOT, "tests", "range_alloc_driver.cpp"), "-o", exe], check=True)

    def run(ops):
        out = subprocess.run([exe], input="\n".join(ops) + "\n", capture_output=True, text=True, check=True).stdout
        offs, states = [], []
        for line in out.splitlines():
            w = line.split()
            if w[0] == "off":
                offs.append(int(w[1]))
            else:
                states.append((int(w[1]), int(w[3])))
        return offs, states
    return run


def test_reference_sequence(driver):
    # range_alloc.zig:61-78: three blocks of 30 in 100, a fourth of 20 does not
    # fit; releasing b then a leaves two free ranges; 60 fits at 0; releasing
    # everything coalesces into one range of 100 with nothing used
    offs, states = driver(["init 100", "alloc 30", "alloc 30", "alloc 30", "alloc 20", "release 30 30",
                           "release 0 30", "alloc 60", "release 0 60", "release 60 30"])
    assert offs == [0, 30, 60, -1, 0]
    assert states[6] == (30, 2)       # after releasing b and a
    assert states[-1] == (0, 1)


def test_random_sequences_match_the_oracle(driver):
    L = O.lib()
    rng = np.random.default_rng(21)
    for trial in range(20):
        cap = int(rng.integers(50, 5000))
        ops = [f"init {cap}"]
        r = O.OrRangeAlloc()
        L.or_range_init(C.byref(r), cap)
        live, want_offs, want_states = [], [], []
        for _ in range(400):
            if live and rng.random() < 0.45:
                off, size = live.pop(int(rng.integers(0, len(live))))
                ops.append(f"release {off} {size}")
                L.or_range_release(C.byref(r), off, size)
            else:
                size = int(rng.integers(1, cap // 6 + 2))
                ops.append(f"alloc {size}")
                off = L.or_range_alloc(C.byref(r), size)
                want_offs.append(off)
                if off >= 0:
                    live.append((off, size))
            want_states.append((r.used, r.n))
        offs, states = driver(ops)
        assert offs == want_offs, f"trial {trial}"
        assert states[1:] == want_states, f"trial {trial}"
        # free ranges plus live blocks tile the capacity exactly
        assert r.used == sum(s for _, s in live)
        L.or_range_free(C.byref(r))
