"""Build step for the native libraries (run by ``__graft_entry__.build()``).

Compiles, in tree and for gfx950 only:
  pyrit_amd/libpyrit.so          build.hip, trace.hip, api.cpp, serve.cpp,
                                 world.cpp, devcache.cpp, dag.cpp, tune.cpp,
                                 palette.hip, pathedit.hip, rtprims.hip
                                 (the C ABI of
                                 include/pyrit.h and include/pyrit_hip_build.h)
  pyrit_amd/libpyrit_terrain.so  terrain.hip (seeded workload generator)

Usage: python -m pyrit_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
# Packed FP32 instructions are disabled (-packed-fp32-ops off) and fp
# contraction is off everywhere: the traversal must round exactly like the
# reference's strict-IEEE float code.
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}",
    "-ffp-contract=off", "-Wno-unused-value", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIBS = {
    "libpyrit.so": ["build.hip", "trace.hip", "api.cpp", "serve.cpp", "world.cpp", "devcache.cpp", "dag.cpp", "tune.cpp", "palette.hip", "pathedit.hip", "rtprims.hip"],
    "libpyrit_terrain.so": ["terrain.hip"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpyrit")


def _compile(src: str, obj: str) -> None:
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [hipcc(), *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stderr}")


def build(jobs: int = 4, verbose: bool = True) -> list[str]:
    out = []
    with tempfile.TemporaryDirectory(prefix="pyrit_build_") as tmp:
        work = []
        for lib, srcs in LIBS.items():
            for s in srcs:
                work.append((s, os.path.join(tmp, s + ".o")))
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            list(ex.map(lambda a: _compile(*a), work))
        for lib, srcs in LIBS.items():
            target = os.path.join(HERE, lib)
            objs = [os.path.join(tmp, s + ".o") for s in srcs]
            tmp_target = target + ".tmp"
            cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_target, *objs]
            res = subprocess.run(cmd, capture_output=True, text=True)
            if res.returncode != 0:
                raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
            os.replace(tmp_target, target)
            out.append(target)
            if verbose:
                print(f"built {os.path.relpath(target, ROOT)}", flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    build(args.jobs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
