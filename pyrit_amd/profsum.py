"""Summarise rocprofv3 CSV output into the short text files kept under
profiles/.

  python -m pyrit_amd.profsum stats <kernel_stats.csv> --out profiles/X.md --cmd "..."
  python -m pyrit_amd.profsum traffic --fetch <fetch_counter_collection.csv>
         --write <write_counter_collection.csv> --voxels N --probe-bytes B
         --out profiles/traffic.json --cmd "..." --cmd "..."

`traffic` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) by what the same
counters read for k_copy_probe, a plain copy kernel that reads and writes
exactly --probe-bytes, and reports HBM bytes per launch per kernel.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re


def short_name(name: str) -> str:
    n = name.split("(")[0].replace("void ", "").replace("pyr::", "").strip()
    return re.sub(r"\s+", "", n)


def stats(path: str, out: str, cmds: list[str], top: int = 25) -> None:
    rows = list(csv.DictReader(open(path)))
    lines = ["# rocprofv3 kernel statistics", "", "Commands:", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        lines.append(f"| {short_name(r['Name']) or '(unnamed)'} | {r['Calls']} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    open(out, "w").write("\n".join(lines) + "\n")


def _counter(path: str, name: str) -> dict:
    agg = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != name:
            continue
        agg.setdefault(short_name(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return agg


def traffic(fetch: str, write: str, voxels: int, probe_bytes: int, out: str, cmds: list[str]) -> None:
    f = _counter(fetch, "FETCH_SIZE")
    w = _counter(write, "WRITE_SIZE")
    probe = "k_copy_probe"
    fk = sum(f[probe]) / len(f[probe])
    wk = sum(w[probe]) / len(w[probe])
    fscale = probe_bytes / (fk * 1024.0)
    wscale = probe_bytes / (wk * 1024.0)
    kernels = []
    for k in f:
        if k == probe or not k:
            continue
        fb = sum(f[k]) / len(f[k]) * 1024.0 * fscale
        wb = sum(w.get(k, [0.0])) / max(1, len(w.get(k, [0.0]))) * 1024.0 * wscale
        kernels.append({"kernel": k, "voxels": voxels, "launches": len(f[k]), "fetch_bytes": fb, "write_bytes": wb,
                        "bytes_per_launch": fb + wb})
    data = {"calibration": {"kernel": probe, "probe_bytes": probe_bytes, "fetch_KiB": fk, "write_KiB": wk,
                            "fetch_scale": fscale, "write_scale": wscale},
            "commands": cmds, "kernels": kernels}
    json.dump(data, open(out, "w"), indent=1)


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--out", required=True)
    s.add_argument("--cmd", action="append", default=[])
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--voxels", type=int, required=True)
    t.add_argument("--probe-bytes", type=int, required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args()
    if a.what == "stats":
        stats(a.csv, a.out, a.cmd)
    else:
        traffic(a.fetch, a.write, a.voxels, a.probe_bytes, a.out, a.cmd)


if __name__ == "__main__":
    main()
