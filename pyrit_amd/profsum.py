"""Summarise rocprofv3 output into the short text files kept under profiles/.

  python -m pyrit_amd.profsum stats <kernel_stats.csv | results.db> [--trace kernel_trace.csv] --out profiles/X.md --cmd "..."
  python -m pyrit_amd.profsum traffic --fetch <counter csv | db> --write <counter csv | db>
         --voxels N --probe-bytes B --out profiles/traffic.json --cmd "..." --cmd "..."
  python -m pyrit_amd.profsum counters <counter csv | db> --kernel NAME [--kernel ...] --out X.json --cmd "..."

Inputs are either rocprofv3's CSV files (`--output-format csv`) or its rocpd
SQLite database (`*_results.db`, the default output format), read with the
standard library's sqlite3.

`traffic` scales FETCH_SIZE and WRITE_SIZE (KiB per dispatch) by what the same
counters read for k_copy_probe, a plain copy kernel that reads and writes
exactly --probe-bytes, and reports HBM bytes per launch per kernel, averaged
over the launches that did work (at least 5 % of the kernel's largest
fetch: a radix pass whose digit is the same for every record returns at once).
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sqlite3


def short_name(name: str) -> str:
    n = name.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "").replace("pyr::", "").strip()
    return re.sub(r"\s+", "", n)


def _kernel_rows(path: str) -> list[dict]:
    """[{name, calls, total_ns, avg_ns, min_ns, max_ns, vgpr, lds}] sorted by total time"""
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        q = ("select name, count(*), sum(duration), avg(duration), min(duration), max(duration), max(vgpr_count), "
             "max(lds_size) from kernels group by name order by 3 desc")
        return [{"name": r[0], "calls": r[1], "total_ns": float(r[2]), "avg_ns": float(r[3]), "min_ns": float(r[4]),
                 "max_ns": float(r[5]), "vgpr": r[6], "lds": r[7]} for r in db.execute(q)]
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append({"name": r["Name"], "calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
                     "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                     "vgpr": None, "lds": None})
    return rows


def _trace_resources(path: str) -> dict:
    """kernel short name -> (arch VGPRs, LDS bytes, [durations ns]) from a kernel trace CSV"""
    res = {}
    for r in csv.DictReader(open(path)):
        e = res.setdefault(short_name(r["Kernel_Name"]), (r.get("VGPR_Count"), r.get("LDS_Block_Size"), []))
        e[2].append(float(r["End_Timestamp"]) - float(r["Start_Timestamp"]))
    return res


def _working_avg(durations: list) -> tuple:
    """(count, mean) of the launches that did work: at least 5 % of the longest"""
    top = max(durations)
    work = [d for d in durations if d >= 0.05 * top]
    return len(work), sum(work) / len(work)


def stats(path: str, out: str, cmds: list[str], top: int = 25, trace: str | None = None) -> None:
    rows = _kernel_rows(path)
    res = _trace_resources(trace) if trace else {}
    for r in rows:
        e = res.get(short_name(r["name"]))
        if e:
            if r["vgpr"] is None:
                r["vgpr"], r["lds"] = e[0], e[1]
            r["work"] = _working_avg(e[2])
    total = sum(r["total_ns"] for r in rows) or 1.0
    lines = ["# rocprofv3 kernel statistics", "", "Commands:", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "| kernel | calls | total ms | avg us | min us | max us | % | VGPRs | LDS B |"
              + (" working calls | avg us of those |" if res else ""),
              "|---|---|---|---|---|---|---|---|---|" + ("---|---|" if res else "")]
    for r in rows[:top]:
        line = (f"| {short_name(r['name']) or '(unnamed)'} | {r['calls']} | {r['total_ns'] / 1e6:.3f} | "
                f"{r['avg_ns'] / 1e3:.1f} | {r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} | "
                f"{100.0 * r['total_ns'] / total:.2f} | {r['vgpr'] if r['vgpr'] is not None else ''} | "
                f"{r['lds'] if r['lds'] is not None else ''} |")
        if res:
            wk = r.get("work")
            line += f" {wk[0]} | {wk[1] / 1e3:.1f} |" if wk else " | |"
        lines.append(line)
    if res:
        lines += ["", "Working calls: launches of at least 5 % of the kernel's longest (a radix pass whose digit "
                  "is the same for every record returns at once)."]
    open(out, "w").write("\n".join(lines) + "\n")


def _counter(path: str, name: str) -> dict:
    """kernel short name -> [counter value per dispatch]"""
    agg = collections.OrderedDict()
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        q = "select kernel_name, value from counters_collection where counter_name = ? order by dispatch_id"
        for k, v in db.execute(q, (name,)):
            agg.setdefault(short_name(k), []).append(float(v))
        return agg
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
        # launches that did work, by fetch; the write counts of the same launches
        # (both runs launch the same sequence)
        top_f = max(f[k])
        work = [i for i, v in enumerate(f[k]) if v >= 0.05 * top_f] or list(range(len(f[k])))
        wl = w.get(k, [])
        fb = sum(f[k][i] for i in work) / len(work) * 1024.0 * fscale
        wsel = [wl[i] for i in work if i < len(wl)]
        wb = sum(wsel) / max(1, len(wsel)) * 1024.0 * wscale
        kernels.append({"kernel": k, "voxels": voxels, "launches": len(work), "fetch_bytes": fb, "write_bytes": wb,
                        "bytes_per_launch": fb + wb})
    data = {"calibration": {"kernel": probe, "probe_bytes": probe_bytes, "fetch_KiB": fk, "write_KiB": wk,
                            "fetch_scale": fscale, "write_scale": wscale},
            "commands": cmds, "kernels": kernels}
    json.dump(data, open(out, "w"), indent=1)


def counters(path: str, kernels: list[str], out: str, cmds: list[str]) -> None:
    """mean per launch of every counter of the given kernels, and the L2 hit
    rate TCC_HIT / (TCC_HIT + TCC_MISS) when both were collected"""
    names = set()
    if path.endswith(".db"):
        db = sqlite3.connect(path)
        names = {r[0] for r in db.execute("select distinct counter_name from counters_collection")}
    else:
        names = {r["Counter_Name"] for r in csv.DictReader(open(path))}
    res = {}
    for c in sorted(names):
        per = _counter(path, c)
        for k in kernels:
            if k in per:
                res.setdefault(k, {"launches": len(per[k])})[c] = sum(per[k]) / len(per[k])
    for k, v in res.items():
        hit, miss = v.get("TCC_HIT_sum"), v.get("TCC_MISS_sum")
        if hit is not None and miss is not None and hit + miss > 0:
            v["l2_hit_rate"] = hit / (hit + miss)
    json.dump({"commands": cmds, "kernels": res}, open(out, "w"), indent=1)


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("input")
    s.add_argument("--out", required=True)
    s.add_argument("--cmd", action="append", default=[])
    s.add_argument("--trace", default=None, help="kernel trace CSV for VGPR and LDS columns")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--voxels", type=int, required=True)
    t.add_argument("--probe-bytes", type=int, required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--cmd", action="append", default=[])
    c = sub.add_parser("counters")
    c.add_argument("input")
    c.add_argument("--kernel", action="append", required=True)
    c.add_argument("--out", required=True)
    c.add_argument("--cmd", action="append", default=[])
    a = ap.parse_args()
    if a.what == "stats":
        stats(a.input, a.out, a.cmd, trace=a.trace)
    elif a.what == "counters":
        counters(a.input, a.kernel, a.out, a.cmd)
    else:
        traffic(a.fetch, a.write, a.voxels, a.probe_bytes, a.out, a.cmd)


if __name__ == "__main__":
    main()
