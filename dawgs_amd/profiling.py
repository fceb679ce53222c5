"""Summaries of rocprofv3 output for the gather-scatter kernels.

    python -m dawgs_amd.profiling stats  <dir-with-*_kernel_stats.csv>
    python -m dawgs_amd.profiling tail <trace dir> --kernel K --last N [--groups G]
    python -m dawgs_amd.profiling curve <trace dir> --kernel K [--every 10]
    python -m dawgs_amd.profiling counters <pmc dir> --counter C --kernel K [--groups G] [--last N]
    python -m dawgs_amd.profiling traffic <FETCH_SIZE run dir> <WRITE_SIZE run dir> \
        --kernel k_gs_k1 --copy-kernel copy --copy-bytes B [--json out.json]

`traffic` turns two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs
into HBM bytes per launch of a kernel.  Each counter is scaled by what it
reads for a plain copy kernel of known size in the same run (on MI355X,
FETCH_SIZE counts half of a copy's read bytes; WRITE_SIZE counts them all).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys


def _rows(d: str, suffix: str):
    files = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    out = []
    for f in files:
        with open(f) as fh:
            out.extend(csv.DictReader(fh))
    return out


def kernel_stats(d: str):
    rows = _rows(d, "kernel_stats.csv")
    return [{"name": r["Name"], "calls": int(r["Calls"]), "avg_ns": float(r["AverageNs"]),
             "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]), "pct": float(r["Percentage"])}
            for r in rows]


def trace_tail_mean(d: str, kernel: str, last: int) -> dict:
    """Mean duration of the last `last` dispatches of a kernel in a
    --kernel-trace run (the timed region of a benchmark that launches its
    timed steps last)."""
    rows = [r for r in _rows(d, "kernel_trace.csv") if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows]
    tail = durs[-last:] if last else durs
    return {"dispatches": len(durs), "tail": len(tail), "tail_mean_ns": sum(tail) / max(len(tail), 1),
            "all_mean_ns": sum(durs) / max(len(durs), 1)}


def trace_curve(d: str, kernel: str, every: int = 10) -> list:
    """A kernel's dispatches in time order, cut into runs of `every`: per run
    the first dispatch's index and start time (ms after the first dispatch),
    the mean duration (us) and the largest idle gap before a dispatch of the
    run (us).  Shows how the kernel time moves after setup and idle gaps."""
    rows = [r for r in _rows(d, "kernel_trace.csv") if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    se = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in rows]
    out = []
    for i in range(0, len(se), every):
        run = se[i:i + every]
        gaps = [(s - se[j - 1][1]) / 1e3 for j, (s, _) in enumerate(run, start=i) if j > 0]
        out.append({"first": i, "t_ms": round((run[0][0] - se[0][0]) / 1e6, 3),
                    "mean_us": round(sum(e - s for s, e in run) / len(run) / 1e3, 1),
                    "max_gap_us": round(max(gaps), 1) if gaps else 0.0})
    return out


def _groups(vals: list, groups: int, last: int) -> list:
    """`vals` (in dispatch order) cut into `groups` equal runs, the mean of
    each run's last `last` values (all when 0)."""
    n = len(vals) // groups
    if n * groups != len(vals) or n == 0:
        raise ValueError(f"{len(vals)} dispatches do not split into {groups} equal groups")
    out = []
    for g in range(groups):
        run = vals[g * n:(g + 1) * n]
        run = run[-last:] if last else run
        out.append(sum(run) / len(run))
    return out


def trace_group_means(d: str, kernel: str, groups: int, last: int) -> list:
    """Mean duration (ns) per group of a kernel's dispatches in time order,
    e.g. one group per degree of a sweep that launches the same count each."""
    rows = [r for r in _rows(d, "kernel_trace.csv") if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    return _groups([int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows], groups, last)


def counter_group_means(d: str, counter: str, kernel: str, groups: int, last: int) -> list:
    """Mean counter value (KiB) per group of a kernel's dispatches in
    dispatch order."""
    v = [(int(r["Dispatch_Id"]), float(r["Counter_Value"])) for r in _rows(d, "counter_collection.csv")
         if r["Counter_Name"] == counter and kernel in r["Kernel_Name"]]
    v.sort()
    return _groups([x for _, x in v], groups, last)


def counter_means(d: str, counter: str, last: int = 0):
    """Mean counter value (KiB as rocprofv3 reports it) per kernel name, over
    each kernel's last `last` dispatches (all when 0)."""
    acc = {}
    for r in _rows(d, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        acc.setdefault(r["Kernel_Name"], []).append((int(r["Dispatch_Id"]), float(r["Counter_Value"])))
    out = {}
    for k, v in acc.items():
        v.sort()
        vals = [x for _, x in (v[-last:] if last else v)]
        out[k] = sum(vals) / len(vals)
    return out


def _pick(means: dict, pattern: str, exclude: str = "rocclr") -> float:
    hits = [v for k, v in means.items() if pattern in k and exclude not in k]
    if len(hits) != 1:
        raise KeyError(f"{len(hits)} kernels match {pattern!r}: {[k for k in means if pattern in k]}")
    return hits[0]


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_kernel: str, copy_bytes: int, last: int = 0) -> dict:
    """HBM bytes per launch of `kernel`, over its last `last` dispatches (all
    when 0: plan creation may launch a kernel too, to time block orders)."""
    f, w = counter_means(fetch_dir, "FETCH_SIZE"), counter_means(write_dir, "WRITE_SIZE")
    fk, wk = counter_means(fetch_dir, "FETCH_SIZE", last), counter_means(write_dir, "WRITE_SIZE", last)
    half = copy_bytes / 2  # a copy reads half and writes half of the bytes it moves
    f_scale = half / (_pick(f, copy_kernel) * 1024)
    w_scale = half / (_pick(w, copy_kernel) * 1024)
    rd = _pick(fk, kernel) * 1024 * f_scale
    wr = _pick(wk, kernel) * 1024 * w_scale
    return {"kernel": kernel, "fetch_bytes": round(rd), "write_bytes": round(wr),
            "hbm_bytes_per_launch": round(rd + wr), "fetch_scale": round(f_scale, 4),
            "write_scale": round(w_scale, 4), "copy_kernel": copy_kernel, "copy_bytes": copy_bytes}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    tl = sub.add_parser("tail")
    tl.add_argument("dir")
    tl.add_argument("--kernel", required=True)
    tl.add_argument("--last", type=int, required=True)
    tl.add_argument("--groups", type=int, default=1, help="split the dispatches into G equal runs")
    cv = sub.add_parser("curve")
    cv.add_argument("dir")
    cv.add_argument("--kernel", required=True)
    cv.add_argument("--every", type=int, default=10)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--counter", required=True)
    c.add_argument("--kernel", required=True)
    c.add_argument("--groups", type=int, default=1)
    c.add_argument("--last", type=int, default=0)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--last", type=int, default=0, help="only each kernel's last N dispatches")
    t.add_argument("--json")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        for k in sorted(kernel_stats(a.dir), key=lambda r: -r["pct"]):
            print(f"{k['avg_ns'] / 1e3:10.1f} us avg  {k['calls']:5d} calls  {k['pct']:6.2f}%  {k['name'][:110]}")
    elif a.cmd == "tail" and a.groups > 1:
        print(json.dumps({"group_tail_mean_ns": trace_group_means(a.dir, a.kernel, a.groups, a.last)}))
    elif a.cmd == "tail":
        print(json.dumps(trace_tail_mean(a.dir, a.kernel, a.last), indent=1))
    elif a.cmd == "curve":
        for r in trace_curve(a.dir, a.kernel, a.every):
            print(f"{r['first']:5d}  t={r['t_ms']:9.3f} ms  mean {r['mean_us']:7.1f} us  max gap {r['max_gap_us']:9.1f} us")
    elif a.cmd == "counters":
        print(json.dumps({"counter": a.counter,
                          "group_mean_KiB": counter_group_means(a.dir, a.counter, a.kernel, a.groups, a.last)}))
    else:
        r = traffic(a.fetch_dir, a.write_dir, a.kernel, a.copy_kernel, a.copy_bytes, a.last)
        print(json.dumps(r, indent=1))
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(r, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
