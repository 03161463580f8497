"""Summaries of rocprofv3 output for profiles/ (run on the GPU host).

    python -m iss_amd.prof stats <kernel_stats.csv>
        kernel-trace summary: calls, total, average and share per kernel
    python -m iss_amd.prof traffic <fetch_dir> <write_dir> [out.json]
        HBM bytes of each kernel's largest dispatch (the full-size launch: the
        autotuner's shorter passes dispatch the same kernels) from two
        `rocprofv3 --pmc` runs (FETCH_SIZE, then WRITE_SIZE, KiB per
        dispatch).  Both counters are scaled so that the largest
        __amd_rocclr_copyBuffer dispatch reads and writes the size
        bench.py --calibration-copy-mib copied.
    python -m iss_amd.prof counters <pmc_dir> [<pmc_dir> ...]
        mean of every counter per kernel
"""
from __future__ import annotations

import collections
import csv
import glob
import json
import os
import sys


def _short(name: str) -> str:
    n = name.split("(")[0].replace("issdev::", "")
    return n.replace("void ", "")


def _counters(d):
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            agg[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return agg


def stats(path):
    rows = list(csv.DictReader(open(path)))
    print("| kernel | calls | total ms | average us | share |")
    print("|---|---|---|---|---|")
    for r in rows:
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
              f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} % |")


def traffic(fetch_dir, write_dir, out=None, copy_mib=256):
    fe, wr = _counters(fetch_dir), _counters(write_dir)
    kib = copy_mib * 1024.0
    cf = max((v for k, d in fe.items() if "copyBuffer" in k for v in d.get("FETCH_SIZE", [])), default=None)
    cw = max((v for k, d in wr.items() if "copyBuffer" in k for v in d.get("WRITE_SIZE", [])), default=None)
    ff = kib / cf if cf else 1.0
    fw = kib / cw if cw else 1.0
    res = {"calibration": {"copy_KiB": kib, "FETCH_SIZE_KiB": cf, "WRITE_SIZE_KiB": cw,
                           "fetch_factor": ff, "write_factor": fw}}
    for k in sorted(set(fe) | set(wr)):
        if "rocclr" in k:
            continue
        f = fe.get(k, {}).get("FETCH_SIZE", [])
        w = wr.get(k, {}).get("WRITE_SIZE", [])
        fb = max(f) * ff * 1024 if f else 0.0
        wb = max(w) * fw * 1024 if w else 0.0
        entry = {"fetch_bytes": fb, "write_bytes": wb, "bytes": fb + wb, "dispatches": len(f)}
        res[k] = entry
        base = k.split("<")[0]
        # a template's base name stands for its most-dispatched instance
        if base != k and (base not in res or res[base]["dispatches"] < len(f)):
            res[base] = entry
    txt = json.dumps(res, indent=1)
    if out:
        open(out, "w").write(txt + "\n")
    print(txt)


def counters(*dirs):
    agg = collections.defaultdict(dict)
    for d in dirs:
        for k, cs in _counters(d).items():
            for c, v in cs.items():
                agg[k][c] = sum(v) / len(v)
    for k in sorted(agg):
        if "rocclr" in k:
            continue
        print(k)
        for c, v in sorted(agg[k].items()):
            print(f"    {c:26s} {v:14.5g}")


if __name__ == "__main__":
    cmd, args = sys.argv[1], sys.argv[2:]
    {"stats": stats, "traffic": traffic, "counters": counters}[cmd](*args)
