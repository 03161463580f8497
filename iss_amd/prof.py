"""Summaries of rocprofv3 output for profiles/ (run on the GPU host).

    python -m iss_amd.prof stats <kernel_stats.csv>
        kernel-trace summary: calls, total, average and share per kernel
    python -m iss_amd.prof window <kernel_trace.csv> <anchor> <steps> [<stop>]
        the timed steps of a run: the dispatches before the first dispatch
        of the kernel `stop` (the first kernel of the work after the steps;
        without it, the whole trace), from the start of the steps-th last
        dispatch of the kernel `anchor` (the first kernel of a step) among
        them; per kernel calls, ms per step and share
    python -m iss_amd.prof traffic <fetch_dir> <write_dir> [out.json]
        HBM bytes of each kernel's largest dispatch (the full-size launch: the
        autotuner's shorter passes dispatch the same kernels) from two
        `rocprofv3 --pmc` runs (FETCH_SIZE, then WRITE_SIZE, KiB per
        dispatch).  Both counters are scaled so that the largest
        __amd_rocclr_copyBuffer dispatch reads and writes the size
        bench.py --calibration-copy-mib copied.
    python -m iss_amd.prof counters <pmc_dir> [<pmc_dir> ...]
        mean of every counter per kernel
    python -m iss_amd.prof util <pmc_dir> [<pmc_dir> ...] [out.json]
        hardware utilisation of each kernel's full-size dispatches from
        `rocprofv3 --pmc` runs (rocpd databases: counters and durations per
        dispatch; each counter averaged over the runs that hold it) with
        SQ_WAVES, SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F32, SQ_INSTS_MFMA,
        SQ_INSTS_VALU_MFMA_MOPS_BF16, SQ_INSTS_VALU_FLOPS_FP32 and
        GRBM_GUI_ACTIVE: clock, VALU issue-slot share, FP32 FLOP/s and BF16
        MFMA FLOP/s against the MI355X peaks
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


def window(path, anchor, steps, stop=None):
    steps = int(steps)
    rows = [(_short(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
            for r in csv.DictReader(open(path))]
    rows.sort(key=lambda r: r[1])
    if stop:
        rows = rows[:next((i for i, r in enumerate(rows) if r[0] == stop), len(rows))]
    starts = [t0 for k, t0, _ in rows if k == anchor]
    if len(starts) < steps:
        raise SystemExit(f"{anchor}: {len(starts)} dispatches, fewer than {steps} steps")
    t_begin = starts[-steps]
    win = [r for r in rows if r[1] >= t_begin]
    agg = collections.defaultdict(lambda: [0, 0])
    for k, t0, t1 in win:
        agg[k][0] += 1
        agg[k][1] += t1 - t0
    busy = sum(v[1] for v in agg.values())
    span = max(t1 for _, _, t1 in win) - t_begin
    print(f"{len(win)} dispatches from the {steps}-th last `{anchor}`: kernels busy {busy / 1e6 / steps:.3f} ms "
          f"per step, span {span / 1e6 / steps:.3f} ms per step (with the host's gaps between launches)")
    print()
    print("| kernel | calls per step | ms per step | average us | share of busy |")
    print("|---|---|---|---|---|")
    for k, (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{k}` | {n / steps:g} | {ns / 1e6 / steps:.3f} | {ns / n / 1e3:.1f} | {100.0 * ns / busy:.1f} % |")


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


# MI355X: 8 XCDs of 32 CUs x 4 SIMDs; a wave64 VALU instruction occupies its
# SIMD for 4 cycles; GRBM_GUI_ACTIVE counts each XCD's active cycles;
# SQ_INSTS_VALU_FLOPS_FP32 counts FLOP per wave instruction (FMA = 2), so x 64
# lanes (an upper bound: inactive lanes count); peaks FP32 vector 157.3
# TFLOP/s, dense BF16 MFMA 2.5 PFLOP/s
N_XCD = 8
N_SIMD = 1024
WAVE = 64
FP32_PEAK = 157.3e12
BF16_PEAK = 2.5e15
MOP_FLOP = 512  # SQ_INSTS_VALU_MFMA_MOPS_*: units of 512 FLOP


def util(*args):
    import sqlite3
    out = args[-1] if args and args[-1].endswith(".json") else None
    dirs = args[:-1] if out else args
    # kernel -> (run, dispatch) -> counters
    per = collections.defaultdict(lambda: collections.defaultdict(dict))
    for run, d in enumerate(dirs):
        for db in glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True):
            con = sqlite3.connect(db)
            for disp, name, cname, val, dur in con.execute(
                    "select dispatch_id, kernel_name, counter_name, value, duration from counters_collection"):
                e = per[_short(name)][(run, db, disp)]
                e[cname] = e.get(cname, 0.0) + float(val)
                e["_dur_ns"] = float(dur)
    res = {"N_SIMD": N_SIMD, "FP32_peak_TFLOPs": FP32_PEAK / 1e12, "BF16_peak_TFLOPs": BF16_PEAK / 1e12}
    print("| kernel | dispatches | us | clock GHz | VALU/wave | VALU issue | trans share | FP32 TFLOP/s | "
          "FP32 frac | BF16 MFMA TFLOP/s | MFMA frac |")
    print("|---|---|---|---|---|---|---|---|---|---|---|")
    for k in sorted(per):
        if "rocclr" in k:
            continue
        ds = list(per[k].values())
        top = max(e["_dur_ns"] for e in ds)
        full = [e for e in ds if e["_dur_ns"] >= 0.5 * top]  # the full-size launches
        names = set().union(*full)
        m = {}
        for c in names:  # each counter over the launches (of any run) that hold it
            v = [e[c] for e in full if c in e]
            m[c] = sum(v) / len(v)
        # rates use each counter's own runs' durations: the full-size launches
        # of one kernel take the same time in every run
        dur = m["_dur_ns"] * 1e-9
        if dur <= 0:
            continue
        clk = m.get("GRBM_GUI_ACTIVE", 0.0) / (N_XCD * dur) if m.get("GRBM_GUI_ACTIVE") else 2.4e9
        valu = m.get("SQ_INSTS_VALU", 0.0)
        waves = max(m.get("SQ_WAVES", 1.0), 1.0)
        issue = valu * 4.0 / (N_SIMD * clk * dur)
        f32 = m.get("SQ_INSTS_VALU_FLOPS_FP32", 0.0) * WAVE / dur
        bf16 = m.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0) * MOP_FLOP / dur
        trans = m.get("SQ_INSTS_VALU_TRANS_F32", 0.0) / valu if valu else 0.0
        res[k] = {"dispatches": len(full), "us": dur * 1e6, "clock_GHz": clk / 1e9, "valu_per_wave": valu / waves,
                  "valu_issue_frac": issue, "trans_share": trans, "fp32_TFLOPs": f32 / 1e12,
                  "fp32_frac": f32 / FP32_PEAK, "bf16_mfma_TFLOPs": bf16 / 1e12, "mfma_frac": bf16 / BF16_PEAK}
        r = res[k]
        print(f"| `{k}` | {r['dispatches']} | {r['us']:.1f} | {r['clock_GHz']:.2f} | {r['valu_per_wave']:.0f} | "
              f"{r['valu_issue_frac']:.3f} | {r['trans_share']:.3f} | {r['fp32_TFLOPs']:.2f} | {r['fp32_frac']:.4f} | "
              f"{r['bf16_mfma_TFLOPs']:.2f} | {r['mfma_frac']:.4f} |")
    if out:
        open(out, "w").write(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    cmd, args = sys.argv[1], sys.argv[2:]
    {"stats": stats, "window": window, "traffic": traffic, "counters": counters, "util": util}[cmd](*args)
