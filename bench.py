"""Benchmark: sampled hadrons/s and yield-pass GB/s on one MI355X.

Workload (the configuration the metric is quoted on): a viscous 3+1D
freeze-out surface of 200,000 cells (tau in [0.6, 15] fm, eta_s in [-5, 5],
T in [0.135, 0.160] GeV, mu_B in [0, 0.3] GeV, |u_perp| <= 3, random
traceless pi with |pi|/(e+P) <= 0.3, bulk Pi/(e+P) in [-0.2, 0], OSU
14-moment bulk + shear delta f), the 321-species UrQMD hadron list, ~1e4
hadrons per event.  One step = one yield pass (FP32 series, FP64 species
CDF) + one sampling pass of 1000 oversampled events, with the surface
already resident in HBM.  `value` = hadrons sampled per second over the
timed steps (yields included).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed); 78.6 unpacked


def yield_bytes(n_cells: int, n_species: int, with_bulk: bool) -> int:
    """Algorithmic HBM bytes of one yield pass: per cell T, muB, muS, muQ,
    u.dsigma, Pi (6 x 4 B) and, with bulk delta f, 3 FP64 coefficients;
    per (cell, species) one FP64 species-CDF entry; per cell its FP64 total."""
    per_cell_in = 6 * 4 + (3 * 8 if with_bulk else 0)
    return n_cells * (per_cell_in + 8) + n_cells * n_species * 8


def sample_flops(n_trials: int, n_inner: int, n_accepted: int) -> float:
    """Algorithmic FP32 ops of the sampling pass (FMA = 2, other ops and
    transcendentals = 1): ~115 per |p| proposal, ~85 per outer trial, ~250
    per accepted hadron (cell/species selection, boost, record)."""
    return 115.0 * n_inner + 85.0 * n_trials + 250.0 * n_accepted


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(cells, sp, seconds_target=12.0):
    """The CPU oracle (a 1-thread restatement of the reference algorithm) on a
    bounded sample of the same workload: the first 2000 cells, 14-moment
    bulk + shear delta f, sampled for as many events as fit ~10 s."""
    from tests import oracle_binding as ob
    from iss_amd import tables  # noqa: F401

    sub = cells[:2000]
    orc = ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=11)
    orc.set_cells_lrf(sub)
    t0 = time.perf_counter()
    _, h = orc.sample(1, 10)
    dt = time.perf_counter() - t0
    rate0 = max(h.size / max(dt, 1e-9), 1.0)
    nev = int(max(10, min(100000, seconds_target * rate0 / max(h.size / 10, 1))))
    t0 = time.perf_counter()
    _, h = orc.sample(2, nev)
    dt = time.perf_counter() - t0
    return {"value": h.size / dt, "unit": "hadrons/s", "cores": 1, "kind": "port",
            "sample": f"oracle (yields + sampling, 1 thread) on the first 2000 cells of the "
                      f"workload, {nev} events, {h.size} hadrons in {dt:.1f} s; host {cpu_model()}, "
                      f"nproc {os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=200000)
    ap.add_argument("--events", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the sampler has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import iss_amd
    from iss_amd import synth, tables

    sp = tables.species_arrays(1)
    cells, _ = synth.viscous_lrf(args.cells, sp, seed=2, muB_max=0.3, target_hadrons=1.0e4)
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surface_lrf(cells)
    # normalise the surface to 1e4 hadrons per event including delta f
    tot = ctx.compute_yields(precision=64)
    cells[:, 4:8] *= np.float32(1.0e4 / tot.sum())
    ctx.set_surface_lrf(cells)
    nc, ns = ctx.num_cells(), ctx.n_species

    def step(i):
        ctx.compute_yields(precision=args.precision)
        n = ctx.sample(args.events, seed=12345, event0=i * args.events, precision=args.precision)
        return n, ctx.stats()

    for i in range(args.warmup):
        step(i)
    tot_h = 0
    ky, ks, trials, inner, stot = [], [], 0, 0, []
    t0 = time.perf_counter()
    for i in range(args.steps):
        n, st = step(args.warmup + i)
        tot_h += n
        ky.append(st.ms_yield_kernel)
        ks.append(st.ms_sample_kernel)
        stot.append(st.ms_sample)
        trials += st.n_trials
        inner += st.n_inner
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_step = 1e3 * elapsed / args.steps
    ky_ms, ks_ms = float(np.mean(ky)), float(np.mean(ks))
    ybytes = yield_bytes(nc, ns, True)
    y_gbs = ybytes / (ky_ms * 1e-3) / 1e9
    s_tflops = sample_flops(trials, inner, tot_h) / args.steps / (ks_ms * 1e-3) / 1e12
    roof_y = {"bound": "hbm", "achieved": y_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": y_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "k_yield",
              "ms": ky_ms, "bytes": ybytes}
    roof_s = {"bound": "valu", "achieved": s_tflops, "peak": FP32_VALU_PEAK_TFLOPS,
              "unit": "TFLOP/s", "frac": s_tflops / FP32_VALU_PEAK_TFLOPS, "traffic": None,
              "kernel": "k_sample", "ms": ks_ms}
    dominant = roof_s if ks_ms >= ky_ms else roof_y
    out = {
        "metric": "sampled hadrons/s (learned proposal vs plain GPU) at fixed chi2; yield-pass GB/s",
        "value": tot_h / elapsed,
        "unit": "hadrons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": "viscous 3+1D surface, plain rejection proposal", "cells": nc,
                   "species": ns, "events_per_step": args.events,
                   "hadrons_per_event": tot_h / (args.steps * args.events),
                   "deltaf": "shear + OSU 14-moment bulk", "proposal": "plain"},
        "roofline": dominant,
        "roofline_yield": roof_y,
        "roofline_sample": roof_s,
        "sample_hadrons_per_s": tot_h / (1e-3 * float(np.sum(stot))),
        "yield_GBs": y_gbs,
        "trials_per_hadron": trials / max(tot_h, 1),
        "inner_per_hadron": inner / max(tot_h, 1),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cells, sp)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
