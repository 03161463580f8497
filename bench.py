"""Benchmark: sampled hadrons/s, learned proposal against the plain GPU
sampler, and yield-pass GB/s, on one MI355X.

Workload (the configuration the metric is quoted on): a viscous 3+1D
freeze-out surface of 200,000 cells (tau in [0.6, 15] fm, eta_s in [-5, 5],
T in [0.135, 0.160] GeV, mu_B in [0, 0.3] GeV, |u_perp| <= 3, random
traceless pi with |pi|/(e+P) <= 0.3, bulk Pi/(e+P) in [-0.2, 0], OSU
14-moment bulk + shear delta f), the 321-species UrQMD hadron list,
normalised to 1e4 hadrons per event.  One step = one yield pass (FP32
series and yield rows; FP64 species totals) + one sampling pass of 1000
oversampled events (~1e7 hadrons; each cell's FP64 species prefix is built
at pick time), surface resident in HBM.  The plain steps run first; the
network is then trained on the 1e7 hadrons the last plain pass accepted
(outside the timed region), and the learned steps run on the same surface.
`value` = hadrons/s of the learned-proposal steps (yields included); the
plain steps' rate and the ratio are reported beside it.  Both samplers use
the physical Fermi-Dirac |p| proposal (fermion_cdf2_quirk = 0) so that every
hadron goes through the proposal being measured.

Beside it (outside the timed region): the resonance feed-down of a step's
hadrons to stable particles, and the smooth Cooper-Frye spectra of pi+, K+ and
p on the reference's (pT, phi, eta_s) quadrature grid over the same surface,
FP32 and FP64.

--workload c3 measures the batched configuration instead: 256 surfaces of
log-normal size (median 3e4 cells, 1e3..5e5), each with its own event count
in [10, 2000]; one step = one yield pass over all of them + one batched
sampling pass (rounds of autotuned size replayed from a captured hipGraph).
--workload c5 is the stress configuration: one 3+1D surface of 2M cells with
mu_B up to 0.45 GeV, 5000 events of ~1e4 hadrons per step, in the same
batched persistent rounds.  Both run the plain proposal, then the learned one
(network trained on ~1e7 hadrons of a plain pass, outside the timed region);
value = the learned rate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--workload c2|c3|c5]
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
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector, packed (78.6 unpacked)

# Algorithmic units of the path's measurement: a sampling pass
# costs 200 FP32 ops per trial (|p| proposal ~115 incl. inner tries, outer
# trial ~85) plus 250 per accepted hadron (CDF refresh ~200, boost and record
# ~50), FMA = 2; one network evaluation 2 (14 x 64 + 64 x 64 + 64 x 8) FLOP
# (8 outputs) on the BF16 MFMA
OPS_PER_TRIAL = 200.0
OPS_PER_ACCEPTED = 250.0
NET_FLOP_PER_PAIR = 2.0 * (14 * 64 + 64 * 64 + 64 * 8)   # 11,008
BF16_PEAK_TFLOPS = 2500.0      # MI355X dense BF16 MFMA
# one smooth-spectrum integrand point as the reference writes it
# (src/emissionfunction.cpp:756-806): pt, pz, p.u, exponent, exp, f0,
# p.dsigma, the ten-term shear contraction, shear and bulk delta f, the
# |delta f| restriction, the result and the weighted sum
OPS_SPECTRA_POINT = 70.0
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector


def yield_bytes(n_cells: int, n_species: int, esz: int = 4) -> int:
    """Algorithmic HBM bytes of one yield pass: per cell T, muB, muS, muQ,
    u.dsigma, Pi (6 x 4 B), per (cell, species) one yield in the working
    precision (esz bytes), and the FP64 species totals."""
    return n_cells * 4 * 6 + n_cells * n_species * esz + 8 * n_species


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(cells, seconds_target=10.0, total_cells=None):
    """The CPU oracle (a restatement of the reference algorithm: yields,
    Poisson multiplicities, |p| tables and rejection, boost) on a bounded
    sample of the same workload: the first 2000 cells of the surface, as many
    events as fill ~10 s, on one thread (the reference is single-threaded) and
    on all host cores (std::thread over events, each thread with its own
    copy and seeds).  `value` is the all-core rate."""
    from tests import oracle_binding as ob

    sub = cells[:2000]
    orc = ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=11, fd_quirk=0)
    orc.set_cells_lrf(sub)
    t0 = time.perf_counter()
    _, h = orc.sample(1, 200)
    dt = time.perf_counter() - t0
    per_event = max(h.size / 200.0, 1e-3)
    rate0 = max(h.size / max(dt, 1e-9), 1.0)
    nev = int(np.clip(seconds_target * rate0 / per_event, 200, 2_000_000))
    t0 = time.perf_counter()
    n1 = orc.o.orc_sample_mt(orc.h, 2, nev, 1)
    dt1 = time.perf_counter() - t0
    # the box's CPU share for one GPU is 16 cores
    cores = max(1, min(os.cpu_count() or 1, 16))
    nev_all = nev * cores
    t0 = time.perf_counter()
    nall = orc.o.orc_sample_mt(orc.h, 3, nev_all, cores)
    dtall = time.perf_counter() - t0
    host = f"host {cpu_model()}, nproc {os.cpu_count()}"
    return {"value": nall / dtall, "unit": "hadrons/s", "cores": cores, "kind": "port",
            "sample": f"oracle (yields + sampling) on the first 2000 of the {total_cells or cells.shape[0]:,} cells: {cores} threads x "
                      f"{nev} events, {nall} hadrons in {dtall:.1f} s; {host}",
            "single_thread": {"value": n1 / dt1, "cores": 1,
                              "sample": f"{nev} events, {n1} hadrons in {dt1:.1f} s"}}


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE and WRITE_SIZE, calibrated), or None."""
    f = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(f) as fh:
            return json.load(fh).get(kernel)
    except (OSError, ValueError):
        return None


def calibration_copy(nbytes):
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def run_steps(ctx, args, proposal, first_event):
    def step(i):
        ctx.compute_yields(precision=args.precision)
        n = ctx.sample(args.events, seed=12345, event0=first_event + i * args.events,
                       precision=args.precision, proposal=proposal)
        return n, ctx.stats()

    for i in range(args.warmup):
        step(i)
    acc = {"hadrons": 0, "trials": 0, "inner": 0, "ms_yield": [], "ms_sample": [], "ms_setup": [],
           "ms_trials": [], "ms_ltrials": [], "ms_rsetup": [], "ms_rtrials": [], "violations": 0, "fallback": 0,
           "pairs": 0, "routed": 0, "routed_trials": 0}
    t0 = time.perf_counter()
    for i in range(args.steps):
        n, st = step(args.warmup + i)
        acc["hadrons"] += n
        acc["trials"] += st.n_trials
        acc["inner"] += st.n_inner
        acc["ms_yield"].append(st.ms_yield_kernel)
        acc["ms_sample"].append(st.ms_sample)
        acc["ms_setup"].append(st.ms_setup_kernel)
        acc["ms_trials"].append(st.ms_trials_kernel)
        acc["violations"] += st.n_violations
        acc["fallback"] += st.n_fallback
        acc["pairs"] += st.n_pairs if proposal == 1 else 0
        acc["routed"] += st.n_routed if proposal == 1 else 0
        acc["routed_trials"] += st.n_routed_trials
        acc["ms_ltrials"].append(st.ms_learned_trials_kernel)
        acc["ms_rsetup"].append(st.ms_routed_setup_kernel)
        acc["ms_rtrials"].append(st.ms_routed_trials_kernel)
    acc["elapsed"] = time.perf_counter() - t0
    for k in ("ms_yield", "ms_sample", "ms_setup", "ms_trials", "ms_ltrials", "ms_rsetup", "ms_rtrials"):
        acc[k] = float(np.mean(acc[k]))
    return acc


def hw_from_profiles(kernel: str):
    """Measured utilisation of `kernel` (VALU issue-slot share, FP32 and BF16
    MFMA FLOP/s against peak) from the committed rocprofv3 counter summary,
    or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "utilisation.json")) as fh:
            u = json.load(fh)
    except (OSError, ValueError):
        return None
    for k, v in u.items():
        if isinstance(v, dict) and k.split("<")[0] == kernel:
            return v
    return None


def roofline_valu(kernel, ops, ms):
    tf = ops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"bound": "valu", "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP32_PEAK_TFLOPS, "traffic": traffic_from_profiles(kernel), "kernel": kernel, "ms": ms,
            "hardware": hw_from_profiles(kernel)}


def roofline_mfma(kernel, flop, ms):
    tf = flop / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"bound": "mfma", "achieved": tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / BF16_PEAK_TFLOPS, "traffic": traffic_from_profiles(kernel), "kernel": kernel, "ms": ms,
            "hardware": hw_from_profiles(kernel)}


def spectra_extra(ctx, sp, n_cells):
    """Smooth spectra of pi+, K+, p on the reference grid, FP32 and FP64."""
    from iss_amd import tables
    b = tables.bin_tables()
    idx = np.array([int(np.nonzero(sp["pid"] == p)[0][0]) for p in (211, 321, 2212)], dtype=np.int32)
    out = {"species": [211, 321, 2212], "grid": [int(b["pT"].size), int(b["phi"].size), int(b["eta"].size)],
           "cells": n_cells}
    points = idx.size * b["pT"].size * b["phi"].size * b["eta"].size * n_cells
    for prec in (32, 64):
        ms = []
        for _ in range(3):
            ctx.spectra(idx, b["pT"], b["phi"], b["eta"], b["eta_weight"], precision=prec)
            ms.append(ctx.spectra_ms())
        t = float(min(ms))
        tf = OPS_SPECTRA_POINT * points / (t * 1e-3) / 1e12
        peak = FP32_PEAK_TFLOPS if prec == 32 else FP64_PEAK_TFLOPS
        out[f"fp{prec}"] = {"ms": t, "points_per_s": points / (t * 1e-3),
                            "roofline": {"bound": "valu", "achieved": tf, "peak": peak, "unit": "TFLOP/s",
                                         "frac": tf / peak, "kernel": "k_spectra"}}
    return out


def run_batched(args):
    """--workload c3 (256 uneven surfaces) or c5 (one 2M-cell surface with
    mu_B up to 0.45 GeV, 5000 events): batched persistent rounds replayed from
    a captured hipGraph, plain proposal, then the learned one on the same
    surfaces; value = the learned rate."""
    import iss_amd
    from iss_amd import synth, tables

    c5 = args.workload == "c5"
    sp = tables.species_arrays(1)
    t0 = time.perf_counter()
    if c5:
        cells, _ = synth.viscous_lrf(2_000_000, sp, seed=5, muB_max=0.45, target_hadrons=1.0e4)
        surfaces, ev = [cells], np.array([5000], dtype=np.int32)
    else:
        surfaces, ev = synth.batched_surfaces(sp, n_surfaces=args.surfaces)
    t_gen = time.perf_counter() - t0
    ncell = np.array([s.shape[0] for s in surfaces])
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1), fermion_cdf2_quirk=0)
    if c5:
        ctx.set_surface_lrf(cells)
        tot = ctx.compute_yields(precision=64)
        cells[:, 4:8] *= np.float32(1.0e4 / tot.sum())
    ctx.set_surfaces_lrf(surfaces)
    sub = surfaces[0][:2000].copy() if c5 else None
    del surfaces
    tune = {}
    if not args.no_tune:
        ctx.compute_yields(precision=args.precision)
        tune["plain"] = ctx.autotune(proposal=0, n_events=2, seed=7, repeats=1)
        # (whole batches of the workload: the round size and the learned
        # routing depend on the hadrons per (cell, species) pair)
        tune["batch_plain"] = ctx.autotune_batch(ev, seed=7, precision=args.precision, proposal=0)
    # the network, trained on ~1e7 hadrons accepted by a plain pass over the
    # whole record (outside the timed region)
    per_event = float(ctx.compute_yields(precision=args.precision).sum())
    n_train_ev = int(np.clip(1.0e7 / max(per_event, 1.0), 1, 1000))
    ctx.sample(n_train_ev, seed=3, precision=args.precision)
    n_train = ctx.mlp_collect(10_000_000)
    ctx.mlp_init(1)
    rep = ctx.mlp_train(batch=8192, steps=args.train_steps, lr=3e-3, seed=5)
    if not args.no_tune:
        tune["learned"] = ctx.autotune(proposal=1, n_events=2, seed=7, repeats=1)
        tune["batch_learned"] = ctx.autotune_batch(ev, seed=7, precision=args.precision, proposal=1)

    def run(proposal):
        def step(i):
            ctx.compute_yields(precision=args.precision)
            n = ctx.sample_batch(ev, seed=100 + i, precision=args.precision, use_graph=1, proposal=proposal)
            return n, ctx.stats(), ctx.batch_ms()

        for i in range(args.warmup):
            step(i)
        acc = {"hadrons": 0, "ms_yield": [], "ms_batch": [], "violations": 0, "fallback": 0, "trials": 0,
               "routed": 0}
        t0 = time.perf_counter()
        for i in range(args.steps):
            n, st, mb = step(args.warmup + i)
            acc["hadrons"] += n
            acc["ms_yield"].append(st.ms_yield_kernel)
            acc["ms_batch"].append(mb)
            acc["violations"] += st.n_violations
            acc["fallback"] += st.n_fallback
            acc["trials"] += st.n_trials
            acc["routed"] += st.n_routed
        acc["elapsed"] = time.perf_counter() - t0
        acc["R"] = ctx.lib.iss_batch_events_per_round(ctx.h)
        return acc

    plain = run(0)
    learned = run(1)
    nc, ns = ctx.num_cells(), ctx.n_species
    # (the batched pass stores yield rows in the working precision)
    ybytes = yield_bytes(nc, ns, 4 if args.precision == 32 else 8)
    my = float(np.mean(learned["ms_yield"]))
    y_gbs = ybytes / (my * 1e-3) / 1e9
    v_learn = learned["hadrons"] / learned["elapsed"]
    v_plain = plain["hadrons"] / plain["elapsed"]
    batch_rate = {k: (r["hadrons"] / args.steps) / (1e-3 * float(np.mean(r["ms_batch"])))
                  for k, r in (("learned", learned), ("plain", plain))}
    if c5:
        config = {"workload": "stress: one 3+1D viscous surface, 2M cells, mu_B <= 0.45 GeV, 5000 events per step, "
                              "learned proposal in batched persistent rounds",
                  "cells": int(nc), "events_per_step": int(ev[0]), "species": ns}
    else:
        config = {"workload": "batched uneven surfaces, learned proposal", "surfaces": int(ncell.size),
                  "cells": int(nc), "cells_min_median_max": [int(ncell.min()), int(np.median(ncell)),
                                                             int(ncell.max())],
                  "events_min_max": [int(ev.min()), int(ev.max())], "species": ns}
    config.update({"hadrons_per_step": learned["hadrons"] / args.steps,
                   "events_per_round": {k: r["R"] for k, r in (("learned", learned), ("plain", plain))},
                   "generation_s": t_gen,
                   "deltaf": "shear + OSU 14-moment bulk", "train_samples": int(n_train),
                   "train_steps": args.train_steps})
    out = {
        "metric": "sampled hadrons/s (learned proposal vs plain GPU) at fixed chi2; yield-pass GB/s",
        "value": v_learn,
        "unit": "hadrons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * learned["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": y_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": y_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "k_yield_rows", "ms": my,
                     "bytes": ybytes},
        "cpu_baseline": None,
        "plain_hadrons_per_s": v_plain,
        "learned_vs_plain": v_learn / v_plain,
        "batch_pass_hadrons_per_s": batch_rate,
        "proposals_per_hadron": {k: r["trials"] / max(r["hadrons"], 1) for k, r in
                                 (("learned", learned), ("plain", plain))},
        "envelope_violations": learned["violations"],
        "fallback_hadrons": learned["fallback"],
        "routed_hadrons_per_step": learned["routed"] / args.steps,
        "routing_min_hadrons": ctx.learned_routing(),
        "training": rep,
        "autotune": tune,
    }
    ctx.close()
    if c5 and not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sub, total_cells=int(nc))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=200000)
    ap.add_argument("--events", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=32)
    ap.add_argument("--train-steps", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-spectra", action="store_true", help="skip the smooth-spectra measurement")
    ap.add_argument("--workload", choices=["c2", "c3", "c5"], default="c2")
    ap.add_argument("--surfaces", type=int, default=256, help="c3: number of surfaces")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a known-size kernel for "
                         "calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the sampler has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("c3", "c5"):
        run_batched(args)
        return

    import iss_amd
    from iss_amd import synth, tables

    if args.calibration_copy_mib > 0:
        calibration_copy(args.calibration_copy_mib << 20)
    sp = tables.species_arrays(1)
    cells, _ = synth.viscous_lrf(args.cells, sp, seed=2, muB_max=0.3, target_hadrons=1.0e4)
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1), fermion_cdf2_quirk=0)
    ctx.set_surface_lrf(cells)
    tot = ctx.compute_yields(precision=64)
    cells[:, 4:8] *= np.float32(1.0e4 / tot.sum())
    ctx.set_surface_lrf(cells)
    nc, ns = ctx.num_cells(), ctx.n_species

    tune = {}
    if not args.no_tune:
        ctx.compute_yields(precision=args.precision)
        # (passes of the workload's size: the learned routing depends on the
        # hadrons per (cell, species) pair of a pass)
        tune["plain"] = ctx.autotune(proposal=0, n_events=args.events, seed=7, repeats=3)
    plain = run_steps(ctx, args, 0, 0)
    # training set: the hadrons accepted by the last plain pass
    n_train = ctx.mlp_collect(10_000_000)
    ctx.mlp_init(1)
    rep = ctx.mlp_train(batch=8192, steps=args.train_steps, lr=3e-3, seed=5)
    if not args.no_tune:
        tune["learned"] = ctx.autotune(proposal=1, n_events=args.events, seed=7, repeats=3)
    learned = run_steps(ctx, args, 1, (args.warmup + args.steps) * args.events)

    v_learn = learned["hadrons"] / learned["elapsed"]
    v_plain = plain["hadrons"] / plain["elapsed"]
    hl, hp = learned["hadrons"] / args.steps, plain["hadrons"] / args.steps
    ybytes = yield_bytes(nc, ns, 4 if args.precision == 32 else 8)
    y_gbs = ybytes / (learned["ms_yield"] * 1e-3) / 1e9
    roof_y = {"bound": "hbm", "achieved": y_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": y_gbs / HBM_PEAK_GBS, "traffic": traffic_from_profiles("k_yield_rows"), "kernel": "k_yield_rows",
              "ms": learned["ms_yield"], "bytes": ybytes, "hardware": hw_from_profiles("k_yield_rows")}
    pairs = learned["pairs"] / args.steps
    routed = learned["routed"] / args.steps
    # every pair routed (the autotuner's choice where no pair pays): the
    # learned step ran the plain pass, whose set-up is the plain one
    plain_step = learned["ms_ltrials"] == 0 and learned["ms_rtrials"] == 0
    if plain_step:
        learned["ms_lsetup"] = 0.0
    else:
        learned["ms_lsetup"] = learned["ms_setup"]
    # the network (one evaluation per pair) against the BF16 MFMA peak; the
    # sampling pass and its trial kernels against the FP32 vector peak
    l_setup = roofline_mfma("k_learn_setup", NET_FLOP_PER_PAIR * pairs, learned["ms_lsetup"])
    lt, pt_ = learned["trials"] / args.steps, plain["trials"] / args.steps
    # the learned step's trial kernels apart: the learned one on the pairs'
    # hadrons, the plain one on the routed and envelope-less hadrons
    rt = learned["routed_trials"] / args.steps
    hr = routed + learned["fallback"] / args.steps
    l_trials = roofline_valu("k_learn_trials", OPS_PER_TRIAL * (lt - rt) + OPS_PER_ACCEPTED * (hl - hr),
                             learned["ms_ltrials"])
    r_trials = roofline_valu("k_trials_routed", OPS_PER_TRIAL * rt + OPS_PER_ACCEPTED * hr, learned["ms_rtrials"])
    p_trials = roofline_valu("k_trials", OPS_PER_TRIAL * pt_ + OPS_PER_ACCEPTED * hp, plain["ms_trials"])
    pass_l = roofline_valu("sampling pass (learned)", OPS_PER_TRIAL * lt + OPS_PER_ACCEPTED * hl, learned["ms_sample"])
    pass_p = roofline_valu("sampling pass (plain)", OPS_PER_TRIAL * pt_ + OPS_PER_ACCEPTED * hp, plain["ms_sample"])
    p_setup = {"kernel": "k_setup", "ms": plain["ms_setup"], "hardware": hw_from_profiles("k_setup")}
    # the dominant kernel of the learned step (the value's step); with every
    # pair routed that step is the plain pass
    cands = [(learned["ms_yield"], roof_y), (learned["ms_lsetup"], l_setup), (learned["ms_ltrials"], l_trials),
             (learned["ms_rtrials"], r_trials)]
    if plain_step:
        cands.append((learned["ms_trials"], roofline_valu("k_trials", OPS_PER_TRIAL * lt + OPS_PER_ACCEPTED * hl,
                                                          learned["ms_trials"])))
    dominant = max(cands, key=lambda t: t[0])[1]
    out = {
        "metric": "sampled hadrons/s (learned proposal vs plain GPU) at fixed chi2; yield-pass GB/s",
        "value": v_learn,
        "unit": "hadrons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * learned["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == 32 else "f64",
        "data": "synthetic",
        "config": {"workload": "viscous 3+1D surface, 200k cells, learned proposal", "cells": nc,
                   "species": ns, "events_per_step": args.events, "hadrons_per_event": hl / args.events,
                   "deltaf": "shear + OSU 14-moment bulk", "proposal": "learned (trained on the plain "
                   "sampler's accepted hadrons)", "train_samples": n_train, "train_steps": rep["steps"]},
        "roofline": dominant,
        "cpu_baseline": None,
        "plain_hadrons_per_s": v_plain,
        "learned_vs_plain": v_learn / v_plain,
        "sample_pass_hadrons_per_s": {"learned": hl / (1e-3 * learned["ms_sample"]),
                                      "plain": hp / (1e-3 * plain["ms_sample"])},
        "proposals_per_hadron": {"learned": learned["inner"] / learned["hadrons"],
                                 "plain": plain["inner"] / plain["hadrons"]},
        "envelope_violations": learned["violations"],
        "learned_pairs_per_step": pairs,
        "hadrons_per_pair": (hl - routed) / max(pairs, 1.0),
        "routed_hadrons_per_step": routed,
        "routing_min_hadrons": ctx.learned_routing(),
        "fallback_hadrons": learned["fallback"],
        "training": rep,
        "autotune": tune,
        "learned_step_is_plain_pass": plain_step,
        "kernels_ms": {"k_yield_rows": learned["ms_yield"], "k_learn_setup": learned["ms_lsetup"],
                       "k_learn_trials": learned["ms_ltrials"], "k_setup (routed)": learned["ms_rsetup"],
                       "k_trials_routed": learned["ms_rtrials"], "k_setup": plain["ms_setup"],
                       "k_trials": plain["ms_trials"]},
        "roofline_yield": roof_y,
        "roofline_sample": {"pass_learned": pass_l, "pass_plain": pass_p, "k_learn_setup": l_setup,
                            "k_learn_trials": l_trials, "k_trials_routed": r_trials, "k_setup": p_setup,
                            "k_trials": p_trials},
        "yield_GBs": y_gbs,
    }
    # resonance feed-down of the last learned step's hadrons (outside the timed region)
    ctx.set_decay_table(tables.decay_table(1))
    n_in = ctx.stats().n_hadrons
    ms_dec = []
    for i in range(3):
        if i:
            ctx.sample(args.events, seed=777 + i, precision=args.precision, proposal=1)
            n_in = ctx.stats().n_hadrons
        n_out = ctx.decay(seed=99 + i)
        ms_dec.append(ctx.stats().ms_decay)
    st = ctx.stats()
    out["decays"] = {"hadrons_in": int(n_in), "hadrons_out": int(n_out), "generations": int(st.n_decay_generations),
                     "ms": float(min(ms_dec)), "hadrons_per_s": n_in / (1e-3 * float(min(ms_dec)))}
    if not args.no_spectra:
        out["spectra"] = spectra_extra(ctx, sp, nc)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cells)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
