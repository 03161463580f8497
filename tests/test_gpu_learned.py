"""Learned proposal on the GPU: kernels against the oracle, exactness, gain.

* Hadron parameters and band envelopes (iss_band_diag) against
  oracle/mlp_oracle.py on realistic viscous cells.
* Network: the kernels' inputs against the oracle's features, the loss and
  its output gradient against the oracle's forward pass and band envelope,
  and every parameter gradient against the oracle's BF16-emulating backward
  pass fed with the kernel's own output gradient.
* Exactness: with a trained network the learned sampler reproduces the plain
  sampler's species, event sizes and spectra (two-sample chi^2 / KS), draws
  no proposal above the envelope, and passes the reference's one-cell
  T^{mu nu} criterion (tests/TestOutputFiles.py:5-13), also with the
  Fermi-Dirac table quirk handing fermions to the plain sampler.
* Gain: proposals per hadron of plain, untrained (every kappa = kappa0/2)
  and trained networks.
"""
import ctypes
import os
import sys

import numpy as np
import pytest
from scipy import stats as sst

import iss_amd
from iss_amd import synth, tables
from tests import oracle_binding as ob
from tests.sample_stats import GROUPS, obs as _obs, quantile_bins as _quantile_bins, two_sample_chi2 as _two_sample_chi2
from tests.test_gpu_sampling import tmunu_from_hadrons

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu

# routing threshold of the routed learned pass (hadrons per (cell, species))
ROUTE = 3


def _ctx(quirk=0, shear=1, bulk=1, bulk_kind=11):
    ctx = iss_amd.Context(0)
    sp = tables.species_arrays(1)
    ctx.set_species(sp)
    kw = dict(shear=shear, bulk=bulk, bulk_kind=bulk_kind, fermion_cdf2_quirk=quirk)
    if bulk and bulk_kind == 11:
        kw["k11"] = tables.read_kind11(1)
    ctx.set_deltaf(**kw)
    return ctx, sp


def _pairs(sp, ncell, n, seed):
    rng = np.random.default_rng(seed)
    light = np.nonzero(sp["mass"] < 1.0)[0]
    c = rng.integers(0, ncell, n).astype(np.int32)
    s = np.where(rng.random(n) < 0.7, rng.choice(light, n), rng.integers(0, len(sp["mass"]), n)).astype(np.int32)
    return c, s


def _oracle_params(orc, lrf, sp, c, s):
    v = np.zeros(3)
    orc.o.orc_k11_coeff(orc.h, float(lrf[c, 13]), float(lrf[c, 16]), v.ctypes.data_as(ctypes.c_void_p))
    cf = np.concatenate([v, np.zeros(3)])
    return mo.hadron_params(lrf[c], cf, sp, s, 1, 11, -1, 0)


@pytest.fixture(scope="module")
def small_surface():
    ctx, sp = _ctx()
    cells, _ = synth.viscous_lrf(500, sp, seed=21, target_hadrons=500.0)
    ctx.set_surface_lrf(cells)
    lrf = ctx.get_surface_lrf()
    orc = ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=11, fd_quirk=0)
    return ctx, sp, lrf, orc


def test_band_envelopes_match_oracle(small_surface):
    ctx, sp, lrf, orc = small_surface
    c, s = _pairs(sp, lrf.shape[0], 300, seed=3)
    rho = np.random.default_rng(4).uniform(0.05, 0.95, (c.size, 8)).astype(np.float32)
    tab, par = ctx.band_diag(c, s, rho)
    n_ok = 0
    for i in range(c.size):
        P = _oracle_params(orc, lrf, sp, int(c[i]), int(s[i]))
        assert int(par[i, 21]) == P["flags"]
        for j, key in enumerate(("T", "mu", "m", "sgn", "x", "tmax", "d0", "dn", "D", "l1", "l2", "l3")):
            assert abs(par[i, j] - P[key]) <= 2e-6 * max(1.0, abs(P[key])) + 2e-7 * abs(P["l1"]) * (j >= 9), (i, key)
        for j, key in ((15, "ps0"), (16, "qi"), (17, "q0"), (18, "q1"), (19, "q2")):
            assert abs(par[i, j] - P[key]) <= 1e-5 * max(1e-3, abs(P[key])), (i, key, par[i, j], P[key])
        if P["flags"]:
            continue
        ok, lm, dz, t = mo.band_setup(P, rho[i].astype(np.float64))
        assert bool(tab[i, 0]) == ok
        assert abs(tab[i, 1] - lm) < 2e-2, (i, tab[i, 1], lm)
        assert int(tab[i, 46]) == t["klast"]
        live = t["mass"] > 0
        for off, key, tol in ((2, "c", 2e-2), (11, "s", 2e-2), (20, "kap", 2e-2)):
            g = tab[i, off:off + 9][live]
            o = t[key][live]
            assert np.all(np.abs(g - o) <= tol * np.maximum(1.0, np.abs(o))), (i, key, g, o)
        assert np.all(np.abs(tab[i, 29:38][live] - t["beta"][live]) <= 1.01 / 255)
        assert np.all(np.abs(tab[i, 38:46] - t["Q"]) < 2e-2)
        n_ok += 1
    assert n_ok > 200


def test_network_passes_match_oracle(small_surface):
    ctx, sp, lrf, orc = small_surface
    ctx.mlp_init(seed=3)
    par = ctx.mlp_get_params().astype(np.float64)
    c, s = _pairs(sp, lrf.shape[0], 256, seed=5)
    r = ctx.mlp_eval(c, s)
    X = r["x"].astype(np.float64)
    valid = np.zeros(c.size, dtype=bool)
    dz_o = np.zeros((c.size, 8))
    lm_o = np.zeros(c.size)
    for i in range(c.size):
        P = _oracle_params(orc, lrf, sp, int(c[i]), int(s[i]))
        if P["flags"]:
            assert r["loss"][i] == 0 and np.all(r["dz"][i] == 0)
            continue
        f = mo.bf16(mo.features(lrf[c[i]], sp, int(s[i]), P))
        assert np.all(np.abs(X[i] - f) <= 2.0 ** -7 * np.abs(f) + 1e-6), (i, X[i], f)
        O, _ = mo.forward(par, X[i:i + 1])
        ok, lm, dz, _ = mo.band_setup(P, 1 / (1 + np.exp(-O[0])))
        valid[i] = ok
        lm_o[i], dz_o[i] = lm, dz
    assert valid.sum() > 200
    assert np.max(np.abs(r["loss"][valid] - lm_o[valid])) < 2e-2
    scale = np.abs(dz_o[valid]).max()
    assert np.max(np.abs(r["dz"][valid] - dz_o[valid])) < 5e-2 * scale + 1e-6
    # backward: the oracle fed with the kernel's own inputs and output gradient
    O, cache = mo.forward(par, X)
    g_o = mo.backward(par, cache, r["dz"].astype(np.float64))
    for lo, hi in ((mo.P_W1, mo.P_B1), (mo.P_B1, mo.P_W2), (mo.P_W2, mo.P_B2), (mo.P_B2, mo.P_W3),
                   (mo.P_W3, mo.P_B3), (mo.P_B3, mo.N_PARAMS)):
        gg, go = r["grad"][lo:hi], g_o[lo:hi]
        assert np.max(np.abs(gg - go)) <= 2e-3 * np.abs(go).max() + 1e-7, (lo, np.max(np.abs(gg - go)), np.abs(go).max())


@pytest.fixture(scope="module")
def trained():
    """Plain pass, training set, training, and three sampling passes of one
    C2-like surface (20k cells)."""
    ctx, sp = _ctx()
    cells, _ = synth.viscous_lrf(20000, sp, seed=2, target_hadrons=2000.0)
    ctx.set_surface_lrf(cells)
    ctx.compute_yields(precision=32)
    nev = 400
    out = {"sp": sp, "nev": nev}
    ctx.sample(nev, seed=11)
    out["plain"] = (ctx.hadrons().copy(), ctx.event_counts().copy(), ctx.stats().as_dict())
    out["n_train"] = ctx.mlp_collect(2_000_000)
    # the network's gain is measured with every pair on the learned proposal
    ctx.set_learned_routing(1)
    ctx.mlp_init(0)
    ctx.sample(nev, seed=11, proposal=1)
    out["zero"] = (ctx.hadrons().copy(), ctx.event_counts().copy(), ctx.stats().as_dict())
    ctx.mlp_init(1)
    out["report"] = ctx.mlp_train(batch=8192, steps=300, lr=3e-3, seed=5)
    ctx.sample(nev, seed=11, proposal=1)
    out["learned"] = (ctx.hadrons().copy(), ctx.event_counts().copy(), ctx.stats().as_dict())
    ctx.sample(nev, seed=11, proposal=1)
    out["learned2"] = ctx.hadrons().copy()
    # routing: pairs of fewer than ROUTE hadrons take the plain sampler
    ctx.set_learned_routing(ROUTE)
    ctx.sample(nev, seed=11, proposal=1)
    out["routed"] = (ctx.hadrons().copy(), ctx.event_counts().copy(), ctx.stats().as_dict())
    ctx.sample(nev, seed=11, proposal=1)
    out["routed2"] = ctx.hadrons().copy()
    print("train", out["report"], "set", out["n_train"])
    for k in ("plain", "zero", "learned", "routed"):
        st = out[k][2]
        print(k, "hadrons", st["n_hadrons"], "proposals/hadron %.3f" % (st["n_inner"] / st["n_hadrons"]),
              "fallback", st["n_fallback"], "routed", st["n_routed"], "pairs", st["n_pairs"],
              "violations", st["n_violations"], "ms %.3f" % st["ms_sample_kernel"])
    return out


def test_training_lowers_envelope_mass(trained):
    rep = trained["report"]
    assert trained["n_train"] > 100000
    assert rep["loss_last"] < rep["loss_first"] - 0.02


@pytest.mark.parametrize("which", ["learned", "routed"])
def test_learned_is_exact_and_deterministic(trained, which):
    hp, cp, sp_ = trained["plain"]
    hl, cl, sl = trained[which]
    assert sl["n_violations"] == 0 and trained["zero"][2]["n_violations"] == 0
    assert sl["n_fallback"] == 0 and sl["n_failed"] == 0
    # the same species draws and event assignment as the plain sampler
    assert np.array_equal(cp, cl)
    assert np.array_equal(np.sort(hp["pid"]), np.sort(hl["pid"]))
    assert np.array_equal(hl, trained[which + "2"])
    # momenta: a different, independent sample of the same distribution
    tot_chi, tot_ndf = 0.0, 0
    nev = trained["nev"]
    for name, pids in GROUPS.items():
        if pids is None:
            allp = sum((v for v in GROUPS.values() if v), [])
            a, b = hl[~np.isin(hl["pid"], allp)], hp[~np.isin(hp["pid"], allp)]
        else:
            a, b = hl[np.isin(hl["pid"], pids)], hp[np.isin(hp["pid"], pids)]
        pa, ya, fa = _obs(a)
        pb, yb, fb = _obs(b)
        chi, ndf = 0.0, 0
        for u, v, lo, hi in ((pa, pb, 0.0, 10.0), (ya, yb, -12.0, 12.0), (fa, fb, -np.pi, np.pi)):
            cc, nn = _two_sample_chi2(u, nev, v, nev, _quantile_bins(u, v, lo, hi, per_bin=200, max_bins=60))
            chi += cc
            ndf += nn
        edges = _quantile_bins(pa, pb, 0.0, 10.0, per_bin=2000, max_bins=12)
        for k in range(len(edges) - 1):
            ca = np.cos(2 * fa[(pa >= edges[k]) & (pa < edges[k + 1])])
            cb = np.cos(2 * fb[(pb >= edges[k]) & (pb < edges[k + 1])])
            chi += (ca.mean() - cb.mean()) ** 2 / (ca.var() / ca.size + cb.var() / cb.size)
            ndf += 1
        ks = sst.ks_2samp(pa, pb).pvalue
        print(name, "chi2/ndf %.3f ndf %d KS p %.3g" % (chi / ndf, ndf, ks))
        assert chi / ndf <= 1 + 4 * np.sqrt(2.0 / ndf)
        assert ks > 0.01
        tot_chi += chi
        tot_ndf += ndf
    print("combined chi2/ndf %.4f ndf %d" % (tot_chi / tot_ndf, tot_ndf))
    assert tot_chi / tot_ndf <= 1.1


def test_routing_counts_small_pairs(trained):
    """Routing sends exactly the (cell, species) pairs of fewer than ROUTE
    hadrons in the pass to the plain sampler: recounted from the output,
    where a cell is its position (unique on this surface) and a species its
    pid."""
    h, _, st = trained["routed"]
    _, _, st1 = trained["learned"]
    key = np.stack([h["x"], h["y"], h["z"], h["t"]], axis=1)
    _, cell_of = np.unique(key, axis=0, return_inverse=True)
    pair = np.stack([cell_of.ravel().astype(np.int64), h["pid"].astype(np.int64)], axis=1)
    _, mult = np.unique(pair, axis=0, return_counts=True)
    print("pairs", mult.size, "of fewer than", ROUTE, "hadrons", int(np.sum(mult < ROUTE)), "routed hadrons",
          st["n_routed"], "learned pairs", st["n_pairs"], "(all pairs without routing:", st1["n_pairs"], ")")
    assert st1["n_routed"] == 0 and st1["n_pairs"] == mult.size
    assert st["n_routed"] == np.sum(mult[mult < ROUTE])
    assert st["n_pairs"] == np.sum(mult >= ROUTE)
    # the routed hadrons go through the plain sampler's set-up, not the fallback count
    assert st["n_fallback"] == 0


def test_learned_needs_fewer_proposals(trained):
    rate = {k: trained[k][2]["n_inner"] / trained[k][2]["n_hadrons"] for k in ("plain", "zero", "learned")}
    print("proposals per hadron", rate)
    assert rate["learned"] < rate["zero"] < rate["plain"]
    assert rate["learned"] < 0.6 * rate["plain"]


@pytest.mark.parametrize("case", list(ob.REF_CASES))
def test_learned_reference_tmunu(case):
    """The one-cell reference cases with the learned proposal (untrained
    network) and the reference's Fermi-Dirac table quirk: fermions go to the
    plain sampler, everything else through the envelope."""
    cfg = ob.case_config(case)
    ctx = iss_amd.Context(0)
    sp = tables.species_arrays(cfg["afterburner"])
    ctx.set_species(sp)
    kw = dict(shear=cfg["shear"], bulk=cfg["bulk"], bulk_kind=cfg["bulk_kind"])
    if cfg["bulk_kind"] == 21:
        kw["neos"] = tables.read_neos(21, cfg["afterburner"])
    if cfg["bulk"] and cfg["bulk_kind"] == 11:
        kw["k11"] = tables.read_kind11(cfg["afterburner"])
    ctx.set_deltaf(**kw)
    lab = tables.read_music_surface(cfg["surface"], bulk=1, rhob=cfg["rhob"])
    ctx.set_surface_lab(lab, eos=tables.read_eos(cfg["iEOS"], cfg["afterburner"]))
    lrf = ctx.get_surface_lrf()
    volume = lrf[0, 4] / lrf[0, 8]
    ctx.mlp_init(0)
    nev = 1000
    n = ctx.sample(nev, seed=30240 + len(case), proposal=1)
    st = ctx.stats()
    assert st.n_failed == 0 and st.n_range_errors == 0 and st.n_violations == 0
    assert st.n_fallback > 0
    h = ctx.hadrons()
    assert h.size == n and np.all(h["pid"] != 0)
    Ts, nBs = tmunu_from_hadrons(h, sp, nev, volume)
    orc = ob.Oracle(afterburner=cfg["afterburner"], shear=cfg["shear"], bulk=cfg["bulk"],
                    bulk_kind=cfg["bulk_kind"])
    orc.read_surface(cfg["surface"], bulk=1, rhob=cfg["rhob"], eos_file=cfg["eos_file"], iEOS=cfg["iEOS"])
    Tf, Qf = orc.surface_tmunu()
    diff = np.concatenate([Tf - Ts, [Qf[0] - nBs]])
    print(case, "fallback", st.n_fallback, "of", n, "mean|diff|", np.mean(np.abs(diff)))
    assert np.mean(np.abs(diff)) < 1e-3


def test_autotuned_launch_shape_keeps_the_samples():
    """The autotuner's launch shapes leave the samples alone: with
    counter-based random numbers the hadrons of a pass are bit-identical
    before and after (at the same routing threshold, which the learned
    proposal's tuning also chooses: it decides which sampler draws a hadron's
    momentum, not its law)."""
    ctx, sp = _ctx()
    cells, _ = synth.viscous_lrf(3000, sp, seed=8, target_hadrons=600.0)
    ctx.set_surface_lrf(cells)
    ctx.compute_yields(precision=32)
    ctx.mlp_init(0)
    route0 = ctx.learned_routing()
    ref = {}
    for prop in (0, 1):
        ctx.sample(200, seed=21, proposal=prop)
        ref[prop] = ctx.hadrons().copy()
    for prop in (0, 1):
        rep = ctx.autotune(proposal=prop, n_events=100, seed=3, repeats=1)
        print("proposal", prop, rep)
        assert rep["n_tried"] >= 10 and rep["ms_best"] <= rep["ms_default"] + 1e-9
        if prop == 1:
            assert rep["learned_min_hadrons"] in (1, 2, 3, 4, 6, 8, 12, 16, 65)  # 65: every pair routed
            assert ctx.learned_routing() == rep["learned_min_hadrons"]
            ctx.set_learned_routing(route0)
        ctx.sample(200, seed=21, proposal=prop)
        assert np.array_equal(ctx.hadrons(), ref[prop])
