"""GPU sampler vs the reference's integration criterion and the CPU oracle.

* The reference's one-cell tests (tests/RunAutoTest{Ideal1-4,Viscous1-2}.sh):
  mean |T^{mu nu}_cell - T^{mu nu}_samples| over the 16 components and the
  net-baryon row below 1e-3 (tests/TestOutputFiles.py:5-13).
* Statistical agreement with the oracle on a small viscous surface:
  per-species multiplicities within 3 sigma of E * N_s (chi^2 over species),
  pT / rapidity / v2 spectra chi^2/ndf <= 1.1 against the oracle's sample,
  KS p > 0.01 on pT.
"""
import numpy as np
import pytest
from scipy import stats as sst

import iss_amd
from iss_amd import synth, tables
from tests import oracle_binding as ob
from tests.sample_stats import GROUPS, group_chi2 as _gc

pytestmark = pytest.mark.gpu


def _gpu_ctx(afterburner, shear, bulk, bulk_kind):
    ctx = iss_amd.Context(0)
    sp = tables.species_arrays(afterburner)
    ctx.set_species(sp)
    kw = dict(shear=shear, bulk=bulk, bulk_kind=bulk_kind)
    if bulk_kind == 21:
        kw["neos"] = tables.read_neos(21, afterburner)
    if bulk and bulk_kind == 11:
        kw["k11"] = tables.read_kind11(afterburner)
    ctx.set_deltaf(**kw)
    return ctx, sp


def tmunu_from_hadrons(h, sp_all, nev, volume):
    p = np.stack([h["E"], h["px"], h["py"], h["pz"]]).astype(np.float64)
    w = 1.0 / p[0]
    T = np.array([[np.dot(p[i] * w, p[j]) for j in range(4)] for i in range(4)]) / (nev * volume)
    Bmap = dict(zip(sp_all["pid"].tolist(), sp_all["B"].tolist()))
    B = np.array([Bmap[x] for x in h["pid"].tolist()], dtype=np.float64) if h.size else np.zeros(0)
    return T.ravel(), B.sum() / (nev * volume)


@pytest.mark.parametrize("case", list(ob.REF_CASES))
def test_reference_tmunu_reconstruction(case):
    cfg = ob.case_config(case)
    ctx, sp = _gpu_ctx(cfg["afterburner"], cfg["shear"], cfg["bulk"], cfg["bulk_kind"])
    lab = tables.read_music_surface(cfg["surface"], bulk=1, rhob=cfg["rhob"])
    assert lab.shape[0] == 1
    ctx.set_surface_lab(lab, eos=tables.read_eos(cfg["iEOS"], cfg["afterburner"]))
    lrf = ctx.get_surface_lrf()
    volume = lrf[0, 4] / lrf[0, 8]
    nev = 1000
    n = ctx.sample(nev, seed=20240 + len(case))
    st = ctx.stats()
    assert st.n_failed == 0 and st.n_range_errors == 0
    h = ctx.hadrons()
    assert h.size == n and np.all(h["pid"] != 0)
    Ts, nBs = tmunu_from_hadrons(h, sp, nev, volume)
    orc = ob.Oracle(afterburner=cfg["afterburner"], shear=cfg["shear"], bulk=cfg["bulk"],
                    bulk_kind=cfg["bulk_kind"])
    orc.read_surface(cfg["surface"], bulk=1, rhob=cfg["rhob"], eos_file=cfg["eos_file"],
                     iEOS=cfg["iEOS"])
    Tf, Qf = orc.surface_tmunu()
    diff = np.concatenate([Tf - Ts, [Qf[0] - nBs]])
    print(case, "hadrons/event", n / nev, "mean|diff|", np.mean(np.abs(diff)))
    assert np.mean(np.abs(diff)) < 1e-3


VISCOUS_DF = dict(shear=1, bulk=1, bulk_kind=11)


@pytest.fixture(scope="module")
def viscous_oracle():
    """The oracle's sample of a small viscous surface (FP64 throughout)."""
    sp = tables.species_arrays(1)
    cells, _ = synth.viscous_lrf(400, sp, seed=31, target_hadrons=800.0)
    orc = ob.Oracle(afterburner=1, **VISCOUS_DF)
    orc.set_cells_lrf(cells)
    nev_o = 10000
    co, ho = orc.sample(99, nev_o)
    return dict(cells=cells, ho=ho, co=co, nev_o=nev_o)


@pytest.fixture(scope="module", params=[64, 32], ids=["fp64_yields", "fp32_yields"])
def viscous_pair(request, viscous_oracle):
    """GPU sample of the same surface against the oracle's: FP64 yields (the
    reference's arithmetic) and the production FP32 yields."""
    ctx, sp = _gpu_ctx(1, **VISCOUS_DF)
    ctx.set_surface_lrf(viscous_oracle["cells"])
    tot = ctx.compute_yields(precision=request.param)
    nev_g = 100000
    ctx.sample(nev_g, seed=77 + request.param, precision=request.param)
    hg = ctx.hadrons()
    cg = ctx.event_counts()
    return dict(sp=sp, tot=tot, hg=hg, cg=cg, nev_g=nev_g, **{k: viscous_oracle[k] for k in ("ho", "co", "nev_o")})


def test_species_multiplicities_within_3sigma(viscous_pair):
    d = viscous_pair
    sp, tot, hg, nev = d["sp"], d["tot"], d["hg"], d["nev_g"]
    idx = {p: i for i, p in enumerate(sp["pid"].tolist())}
    cnt = np.bincount([idx[p] for p in hg["pid"].tolist()], minlength=len(sp["pid"]))
    mu = nev * tot
    sel = mu > 0
    z = (cnt[sel] - mu[sel]) / np.sqrt(mu[sel])
    chi2 = np.sum(z ** 2)
    pval = sst.chi2.sf(chi2, sel.sum())
    print("species chi2/ndf", chi2 / sel.sum(), "p", pval, "max |z|", np.abs(z).max())
    assert pval > 1e-3
    assert np.mean(np.abs(z) > 3) < 0.02
    assert np.abs(z).max() < 5


def test_event_multiplicity_is_poisson(viscous_pair):
    d = viscous_pair
    c = d["cg"]
    lam = d["tot"].sum()
    assert abs(c.mean() - lam) < 4 * np.sqrt(lam / c.size)
    assert abs(c.var() / lam - 1) < 0.1


def test_spectra_chi2_and_ks_vs_oracle(viscous_pair):
    """dN/dy, pT, phi and v2(pT) of pions, kaons, protons and the rest:
    combined chi^2/ndf <= 1.1; each group within 4 sigma of chi^2/ndf = 1;
    KS p > 0.01 on every group's pT."""
    tot_chi, tot_ndf = 0.0, 0
    for name, pids in GROUPS.items():
        chi, ndf, ks, ng, no = _gc(viscous_pair, pids)
        print(name, "chi2/ndf %.3f ndf %d KS p %.3g N %d %d" % (chi / ndf, ndf, ks, ng, no))
        assert chi / ndf <= 1 + 4 * np.sqrt(2.0 / ndf)
        assert ks > 0.01
        tot_chi += chi
        tot_ndf += ndf
    print("combined chi2/ndf %.4f ndf %d" % (tot_chi / tot_ndf, tot_ndf))
    assert tot_chi / tot_ndf <= 1.1


def test_same_seed_same_output():
    ctx, sp = _gpu_ctx(1, 1, 1, 11)
    cells, _ = synth.viscous_lrf(300, sp, seed=41, target_hadrons=300.0)
    ctx.set_surface_lrf(cells)
    ctx.sample(50, seed=5)
    a = ctx.hadrons().copy()
    ca = ctx.event_counts().copy()
    ctx.sample(50, seed=5)
    b = ctx.hadrons()
    assert np.array_equal(a, b) and np.array_equal(ca, ctx.event_counts())
    ctx.sample(50, seed=6)
    assert not np.array_equal(a, ctx.hadrons()[: a.size]) or a.size != ctx.stats().n_hadrons


@pytest.mark.parametrize("n_events", [1, 3, 2048, 8192, 8193, 20000])
def test_event_counts_split_sizes(n_events):
    """The per-event split of the hadron total (up to 8192 events a uniform
    event per slot and histograms, beyond that a binomial tree, one launch
    per level): counts add up to the total, and are Poisson with the mean of
    the total per event (chi^2 of the count histogram, p > 1e-3)."""
    sp = tables.species_arrays(1)
    cells, _ = synth.viscous_lrf(300, sp, seed=8, target_hadrons=30.0)
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surface_lrf(cells)
    lam = ctx.compute_yields(precision=32).sum()
    n = ctx.sample(n_events, seed=21)
    c = ctx.event_counts()
    assert c.size == n_events and c.sum() == n
    if n_events >= 2048:
        kmax = int(c.max()) + 1
        obs = np.bincount(c, minlength=kmax + 1).astype(np.float64)
        pmf = sst.poisson.pmf(np.arange(kmax + 1), lam)
        pmf[-1] += sst.poisson.sf(kmax, lam)
        exp = n_events * pmf
        ok = exp > 10
        chi = float(np.sum((obs[ok] - exp[ok]) ** 2 / exp[ok]))
        pv = sst.chi2.sf(chi, int(ok.sum()) - 1)
        print(n_events, "mean", c.mean(), "lambda", lam, "chi2 p", pv)
        assert pv > 1e-3
    ctx.sample(n_events, seed=21)
    assert np.array_equal(ctx.event_counts(), c)
    ctx.close()


def test_event_placement_writes_every_slot_once():
    """Slots go to output positions through a keyed permutation of [0, K)
    (a mixed-radix Feistel network with cycle walking): for totals K from 1
    up to ~1e5, over many seeds (K squares, one off a square, primes and
    everything between occur), every output record is written exactly once,
    so none is left empty, and the event sizes add up to K."""
    sp = tables.species_arrays(1)
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    seen = set()
    for ncell, target, nev, seeds in ((1, 0.5, 4, range(40)), (3, 4.0, 7, range(40)),
                                      (50, 30.0, 31, range(10)), (2000, 200.0, 500, range(3))):
        cells, _ = synth.viscous_lrf(ncell, sp, seed=ncell, target_hadrons=target)
        ctx.set_surface_lrf(cells)
        for seed in seeds:
            n = ctx.sample(nev, seed=1000 + seed)
            st = ctx.stats()
            assert st.n_failed == 0 and st.n_range_errors == 0
            h = ctx.hadrons()
            assert h.size == n and ctx.event_counts().sum() == n
            # (an unwritten record is all zeros; a written one has pid != 0 and E > 0)
            assert np.all(h["pid"] != 0) and np.all(h["E"] > 0.)
            seen.add(n)
    print("totals sampled", len(seen), "from", min(seen), "to", max(seen))
    assert min(seen) <= 3 and max(seen) > 50000
    ctx.close()
