"""The benchmark configurations at full size on the GPU.

* C1, the plain ideal boost-invariant surface (tau = 10 fm, 40 x 50 cells of
  0.3 fm, radial flow v = 0.6 tanh(r / 7 fm), T = 0.150 GeV, mu = 0), read as
  lab-frame cells, regulated and boosted on the GPU, FP64, 100+ events with
  rapidity drawn in [-5, 5] (src/FSSW.cpp:950-958, 1024-1029): every
  (cell, species) yield within 1e-6 of the oracle; pT / y / phi / v2 against
  the oracle's sample at chi^2/ndf <= 1.1 with KS p > 0.01; dN/dy flat over the
  window; per-species multiplicities against E (y_RB - y_LB) N_s.
* C3, 256 batched surfaces of 1e3 - 5e5 cells (log-normal, median 3e4) with
  10 - 2000 events each: every surface's event sizes against its yields
  (Poisson mean and variance), from the batch layout; with the events cut
  a hundredfold, every surface's species counts against its yields.
* C5, the 2,000,000-cell viscous surface with mu_B up to 0.45 GeV: FP32 yields
  within 1e-5 of FP64 on every (cell, species); 5,000 events of ~1e4 hadrons
  with Poisson event sizes, species counts against the yields, no failed or
  out-of-range slot; the learned proposal on the same pass draws no momentum
  above its envelope and keeps the species and event sizes.
"""
import numpy as np
import pytest
from scipy import stats as sst

import iss_amd
from iss_amd import synth, tables
from tests import oracle_binding as ob
from tests.sample_stats import GROUPS, group_chi2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


def _species_counts(h, sp):
    pids = sp["pid"]
    order = np.argsort(pids)
    pos = np.searchsorted(pids[order], h["pid"])
    pos = np.minimum(pos, pids.size - 1)
    assert np.all(pids[order][pos] == h["pid"]), "hadron of a species outside the table"
    return np.bincount(order[pos], minlength=pids.size)


def _species_chi2(h, sp, mu):
    cnt = _species_counts(h, sp)
    sel = mu > 5
    z = (cnt[sel] - mu[sel]) / np.sqrt(mu[sel])
    chi2 = float(np.sum(z ** 2))
    # counts of species with tiny means go into one bin
    rest_o, rest_e = cnt[~sel].sum(), mu[~sel].sum()
    if rest_e > 0:
        chi2 += (rest_o - rest_e) ** 2 / rest_e
    ndf = int(sel.sum()) + (1 if rest_e > 0 else 0)
    return chi2, ndf, np.abs(z).max() if z.size else 0.0


# ------------------------------------------------------------------- C1
@pytest.fixture(scope="module")
def c1(species):
    lab = synth.ideal_binv_lab(species)
    assert lab.shape == (2000, 34)
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_deltaf(shear=0, bulk=0, bulk_kind=1)
    ctx.set_surface_lab(lab)
    return ctx, lab


def test_c1_fp64_yields_every_cell(c1, species):
    ctx, lab = c1
    assert ctx.num_cells() == 2000
    tot, per = ctx.compute_yields(precision=64, per_cell=True, hydro_mode=1)
    orc = ob.Oracle(afterburner=1, hydro_mode=1)
    orc.set_cells_lab(lab)
    ref = orc.yields()
    assert per.shape == ref.shape == (2000, len(species["pid"]))
    nz = ref > 0
    assert np.all(per[~nz] == 0)
    rel = np.abs(per[nz] - ref[nz]) / ref[nz]
    print("C1 FP64 max rel", rel.max(), "sum u.dsigma", ctx.get_surface_lrf()[:, 4].sum())
    assert rel.max() < 1e-6
    # dN/dy of the whole surface (the estimate: ~540 hadrons per unit rapidity)
    assert 300 < tot.sum() < 800


def test_c1_sampled_spectra_match_oracle(c1, species):
    ctx, lab = c1
    tot = ctx.compute_yields(precision=64, hydro_mode=1)
    nev_g = 2000
    n = ctx.sample(nev_g, seed=101, hydro_mode=1, y_LB=-5.0, y_RB=5.0, precision=64)
    st = ctx.stats()
    assert st.n_failed == 0 and st.n_range_errors == 0
    hg = ctx.hadrons()
    counts = ctx.event_counts()
    assert counts.sum() == n and hg.size == n
    lam = 10.0 * tot.sum()
    print("C1 hadrons/event", n / nev_g, "expected", lam)
    assert abs(counts.mean() - lam) < 4 * np.sqrt(lam / nev_g)
    assert abs(counts.var(ddof=1) / lam - 1) < 5 * np.sqrt(2.0 / nev_g)
    # species
    chi2, ndf, zmax = _species_chi2(hg, species, nev_g * 10.0 * tot)
    print("C1 species chi2/ndf", chi2 / ndf, "max |z|", zmax)
    assert sst.chi2.sf(chi2, ndf) > 1e-3 and zmax < 5
    # boost invariance: rapidity uniform on [-5, 5]
    pT = np.hypot(hg["px"], hg["py"]).astype(np.float64)
    mT = np.sqrt(pT ** 2 + hg["mass"].astype(np.float64) ** 2)
    y = 0.5 * np.log((hg["E"].astype(np.float64) + hg["pz"]) / (hg["E"].astype(np.float64) - hg["pz"]))
    assert np.all(np.abs(y) < 5.0 + 1e-3)
    hy, _ = np.histogram(y, bins=20, range=(-5, 5))
    chi_y = np.sum((hy - hy.mean()) ** 2 / hy.mean())
    print("C1 dN/dy flatness chi2/ndf", chi_y / 19)
    assert sst.chi2.sf(chi_y, 19) > 1e-3
    # spectra against the oracle's sample of the same surface
    orc = ob.Oracle(afterburner=1, hydro_mode=1, y_LB=-5.0, y_RB=5.0)
    orc.set_cells_lab(lab)
    nev_o = 150
    co, ho = orc.sample(202, nev_o)
    d = dict(hg=hg, ho=ho, nev_g=nev_g, nev_o=nev_o)
    tot_chi, tot_ndf = 0.0, 0
    for name, pids in GROUPS.items():
        chi, nd, ks, ng, no = group_chi2(d, pids)
        print("C1", name, "chi2/ndf %.3f ndf %d KS p %.3g N %d %d" % (chi / nd, nd, ks, ng, no))
        assert ks > 0.01
        tot_chi += chi
        tot_ndf += nd
    print("C1 combined chi2/ndf %.4f ndf %d" % (tot_chi / tot_ndf, tot_ndf))
    assert tot_chi / tot_ndf <= 1.1


# ------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3(species):
    surfaces, ev = synth.batched_surfaces(species)
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surfaces_lrf(surfaces)
    _, per = ctx.compute_yields(precision=32, per_cell=True)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in surfaces])])
    lam_s = np.stack([per[off[k]:off[k + 1]].astype(np.float64).sum(axis=0) for k in range(len(surfaces))])
    del per
    return ctx, surfaces, ev, lam_s


def test_c3_full_batch_event_sizes_follow_every_surface(c3):
    ctx, surfaces, ev, lam_s = c3
    assert len(surfaces) == 256
    ncells = np.array([s.shape[0] for s in surfaces])
    assert ncells.min() >= 1000 and ncells.max() <= 500000 and ev.min() >= 10 and ev.max() <= 2000
    n = ctx.sample_batch(ev, seed=33)
    st = ctx.stats()
    assert st.n_failed == 0 and st.n_range_errors == 0
    seg, evs = ctx.batch_layout()
    R = evs.shape[2]
    print("C3 hadrons", n, "rounds", seg.shape[0], "R", R, "ms", ctx.batch_ms())
    assert seg.reshape(-1, 2)[:, 1].sum() == n
    zm, zv = [], []
    for k in range(len(surfaces)):
        # rounds before a surface's last one carry R of its events each
        sizes = evs[:, k, :].ravel()[: int(ev[k])]
        assert sizes.size == ev[k] and sizes.sum() == seg[:, k, 1].sum()
        lam = lam_s[k].sum()
        zm.append((sizes.mean() - lam) / np.sqrt(lam / ev[k]))
        if ev[k] >= 100:
            # a Poisson sample variance has variance ~ lam (2 lam + 1) / (E - 1)
            zv.append((sizes.var(ddof=1) - lam) / np.sqrt(lam * (2 * lam + 1) / (ev[k] - 1)))
    zm, zv = np.array(zm), np.array(zv)
    print("C3 per-surface mean z: chi2/ndf %.3f max %.2f; variance z: chi2/ndf %.3f max %.2f"
          % ((zm ** 2).mean(), np.abs(zm).max(), (zv ** 2).mean(), np.abs(zv).max()))
    assert sst.chi2.sf((zm ** 2).sum(), zm.size) > 1e-3 and np.abs(zm).max() < 5
    assert sst.chi2.sf((zv ** 2).sum(), zv.size) > 1e-4 and np.abs(zv).max() < 5.5


def test_c3_species_of_every_surface(c3, species):
    ctx, surfaces, ev, lam_s = c3
    ev_small = np.maximum(1, ev // 100).astype(np.int32)
    ctx.sample_batch(ev_small, seed=34)
    h = ctx.hadrons()
    layout = ctx.batch_layout()
    chi_tot, ndf_tot, zmax = 0.0, 0, 0.0
    for k in range(len(surfaces)):
        hk, sizes = ctx.batch_surface(k, hadrons=h, layout=layout)
        assert sizes.size == ev_small[k] and sizes.sum() == hk.size
        chi, nd, zm = _species_chi2(hk, species, ev_small[k] * lam_s[k])
        chi_tot += chi
        ndf_tot += nd
        zmax = max(zmax, zm)
    print("C3 species chi2/ndf %.4f ndf %d max |z| %.2f" % (chi_tot / ndf_tot, ndf_tot, zmax))
    assert sst.chi2.sf(chi_tot, ndf_tot) > 1e-3 and zmax < 5.5


# ------------------------------------------------------------------- C5
@pytest.fixture(scope="module")
def c5(species):
    cells, _ = synth.viscous_lrf(2_000_000, species, seed=55, muB_max=0.45, target_hadrons=1.0e4)
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surface_lrf(cells)
    return ctx


def test_c5_fp32_yields_within_1e5_of_fp64_everywhere(c5):
    ctx = c5
    t64, p64 = ctx.compute_yields(precision=64, per_cell=True)
    t32, p32 = ctx.compute_yields(precision=32, per_cell=True)
    relt = np.abs(t32 - t64) / np.maximum(t64, 1e-300)
    worst, where, zeros_ok = 0.0, None, True
    for a in range(0, p64.shape[0], 250_000):
        b64 = p64[a:a + 250_000]
        b32 = p32[a:a + 250_000].astype(np.float64)
        rel = np.where(b64 > 1e-25, np.abs(b32 - b64) / np.maximum(b64, 1e-300), 0.0)
        k = int(np.argmax(rel))
        if rel.flat[k] > worst:
            worst, where = float(rel.flat[k]), (a + k // rel.shape[1], k % rel.shape[1])
        zeros_ok &= bool(np.all(b32[b64 == 0] == 0))
    print("C5 FP32 vs FP64: max rel per (cell, species) %.3g at (cell, species) %s, species totals %.3g, "
          "hadrons/event %.1f" % (worst, where, relt.max(), t64.sum()))
    assert worst < 1e-5 and relt.max() < 1e-5 and zeros_ok
    assert 5e3 < t64.sum() < 2e4


def test_c5_sampling_pass(c5, species):
    ctx = c5
    tot = ctx.compute_yields(precision=32)
    E = 5000
    n = ctx.sample(E, seed=505)
    st = ctx.stats()
    print("C5 plain: hadrons", n, "per event", n / E, "sample ms", st.ms_sample)
    assert st.n_failed == 0 and st.n_range_errors == 0
    counts = ctx.event_counts()
    lam = tot.sum()
    assert counts.sum() == n
    assert abs(counts.mean() - lam) < 4 * np.sqrt(lam / E)
    assert abs(counts.var(ddof=1) / lam - 1) < 5 * np.sqrt(2.0 / E)
    h = ctx.hadrons()
    assert np.all(h["pid"] != 0)
    chi2, ndf, zmax = _species_chi2(h, species, E * tot)
    print("C5 species chi2/ndf", chi2 / ndf, "max |z|", zmax)
    assert sst.chi2.sf(chi2, ndf) > 1e-3 and zmax < 5
    cnt_plain = _species_counts(h, species)
    del h
    # the learned proposal on the same pass: exact envelope, same species and events
    ctx.mlp_init(7)
    n2 = ctx.sample(E, seed=505, proposal=1)
    st2 = ctx.stats()
    print("C5 learned: proposals/hadron", st2.n_inner / max(n2, 1), "fallback", st2.n_fallback,
          "violations", st2.n_violations, "sample ms", st2.ms_sample)
    assert n2 == n and st2.n_violations == 0 and st2.n_failed == 0 and st2.n_range_errors == 0
    assert np.array_equal(ctx.event_counts(), counts)
    assert np.array_equal(_species_counts(ctx.hadrons(), species), cnt_plain)
