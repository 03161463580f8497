"""Local charge conservation (local_charge_conservation = 1,
src/FSSW.cpp:931-938, 1035-1048) on the GPU: negative species are not
sampled, and every positive hadron is followed in its event by its
antiparticle, sampled from the same cell with the negated charges.

  * structure: every positive hadron is followed by its partner (pid
    negated, same cell position, same t and z: same eta_s, also on
    boost-invariant surfaces where eta_s is drawn), negative hadrons occur
    only as partners, every event has zero net charge, the same seed gives
    the same list;
  * law: the positive species' counts against their yields (Poisson, 5
    sigma); the pi+ parents' and the pi- partners' (pT, phi) histograms
    against the smooth spectra of pi+ and pi- (equal per-cell yields at
    mu_Q = mu_S = 0), chi^2 p > 1e-3;
  * a species table without a positive species' antiparticle is an error.
"""
import numpy as np
import pytest
from scipy import stats

import iss_amd
from iss_amd import synth, tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


def _ctx(sp):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    return ctx


def _check_structure(h, counts, sp):
    q = dict(zip(sp["pid"].tolist(), sp["Q"].tolist()))
    off = np.concatenate([[0], np.cumsum(counts)])
    npos = 0
    for e in range(counts.size):
        ev = h[off[e]:off[e + 1]]
        qs = np.array([q[int(p)] for p in ev["pid"]])
        assert qs.sum() == 0, e
        k = 0
        while k < ev.size:
            if qs[k] > 0:
                npos += 1
                assert k + 1 < ev.size, e
                a, b = ev[k], ev[k + 1]
                assert b["pid"] == -a["pid"] and b["mass"] == a["mass"]
                for f in ("x", "y", "t", "z"):
                    assert b[f] == a[f], (e, k, f)
                k += 2
            else:
                assert qs[k] == 0, (e, k, int(ev["pid"][k]))
                k += 1
    return npos


@pytest.mark.parametrize("hydro_mode", [2, 1])
def test_partners_follow_parents_and_events_are_neutral(species, hydro_mode):
    cells, _ = synth.viscous_lrf(5000, species, seed=41, target_hadrons=500.0)
    if hydro_mode == 1:
        cells = cells.copy()
        cells[:, 3] = 0.0          # boost-invariant surfaces sit at eta = 0
    ctx = _ctx(species)
    ctx.set_surface_lrf(cells)
    n = ctx.sample(40, seed=5, hydro_mode=hydro_mode, local_charge_conservation=1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    assert n == h.size == counts.sum()
    npos = _check_structure(h, counts, species)
    print("hadrons", n, "positive parents", npos)
    assert npos > 1000
    for f in ("E", "px", "py", "pz", "t", "x", "y", "z"):
        assert np.all(np.isfinite(h[f])), f
    E = h["E"].astype(np.float64)
    m2 = E ** 2 - h["px"].astype(np.float64) ** 2 - h["py"].astype(np.float64) ** 2 - h["pz"].astype(np.float64) ** 2
    bad = np.abs(m2 - h["mass"].astype(np.float64) ** 2) >= 1e-6 * E ** 2 + 1e-4
    if bad.any():
        print("off shell", int(bad.sum()), h[np.nonzero(bad)[0][:5]], np.nonzero(bad)[0][:5])
    assert not bad.any()
    ctx.sample(40, seed=5, hydro_mode=hydro_mode, local_charge_conservation=1)
    assert np.array_equal(ctx.hadrons(), h)
    ctx.close()


def _hist_chi2(obs, dens, n_scale):
    exp = n_scale * dens
    ok = exp > 20
    return float(np.sum((obs[ok] - exp[ok]) ** 2 / exp[ok])), int(ok.sum())


def test_law_of_parents_and_partners(species):
    cells, _ = synth.viscous_lrf(20000, species, seed=21, muB_max=0.2, target_hadrons=4000.0)
    cells = cells.copy()
    epp = cells[:, 12] + cells[:, 14]
    cells[:, 20:25] *= 0.05
    cells[:, 19] = np.maximum(cells[:, 19], -0.01 * epp)
    ctx = _ctx(species)
    ctx.set_surface_lrf(cells)
    tot, _ = ctx.compute_yields(precision=64, per_cell=True)
    n_ev = 300
    ctx.sample(n_ev, seed=91, local_charge_conservation=1)
    h = ctx.hadrons()
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    # positive species' counts against their yields
    for pid in (211, 321, 2212):
        k = np.count_nonzero(h["pid"] == pid)
        mu = n_ev * tot[idx[pid]]
        assert abs(k - mu) < 5 * np.sqrt(mu) + 1, (pid, k, mu)
        assert np.count_nonzero(h["pid"] == -pid) == k
    # (pT, phi) shapes: pi+ parents and pi- partners against the smooth spectra
    pt_edges = np.linspace(0.0, 2.0, 21)
    ph_edges = np.linspace(0.0, 2 * np.pi, 9)
    xg, wg = np.polynomial.legendre.leggauss(6)
    pts = (0.5 * (pt_edges[1:] + pt_edges[:-1])[:, None] + 0.5 * np.diff(pt_edges)[:, None] * xg[None]).ravel()
    wpt = (0.5 * np.diff(pt_edges)[:, None] * wg[None]).ravel()
    phs = (0.5 * (ph_edges[1:] + ph_edges[:-1])[:, None] + 0.5 * np.diff(ph_edges)[:, None] * xg[None]).ravel()
    wph = (0.5 * np.diff(ph_edges)[:, None] * wg[None]).ravel()
    eta = np.arange(-10.0, 10.0 + 1e-9, 0.05)
    chi, ndf = 0.0, 0
    for pid in (211, -211):
        sel = h["pid"] == pid
        px, py = h["px"][sel].astype(np.float64), h["py"][sel].astype(np.float64)
        obs, _, _ = np.histogram2d(np.hypot(px, py), np.mod(np.arctan2(py, px), 2 * np.pi),
                                   bins=[pt_edges, ph_edges])
        dn = ctx.spectra(np.array([idx[pid]], dtype=np.int32), pts, phs, eta, np.full(eta.size, 0.05),
                         precision=64, all_cells=0)[0]
        dens = (dn * (pts * wpt)[:, None] * wph[None]).reshape(pt_edges.size - 1, 6, ph_edges.size - 1, 6).sum(
            axis=(1, 3))
        c2, d = _hist_chi2(obs, dens, n_ev)
        print(pid, "sampled", int(sel.sum()), "chi2/ndf", c2 / d)
        chi += c2
        ndf += d
    p = stats.chi2.sf(chi, ndf)
    print("chi2/ndf", chi / ndf, "ndf", ndf, "p", p)
    assert ndf > 200 and p > 1e-3
    ctx.close()


def test_missing_antiparticle_is_an_error(species):
    keep = species["pid"] != -211
    sp = {k: np.asarray(v)[keep] for k, v in species.items()}
    cells, _ = synth.viscous_lrf(500, species, seed=3, target_hadrons=50.0)
    ctx = _ctx(sp)
    ctx.set_surface_lrf(cells)
    with pytest.raises(iss_amd.ISSError, match="antiparticle of pid 211"):
        ctx.sample(2, seed=1, local_charge_conservation=1)
    assert ctx.sample(2, seed=1) > 0
    ctx.close()


def test_empty_surface_and_single_cell_with_lcc(species):
    ctx = _ctx(species)
    ctx.set_surface_lrf(np.zeros((0, 28), dtype=np.float32))
    assert ctx.sample(4, seed=1, local_charge_conservation=1) == 0
    assert np.all(ctx.event_counts() == 0)
    cells, _ = synth.viscous_lrf(1, species, seed=9, target_hadrons=80.0)
    ctx.set_surface_lrf(cells)
    n = ctx.sample(50, seed=2, local_charge_conservation=1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    assert n == h.size == counts.sum() > 0
    _check_structure(h, counts, species)
    # the switch also turns back off: negative species are sampled again
    ctx.sample(50, seed=2)
    q = dict(zip(species["pid"].tolist(), species["Q"].tolist()))
    assert any(q[int(p)] < 0 for p in ctx.hadrons()["pid"])
    ctx.close()
