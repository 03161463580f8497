"""dN_dy_sampling_model 1, 10 and 20 on the GPU (src/FSSW.cpp:250-309,
884-1070): per event and species the number of hadrons comes from the
species' yield dN (times y_RB - y_LB on boost-invariant surfaces), then each
hadron's cell is drawn from that species' cell yields and its momentum from
the cell, as in the Poisson pass.

  * model 1: every (event, species) count is floor(dN) or floor(dN) + 1, and
    the second with the probability dN - floor(dN) (binomial, 5 sigma);
  * model 20: the counts follow GSL's negative binomial with n = para1 dN and
    p = 1 / (1 + para1), whose mean n (1 - p) / p is para1^2 dN (chi^2
    against scipy's nbinom, which has the same parametrisation, p > 1e-3;
    mean within 5 sigma), for para1 = 1 and 0.5;
  * model 10: count - floor(dN) follows the negative binomial with
    n = para1 (dN - floor(dN));
  * events hold their hadrons in species order (the reference's species
    loop is the outer one); the cell of every hadron of a species follows
    the species' per-cell yields (chi^2, p > 1e-3); the pT spectrum of pi+
    follows the smooth spectrum (chi^2, p > 1e-3); the same seed gives the
    same list;
  * with local charge conservation every positive hadron is followed by its
    antiparticle and negative species are not counted;
  * unknown models, a non-positive para1 and the learned proposal are errors.
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


def _ctx(sp, cells):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surface_lrf(cells)
    return ctx


def _per_event_counts(h, counts, n_species, idx):
    """[E, Ns] hadron counts; also checks species order inside each event."""
    off = np.concatenate([[0], np.cumsum(counts)])
    sidx = np.array([idx[int(p)] for p in h["pid"]], dtype=np.int64) if h.size else np.zeros(0, np.int64)
    out = np.zeros((counts.size, n_species), dtype=np.int64)
    for e in range(counts.size):
        s = sidx[off[e]:off[e + 1]]
        assert np.all(np.diff(s) >= 0), e
        out[e] = np.bincount(s, minlength=n_species)
    return out


def _nb_chi2(k, n, p):
    """chi^2 p-value of integer samples k against nbinom(n, p), bins with >= 10 expected."""
    kmax = int(k.max()) + 1
    obs = np.bincount(k, minlength=kmax + 1).astype(np.float64)
    pmf = stats.nbinom.pmf(np.arange(kmax + 1), n, p)
    pmf[-1] += stats.nbinom.sf(kmax, n, p)
    exp = k.size * pmf
    # pool the tails into neighbours until every bin expects >= 10
    ob, ex, acc_o, acc_e = [], [], 0.0, 0.0
    for o, e in zip(obs, exp):
        acc_o += o
        acc_e += e
        if acc_e >= 10:
            ob.append(acc_o)
            ex.append(acc_e)
            acc_o = acc_e = 0.0
    if acc_e > 0 and ex:
        ob[-1] += acc_o
        ex[-1] += acc_e
    ob, ex = np.array(ob), np.array(ex)
    chi = float(np.sum((ob - ex) ** 2 / ex))
    return stats.chi2.sf(chi, max(ob.size - 1, 1)), ob.size


@pytest.mark.parametrize("hydro_mode", [2, 1])
def test_model1_counts_are_floor_or_floor_plus_one(species, hydro_mode):
    cells, _ = synth.viscous_lrf(3000, species, seed=17, target_hadrons=300.0)
    if hydro_mode == 1:
        cells = cells.copy()
        cells[:, 3] = 0.0
    ctx = _ctx(species, cells)
    tot = ctx.compute_yields(precision=32)
    n_ev = 400
    n = ctx.sample(n_ev, seed=3, hydro_mode=hydro_mode, y_LB=-0.5, y_RB=0.75, sampling_model=1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    assert n == h.size == counts.sum()
    assert np.all(h["pid"] != 0)
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    per = _per_event_counts(h, counts, tot.size, idx)
    dN = tot * (1.25 if hydro_mode == 1 else 1.0)
    fl = np.floor(dN)
    extra = per - fl[None, :]
    assert np.all((extra == 0) | (extra == 1))
    frac = dN - fl
    got = extra.sum(axis=0)
    sd = np.sqrt(n_ev * frac * (1 - frac))
    assert np.all(np.abs(got - n_ev * frac) <= 5 * sd + 1e-9), np.max(np.abs(got - n_ev * frac) - 5 * sd)
    # same seed, same list
    ctx.sample(n_ev, seed=3, hydro_mode=hydro_mode, y_LB=-0.5, y_RB=0.75, sampling_model=1)
    assert np.array_equal(ctx.hadrons(), h)
    ctx.close()


@pytest.mark.parametrize("para1", [1.0, 0.5])
def test_model20_counts_follow_the_negative_binomial(species, para1):
    cells, _ = synth.viscous_lrf(400, species, seed=23, target_hadrons=40.0)
    ctx = _ctx(species, cells)
    tot = ctx.compute_yields(precision=32)
    n_ev = 4000
    ctx.sample(n_ev, seed=11, sampling_model=20, sampling_para1=para1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    per = _per_event_counts(h, counts, tot.size, idx)
    p = 1.0 / (1.0 + para1)
    for pid in (211, -211, 321, 2212):
        s = idx[pid]
        nk = para1 * tot[s]
        mean = nk * (1 - p) / p
        var = nk * (1 - p) / p ** 2
        m = per[:, s].mean()
        assert abs(m - mean) < 5 * np.sqrt(var / n_ev), (pid, m, mean)
        pv, nb = _nb_chi2(per[:, s], nk, p)
        print(pid, "dN", tot[s], "mean", m, "expected", mean, "chi2 p", pv, "bins", nb)
        assert pv > 1e-3 and nb >= 3, pid
    ctx.close()


def test_model10_fraction_follows_the_negative_binomial(species):
    cells, _ = synth.viscous_lrf(400, species, seed=29, target_hadrons=40.0)
    ctx = _ctx(species, cells)
    tot = ctx.compute_yields(precision=32)
    n_ev = 4000
    para1 = 2.0
    ctx.sample(n_ev, seed=13, sampling_model=10, sampling_para1=para1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    per = _per_event_counts(h, counts, tot.size, idx)
    fl = np.floor(tot)
    assert np.all(per >= fl[None, :])
    p = 1.0 / (1.0 + para1)
    tested = 0
    for s in np.argsort(-(tot - fl))[:4]:
        nk = para1 * (tot[s] - fl[s])
        pv, nb = _nb_chi2(per[:, s] - int(fl[s]), nk, p)
        print(int(species["pid"][s]), "dN", tot[s], "chi2 p", pv, "bins", nb)
        if nb >= 3:
            assert pv > 1e-3, int(species["pid"][s])
            tested += 1
    assert tested >= 2
    ctx.close()


def test_cells_and_spectra_follow_the_yields(species):
    n_cells = 60
    cells, _ = synth.viscous_lrf(n_cells, species, seed=31, target_hadrons=30.0)
    # small delta f, where the spectra's delta f restriction and the
    # sampler's agree
    cells = cells.copy()
    epp = cells[:, 12] + cells[:, 14]
    cells[:, 20:25] *= 0.05
    cells[:, 19] = np.maximum(cells[:, 19], -0.01 * epp)
    ctx = _ctx(species, cells)
    tot, pc = ctx.compute_yields(precision=64, per_cell=True)
    n_ev = 3000
    ctx.sample(n_ev, seed=7, sampling_model=20, sampling_para1=1.0)
    h = ctx.hadrons()
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    key = {(np.float32(cells[c, 1]).tobytes(), np.float32(cells[c, 2]).tobytes()): c for c in range(n_cells)}
    assert len(key) == n_cells
    chi, ndf = 0.0, 0
    for pid in (211, 321, 2212):
        sel = h[h["pid"] == pid]
        cidx = np.array([key[(x.tobytes(), y.tobytes())] for x, y in zip(sel["x"], sel["y"])])
        obs = np.bincount(cidx, minlength=n_cells).astype(np.float64)
        w = pc[:, idx[pid]]
        exp = sel.size * w / w.sum()
        ok = exp > 10
        chi += float(np.sum((obs[ok] - exp[ok]) ** 2 / exp[ok]))
        ndf += int(ok.sum()) - 1
    pv = stats.chi2.sf(chi, ndf)
    print("cells chi2/ndf", chi / ndf, "ndf", ndf, "p", pv)
    assert ndf > 60 and pv > 1e-3
    # pi+ pT spectrum against the smooth spectrum
    pt_edges = np.linspace(0.0, 2.0, 21)
    xg, wg = np.polynomial.legendre.leggauss(6)
    pts = (0.5 * (pt_edges[1:] + pt_edges[:-1])[:, None] + 0.5 * np.diff(pt_edges)[:, None] * xg[None]).ravel()
    wpt = (0.5 * np.diff(pt_edges)[:, None] * wg[None]).ravel()
    phs = np.linspace(0.0, 2 * np.pi, 17)[:-1] + np.pi / 16
    eta = np.arange(-10.0, 10.0 + 1e-9, 0.05)
    sel = h["pid"] == 211
    pt = np.hypot(h["px"][sel].astype(np.float64), h["py"][sel].astype(np.float64))
    obs, _ = np.histogram(pt, bins=pt_edges)
    dn = ctx.spectra(np.array([idx[211]], dtype=np.int32), pts, phs, eta, np.full(eta.size, 0.05),
                     precision=64, all_cells=0)[0]
    dens = (dn.mean(axis=1) * 2 * np.pi * pts * wpt).reshape(pt_edges.size - 1, 6).sum(axis=1)
    exp = obs.sum() * dens / dens.sum()   # the shape inside 0 < pT < 2 GeV
    ok = exp > 20
    c2 = float(np.sum((obs[ok] - exp[ok]) ** 2 / exp[ok]))
    pv = stats.chi2.sf(c2, int(ok.sum()) - 1)
    print("pi+ pT chi2/ndf", c2 / (ok.sum() - 1), "p", pv)
    assert ok.sum() > 10 and pv > 1e-3
    ctx.close()


def test_model1_with_local_charge_conservation(species):
    cells, _ = synth.viscous_lrf(2000, species, seed=37, target_hadrons=200.0)
    ctx = _ctx(species, cells)
    n_ev = 100
    n = ctx.sample(n_ev, seed=5, sampling_model=1, local_charge_conservation=1)
    h = ctx.hadrons()
    counts = ctx.event_counts()
    assert n == h.size == counts.sum() > 0
    tot = ctx.compute_yields(precision=32)
    q = dict(zip(species["pid"].tolist(), species["Q"].tolist()))
    idx = {int(p): i for i, p in enumerate(species["pid"])}
    off = np.concatenate([[0], np.cumsum(counts)])
    fl = np.floor(tot)
    for e in range(n_ev):
        ev = h[off[e]:off[e + 1]]
        qs = np.array([q[int(p)] for p in ev["pid"]])
        assert qs.sum() == 0
        k = 0
        parents = np.zeros(tot.size, dtype=np.int64)
        while k < ev.size:
            if qs[k] > 0:
                a, b = ev[k], ev[k + 1]
                assert b["pid"] == -a["pid"]
                for f in ("x", "y", "t", "z"):
                    assert b[f] == a[f]
                parents[idx[int(a["pid"])]] += 1
                k += 2
            else:
                assert qs[k] == 0
                parents[idx[int(ev["pid"][k])]] += 1
                k += 1
        # counts of the sampled (non-negative) species: floor or floor + 1
        pos = np.array([q[int(p)] >= 0 for p in species["pid"]])
        d = parents[pos] - fl[pos]
        assert np.all((d == 0) | (d == 1)), e
    ctx.close()


def test_model_errors(species):
    cells, _ = synth.viscous_lrf(200, species, seed=3, target_hadrons=20.0)
    ctx = _ctx(species, cells)
    with pytest.raises(iss_amd.ISSError, match="dN_dy_sampling_model"):
        ctx.sample(2, seed=1, sampling_model=5)
    with pytest.raises(iss_amd.ISSError, match="para1"):
        ctx.sample(2, seed=1, sampling_model=20, sampling_para1=0.0)
    ctx.mlp_init(seed=1)
    with pytest.raises(iss_amd.ISSError, match="learned proposal"):
        ctx.sample(2, seed=1, sampling_model=10, proposal=1)
    # an empty surface gives empty events
    ctx.set_surface_lrf(np.zeros((0, 28), dtype=np.float32))
    assert ctx.sample(3, seed=1, sampling_model=1) == 0
    assert np.all(ctx.event_counts() == 0)
    ctx.close()
