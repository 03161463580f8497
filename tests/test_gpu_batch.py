"""Batched surfaces (iss_set_surfaces_lrf / iss_sample_batch): many surfaces
of uneven size, each with its own event count, sampled in rounds replayed
from one captured hipGraph.

Checks:
  * the captured-and-replayed rounds give bit-identical output to the same
    rounds launched one by one, and the same seed gives the same output;
  * the layout: segments tile the hadron list, event sizes add up to their
    segment, every surface gets exactly its events;
  * the law of every surface: hadrons per event against the surface's yields
    (Poisson mean and variance), species counts against its per-species yields;
  * one surface taken out of the batch against the same surface sampled alone
    (two-sample chi^2 of pT, y and species);
  * the batch autotuner returns a tuned round size per proposal that the
    proposal's batches then use;
  * decays of a batch's events equal the decays of the same list given as
    a plain hadron list, and the layout follows the decayed list.
"""
import numpy as np
import pytest
from scipy import stats

import iss_amd
from iss_amd import synth, tables

pytestmark = pytest.mark.gpu

SIZES = [300, 5000, 1200, 20000, 0, 800, 9000]
EVENTS = [3, 17, 1, 40, 5, 0, 12]


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


@pytest.fixture(scope="module")
def surfaces(species):
    out = []
    for k, n in enumerate(SIZES):
        if n == 0:
            out.append(np.zeros((0, 28), dtype=np.float32))
            continue
        c, _ = synth.viscous_lrf(n, species, seed=1000 + k, target_hadrons=0.05 * n)
        out.append(c)
    return out


def _ctx(species, surfaces):
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_surfaces_lrf(surfaces)
    return ctx


def test_graph_replay_equals_launches_and_is_deterministic(species, surfaces):
    ctx = _ctx(species, surfaces)
    n1 = ctx.sample_batch(EVENTS, seed=9, events_per_round=8, use_graph=1)
    h1 = ctx.hadrons()
    seg1, ev1 = ctx.batch_layout()
    n2 = ctx.sample_batch(EVENTS, seed=9, events_per_round=8, use_graph=0)
    h2 = ctx.hadrons()
    seg2, ev2 = ctx.batch_layout()
    n3 = ctx.sample_batch(EVENTS, seed=9, events_per_round=8, use_graph=1)
    h3 = ctx.hadrons()
    print("batch hadrons", n1, "rounds", seg1.shape[0], "ms", ctx.batch_ms())
    assert n1 == n2 == n3 > 0
    assert np.array_equal(h1, h2) and np.array_equal(h1, h3)
    assert np.array_equal(seg1, seg2) and np.array_equal(ev1, ev2)
    st = ctx.stats()
    assert st.n_failed == 0 and st.n_range_errors == 0
    ctx.close()


def test_layout_tiles_the_output(species, surfaces):
    ctx = _ctx(species, surfaces)
    R = 8
    n = ctx.sample_batch(EVENTS, seed=3, events_per_round=R)
    seg, ev = ctx.batch_layout()
    assert seg.shape == (int(np.ceil(max(EVENTS) / R)), len(SIZES), 2)
    flat = seg.reshape(-1, 2)
    assert flat[0, 0] == 0 and np.all(flat[1:, 0] == flat[:-1, 0] + flat[:-1, 1])
    assert flat[-1, 0] + flat[-1, 1] == n
    assert np.array_equal(ev.sum(axis=2), seg[:, :, 1])
    h = ctx.hadrons()
    assert np.all(h["pid"] != 0)
    E = h["E"].astype(np.float64)
    m2 = E ** 2 - h["px"].astype(np.float64) ** 2 - h["py"].astype(np.float64) ** 2 - h["pz"].astype(np.float64) ** 2
    # float E and p of hadrons at |y| up to 7: E^2 - p^2 holds ~1e-7 E^2
    assert np.all(np.abs(m2 - h["mass"].astype(np.float64) ** 2) < 1e-6 * E ** 2 + 1e-4)
    for k, E in enumerate(EVENTS):
        hk, sizes = ctx.batch_surface(k, hadrons=h, layout=(seg, ev))
        assert sizes.size == E and sizes.sum() == hk.size
        if SIZES[k] == 0:
            assert hk.size == 0
    ctx.close()


@pytest.mark.parametrize("precision", [32, 64])
def test_every_surface_follows_its_yields(species, surfaces, precision):
    """Rounds draw every slot's species from its cell's yield row in the
    working precision (FP32 rows, or FP64 rows with precision 64)."""
    ctx = _ctx(species, surfaces)
    _, per = ctx.compute_yields(precision=precision, per_cell=True)
    ev = np.array([400, 60, 1000, 30, 7, 0, 90], dtype=np.int32)
    n = ctx.sample_batch(ev, seed=21, events_per_round=16, precision=precision)
    h = ctx.hadrons()
    layout = ctx.batch_layout()
    off = np.concatenate([[0], np.cumsum(SIZES)])
    pids = species["pid"]
    chi2_tot, ndf_tot = 0.0, 0
    for k in range(len(SIZES)):
        hk, sizes = ctx.batch_surface(k, hadrons=h, layout=layout)
        lam_s = per[off[k]:off[k + 1]].astype(np.float64).sum(axis=0)
        lam = lam_s.sum()
        if ev[k] == 0 or lam == 0:
            assert hk.size == 0
            continue
        mean = sizes.mean()
        assert abs(mean - lam) < 4 * np.sqrt(lam / ev[k]) + 1e-9, (k, mean, lam)
        if ev[k] >= 60:
            var = sizes.var(ddof=1)
            # variance of a Poisson sample variance ~ lam / E * (2 lam + 1) (large lam)
            assert abs(var - lam) < 5 * np.sqrt((2 * lam * lam + lam) / ev[k]), (k, var, lam)
        exp = ev[k] * lam_s
        obs = np.array([np.count_nonzero(hk["pid"] == p) for p in pids])
        big = exp > 20
        chi2_tot += float(np.sum((obs[big] - exp[big]) ** 2 / exp[big]))
        ndf_tot += int(big.sum())
    p = stats.chi2.sf(chi2_tot, ndf_tot)
    print("species chi2/ndf", chi2_tot / ndf_tot, "ndf", ndf_tot, "p", p)
    assert p > 1e-3
    ctx.close()


def test_surface_in_batch_matches_surface_alone(species, surfaces):
    k = 3
    ctx = _ctx(species, surfaces)
    ev = np.zeros(len(SIZES), dtype=np.int32)
    ev[k] = 300
    ev[1] = 50
    ctx.sample_batch(ev, seed=5, events_per_round=32)
    hb, _ = ctx.batch_surface(k)
    alone = iss_amd.Context(0)
    alone.set_species(species)
    alone.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    alone.set_surface_lrf(surfaces[k])
    alone.sample(300, seed=6)
    ha = alone.hadrons()

    def chi2_two(a, b, edges):
        ca, _ = np.histogram(a, edges)
        cb, _ = np.histogram(b, edges)
        ok = (ca + cb) > 10
        ra, rb = np.sqrt(cb.sum() / ca.sum()), np.sqrt(ca.sum() / cb.sum())
        c2 = np.sum((ra * ca[ok] - rb * cb[ok]) ** 2 / (ca[ok] + cb[ok]))
        return c2, int(ok.sum()) - 1

    def pt(h):
        return np.hypot(h["px"].astype(np.float64), h["py"])

    def rap(h):
        E, pz = h["E"].astype(np.float64), h["pz"].astype(np.float64)
        return 0.5 * np.log((E + pz) / (E - pz))

    tot, ndf = 0.0, 0
    for a, b, edges in ((pt(hb), pt(ha), np.linspace(0, 3, 31)), (rap(hb), rap(ha), np.linspace(-7, 7, 29))):
        c2, d = chi2_two(a, b, edges)
        tot += c2
        ndf += d
    ua = np.unique(np.concatenate([hb["pid"], ha["pid"]]))
    cb_ = np.array([np.count_nonzero(hb["pid"] == p) for p in ua])
    ca_ = np.array([np.count_nonzero(ha["pid"] == p) for p in ua])
    ok = (ca_ + cb_) > 20
    rb_, ra_ = np.sqrt(ca_.sum() / cb_.sum()), np.sqrt(cb_.sum() / ca_.sum())
    tot += float(np.sum((rb_ * cb_[ok] - ra_ * ca_[ok]) ** 2 / (ca_[ok] + cb_[ok])))
    ndf += int(ok.sum()) - 1
    p = stats.chi2.sf(tot, ndf)
    print("batch vs alone chi2/ndf", tot / ndf, "p", p, "hadrons", hb.size, ha.size)
    assert p > 1e-3
    # same law of the event count: 300 events each
    assert abs(hb.size - ha.size) < 5 * np.sqrt(hb.size + ha.size)
    ctx.close()
    alone.close()


def test_autotuned_round_size_is_used(species, surfaces):
    """one tuned round size per proposal, each used by its own batches; the
    learned proposal also times its routing thresholds and keeps one"""
    ctx = _ctx(species, surfaces)
    cand = (4, 16, 64, 256, 1024, 4096, 16384)
    routes = (1, 2, 3, 4, 6, 8, 12, 16, 65)  # 65: every pair routed
    rep = ctx.autotune_batch(EVENTS, seed=2)
    ctx.mlp_init(1)
    rep_l = ctx.autotune_batch(EVENTS, seed=2, proposal=1)
    print("batch autotune", rep, "learned", rep_l)
    assert rep["n_tried"] == len(cand) and rep["learned_min_hadrons"] == 0
    assert rep_l["n_tried"] == len(cand) + len(routes)
    assert rep_l["learned_min_hadrons"] in routes and ctx.learned_routing() == rep_l["learned_min_hadrons"]
    for r in (rep, rep_l):
        assert r["refill_min"] in cand
    for proposal, r in ((0, rep), (1, rep_l)):
        ctx.sample_batch(EVENTS, seed=2, proposal=proposal)
        seg, ev = ctx.batch_layout()
        assert ev.shape[2] == r["refill_min"]
        assert seg.shape[0] == int(np.ceil(max(EVENTS) / r["refill_min"]))
    ctx.close()


def test_decays_of_a_batch(species, surfaces):
    """iss_decay after iss_sample_batch: the batch's events decay as one list
    (bit-identical to decaying the same hadrons and event sizes given through
    iss_set_hadrons with the same seed), and the batch layout is rewritten:
    segments tile the decayed list and every surface keeps its events.
    (Charges are not checked per event: a channel of more than three bodies
    leaves no daughters, as in the reference.)"""
    ctx = _ctx(species, surfaces)
    ctx.set_decay_table(tables.decay_table(1))
    R = 8
    n0 = ctx.sample_batch(EVENTS, seed=5, events_per_round=R)
    h0 = ctx.hadrons()
    seg0, ev0 = ctx.batch_layout()
    sizes0 = np.concatenate([ev0[r, k, :max(0, min(R, EVENTS[k] - r * R))]
                             for r in range(seg0.shape[0]) for k in range(len(EVENTS))])
    assert sizes0.sum() == n0
    n1 = ctx.decay(seed=23)
    h1 = ctx.hadrons()
    seg1, ev1 = ctx.batch_layout()
    print("batch decays", n0, "->", n1, "events", sizes0.size)
    assert n1 == h1.size > n0
    flat = seg1.reshape(-1, 2)
    assert flat[0, 0] == 0 and np.all(flat[1:, 0] == flat[:-1, 0] + flat[:-1, 1])
    assert flat[-1, 0] + flat[-1, 1] == n1
    assert np.array_equal(ev1.sum(axis=2), seg1[:, :, 1])
    sizes1 = np.concatenate([ev1[r, k, :max(0, min(R, EVENTS[k] - r * R))]
                             for r in range(seg1.shape[0]) for k in range(len(EVENTS))])
    for k, E in enumerate(EVENTS):
        hk, sk = ctx.batch_surface(k, hadrons=h1, layout=(seg1, ev1))
        assert sk.size == E and sk.sum() == hk.size
    # the same list decayed through iss_set_hadrons
    ctx2 = iss_amd.Context(0)
    ctx2.set_species(species)
    ctx2.set_decay_table(tables.decay_table(1))
    ctx2.set_hadrons(sizes0, h0)
    assert ctx2.decay(seed=23) == n1
    assert np.array_equal(ctx2.hadrons(), h1)
    assert np.array_equal(ctx2.event_counts(), sizes1)
    ctx2.close()
    ctx.close()
