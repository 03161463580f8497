"""GPU resonance feed-down (iss_set_decay_table / iss_decay) against the CPU
oracle's restatement of the reference's decayer (itself pinned draw for draw
to the reference's particle_decay sources).  The random streams differ
(Philox per particle and generation against one mt19937), so parity is by law:

  * only stable particles come out, events keep their order and grouping, the
    same seed gives the same list;
  * energy, momentum, charge, baryon number and strangeness are conserved per
    event through whole decay chains (mothers whose chains are all two- and
    three-body);
  * the decayed species counts and the pi+-, p pT and y spectra of a sampled
    viscous surface agree with the oracle decaying an independent sample of
    the same surface (two-sample chi^2);
  * a channel closed at the pole mass is reported as an error, where the
    reference exits.
"""
import numpy as np
import pytest
from scipy import stats

import iss_amd
from iss_amd import synth, tables
from tests import oracle_binding as ob
from tests.test_decay_oracle import _closed, _hadrons

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


def _ctx(species, ab=1):
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1))
    ctx.set_decay_table(tables.decay_table(ab))
    return ctx


def _two_sample(a, b):
    ok = (a + b) > 10
    ra, rb = np.sqrt(b.sum() / a.sum()), np.sqrt(a.sum() / b.sum())
    return float(np.sum((ra * a[ok] - rb * b[ok]) ** 2 / (a[ok] + b[ok]))), int(ok.sum()) - 1


def test_conservation_grouping_and_determinism(species):
    t = ob.decay_table(1)
    ok = _closed(t) & (t["stable"] == 0)
    rng = np.random.default_rng(5)
    counts = rng.integers(1, 12, 300).astype(np.int64)
    h = _hadrons(t, rng.choice(np.nonzero(ok)[0], counts.sum()), rng)
    ctx = _ctx(species)
    ctx.set_hadrons(counts, h)
    n = ctx.decay(seed=17)
    st = ctx.stats()
    out = ctx.hadrons()
    ev = ctx.event_counts()
    print("decayed", counts.sum(), "->", n, "generations", st.n_decay_generations, "ms", st.ms_decay)
    assert out.size == n == ev.sum() and ev.size == counts.size
    stable = set(t["monval"][t["stable"] == 1].tolist())
    assert all(int(p) in stable for p in np.unique(out["pid"]))
    q = {int(m): (t["B"][i], t["S"][i], t["Q"][i]) for i, m in reversed(list(enumerate(t["monval"])))}
    oi = np.concatenate([[0], np.cumsum(counts)])
    oo = np.concatenate([[0], np.cumsum(ev)])
    for e in range(counts.size):
        hin, hout = h[oi[e]:oi[e + 1]], out[oo[e]:oo[e + 1]]
        for f in ("E", "px", "py", "pz"):
            tol = 3e-5 * max(1.0, float(hin["E"].astype(np.float64).sum())) * max(hout.size, 1)
            assert abs(hout[f].astype(np.float64).sum() - hin[f].astype(np.float64).sum()) < tol, (e, f)
        qi = np.array([q[int(p)] for p in hin["pid"]]).sum(axis=0)
        qo = np.array([q[int(p)] for p in hout["pid"]]).sum(axis=0)
        assert np.array_equal(qi, qo), e
        m2 = hout["E"].astype(np.float64) ** 2 - hout["px"] ** 2 - hout["py"] ** 2 - hout["pz"] ** 2
        assert np.all(np.abs(m2 - hout["mass"].astype(np.float64) ** 2) < 1e-4 * hout["E"] ** 2 + 1e-4)
    # same seed, same list
    ctx.set_hadrons(counts, h)
    ctx.decay(seed=17)
    assert np.array_equal(ctx.hadrons(), out)
    ctx.close()


def test_decayed_sample_matches_oracle_law(species):
    cells, _ = synth.viscous_lrf(20000, species, seed=31, target_hadrons=1000.0)
    ctx = _ctx(species)
    ctx.set_surface_lrf(cells)
    # two independent samples of the same surface: one decayed on the GPU,
    # the other by the oracle (decaying one list twice would correlate them)
    ctx.sample(60, seed=1008)
    hb = ctx.hadrons()
    cb0 = ctx.event_counts()
    ctx.sample(60, seed=8)
    h0 = ctx.hadrons()
    c0 = ctx.event_counts()
    n = ctx.decay(seed=9)
    hg = ctx.hadrons()
    cg = ctx.event_counts()
    assert cg.size == c0.size and cg.sum() == n
    co, ho = ob.decay(cb0, hb, 10)
    print("sampled", h0.size, "decayed GPU", hg.size, "oracle", ho.size)
    # species counts
    pids = np.union1d(np.unique(hg["pid"]), np.unique(ho["pid"]))
    a = np.array([np.count_nonzero(hg["pid"] == p) for p in pids], dtype=np.float64)
    b = np.array([np.count_nonzero(ho["pid"] == p) for p in pids], dtype=np.float64)
    c2, nd = _two_sample(a, b)
    # pi+ pT and rapidity
    for pid in (211, -211, 2212):
        sg, so = hg[hg["pid"] == pid], ho[ho["pid"] == pid]
        for f, edges in ((lambda x: np.hypot(x["px"].astype(np.float64), x["py"]), np.linspace(0, 3, 31)),
                         (lambda x: 0.5 * np.log((x["E"].astype(np.float64) + x["pz"]) /
                                                 (x["E"].astype(np.float64) - x["pz"])), np.linspace(-7, 7, 29))):
            ca, _ = np.histogram(f(sg), edges)
            cb, _ = np.histogram(f(so), edges)
            x, d = _two_sample(ca.astype(np.float64), cb.astype(np.float64))
            c2 += x
            nd += d
    p = stats.chi2.sf(c2, nd)
    print("GPU vs oracle decays: chi2/ndf", c2 / nd, "ndf", nd, "p", p)
    assert p > 1e-3
    assert abs(hg.size - ho.size) < 5 * np.sqrt(hg.size)
    ctx.close()


def test_closed_channel_is_an_error(species):
    """a mother lighter than its only channel's daughters: the reference
    exits (src/particle_decay.cpp:351-357); here the call reports it"""
    t = {"monval": [113, 2212, -2212], "mass": [0.775, 0.938, 0.938], "width": [0.149, 0.0, 0.0],
         "stable": [0, 1, 1], "first_channel": [0, 1, 2], "n_channels": [1, 1, 1], "npart": [2, 1, 1],
         "branching_ratio": [1.0, 1.0, 1.0],
         "daughters": [[2212, -2212, 0, 0, 0], [2212, 0, 0, 0, 0], [-2212, 0, 0, 0, 0]]}
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_decay_table(t)
    h = np.zeros(4, dtype=iss_amd.HADRON_DTYPE)
    h["pid"], h["mass"], h["E"], h["px"] = 113, 0.775, 1.0, 0.6
    ctx.set_hadrons(np.array([4]), h)
    with pytest.raises(iss_amd.ISSError, match="pole mass below"):
        ctx.decay(seed=1)
    assert ctx.event_counts().sum() == 0
    ctx.close()


def test_empty_list_and_stable_only(species):
    """no hadrons, and a list of stable hadrons only: nothing to decay"""
    ctx = _ctx(species)
    ctx.set_hadrons(np.array([0, 0, 0]), np.zeros(0, dtype=iss_amd.HADRON_DTYPE))
    assert ctx.decay(seed=1) == 0
    assert np.array_equal(ctx.event_counts(), [0, 0, 0])
    t = ob.decay_table(1)
    st = np.nonzero(t["stable"] == 1)[0]
    rng = np.random.default_rng(2)
    h = _hadrons(t, rng.choice(st, 30), rng)
    ctx.set_hadrons(np.array([10, 0, 20]), h)
    assert ctx.decay(seed=1) == 30
    assert np.array_equal(ctx.event_counts(), [10, 0, 20])
    assert np.array_equal(ctx.hadrons(), h)
    assert ctx.stats().n_decay_generations == 1
    ctx.close()
