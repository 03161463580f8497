"""The learned proposal in batched rounds (iss_sample_batch, proposal 1).

Under both proposals the slots of a round draw their species from the same
stream, so a learned batch must reproduce the plain batch of the same seed
exactly in everything but the momenta: hadron count, segments, event sizes
and the species of every slot.  The momenta follow the same law: two-sample
chi^2 of pT, y, phi and v2(pT) per species group, learned against plain, on
the large surfaces.  The envelope is exact (no draw above it), replayed
rounds equal launched rounds, and the seed fixes the output.
"""
import numpy as np
import pytest
from scipy import stats

import iss_amd
from iss_amd import synth, tables
from tests import sample_stats as ss

pytestmark = pytest.mark.gpu

SIZES = [300, 5000, 1200, 20000, 0, 800, 9000]
EVENTS = [30, 160, 10, 300, 5, 0, 120]


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


@pytest.fixture(scope="module")
def ctx(species):
    surfaces = []
    for k, n in enumerate(SIZES):
        if n == 0:
            surfaces.append(np.zeros((0, 28), dtype=np.float32))
            continue
        c, _ = synth.viscous_lrf(n, species, seed=2000 + k, target_hadrons=0.05 * n)
        surfaces.append(c)
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    # (without the reference's Fermi-Dirac table quirk, whose fermions the
    # learned sampler hands to the plain one)
    ctx.set_deltaf(shear=1, bulk=1, bulk_kind=11, k11=tables.read_kind11(1), fermion_cdf2_quirk=0)
    ctx.set_surfaces_lrf(surfaces)
    # the network, trained on a plain pass over the whole record
    ctx.sample(40, seed=5)
    assert ctx.mlp_collect(2_000_000) > 0
    ctx.mlp_init(1)
    ctx.mlp_train(batch=8192, steps=60, lr=3e-3, seed=5)
    yield ctx
    ctx.close()


def _batch(ctx, proposal, seed=11, use_graph=1):
    n = ctx.sample_batch(EVENTS, seed=seed, events_per_round=16, use_graph=use_graph, proposal=proposal)
    h = ctx.hadrons()
    seg, ev = ctx.batch_layout()
    return n, h, seg, ev, ctx.stats()


def test_learned_batch_matches_plain_slot_by_slot(ctx):
    n0, h0, seg0, ev0, st0 = _batch(ctx, 0)
    n1, h1, seg1, ev1, st1 = _batch(ctx, 1)
    print("batch hadrons", n0, n1, "learned fallback", st1.n_fallback, "violations", st1.n_violations)
    assert n0 == n1 > 0
    assert np.array_equal(seg0, seg1) and np.array_equal(ev0, ev1)
    assert np.array_equal(h0["pid"], h1["pid"])
    assert st0.n_failed == 0 and st1.n_failed == 0
    assert st1.n_violations == 0
    assert st1.n_fallback < 0.01 * n1
    # the momenta differ (independent draws), the positions of the cells do not
    assert not np.array_equal(h0["px"], h1["px"])
    assert np.array_equal(h0["x"], h1["x"]) and np.array_equal(h0["t"], h1["t"])
    # on shell to the FP32 rounding of E and p (large E at large rapidity)
    m = h1["mass"].astype(np.float64)
    E = h1["E"].astype(np.float64)
    e2 = E ** 2 - (h1["px"].astype(np.float64) ** 2 + h1["py"].astype(np.float64) ** 2 +
                   h1["pz"].astype(np.float64) ** 2)
    assert np.all(np.abs(e2 - m ** 2) <= 4e-6 * E ** 2 + 1e-5)


def test_learned_batch_replay_and_seed(ctx):
    n1, h1, seg1, _, _ = _batch(ctx, 1, seed=4, use_graph=1)
    n2, h2, seg2, _, _ = _batch(ctx, 1, seed=4, use_graph=0)
    n3, h3, _, _, _ = _batch(ctx, 1, seed=4, use_graph=1)
    assert n1 == n2 == n3
    assert np.array_equal(h1, h2) and np.array_equal(h1, h3)
    assert np.array_equal(seg1, seg2)


@pytest.mark.parametrize("k", [1, 3, 6])
def test_learned_batch_law_equals_plain(ctx, k):
    ctx.sample_batch(EVENTS, seed=21, events_per_round=16, proposal=0)
    hp, _ = ctx.batch_surface(k)
    ctx.sample_batch(EVENTS, seed=22, events_per_round=16, proposal=1)
    hl, _ = ctx.batch_surface(k)
    d = {"hg": hl, "ho": hp, "nev_g": EVENTS[k], "nev_o": EVENTS[k]}
    chi, ndf, ks_min = 0.0, 0, 1.0
    for pids in ss.GROUPS.values():
        c, n, ks, _, _ = ss.group_chi2(d, pids)
        chi += c
        ndf += n
        ks_min = min(ks_min, ks)
    p = stats.chi2.sf(chi, ndf)
    print(f"surface {k}: {hl.size} learned, {hp.size} plain hadrons, chi2/ndf {chi / ndf:.3f} (ndf {ndf}), p {p:.3g},"
          f" min KS p {ks_min:.3g}")
    assert ndf > 50
    assert p > 1e-3 and ks_min > 1e-4
