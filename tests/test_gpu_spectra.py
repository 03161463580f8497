"""GPU smooth spectra (iss_spectra) and flows against the CPU oracle.

The oracle evaluates the reference's lab-frame Milne integrand
(src/emissionfunction.cpp:686-806) on the regulated lab cells; the GPU
contracts the boosted momentum with the float rest-frame records the sampler
keeps.  The rest-frame records hold 24-bit mantissas, and the exponent
(p.u - mu)/T reaches 50 at the top of the pT table, so the bar for FP64 is
5e-5 of the largest value at the same pT (the float rounding of the records,
amplified by E/T and by u^2 in pi_LRF), and 1e-9 where the records are exact
(static cells: closed forms).  FP32 is held to 1e-5 of FP64 (relative, on
values above 1e-6 of the row maximum).

The sampled-vs-smooth gate: the rapidity-integrated smooth spectrum (eta_s
table wide enough to be the full integral), with |delta f| rescaled to 1,
is the law the FSSW sampler draws from when delta f stays inside (-1, 1);
chi^2 of sampled (pT, phi) histograms against it.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

import iss_amd
from iss_amd import synth, tables
from tests import neos22_table
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

HBARC = 0.197327053
PREF = 1.0 / (8.0 * math.pi ** 3) / HBARC ** 3
FP64_ROW_TOL = 5e-5
FP32_RTOL = 1e-5

MODES = [
    ("shear", dict(shear=1, bulk=0, bulk_kind=1)),
    ("rta_bulk", dict(shear=1, bulk=1, bulk_kind=1)),
    ("osu14", dict(shear=1, bulk=1, bulk_kind=11)),
    ("ce21", dict(shear=1, bulk=1, bulk_kind=21)),
    ("mom22", dict(shear=1, bulk=1, bulk_kind=20)),   # synthetic table (tests/neos22_table.py)
]
PIDS = [211, -211, 321, 2212, -2212, 3122]


def _ctx(sp, df):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    kw = dict(df)
    if df["bulk"] and df["bulk_kind"] == 11:
        kw["k11"] = tables.read_kind11(1)
    if df["bulk_kind"] == 21:
        kw["neos"] = tables.read_neos(21, 1)
    if df["bulk_kind"] == 20:
        kw["neos"] = neos22_table.table()
    ctx.set_deltaf(**kw)
    return ctx


@pytest.fixture(scope="module")
def table_dir22(tmp_path_factory):
    return neos22_table.table_dir(str(tmp_path_factory.mktemp("neos22")))


def _oracle(df, td22):
    return ob.Oracle(afterburner=1, table_dir=td22, **df) if df["bulk_kind"] == 20 else ob.Oracle(afterburner=1, **df)


def _idx(sp, pids):
    return np.array([int(np.nonzero(sp["pid"] == p)[0][0]) for p in pids], dtype=np.int32)


def _row_err(a, ref):
    rowmax = np.abs(ref).max(axis=-1, keepdims=True)
    return np.max(np.abs(a - ref) / np.maximum(rowmax, 1e-300))


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


@pytest.fixture(scope="module")
def bins():
    return tables.bin_tables()


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp64_spectra_match_oracle_reference_grid(species, bins, name, df, table_dir22):
    lab = synth.viscous_lab(300, species, seed=11)
    ctx = _ctx(species, df)
    ctx.set_surface_lab(lab)
    assert 0 < ctx.num_cells() < 300   # some cells have u.dsigma < 0
    idx = _idx(species, PIDS)
    args = (idx, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"])
    orc = _oracle(df, table_dir22)
    orc.set_cells_lab(lab)
    for restrict in (1, 0):
        g = ctx.spectra(*args, precision=64, all_cells=1, restrict_deltaf=restrict)
        r = orc.spectra(*args, restrict_deltaf=restrict)
        err = _row_err(g, r)
        print(name, "restrict", restrict, "FP64 row-normalised error", err, "ms", ctx.spectra_ms())
        assert err < FP64_ROW_TOL
    # the sampled cells only
    lrf, kept = ob.transform(ob.regulate(lab))
    orc_k = _oracle(df, table_dir22)
    orc_k.set_cells_lab(lab[kept])
    g = ctx.spectra(*args, precision=64, all_cells=0)
    assert _row_err(g, orc_k.spectra(*args)) < FP64_ROW_TOL
    ctx.close()


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp32_spectra_within_1e5_of_fp64(species, bins, name, df):
    lab = synth.viscous_lab(2000, species, seed=12, frac_spacelike=0.0)
    ctx = _ctx(species, df)
    ctx.set_surface_lab(lab)
    idx = _idx(species, PIDS)
    args = (idx, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"])
    g64 = ctx.spectra(*args, precision=64)
    g32 = ctx.spectra(*args, precision=32)
    rowmax = np.abs(g64).max(axis=-1, keepdims=True)
    big = np.abs(g64) > 1e-6 * rowmax
    rel = np.abs(g32 - g64)[big] / np.abs(g64)[big]
    print(name, "FP32 max rel", rel.max(), "median", np.median(rel))
    assert rel.max() < FP32_RTOL
    ctx.close()


def test_fixed_rapidity_mode_matches_oracle(species, bins):
    df = dict(shear=1, bulk=1, bulk_kind=1)
    lab = synth.viscous_lab(200, species, seed=13)
    ctx = _ctx(species, df)
    ctx.set_surface_lab(lab)
    idx = _idx(species, [211, 2212])
    ys = np.array([-2.0, -0.5, 0.0, 0.3, 1.7])
    g = ctx.spectra(idx, bins["pT"], bins["phi"], ys, mode=1, precision=64)
    orc = ob.Oracle(afterburner=1, **df)
    orc.set_cells_lab(lab)
    r = orc.spectra(idx, bins["pT"], bins["phi"], ys, mode=1)
    assert g.shape == (2, ys.size, bins["pT"].size, bins["phi"].size)
    err = _row_err(g, r)
    print("mode 1 FP64 row-normalised error", err)
    assert err < FP64_ROW_TOL
    ctx.close()


def test_options_positive_only_and_ratio(species, bins):
    df = dict(shear=1, bulk=0, bulk_kind=1)
    lab = synth.viscous_lab(150, species, seed=14, bulk=False)
    ctx = _ctx(species, df)
    ctx.set_surface_lab(lab)
    orc = ob.Oracle(afterburner=1, **df)
    orc.set_cells_lab(lab)
    idx = _idx(species, [211, 2212])
    args = (idx, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"])
    for kw in (dict(use_pos_dN_only=1, restrict_deltaf=0), dict(restrict_deltaf=1, deltaf_max_ratio=0.5)):
        err = _row_err(ctx.spectra(*args, precision=64, **kw), orc.spectra(*args, **kw))
        print(kw, err)
        assert err < FP64_ROW_TOL
    ctx.close()


def test_static_cell_closed_form_on_gpu(species):
    """records exact in float: the GPU FP64 result is the K1 series to 1e-9"""
    ctx = _ctx(species, dict(shear=0, bulk=0, bulk_kind=1))
    c = np.zeros(34)
    T = float(np.float32(0.15))
    c[0], c[4], c[8], c[13], c[12], c[14] = 2.0, 1.5, 1.0, T, 0.3, 0.1
    ctx.set_surface_lab(c[None].astype(np.float32))
    idx = _idx(species, [211, 321, 2212])
    pT = np.array([0.05, 0.5, 1.5, 3.0])
    eta = np.arange(-12.0, 12.0 + 1e-9, 0.02)
    w = np.full(eta.size, 0.02)
    g = ctx.spectra(idx, pT, np.array([0.0, 1.0]), eta, w, precision=64, restrict_deltaf=0)
    for k, s in enumerate(idx):
        m, deg, sign = species["mass"][s], species["g"][s], species["sign"][s]
        mT = np.sqrt(m * m + pT * pT)
        ser = np.zeros_like(mT)
        for n in range(1, 40):
            th = 1.0 if sign < 0 else (-1.0) ** (n + 1)
            ser += th * 2 * mT * special.kn(1, n * mT / T)
        np.testing.assert_allclose(g[k], np.repeat((PREF * deg * 2.0 * 1.5 * ser)[:, None], 2, 1), rtol=1e-9)
    ctx.close()


def test_flows_of_gpu_spectra_match_oracle_flows(species, bins):
    df = dict(shear=1, bulk=1, bulk_kind=1)
    lab = synth.viscous_lab(300, species, seed=15)
    ctx = _ctx(species, df)
    ctx.set_surface_lab(lab)
    s = _idx(species, [211])
    g = ctx.spectra(s, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"], precision=64)[0]
    orc = ob.Oracle(afterburner=1, **df)
    orc.set_cells_lab(lab)
    r = orc.spectra(s, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"])[0]
    m = species["mass"][s[0]]
    d1, i1 = iss_amd.flows(g, bins["pT"], bins["pT_weight"], bins["phi"], bins["phi_weight"], m, 9)
    d2, i2 = ob.flows(r, bins["pT"], bins["pT_weight"], bins["phi"], bins["phi_weight"], m, 9)
    # v_n are ratios of integrals: absolute agreement to the spectra's bar
    assert np.max(np.abs(d1[:, 3:] - d2[:, 3:])) < 1e-4
    np.testing.assert_allclose(d1[:, 2], d2[:, 2], rtol=FP64_ROW_TOL)
    np.testing.assert_allclose(i1[:, 3:], i2[:, 3:], atol=1e-5)
    ctx.close()


def test_sampled_histograms_match_smooth_spectrum(species):
    """chi^2 of sampled pi+ (pT, phi) counts against the smooth spectrum"""
    cells, _ = synth.viscous_lrf(20000, species, seed=21, muB_max=0.2, target_hadrons=4000.0)
    cells = cells.copy()
    epp = cells[:, 12] + cells[:, 14]
    cells[:, 20:25] *= 0.05        # |pi| / (e+P) <= 0.015: |delta f| < 1 where the weight matters
    cells[:, 19] = np.maximum(cells[:, 19], -0.01 * epp)
    df = dict(shear=1, bulk=1, bulk_kind=1)
    ctx = _ctx(species, df)
    ctx.set_surface_lrf(cells)
    n_ev = 400
    ctx.sample(n_ev, seed=77)
    h = ctx.hadrons()
    sel = h["pid"] == 211
    px, py = h["px"][sel].astype(np.float64), h["py"][sel].astype(np.float64)
    pt = np.hypot(px, py)
    ph = np.mod(np.arctan2(py, px), 2 * np.pi)
    pt_edges = np.linspace(0.0, 2.0, 21)
    ph_edges = np.linspace(0.0, 2 * np.pi, 9)
    counts, _, _ = np.histogram2d(pt, ph, bins=[pt_edges, ph_edges])
    # expected: E * sum over Gauss nodes of pT w_pT w_phi dN/(pT dpT dphi) (all rapidities)
    xg, wg = np.polynomial.legendre.leggauss(6)
    pts = (0.5 * (pt_edges[1:] + pt_edges[:-1])[:, None] + 0.5 * np.diff(pt_edges)[:, None] * xg[None]).ravel()
    wpt = (0.5 * np.diff(pt_edges)[:, None] * wg[None]).ravel()
    phs = (0.5 * (ph_edges[1:] + ph_edges[:-1])[:, None] + 0.5 * np.diff(ph_edges)[:, None] * xg[None]).ravel()
    wph = (0.5 * np.diff(ph_edges)[:, None] * wg[None]).ravel()
    eta = np.arange(-10.0, 10.0 + 1e-9, 0.05)
    s = _idx(species, [211])
    dn = ctx.spectra(s, pts, phs, eta, np.full(eta.size, 0.05), precision=64, all_cells=0)[0]
    dens = dn * (pts * wpt)[:, None] * wph[None]
    exp = n_ev * dens.reshape(pt_edges.size - 1, 6, ph_edges.size - 1, 6).sum(axis=(1, 3))
    ok = exp > 20
    chi2 = float(np.sum((counts[ok] - exp[ok]) ** 2 / exp[ok]))
    ndf = int(ok.sum())
    p = stats.chi2.sf(chi2, ndf)
    print("pi+ sampled", int(sel.sum()), "chi2/ndf", chi2 / ndf, "ndf", ndf, "p", p)
    assert ndf > 100
    assert p > 1e-3
    ctx.close()
