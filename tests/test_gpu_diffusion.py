"""GPU baryon-diffusion delta f (include_deltaf_diffusion = 1) against the
CPU oracle, through the C-ABI.

  * FP64 per-cell yields within 1e-6 of the oracle, and the diffusion part
    alone (yield with diffusion minus yield without) within 1e-6 of the
    oracle's, so the bar pins the correction itself and not only the total;
  * FP32 species totals within 1e-5 of FP64;
  * FP64 smooth spectra within the lab-spectra bar of the oracle, the
    correction itself well above that bar;
  * sampled p and pbar (pT, phi) histograms against the GPU's own smooth
    spectrum of a surface whose rest-frame current points one way (a directed
    distortion the diffusion-free spectrum fails by far).
"""
import numpy as np
import pytest
from scipy import stats

import iss_amd
from iss_amd import synth, tables
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-6
FP32_RTOL = 1e-5
FP64_ROW_TOL = 5e-5

MODES = [
    ("rta_bulk", dict(shear=1, bulk=1, bulk_kind=1)),
    ("osu14", dict(shear=1, bulk=1, bulk_kind=11)),
]


def _ctx(sp, df, diffusion=1):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    kw = dict(df)
    if df["bulk"] and df["bulk_kind"] == 11:
        kw["k11"] = tables.read_kind11(1)
    ctx.set_deltaf(**kw, diffusion=diffusion, qmu_table=tables.read_qmu_table() if diffusion else None)
    return ctx


def _idx(sp, pids):
    return np.array([int(np.nonzero(sp["pid"] == p)[0][0]) for p in pids], dtype=np.int32)


def _row_err(a, ref):
    rowmax = np.abs(ref).max(axis=-1, keepdims=True)
    return np.max(np.abs(a - ref) / np.maximum(rowmax, 1e-300))


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp64_diffusion_yields_match_oracle(species, name, df):
    cells, _ = synth.viscous_lrf(3000, species, seed=17, muB_max=0.45, diffusion=1.0)
    per = {}
    ref = {}
    for d in (0, 1):
        ctx = _ctx(species, df, d)
        ctx.set_surface_lrf(cells)
        _, per[d] = ctx.compute_yields(precision=64, per_cell=True)
        ctx.close()
        orc = ob.Oracle(afterburner=1, diffusion=d, **df)
        orc.set_cells_lrf(cells)
        ref[d] = orc.yields()
    nz = ref[1] > 0
    assert np.all(per[1][~nz] == 0.0)
    rel = np.abs(per[1][nz] - ref[1][nz]) / ref[1][nz]
    # the diffusion part alone, relative to the yield it corrects
    both = (ref[0] > 0) & nz
    dg, do = per[1] - per[0], ref[1] - ref[0]
    reld = np.abs(dg[both] - do[both]) / ref[1][both]
    print(name, "max rel FP64", rel.max(), "correction: max rel", reld.max(),
          "median |correction|/yield", np.median(np.abs(do[both]) / ref[1][both]))
    assert rel.max() < FP64_RTOL
    assert reld.max() < FP64_RTOL
    assert np.median(np.abs(do[both]) / ref[1][both]) > 1e-3


def test_fp32_diffusion_yields_within_1e5_of_fp64(species):
    cells, _ = synth.viscous_lrf(3000, species, seed=18, muB_max=0.3, diffusion=1.0)
    ctx = _ctx(species, MODES[1][1])
    ctx.set_surface_lrf(cells)
    t64, _ = ctx.compute_yields(precision=64, per_cell=True)
    t32, _ = ctx.compute_yields(precision=32, per_cell=True)
    relt = np.abs(t32 - t64) / np.maximum(t64, 1e-300)
    print("max rel FP32 species total", relt.max())
    assert relt.max() < FP32_RTOL
    ctx.close()


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp64_diffusion_spectra_match_oracle(species, name, df):
    bins = tables.bin_tables()
    lab = synth.viscous_lab(300, species, seed=13, diffusion=0.3)
    idx = _idx(species, [211, 321, 2212, -2212, 3122])
    args = (idx, bins["pT"], bins["phi"], bins["eta"], bins["eta_weight"])
    g = {}
    for d in (0, 1):
        ctx = _ctx(species, df, d)
        ctx.set_surface_lab(lab)
        g[d] = ctx.spectra(*args, precision=64, all_cells=1)
        ctx.close()
    orc = ob.Oracle(afterburner=1, diffusion=1, **df)
    orc.set_cells_lab(lab)
    r = orc.spectra(*args)
    err = _row_err(g[1], r)
    eff = _row_err(g[0][2:4], g[1][2:4])   # p and pbar
    print(name, "FP64 row-normalised error", err, "diffusion effect on p, pbar", eff)
    assert err < FP64_ROW_TOL
    assert eff > 100 * FP64_ROW_TOL


def _coeff_D(T, muB):
    """the diffusion coefficient table's bilinear value (src/FSSW.cpp:1571-1606)"""
    qt = tables.read_qmu_table()
    x, y = (T - 0.05) / 0.001, muB / 0.007892
    iT, im = x.astype(int), y.astype(int)
    xf, yf = x - iT, y - im
    return (qt[iT, im] * (1 - xf) * (1 - yf) + qt[iT, im + 1] * (1 - xf) * yf + qt[iT + 1, im + 1] * xf * yf
            + qt[iT + 1, im] * xf * (1 - yf))


def _directed(species, amp, seed):
    """rest-frame current along x with q_x / D in [amp / 2, amp]: the
    diffusion delta f stays well inside (-1, 1) where hadrons are, so the
    sampler's Poisson mean (the unclamped yield) and the smooth spectrum
    describe the same law"""
    cells, _ = synth.viscous_lrf(20000, species, seed=seed, muB_max=0.3, target_hadrons=4000.0)
    cells = cells.copy()
    epp = cells[:, 12] + cells[:, 14]
    cells[:, 20:25] *= 0.05
    cells[:, 19] = np.maximum(cells[:, 19], -0.01 * epp)
    rng = np.random.default_rng(seed)
    D = _coeff_D(cells[:, 13].astype(np.float64), cells[:, 16].astype(np.float64))
    cells[:, 25] = amp * rng.uniform(0.5, 1, cells.shape[0]) * D
    cells[:, 26:28] = 0
    return cells


def test_sampled_baryons_match_smooth_spectrum_with_diffusion(species):
    """chi^2 of sampled p and pbar (pT, phi) counts against the smooth
    spectrum with diffusion; the diffusion-free spectrum must fail"""
    cells = _directed(species, 0.5, 23)
    df = dict(shear=1, bulk=1, bulk_kind=1)
    ctx = _ctx(species, df)
    ctx.set_surface_lrf(cells)
    ctx0 = _ctx(species, df, 0)
    ctx0.set_surface_lrf(cells)
    n_ev = 2000
    ctx.sample(n_ev, seed=78)
    h = ctx.hadrons()
    pt_edges = np.linspace(0.0, 2.5, 11)
    ph_edges = np.linspace(0.0, 2 * np.pi, 9)
    xg, wg = np.polynomial.legendre.leggauss(6)
    pts = (0.5 * (pt_edges[1:] + pt_edges[:-1])[:, None] + 0.5 * np.diff(pt_edges)[:, None] * xg[None]).ravel()
    wpt = (0.5 * np.diff(pt_edges)[:, None] * wg[None]).ravel()
    phs = (0.5 * (ph_edges[1:] + ph_edges[:-1])[:, None] + 0.5 * np.diff(ph_edges)[:, None] * xg[None]).ravel()
    wph = (0.5 * np.diff(ph_edges)[:, None] * wg[None]).ravel()
    eta = np.arange(-10.0, 10.0 + 1e-9, 0.05)
    chi = {0: 0.0, 1: 0.0}
    ndf = 0
    for pid in (2212, -2212):
        sel = h["pid"] == pid
        px, py = h["px"][sel].astype(np.float64), h["py"][sel].astype(np.float64)
        counts, _, _ = np.histogram2d(np.hypot(px, py), np.mod(np.arctan2(py, px), 2 * np.pi),
                                      bins=[pt_edges, ph_edges])
        s = _idx(species, [pid])
        exp = {}
        for d, c in ((1, ctx), (0, ctx0)):
            dn = c.spectra(s, pts, phs, eta, np.full(eta.size, 0.05), precision=64, all_cells=0)[0]
            dens = dn * (pts * wpt)[:, None] * wph[None]
            exp[d] = n_ev * dens.reshape(pt_edges.size - 1, 6, ph_edges.size - 1, 6).sum(axis=(1, 3))
        ok = exp[1] > 20
        for d in (0, 1):
            chi[d] += float(np.sum((counts[ok] - exp[d][ok]) ** 2 / exp[d][ok]))
        ndf += int(ok.sum())
        print(pid, "sampled", int(sel.sum()))
    p1, p0 = stats.chi2.sf(chi[1], ndf), stats.chi2.sf(chi[0], ndf)
    print("with diffusion chi2/ndf", chi[1] / ndf, "p", p1, "| without:", chi[0] / ndf, "p", p0, "ndf", ndf)
    assert ndf > 80
    assert p1 > 1e-3
    assert p0 < 1e-6
    ctx.close()
    ctx0.close()


def test_zero_current_equals_no_diffusion(species):
    """q = 0 everywhere: the diffusion terms vanish identically (yields in
    both precisions, the sampled list of the same seed)"""
    cells, _ = synth.viscous_lrf(2000, species, seed=19, muB_max=0.3)
    cells = cells.copy()
    cells[:, 25:28] = 0.0
    out = {}
    for d in (0, 1):
        ctx = _ctx(species, MODES[0][1], d)
        ctx.set_surface_lrf(cells)
        out[d] = [ctx.compute_yields(precision=p, per_cell=True)[1] for p in (64, 32)]
        ctx.sample(5, seed=3)
        out[d].append(ctx.hadrons())
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
