"""Two-sample statistics of sampled hadron lists (test infrastructure):
chi^2 of histograms with equal-probability bins, v2(pT) differences and KS
tests per species group."""
import numpy as np
from scipy import stats as sst


def two_sample_chi2(a, wa, b, wb, bins):
    """Two-sample chi^2 of histograms of samples a, b drawn from wa and wb
    events (scaled for unequal sizes); returns (chi2, ndf)."""
    ha, _ = np.histogram(a, bins=bins)
    hb, _ = np.histogram(b, bins=bins)
    k1, k2 = np.sqrt(wb / wa), np.sqrt(wa / wb)
    m = (ha + hb) > 0
    chi2 = np.sum((k1 * ha[m] - k2 * hb[m]) ** 2 / (ha[m] + hb[m]))
    return chi2, m.sum() - 1


def quantile_bins(a, b, lo, hi, per_bin=25, max_bins=60):
    """equal-probability bins of the pooled sample, >= per_bin oracle counts"""
    nb = int(np.clip(min(a.size, b.size) // per_bin, 5, max_bins))
    pool = np.concatenate([a, b])
    pool = pool[(pool > lo) & (pool < hi)]
    edges = np.unique(np.quantile(pool, np.linspace(0, 1, nb + 1)))
    edges[0], edges[-1] = lo, hi
    return edges


def obs(h):
    pT = np.hypot(h["px"], h["py"]).astype(np.float64)
    mT = np.sqrt(pT ** 2 + h["mass"].astype(np.float64) ** 2)
    y = np.arcsinh(h["pz"] / mT)
    phi = np.arctan2(h["py"], h["px"])
    return pT, y, phi


GROUPS = {"pion": [211, -211, 111], "kaon": [321, -321, 311, -311], "proton": [2212, -2212],
          "other": None}


def group_chi2(d, pids):
    hg, ho = d["hg"], d["ho"]
    if pids is None:
        allp = sum((v for v in GROUPS.values() if v), [])
        hg = hg[~np.isin(hg["pid"], allp)]
        ho = ho[~np.isin(ho["pid"], allp)]
    else:
        hg = hg[np.isin(hg["pid"], pids)]
        ho = ho[np.isin(ho["pid"], pids)]
    pg, yg, fg = obs(hg)
    po, yo, fo = obs(ho)
    chi, ndf = 0.0, 0
    for a, b, lo, hi in ((pg, po, 0.0, 10.0), (yg, yo, -12.0, 12.0), (fg, fo, -np.pi, np.pi)):
        c, n = two_sample_chi2(a, d["nev_g"], b, d["nev_o"], quantile_bins(a, b, lo, hi))
        chi += c
        ndf += n
    # v2(pT): <cos 2 phi> in pT bins, difference over the combined error
    edges = quantile_bins(pg, po, 0.0, 10.0, per_bin=400, max_bins=12)
    for k in range(len(edges) - 1):
        cg = np.cos(2 * fg[(pg >= edges[k]) & (pg < edges[k + 1])])
        co = np.cos(2 * fo[(po >= edges[k]) & (po < edges[k + 1])])
        if cg.size > 1 and co.size > 1:
            chi += (cg.mean() - co.mean()) ** 2 / (cg.var() / cg.size + co.var() / co.size)
            ndf += 1
    ks = sst.ks_2samp(pg, po).pvalue
    return chi, ndf, ks, pg.size, po.size
