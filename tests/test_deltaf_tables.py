"""The oracle's delta-f coefficient lookups pinned against the raw tables.

The GPU's per-cell coefficients agree with the oracle's to 1e-6 through the
yield tests; these tests pin the oracle's lookups themselves, so that a
transposed (T, mu_B) index, a c0/c2 swap in the 14-moment combination, a wrong
NEoS row stride or a dropped polynomial term fails here:

* kind 1 (RTA, src/FSSW.cpp:1109-1136): the polynomial constants read from the
  reference's source text (tests/golden/rta_bulk_poly.json, written by
  tests/golden/make_deltaf_golden.py), evaluated here in numpy;
* kind 11 (OSU 14-moment, src/FSSW.cpp:1379-1430): c0/c1/c2.dat parsed here
  independently; bilinear interpolation returns node values at the nodes, the
  mean of the four corners at cell centres, and the edge node beyond the table
  (clamped indices);
* kind 21 (Chapman-Enskog NEoS, src/FSSW.cpp:1433-1487): the same on the
  200 x 200 (e, n_B) table, whose n_B spacing differs per e row;
* kind 20 (22-moment NEoS, src/FSSW.cpp:1490-1543): a synthetic table of the
  reference's layout (its own file is not shipped), node values.
"""
import json
import os
import tempfile

import numpy as np
import pytest

from tests import neos22_table
from tests import oracle_binding as ob

HBARC = 0.197327053
DT = os.path.join(ob.TABLE_DIR, "deltaf_tables", "urqmd")


def _coef(fn, h, *args, n=3):
    out = np.zeros(n)
    fn(h, *args, out.ctypes.data_as(ob.C.c_void_p))
    return out


# ------------------------------------------------------------------ kind 1
def test_kind1_polynomials_match_reference_source():
    poly = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rta_bulk_poly.json")))
    orc = ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=1)
    for T in np.linspace(0.100, 0.180, 33):
        x = T / HBARC
        got = _coef(orc.o.orc_k1_coeff, orc.h, T)
        for k in (0, 1):
            a = np.array(poly[f"c{k}"])
            terms = a * x ** np.arange(11)
            ref = terms.sum()
            # the sum cancels: compare against the size of its terms
            assert abs(got[k] - ref) <= 1e-13 * np.abs(terms).sum(), (T, k, got[k], ref)
    # a value of the fit (c1 = tau_Pi / (zeta/s)-like ratio is O(1), c0 positive)
    got = _coef(orc.o.orc_k1_coeff, orc.h, 0.150)
    assert got[0] > 0 and 0.0 < got[1] < 10.0


# ----------------------------------------------------------------- kind 11
def _raw_k11():
    """c0, c1, c2 as [iT][imu] arrays plus the T and mu_B grids, parsed from
    the raw files: three header lines, then rows (T, mu_B, value) with mu_B
    the outer and T the inner index."""
    out = {}
    for k in ("c0", "c1", "c2"):
        lines = open(os.path.join(DT, k + ".dat")).read().splitlines()[3:]
        v = np.array([[float(t) for t in ln.split()] for ln in lines if ln.strip()])
        Ts, mus = np.unique(v[:, 0]), np.unique(v[:, 1])
        grid = np.full((Ts.size, mus.size), np.nan)
        grid[np.searchsorted(Ts, v[:, 0]), np.searchsorted(mus, v[:, 1])] = v[:, 2]
        assert not np.isnan(grid).any()
        out[k] = grid
    out["T"], out["mu"] = Ts, mus
    return out


def _k11_expected(c0, c1, c2, T):
    return np.array([(c0 - c2) / T ** 4, c1 / T ** 3, (4 * c2 - c0) / T ** 4])


@pytest.fixture(scope="module")
def k11():
    return _raw_k11(), ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=11)


def test_kind11_grid_is_rectangular_and_distinct(k11):
    raw, _ = k11
    # a transposed lookup could not hide on a square grid: this one is not
    assert raw["T"].size == 101 and raw["mu"].size == 81
    assert not np.allclose(raw["c0"], raw["c2"])


def test_kind11_nodes(k11):
    raw, orc = k11
    rng = np.random.default_rng(3)
    idx = [(0, 0), (100, 80), (0, 80), (100, 0), (50, 40)] + \
          [(int(i), int(j)) for i, j in zip(rng.integers(0, 101, 40), rng.integers(0, 81, 40))]
    for i, j in idx:
        T, mu = raw["T"][i], raw["mu"][j]
        got = _coef(orc.o.orc_k11_coeff, orc.h, T, mu)
        ref = _k11_expected(raw["c0"][i, j], raw["c1"][i, j], raw["c2"][i, j], T)
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max()), (i, j, got, ref)


def test_kind11_cell_centres_are_corner_means(k11):
    raw, orc = k11
    rng = np.random.default_rng(4)
    for i, j in zip(rng.integers(0, 100, 40), rng.integers(0, 80, 40)):
        T = 0.5 * (raw["T"][i] + raw["T"][i + 1])
        mu = 0.5 * (raw["mu"][j] + raw["mu"][j + 1])
        c = [raw[k][i:i + 2, j:j + 2].mean() for k in ("c0", "c1", "c2")]
        got = _coef(orc.o.orc_k11_coeff, orc.h, T, mu)
        ref = _k11_expected(*c, T)
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max()), (i, j, got, ref)


def test_kind11_beyond_the_table_takes_the_edge_node(k11):
    raw, orc = k11
    dT, dmu = raw["T"][1] - raw["T"][0], raw["mu"][1] - raw["mu"][0]
    cases = [(raw["T"][-1] + 3.5 * dT, raw["mu"][10], -1, 10),
             (raw["T"][0] - 4.5 * dT, raw["mu"][20], 0, 20),
             (raw["T"][30], raw["mu"][-1] + 2.5 * dmu, 30, -1),
             (raw["T"][-1] + 5 * dT, raw["mu"][-1] + 5 * dmu, -1, -1)]
    for T, mu, i, j in cases:
        got = _coef(orc.o.orc_k11_coeff, orc.h, T, mu)
        ref = _k11_expected(raw["c0"][i, j], raw["c1"][i, j], raw["c2"][i, j], T)
        assert np.allclose(got, ref, rtol=1e-12), (T, mu, got, ref)


# ------------------------------------------------------------- kinds 21, 20
def _raw_neos(path, ncol):
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.strip() and not ln.startswith("#")]
    a = np.array(rows, dtype=np.float64)
    assert a.shape == (40000, ncol)
    return a.reshape(200, 200, ncol)


def _ce_expected(r):
    # (1 / zeta_hat, 1/3 - c_s^2-hat, eta_hat) from columns (c2_hat, zeta_hat, eta_hat)
    return np.array([1.0 / r[3], 1.0 / 3.0 - r[2], r[4]])


@pytest.fixture(scope="module")
def ce():
    return _raw_neos(os.path.join(DT, "NEoSBQS_CE_deltafCoeff.dat"), 5), \
        ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=21)


def test_ce_nodes(ce):
    raw, orc = ce
    rng = np.random.default_rng(5)
    idx = [(0, 0), (198, 198), (0, 150), (120, 0)] + \
          [(int(i), int(j)) for i, j in zip(rng.integers(0, 199, 60), rng.integers(0, 199, 60))]
    for i, j in idx:
        got = _coef(orc.o.orc_neos_coeff, orc.h, raw[i, j, 0], raw[i, j, 1])
        ref = _ce_expected(raw[i, j])
        assert np.allclose(got, ref, rtol=1e-6), (i, j, got, ref)


def test_ce_midpoints(ce):
    raw, orc = ce
    rng = np.random.default_rng(6)
    for i, j in zip(rng.integers(0, 199, 40), rng.integers(0, 198, 40)):
        # n_B halfway along an e row (that row's own spacing)
        got = _coef(orc.o.orc_neos_coeff, orc.h, raw[i, j, 0], 0.5 * (raw[i, j, 1] + raw[i, j + 1, 1]))
        ref = _ce_expected(0.5 * (raw[i, j] + raw[i, j + 1]))
        assert np.allclose(got, ref, rtol=1e-6), (i, j, got, ref)
        # e halfway between two rows at n_B = 0
        got = _coef(orc.o.orc_neos_coeff, orc.h, 0.5 * (raw[i, 0, 0] + raw[i + 1, 0, 0]), 0.0)
        ref = _ce_expected(0.5 * (raw[i, 0] + raw[i + 1, 0]))
        assert np.allclose(got, ref, rtol=1e-6), (i, got, ref)


def test_ce_n_b_beyond_a_row_takes_its_last_entry(ce):
    raw, orc = ce
    for i in (0, 37, 150, 198):
        got = _coef(orc.o.orc_neos_coeff, orc.h, raw[i, 0, 0], 2.0 * raw[i, 199, 1])
        assert np.allclose(got, _ce_expected(raw[i, 199]), rtol=1e-9), (i, got)


def test_22mom_synthetic_nodes():
    rows = neos22_table.table()
    raw = rows.reshape(200, 200, 8)
    with tempfile.TemporaryDirectory() as tmp:
        td = neos22_table.table_dir(tmp, rows)
        orc = ob.Oracle(afterburner=1, shear=1, bulk=1, bulk_kind=20, table_dir=td)
        rng = np.random.default_rng(7)
        idx = [(0, 0), (198, 198), (60, 3)] + \
              [(int(i), int(j)) for i, j in zip(rng.integers(0, 199, 40), rng.integers(0, 199, 40))]
        for i, j in idx:
            got = _coef(orc.o.orc_neos_coeff, orc.h, raw[i, j, 0], raw[i, j, 1], n=6)
            assert np.allclose(got, raw[i, j, 2:], rtol=1e-9), (i, j, got, raw[i, j, 2:])
        # n_B halfway: the mean of the two entries of the row
        for i, j in zip(rng.integers(0, 199, 20), rng.integers(0, 198, 20)):
            got = _coef(orc.o.orc_neos_coeff, orc.h, raw[i, j, 0], 0.5 * (raw[i, j, 1] + raw[i, j + 1, 1]), n=6)
            assert np.allclose(got, 0.5 * (raw[i, j, 2:] + raw[i, j + 1, 2:]), rtol=1e-9)
