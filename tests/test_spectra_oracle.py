"""CPU checks that pin the smooth-spectra and flow parts of the oracle
(oracle/fssw_oracle.cpp: smooth_spectra, flows) by closed forms, and the
product's host-side flows (iss_flows) against the oracle.

Closed forms used (single cells, mode 0 with a wide fine eta_s table so the
sum is the integral to 1e-12):
  * static cell, dsigma along u:  sum_n th_n e^{n mu/T} 2 mT K1(n mT/T) with
    th_n = 1 (Boltzmann), 1 (BE), (-1)^{n+1} (FD);
  * transverse flow u = gamma (1, v, 0, 0), Boltzmann:
    2 mT exp(gamma v pT cos phi / T) K1(gamma mT / T); its v_n(pT) are
    I_n(a) / I_0(a), a = gamma v pT / T;
  * shear pi^xx = -pi^yy: a factor 1 + pT^2 pi cos(2 phi) / (2 T^2 (e+P)),
    i.e. v_2 = pT^2 pi / (4 T^2 (e+P)); the |delta f| <= 1 rescaling and the
    dropping of negative contributions follow from the same factor;
  * bulk kind 1: 2 mT K1(a) - c0 Pi (2 m^2 K0(a) / (3T) - c1 mT^2 (K0(a) + K2(a)) / T);
  * mode 1 (fixed rapidity, cell at eta_s): mT cosh(y - eta_s) f0(mT cosh(y - eta_s)).
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy import special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from tests import oracle_binding as ob  # noqa: E402
from iss_amd import tables  # noqa: E402

HBARC = 0.197327053
PREF = 1.0 / (8.0 * math.pi ** 3) / HBARC ** 3


def _cell(tau=1.0, eta=0.0, area=2.0, u=(0.0, 0.0, 0.0), T=0.15, e=0.3, P=0.1, muB=0.0, pi=None, bulkPi=0.0):
    """one lab-frame record (34 floats), dsigma = (area, 0, 0, 0) / tau along tau"""
    c = np.zeros(34)
    c[0], c[3] = tau, eta
    c[4] = area
    c[8] = math.sqrt(1 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    c[9], c[10], c[11] = u
    c[12], c[13], c[14], c[16] = e, T, P, muB
    if pi is not None:
        for k, v in pi.items():
            c[19 + ["00", "01", "02", "03", "11", "12", "13", "22", "23", "33"].index(k)] = v
    c[29] = bulkPi
    return c


def f32(x):
    """the value a float field of a cell record holds"""
    return float(np.float32(x))


def _species_index(orc, pid):
    sp = orc.species()
    return int(np.nonzero(sp["pid"] == pid)[0][0]), sp


ETA_FINE = np.arange(-12.0, 12.0 + 1e-9, 0.02)
W_FINE = np.full(ETA_FINE.size, 0.02)


def _series(m, pT, T, mu, sign, a_scale=1.0, nterms=40):
    mT = np.sqrt(m * m + pT * pT)
    s = np.zeros_like(mT)
    for n in range(1, nterms + 1):
        th = 1.0 if sign == 0 and n == 1 else (0.0 if sign == 0 else (1.0 if sign < 0 else (-1.0) ** (n + 1)))
        if th == 0.0:
            continue
        s += th * math.exp(n * mu / T) * 2 * mT * special.kn(1, n * a_scale * mT / T)
    return s


@pytest.mark.parametrize("qs,pid", [(0, 211), (1, 211), (1, 2212), (1, 321)])
def test_static_cell_quantum_series(qs, pid):
    orc = ob.Oracle(afterburner=1, qs=qs, shear=0, bulk=0)
    si, sp = _species_index(orc, pid)
    T, muB = f32(0.15), f32(0.1)
    orc.set_cells_lab(_cell(tau=3.0, area=2.0, T=T, muB=muB)[None])
    pT = np.array([0.05, 0.3, 1.0, 2.5])
    phi = np.array([0.0, 1.0, 2.5])
    out = orc.spectra(si, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=0)[0]
    m, g, sign, B = sp["mass"][si], sp["g"][si], sp["sign"][si], sp["B"][si]
    ref = PREF * g * 3.0 * 2.0 * _series(m, pT, T, B * muB, sign)
    np.testing.assert_allclose(out, np.repeat(ref[:, None], phi.size, 1), rtol=1e-10)


def test_transverse_flow_blast_wave_and_flows():
    orc = ob.Oracle(afterburner=1, qs=0, shear=0, bulk=0)
    si, sp = _species_index(orc, 211)
    T = f32(0.14)
    ux = f32(0.5 / math.sqrt(0.75))      # v = 0.5
    gam = f32(math.sqrt(1 + ux * ux))    # u^tau as the regulated record holds it
    v = ux / gam
    orc.set_cells_lab(_cell(tau=1.0, area=1.0, u=(ux, 0.0, 0.0), T=T)[None])
    pT = np.array([0.1, 0.4, 0.9, 1.6])
    nphi = 256
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    phw = np.full(nphi, 2 * math.pi / nphi)
    out = orc.spectra(si, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=0)[0]
    m, g = sp["mass"][si], sp["g"][si]
    mT = np.sqrt(m * m + pT * pT)
    ref = PREF * g * 2 * mT[:, None] * np.exp(gam * v * pT[:, None] * np.cos(phi)[None] / T) * \
        special.kn(1, gam * mT / T)[:, None]
    np.testing.assert_allclose(out, ref, rtol=1e-10)
    diff, inte = ob.flows(out, pT, np.ones_like(pT), phi, phw, m, to_order=4)
    a = gam * v * pT / T
    for n in range(1, 5):
        vn = special.iv(n, a) / special.iv(0, a)
        np.testing.assert_allclose(diff[:, 3 + 3 * (n - 1)], vn, rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(diff[:, 4 + 3 * (n - 1)], 0.0, atol=1e-12)
        np.testing.assert_allclose(diff[:, 5 + 3 * (n - 1)], vn, rtol=1e-9, atol=1e-14)
    # dN/(pT dpT dy) column: the phi integral / 2 pi
    np.testing.assert_allclose(diff[:, 2], PREF * g * 2 * mT * special.kn(1, gam * mT / T) * special.iv(0, a),
                               rtol=1e-9)
    np.testing.assert_allclose(diff[:, 1], mT - m, rtol=1e-12)


def test_shear_cos2phi_restrict_and_positive_only():
    orc = ob.Oracle(afterburner=1, qs=0, shear=1, bulk=0, bulk_kind=1)
    si, sp = _species_index(orc, 211)
    T, e, P = f32(0.15), f32(0.3), f32(0.1)
    pis = f32(0.02)
    orc.set_cells_lab(_cell(T=T, e=e, P=P, pi={"11": pis, "22": -pis})[None])
    pT = np.array([0.2, 0.8, 1.5, 3.0])
    nphi = 64
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    m, g = sp["mass"][si], sp["g"][si]
    mT = np.sqrt(m * m + pT * pT)
    base = PREF * g * 2.0 * 2 * mT * special.kn(1, mT / T)
    df = pT[:, None] ** 2 * pis * np.cos(2 * phi)[None] / (2 * T * T * (e + P))
    assert np.abs(df).max() > 1.5  # the rescaling and the positive cut both act at large pT
    out = orc.spectra(si, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=0)[0]
    np.testing.assert_allclose(out, base[:, None] * (1 + df), rtol=1e-9, atol=1e-12 * base.max())
    small = pT < 1.0
    diff, _ = ob.flows(out, pT, np.ones_like(pT), phi, np.full(nphi, 2 * math.pi / nphi), m, to_order=3)
    np.testing.assert_allclose(diff[small, 6], pT[small] ** 2 * pis / (4 * T * T * (e + P)), rtol=1e-9)
    np.testing.assert_allclose(diff[small, 3], 0.0, atol=1e-12)
    # |delta f| rescaled to at most 0.7
    out_r = orc.spectra(si, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=1, deltaf_max_ratio=0.7)[0]
    dfr = df * np.minimum(1.0, 0.7 / (np.abs(df) + 1e-10))
    np.testing.assert_allclose(out_r, base[:, None] * (1 + dfr), rtol=1e-9)
    # negative contributions dropped
    out_p = orc.spectra(si, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=0, use_pos_dN_only=1)[0]
    np.testing.assert_allclose(out_p, base[:, None] * np.maximum(1 + df, 0.0), rtol=1e-9,
                               atol=1e-12 * base.max())


def test_bulk_kind1_bessel_form():
    orc = ob.Oracle(afterburner=1, qs=0, shear=0, bulk=1, bulk_kind=1)
    si, sp = _species_index(orc, 2212)
    T, Pi = f32(0.15), f32(-0.02)
    orc.set_cells_lab(_cell(T=T, bulkPi=Pi)[None])
    c = np.zeros(3)
    ob.lib().orc_k1_coeff(orc.h, T, ob._p(c))
    pT = np.array([0.1, 0.5, 1.2, 2.0])
    out = orc.spectra(si, pT, np.array([0.3]), ETA_FINE, W_FINE, restrict_deltaf=0)[0][:, 0]
    m, g = sp["mass"][si], sp["g"][si]
    mT = np.sqrt(m * m + pT * pT)
    a = mT / T
    Pl = Pi / HBARC
    ref = 2 * mT * special.kn(1, a) - c[0] * Pl * (2 * m * m * special.kn(0, a) / (3 * T)
                                                   - c[1] * mT ** 2 * (special.kn(0, a) + special.kn(2, a)) / T)
    np.testing.assert_allclose(out, PREF * g * 2.0 * ref, rtol=1e-9)
    assert abs(c[0] * Pl) > 1e-3  # the bulk term is not negligible


def test_fixed_rapidity_mode():
    orc = ob.Oracle(afterburner=1, qs=0, shear=0, bulk=0)
    si, sp = _species_index(orc, 321)
    T, eta_s, tau = f32(0.15), f32(0.7), 2.0
    orc.set_cells_lab(_cell(tau=tau, eta=eta_s, area=1.5, T=T)[None])
    pT = np.array([0.3, 1.1])
    ys = np.array([-1.0, 0.0, 0.7, 2.0])
    out = orc.spectra(si, pT, np.array([0.0, 2.0]), ys, mode=1, restrict_deltaf=0)[0]
    m, g = sp["mass"][si], sp["g"][si]
    mT = np.sqrt(m * m + pT * pT)
    for iy, y in enumerate(ys):
        E = mT * math.cosh(y - eta_s)
        ref = PREF * g * tau * 1.5 * E * np.exp(-E / T)
        np.testing.assert_allclose(out[iy], np.repeat(ref[:, None], 2, 1), rtol=1e-12)


def test_flows_pure_harmonics_reference_grid():
    # the tables hold 8 significant digits: Gauss quadrature exact to ~1e-8
    b = tables.bin_tables()
    pT, phi = b["pT"], b["phi"]
    rng = np.random.default_rng(1)
    v = rng.uniform(0, 0.2, (pT.size, 4))
    psi = rng.uniform(0, math.pi, (pT.size, 4))
    dN = np.ones((pT.size, phi.size))
    for n in range(1, 5):
        dN += 2 * v[:, n - 1:n] * np.cos(n * (phi[None] - psi[:, n - 1:n]))
    dN *= np.exp(-pT)[:, None]
    diff, inte = ob.flows(dN, pT, b["pT_weight"], phi, b["phi_weight"], 0.13957, to_order=6)
    for n in range(1, 5):
        np.testing.assert_allclose(diff[:, 3 + 3 * (n - 1)], v[:, n - 1] * np.cos(n * psi[:, n - 1]), atol=1e-7)
        np.testing.assert_allclose(diff[:, 4 + 3 * (n - 1)], v[:, n - 1] * np.sin(n * psi[:, n - 1]), atol=1e-7)
        np.testing.assert_allclose(diff[:, 5 + 3 * (n - 1)], v[:, n - 1], atol=1e-7)
    np.testing.assert_allclose(diff[:, 3 + 3 * 4:], 0.0, atol=1e-7)  # orders 5, 6
    np.testing.assert_allclose(diff[:, 2], np.exp(-pT), rtol=1e-7)
    # integrated: row 0 normalisation, rows n weighted by pT dpT
    wt = pT * b["pT_weight"]
    np.testing.assert_allclose(inte[0], [0, np.sum(2 * math.pi * np.exp(-pT) * wt), 0, 1, 0, 1], rtol=1e-7)
    num = np.sum(2 * math.pi * np.exp(-pT) * v[:, 1] * np.cos(2 * psi[:, 1]) * wt)
    np.testing.assert_allclose(inte[2, 1], num, rtol=1e-6)
    np.testing.assert_allclose(inte[:, 0], np.arange(7))


def test_product_flows_match_oracle():
    import iss_amd
    b = tables.bin_tables()
    rng = np.random.default_rng(5)
    dN = rng.uniform(0.5, 2.0, (b["pT"].size, b["phi"].size))
    d1, i1 = ob.flows(dN, b["pT"], b["pT_weight"], b["phi"], b["phi_weight"], 0.938, to_order=9)
    d2, i2 = iss_amd.flows(dN, b["pT"], b["pT_weight"], b["phi"], b["phi_weight"], 0.938, to_order=9)
    np.testing.assert_allclose(d2, d1, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(i2, i1, rtol=1e-13, atol=1e-15)
