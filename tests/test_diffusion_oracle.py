"""CPU checks that pin the baryon-diffusion part of the oracle
(oracle/fssw_oracle.cpp: expint_en, DeltafTables::qmu_coeff, the diffusion
terms of yields_one_species):

  * E_n(x) against scipy.special.expn over the table range and outside it;
  * the diffusion correction of a cell's yield against direct quadrature of
    g/(2 pi^2) (dsigma.q)/D int d|p| |p|^4/(3E) f0(1 - s f0)(-n_B/(e+P) + B/E):
    the reference's series (its I_1 expansion in E_2 ... E_18, src/FSSW.cpp:
    783-809) reproduces the integral to its own truncation error, measured
    below 0.3 % of the correction for light mesons and 3 % for the heaviest
    baryon; a dropped term, a wrong sign or index fails by far more;
  * the coefficient table lookup: the reference's bilinear rule, 1e30 (no
    correction) outside the table.
"""
import os
import sys

import numpy as np
import pytest
from scipy import integrate, special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from tests import oracle_binding as ob  # noqa: E402
from iss_amd import tables  # noqa: E402

HBARC = 0.197327053


def f32(x):
    return float(np.float32(x))


def test_expint_matches_scipy():
    o = ob.lib()
    for n in (2, 4, 6, 8, 10, 12, 14, 16, 18):
        for x in (0.05, 0.3, 0.5, 0.99, 1.0, 2.7, 10.0, 50.0, 200.0, 399.9, 600.0):
            a, b = o.orc_expint(n, x), special.expn(n, x)
            assert abs(a - b) <= 1e-12 * b, (n, x, a, b)


def _qcoeff(T, muB):
    qt = tables.read_qmu_table()
    T0, mu0, dT, dmu = 0.05, 0.0, 0.001, 0.007892
    iT, im = int((T - T0) / dT), int((muB - mu0) / dmu)
    if iT < 0 or im < 0 or iT > 148 or im > 98:
        return 1e30
    xf, yf = (T - T0) / dT - iT, (muB - mu0) / dmu - im
    return (qt[iT, im] * (1 - xf) * (1 - yf) + qt[iT, im + 1] * (1 - xf) * yf + qt[iT + 1, im + 1] * xf * yf
            + qt[iT + 1, im] * xf * (1 - yf))


def _cell(T=0.15, muB=0.2):
    c = np.zeros(28, dtype=np.float32)
    c[0] = 1
    c[4:8] = (1.0, 0.2, 0.1, 0.3)
    c[8] = 1
    c[12], c[13], c[14], c[15], c[16] = 0.4, T, 0.12, 0.05, muB
    c[25:28] = (0.02, 0.01, -0.015)
    return c


@pytest.mark.parametrize("pid,tol", [(211, 0.005), (-211, 0.005), (321, 0.005), (2212, 0.01), (-2212, 0.01),
                                     (3122, 0.02), (3334, 0.04)])
def test_diffusion_yield_against_quadrature(pid, tol):
    c = _cell()
    ys = []
    for diff in (0, 1):
        orc = ob.Oracle(afterburner=1, shear=0, bulk=0, diffusion=diff)
        orc.set_cells_lrf(c[None])
        ys.append(orc.yields()[0])
    sp = orc.species()
    s = int(np.nonzero(sp["pid"] == pid)[0][0])
    m, g, sg, B = sp["mass"][s], sp["g"][s], sp["sign"][s], sp["B"][s]
    T, muB = f32(0.15), f32(0.2)
    mu = B * muB
    dsq = sum(f32(c[25 + k]) * f32(c[5 + k]) for k in range(3))
    pq = f32(0.05) / (f32(0.4) + f32(0.12))

    def a(k):
        def fun(p):
            E = np.sqrt(p * p + m * m)
            f0 = 1 / (np.exp((E - mu) / T) + sg)
            return p ** 4 / E ** k * f0 * (1 - sg * f0)
        return integrate.quad(fun, 0, 80 * T + m, limit=500, epsabs=0, epsrel=1e-12)[0] / 3

    quad = g / (2 * np.pi ** 2) / HBARC ** 3 * dsq / _qcoeff(T, muB) * (-pq * a(1) + B * a(2))
    series = ys[1][s] - ys[0][s]
    print(pid, series, quad, (series - quad) / abs(quad))
    assert abs(series - quad) < tol * abs(quad)
    assert abs(quad) > 1e-3 * ys[0][s]   # the correction is not negligible


def test_coefficient_outside_table_switches_the_correction_off():
    c = _cell(T=0.30)          # beyond the table's T = 0.199
    ys = []
    for diff in (0, 1):
        orc = ob.Oracle(afterburner=1, shear=0, bulk=0, diffusion=diff)
        orc.set_cells_lrf(c[None])
        ys.append(orc.yields()[0])
    np.testing.assert_allclose(ys[1], ys[0], rtol=1e-20, atol=1e-30)


@pytest.mark.parametrize("pid", [211, 2212, -2212, 3122])
def test_static_cell_spectrum_with_diffusion_closed_form(pid):
    """a static cell with n_B and a lab current q = (0, q_x, 0, 0): with
    f0 (1 - s f0) = sum_n n th_n e^{-n (E - mu)/T} the eta_s integral gives
    tau A sum_n e^{n mu/T} [th_n 2 mT K1(a_n)
        - (pT cos(phi) q_x / D) n th_n (n_B/(e+P) 2 mT K1(a_n) - 2 B K0(a_n))],
    a_n = n mT / T, th_n = (-s)^(n-1)"""
    from tests.test_spectra_oracle import ETA_FINE, PREF, W_FINE, _cell
    T, muB, e, P, nB, qx = f32(0.15), f32(0.2), f32(0.4), f32(0.12), f32(0.05), f32(0.002)
    c = _cell(tau=3.0, area=2.0, T=T, e=e, P=P, muB=muB)
    c[15] = nB
    c[31] = qx
    orc = ob.Oracle(afterburner=1, qs=1, shear=0, bulk=0, diffusion=1)
    orc.set_cells_lab(c[None])
    sp = orc.species()
    s = int(np.nonzero(sp["pid"] == pid)[0][0])
    m, g, sg, B = sp["mass"][s], sp["g"][s], sp["sign"][s], sp["B"][s]
    pT = np.array([0.05, 0.3, 1.0, 2.0])
    phi = np.array([0.0, 1.0, 2.5, 4.0])
    out = orc.spectra(s, pT, phi, ETA_FINE, W_FINE, restrict_deltaf=0)[0]
    mT = np.sqrt(m * m + pT * pT)
    pq = nB / (e + P)
    D = _qcoeff(T, muB)
    eq = np.zeros_like(mT)
    dq = np.zeros_like(mT)
    for n in range(1, 41):
        th = (-sg) ** (n - 1)
        a = n * mT / T
        w = np.exp(n * B * muB / T)
        eq += th * w * 2 * mT * special.kn(1, a)
        dq += n * th * w * (pq * 2 * mT * special.kn(1, a) - 2 * B * special.kn(0, a))
    ref = PREF * g * 3.0 * 2.0 * (eq[:, None] - (pT[:, None] * np.cos(phi)[None] * qx / D) * dq[:, None])
    rel_effect = np.max(np.abs(ref - PREF * g * 6.0 * eq[:, None]) / np.abs(ref))
    print(pid, "diffusion part up to", rel_effect, "of the spectrum")
    assert rel_effect > 1e-2
    np.testing.assert_allclose(out, ref, rtol=1e-9)
