"""CPU checks that pin the learned-proposal oracle (oracle/mlp_oracle.py).

* the network's backward pass against central differences of its FP64
  forward pass;
* d log M / d z of the band envelope against differences of log M;
* the envelope against the exact rest-frame target on random momenta and
  directions in every band (no draw may exceed it);
* the selection thresholds: every band of positive mass has a non-empty
  range and the last one ends at 2^16.
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import mlp_oracle as mo  # noqa: E402
from iss_amd import synth  # noqa: E402


def _xavier(seed):
    rng = np.random.default_rng(seed)
    par = np.zeros(mo.N_PARAMS)
    for off, fo, fi, used in ((mo.P_W1, 64, 16, 14), (mo.P_W2, 64, 64, 64), (mo.P_W3, 8, 64, 64)):
        lim = math.sqrt(6.0 / (fi + fo))
        w = rng.uniform(-lim, lim, size=(fo, fi))
        w[:, used:] = 0.0
        par[off:off + fo * fi] = w.ravel()
    par[mo.P_B1:mo.P_W2] = rng.normal(0, 0.1, 64)
    par[mo.P_B2:mo.P_W3] = rng.normal(0, 0.1, 64)
    par[mo.P_B3:] = rng.normal(0, 0.1, 8)
    return par


def test_backward_matches_central_differences():
    par = _xavier(3)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(12, 16))
    X[:, 14:] = 0
    dO = rng.normal(size=(12, 8))
    O, cache = mo.forward(par, X, emulate=False)
    g = mo.backward(par, cache, dO, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(mo.P_W1, mo.P_B1), 10), rng.choice(np.arange(mo.P_B1, mo.P_W2), 5),
                          rng.choice(np.arange(mo.P_W2, mo.P_B2), 15), rng.choice(np.arange(mo.P_B2, mo.P_W3), 5),
                          rng.choice(np.arange(mo.P_W3, mo.P_B3), 10), np.arange(mo.P_B3, mo.N_PARAMS)])
    h = 1e-6
    for i in idx:
        pp, pm = par.copy(), par.copy()
        pp[i] += h
        pm[i] -= h
        fd = (np.sum(dO * mo.forward(pp, X, emulate=False)[0]) - np.sum(dO * mo.forward(pm, X, emulate=False)[0])) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even (1)
    assert mo.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert mo.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    assert mo.bf16(np.float32(-2.7182817)) == -2.71875


def _hadrons(n, seed, bulk=11, shear=1, neos=-1, coef_fn=None):
    """Realistic rest-frame cells and species; coef via the C++ oracle."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import Oracle
    o = Oracle(afterburner=1, qs=1, shear=shear, bulk=1 if bulk else 0, bulk_kind=bulk if bulk else 1,
               hydro_mode=2)
    sp = o.species()
    cells, _ = synth.viscous_lrf(64, sp, seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    light = np.nonzero(sp["mass"] < 1.0)[0]
    for i in range(n):
        c = rng.integers(0, cells.shape[0])
        s = int(rng.choice(light)) if rng.random() < 0.7 else int(rng.integers(0, len(sp["mass"])))
        cf = np.zeros(6)
        if bulk == 11:
            v = np.zeros(3)
            o.o.orc_k11_coeff(o.h, float(np.float32(cells[c, 13])), float(np.float32(cells[c, 16])),
                              v.ctypes.data_as(ctypes.c_void_p))
            cf[:3] = v
        P = mo.hadron_params(cells[c], cf, sp, s, shear, bulk, neos)
        out.append((cells[c], sp, s, P))
    return out


def test_band_gradient_matches_differences():
    rng = np.random.default_rng(5)
    checked = 0
    for cell, sp, s, P in _hadrons(40, seed=11):
        if P["flags"]:
            continue
        rho = rng.uniform(0.1, 0.9, 8)
        ok, lm, dz, _ = mo.band_setup(P, rho, quantise=False)
        assert ok
        z = np.log(rho / (1 - rho))
        h = 1e-6
        for k in range(8):
            zp, zm = z.copy(), z.copy()
            zp[k] += h
            zm[k] -= h
            lp = mo.band_setup(P, 1 / (1 + np.exp(-zp)), quantise=False)[1]
            lmn = mo.band_setup(P, 1 / (1 + np.exp(-zm)), quantise=False)[1]
            fd = (lp - lmn) / (2 * h)
            assert abs(fd - dz[k]) <= 1e-7 + 1e-4 * abs(fd), (k, fd, dz[k])
        checked += 1
    assert checked >= 20


@pytest.mark.parametrize("bulk,neos", [(11, -1), (0, -1)])
def test_envelope_dominates_target(bulk, neos):
    rng = np.random.default_rng(6)
    worst = 0.0
    n_checked = 0
    for cell, sp, s, P in _hadrons(60, seed=12, bulk=bulk, neos=neos):
        if P["flags"]:
            continue
        rho = rng.uniform(0.0, 1.0, 8)
        ok, lm, dz, tab = mo.band_setup(P, rho, quantise=True)
        assert ok
        for k in range(mo.N_BANDS):
            a, b = tab["a"][k], tab["b"][k]
            if not (b > a) or tab["mass"][k] <= 0:
                continue
            tau = rng.uniform(a, b, 1500)
            tau[:2] = [a, b]
            n = rng.normal(size=(1500, 3))
            n /= np.linalg.norm(n, axis=1, keepdims=True)
            t, W, y = mo.target(P, tau, n)
            env = mo.envelope(P, tab, k, tau, W, y)
            worst = max(worst, float(np.max(t / env)))
        n_checked += 1
    assert n_checked >= 30
    assert worst <= 1.0, worst


def test_thresholds_cover_every_band():
    rng = np.random.default_rng(7)
    for cell, sp, s, P in _hadrons(30, seed=13):
        if P["flags"]:
            continue
        ok, lm, dz, tab = mo.band_setup(P, rng.uniform(0, 1, 8))
        Qk = tab["Qint"]
        prev = 0
        for k in range(mo.N_BANDS):
            width = Qk[k] - prev
            assert (width > 0) == (tab["mass"][k] > 0 or False) or k > tab["klast"]
            prev = Qk[k]
        assert Qk[tab["klast"]] == 2 ** 16
