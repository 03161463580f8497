"""CPU oracle of the learned proposal (test infrastructure only).

Nothing in the product imports this file: tests/ use it to check the HIP
kernels of iss_amd/csrc/learned_kernels.h.  The learned proposal has no
counterpart in the reference; what it must preserve is the reference's
rest-frame target, the density that the rejection loop of
src/FSSW.cpp:1852-1966 (with src/MomentumSamplerBase.cpp:20-58) accepts:

    pi(t, n) ~ p E f0(E) w1 w2,   t = (E - mu)/T in [x, t_max],
    w1 = clamp(p.dsigma / (E (|ds0| + |ds|)), 0, 1)      (src/FSSW.cpp:1867-1891)
    w2 = clamp((1 + df_shear + df_bulk) / 2, 0, 1)        (src/FSSW.cpp:1897-1942)

This file restates, in FP64 numpy:

* hadron_params: the per-hadron parameters (mu' = min(m, B muB + ...) as at
  src/FSSW.cpp:1861-1863 with the float sum of the device, the delta-f
  coefficients of src/FSSW.cpp:1897-1928 written as polynomials in E);
* features: the 14 network inputs;
* forward / backward: the 14 -> 64 -> 64 -> 8 tanh-GELU perceptron, with
  BF16 rounding at the points the kernels round (inputs, weights, hidden
  activations, output and hidden gradients) when emulate=True, and pure
  FP64 when emulate=False;
* band_setup: the band envelope (bounds, fp16 / u8 quantisation, selection
  thresholds) and d log M / d z;
* target / envelope: pointwise target and envelope, for brute-force checks.

Pinning (tests/test_learned_oracle.py): the backward pass against central
differences of the FP64 forward, d log M / d z against differences of
band_setup, and the envelope against the target on random momenta and
directions (every draw below the envelope).  The target's delta-f forms are
pinned end to end: the learned sampler's output is compared statistically
with the plain sampler's and with the reference's T^{mu nu} criterion.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_OUT, N_BANDS = 16, 64, 8, 9
P_W1 = 0
P_B1 = P_W1 + N_HID * N_IN
P_W2 = P_B1 + N_HID
P_B2 = P_W2 + N_HID * N_HID
P_W3 = P_B2 + N_HID
P_B3 = P_W3 + N_OUT * N_HID
N_PARAMS = P_B3 + N_OUT
LOG_MARGIN = 1e-4
HBARC = 0.197327053
# band edges in t - x: 2 (-log(1 - k/8)), k = 0..7, then +4 for the far tail
EDGES = [2.0 * -math.log(1.0 - k / 8.0) for k in range(8)] + [2.0 * -math.log(1.0 / 8.0) + 4.0]

# |p|-sampler tables (t = 0..2 Bose-Einstein, 3..5 Fermi-Dirac): first grid
# point, step and length, as the library builds them
_M0 = [0.05, 30.0, 50.0, 0.0, 30.0, 50.0]


def table_grid(tb):
    fer = tb >= 3
    emin = _M0[tb] if fer else _M0[tb] + 0.05
    emax = emin + (40.0 if fer else 50.0)
    de = 0.02 if fer else 0.05
    n = int((emax - emin) / de + 1)
    return emin, de, n


# --------------------------------------------------------------- rounding
def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f16(x):
    return float(np.float16(np.float32(x)))


def f16_up(x):
    """Smallest fp16 value >= x."""
    h = np.float16(np.float32(x))
    if float(h) < x:
        h = np.nextafter(h, np.float16(np.inf))
    return float(h)


# ---------------------------------------------------------------- network
def unpack(par):
    par = np.asarray(par, dtype=np.float64)
    W1 = par[P_W1:P_B1].reshape(N_HID, N_IN)
    b1 = par[P_B1:P_W2]
    W2 = par[P_W2:P_B2].reshape(N_HID, N_HID)
    b2 = par[P_B2:P_W3]
    W3 = par[P_W3:P_B3].reshape(N_OUT, N_HID)
    b3 = par[P_B3:N_PARAMS]
    return W1, b1, W2, b2, W3, b3


def _gelu_t(x):
    return np.tanh(0.7978845608028654 * (x + 0.044715 * x ** 3))


def gelu(x):
    return 0.5 * x * (1.0 + _gelu_t(x))


def gelu_d(x):
    th = _gelu_t(x)
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * 0.7978845608028654 * (1.0 + 3 * 0.044715 * x * x)


def forward(par, X, emulate=True):
    """X: (n, 16) inputs -> (outputs (n, 8), cache)."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = unpack(par)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    Xb = r(X)
    Z1 = Xb @ W1.T + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2.T + b2
    H2 = r(gelu(Z2))
    O = H2 @ W3.T + b3
    return O, dict(Xb=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, W3=W3)


def backward(par, cache, dO, emulate=True):
    """Gradient of sum_s <dO_s, O_s> with respect to the parameters."""
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    dOb = r(dO)
    g = np.zeros(N_PARAMS)
    g[P_W3:P_B3] = (dOb.T @ cache["H2"]).ravel()
    g[P_B3:N_PARAMS] = np.asarray(dO, dtype=np.float64).sum(0)
    dZ2 = (dOb @ cache["W3"]) * gelu_d(cache["Z2"])
    dZ2b = r(dZ2)
    g[P_W2:P_B2] = (dZ2b.T @ cache["H1"]).ravel()
    g[P_B2:P_W3] = dZ2.sum(0)
    dZ1 = (dZ2b @ cache["W2"]) * gelu_d(cache["Z1"])
    dZ1b = r(dZ1)
    g[P_W1:P_B1] = (dZ1b.T @ cache["Xb"]).ravel()
    g[P_B1:P_W2] = dZ1.sum(0)
    return g


# ------------------------------------------------------- hadron parameters
def hadron_params(cell, coef, sp, s, shear, bulk, neos, quirk=0):
    """cell: 28 rest-frame floats; coef: 6 delta-f coefficients of the cell;
    sp: species arrays; bulk: 0, 1 (RTA), 11 (14-moment), 20, 21."""
    f = np.float32
    cell = np.asarray(cell, dtype=np.float32)
    T = f(cell[13])
    m = float(sp["mass"][s])
    Bq, Sq, Qq = f(sp["B"][s]), f(sp["S"][s]), f(sp["Q"][s])
    mu_f = f(f(f(Bq * cell[16]) + f(Sq * cell[17])) + f(Qq * cell[18]))
    mu = min(m, float(mu_f))
    Td = float(T)
    x = m / Td - mu / Td
    sign = int(sp["sign"][s])
    tb = (0 if sign == -1 else 3) + (0 if x < 30.0 else (1 if x < 50.0 else 2))
    et0, de, n = table_grid(tb)
    i0 = int((x - et0) / de)
    flags = 1 if (i0 < 0 or i0 >= n - 1) else 0
    if (sign < 0 and x < 1e-3) or (quirk and tb >= 3):
        flags |= 2
    pim = np.array([[cell[20], cell[21], cell[22]], [cell[21], cell[23], cell[24]],
                    [cell[22], cell[24], -float(cell[20]) - float(cell[23])]], dtype=np.float64)
    lam, V = np.linalg.eigh(pim)
    order = np.argsort(lam)[::-1]
    lam, V = lam[order], V[:, order]
    d = np.array(cell[5:8], dtype=np.float64)
    dn = float(np.linalg.norm(d))
    dh = V.T @ d / dn if dn > 0 else np.zeros(3)
    P = dict(T=Td, mu=mu, m=m, sgn=float(sign), x=x, tmax=et0 + (n - 1) * de,
             d0=float(cell[4]), dn=dn, D=abs(float(cell[4])) + dn,
             l1=lam[0], l2=lam[1], l3=lam[2], flags=flags, tb=tb, dh=dh, R=V)
    cf = [float(np.float32(v)) for v in coef]
    P["smode"], P["ps0"] = 0, 0.0
    if shear:
        if neos == 1:
            P["smode"], P["ps0"] = 2, 1.0 / (2.0 * cf[2]) / Td
        elif neos == 0:
            P["smode"], P["ps0"] = 1, cf[0]
        else:
            P["smode"], P["ps0"] = 1, 1.0 / (2.0 * Td * Td * (float(cell[12]) + float(cell[14])))
    qi = q0 = q1 = q2 = 0.0
    if bulk:
        Pi = float(cell[19])
        if bulk == 1:
            Pi = Pi / HBARC
        B, S, Q = float(sp["B"][s]), float(sp["S"][s]), float(sp["Q"][s])
        if bulk in (1, 21):
            qi = -cf[0] * Pi * m * m / (3.0 * Td)
            q1 = cf[0] * cf[1] * Pi / Td
        elif bulk == 11:
            q0, q1, q2 = Pi * cf[0] * m * m, Pi * cf[1] * B, Pi * cf[2]
        elif bulk == 20:
            q0, q1, q2 = Pi * m * m * cf[2], Pi * (B * cf[3] + S * cf[4] + Q * cf[5]), Pi * (cf[1] - cf[2])
    P.update(qi=qi, q0=q0, q1=q1, q2=q2)
    return P


def ps_of(P, E):
    if P["smode"] == 2:
        return P["ps0"] / E
    return P["ps0"] if P["smode"] == 1 else 0.0


def qb_of(P, E):
    return P["qi"] / E + P["q0"] + E * (P["q1"] + P["q2"] * E)


def features(cell, sp, s, P):
    cell = np.asarray(cell, dtype=np.float64)
    epp = cell[12] + cell[14]
    ie = 5.0 / epp if epp > 0 else 0.0
    Eth = P["m"] + 2.0 * P["T"]
    f = np.zeros(N_IN)
    f[0] = (P["T"] / 0.15 - 1.0) * 10.0
    f[1] = cell[16] / 0.3
    f[2:7] = cell[20:25] * ie
    f[7] = cell[19] * ie
    f[8] = P["m"]
    f[9] = sp["g"][s] * 0.25
    f[10] = sp["B"][s]
    f[11] = P["sgn"]
    f[12] = (Eth * Eth - P["m"] ** 2) * ps_of(P, Eth) * (P["l1"] - P["l3"])
    f[13] = qb_of(P, Eth)
    return f


# ----------------------------------------------------------- band envelope
def _clamp01(v):
    return min(max(v, 0.0), 1.0)


def band_setup(P, rho, quantise=True):
    """Envelope of one hadron.  Returns (ok, log M, dz[8], table) where
    table = dict(c, s, kap, beta (per band), Q (8 thresholds / 2^16), klast,
    a, b (band ends in t - x), mass)."""
    T, mu, m, sg, x = P["T"], P["mu"], P["m"], P["sgn"], P["x"]
    tau_max = P["tmax"] - x
    mk = np.zeros(N_BANDS)
    dS2 = np.zeros(N_OUT)
    MtS1 = np.zeros(N_OUT)
    tab = dict(c=np.zeros(N_BANDS), s=np.zeros(N_BANDS), kap=np.zeros(N_BANDS), beta=np.zeros(N_BANDS),
               a=np.zeros(N_BANDS), b=np.zeros(N_BANDS))
    ok = tau_max > 0
    for k in range(N_BANDS):
        a = min(EDGES[k], tau_max)
        b = min(EDGES[k + 1], tau_max) if k < N_BANDS - 1 else tau_max
        tab["a"][k], tab["b"][k] = a, b
        w = b - a
        if not w > 0:
            continue
        ta, tb = x + a, x + b
        Elo, Ehi = T * ta + mu, T * tb + mu
        plo = math.sqrt(max(Elo * Elo - m * m, 0.0))
        phi = math.sqrt(max(Ehi * Ehi - m * m, 0.0))
        if not (phi > 0 and Ehi > 0):
            continue
        vlo = plo / Elo if Elo > 0 else 0.0
        vhi = phi / Ehi
        beta = 0.0
        if P["dn"] > 0:
            beta = min(0.98, vlo * P["dn"] / P["d0"]) if P["d0"] > 0 else 0.98
        if quantise:
            beta = math.floor(beta * 255.0) / 255.0
        S1 = 0.0
        if P["D"] > 0:
            S1 = max(_clamp01((P["d0"] + vhi * P["dn"]) / P["D"]) / (1 + beta),
                     _clamp01((P["d0"] - vlo * P["dn"]) / P["D"]) / (1 - beta))
        ea, eb = math.exp(-ta), math.exp(-tb)
        oma, omb = 1 / (1 + sg * ea), 1 / (1 + sg * eb)
        om_lo, om_hi = min(oma, omb), max(oma, omb)
        Qmax = max(qb_of(P, Elo), qb_of(P, Ehi))
        if P["qi"] == 0 and P["q2"] != 0:
            Es = -P["q1"] / (2 * P["q2"])
            if Elo < Es < Ehi:
                Qmax = max(Qmax, qb_of(P, Es))
        elif P["qi"] != 0 and P["q2"] == 0 and P["q1"] != 0 and P["qi"] / P["q1"] > 0:
            Es = math.sqrt(P["qi"] / P["q1"])
            if Elo < Es < Ehi:
                Qmax = max(Qmax, qb_of(P, Es))
        Amax = 1 + (om_hi * Qmax if Qmax >= 0 else om_lo * Qmax)
        shear = P["smode"] != 0 and P["l1"] > 0
        cmin = cmax = 0.0
        if shear:
            pa, pb = ps_of(P, max(Elo, 1e-30)), ps_of(P, Ehi)
            psmin, psmax = min(pa, pb), max(pa, pb)
            if psmin >= 0:
                cmin, cmax = om_lo * plo * plo * psmin, om_hi * phi * phi * psmax
            elif psmax <= 0:
                cmin, cmax = om_hi * phi * phi * psmin, om_lo * plo * plo * psmax
            else:
                cmin, cmax = om_hi * phi * phi * psmin, om_hi * phi * phi * psmax
        kap, kap0, clipped = 0.0, 0.0, True
        if k < N_OUT and shear and P["l3"] < 0:
            tm = x + 0.5 * (a + b)
            Em = T * tm + mu
            pm2 = max(Em * Em - m * m, 0.0)
            omm = 1 / (1 + sg * math.exp(-tm))
            cm = omm * pm2 * ps_of(P, Em)
            am = 1 + omm * qb_of(P, Em)
            kap0 = cm / am if am > 0 else 0.0
            raw = rho[k] * kap0
            lo, hi = -0.98 / P["l1"], 0.98 / (-P["l3"])
            kap = min(max(raw, lo), hi)
            if quantise:
                kap = f16(kap)
            if 1 + kap * P["l1"] < 0.01 or 1 + kap * P["l3"] < 0.01:
                kap = 0.0
            clipped = not (lo < raw < hi)
        cands = [P["l3"] if shear else 0.0, 0.0, P["l1"] if shear else 0.0]
        if shear and cmax != 0:
            wc = (2 - Amax) / cmax
            if 0 < wc < P["l1"]:
                cands.append(wc)
        if shear and cmin != 0:
            wc = (2 - Amax) / cmin
            if P["l3"] < wc < 0:
                cands.append(wc)
        S2, Wst, Nst = 0.0, 0.0, 0.0
        for Wc in cands:
            num = _clamp01(0.5 * (Amax + (cmax * Wc if Wc >= 0 else cmin * Wc)))
            v = num / (1 + kap * Wc)
            if v > S2:
                S2, Wst, Nst = v, Wc, num
        # |p| envelope
        U = (-T * T * (Ehi * Ehi + m * m) / phi ** 4 - T * T / (Ehi * Ehi)
             + (ea / (1 - ea) ** 2 if sg < 0 else 0.0))
        tau_s = 0.5 * (a + b) if k < N_BANDS - 1 else min(a + 1.0, b)
        ts = x + tau_s
        Es = T * ts + mu
        ps2 = Es * Es - m * m
        if U <= 0 and ps2 > 0:
            es = math.exp(-ts)
            Gs = math.log(math.sqrt(ps2) * Es) - tau_s - math.log1p(sg * es)
            ss = T * Es / ps2 + T / Es - 1 + sg * es / (1 + sg * es)
            cc = Gs - ss * tau_s
        else:
            ss = -1.0
            cc = math.log(phi * Ehi * om_hi) if om_hi < math.inf else math.inf
        if not (math.isfinite(cc) and math.isfinite(ss)):
            ok = False
            continue
        Mt = math.exp(cc + ss * a) * (math.expm1(ss * w) / ss if abs(ss * w) > 1e-6 else w)
        if not S1 * S2 > 0:
            continue
        c0 = cc + math.log(S1 * S2) + LOG_MARGIN
        if quantise:
            s16 = f16(ss)
            c1 = c0 + max(0.0, (ss - s16) * a, (ss - s16) * b)
            c16 = f16_up(c1)
        else:
            s16, c16 = ss, c0
        mk[k] = math.exp(c16 + s16 * a) * (math.expm1(s16 * w) / s16 if abs(s16 * w) > 1e-6 else w)
        tab["c"][k], tab["s"][k], tab["kap"][k], tab["beta"][k] = c16, s16, kap, beta
        if k < N_OUT:
            MtS1[k] = Mt * S1 * math.exp(LOG_MARGIN)
            dS2[k] = 0.0 if (clipped or not shear) else -Nst * Wst / (1 + kap * Wst) ** 2 * kap0
    M = mk.sum()
    if not (ok and M > 0 and math.isfinite(M)):
        return False, 0.0, np.zeros(N_OUT), tab
    logM = math.log(M)
    rho = np.asarray(rho, dtype=np.float64)
    dz = MtS1 * dS2 / M * rho * (1 - rho)
    # selection thresholds, in units of 2^-16
    two32 = 65536
    klast = int(np.nonzero(mk > 0)[0].max())
    Qk = np.zeros(N_BANDS, dtype=np.int64)
    Qp, cum = 0, 0.0
    for k in range(N_BANDS):
        cum += mk[k]
        if k >= klast:
            q = two32
        else:
            q = int(math.floor(cum / M * two32))
            q = max(q, Qp + 1) if mk[k] > 0 else Qp
            after = int(np.sum(mk[k + 1:klast + 1] > 0))
            q = min(q, two32 - after)
        Qk[k] = q
        Qp = q
    cs = np.diff(np.concatenate([[0], Qk])) / two32
    R = max(mk[k] / cs[k] for k in range(N_BANDS) if mk[k] > 0 and cs[k] > 0)
    for k in range(N_BANDS):
        if mk[k] > 0 and cs[k] > 0:
            cg = tab["c"][k] - math.log(mk[k] / (cs[k] * R))
            tab["c"][k] = f16_up(max(cg, tab["c"][k])) if quantise else max(cg, tab["c"][k])
    tab["Q"] = np.minimum(Qk[:8], two32 - 1) / two32
    tab["Qint"] = Qk
    tab["klast"] = klast
    tab["mass"] = mk
    return True, logM, dz, tab


# ------------------------------------------------------- pointwise checks
def target(P, tau, n_eig):
    """Unnormalised target e^G w1 w2 at t = x + tau and unit direction n in
    the shear eigenframe (n_eig: (..., 3)); d-hat in that frame is P['dh']."""
    T, mu, m, sg = P["T"], P["mu"], P["m"], P["sgn"]
    t = P["x"] + tau
    E = T * t + mu
    p2 = np.maximum(E * E - m * m, 0.0)
    p = np.sqrt(p2)
    fe = 1.0 / (1.0 + sg * np.exp(-t))
    G = np.log(np.maximum(p * E, 1e-300)) - tau + np.log(fe)
    y = n_eig @ P["dh"]
    W = P["l1"] * n_eig[..., 0] ** 2 + P["l2"] * n_eig[..., 1] ** 2 + P["l3"] * n_eig[..., 2] ** 2
    w1 = np.clip((E * P["d0"] + p * P["dn"] * y) / (E * P["D"]), 0, 1) if P["D"] > 0 else 0 * E
    dfs = fe * p2 * W * ps_of(P, E)
    dfb = fe * qb_of(P, E)
    w2 = np.clip(0.5 * (1 + dfs + dfb), 0, 1)
    return np.exp(G) * w1 * w2, W, y


def envelope(P, tab, k, tau, W, y):
    return np.exp(tab["c"][k] + tab["s"][k] * tau) * (1 + tab["kap"][k] * W) * (1 + tab["beta"][k] * y)
