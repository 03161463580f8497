"""Readers for the iSS data tables (host plumbing for the Python binding).

Formats follow the reference readers: the PDG list with anti-baryons
appended (src/readindata.cpp:971-1116), the chosen-particle list and the mass
sort of the sampler (src/FSSW.cpp:116-162), the HRG EoS tables
(src/readindata.cpp:913-968) and the delta-f coefficient tables
(src/FSSW.cpp:1215-1376).
"""
from __future__ import annotations

import os

import numpy as np

from . import TABLE_DIR

AFTERBURNER_FILES = {
    1: ("pdg-urqmd_v3.3+.dat", "chosen_particles_urqmd_v3.3+.dat", "urqmd"),
    2: ("pdg-SMASH.dat", "chosen_particles_SMASH.dat", "smash"),
    0: ("pdg-s95pv1.dat", "chosen_particles_s95p-v1.dat", "urqmd"),
}


def read_pdg(path: str, quantum_statistics: bool = True) -> list[dict]:
    """PDG list: one particle line followed by its decay lines; anti-baryons
    are appended right after their baryon; sign -1 mesons (BE), +1 baryons
    (FD), 0 without quantum statistics."""
    toks = open(path).read().split()
    out = []
    i = 0
    while i + 12 <= len(toks):
        monval, name = int(toks[i]), toks[i + 1]
        mass, width = float(toks[i + 2]), float(toks[i + 3])
        g, B, S, charm, bottom, giso, Q, ndec = (int(t) for t in toks[i + 4:i + 12])
        i += 12 + 8 * ndec
        p = dict(pid=monval, name=name, mass=mass, width=width, g=g, B=B, S=S, Q=Q)
        out.append(p)
        if B > 0:
            out.append(dict(pid=-monval, name="Anti-" + name, mass=mass, width=width, g=g,
                            B=-B, S=-S, Q=-Q))
    for p in out:
        p["sign"] = (-1 if p["B"] == 0 else 1) if quantum_statistics else 0
    return out


def read_chosen(path: str) -> list[int]:
    out = []
    for line in open(path):
        t = line.split()
        if t:
            out.append(int(float(t[0])))
    return out


def sampling_order(pdg: list[dict], chosen: list[int]) -> list[int]:
    """Chosen Monte-Carlo ids -> first PDG index, then a stable sort by mass."""
    first = {}
    for i, p in enumerate(pdg):
        first.setdefault(p["pid"], i)
    idx = [first[m] for m in chosen if m in first]
    return sorted(idx, key=lambda i: pdg[i]["mass"])  # sorted() is stable


def species_arrays(afterburner: int = 1, quantum_statistics: bool = True,
                   table_dir: str = TABLE_DIR) -> dict:
    pdg_f, chosen_f, _ = AFTERBURNER_FILES[afterburner]
    pdg = read_pdg(os.path.join(table_dir, pdg_f), quantum_statistics)
    order = sampling_order(pdg, read_chosen(os.path.join(table_dir, chosen_f)))
    sp = {k: np.array([pdg[i][k] for i in order]) for k in ("pid", "mass", "g", "B", "S", "Q", "sign")}
    sp["mass"] = sp["mass"].astype(np.float64)
    sp["names"] = [pdg[i]["name"] for i in order]
    return sp


def decay_table(afterburner: int = 1, table_dir: str = TABLE_DIR) -> dict:
    """The PDG list with its decay channels as particle_decay reads it
    (src/particle_decay.cpp:33-170): an anti-baryon entry after every baryon
    with B = 1, its daughters conjugated unless neutral in B, Q and S; stable
    when the first channel has one daughter."""
    toks = open(os.path.join(table_dir, AFTERBURNER_FILES[afterburner][0])).read().split()
    ent, chans = [], []
    i = 0
    while i + 12 <= len(toks):
        mon = int(toks[i])
        mass, width = float(toks[i + 2]), float(toks[i + 3])
        B, S, Q, nd = int(toks[i + 5]), int(toks[i + 6]), int(toks[i + 10]), int(toks[i + 11])
        i += 12
        ch = []
        for _ in range(nd):
            c = toks[i:i + 8]
            i += 8
            ch.append((int(c[1]), float(c[2]), [int(x) for x in c[3:8]]))
        ent.append(dict(monval=mon, mass=mass, width=width, B=B, S=S, Q=Q, ch=ch))
        if B == 1:
            conj = []
            for npart, br, parts in ch:
                new = []
                for d in parts:
                    if d == 0:
                        new.append(0)
                        continue
                    q = next((e for e in ent if e["monval"] == d), None)
                    neutral = q is not None and q["B"] == 0 and q["Q"] == 0 and q["S"] == 0
                    new.append(d if neutral else -d)
                conj.append((npart, br, new))
            ent.append(dict(monval=-mon, mass=mass, width=width, B=-B, S=-S, Q=-Q, ch=conj))
    out = {k: [] for k in ("monval", "mass", "width", "stable", "first_channel", "n_channels", "baryon",
                           "strange", "charge")}
    npart, br, dau = [], [], []
    for e in ent:
        out["monval"].append(e["monval"])
        out["mass"].append(e["mass"])
        out["width"].append(e["width"])
        out["stable"].append(1 if e["ch"] and e["ch"][0][0] == 1 else 0)
        out["first_channel"].append(len(npart))
        out["n_channels"].append(len(e["ch"]))
        out["baryon"].append(e["B"])
        out["strange"].append(e["S"])
        out["charge"].append(e["Q"])
        for c in e["ch"]:
            npart.append(c[0])
            br.append(c[1])
            dau.append(c[2])
    res = {k: np.array(v, dtype=np.float64 if k in ("mass", "width") else np.int32) for k, v in out.items()}
    res["npart"] = np.array(npart, dtype=np.int32)
    res["branching_ratio"] = np.array(br, dtype=np.float64)
    res["daughters"] = np.array(dau, dtype=np.int32).reshape(-1, 5)
    return res


def bin_tables(table_dir: str = TABLE_DIR) -> dict:
    """Quadrature grids of the smooth spectra (src/iSS.cpp:151-154): Gauss pT
    (15 nodes, GeV) and phi (48 nodes) with weights, and the uniform eta_s
    table (-4..4, step 0.1) with its weights."""
    d = os.path.join(table_dir, "bin_tables")
    pt = np.loadtxt(os.path.join(d, "pT_gauss_table.dat"))
    ph = np.loadtxt(os.path.join(d, "phi_gauss_table.dat"))
    et = np.loadtxt(os.path.join(d, "eta_uni_table.dat"))
    return dict(pT=pt[:, 0].copy(), pT_weight=pt[:, 1].copy(), phi=ph[:, 0].copy(), phi_weight=ph[:, 1].copy(),
                eta=et[:, 0].copy(), eta_weight=et[:, 1].copy())


def read_eos(iEOS: int, afterburner: int, table_dir: str = TABLE_DIR) -> dict:
    """HRG EoS rows {e, nB, P, T, muB, muS, muQ} for iEOS 9/91 (1-D in e) or
    12/14 (200 n_B per e)."""
    if iEOS in (9, 91):
        stem = "HRGEOS_PST-"
    elif iEOS == 12:
        stem = "HRGNEOS_B-"
    elif iEOS == 14:
        stem = "HRGNEOS_BQS-"
    else:
        raise ValueError("EoS regulation needs iEOS 9, 91, 12 or 14")
    if iEOS == 9:
        afterburner = 1
    if iEOS == 91:
        afterburner = 2
    suffix = {1: "urqmd_v3.3+.dat", 2: "SMASH.dat"}[afterburner]
    raw = np.loadtxt(os.path.join(table_dir, "EOS_tables", stem + suffix), comments="#")
    rows = np.zeros((raw.shape[0], 7))
    if iEOS in (9, 91):
        rows[:, 0], rows[:, 2], rows[:, 3] = raw[:, 0], raw[:, 1], raw[:, 3]
    else:
        rows[:, :raw.shape[1]] = raw
    return {"iEOS": iEOS, "rows": rows}


def read_qmu_table(table_dir: str = TABLE_DIR) -> np.ndarray:
    """Coefficients_RTA_diffusion.dat as FSSW::load_deltaf_qmu_coeff_table
    reads it (src/FSSW.cpp:1546-1566): rows (T, mu_B, D), mu_B outer and T
    inner; returned as [150 T][100 mu_B]."""
    v = np.loadtxt(os.path.join(table_dir, "deltaf_tables", "Coefficients_RTA_diffusion.dat"))
    return np.ascontiguousarray(v[:, 2].reshape(100, 150).T)


def read_kind11(afterburner: int = 1, table_dir: str = TABLE_DIR) -> dict:
    """OSU 14-moment bulk coefficients c0, c1, c2 on a (T, muB) grid, all three
    loaded (the reference's c0 read fails, src/FSSW.cpp:1246-1278)."""
    sub = "smash_box" if afterburner == 2 else "urqmd"
    d = os.path.join(table_dir, "deltaf_tables", sub)
    out = {}
    for k in ("c0", "c1", "c2"):
        with open(os.path.join(d, k + ".dat")) as f:
            nT = int(f.readline())
            nmu = int(f.readline())
            f.readline()
            v = np.loadtxt(f)
        # file order: mu outer, T inner
        out[k] = v[:, 2].reshape(nmu, nT).T.copy().ravel()
        T, mu = v[:, 0].reshape(nmu, nT), v[:, 1].reshape(nmu, nT)
        out.update(nT=nT, nmu=nmu, T0=T[0, 0], mu0=mu[0, 0], dT=T[0, 1] - T[0, 0],
                   dmu=mu[1, 0] - mu[0, 0])
    return out


def read_neos(kind: int, afterburner: int = 1, table_dir: str = TABLE_DIR) -> np.ndarray:
    sub = "smash" if afterburner == 2 else "urqmd"
    name = "NEoSBQS_CE_deltafCoeff.dat" if kind == 21 else "NEoSBQS_22mom_deltafCoeff.dat"
    a = np.loadtxt(os.path.join(table_dir, "deltaf_tables", sub, name), comments="#")
    assert a.shape[0] == 40000
    return a


def read_music_surface(path: str, binary: bool = False, boost_invariant: bool = False,
                       bulk: int = 0, rhob: int = 0, diff: int = 0) -> np.ndarray:
    """MUSIC freeze-out surface as lab-frame records (34 floats, FO_surf order,
    src/readindata.cpp:395-546, 626-765): hbarC factors applied, P from
    (e+P)/T, cells with T <= 0.01 GeV dropped.  No regulation."""
    hc = 0.197327053
    if binary:
        raw = np.fromfile(path, dtype=np.float32)
        raw = raw[: raw.size // 34 * 34].reshape(-1, 34).astype(np.float64)
        cols = raw
        has_bulk = has_rhob = has_diff = True
    else:
        toks = open(path).read().split()
        per = 28 + (bulk == 1) + (rhob == 1) + 4 * (diff == 1)
        n = len(toks) // per
        cols = np.array(toks[: n * per], dtype=np.float64).reshape(n, per)
        has_bulk, has_rhob, has_diff = bulk == 1, rhob == 1, diff == 1
    n = cols.shape[0]
    out = np.zeros((n, 34), dtype=np.float64)
    out[:, 0:12] = cols[:, 0:12]
    if boost_invariant:
        out[:, 3] = 0.0
        out[:, 7] = 0.0
    out[:, 12] = cols[:, 12] * hc          # e
    out[:, 13] = cols[:, 13] * hc          # T
    out[:, 16] = cols[:, 14] * hc          # muB
    out[:, 17] = cols[:, 15] * hc          # muS
    out[:, 18] = cols[:, 16] * hc          # muQ
    out[:, 14] = cols[:, 17] * out[:, 13] - out[:, 12]  # P = (e+P)/T * T - e
    out[:, 19:29] = cols[:, 18:28] * hc    # pi^{mu nu}
    k = 28
    if has_bulk:
        out[:, 29] = cols[:, k] * hc
        k += 1
    if has_rhob:
        out[:, 15] = cols[:, k]
        k += 1
    if has_diff:
        out[:, 30:34] = cols[:, k:k + 4]
    out = out.astype(np.float32)
    return out[out[:, 13] > 0.01]
