"""The drop-in iSS.e run as the reference's integration tests run it
(tests/RunAutoTest*.sh): copy music_input_<EoS> to music_input, run
`iSS.e <parameter file> <work dir> <surface file>`, then apply
tests/TestOutputFiles.py's criterion to checkReconstructedTmunu.dat:
mean |diff| over the 16 T^{mu nu} components and the net-baryon row < 1e-3.

The ideal tests' parameter file names bulk_deltaf_kind = 20, whose 22-moment
table is not part of this snapshot; ideal cells carry no Pi and no pi, so the
kind is overridden to 1 there (it cannot change those samples)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "iss_amd", "iSS.e")
FIX = os.path.join(ROOT, "tests", "golden", "ref_fixtures")

RUNS = {  # name: (music_input, parameter file, surface)
    "Ideal1": ("music_input_91", "iSS_parameters_ideal.dat", "testIdealOneFluidCell1.dat"),
    "Ideal2": ("music_input_9", "iSS_parameters_ideal.dat", "testIdealOneFluidCell2.dat"),
    "Ideal3": ("music_input_12", "iSS_parameters_ideal.dat", "testIdealOneFluidCell3.dat"),
    "Ideal4": ("music_input_14", "iSS_parameters_ideal.dat", "testIdealOneFluidCell4.dat"),
    "Viscous1": ("music_input_91", "iSS_parameters_CEdeltaf.dat", "testViscousOneFluidCell1.dat"),
    "Viscous2": ("music_input_9", "iSS_parameters_CEdeltaf.dat", "testViscousOneFluidCell2.dat"),
}


def _run(tmp_path, name, extra=(), surface_text=None, music_text=None):
    mi, params, surf = RUNS[name]
    work = tmp_path / "tests"
    work.mkdir()
    if music_text is None:
        shutil.copy(os.path.join(FIX, mi), work / "music_input")
    else:
        (work / "music_input").write_text(music_text)
    if surface_text is None:
        shutil.copy(os.path.join(FIX, surf), work / surf)
    else:
        (work / surf).write_text(surface_text)
    shutil.copy(os.path.join(FIX, params), tmp_path / params)
    os.symlink(os.path.join(ROOT, "iss_amd", "tables"), tmp_path / "iSS_tables")
    args = [EXE, params, "tests", surf, "randomSeed=7", "number_of_repeated_sampling=1000"]
    if "ideal" in params:
        args.append("bulk_deltaf_kind=1")
    args += list(extra)
    r = subprocess.run(args, cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:]
    return r


@pytest.mark.parametrize("name", list(RUNS))
def test_dropin_cli_passes_reference_criterion(tmp_path, name):
    _run(tmp_path, name)
    data = np.loadtxt(tmp_path / "checkReconstructedTmunu.dat")
    assert data.shape == (17, 3)
    res = np.mean(np.abs(data[:, 2]))
    print(name, "mean |diff|", res)
    assert res < 1e-3


def test_dropin_binary_output_format(tmp_path):
    """particle_samples.bin: per event int32 N, then per hadron int32 pid and
    9 float32 (mass, t, x, y, z, E, px, py, pz) (src/FSSW.cpp:529-561)."""
    _run(tmp_path, "Ideal1", ["use_binary_format=1", "number_of_repeated_sampling=20",
                              "perform_checks=0"])
    raw = (tmp_path / "particle_samples.bin").read_bytes()
    off, nev, ntot = 0, 0, 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        off += 4
        for _ in range(n):
            pid = struct.unpack_from("<i", raw, off)[0]
            m, t, x, y, z, E, px, py, pz = struct.unpack_from("<9f", raw, off + 4)
            assert pid != 0 and E > 0 and abs(E * E - px * px - py * py - pz * pz - m * m) < 1e-3 * E * E
            off += 40
        nev += 1
        ntot += n
    assert nev == 20 and off == len(raw) and ntot > 20 * 10000


def test_dropin_oscar_and_gzip(tmp_path):
    import gzip
    _run(tmp_path, "Ideal2", ["use_OSCAR_format=1", "number_of_repeated_sampling=3",
                              "perform_checks=0"])
    lines = (tmp_path / "OSCAR.DAT").read_text().splitlines()
    assert lines[0].startswith("OSC1997A")
    hdr = [l for l in lines if len(l.split()) == 4]
    assert len(hdr) >= 3
    (tmp_path / "tests").rename(tmp_path / "t_old")
    for f in ("iSS_tables", "iSS_parameters_ideal.dat"):
        os.remove(tmp_path / f)
    _run(tmp_path, "Ideal2", ["use_gzip_format=1", "number_of_repeated_sampling=3",
                              "perform_checks=0"])
    with gzip.open(tmp_path / "particle_samples.gz", "rt") as g:
        first = g.readline().split()
        n = int(first[0])
        rec = g.readline().split()
    assert n > 10000 and len(rec) == 10


@pytest.mark.parametrize("name", ["Viscous1", "Ideal2"])
def test_dropin_learned_proposal(tmp_path, name):
    """sampling_proposal = 1: the first events come from the plain sampler,
    the network is trained on their hadrons and samples the rest; the
    reference's T^{mu nu} criterion holds for the whole run (the Fermi-Dirac
    table quirk off, so every species goes through the envelope)."""
    r = _run(tmp_path, name, ["sampling_proposal=1", "fermion_cdf2_reference_quirk=0",
                              "number_of_repeated_sampling=1000"])
    assert "learned proposal trained" in r.stdout, r.stdout[-2000:]
    data = np.loadtxt(tmp_path / "checkReconstructedTmunu.dat")
    res = np.mean(np.abs(data[:, 2]))
    print(name, "learned mean |diff|", res)
    assert res < 1e-3


def _table_blocks(path, rows):
    a = np.loadtxt(path)
    return a.reshape(-1, rows, a.shape[1])


@pytest.mark.parametrize("historic", [0, 1])
def test_dropin_calculate_vn_matches_oracle(tmp_path, historic):
    """MC_sampling = 0, calculate_vn = 1: the reference's smooth-spectra
    outputs (src/emissionfunction.cpp:1129-1218, historic :1031-1120) from
    one GPU pass, against the oracle's lab-frame integral and flows on the
    same regulated cell.  Bar: the float rest-frame record (5e-5 of the row
    maximum; the files hold 9 significant digits)."""
    order = 4
    _run(tmp_path, "Viscous1", ["MC_sampling=0", "calculate_vn=1", "perform_checks=0",
                                f"calculate_vn_to_order={order}",
                                f"use_historic_flow_output_format={historic}"])
    cfg = ob.case_config("Viscous1")
    orc = ob.Oracle(afterburner=cfg["afterburner"], shear=cfg["shear"], bulk=cfg["bulk"],
                    bulk_kind=cfg["bulk_kind"])
    orc.read_surface(cfg["surface"], bulk=1, rhob=cfg["rhob"], eos_file=cfg["eos_file"], iEOS=cfg["iEOS"])
    from iss_amd import tables
    b = tables.bin_tables()
    sp = orc.species()
    parts = orc.particles()
    work = tmp_path / "tests"
    dn = _table_blocks(work / "dN_pTdpTdphidy.dat", b["phi"].size)
    assert dn.shape == (parts["pid"].size, b["phi"].size, b["pT"].size)
    checked = 0
    for pid in (211, -211, 321, 2212, -2212, 3334):
        s = int(np.nonzero(sp["pid"] == pid)[0][0])
        ref = orc.spectra(s, b["pT"], b["phi"], b["eta"], b["eta_weight"])[0]
        n = int(np.nonzero(parts["pid"] == pid)[0][0])
        got = dn[n].T
        rowmax = np.abs(ref).max(axis=1, keepdims=True)
        assert np.max(np.abs(got - ref) / rowmax) < 5e-5, pid
        d_ref, i_ref = ob.flows(ref, b["pT"], b["pT_weight"], b["phi"], b["phi_weight"], sp["mass"][s], order)
        if historic:
            continue
        d_got = np.loadtxt(work / f"thermal_{pid}_vndata.dat")
        i_got = np.loadtxt(work / f"thermal_{pid}_integrated__vndata.dat")
        assert d_got.shape == (b["pT"].size, 3 + 3 * order) and i_got.shape == (order + 1, 6)
        np.testing.assert_allclose(d_got[:, :2], d_ref[:, :2], rtol=1e-7)
        np.testing.assert_allclose(d_got[:, 2], d_ref[:, 2], rtol=5e-5)
        assert np.max(np.abs(d_got[:, 3:] - d_ref[:, 3:])) < 1e-4
        np.testing.assert_allclose(i_got[:, 3:], i_ref[:, 3:], atol=1e-5)
        checked += 1
    if historic:
        txt = (work / "v2data.dat").read_text()
        assert txt.count("# Output for particle:") == parts["pid"].size
        inte = (work / "v2data-inte.dat").read_text()
        assert inte.count("# For:") == parts["pid"].size
    else:
        assert checked == 6
    # non-chosen list entries are zero tables
    chosen = set(sp["pid"].tolist())
    for n in range(parts["pid"].size):
        if parts["pid"][n] not in chosen:
            assert not dn[n].any()


def _read_bin(path):
    raw = path.read_bytes()
    off, evs = 0, []
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        off += 4
        ev = []
        for _ in range(n):
            pid = struct.unpack_from("<i", raw, off)[0]
            vals = struct.unpack_from("<9f", raw, off + 4)
            ev.append((pid,) + vals)
            off += 40
        evs.append(ev)
    return evs


def test_dropin_perform_decays(tmp_path):
    """perform_decays = 1: the sampled hadrons are replaced by their stable
    descendants (FSSW::perform_resonance_feed_down, src/FSSW.cpp:1746-1779)
    before the writer; charge is conserved on average against the primary
    sample of the same seed (decays conserve it, the vanishing 4-body and
    '-2' channels of the table aside)."""
    _run(tmp_path, "Ideal2", ["use_binary_format=1", "number_of_repeated_sampling=20", "perform_checks=0",
                              "perform_decays=1"])
    dec = _read_bin(tmp_path / "particle_samples.bin")
    (tmp_path / "tests").rename(tmp_path / "t_dec")
    for f in ("iSS_tables", "iSS_parameters_ideal.dat", "particle_samples.bin"):
        os.remove(tmp_path / f)
    _run(tmp_path, "Ideal2", ["use_binary_format=1", "number_of_repeated_sampling=20", "perform_checks=0"])
    prim = _read_bin(tmp_path / "particle_samples.bin")
    assert len(dec) == len(prim) == 20
    t = ob.decay_table(1)
    stable = set(t["monval"][t["stable"] == 1].tolist())
    q = {int(m): int(t["Q"][i]) for i, m in reversed(list(enumerate(t["monval"])))}
    nd = sum(len(e) for e in dec)
    npr = sum(len(e) for e in prim)
    assert nd > 1.3 * npr
    assert all(h[0] in stable for e in dec for h in e)
    for e in dec[:3]:
        for h in e:
            pid, m, tt, x, y, z, E, px, py, pz = h
            assert abs(E * E - px * px - py * py - pz * pz - m * m) < 1e-3 * E * E + 1e-4
    qd = np.mean([sum(q[h[0]] for h in e) for e in dec])
    qp = np.mean([sum(q.get(h[0], 0) for h in e) for e in prim])
    print("hadrons primary", npr, "decayed", nd, "mean charge", qp, qd)
    assert abs(qd - qp) < 5 * np.sqrt(npr / 20) / np.sqrt(20) + 1


def test_dropin_diffusion_spectra_match_oracle(tmp_path):
    """include_deltaf_diffusion = 1 with a surface that carries n_B and q^mu
    (turn_on_baryon_diffusion 1 in music_input): the drop-in reads
    Coefficients_RTA_diffusion.dat (src/FSSW.cpp:191-195) and its smooth
    spectra match the oracle's on the same file.  The cell is at rest with a
    transverse q, so the diffusion term shows as a directed (cos phi)
    modulation of p and pbar that the diffusion-free spectrum lacks."""
    name = "Viscous1"
    mi, _, surf = RUNS[name]
    base = open(os.path.join(FIX, surf)).read().split()
    line = " ".join(base[:28] + [base[28], "0.05", "0.0", "0.0006", "-0.0003", "0.0"])
    music = open(os.path.join(FIX, mi)).read()
    music = music.replace("Include_Rhob_Yes_1_No_0  0", "Include_Rhob_Yes_1_No_0  1")
    music = music.replace("turn_on_baryon_diffusion  0", "turn_on_baryon_diffusion  1")
    _run(tmp_path, name, ["MC_sampling=0", "calculate_vn=1", "perform_checks=0", "calculate_vn_to_order=2",
                          "include_deltaf_diffusion=1", "turn_on_rhob=1", "turn_on_diff=1"],
         surface_text=line + "\n", music_text=music)
    cfg = ob.case_config(name)
    orc = {}
    for d in (0, 1):
        orc[d] = ob.Oracle(afterburner=cfg["afterburner"], shear=cfg["shear"], bulk=cfg["bulk"],
                           bulk_kind=cfg["bulk_kind"], diffusion=d)
        orc[d].read_surface(str(tmp_path / "tests" / surf), bulk=1, rhob=1, diff=1, eos_file=cfg["eos_file"],
                            iEOS=cfg["iEOS"])
    from iss_amd import tables
    b = tables.bin_tables()
    sp = orc[1].species()
    parts = orc[1].particles()
    dn = _table_blocks(tmp_path / "tests" / "dN_pTdpTdphidy.dat", b["phi"].size)
    for pid in (211, 2212, -2212):
        s = int(np.nonzero(sp["pid"] == pid)[0][0])
        ref = orc[1].spectra(s, b["pT"], b["phi"], b["eta"], b["eta_weight"])[0]
        ref0 = orc[0].spectra(s, b["pT"], b["phi"], b["eta"], b["eta_weight"])[0]
        got = dn[int(np.nonzero(parts["pid"] == pid)[0][0])].T
        rowmax = np.abs(ref).max(axis=1, keepdims=True)
        err = np.max(np.abs(got - ref) / rowmax)
        eff = np.max(np.abs(ref0 - ref) / rowmax)
        v1 = np.loadtxt(tmp_path / "tests" / f"thermal_{pid}_integrated__vndata.dat")[1]
        print(pid, "row error", err, "diffusion effect", eff, "integrated v1 row", v1)
        assert err < 5e-5, pid
        if pid != 211:
            assert eff > 1e-2, pid


def test_dropin_local_charge_conservation(tmp_path):
    """local_charge_conservation = 1 (src/FSSW.cpp:931-938, 1035-1048): no
    negative species is sampled on its own; each positive hadron is followed
    by its antiparticle from the same cell, so every event is neutral."""
    _run(tmp_path, "Ideal2", ["use_binary_format=1", "number_of_repeated_sampling=20", "perform_checks=0",
                              "local_charge_conservation=1"])
    evs = _read_bin(tmp_path / "particle_samples.bin")
    assert len(evs) == 20
    t = ob.decay_table(1)
    q = {int(m): int(t["Q"][i]) for i, m in reversed(list(enumerate(t["monval"])))}
    npos = 0
    for e in evs:
        assert sum(q[h[0]] for h in e) == 0
        k = 0
        while k < len(e):
            if q[e[k][0]] > 0:
                npos += 1
                assert e[k + 1][0] == -e[k][0] and e[k + 1][2:5] == e[k][2:5]  # t, x, y
                k += 2
            else:
                assert q[e[k][0]] == 0
                k += 1
    print("positive parents", npos)
    assert npos > 100


def _event_sizes(path):
    """hadrons per event of a binary particle_samples file, without parsing them"""
    raw = path.read_bytes()
    off, sizes = 0, []
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        sizes.append(n)
        off += 4 + 40 * n
    return np.array(sizes, dtype=np.float64)


def test_dropin_sampling_models(tmp_path):
    """dN_dy_sampling_model 1 and 20 (src/FSSW.cpp:250-309) through the
    parameter file: with model 1 every species' count varies by at most one
    between events; with model 20 and dN_dy_sampling_para1 = 1 the event
    multiplicity has the Poisson pass's mean and twice its variance (a sum of
    negative binomials with n = dN, p = 1/2)."""
    # the Ideal2 cell with 1/100 of its d sigma: about 190 hadrons per event
    cols = open(os.path.join(FIX, RUNS["Ideal2"][2])).read().split()
    cols[4] = repr(float(cols[4]) / 100.0)
    line = "  ".join(cols) + "\n"

    def run(sub, n_ev, extra):
        d = tmp_path / sub
        d.mkdir()
        _run(d, "Ideal2", ["use_binary_format=1", f"number_of_repeated_sampling={n_ev}", "perform_checks=0"] + extra,
             surface_text=line)
        return d / "particle_samples.bin"

    ev1 = _read_bin(run("m1", 200, ["dN_dy_sampling_model=1"]))
    assert len(ev1) == 200
    pids = sorted({h[0] for e in ev1 for h in e})
    per = np.array([[sum(1 for h in e if h[0] == p) for p in pids] for e in ev1])
    assert np.all(per.max(axis=0) - per.min(axis=0) <= 1)
    n30 = _event_sizes(run("m30", 2000, ["dN_dy_sampling_model=30"]))
    n20 = _event_sizes(run("m20", 2000, ["dN_dy_sampling_model=20", "dN_dy_sampling_para1=1"]))
    assert n30.size == n20.size == 2000
    mean = n30.mean()
    print("model 30 mean", mean, "var", n30.var(), "model 20 mean", n20.mean(), "var", n20.var())
    assert abs(n20.mean() - mean) < 5 * np.sqrt(2 * mean / n20.size) + 5 * np.sqrt(mean / n30.size)
    assert 1.7 < n20.var() / mean < 2.3
