"""MUSIC surface input: 34-float binary records and the boost-invariant reader.

The drop-in's reader (iss_amd/csrc/host/surface_io.cpp, built here into a small
probe with g++, no GPU) reads the same cells from a text surface and from its
34-float binary form (src/readindata.cpp:395-546 boost invariant, 626-765
3+1D; record layout :646-689), bit for bit when the text holds exactly the
floats of the binary file.  The oracle's reader and, where it was built here,
the reference's own reader (oracle/_ref) read the binary files to the same
cells.  Boost-invariant surfaces (hydro_mode 1) get eta_s = 0 and dsigma_eta
= 0 whatever the file holds; cells at T <= 0.01 GeV are dropped in every path.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import oracle_binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "iss_amd", "csrc", "host")
HBARC = 0.197327053


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    so = os.path.join(str(d), "libprobe.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + HOST, "-I" + os.path.join(ROOT, "include"),
                    "-o", so, os.path.join(ROOT, "tests", "reader_probe.cpp"), os.path.join(HOST, "surface_io.cpp"),
                    "-lz"], check=True)
    lib = C.CDLL(so)
    lib.probe_read.restype = C.c_long
    lib.probe_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int,
                               C.c_void_p, C.c_long]
    return lib


def _read(probe, work, surface, mode, bulk, rhob, diff, eos="", iEOS=0):
    out = np.zeros((4096, 34), dtype=np.float32)
    n = probe.probe_read(work.encode(), surface.encode(), mode, bulk, rhob, diff, eos.encode(), iEOS,
                         out.ctypes.data_as(C.c_void_p), out.shape[0])
    assert n >= 0, n
    return out[:n].copy()


def _music_input(work, iEOS, bulk, rhob, diff, binary):
    with open(os.path.join(work, "music_input"), "w") as f:
        f.write(f"EOS_to_use {iEOS}\nInclude_Bulk_Visc_Yes_1_No_0  {bulk}\nInclude_Rhob_Yes_1_No_0  {rhob}\n"
                f"turn_on_baryon_diffusion  {diff}\nfreeze_surface_in_binary  {binary}\nEndOfData\n")


def _binary_records(cols, bulk, rhob, diff):
    """Text columns (file units) -> 34-float binary records: 28 fields, then
    bulk Pi, n_B and q^mu in their slots (zero where the text has none)."""
    n = cols.shape[0]
    rec = np.zeros((n, 34), dtype=np.float32)
    rec[:, :28] = cols[:, :28]
    k = 28
    if bulk:
        rec[:, 28] = cols[:, k]
        k += 1
    if rhob:
        rec[:, 29] = cols[:, k]
        k += 1
    if diff:
        rec[:, 30:34] = cols[:, k:k + 4]
    return rec


def _write_text(path, cols):
    # %.17g of float32 values: the text parses back to exactly those floats
    with open(path, "w") as f:
        for row in cols.astype(np.float32).astype(np.float64):
            f.write(" ".join("%.17g" % v for v in row) + "\n")


def _synthetic_cols(n, seed):
    """A surface in MUSIC file units with bulk, n_B and q^mu columns (34 text
    columns); non-zero eta_s and dsigma_eta; a few cold cells (T < 0.01 GeV)."""
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 34))
    c[:, 0] = rng.uniform(0.6, 15, n)
    c[:, 1:3] = rng.uniform(-10, 10, (n, 2))
    c[:, 3] = rng.uniform(-3, 3, n)
    c[:, 4] = rng.uniform(0.5, 2.0, n)
    c[:, 5:8] = rng.uniform(-0.3, 0.3, (n, 3))
    c[:, 9:12] = rng.uniform(-1.5, 1.5, (n, 3))
    c[:, 8] = np.sqrt(1 + (c[:, 9:12] ** 2).sum(1))
    T = rng.uniform(0.135, 0.16, n) / HBARC
    T[::17] = 0.005 / HBARC
    e = rng.uniform(0.2, 0.6, n) / HBARC
    c[:, 12], c[:, 13] = e, T
    c[:, 14] = rng.uniform(0, 0.3, n) / HBARC
    c[:, 17] = rng.uniform(2.0, 6.0, n)
    c[:, 18:28] = rng.normal(0, 0.02, (n, 10)) / HBARC
    c[:, 28] = rng.uniform(-0.05, 0, n) / HBARC
    c[:, 29] = rng.uniform(0, 0.1, n)
    c[:, 30:34] = rng.normal(0, 0.01, (n, 4))
    return c


CASES = {name: c for name, c in ob.REF_CASES.items()}


def _case_cols(name):
    cfg = ob.case_config(name)
    toks = open(cfg["surface"]).read().split()
    per = 28 + 1 + cfg["rhob"]
    return np.array(toks[: len(toks) // per * per], dtype=np.float64).reshape(-1, per), cfg


@pytest.mark.parametrize("mode", [2, 1], ids=["3+1D", "boost_invariant"])
@pytest.mark.parametrize("name", list(CASES))
def test_reference_cases_binary_equals_text(probe, name, mode):
    cols, cfg = _case_cols(name)
    with tempfile.TemporaryDirectory() as w:
        # the original text, and the same values as exact-float text
        _music_input(w, cfg["iEOS"], 1, cfg["rhob"], 0, 0)
        open(os.path.join(w, "orig.dat"), "w").write(open(cfg["surface"]).read())
        _write_text(os.path.join(w, "exact.dat"), cols)
        t_orig = _read(probe, w, "orig.dat", mode, 1, cfg["rhob"], 0, cfg["eos_file"], cfg["iEOS"])
        t_exact = _read(probe, w, "exact.dat", mode, 1, cfg["rhob"], 0, cfg["eos_file"], cfg["iEOS"])
        _music_input(w, cfg["iEOS"], 1, cfg["rhob"], 0, 1)
        _binary_records(cols, 1, cfg["rhob"], 0).tofile(os.path.join(w, "surf.bin"))
        b = _read(probe, w, "surf.bin", mode, 1, cfg["rhob"], 0, cfg["eos_file"], cfg["iEOS"])
        assert b.shape == t_exact.shape == (1, 34)
        assert np.array_equal(b.view(np.uint32), t_exact.view(np.uint32))
        assert np.allclose(b, t_orig, rtol=3e-7, atol=1e-30)
        # the oracle's binary reader (pinned to the reference's text reader) agrees
        o = ob.Oracle(afterburner=cfg["afterburner"])
        assert o.read_surface(os.path.join(w, "surf.bin"), binary=1, boost_inv=int(mode == 1), bulk=1,
                              rhob=cfg["rhob"], eos_file=cfg["eos_file"], iEOS=cfg["iEOS"]) >= 0
        lab = np.zeros((1, 34), dtype=np.float32)
        ob.lib().orc_get_cells_lab(o.h, lab.ctypes.data_as(C.c_void_p))
        assert np.allclose(b, lab, rtol=1e-6, atol=1e-12)
        # the reference's own reader of the binary file
        r = ob.ref_lib()
        if r is not None:
            n = r.ref_read_surface(w.encode(), ob.TABLE_DIR.encode(), b"surf.bin", mode, cfg["afterburner"], 1, 1,
                                   cfg["rhob"], 0)
            assert n == 1
            ref = np.zeros((1, 34), dtype=np.float32)
            r.ref_get_cells(ref.ctypes.data_as(C.c_void_p))
            assert np.allclose(b, ref, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("mode", [2, 1], ids=["3+1D", "boost_invariant"])
def test_synthetic_surface_all_columns(probe, mode):
    cols = _synthetic_cols(300, seed=5)
    eos = os.path.join(ob.TABLE_DIR, "EOS_tables", "HRGNEOS_BQS-urqmd_v3.3+.dat")
    with tempfile.TemporaryDirectory() as w:
        _music_input(w, 14, 1, 1, 1, 0)
        _write_text(os.path.join(w, "surf.dat"), cols)
        t = _read(probe, w, "surf.dat", mode, 1, 1, 1, eos, 14)
        _music_input(w, 14, 1, 1, 1, 1)
        _binary_records(cols, 1, 1, 1).tofile(os.path.join(w, "surf.bin"))
        b = _read(probe, w, "surf.bin", mode, 1, 1, 1, eos, 14)
        keep = cols[:, 13].astype(np.float32) * HBARC > 0.01
        assert b.shape[0] == t.shape[0] == int(keep.sum()) < cols.shape[0]
        assert np.array_equal(b.view(np.uint32), t.view(np.uint32))
        kept = cols[keep].astype(np.float32)
        # position, dsigma, u straight from the file (boost invariant: eta, dsigma_eta = 0)
        for f in (0, 1, 2, 4, 5, 6):
            assert np.array_equal(b[:, f], kept[:, f])
        if mode == 1:
            assert np.all(b[:, 3] == 0) and np.all(b[:, 7] == 0)
        else:
            assert np.array_equal(b[:, 3], kept[:, 3]) and np.array_equal(b[:, 7], kept[:, 7])
        # n_B and q^i keep file units (q^0 is regulated to u.q = 0); bulk Pi is scaled by hbar c
        assert np.array_equal(b[:, 15], kept[:, 29])
        assert np.array_equal(b[:, 31:34], kept[:, 31:34])
        assert np.allclose(b[:, 29], kept[:, 28] * HBARC, rtol=1e-6)
        o = ob.Oracle(afterburner=1)
        assert o.read_surface(os.path.join(w, "surf.bin"), binary=1, boost_inv=int(mode == 1), bulk=1, rhob=1,
                              diff=1, eos_file=eos, iEOS=14) >= 0
        lab = np.zeros(b.shape, dtype=np.float32)
        ob.lib().orc_get_cells_lab(o.h, lab.ctypes.data_as(C.c_void_p))
        assert np.allclose(b, lab, rtol=1e-5, atol=1e-9)
        r = ob.ref_lib()
        if r is not None:
            n = r.ref_read_surface(w.encode(), ob.TABLE_DIR.encode(), b"surf.bin", mode, 1, 1, 1, 1, 1)
            assert n == b.shape[0]
            ref = np.zeros(b.shape, dtype=np.float32)
            r.ref_get_cells(ref.ctypes.data_as(C.c_void_p))
            assert np.allclose(b, ref, rtol=1e-5, atol=1e-9)
