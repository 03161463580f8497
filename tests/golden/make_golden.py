"""Writes the golden fixtures the oracle is pinned against.

* kn_golden.npz: K_1, K_2, K_3 at the reference's table grid x = 0.5 + 0.05 i
  (i < 7991, src/FSSW.cpp:1609-1615) from scipy.special.kn, an independent
  library (GSL, which fills the reference's table, is absent here).
* ref_draws.npz: draws of the reference's own |p| samplers
  (src/MomentumSamplerShell.cpp, std::mt19937 via src/Random.cpp) and of its
  inverse-CDF cell pick (src/RandomVariable1DArray.cpp), from the reference
  sources compiled in place (oracle/ref/Makefile -> oracle/_ref/libissref.so).
* ref_reader.npz: the reference reader's regulated cells and PDG list for the
  one-cell test surfaces (src/readindata.cpp), same build.

Run from the repository root: python tests/golden/make_golden.py
"""
import ctypes as C
import os
import sys
import tempfile

import numpy as np
from scipy.special import kn

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PSAMPLE_CASES = [  # (m, T, mu, sign): light/heavy bosons and fermions, all 6 tables
    (0.138, 0.150, 0.0, -1), (0.494, 0.135, 0.05, -1), (0.938, 0.140, 0.30, 1),
    (0.938, 0.150, 0.45, 1), (1.5, 0.04, 0.0, -1), (2.2, 0.06, 0.0, 1),
    (5.0, 0.12, 0.0, 1), (3.0, 0.05, 0.0, -1), (7.0, 0.12, 0.0, -1), (1.116, 0.155, 0.2, 1)]


def kn_fixture():
    x = 0.5 + 0.05 * np.arange(7991)
    np.savez_compressed(os.path.join(HERE, "kn_golden.npz"), x=x, K1=kn(1, x), K2=kn(2, x), K3=kn(3, x))


def ref_fixtures():
    from tests import oracle_binding as ob
    r = ob.ref_lib()
    if r is None:
        print("reference build missing: run make -C oracle/ref first")
        return
    draws = {}
    for k, (m, T, mu, s) in enumerate(PSAMPLE_CASES):
        out = np.zeros(4000)
        r.ref_psample(1000 + k, m, T, mu, s, out.size, out.ctypes.data_as(C.c_void_p))
        draws[f"p{k}"] = out
    draws["cases"] = np.array(PSAMPLE_CASES)
    vals = np.random.default_rng(1).random(1000)
    vals[::7] = 0.0
    picks = np.zeros(20000, dtype=np.int64)
    r.ref_cdf_pick(vals.ctypes.data_as(C.c_void_p), vals.size, 3, picks.size,
                   picks.ctypes.data_as(C.c_void_p))
    draws["cdf_vals"] = vals
    draws["cdf_picks"] = picks
    np.savez_compressed(os.path.join(HERE, "ref_draws.npz"), **draws)

    reader = {}
    with tempfile.TemporaryDirectory(prefix="issref_") as work:
        for name, c in ob.REF_CASES.items():
            reader.update(_read_case(r, ob, work, name, c))
    np.savez_compressed(os.path.join(HERE, "ref_reader.npz"), **reader)


def _read_case(r, ob, work, name, c):
    """One fixture surface through the reference reader, in scratch dir `work`."""
    reader = {}
    cfg = ob.case_config(name)
    with open(os.path.join(work, "music_input"), "w") as f:
        f.write(f"EOS_to_use {cfg['iEOS']}\nInclude_Bulk_Visc_Yes_1_No_0  1\n"
                f"Include_Rhob_Yes_1_No_0  {cfg['rhob']}\nturn_on_baryon_diffusion  0\n"
                f"freeze_surface_in_binary  0\nEndOfData\n")
    src = os.path.join(ob.FIXTURES, c["surface"])
    dst = os.path.join(work, c["surface"])
    open(dst, "w").write(open(src).read())
    ab = 2 if c["params"] == "CE" else 1
    n = r.ref_read_surface(work.encode(), ob.TABLE_DIR.encode(), c["surface"].encode(), 2, ab, 1,
                           1, cfg["rhob"], 0)
    cells = np.zeros((n, 34), dtype=np.float32)
    r.ref_get_cells(cells.ctypes.data_as(C.c_void_p))
    reader[name] = cells
    npart = r.ref_num_particles()
    arrs = [np.zeros(npart, dtype=np.int32) for _ in range(7)]
    arrs[1] = np.zeros(npart)
    r.ref_particles(*[a.ctypes.data_as(C.c_void_p) for a in arrs])
    reader[name + "_pdg"] = np.stack([arrs[0], arrs[2], arrs[3], arrs[4], arrs[5], arrs[6]])
    reader[name + "_mass"] = arrs[1]
    return reader


if __name__ == "__main__":
    kn_fixture()
    ref_fixtures()
    print("fixtures written to", HERE)
