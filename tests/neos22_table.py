"""A synthetic 22-moment NEoS delta-f table and a table directory that holds it.

The reference reads bulk_deltaf_kind 20 coefficients from
deltaf_tables/<afterburner>/NEoSBQS_22mom_deltafCoeff.dat (src/FSSW.cpp:
1340-1376): one header line, then 200 x 200 rows of 8 columns, e outer and n_B
inner, n_B starting at 0 with its own spacing per e row; columns 2..7 are the
six coefficients get22momNEOSBQSCoefficients interpolates (src/FSSW.cpp:
1490-1543).  That file is not shipped with the reference's tables, so the tests
build one of the same layout with smooth, distinct columns of realistic size:
column 2 multiplies the shear tensor (fm^3/GeV^3, near 1/(2 T^2 (e+P))), columns
3..7 the bulk pressure (c1, c2 and the B, S, Q coefficients).
"""
from __future__ import annotations

import os

import numpy as np

from tests import oracle_binding as ob

NE = NB = 200
FILE = "NEoSBQS_22mom_deltafCoeff.dat"


def table() -> np.ndarray:
    """[40000][8] rows (e, n_B, c0..c5)."""
    e = 0.01 * (np.arange(NE) + 1.0)
    nbmax = 0.75 * e ** 0.85
    rows = np.zeros((NE, NB, 8))
    for i in range(NE):
        nb = np.arange(NB) * (nbmax[i] / (NB - 1))
        x = nb / nbmax[i]
        T = 0.15 * (e[i] / 0.4) ** 0.25
        rows[i, :, 0] = e[i]
        rows[i, :, 1] = nb
        rows[i, :, 2] = 1.0 / (2.0 * T * T * 1.3 * e[i]) * (1.0 + 0.1 * x)
        rows[i, :, 3] = 2.0 + 0.5 * x + 0.3 * e[i]
        rows[i, :, 4] = 0.8 + 0.2 * np.sin(3.0 * e[i]) - 0.1 * x
        rows[i, :, 5] = 0.4 * (1.0 + e[i]) + 0.2 * x
        rows[i, :, 6] = 0.1 + 0.05 * x * x
        rows[i, :, 7] = -0.05 * (1.0 + x)
    return rows.reshape(NE * NB, 8)


def write(path: str, rows: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("# e(GeV/fm^3) rho_B(1/fm^3) c0 c1 c2 c3 c4 c5 (synthetic)\n")
        np.savetxt(f, rows, fmt="%.17e")


def table_dir(tmp: str, rows: np.ndarray | None = None) -> str:
    """A copy of the vendored table directory (symbolic links) whose urqmd
    delta-f folder also holds the synthetic 22-moment table."""
    rows = table() if rows is None else rows
    src = ob.TABLE_DIR
    dst = os.path.join(tmp, "tables")
    os.makedirs(dst, exist_ok=True)
    for name in os.listdir(src):
        if name != "deltaf_tables":
            os.symlink(os.path.join(src, name), os.path.join(dst, name))
    dsrc, ddst = os.path.join(src, "deltaf_tables"), os.path.join(dst, "deltaf_tables")
    os.makedirs(ddst)
    for name in os.listdir(dsrc):
        if name != "urqmd":
            os.symlink(os.path.join(dsrc, name), os.path.join(ddst, name))
    usrc, udst = os.path.join(dsrc, "urqmd"), os.path.join(ddst, "urqmd")
    os.makedirs(udst)
    for name in os.listdir(usrc):
        os.symlink(os.path.join(usrc, name), os.path.join(udst, name))
    write(os.path.join(udst, FILE), rows)
    return dst
