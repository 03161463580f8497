"""GPU yield pass vs the CPU oracle (through the C-ABI).

Bars: per-cell, per-species yields in FP64 mode within 1e-6 relative of the
oracle; FP32 mode within 1e-5 relative of FP64 (tolerances below); FP64
species totals equal to the sum of the per-cell yields.
"""
import numpy as np
import pytest

import iss_amd
from iss_amd import synth, tables
from tests import neos22_table
from tests import oracle_binding as ob

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-6
FP32_RTOL = 1e-5

# (bulk kind, shear, bulk) combinations of the FSSW path
MODES = [
    ("ideal", dict(shear=0, bulk=0, bulk_kind=1)),
    ("shear", dict(shear=1, bulk=0, bulk_kind=1)),
    ("rta_bulk", dict(shear=1, bulk=1, bulk_kind=1)),
    ("osu14", dict(shear=1, bulk=1, bulk_kind=11)),
    ("ce21", dict(shear=1, bulk=1, bulk_kind=21)),
    # 22-moment NEoS: a synthetic table of the reference's layout (its own
    # file is not shipped; tests/neos22_table.py)
    ("mom22", dict(shear=1, bulk=1, bulk_kind=20)),
]


def _ctx(sp, df):
    ctx = iss_amd.Context(0)
    ctx.set_species(sp)
    kw = dict(df)
    if df["bulk"] and df["bulk_kind"] == 11:
        kw["k11"] = tables.read_kind11(1)
    if df["bulk_kind"] == 21:
        kw["neos"] = tables.read_neos(21, 1)
    if df["bulk_kind"] == 20:
        kw["neos"] = neos22_table.table()
    ctx.set_deltaf(**kw)
    return ctx


def _oracle(df, tmp_path_factory):
    if df["bulk_kind"] == 20:
        td = neos22_table.table_dir(str(tmp_path_factory.mktemp("neos22")))
        return ob.Oracle(afterburner=1, table_dir=td, **df)
    return ob.Oracle(afterburner=1, **df)


@pytest.fixture(scope="module")
def species():
    return tables.species_arrays(1)


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp64_per_cell_yields_match_oracle(species, name, df, tmp_path_factory):
    cells, _ = synth.viscous_lrf(3000, species, seed=7, muB_max=0.45)
    ctx = _ctx(species, df)
    ctx.set_surface_lrf(cells)
    tot, per = ctx.compute_yields(precision=64, per_cell=True)
    orc = _oracle(df, tmp_path_factory)
    orc.set_cells_lrf(cells)
    ref = orc.yields()
    assert per.shape == ref.shape
    nz = ref > 0
    # clamped entries are exactly zero in both
    assert np.all(per[~nz] == 0.0)
    rel = np.abs(per[nz] - ref[nz]) / ref[nz]
    print(name, "max rel FP64", rel.max())
    assert rel.max() < FP64_RTOL
    # FP64 species totals = sum over cells
    assert np.allclose(tot, per.sum(axis=0), rtol=1e-12, atol=0)


@pytest.mark.parametrize("name,df", MODES, ids=[m[0] for m in MODES])
def test_fp32_yields_within_1e5_of_fp64(species, name, df):
    cells, _ = synth.viscous_lrf(3000, species, seed=8, muB_max=0.3)
    ctx = _ctx(species, df)
    ctx.set_surface_lrf(cells)
    t64, p64 = ctx.compute_yields(precision=64, per_cell=True)
    t32, p32 = ctx.compute_yields(precision=32, per_cell=True)
    relt = np.abs(t32 - t64) / np.maximum(t64, 1e-300)
    print(name, "max rel FP32 species total", relt.max())
    assert relt.max() < FP32_RTOL
    # every per-cell yield FP32 can represent
    big = p64 > 1e-25
    relc = np.abs(p32[big] - p64[big]) / p64[big]
    print(name, "max rel FP32 per cell", relc.max(), "99.9%", np.quantile(relc, 0.999))
    assert relc.max() < FP32_RTOL
    assert np.all(p32[p64 == 0] == 0)


def _lrf_numpy(lab):
    """Regulation + rest-frame transform in float64 (restating
    src/readindata.cpp:802-842, 1216-1246 and src/iSS.cpp:170-293)."""
    L_ = lab.astype(np.float64)
    n = L_.shape[0]
    u = np.stack([np.sqrt(1 + L_[:, 9] ** 2 + L_[:, 10] ** 2 + L_[:, 11] ** 2),
                  L_[:, 9], L_[:, 10], L_[:, 11]], axis=1)
    idx = [[19, 20, 21, 22], [20, 23, 24, 25], [21, 24, 26, 27], [22, 25, 27, 28]]
    W = L_[:, idx]
    g = np.array([-1.0, 1, 1, 1])
    udp = np.einsum("na,nab->nb", u * g, W) * 1.0
    udp = -u[:, :1] * W[:, 0, :] + np.einsum("na,nab->nb", u[:, 1:], W[:, 1:, :])
    tr = -W[:, 0, 0] + W[:, 1, 1] + W[:, 2, 2] + W[:, 3, 3]
    ul = u * g
    upu = np.einsum("na,nab,nb->n", ul, W, ul)
    G = np.diag(g)
    pim = (W + u[:, :, None] * udp[:, None, :] + u[:, None, :] * udp[:, :, None]
           + u[:, :, None] * u[:, None, :] * upu[:, None, None]
           - (G[None] + u[:, :, None] * u[:, None, :]) * ((tr + upu) / 3)[:, None, None])
    ch, sh = np.cosh(L_[:, 3]), np.sinh(L_[:, 3])
    R = np.zeros((n, 4, 4))
    R[:, 0, 0] = ch; R[:, 0, 3] = sh; R[:, 3, 0] = sh; R[:, 3, 3] = ch
    R[:, 1, 1] = 1; R[:, 2, 2] = 1
    ptz = np.einsum("nab,nbc,ndc->nad", R, pim, R)
    ut = u[:, 0] * ch + u[:, 3] * sh
    uz = u[:, 3] * ch + u[:, 0] * sh
    uv = np.stack([L_[:, 9], L_[:, 10], uz], axis=1)
    Lm = np.zeros((n, 4, 4))
    Lm[:, 0, 0] = ut
    Lm[:, 0, 1:] = -uv
    Lm[:, 1:, 0] = -uv
    Lm[:, 1:, 1:] = np.eye(3)[None] + uv[:, :, None] * uv[:, None, :] / (ut + 1)[:, None, None]
    tau = L_[:, 0]
    dup = np.stack([tau * L_[:, 4] * ch - L_[:, 7] * sh, -tau * L_[:, 5], -tau * L_[:, 6],
                    -L_[:, 7] * ch + tau * L_[:, 4] * sh], axis=1)
    dl = np.einsum("nab,nb->na", Lm, dup)
    pl = np.einsum("nab,nbc,ncd->nad", Lm, ptz, Lm)
    out = np.zeros((n, 28))
    out[:, :4] = L_[:, :4]
    out[:, 4] = dl[:, 0]
    out[:, 5:8] = -dl[:, 1:]
    out[:, 8], out[:, 9], out[:, 10], out[:, 11] = ut, L_[:, 9], L_[:, 10], uz
    out[:, 12:19] = L_[:, 12:19]
    out[:, 19] = L_[:, 29]
    out[:, 20:25] = np.stack([pl[:, 1, 1], pl[:, 1, 2], pl[:, 1, 3], pl[:, 2, 2], pl[:, 2, 3]], 1)
    keep = out[:, 4].astype(np.float32) >= 0
    return out[keep], np.nonzero(keep)[0], W, u, tau


def test_lab_surface_transform(species):
    rng = np.random.default_rng(3)
    n = 500
    lab = np.zeros((n, 34), dtype=np.float32)
    lab[:, 0] = rng.uniform(0.6, 12, n)
    lab[:, 1:3] = rng.uniform(-8, 8, (n, 2))
    lab[:, 3] = rng.uniform(-3, 3, n)
    lab[:, 4] = rng.uniform(-0.2, 1.0, n)           # some u.dsigma < 0 cells
    lab[:, 5:8] = rng.uniform(-0.3, 0.3, (n, 3))
    lab[:, 9:12] = rng.uniform(-1.5, 1.5, (n, 3))
    lab[:, 12] = rng.uniform(0.2, 0.5, n)
    lab[:, 13] = rng.uniform(0.13, 0.16, n)
    lab[:, 14] = 0.3 * lab[:, 12]
    lab[:, 19:29] = rng.normal(scale=0.02, size=(n, 10))
    lab[:, 29] = -rng.uniform(0, 0.02, n)
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_surface_lab(lab)
    got = ctx.get_surface_lrf().astype(np.float64)
    ref, kept, W, u, tau = _lrf_numpy(lab)
    assert got.shape == ref.shape and 0 < got.shape[0] < n
    # (1) FP64 restatement: float storage rounding only
    scale = np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max(axis=0, keepdims=True) + 1e-30)
    err = np.abs(got - ref) / scale
    print("max rel vs FP64 restatement", err.max())
    assert err.max() < 1e-6
    # (2) the oracle mirrors the reference's float boost: agree within its
    # rounding, amplified by (u^t)^2 for pi and u^t for dsigma
    orc, okept = ob.transform(ob.regulate(lab))
    assert np.array_equal(okept, kept)
    orc = orc.astype(np.float64)
    ut = ref[:, 8:9]
    pinorm = np.abs(ref[:, 20:25]).max(axis=1, keepdims=True)
    tol_pi = 5e-5 * ut ** 2 * pinorm + 1e-12
    assert np.all(np.abs(orc[:, 20:25] - got[:, 20:25]) < tol_pi)
    dnorm = np.abs(ref[:, 4:8]).max(axis=1, keepdims=True)
    assert np.all(np.abs(orc[:, 4:8] - got[:, 4:8]) < 2e-5 * ut * dnorm + 1e-12)
    # (3) invariants: u.dsigma = dsigma_LRF^0; pi:pi is frame independent
    tk = tau[kept]
    dsig = np.stack([tk * lab[kept, 4], tk * lab[kept, 5], tk * lab[kept, 6], lab[kept, 7]], 1)
    udsig = (u[kept] * dsig).sum(axis=1)  # lower-index Milne dsigma, u^mu contravariant
    assert np.allclose(got[:, 4], udsig, rtol=2e-6, atol=1e-7)
    pl = got[:, 20:25]
    pizz = -pl[:, 0] - pl[:, 3]
    pipi_lrf = pl[:, 0] ** 2 + pl[:, 3] ** 2 + pizz ** 2 + 2 * (pl[:, 1] ** 2 + pl[:, 2] ** 2 + pl[:, 4] ** 2)
    reg = ob.regulate(lab)[kept].astype(np.float64)
    Wr = reg[:, [[19, 20, 21, 22], [20, 23, 24, 25], [21, 24, 26, 27], [22, 25, 27, 28]]]
    g = np.diag([1.0, -1, -1, -1])
    pipi_lab = np.einsum("nab,ac,bd,ncd->n", Wr, g, g, Wr)
    assert np.allclose(pipi_lrf, pipi_lab, rtol=1e-3, atol=1e-9)


def test_empty_surface_and_single_cell(species):
    ctx = iss_amd.Context(0)
    ctx.set_species(species)
    ctx.set_surface_lrf(np.zeros((0, 28), dtype=np.float32))
    tot = ctx.compute_yields(precision=32)
    assert np.all(tot == 0)
    assert ctx.sample(5, seed=1) == 0
    assert np.all(ctx.event_counts() == 0)
    cells, _ = synth.viscous_lrf(1, species, seed=9, target_hadrons=50.0)
    ctx.set_surface_lrf(cells)
    tot, per = ctx.compute_yields(precision=64, per_cell=True)
    orc = ob.Oracle(afterburner=1)
    orc.set_cells_lrf(cells)
    ref = orc.yields()
    nz = ref > 0
    assert np.max(np.abs(per[nz] - ref[nz]) / ref[nz]) < FP64_RTOL


@pytest.mark.parametrize("name,df", MODES[:2], ids=[m[0] for m in MODES[:2]])
def test_bessel_arguments_outside_the_staged_table(species, name, df):
    """Hot cells (T = 0.3 GeV: pion m/T < 0.5, below the K_n table) and cold
    ones (T = 16 MeV: heavy m/T beyond the rows staged in LDS) make the pass
    rerun with the exact K_n and the whole global table
    (src/FSSW.cpp:1646-1703 calls GSL outside the table): FP64 against the
    oracle, FP32 against FP64, on the same bars as the other cells."""
    cells, _ = synth.viscous_lrf(400, species, seed=12, muB_max=0.2)
    cells = cells.copy()
    cells[:20, 13] = 0.30
    cells[20:40, 13] = 0.016
    cells[20:40, 16] = 0.0
    ctx = _ctx(species, df)
    ctx.set_surface_lrf(cells)
    t64, p64 = ctx.compute_yields(precision=64, per_cell=True)
    orc = ob.Oracle(afterburner=1, **df)
    orc.set_cells_lrf(cells)
    ref = orc.yields()
    nz = ref > 0
    rel = np.abs(p64[nz] - ref[nz]) / ref[nz]
    print(name, "max rel FP64", rel.max())
    assert rel.max() < FP64_RTOL
    assert np.allclose(t64, p64.sum(axis=0), rtol=1e-12, atol=0)
    t32, p32 = ctx.compute_yields(precision=32, per_cell=True)
    big = p64 > 1e-25
    relc = np.abs(p32[big] - p64[big]) / p64[big]
    print(name, "max rel FP32 per cell", relc.max())
    assert relc.max() < FP32_RTOL
    ctx.close()
