"""Synthetic freeze-out surfaces for the benchmark and tests.

Configurations (all seeded, generated in memory):
  C1  plain ideal, boost-invariant constant-tau surface: tau = 10 fm on a
      40 x 50 (x, y) grid, dx = dy = 0.3 fm, radial flow u_r = gamma v with
      v = 0.6 tanh(r / 7 fm), T = 0.150 GeV, mu_B = 0.
  C2  viscous 3+1D: N cells with tau in [0.6, 15] fm, eta_s in [-5, 5],
      T in [0.135, 0.160] GeV, mu_B in [0, 0.3] GeV, |u_perp| <= 3, random
      traceless transverse pi^{mu nu} with |pi|/(e+P) <= 0.3, bulk
      Pi/(e+P) in [-0.2, 0], normalisation such that <hadrons/event> ~ 1e4.
  C3  256 surfaces generated as C2 with seeds 1000 + k, cell counts
      clip(round(exp(N(ln 3e4, 1))), 1e3, 5e5), oversampled events
      U_int[10, 2000], and C2's hadron density per cell (0.05 per event).
  C5  as C2 with 2,000,000 cells and mu_B up to 0.45 GeV.

  LAB viscous lab-frame cells (viscous_lab): eta_s in [-1.5, 1.5], flow with
      longitudinal and transverse components, time-like and space-like
      dsigma (a few per cent with u.dsigma < 0), pi^{mu nu} orthogonal to u
      and traceless, for the smooth-spectra checks.

Cells are returned as rest-frame records (28 floats, FO_surf_LRF order) or,
for C1 and LAB, as lab-frame MUSIC records (34 floats, FO_surf order).  e, P, n_B are
those of an ideal hadron-resonance gas of the same species list.
"""
from __future__ import annotations

import numpy as np
from scipy.special import kn

HBARC = 0.197327053


def hrg_thermo(T, muB, sp, nterms=10):
    """(e [GeV/fm^3], P [GeV/fm^3], n_B [1/fm^3]) of an ideal BE/FD gas of the
    species `sp` (mass, g, B, sign) as Bessel series; T, muB arrays."""
    T = np.asarray(T, dtype=np.float64)
    muB = np.asarray(muB, dtype=np.float64)
    e = np.zeros_like(T)
    P = np.zeros_like(T)
    nB = np.zeros_like(T)
    for m, g, B, s in zip(sp["mass"], sp["g"], sp["B"], sp["sign"]):
        nt = nterms if m < 0.7 else 1
        lam = np.exp(B * muB / T)
        for n in range(1, nt + 1):
            th = (-s) ** (n - 1) if s != 0 else (1.0 if n == 1 else 0.0)
            if th == 0.0:
                continue
            x = n * m / T
            f = th * lam ** n
            k2 = kn(2, x)
            k1 = kn(1, x)
            pn = g / (2 * np.pi ** 2) * m * m * T * T / (n * n) * k2
            P += f * pn
            e += f * g / (2 * np.pi ** 2) * (3 * m * m * T * T / (n * n) * k2 + m ** 3 * T / n * k1)
            nB += B * f * g / (2 * np.pi ** 2) * m * m * T / n * k2
    c3 = HBARC ** 3
    return e / c3, P / c3, nB / c3


def hrg_grid(sp, t0, t1, m0, m1, n=64):
    """hrg_thermo tabulated on an n x n (T, mu_B) grid (for hrg_thermo_fast)."""
    tg = np.linspace(t0, t1, n)
    mg = np.linspace(m0, m1, n)
    TT, MM = np.meshgrid(tg, mg, indexing="ij")
    return (t0, t1, m0, m1, n, hrg_thermo(TT.ravel(), MM.ravel(), sp))


def hrg_thermo_fast(T, muB, sp, grid=None):
    """hrg_thermo on a 64 x 64 (T, mu_B) grid with bilinear interpolation for
    large surfaces (relative interpolation error ~1e-5); `grid` (hrg_grid)
    reuses one table for many surfaces."""
    T = np.asarray(T, dtype=np.float64)
    muB = np.asarray(muB, dtype=np.float64)
    if grid is None and T.size < 5000:
        return hrg_thermo(T, muB, sp)
    if grid is None:
        grid = hrg_grid(sp, T.min(), T.max() + 1e-9, muB.min(), muB.max() + 1e-9)
    t0, t1, m0, m1, nt, grids = grid
    nm = nt
    ft = (T - t0) / (t1 - t0) * (nt - 1)
    fm = (muB - m0) / (m1 - m0) * (nm - 1)
    it = np.clip(ft.astype(int), 0, nt - 2)
    im = np.clip(fm.astype(int), 0, nm - 2)
    a, b = ft - it, fm - im
    out = []
    for g in grids:
        g = g.reshape(nt, nm)
        out.append((1 - a) * (1 - b) * g[it, im] + a * (1 - b) * g[it + 1, im]
                   + (1 - a) * b * g[it, im + 1] + a * b * g[it + 1, im + 1])
    return tuple(out)


def _lrf_shear(rng, n, epp, frac_max=0.3):
    """5 independent components of a traceless symmetric rest-frame pi with
    sqrt(pi:pi)/(e+P) uniform in [0, frac_max]."""
    comp = rng.normal(size=(n, 5))
    xx, xy, xz, yy, yz = comp.T
    zz = -xx - yy
    norm = np.sqrt(xx ** 2 + yy ** 2 + zz ** 2 + 2 * (xy ** 2 + xz ** 2 + yz ** 2))
    scale = rng.uniform(0, frac_max, n) * epp / np.maximum(norm, 1e-12)
    return comp * scale[:, None]


def viscous_lrf(n_cells, sp, seed=2, muB_max=0.3, target_hadrons=1.0e4, bulk=True, shear=True, grid=None,
                diffusion=0.0):
    """C2-style rest-frame surface (28 floats per cell) and the area scale."""
    rng = np.random.default_rng(seed)
    n = int(n_cells)
    c = np.zeros((n, 28), dtype=np.float64)
    tau = rng.uniform(0.6, 15.0, n)
    eta = rng.uniform(-5.0, 5.0, n)
    c[:, 0] = tau
    c[:, 1] = rng.uniform(-10, 10, n)
    c[:, 2] = rng.uniform(-10, 10, n)
    c[:, 3] = eta
    T = rng.uniform(0.135, 0.160, n)
    muB = rng.uniform(0.0, muB_max, n)
    # dsigma in the rest frame: A (1, a), |a| <= 0.3 uniform in the ball
    a = rng.normal(size=(n, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    a *= 0.3 * rng.uniform(0, 1, n)[:, None] ** (1 / 3)
    # flow: |u_perp| uniform in [0, 3], u^eta = 0 -> (t, z) components
    ur = rng.uniform(0, 3.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    ux, uy = ur * np.cos(ph), ur * np.sin(ph)
    utau = np.sqrt(1 + ux ** 2 + uy ** 2)
    c[:, 8] = utau * np.cosh(eta)
    c[:, 9] = ux
    c[:, 10] = uy
    c[:, 11] = utau * np.sinh(eta)
    e, P, nB = hrg_thermo_fast(T, muB, sp, grid)
    c[:, 12], c[:, 13], c[:, 14], c[:, 15], c[:, 16] = e, T, P, nB, muB
    epp = e + P
    if bulk:
        c[:, 19] = -rng.uniform(0, 0.2, n) * epp
    if shear:
        c[:, 20:25] = _lrf_shear(rng, n, epp)
    if diffusion > 0:
        # rest-frame baryon diffusion current, |q| / n_B <= diffusion
        q = rng.normal(0, 1, (n, 3))
        q *= (diffusion * rng.uniform(0, 1, n) * np.maximum(nB, 1e-3) / np.linalg.norm(q, axis=1))[:, None]
        c[:, 25:28] = q
    # area: Poisson mean per cell ~ n_eq * A; choose A for the target
    dens = P / T  # Boltzmann number density [1/fm^3]
    w = rng.uniform(0.5, 1.5, n)
    A = target_hadrons / float(np.sum(dens * w))
    c[:, 4] = A * w
    c[:, 5:8] = (A * w)[:, None] * a
    return c.astype(np.float32), A


def ideal_binv_lab(sp=None):
    """C1 lab-frame boost-invariant surface (34 floats per cell)."""
    nx, ny, d = 40, 50, 0.3
    xs = (np.arange(nx) - 19.5) * d
    ys = (np.arange(ny) - 24.5) * d
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    X, Y = X.ravel(), Y.ravel()
    n = X.size
    r = np.hypot(X, Y)
    v = 0.6 * np.tanh(r / 7.0)
    gam = 1 / np.sqrt(1 - v ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(r > 0, X / r, 0.0)
        cy = np.where(r > 0, Y / r, 0.0)
    c = np.zeros((n, 34), dtype=np.float64)
    c[:, 0] = 10.0
    c[:, 1], c[:, 2] = X, Y
    c[:, 4] = d * d * 1.0  # dsigma_tau = dx dy deta, times tau in the transform
    c[:, 8] = gam
    c[:, 9], c[:, 10] = gam * v * cx, gam * v * cy
    T = np.full(n, 0.150)
    c[:, 13] = T
    if sp is not None:
        e, P, nB = hrg_thermo(T, np.zeros(n), sp)
        c[:, 12], c[:, 14] = e, P
    return c.astype(np.float32)


def _boost_to_lab(u):
    """Lorentz matrices taking rest-frame contravariant vectors to the frame
    where the fluid moves with 4-velocity u (n x 4)."""
    n = u.shape[0]
    B = np.zeros((n, 4, 4))
    B[:, 0, 0] = u[:, 0]
    B[:, 0, 1:] = u[:, 1:]
    B[:, 1:, 0] = u[:, 1:]
    B[:, 1:, 1:] = np.eye(3)[None] + u[:, 1:, None] * u[:, None, 1:] / (u[:, 0] + 1.0)[:, None, None]
    return B


def viscous_lab(n_cells, sp, seed=7, muB_max=0.2, shear=True, bulk=True, frac_spacelike=0.3,
                eta_max=1.5, area=1.0, diffusion=0.0):
    """Lab-frame MUSIC records (34 floats, FO_surf order) of a viscous
    surface in Milne coordinates.  Components in the local orthonormal frame
    (tau, x, y, eta): u3 = tau u^eta, da3 the eta_s component (p.dsigma =
    tau (pt da0 + px da1 + py da2) + pz da3), pi^{mu nu} contravariant with
    pi03 = tau pi^{tau eta}, pi33 = tau^2 pi^{eta eta}."""
    rng = np.random.default_rng(seed)
    n = int(n_cells)
    c = np.zeros((n, 34), dtype=np.float64)
    tau = rng.uniform(0.8, 12.0, n)
    c[:, 0] = tau
    c[:, 1] = rng.uniform(-8, 8, n)
    c[:, 2] = rng.uniform(-8, 8, n)
    c[:, 3] = rng.uniform(-eta_max, eta_max, n)
    ux = rng.normal(0, 0.8, n)
    uy = rng.normal(0, 0.8, n)
    ue = rng.normal(0, 0.3, n)
    u0 = np.sqrt(1 + ux ** 2 + uy ** 2 + ue ** 2)
    u = np.stack([u0, ux, uy, ue], axis=1)
    c[:, 8:12] = u
    # dsigma_mu (orthonormal, lower index): mostly time-like along u, some space-like
    D = np.zeros((n, 4))
    D[:, 0] = area * rng.uniform(0.5, 1.5, n)
    D[:, 1:] = area * rng.normal(0, 0.25, (n, 3))
    sl = rng.uniform(0, 1, n) < frac_spacelike
    D[sl, 1:] = area * rng.normal(0, 1.2, (sl.sum(), 3))
    D[sl, 0] = area * rng.normal(0, 0.4, sl.sum())
    c[:, 4] = D[:, 0] / tau
    c[:, 5] = D[:, 1] / tau
    c[:, 6] = D[:, 2] / tau
    c[:, 7] = D[:, 3]
    T = rng.uniform(0.135, 0.160, n)
    muB = rng.uniform(0.0, muB_max, n)
    e, P, nB = hrg_thermo(T, muB, sp)
    c[:, 12], c[:, 13], c[:, 14], c[:, 15], c[:, 16] = e, T, P, nB, muB
    epp = e + P
    if shear:
        xx, xy, xz, yy, yz = _lrf_shear(rng, n, epp).T
        PL = np.zeros((n, 4, 4))
        PL[:, 1, 1], PL[:, 1, 2], PL[:, 1, 3] = xx, xy, xz
        PL[:, 2, 1], PL[:, 2, 2], PL[:, 2, 3] = xy, yy, yz
        PL[:, 3, 1], PL[:, 3, 2], PL[:, 3, 3] = xz, yz, -xx - yy
        B = _boost_to_lab(u)
        Pm = np.einsum("nab,nbc,ndc->nad", B, PL, B)
        idx = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
        for k, (a, b) in enumerate(idx):
            c[:, 19 + k] = Pm[:, a, b]
    if bulk:
        c[:, 29] = -rng.uniform(0, 0.15, n) * epp
    if diffusion > 0:
        # baryon diffusion current orthogonal to u: (0, q) in the rest frame,
        # |q| / n_B <= diffusion, boosted to the lab
        ql = np.zeros((n, 4))
        ql[:, 1:] = rng.normal(0, 1, (n, 3))
        ql[:, 1:] *= (diffusion * rng.uniform(0, 1, n) * np.maximum(nB, 1e-3) /
                      np.linalg.norm(ql[:, 1:], axis=1))[:, None]
        qm = np.einsum("nab,nb->na", _boost_to_lab(u), ql)
        c[:, 30:34] = qm
    return c.astype(np.float32)


def batched_surfaces(sp, n_surfaces=256, seed=3, median_cells=3.0e4, sigma=1.0, min_cells=1000,
                     max_cells=500000, events=(10, 2000), hadrons_per_cell=0.05):
    """C3: (list of rest-frame surfaces, events per surface)."""
    rng = np.random.default_rng(seed)
    n = np.clip(np.round(np.exp(rng.normal(np.log(median_cells), sigma, n_surfaces))), min_cells,
                max_cells).astype(np.int64)
    ev = rng.integers(events[0], events[1] + 1, n_surfaces).astype(np.int32)
    grid = hrg_grid(sp, 0.135, 0.160 + 1e-9, 0.0, 0.3 + 1e-9)
    surfaces = []
    for k in range(n_surfaces):
        c, _ = viscous_lrf(int(n[k]), sp, seed=1000 + k, target_hadrons=hadrons_per_cell * float(n[k]), grid=grid)
        surfaces.append(c)
    return surfaces, ev
