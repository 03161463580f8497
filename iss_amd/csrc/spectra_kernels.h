// Smooth Cooper-Frye spectra dN / (pT dpT dphi dy) of chosen species on a
// (pT, phi) grid, summed over the surface cells: the reference's
// EmissionFunctionArray::calculate_dN_pTdpTdphidy (src/emissionfunction.cpp:
// 625-844).  (one translation unit: included by iss_hip.hip)
//
// The reference contracts the lab momentum with the lab-frame Milne cell
// (p.u, p.dsigma, p_mu p_nu pi^{mu nu}).  These are Lorentz invariants, so
// here the lab momentum (mT cosh y, pT cos phi, pT sin phi, mT sinh y) is
// boosted into each cell's rest frame and contracted with the resident
// rest-frame cell record, whose dsigma already carries the factor tau.  The
// delta f is the sampler's (shear with the three coefficient choices, bulk
// kinds 1/11/20/21; src/FSSW.cpp:1897-1928); for shear and bulk kind 1 this is
// the reference's smooth-path delta f (src/emissionfunction.cpp:768-776,
// 4156-4186).  As in the reference, delta f is not clamped into [0, 2]; it is
// optionally rescaled to |delta f| <= deltaf_max_ratio (flag_restrict_deltaf,
// :791-799) and negative contributions are optionally dropped
// (use_pos_dN_only, :804-806).
//
// Layout: a per-cell FP64 record of kSpecRec values (the species-independent
// part of the integrand, k_spectra_rec) is read with wave-uniform (scalar)
// loads, so the per-point work is pure VALU.  One wave evaluates 64 output
// points over one contiguous chunk of cells; chunk sums go to
// partial[chunk][species][point] in FP64 and k_spectra_reduce adds them in
// chunk order (deterministic).
#pragma once

#include "kernels.h"
#include "yield_kernels.h"

namespace issdev {

constexpr int kSpecRec = 36;
enum SpecSlot {
    RS_IT = 0,                          // 1 / T
    RS_MB, RS_MS, RS_MQ,                // mu_B / T, mu_S / T, mu_Q / T
    RS_U0, RS_U1, RS_U2, RS_U3, RS_IU,  // u (t, x, y, z), 1 / (u0 + 1)
    RS_D0, RS_D1, RS_D2, RS_D3,         // rest-frame dsigma_mu (includes tau)
    RS_XX, RS_XY2, RS_XZ2, RS_YY, RS_YZ2,  // rest-frame pi: xx, 2xy, 2xz, yy, 2yz
    RS_SA, RS_SB,                       // shear: omf * W * (SA + SB / E)
    RS_BM2, RS_BE0, RS_BB, RS_BS, RS_BQ, RS_BE2, RS_BM2E,
    // bulk: omf * (BM2 m^2 + (BE0 + B BB + S BS + Q BQ) E + BE2 E^2 + BM2E m^2 / E)
    RS_CHE, RS_SHE,                     // cosh, sinh of the cell's eta_s
    RS_QX, RS_QY, RS_QZ, RS_PQ,         // diffusion: q / D_qmu, n_B / (e + P)
    kSpecUsed
};
static_assert(kSpecUsed <= kSpecRec, "spectra record");

__global__ void k_spectra_rec(const float* __restrict__ cells, const double* __restrict__ coef, int64_t N,
                              int shear, int bulk, int neos, int diff, double* __restrict__ rec) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double r[kSpecRec];
    for (int k = 0; k < kSpecRec; k++) r[k] = 0.;
    const double T = cells[F_T * N + i];
    r[RS_IT] = 1. / T;
    r[RS_MB] = cells[F_MUB * N + i] / T;
    r[RS_MS] = cells[F_MUS * N + i] / T;
    r[RS_MQ] = cells[F_MUQ * N + i] / T;
    const double u0 = cells[F_U0 * N + i];
    r[RS_U0] = u0;
    r[RS_U1] = cells[F_U1 * N + i];
    r[RS_U2] = cells[F_U2 * N + i];
    r[RS_U3] = cells[F_U3 * N + i];
    r[RS_IU] = 1. / (u0 + 1.);
    r[RS_D0] = cells[F_DA0 * N + i];
    r[RS_D1] = cells[F_DA1 * N + i];
    r[RS_D2] = cells[F_DA2 * N + i];
    r[RS_D3] = cells[F_DA3 * N + i];
    double c[6];
    for (int k = 0; k < 6; k++) c[k] = coef[k * N + i];
    if (shear) {
        r[RS_XX] = cells[F_PIXX * N + i];
        r[RS_XY2] = 2. * cells[F_PIXY * N + i];
        r[RS_XZ2] = 2. * cells[F_PIXZ * N + i];
        r[RS_YY] = cells[F_PIYY * N + i];
        r[RS_YZ2] = 2. * cells[F_PIYZ * N + i];
        if (neos == 1) r[RS_SB] = 1. / (2. * c[2] * T);
        else if (neos == 0) r[RS_SA] = c[0];
        else r[RS_SA] = 1. / (2. * T * T * (static_cast<double>(cells[F_E * N + i]) + cells[F_P * N + i]));
    }
    if (bulk != BULK_NONE) {
        double Pi = cells[F_PI * N + i];
        if (bulk == BULK_RTA) Pi = Pi / kHbarC;
        if (bulk == BULK_RTA || bulk == BULK_CE) {
            // -c0 Pi (m^2 / (3 T E) - c1 E / T)
            r[RS_BM2E] = -c[0] * Pi / (3. * T);
            r[RS_BE0] = c[0] * c[1] * Pi / T;
        } else if (bulk == BULK_14MOM) {
            r[RS_BM2] = Pi * c[0];
            r[RS_BB] = Pi * c[1];
            r[RS_BE2] = Pi * c[2];
        } else if (bulk == BULK_22MOM) {
            r[RS_BM2] = Pi * c[2];
            r[RS_BB] = Pi * c[3];
            r[RS_BS] = Pi * c[4];
            r[RS_BQ] = Pi * c[5];
            r[RS_BE2] = Pi * (c[1] - c[2]);
        }
    }
    const double eta = cells[F_ETA * N + i];
    r[RS_CHE] = cosh(eta);
    r[RS_SHE] = sinh(eta);
    if (diff) {
        // (src/emissionfunction.cpp:777-786): (1 - s f0) (n_B/(e+P) - B/E) p.q / D_qmu
        const double iq = coef[6 * N + i];
        r[RS_QX] = cells[F_QX * N + i] * iq;
        r[RS_QY] = cells[F_QY * N + i] * iq;
        r[RS_QZ] = cells[F_QZ * N + i] * iq;
        r[RS_PQ] = coef[7 * N + i];
    }
    for (int k = 0; k < kSpecRec; k++) rec[i * kSpecRec + k] = r[k];
}

struct SpectraArgs {
    const double* rec;        // [N][kSpecRec]
    int64_t N;
    int64_t cells_per_chunk;
    SpeciesDev sp;
    const int32_t* species;   // [ns] species indices
    int ns;
    const double* pT;         // [npT]
    int npT;
    const double* cphi;       // [nphi]
    const double* sphi;
    int nphi;
    // mode 0: nodes of y - eta_s with weights (cosh, sinh, weight);
    // mode 1: rapidities (cosh, sinh), one output plane per rapidity
    const double* ych;
    const double* ysh;
    const double* yw;
    int ny;
    int mode;
    int positive_only;
    double restrict_ratio;    // <= 0: no rescaling of delta f
    int waves_per_species;
    double* partial;          // [chunks][ns][npts]
};

constexpr double kLog2e = 1.4426950408889634;
constexpr int kSpecPts = 4;   // output points per lane

__device__ __forceinline__ double fexp2(double x) { return exp2(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }

// per (cell, species) constants of the delta f, in the working type R
template <typename R>
struct SpecCell {
    R xx, xy2, xz2, yy, yz2, sa, sb, cm, be, be2, cme, sg, lo, hi;
    R qx, qy, qz, pq, bq;   // diffusion (zero when off)
};

// f0 p.dsigma (1 + delta f) from the rest-frame energy E, momentum (lx, ly,
// lz), p.dsigma and x2 = (E - mu) / T * log2(e).  |delta f| is clamped at
// deltaf_max_ratio (lo = -ratio, hi = ratio, or -+inf when off): the
// reference's rescaling df * min(1, ratio / (|df| + 1e-10)) differs from the
// clamp by at most ratio * 1e-10 / |df| in 1 + delta f.
template <typename R>
__device__ __forceinline__ R spectra_integrand(const SpecCell<R>& c, R E, R lx, R ly, R lz, R pds, R x2,
                                               int positive_only) {
    const R f0 = frcp(fexp2(x2) + c.sg);
    const R omf = R(1) - c.sg * f0;
    const R lz2 = lz * lz;
    const R W = c.xx * (lx * lx - lz2) + c.yy * (ly * ly - lz2) + lx * (c.xy2 * ly + c.xz2 * lz) + c.yz2 * ly * lz;
    const R iE = frcp(E);
    R df = omf * (W * (c.sa + c.sb * iE) + c.cm + E * (c.be + c.be2 * E) + c.cme * iE -
                  (c.pq - c.bq * iE) * (lx * c.qx + ly * c.qy + lz * c.qz));
    df = fmin(fmax(df, c.lo), c.hi);
    R r = f0 * pds * (R(1) + df);
    if (positive_only && r < R(0)) r = R(0);
    return r;
}

// grid (chunks, ns * waves_per_species), 64 threads, kSpecPts points per lane
// (points base + 64 q).
//
// mode 0: the lab momentum at rapidity eta_cell + (y - eta_s)_k is
// (mT ch_k, px, py, mT sh_k); its rest-frame components are linear in mT with
// per-(cell, k) coefficients and offsets per (cell, point):
//   E = mT a0 + Pe,  l_i = mT a_i + P_i,  p.dsigma = mT dk + Pds,
//   a0 = u0 ch - u3 sh, g = sh u3 / (u0 + 1) - ch, (a1, a2, a3) = (u1 g, u2 g, sh + u3 g),
//   Pe = -(px u1 + py u2), P_i = p_i - Pe u_i / (u0 + 1) (p_z = 0).
// The wave-uniform coefficients are formed in FP64 (ch_k, sh_k of boosted
// cells are differences of large terms), the per-lane work runs in R.
template <typename R>
__global__ __launch_bounds__(64) void k_spectra(SpectraArgs a) {
    const int npl = a.npT * a.nphi;
    const int npts = npl * (a.mode == 1 ? a.ny : 1);
    const int si = blockIdx.y / a.waves_per_species;
    const int base = (blockIdx.y % a.waves_per_species) * 64 * kSpecPts + threadIdx.x;
    const int s = a.species[si];
    const double m = a.sp.mass[s];
    const double B = a.sp.B[s], S = a.sp.S[s], Q = a.sp.Q[s];
    double mT[kSpecPts], px[kSpecPts], py[kSpecPts], p0f[kSpecPts], pzf[kSpecPts];
#pragma unroll
    for (int q = 0; q < kSpecPts; q++) {
        const int pt = min(base + 64 * q, npts - 1);
        const int iphi = pt % a.nphi;
        const int ipt = (pt / a.nphi) % a.npT;
        const double pT = a.pT[ipt];
        mT[q] = sqrt(pT * pT + m * m);
        px[q] = pT * a.cphi[iphi];
        py[q] = pT * a.sphi[iphi];
        p0f[q] = pzf[q] = 0.;
        if (a.mode == 1) {
            const int iy = pt / npl;
            p0f[q] = mT[q] * a.ych[iy];
            pzf[q] = mT[q] * a.ysh[iy];
        }
    }
    SpecCell<R> c;
    c.sg = static_cast<R>(a.sp.sign[s]);
    c.bq = static_cast<R>(B);
    const R inf = static_cast<R>(INFINITY);
    c.hi = a.restrict_ratio > 0. ? static_cast<R>(a.restrict_ratio) : inf;
    c.lo = -c.hi;
    const int64_t l0 = static_cast<int64_t>(blockIdx.x) * a.cells_per_chunk;
    const int64_t l1 = min<int64_t>(a.N, l0 + a.cells_per_chunk);
    double acc[kSpecPts];
#pragma unroll
    for (int q = 0; q < kSpecPts; q++) acc[q] = 0.;
    for (int64_t l = l0; l < l1; l++) {
        const double* v = a.rec + l * kSpecRec;
        const double u0 = v[RS_U0], u1 = v[RS_U1], u2 = v[RS_U2], u3 = v[RS_U3], iu = v[RS_IU];
        const double d0 = v[RS_D0], d1 = v[RS_D1], d2 = v[RS_D2], d3 = v[RS_D3];
        const double it = v[RS_IT];
        const double mu = B * v[RS_MB] + S * v[RS_MS] + Q * v[RS_MQ];
        c.xx = static_cast<R>(v[RS_XX]); c.xy2 = static_cast<R>(v[RS_XY2]); c.xz2 = static_cast<R>(v[RS_XZ2]);
        c.yy = static_cast<R>(v[RS_YY]); c.yz2 = static_cast<R>(v[RS_YZ2]);
        c.sa = static_cast<R>(v[RS_SA]); c.sb = static_cast<R>(v[RS_SB]);
        c.cm = static_cast<R>(v[RS_BM2] * m * m);
        c.be = static_cast<R>(v[RS_BE0] + B * v[RS_BB] + S * v[RS_BS] + Q * v[RS_BQ]);
        c.be2 = static_cast<R>(v[RS_BE2]);
        c.cme = static_cast<R>(v[RS_BM2E] * m * m);
        c.qx = static_cast<R>(v[RS_QX]);
        c.qy = static_cast<R>(v[RS_QY]);
        c.qz = static_cast<R>(v[RS_QZ]);
        c.pq = static_cast<R>(v[RS_PQ]);
        if (a.mode == 0) {
            R mTr[kSpecPts], Pe[kSpecPts], Px[kSpecPts], Py[kSpecPts], Pz[kSpecPts], Pds[kSpecPts], X0[kSpecPts];
            R cell[kSpecPts];
#pragma unroll
            for (int q = 0; q < kSpecPts; q++) {
                const double pe = -(px[q] * u1 + py[q] * u2);
                const double qx = px[q] - pe * iu * u1, qy = py[q] - pe * iu * u2, qz = -pe * iu * u3;
                mTr[q] = static_cast<R>(mT[q]);
                Pe[q] = static_cast<R>(pe);
                Px[q] = static_cast<R>(qx);
                Py[q] = static_cast<R>(qy);
                Pz[q] = static_cast<R>(qz);
                Pds[q] = static_cast<R>(pe * d0 + qx * d1 + qy * d2 + qz * d3);
                X0[q] = static_cast<R>((pe * it - mu) * kLog2e);
                cell[q] = R(0);
            }
            const double che = v[RS_CHE], she = v[RS_SHE];
            for (int k = 0; k < a.ny; k++) {
                const double ch = che * a.ych[k] + she * a.ysh[k];
                const double sh = she * a.ych[k] + che * a.ysh[k];
                const double a0 = u0 * ch - u3 * sh;
                const double g = sh * u3 * iu - ch;
                const double a1 = u1 * g, a2 = u2 * g, a3 = sh + u3 * g;
                const R A0 = static_cast<R>(a0), A1 = static_cast<R>(a1), A2 = static_cast<R>(a2),
                        A3 = static_cast<R>(a3);
                const R DK = static_cast<R>(a0 * d0 + a1 * d1 + a2 * d2 + a3 * d3);
                const R AX = static_cast<R>(a0 * it * kLog2e);
                const R wk = static_cast<R>(a.yw[k]);
#pragma unroll
                for (int q = 0; q < kSpecPts; q++) {
                    const R E = mTr[q] * A0 + Pe[q];
                    const R lx = mTr[q] * A1 + Px[q], ly = mTr[q] * A2 + Py[q], lz = mTr[q] * A3 + Pz[q];
                    const R pds = mTr[q] * DK + Pds[q];
                    const R x2 = mTr[q] * AX + X0[q];
                    cell[q] += wk * spectra_integrand(c, E, lx, ly, lz, pds, x2, a.positive_only);
                }
            }
#pragma unroll
            for (int q = 0; q < kSpecPts; q++) acc[q] += static_cast<double>(cell[q]);
        } else {
#pragma unroll
            for (int q = 0; q < kSpecPts; q++) {
                const double pdu = px[q] * u1 + py[q] * u2 + pzf[q] * u3;
                const double E = u0 * p0f[q] - pdu;
                const double kk = pdu * iu - p0f[q];
                const double lx = px[q] + kk * u1, ly = py[q] + kk * u2, lz = pzf[q] + kk * u3;
                const double pds = E * d0 + lx * d1 + ly * d2 + lz * d3;
                acc[q] += static_cast<double>(spectra_integrand(
                    c, static_cast<R>(E), static_cast<R>(lx), static_cast<R>(ly), static_cast<R>(lz),
                    static_cast<R>(pds), static_cast<R>((E * it - mu) * kLog2e), a.positive_only));
            }
        }
    }
    const double g = static_cast<double>(a.sp.g[s]);
    const double pref = g / (8. * kPi * kPi * kPi) / (kHbarC * kHbarC * kHbarC);
#pragma unroll
    for (int q = 0; q < kSpecPts; q++) {
        const int pt = base + 64 * q;
        if (pt < npts) a.partial[(static_cast<int64_t>(blockIdx.x) * a.ns + si) * npts + pt] = acc[q] * pref;
    }
}

// out[i] = sum over chunks, in chunk order
__global__ void k_spectra_reduce(const double* __restrict__ partial, int64_t nchunks, int64_t n,
                                 double* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (int64_t t = 0; t < nchunks; t++) s += partial[t * n + i];
    out[i] = s;
}

}  // namespace issdev
