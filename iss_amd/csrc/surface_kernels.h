// Special-function table, surface regulation / rest-frame transform and per-cell delta-f coefficients.
// (one translation unit: included by iss_hip.hip)
#pragma once

#include "expint.h"
#include "kernels.h"

namespace issdev {

// ============================================================ device math

// K_n(x) = int_0^inf exp(-x cosh t) cosh(n t) dt by the trapezoid rule.  The
// reference fills its table with GSL (src/FSSW.cpp:1626-1633) and calls GSL
// outside [0.5, 399.95] (src/FSSW.cpp:1648-1656); GSL is not used here.
__device__ double kn_exact(int n, double x) {
    const double h = fmin(0.1, 0.25 / sqrt(x));
    const double tpeak = asinh(n / x);
    double sum = 0.5;
    for (int k = 1; k < 100000; k++) {
        const double t = k * h;
        const double f = exp(-x * (cosh(t) - 1.)) * cosh(n * t);
        sum += f;
        if (t > tpeak && f < 1e-18 * sum) break;
    }
    return h * sum * exp(-x);
}

__global__ void k_bessel_table(double* ktab, float2* ktabf, int with_bulk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kKLen) return;
    const double x = kKx0 + i * kKdx;
    for (int o = 1; o <= 3; o++) {
        double v = 0.;
        if (o == 2 || with_bulk) v = kn_exact(o, x);
        ktab[(o - 1) * kKLen + i] = v;
    }
}

__global__ void k_bessel_table_f32(const double* ktab, float2* ktabf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kKLen) return;
    for (int o = 0; o < 3; o++) {
        const double a = ktab[o * kKLen + i];
        const double b = (i + 1 < kKLen) ? ktab[o * kKLen + i + 1] : a;
        ktabf[o * kKLen + i] = make_float2(static_cast<float>(a), static_cast<float>(b - a));
    }
}

// ===================================================== surface preparation

struct EosDev {
    int iEOS;
    int64_t n_rows;
    const double* rows;  // [n_rows][7]
};

// getValuesFromHRGEOS (src/readindata.cpp:1249-1309)
__device__ int eos_lookup(const EosDev& e, double ed, double nB, double out[5]) {
    const int nBlen = (e.iEOS == 12 || e.iEOS == 14) ? 200 : 1;
    auto R = [&](int64_t r, int c) { return e.rows[r * 7 + c]; };
    const double de = R(nBlen, 0) - R(0, 0);
    const int ie = static_cast<int>((ed - R(0, 0)) / de);
    if (ie < 0 || ie >= static_cast<int>(e.n_rows / nBlen) - 2) return -1;
    const int64_t i1 = static_cast<int64_t>(ie) * nBlen, i2 = static_cast<int64_t>(ie + 1) * nBlen;
    const double ef = (ed - R(i1, 0)) / de;
    double f1 = 0., f2 = 0.;
    int n1 = 0, n2 = 0;
    if (nBlen > 1) {
        const double d1 = R(i1 + 1, 1), d2 = R(i2 + 1, 1);
        n1 = min(nBlen - 2, static_cast<int>(nB / d1));
        n2 = min(nBlen - 2, static_cast<int>(nB / d2));
        f1 = fmin(1., (nB - R(i1 + n1, 1)) / d1);
        f2 = fmin(1., (nB - R(i2 + n2, 1)) / d2);
    }
    auto col = [&](int c) {
        const double a = R(i1 + n1, c) * (1. - f1) + R(i1 + n1 + 1, c) * f1;
        const double b = R(i2 + n2, c) * (1. - f2) + R(i2 + n2 + 1, c) * f2;
        return a * (1. - ef) + b * ef;
    };
    out[0] = col(2);
    out[1] = col(3);
    out[2] = out[3] = out[4] = 0.;
    if (nBlen > 1)
        for (int c = 4; c < 7; c++) out[c - 2] = col(c);
    return 0;
}

// One lab-frame MUSIC cell (34 floats, FO_surf order) -> regulation
// (src/readindata.cpp:768-842, 1216-1246) -> rest-frame record
// (src/iSS.cpp:170-293).  Computed in FP64 throughout and stored as float:
// the reference does the boost in float, whose rounding is amplified by
// (u^t)^2 in pi_LRF at large boosts.
__global__ void k_lab_to_lrf(const float* __restrict__ lab, int64_t n, EosDev eos, int use_eos,
                             float* __restrict__ tmp, uint32_t* __restrict__ keep) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* s = lab + 34 * i;
    const double tau = s[0], eta = s[3];
    const double da0 = s[4], da1 = s[5], da2 = s[6], da3 = s[7];
    const double u1 = s[9], u2 = s[10], u3 = s[11];
    double E = s[12], T = s[13], P = s[14];
    const double nB = s[15];
    double muB = s[16], muS = s[17], muQ = s[18];
    const double q1 = s[31], q2 = s[32], q3 = s[33];
    if (use_eos) {
        double v[5];
        if (eos_lookup(eos, E, nB, v) == 0) {
            T = v[1]; muB = v[2]; muS = v[3]; muQ = v[4]; P = v[0];
        }
    }
    const double u0 = sqrt(1. + u1 * u1 + u2 * u2 + u3 * u3);
    const double q0 = (u1 * q1 + u2 * q2 + u3 * q3) / u0;
    // transverse-traceless projection of pi (regulate_Wmunu)
    double pim[4][4];
    {
        const double W[4][4] = {{s[19], s[20], s[21], s[22]},
                                {s[20], s[23], s[24], s[25]},
                                {s[21], s[24], s[26], s[27]},
                                {s[22], s[25], s[27], s[28]}};
        const double u[4] = {u0, u1, u2, u3};
        const double g[4] = {-1., 1., 1., 1.};
        double udp[4];
        for (int a = 0; a < 4; a++)
            udp[a] = -u[0] * W[0][a] + u[1] * W[1][a] + u[2] * W[2][a] + u[3] * W[3][a];
        const double tr = -W[0][0] + W[1][1] + W[2][2] + W[3][3];
        double upu = 0.;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) upu += g[a] * u[a] * W[a][b] * g[b] * u[b];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                const double gab = (a == b) ? g[a] : 0.;
                pim[a][b] = W[a][b] + u[a] * udp[b] + u[b] * udp[a] + u[a] * u[b] * upu -
                            1. / 3. * (gab + u[a] * u[b]) * (tr + upu);
            }
    }
    const double ch = cosh(eta), sh = sinh(eta);
    const double ut = u0 * ch + u3 * sh, uz = u3 * ch + u0 * sh, ux = u1, uy = u2;
    const double ut1 = ut + 1.;
    const double L[4][4] = {
        {ut, -ux, -uy, -uz},
        {-ux, 1. + ux * ux / ut1, ux * uy / ut1, ux * uz / ut1},
        {-uy, ux * uy / ut1, 1. + uy * uy / ut1, uy * uz / ut1},
        {-uz, ux * uz / ut1, uy * uz / ut1, 1. + uz * uz / ut1}};
    const double dup[4] = {tau * da0 * ch - da3 * sh, -tau * da1, -tau * da2,
                           -da3 * ch + tau * da0 * sh};
    double dl[4];
    for (int a = 0; a < 4; a++) {
        double acc = 0.;
        for (int b = 0; b < 4; b++) acc += L[a][b] * dup[b];
        dl[a] = (a == 0) ? acc : -acc;
    }
    keep[i] = (static_cast<float>(dl[0]) < 0.f) ? 0u : 1u;
    const double qtz[4] = {q0 * ch + q3 * sh, q1, q2, q3 * ch + q0 * sh};
    double ql[4];
    for (int a = 0; a < 4; a++) {
        double acc = 0.;
        for (int b = 0; b < 4; b++) acc += L[a][b] * qtz[b];
        ql[a] = acc;
    }
    double Pt[4][4];
    Pt[0][0] = pim[0][0] * ch * ch + 2. * pim[0][3] * ch * sh + pim[3][3] * sh * sh;
    Pt[0][1] = pim[0][1] * ch + pim[1][3] * sh;
    Pt[0][2] = pim[0][2] * ch + pim[2][3] * sh;
    Pt[0][3] = pim[0][0] * ch * sh + pim[0][3] * (ch * ch + sh * sh) + pim[3][3] * sh * ch;
    Pt[1][0] = Pt[0][1];
    Pt[1][1] = pim[1][1];
    Pt[1][2] = pim[1][2];
    Pt[1][3] = pim[0][1] * sh + pim[1][3] * ch;
    Pt[2][0] = Pt[0][2];
    Pt[2][1] = pim[1][2];
    Pt[2][2] = pim[2][2];
    Pt[2][3] = pim[0][2] * sh + pim[2][3] * ch;
    Pt[3][0] = Pt[0][3];
    Pt[3][1] = Pt[1][3];
    Pt[3][2] = Pt[2][3];
    Pt[3][3] = pim[0][0] * sh * sh + 2. * pim[0][3] * sh * ch + pim[3][3] * ch * ch;
    double PL[3][3];
    for (int a = 1; a < 4; a++)
        for (int b = a; b < 4; b++) {
            double acc = 0.;
            for (int c = 0; c < 4; c++)
                for (int d = 0; d < 4; d++) acc += L[a][c] * Pt[c][d] * L[d][b];
            PL[a - 1][b - 1] = acc;
        }
    const double rec[kCellFields] = {tau, s[1], s[2], eta, dl[0], dl[1], dl[2], dl[3],
                                     ut, ux, uy, uz, E, T, P, nB, muB, muS, muQ, s[29],
                                     PL[0][0], PL[0][1], PL[0][2], PL[1][1], PL[1][2],
                                     ql[1], ql[2], ql[3]};
    for (int f = 0; f < kCellFields; f++) tmp[f * n + i] = static_cast<float>(rec[f]);
}

// order-preserving compaction of kept cells into the field-major record
__global__ void k_compact(const float* __restrict__ tmp, const uint32_t* __restrict__ keep,
                          const int64_t* __restrict__ pos, int64_t n, int64_t nk,
                          float* __restrict__ cells) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int64_t o = pos[i];
    for (int f = 0; f < kCellFields; f++) cells[f * nk + o] = tmp[f * n + i];
}

// the cells dropped for u.dsigma < 0, in input order, into their own record
__global__ void k_compact_dropped(const float* __restrict__ tmp, const uint32_t* __restrict__ keep,
                                  const int64_t* __restrict__ pos, int64_t n, int64_t nd,
                                  float* __restrict__ cells) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || keep[i]) return;
    const int64_t o = i - pos[i];
    for (int f = 0; f < kCellFields; f++) cells[f * nd + o] = tmp[f * n + i];
}

// AoS (28 floats, FO_surf_LRF order) -> field-major
__global__ void k_aos_to_soa(const float* __restrict__ aos, int64_t n, float* __restrict__ cells) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int f = 0; f < kCellFields; f++) cells[f * n + i] = aos[28 * i + f];
}

__global__ void k_soa_to_aos(const float* __restrict__ cells, int64_t n, float* __restrict__ aos) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int f = 0; f < kCellFields; f++) aos[28 * i + f] = cells[f * n + i];
}

// ================================================== per-cell coefficients

struct DeltafDev {
    int bulk_kind, include_bulk, neos;  // neos: -1, 0 (22-mom), 1 (CE)
    int include_diffusion;
    const double* qmu_tab;  // [150][100] diffusion coefficients [iT][imu]
    int k11_nT, k11_nmu;
    double k11_T0, k11_mu0, k11_dT, k11_dmu;
    const double* k11;   // [3][nT*nmu]
    int neos_ncol;
    const double* neos_tab;
    double neos_de;
};

// getbulkvisCoefficients(T) for kind 1 (src/FSSW.cpp:1108-1136): the
// relaxation-time fit polynomials in T [1/fm]
__device__ void coef_rta(double T, double c[6]) {
    // the fit's alternating terms reach 1e8 while the result is O(10): keep
    // the reference's unfused multiply-then-add order
#pragma clang fp contract(off)
    const double a0[11] = {642096.624265727, -8163329.49562861, 47162768.4292073,
                           -162590040.002683, 369637951.096896, -578181331.809836,
                           629434830.225675, -470493661.096657, 230936465.421,
                           -67175218.4629078, 8789472.32652964};
    const double a1[11] = {1.18171174036192, -17.6740645873717, 136.298469057177,
                           -635.999435106846, 1918.77100633321, -3836.32258307711,
                           5136.35746882372, -4566.22991441914, 2593.45375240886,
                           -853.908199724349, 124.260460450113};
    const double x = T / kHbarC;
    double pw = 1., s0 = a0[0], s1 = a1[0];
    for (int i = 1; i < 11; i++) {
        pw *= x;
        s0 += a0[i] * pw;
        s1 += a1[i] * pw;
    }
    c[0] = s0;
    c[1] = s1;
}

// getbulkvisCoefficients(T, muB) (src/FSSW.cpp:1379-1430)
__device__ void coef_14mom(const DeltafDev& d, double T, double muB, double c[6]) {
    const int iT = static_cast<int>((T - d.k11_T0) / d.k11_dT);
    const int im = static_cast<int>((muB - d.k11_mu0) / d.k11_dmu);
    const double xf = (T - d.k11_T0) / d.k11_dT - iT;
    const double yf = (muB - d.k11_mu0) / d.k11_dmu - im;
    const int T1 = max(0, min(d.k11_nT - 1, iT)), T2 = max(0, min(d.k11_nT - 1, iT + 1));
    const int m1 = max(0, min(d.k11_nmu - 1, im)), m2 = max(0, min(d.k11_nmu - 1, im + 1));
    const int n = d.k11_nmu, len = d.k11_nT * d.k11_nmu;
    double v[3];
    for (int k = 0; k < 3; k++) {
        const double* t = d.k11 + k * len;
        v[k] = t[T1 * n + m1] * (1. - xf) * (1. - yf) + t[T1 * n + m2] * (1. - xf) * yf +
               t[T2 * n + m2] * xf * yf + t[T2 * n + m1] * xf * (1. - yf);
    }
    const double T3 = T * T * T, T4 = T3 * T;
    c[0] = (v[0] - v[2]) / T4;
    c[1] = v[1] / T3;
    c[2] = (4. * v[2] - v[0]) / T4;
}

// getCENEOSBQSCoefficients / get22momNEOSBQSCoefficients (src/FSSW.cpp:1433-1543)
__device__ void coef_neos(const DeltafDev& d, double Ed, double nB, double c[6]) {
    const int nc = d.neos_ncol;
    auto at = [&](int r, int k) { return d.neos_tab[r * nc + k]; };
    int ie = static_cast<int>((Ed - at(0, 0)) / d.neos_de);
    ie = max(0, min(198, ie));
    const int r1 = ie * 200, r2 = (ie + 1) * 200;
    const double ef = (Ed - at(r1, 0)) / d.neos_de;
    const double d1 = at(r1 + 1, 1), d2 = at(r2 + 1, 1);
    const int n1 = min(198, static_cast<int>(nB / d1)), n2 = min(198, static_cast<int>(nB / d2));
    const double f1 = fmin(1., (nB - at(r1 + n1, 1)) / d1);
    const double f2 = fmin(1., (nB - at(r2 + n2, 1)) / d2);
    double ic[6];
    for (int i = 2; i < nc; i++) {
        const double a = at(r1 + n1, i) * (1 - f1) + at(r1 + n1 + 1, i) * f1;
        const double b = at(r2 + n2, i) * (1 - f2) + at(r2 + n2 + 1, i) * f2;
        ic[i - 2] = a * (1. - ef) + b * ef;
    }
    if (nc == 5) {
        c[0] = 1. / ic[1];
        c[1] = 1. / 3. - ic[0];
        c[2] = ic[2];
    } else {
        for (int i = 0; i < 6; i++) c[i] = ic[i];
    }
}

// get_deltaf_qmu_coeff (src/FSSW.cpp:1569-1606): bilinear in (T, mu_B) on
// T = 0.05 + 0.001 i (150 rows), mu_B = 0.007892 j (100 columns); 1e30
// (no diffusion correction) outside the table
__device__ double coef_qmu(const DeltafDev& d, double T, double muB) {
    const double T0 = 0.05, dT = 0.001, mu0 = 0.0, dmu = 0.007892;
    const int nT = 150, nmu = 100;
    const int iT = static_cast<int>((T - T0) / dT);
    const int imu = static_cast<int>((muB - mu0) / dmu);
    const double xf = (T - T0) / dT - iT, yf = (muB - mu0) / dmu - imu;
    if (imu > nmu - 2 || iT > nT - 2 || imu < 0 || iT < 0) return 1e30;
    const double* t = d.qmu_tab;
    const double f1 = t[iT * nmu + imu], f2 = t[iT * nmu + imu + 1];
    const double f3 = t[(iT + 1) * nmu + imu + 1], f4 = t[(iT + 1) * nmu + imu];
    return f1 * (1. - xf) * (1. - yf) + f2 * (1. - xf) * yf + f3 * xf * yf + f4 * xf * (1. - yf);
}

// E_2, E_4, ..., E_18 on the K-table grid (src/FSSW.cpp:1637-1642)
__global__ void k_expint_table(double* etab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kKLen) return;
    const double x = kKx0 + i * kKdx;
    etab[i] = expint_en(2, x);
    for (int k = 1; k < kEnTerms; k++) etab[k * kKLen + i] = expint_en(2 * k + 2, x);
}

// Per-cell coefficients as the yield loop selects them (src/FSSW.cpp:597-637);
// rows 6, 7: 1 / deltaf_qmu_coeff and n_B / (e + P) of the diffusion delta f
__global__ void k_cell_prep(const float* __restrict__ cells, int64_t Nc, DeltafDev d,
                            double* __restrict__ coef) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= Nc) return;
    double c[6] = {0., 0., 0., 0., 0., 0.};
    const double T = cells[F_T * Nc + i], muB = cells[F_MUB * Nc + i];
    const double Ed = cells[F_E * Nc + i], nB = cells[F_NB * Nc + i];
    if (d.neos >= 0) coef_neos(d, Ed, nB, c);
    if (d.include_bulk) {
        if (d.bulk_kind == 11) coef_14mom(d, T, muB, c);
        else if (d.bulk_kind == 1) coef_rta(T, c);
    }
    for (int k = 0; k < 6; k++) coef[k * Nc + i] = c[k];
    double iq = 0., pq = 0.;
    if (d.include_diffusion) {
        iq = 1. / coef_qmu(d, T, muB);
        pq = nB / (Ed + static_cast<double>(cells[F_P * Nc + i]));
    }
    coef[6 * Nc + i] = iq;
    coef[7 * Nc + i] = pq;
}


}  // namespace issdev
