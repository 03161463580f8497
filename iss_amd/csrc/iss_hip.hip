// gfx950 Cooper-Frye sampler: kernels and the C-ABI of include/iss_hip.h.
//
// Pipeline per context (all on one HIP stream):
//   k_lab_to_lrf + k_compact  regulation and rest-frame transform of the surface
//   k_cell_prep               per-cell delta-f coefficients (FP64)
//   k_yield<R, MODE>          per-(cell, species) yields, the FP64 species CDF of
//                             every cell and per-block FP64 species partials
//   k_species_reduce          FP64 per-species totals
//   k_counts                  hadrons per cell over all events: Poisson(E * N_c)
//   scan                      cell offsets (int64)
//   k_multinomial_level       hadrons per event: Multinomial(K, 1/E), binomial tree
//   k_sample                  species pick, LRF momentum rejection, boost, output
//
// Poisson splitting: the reference draws N_{e,s} ~ Poisson(sum_c N_{c,s}) per
// event and species and then a cell per hadron from the species' cell CDF
// (src/FSSW.cpp:943-979).  That is the same law as independent
// Poisson(N_{c,s}) per (event, cell, species), which is sampled here as
// Poisson(E * N_c) hadrons per cell, a species per hadron from the cell's
// species CDF, and an event per hadron through a keyed permutation of the
// cell-ordered hadron list onto event slots whose sizes are
// Multinomial(K, 1/E).  Same output law, cache-friendly access, and a fully
// deterministic output for a given seed without atomics or sorting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/iss_hip.h"
#include "kernels.h"

using namespace issdev;

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

// Per-cell coefficients as the yield loop selects them (src/FSSW.cpp:597-620)
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
}

// ============================================================== yield pass

template <typename R>
struct KLds;
template <>
struct KLds<double> {
    using T = double;
};
template <>
struct KLds<float> {
    using T = float2;
};

// LDS slot of K_order: K_2 first (every mode), then K_1, then K_3
__device__ __forceinline__ int kslot(int order) { return order == 2 ? 0 : (order == 1 ? 1 : 2); }

// K_n lookup, FP64: the reference's interpolation verbatim
// (src/FSSW.cpp:1662-1666); rows below kKLds from LDS when staged.  Outside
// [0.5, 399.95] the reference calls the exact function; the fast variant
// (EXACT = false) raises *oob instead and the pass is rerun with EXACT = true.
template <bool EXACT>
struct KLookup64 {
    const double* lds;   // may be null
    const double* glob;
    int* oob;
    __device__ __forceinline__ double operator()(int order, double arg) const {
        if (arg < kKx0 || arg > kKxmax - kKdx) {
            if constexpr (EXACT) return kn_exact(order, arg);
            *oob = 1;
            return 0.;
        }
        const int idx = static_cast<int>((arg - kKx0) / kKdx);
        const double fr = (arg - kKx0 - idx * kKdx) / kKdx;
        const double* t = (lds && idx + 1 < kKLds) ? lds + kslot(order) * kKLds
                                                   : glob + (order - 1) * kKLen;
        return (1. - fr) * t[idx] + fr * t[idx + 1];
    }
};

// K_n lookup, FP32: same grid, K_i + f (K_{i+1} - K_i)
template <bool EXACT>
struct KLookup32 {
    const float2* lds;
    const float2* glob;
    int* oob;
    __device__ __forceinline__ float operator()(int order, float arg) const {
        if (arg < 0.5f || arg > 399.95f) {
            if constexpr (EXACT) return static_cast<float>(kn_exact(order, arg));
            *oob = 1;
            return 0.f;
        }
        const float t = (arg - 0.5f) * 20.f;
        const int idx = static_cast<int>(t);
        const float fr = t - static_cast<float>(idx);
        const float2 v = (idx < kKLds) ? lds[kslot(order) * kKLds + idx] : glob[(order - 1) * kKLen + idx];
        return fmaf(fr, v.y, v.x);
    }
};

// calculate_dN_analytic + the per-cell combination and clamp
// (src/FSSW.cpp:645-706, 719-848) for one (cell, species).  `ill` reports
// (|N_eq| + |dN_bulk|) > 4 |N|: rounding errors are amplified when the bulk
// correction nearly cancels the equilibrium yield.

// exp(x) and 1/n: correctly rounded in FP64 (parity with the reference);
// the hardware exp2 / reciprocal in FP32 (within the 1e-5 FP32 budget).
__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ double frcp(double x) { return 1. / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename R, int MODE, typename KL>
__device__ __forceinline__ R yield_one(R mass, R pref, R Bq, R Sq, R Qq, R sgn, R T, R beta,
                                       R muB, R muS, R muQ, R dsu, R Pi, const R* c,
                                       const KL& kl, bool* ill) {
    const R mu = Bq * muB + Sq * muS + Qq * muQ;
    const R lambda = fexp(beta * mu);
    const int nt = (mass < R(0.7) && T > R(0.05)) ? 10 : 1;
    const R mb = mass * beta;
    R r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    R th = 1, fug = lambda;
#pragma unroll 1
    for (int n = 1; n <= nt; n++) {
        const R rn = static_cast<R>(n);
        const R inv = (sizeof(R) == sizeof(double)) ? R(1) : frcp(rn);
        const R arg = rn * mass * beta;
        const R K2 = kl(2, arg);
        // FP64: theta / n as the reference divides; FP32: times 1/n
        const R thn = (sizeof(R) == sizeof(double)) ? th / rn : th * inv;
        const R k2n = (sizeof(R) == sizeof(double)) ? 3 * K2 / rn : 3 * K2 * inv;
        r0 += thn * fug * K2;
        if (MODE == BULK_RTA || MODE == BULK_CE) {
            const R K1 = kl(1, arg);
            r1 += th * fug * (mb * K1 + k2n);
            r2 += th * fug * K1;
        } else if (MODE == BULK_14MOM || MODE == BULK_22MOM) {
            const R K1 = kl(1, arg);
            const R K3 = kl(3, arg);
            const R k3n = (sizeof(R) == sizeof(double)) ? 3 * K3 / rn : 3 * K3 * inv;
            r1 += th * fug * K2;
            r2 += th * fug * (mb * K1 + k2n);
            r3 += th * fug * (mb * K2 + k3n);
        }
        th *= -sgn;
        fug *= lambda;
    }
    const R m2 = mass * mass;
    const R pd = pref * dsu;
    const R Neq = pd * (m2 * T * r0);
    R dN = 0;
    if (MODE == BULK_RTA || MODE == BULK_CE) {
        const R b1 = m2 * T * r1, b2 = m2 * mass / R(3) * r2;
        dN = pd * (-Pi * c[0]) * (-c[1] * b1 + b2);
    } else if (MODE == BULK_14MOM) {
        const R T2 = T * T;
        const R b1 = m2 * T * r1, b2 = m2 * T2 * r2, b3 = m2 * mass * T2 * r3;
        dN = pd * Pi * (b1 * m2 * c[0] + b2 * Bq * c[1] + b3 * c[2]);
    } else if (MODE == BULK_22MOM) {
        const R T2 = T * T;
        const R b1 = m2 * T * r1, b2 = m2 * T2 * r2, b3 = m2 * mass * T2 * r3;
        dN = pd * Pi * (b1 * m2 * c[2] + b2 * (Bq * c[3] + Sq * c[4] + Qq * c[5]) + b3 * (c[1] - c[2]));
    }
    const R N = Neq + dN;
    // ill-conditioned: the bulk correction nearly cancels N_eq
    if (ill) *ill = (fabs(Neq) + fabs(dN)) > R(4) * fabs(N);
    return N > R(0) ? N : R(0);
}

// FP32 yields whose bulk correction nearly cancels N_eq ((|N_eq| + |dN|) >
// 4 |N|) are recomputed in FP64 so that every FP32 yield stays within 1e-5
// of FP64.

// FP64 yield of one (cell, species) from global memory only: the fallback
// for ill-conditioned FP32 entries (kept out of line: it is rare).
template <int MODE, bool EXACT>
__device__ double yield_fp64_global(const YieldArgs& a, int64_t c, int s) {
    const int64_t Nc = a.Nc;
    const KLookup64<EXACT> k64{nullptr, a.ktab, a.oob};
    const double T = a.cells[F_T * Nc + c];
    double cd[6];
    for (int k = 0; k < 6; k++) cd[k] = a.coef[k * Nc + c];
    double Pi = a.cells[F_PI * Nc + c];
    if (MODE == BULK_RTA) Pi = Pi / kHbarC;
    const double pref = static_cast<double>(a.sp.g[s]) * (1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi));
    return yield_one<double, MODE>(a.sp.mass[s], pref, a.sp.B[s], a.sp.S[s], a.sp.Q[s], a.sp.sign[s], T,
                                   1. / T, (double)a.cells[F_MUB * Nc + c], (double)a.cells[F_MUS * Nc + c],
                                   (double)a.cells[F_MUQ * Nc + c], (double)a.cells[F_DA0 * Nc + c], Pi, cd,
                                   k64, (bool*)nullptr);
}

constexpr int kYieldThreads = 256;
// ill-conditioned FP32 items recomputed in FP64 after each tile; a tile with
// more than this many falls back to recomputing inline
constexpr int kFixupCap = 1024;

__host__ __device__ constexpr int ktables_for(int mode) {
    return mode == BULK_NONE ? 1 : ((mode == BULK_RTA || mode == BULK_CE) ? 2 : 3);
}

// LDS bytes of one k_yield workgroup for nsp (padded) species
__host__ __device__ constexpr size_t yield_lds_bytes(int esz, int ntab, int tc, int nsp) {
    return static_cast<size_t>(nsp) * (2 * esz + 4)                 // mass, pref, charges
           + static_cast<size_t>(ntab) * kKLds * 8                  // K rows
           + static_cast<size_t>(tc) * nsp * esz                    // yield tile
           + static_cast<size_t>(kYieldThreads / kWave) * nsp * 8   // FP64 row buffers
           + static_cast<size_t>(tc) * 16 * esz                     // cell parameters
           + static_cast<size_t>(kFixupCap + 1) * 4;                // deferred FP64 items
}

__host__ __device__ constexpr int pad_species(int ns) { return (ns + 3) & ~3; }

// Workgroup-tiled yield pass.  A workgroup owns a tile of TC cells at a time
// (persistent over tiles):
//   1. items (species s, cell c) in species-major order: 16 consecutive lanes
//      share a species (LDS broadcast) and the 10-term light species fill
//      whole waves; raw yields go to an LDS tile;
//   2. one wave per cell row: lanes take contiguous species blocks, an FP64
//      local prefix plus one wave scan gives the species CDF, which is staged
//      in an LDS row buffer and stored coalesced;
//   3. FP64 column sums of the tile feed per-thread species accumulators,
//      written once per workgroup as partials.
template <typename R, int MODE, bool RAW, int TC, bool EXACT>
__global__ __launch_bounds__(kYieldThreads) void k_yield(YieldArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = ktables_for(MODE);
    using KT = typename KLds<R>::T;
    const int Ns = a.Ns;
    const int nsp = pad_species(Ns);
    R* s_mass = reinterpret_cast<R*>(smem);
    R* s_pref = s_mass + nsp;
    char4* s_q = reinterpret_cast<char4*>(s_pref + nsp);
    KT* s_k = reinterpret_cast<KT*>(s_q + nsp);
    R* s_tile = reinterpret_cast<R*>(s_k + NT * kKLds);
    double* s_row = reinterpret_cast<double*>(s_tile + TC * nsp);
    R* s_cell = reinterpret_cast<R*>(s_row + (kYieldThreads / kWave) * nsp);
    int* s_fix = reinterpret_cast<int*>(s_cell + TC * 16);  // [0] count, then (cl << 16 | s)

    const int64_t Nc = a.Nc;
    const R uf = static_cast<R>(1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi));
    for (int s = threadIdx.x; s < Ns; s += blockDim.x) {
        s_mass[s] = static_cast<R>(a.sp.mass[s]);
        s_pref[s] = static_cast<R>(a.sp.g[s]) * uf;
        s_q[s] = make_char4(a.sp.B[s], a.sp.S[s], a.sp.Q[s], a.sp.sign[s]);
    }
    // K_2 first, then K_1, K_3 as the mode needs
    for (int i = threadIdx.x; i < NT * kKLds; i += blockDim.x) {
        const int t = i / kKLds, r = i % kKLds;
        const int o = (t == 0) ? 2 : (t == 1 ? 1 : 3);
        if constexpr (sizeof(R) == sizeof(double))
            s_k[i] = a.ktab[(o - 1) * kKLen + r];
        else
            s_k[i] = a.ktabf[(o - 1) * kKLen + r];
    }
    // this thread's two species accumulators (Ns <= 2 * blockDim)
    double acc0 = 0., acc1 = 0.;
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = threadIdx.x / kWave;
    const int64_t ntiles = (Nc + TC - 1) / TC;
    const int items = TC * Ns;
    // lane-contiguous species block for the row scan
    const int per = (Ns + kWave - 1) / kWave;
    const int s_lo = min(lane * per, Ns), s_hi = min(s_lo + per, Ns);

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t c0 = tile * TC;
        const int nc = static_cast<int>(min<int64_t>(TC, Nc - c0));
        __syncthreads();  // previous tile's LDS reads are done
        if (threadIdx.x == 0) s_fix[0] = 0;
        if (threadIdx.x < nc) {
            // parameter-major [k][cell]: the 16 lanes of a species read 16
            // consecutive banks
            const int64_t c = c0 + threadIdx.x;
            R* p = s_cell + threadIdx.x;
            const R T = a.cells[F_T * Nc + c];
            p[0 * TC] = T;
            p[1 * TC] = R(1) / T;
            p[2 * TC] = a.cells[F_MUB * Nc + c];
            p[3 * TC] = a.cells[F_MUS * Nc + c];
            p[4 * TC] = a.cells[F_MUQ * Nc + c];
            p[5 * TC] = a.cells[F_DA0 * Nc + c];
            R Pi = 0;
            if (MODE != BULK_NONE) {
                Pi = a.cells[F_PI * Nc + c];
                if (MODE == BULK_RTA) Pi = Pi / static_cast<R>(kHbarC);
                for (int k = 0; k < 6; k++) p[(7 + k) * TC] = static_cast<R>(a.coef[k * Nc + c]);
            }
            p[6 * TC] = Pi;
        }
        __syncthreads();
        // 1. yields into the LDS tile
        for (int it = threadIdx.x; it < items; it += blockDim.x) {
            const int s = it / TC, cl = it % TC;
            R y = 0;
            if (cl < nc) {
                const R* q0 = s_cell + cl;
                const R p[13] = {q0[0], q0[TC], q0[2 * TC], q0[3 * TC], q0[4 * TC], q0[5 * TC], q0[6 * TC],
                                 q0[7 * TC], q0[8 * TC], q0[9 * TC], q0[10 * TC], q0[11 * TC], q0[12 * TC]};
                const char4 q = s_q[s];
                if constexpr (sizeof(R) == sizeof(double)) {
                    const KLookup64<EXACT> kl{reinterpret_cast<const double*>(s_k), a.ktab, a.oob};
                    y = yield_one<R, MODE>(s_mass[s], s_pref[s], q.x, q.y, q.z, q.w, p[0], p[1], p[2],
                                           p[3], p[4], p[5], p[6], p + 7, kl, (bool*)nullptr);
                } else {
                    const KLookup32<EXACT> kl{reinterpret_cast<const float2*>(s_k), a.ktabf, a.oob};
                    bool ill = false;
                    y = yield_one<R, MODE>(s_mass[s], s_pref[s], q.x, q.y, q.z, q.w, p[0], p[1], p[2],
                                           p[3], p[4], p[5], p[6], p + 7, kl,
                                           MODE == BULK_NONE ? (bool*)nullptr : &ill);
                    if (MODE != BULK_NONE && ill) {
                        const int k = atomicAdd(&s_fix[0], 1);
                        if (k < kFixupCap) s_fix[1 + k] = (cl << 16) | s;
                        else y = static_cast<R>(yield_fp64_global<MODE, EXACT>(a, c0 + cl, s));
                    }
                }
            }
            s_tile[cl * nsp + s] = y;
        }
        __syncthreads();
        if constexpr (sizeof(R) == sizeof(float) && MODE != BULK_NONE) {
            // 1b. deferred FP64 recomputation of the ill-conditioned items
            const int nfix = min(s_fix[0], kFixupCap);
            for (int k = threadIdx.x; k < nfix; k += blockDim.x) {
                const int e = s_fix[1 + k];
                const int cl = e >> 16, sp = e & 0xFFFF;
                s_tile[cl * nsp + sp] = static_cast<R>(yield_fp64_global<MODE, EXACT>(a, c0 + cl, sp));
            }
            if (nfix > 0) __syncthreads();
        }
        // 2. species CDF rows: one wave per row
        double* rowbuf = s_row + wib * nsp;
        for (int cl = wib; cl < nc; cl += blockDim.x / kWave) {
            const R* yrow = s_tile + cl * nsp;
            double run = 0.;
            for (int s = s_lo; s < s_hi; s++) run += static_cast<double>(yrow[s]);
            const double excl = wave_inclusive_scan(run) - run;
            double pre = excl;
            for (int s = s_lo; s < s_hi; s++) {
                pre += static_cast<double>(yrow[s]);
                rowbuf[s] = pre;
            }
            const double total = __shfl(excl + run, kWave - 1, kWave);
            // rowbuf is written and read by this wave only
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int64_t c = c0 + cl;
            double* row = a.cdf + c * Ns;
            for (int s = lane; s < Ns; s += kWave) {
                row[s] = rowbuf[s];
                if (RAW) static_cast<R*>(a.raw)[c * Ns + s] = yrow[s];
            }
            if (lane == 0) a.ctot[c] = total;
            __builtin_amdgcn_wave_barrier();
        }
        // 3. FP64 column sums into the species accumulators
        for (int s = threadIdx.x; s < Ns; s += blockDim.x) {
            double v = 0.;
            for (int cl = 0; cl < nc; cl++) v += static_cast<double>(s_tile[cl * nsp + s]);
            if (s == static_cast<int>(threadIdx.x)) acc0 += v; else acc1 += v;
        }
    }
    for (int s = threadIdx.x; s < Ns; s += blockDim.x)
        a.partial[static_cast<int64_t>(blockIdx.x) * Ns + s] = (s == static_cast<int>(threadIdx.x)) ? acc0 : acc1;
}

// FP64 per-species totals: one workgroup per species, fixed summation order
__global__ void k_species_reduce(const double* __restrict__ partial, int nblk, int Ns,
                                 double* __restrict__ out) {
    __shared__ double red[256];
    const int s = blockIdx.x;
    double v = 0.;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) v += partial[static_cast<int64_t>(b) * Ns + s];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = red[0];
}

// ============================================================ counting pass

enum StreamTag : uint32_t { TAG_COUNT = 1, TAG_MULTI = 2, TAG_SETUP = 3, TAG_TRIAL = 4 };

__device__ __forceinline__ uint32_t tag_word(uint32_t tag, uint32_t chunk) {
    return (tag << 24) ^ (chunk & 0x00FFFFFFu);
}

// hadrons of cell c over all E events of this pass: Poisson(E * scale * N_c)
__global__ void k_counts(const double* __restrict__ ctot, int64_t Nc, double mult, uint64_t seed,
                         uint32_t chunk, uint32_t* __restrict__ counts) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= Nc) return;
    Stream st(seed, static_cast<uint32_t>(c), static_cast<uint32_t>(c >> 32), tag_word(TAG_COUNT, chunk));
    const int64_t k = poisson(mult * ctot[c], st);
    counts[c] = static_cast<uint32_t>(min<int64_t>(k, 0xFFFFFFFFll));
}

// ----- exclusive scan of uint32 -> int64 (out[n] = total), 3 phases
constexpr int kScanBlock = 256, kScanItems = 8, kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[kScanBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int64_t incl = wave_inclusive_scan(v);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int64_t pre = 0, tot = 0;
    for (int i = 0; i < kScanBlock / kWave; i++) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return pre + incl - v;
}

__global__ void k_scan_tiles(const uint32_t* __restrict__ in, int64_t n, int64_t* __restrict__ tile_sums) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    int64_t v = 0;
    for (int k = 0; k < kScanItems; k++)
        if (base + k < n) v += in[base + k];
    int64_t tot;
    block_exclusive_scan(v, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_top(int64_t* __restrict__ tile_sums, int64_t ntiles) {
    // one workgroup: exclusive scan in place, loops over the tile sums
    __shared__ int64_t carry_s;
    if (threadIdx.x == 0) carry_s = 0;
    __syncthreads();
    for (int64_t b = 0; b < ntiles; b += kScanBlock) {
        const int64_t i = b + threadIdx.x;
        const int64_t v = (i < ntiles) ? tile_sums[i] : 0;
        int64_t tot;
        const int64_t ex = block_exclusive_scan(v, &tot);
        const int64_t carry = carry_s;
        if (i < ntiles) tile_sums[i] = ex + carry;
        __syncthreads();
        if (threadIdx.x == 0) carry_s = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sums[ntiles] = carry_s;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ in, int64_t n,
                             const int64_t* __restrict__ tile_sums, int64_t* __restrict__ out) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    int64_t loc[kScanItems];
    int64_t v = 0;
    for (int k = 0; k < kScanItems; k++) {
        loc[k] = (base + k < n) ? in[base + k] : 0;
        v += loc[k];
    }
    int64_t tot;
    int64_t ex = block_exclusive_scan(v, &tot) + tile_sums[blockIdx.x];
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += loc[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = tile_sums[gridDim.x];
}

// ----- hadrons per event: Multinomial(K, 1/E) as a binomial tree.  Level d
// has 2^d nodes; node k covers events [floor(kE/2^d), floor((k+1)E/2^d)).
__global__ void k_multinomial_level(const int64_t* __restrict__ parent, int64_t* __restrict__ child,
                                    int d, int E, uint64_t seed, uint32_t chunk) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nodes = 1ll << d;
    if (k >= nodes) return;
    const int64_t n = parent[k];
    const int64_t lo = (k * E) >> d, hi = ((k + 1) * E) >> d;
    const int64_t mid = ((2 * k + 1) * E) >> (d + 1);
    int64_t nl = 0;
    if (hi > lo && n > 0) {
        const double p = static_cast<double>(mid - lo) / static_cast<double>(hi - lo);
        Stream st(seed, static_cast<uint32_t>(k), static_cast<uint32_t>(d), tag_word(TAG_MULTI, chunk));
        nl = binomial(n, p, st);
    }
    child[2 * k] = nl;
    child[2 * k + 1] = n - nl;
}

// leaves (level D, 2^D >= E): node k covers at most one event
__global__ void k_multinomial_leaves(const int64_t* __restrict__ leaf, int D, int E,
                                     uint32_t* __restrict__ evcount) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= (1ll << D)) return;
    const int64_t lo = (k * E) >> D, hi = ((k + 1) * E) >> D;
    if (hi - lo == 1) evcount[lo] = static_cast<uint32_t>(leaf[k]);
}

// ============================================================ sampling pass

struct Hadron {
    int pid;
    float mass, E, px, py, pz, t, x, y, z;
};

// CDF_0, CDF_1, CDF_2 of one |p| table at E~ together: the per-term
// exponentials become powers of exp(m0 - E~) and exp(-m0), so one hadron's
// normalisation costs two exponentials and a logarithm.
__device__ void pcdf_all(const PTabDev& pt, int tb, double E, double out[3]) {
    const int fer = pt.fermion[tb], trunc = pt.trunc[tb], quirk = pt.quirk;
    const double m = pt.m0[tb];
    const double r = exp(m - E), em1 = pt.em1[tb];
    double c0 = 0., c1 = 0., c2 = 0., sign = 1., em = 1., ee = r;
    for (int n = 0; n < trunc; n++) {
        const int n1 = n + 1;
        const double dn = static_cast<double>(n1);
        if (trunc <= 5)
            c0 += (fer ? static_cast<double>(static_cast<int>(sign) / n1) : 1. / dn) * em * (1. - ee);
        const double s1 = fer ? sign : 1.;
        c1 += s1 / (dn * dn) * em * ((m * dn + 1) - ee * (E * dn + 1));
        const double s2 = (fer && !quirk) ? sign : 1.;
        c2 += s2 / (dn * dn * dn) * em * ((m * dn * (m * dn + 2) + 2) - ee * (E * dn * (E * dn + 2) + 2));
        sign = -sign;
        em *= em1;
        ee *= r;
    }
    if (trunc > 5) {
        const double eE = exp(-E);
        c0 = fer ? -pt.ep[tb] * log((1. + eE) / (1. + em1)) : pt.ep[tb] * log((1. - eE) / (1. - em1));
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
}

// Persistent sampling kernel with per-lane refill.  Each wavefront owns a
// contiguous range of the cell-ordered hadron slots; a lane that finishes its
// hadron (accepted, or out of trials) takes the next slot of the range, so
// the geometric tails of the rejection loops do not idle the other 63 lanes.
// Slot i of cell c (the j-th of the cell) draws its species from the cell's
// FP64 species CDF, its event slot through the permutation, and its
// rest-frame momentum by the reference's nested rejection
// (src/FSSW.cpp:1852-1966, src/MomentumSamplerBase.cpp:20-58) flattened into
// one loop: an iteration is one |p| proposal plus, when the |p| test passes,
// one outer trial.  Iteration `it` of slot (c, j) uses Philox block
// (c, j, tag, it): word 0 is the 32-bit CDF draw, the other 96 bits give
// four 24-bit uniforms.  The stream depends on (c, j) only, so the output
// does not depend on the launch shape.
// free lanes that trigger a refill: setup then runs at least this wide
constexpr int kRefillMin = 16;

__global__ __launch_bounds__(256) void k_sample(SampleArgs a) {
    __shared__ unsigned long long s_stats[4];
    if (threadIdx.x < 4) s_stats[threadIdx.x] = 0ull;
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    const int64_t wid = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int64_t base = a.K / nwaves, rem = a.K % nwaves;
    const int64_t beg = wid * base + min(wid, rem);
    const int64_t end = beg + base + (wid < rem ? 1 : 0);
    const int64_t Nc = a.Nc;
    const PTabDev& pt = a.pt;
    const uint32_t key0 = static_cast<uint32_t>(a.seed), key1 = static_cast<uint32_t>(a.seed >> 32);
    const uint32_t tw = tag_word(TAG_TRIAL, a.chunk);
    const int max_it = a.max_trials * 64;

    int64_t next = beg;
    int64_t hint = (beg < end) ? upper_bound(a.off + 1, int64_t(0), Nc, beg) : 0;
    unsigned long long n_trial = 0, n_inner = 0, n_fail = 0, n_range = 0;

    // per-lane hadron state
    bool active = false;
    int64_t slot = 0, c = 0;
    uint32_t jj = 0;
    int s = 0, tb = 0, i0 = 0, imax = 0, it = 0, trials = 1;
    double mut = 0., aa = 0., mterm = 0., cmax = 0., rmin0 = 0., x = 0., mu = 0., Td = 1.;
    float T = 1.f, fm = 0.f, fmu = 0.f, fsg = 0.f, dsfac = 1.f, dfpref = 0.f;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f, Pi = 0.f, Bq = 0.f, Sq = 0.f, Qq = 0.f;
    float pixx = 0.f, pixy = 0.f, pixz = 0.f, piyy = 0.f, piyz = 0.f;
    float cf0 = 0.f, cf1 = 0.f, cf2 = 0.f, cf3 = 0.f, cf4 = 0.f, cf5 = 0.f;

    while (true) {
        // ---- refill lanes without a hadron from this wave's slot range
        const unsigned long long need = __ballot(!active);
        if (need && next < end && (__popcll(need) >= kRefillMin || need == ~0ull)) {
            const int64_t navail = end - next;
            const int rank = __popcll(need & ((1ull << lane) - 1));
            const bool take = !active && rank < navail;
            const unsigned long long took = __ballot(take);
            if (take) {
                slot = next + rank;
                // cell of the slot: exponential then binary search from the hint
                int64_t lo = hint, step = 1, hi = hint + 1;
                while (hi < Nc && a.off[hi + 1] <= slot) {
                    lo = hi;
                    step <<= 1;
                    hi = min<int64_t>(Nc - 1, hi + step);
                }
                hi = min<int64_t>(hi + 1, Nc);
                c = upper_bound(a.off + 1, lo, hi, slot);
                jj = static_cast<uint32_t>(slot - a.off[c]);
                Stream st(a.seed, static_cast<uint32_t>(c), jj, tag_word(TAG_SETUP, a.chunk));
                const double* row = a.cdf + c * a.Ns;
                s = upper_bound(row, 0, a.Ns, st.uniform() * a.ctot[c]);
                if (s >= a.Ns) s = a.Ns - 1;
                T = a.cells[F_T * Nc + c];
                const float muB = a.cells[F_MUB * Nc + c], muS = a.cells[F_MUS * Nc + c],
                            muQ = a.cells[F_MUQ * Nc + c];
                const float Ed = a.cells[F_E * Nc + c], Pd = a.cells[F_P * Nc + c];
                d0 = a.cells[F_DA0 * Nc + c]; d1 = a.cells[F_DA1 * Nc + c];
                d2 = a.cells[F_DA2 * Nc + c]; d3 = a.cells[F_DA3 * Nc + c];
                const double mass = a.sp.mass[s];
                const int sign = a.sp.sign[s];
                Bq = a.sp.B[s]; Sq = a.sp.S[s]; Qq = a.sp.Q[s];
                mu = fmin(mass, static_cast<double>(Bq * muB + Sq * muS + Qq * muQ));
                Td = T;
                const double mt = mass / Td;
                mut = mu / Td;
                x = mt - mut;
                tb = ((sign == -1) ? 0 : 3) + ((x < 30.) ? 0 : (x < 50. ? 1 : 2));
                aa = mut * mut - mt * mt / 2.;
                double cm[3];
                pcdf_all(pt, tb, x, cm);
                mterm = cm[2] + 2. * mut * cm[1] + aa * cm[0];
                imax = pt.len[tb] - 1;
                i0 = static_cast<int>((x - pt.Et0[tb]) / pt.dEt[tb]);
                fm = static_cast<float>(mass);
                fmu = static_cast<float>(mu);
                fsg = static_cast<float>(sign);
                dsfac = fabsf(d0) + sqrtf(d1 * d1 + d2 * d2 + d3 * d3);
                dfpref = 1.0f / (2.0f * T * T * (Ed + Pd));
                if (a.shear) {
                    pixx = a.cells[F_PIXX * Nc + c]; pixy = a.cells[F_PIXY * Nc + c];
                    pixz = a.cells[F_PIXZ * Nc + c]; piyy = a.cells[F_PIYY * Nc + c];
                    piyz = a.cells[F_PIYZ * Nc + c];
                }
                if (a.bulk != BULK_NONE) {
                    Pi = a.cells[F_PI * Nc + c];
                    if (a.bulk == BULK_RTA) Pi = Pi / static_cast<float>(kHbarC);
                }
                if (a.bulk != BULK_NONE || a.neos >= 0) {
                    cf0 = a.coef[c]; cf1 = a.coef[Nc + c]; cf2 = a.coef[2 * Nc + c];
                    cf3 = a.coef[3 * Nc + c]; cf4 = a.coef[4 * Nc + c]; cf5 = a.coef[5 * Nc + c];
                }
                it = 0;
                trials = 1;
                active = true;
                if (i0 < 0 || i0 >= imax) {
                    // m/T - mu/T outside the table: the reference exits
                    // (src/MomentumSamplerBase.cpp:33-43); leave a hole
                    n_range++;
                    active = false;
                    Hadron h{};
                    static_cast<Hadron*>(a.out)[a.perm(static_cast<uint64_t>(slot))] = h;
                } else {
                    const double* e0 = pt.C + (static_cast<int64_t>(tb) * kPTabLen + i0) * 4;
                    const double* em = pt.C + (static_cast<int64_t>(tb) * kPTabLen + imax) * 4;
                    rmin0 = e0[2] + 2. * mut * e0[1] + aa * e0[0] - mterm;
                    cmax = em[2] + 2. * mut * em[1] + aa * em[0] - mterm;
                }
            }
            const int ntook = __popcll(took);
            next += ntook;
            if (ntook) hint = __shfl(c, 63 - __clzll(took), kWave);
        }
        if (!__any(active)) {
            if (next >= end) break;
            continue;
        }
        if (!active) continue;
        // ---- one iteration for this lane's hadron
        const u32x4 w = philox4x32(u32x4{static_cast<uint32_t>(c), jj, tw, static_cast<uint32_t>(it)}, key0, key1);
        it++;
        n_inner++;
        const double* C = pt.C + static_cast<int64_t>(tb) * kPTabLen * 4;
        const double r = u32d(w.x) * cmax;
        // the bracket comb(lo) <= r < comb(hi = lo + 1) is unique, so any
        // search reproduces the reference's bisection (inverse_CDF); a 4-ary
        // search issues its three loads together and halves the latency chain
        int lo = i0, hi = imax;
        double rlo = rmin0, rhi = cmax;
        auto comb = [&](int q) {
            const double* e = C + 4 * q;
            return e[2] + 2. * mut * e[1] + aa * e[0] - mterm;
        };
        while (hi - lo > 3) {
            const int q = (hi - lo) >> 2;
            const int m1 = lo + q, m2 = lo + 2 * q, m3 = lo + 3 * q;
            const double v1 = comb(m1), v2 = comb(m2), v3 = comb(m3);
            if (r < v2) {
                if (r < v1) { hi = m1; rhi = v1; } else { lo = m1; rlo = v1; hi = m2; rhi = v2; }
            } else {
                if (r < v3) { lo = m2; rlo = v2; hi = m3; rhi = v3; } else { lo = m3; rlo = v3; }
            }
        }
        while (hi - lo > 1) {
            const int mid = (hi + lo) >> 1;
            const double rm = comb(mid);
            if (r < rm) { hi = mid; rhi = rm; } else { lo = mid; rlo = rm; }
        }
        const double Et0 = pt.Et0[tb], dEt = pt.dEt[tb];
        double E0 = Et0 + lo * dEt;
        if (E0 < x) { E0 = x; rlo = 0.; }
        const double Ehi = Et0 + hi * dEt;
        const double Etl = E0 + (Ehi - E0) / fmax(1e-16, rhi - rlo) * (r - rlo);
        const float Ep = static_cast<float>(Td * Etl + mu);
        const float pm = sqrtf(fmaxf(Ep * Ep - fm * fm, 0.f));
        const float ratio = (pm / Ep) / (1.f - fm * fm / (2.f * Ep * Ep));
        bool done = false;
        if (u24(w.y) <= ratio) {
            n_trial++;
            const float phi = 2.f * static_cast<float>(kPi) * u24(w.z);
            const float cth = 2.f * u24(w.w) - 1.f;
            const float sth = sqrtf(fmaxf(1.f - cth * cth, 0.f));
            float sphi, cphi;
            __sincosf(phi, &sphi, &cphi);
            const float pTc = pm * sth;
            const float px = pTc * cphi, py = pTc * sphi, pz = pm * cth;
            const float p0 = sqrtf(fm * fm + pm * pm);
            const float pds = p0 * d0 + px * d1 + py * d2 + pz * d3;
            const float f0 = 1.f / (__expf((p0 - fmu) / T) + fsg);
            const float omf = 1.f - fsg * f0;
            float dfs = 0.f, dfb = 0.f;
            if (a.shear) {
                const float W = px * px * pixx + 2.f * px * py * pixy + 2.f * px * pz * pixz +
                                py * py * piyy + 2.f * py * pz * piyz + pz * pz * (-pixx - piyy);
                if (a.neos == 1) dfs = omf * W / (2.f * cf2) / (p0 * T);
                else if (a.neos == 0) dfs = omf * W * cf0;
                else dfs = omf * W * dfpref;
            }
            if (a.bulk == BULK_RTA || a.bulk == BULK_CE) {
                const float Eo = p0 / T, mo = fm / T;
                dfb = -omf * cf0 * (mo * mo / (3.f * Eo) - cf1 * Eo) * Pi;
            } else if (a.bulk == BULK_14MOM) {
                dfb = omf * Pi * (cf0 * fm * fm + cf1 * Bq * p0 + cf2 * p0 * p0);
            } else if (a.bulk == BULK_22MOM) {
                dfb = omf * Pi * (fm * fm * cf2 + p0 * (Bq * cf3 + Sq * cf4 + Qq * cf5) + p0 * p0 * (cf1 - cf2));
            }
            const float w1 = fminf(fmaxf(pds / p0 / dsfac, 0.f), 1.f);
            const float w2 = fminf(fmaxf((1.f + dfs + dfb) * 0.5f, 0.f), 1.f);
            const uint32_t u4 = (w.y & 0xFFu) | ((w.z & 0xFFu) << 8) | ((w.w & 0xFFu) << 16);
            const float ua = (static_cast<float>(u4) + 0.5f) * (1.0f / 16777216.0f);
            if (ua < w1 * w2) {
                // boost to the lab frame (src/FSSW.cpp:2001-2010)
                const float u0 = a.cells[F_U0 * Nc + c], u1 = a.cells[F_U1 * Nc + c],
                            u2 = a.cells[F_U2 * Nc + c], u3 = a.cells[F_U3 * Nc + c];
                const float pdu = px * u1 + py * u2 + pz * u3;
                const float k1 = pdu / (u0 + 1.f) + p0;
                const float lx = px + k1 * u1, ly = py + k1 * u2, lz = pz + k1 * u3;
                const float mT = sqrtf(lx * lx + ly * ly + fm * fm);
                const float tau = a.cells[F_TAU * Nc + c], eta = a.cells[F_ETA * Nc + c];
                float eta_s, E, pzl;
                if (a.boost_invariant) {
                    // y uniform in [y_LB, y_RB], eta_s = y - (y - eta_s)
                    // (src/FSSW.cpp:1024-1029)
                    Stream st(a.seed, static_cast<uint32_t>(c), jj, tag_word(TAG_SETUP, a.chunk) ^ 0x00800000u);
                    const float rap = a.yLB + (a.yRB - a.yLB) * static_cast<float>(st.uniform());
                    eta_s = rap - (asinhf(lz / mT) - eta);
                    E = mT * coshf(rap);
                    pzl = mT * sinhf(rap);
                } else {
                    eta_s = eta;
                    E = sqrtf(mT * mT + lz * lz);
                    pzl = lz;
                }
                Hadron h;
                h.pid = a.sp.pid[s];
                h.mass = fm;
                h.E = E; h.px = lx; h.py = ly; h.pz = pzl;
                h.t = tau * coshf(eta_s);
                h.x = a.cells[F_X * Nc + c];
                h.y = a.cells[F_Y * Nc + c];
                h.z = tau * sinhf(eta_s);
                static_cast<Hadron*>(a.out)[a.perm(static_cast<uint64_t>(slot))] = h;
                done = true;
            } else {
                trials++;
            }
        }
        if (!done && (trials >= a.max_trials || it >= max_it)) {
            n_fail++;
            Hadron h{};
            static_cast<Hadron*>(a.out)[a.perm(static_cast<uint64_t>(slot))] = h;
            done = true;
        }
        if (done) active = false;
    }
    // statistics: wave sums, then one atomic per workgroup
    n_trial = wave_sum(n_trial);
    n_inner = wave_sum(n_inner);
    n_fail = wave_sum(n_fail);
    n_range = wave_sum(n_range);
    if (lane == 0) {
        atomicAdd(&s_stats[0], n_trial);
        atomicAdd(&s_stats[1], n_inner);
        atomicAdd(&s_stats[2], n_fail);
        atomicAdd(&s_stats[3], n_range);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_stats[threadIdx.x]) atomicAdd(&a.stats[threadIdx.x], s_stats[threadIdx.x]);
}

// =================================================================== host

namespace {

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
            return ISS_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

}  // namespace

struct iss_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    std::string err;
    int num_cu = 256;
    // species
    int Ns = 0;
    std::vector<int> h_pid;
    DBuf<double> d_mass;
    DBuf<int> d_ints;  // g, B, S, Q, sign, pid
    // delta f
    iss_deltaf_desc df{};
    int bulk_mode = BULK_NONE, neos = -1;
    DBuf<double> d_k11, d_neos;
    double neos_de = 0.;
    bool tables_ready = false;
    DBuf<double> d_ktab;
    DBuf<float2> d_ktabf;
    int ktab_bulk = -1;
    DBuf<double> d_ptab;
    PTabDev pt{};
    // surface
    int64_t Nc = 0;
    DBuf<float> d_cells, d_tmp;
    DBuf<double> d_coef;
    DBuf<uint32_t> d_u32;
    DBuf<int64_t> d_i64, d_scan;
    DBuf<double> d_eos;
    // yields
    bool yields_valid = false;
    int yield_precision = 0;
    DBuf<double> d_cdf, d_ctot, d_partial, d_stot;
    DBuf<int> d_oob;
    DBuf<char> d_raw;
    // sampling
    DBuf<uint32_t> d_counts, d_evcount;
    DBuf<int64_t> d_off, d_evoff, d_tree0, d_tree1;
    DBuf<Hadron> d_out;
    DBuf<unsigned long long> d_stats;
    int64_t K = 0;
    int E = 0;
    iss_stats stats{};
    int yield_grid = 1;
};

namespace {

int set_err(iss_ctx* ctx, int code, const std::string& m) {
    ctx->err = m;
    return code;
}

int launch_scan(iss_ctx* ctx, const uint32_t* in, int64_t n, int64_t* out) {
    const int64_t ntiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    HIPCHK(ctx->d_scan.ensure(ntiles + 1));
    k_scan_tiles<<<ntiles, kScanBlock, 0, ctx->stream>>>(in, n, ctx->d_scan.p);
    k_scan_top<<<1, kScanBlock, 0, ctx->stream>>>(ctx->d_scan.p, ntiles);
    k_scan_apply<<<ntiles, kScanBlock, 0, ctx->stream>>>(in, n, ctx->d_scan.p, out);
    HIPCHK(hipGetLastError());
    return ISS_OK;
}

// |p|-sampler tables (src/BosonMomentumSampler.cpp:7-31,
// src/FermionMomentumSampler.cpp:8-31, src/MomentumSamplerShell.cpp:6-21),
// built on the host in FP64 with the same closed forms as the device pcdf.
double host_pcdf(int fermion, int quirk, int trunc, double m, int k, double E) {
    if (k == 0 && trunc > 5) {
        if (fermion) return -std::exp(m) * std::log((1. + std::exp(-E)) / (1. + std::exp(-m)));
        return std::exp(m) * std::log((1. - std::exp(-E)) / (1. - std::exp(-m)));
    }
    double res = 0., sign = 1.;
    for (int n = 0; n < trunc; n++) {
        const int n1 = n + 1;
        const double em = std::exp(-m * n), ee = std::exp((m - E) * n1);
        if (k == 0) {
            const double c0 = fermion ? static_cast<double>(static_cast<int>(sign) / n1) : 1. / n1;
            res += c0 * em * (1. - ee);
        } else if (k == 1) {
            const double s = fermion ? sign : 1.;
            res += s / (n1 * n1) * em * ((m * n1 + 1) - ee * (E * n1 + 1));
        } else {
            const double s = (fermion && !quirk) ? sign : 1.;
            res += s / (static_cast<double>(n1) * n1 * n1) * em *
                   ((m * n1 * (m * n1 + 2) + 2) - ee * (E * n1 * (E * n1 + 2) + 2));
        }
        sign = -sign;
    }
    return res;
}

int build_ptables(iss_ctx* ctx, int quirk) {
    const double m0s[6] = {0.05, 30., 50., 0., 30., 50.};
    const int tr[6] = {10, 2, 1, 10, 2, 1};
    std::vector<double> h(6ull * kPTabLen * 4, 0.);
    PTabDev& pt = ctx->pt;
    pt.quirk = quirk;
    for (int t = 0; t < 6; t++) {
        const int fer = t >= 3;
        const double Emin = fer ? m0s[t] : m0s[t] + 0.05;
        const double Emax = Emin + (fer ? 40. : 50.);
        const double dE = fer ? 0.02 : 0.05;
        const int np = static_cast<int>((Emax - Emin) / dE + 1);
        pt.Et0[t] = Emin;
        pt.dEt[t] = dE;
        pt.m0[t] = m0s[t];
        pt.len[t] = np;
        pt.trunc[t] = tr[t];
        pt.fermion[t] = fer;
        pt.ep[t] = std::exp(m0s[t]);
        pt.em1[t] = std::exp(-m0s[t]);
        for (int i = 0; i < np; i++) {
            const double e = Emin + i * dE;
            double* r = &h[(static_cast<size_t>(t) * kPTabLen + i) * 4];
            r[0] = host_pcdf(fer, quirk, tr[t], m0s[t], 0, e);
            r[1] = host_pcdf(fer, quirk, tr[t], m0s[t], 1, e);
            r[2] = host_pcdf(fer, quirk, tr[t], m0s[t], 2, e);
        }
    }
    HIPCHK(ctx->d_ptab.ensure(h.size()));
    HIPCHK(hipMemcpyAsync(ctx->d_ptab.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    pt.C = ctx->d_ptab.p;
    return ISS_OK;
}

int ensure_ktab(iss_ctx* ctx) {
    const int with_bulk = 1;  // K1 and K3 are cheap to keep
    if (ctx->ktab_bulk == with_bulk) return ISS_OK;
    HIPCHK(ctx->d_ktab.ensure(3 * kKLen));
    HIPCHK(ctx->d_ktabf.ensure(3 * kKLen));
    k_bessel_table<<<(kKLen + 127) / 128, 128, 0, ctx->stream>>>(ctx->d_ktab.p, ctx->d_ktabf.p, with_bulk);
    k_bessel_table_f32<<<(kKLen + 127) / 128, 128, 0, ctx->stream>>>(ctx->d_ktab.p, ctx->d_ktabf.p);
    HIPCHK(hipGetLastError());
    ctx->ktab_bulk = with_bulk;
    return ISS_OK;
}

DeltafDev deltaf_dev(iss_ctx* ctx) {
    DeltafDev d{};
    d.bulk_kind = ctx->df.bulk_kind;
    d.include_bulk = ctx->df.include_bulk;
    d.neos = ctx->neos;
    d.k11_nT = ctx->df.k11_nT;
    d.k11_nmu = ctx->df.k11_nmu;
    d.k11_T0 = ctx->df.k11_T0;
    d.k11_mu0 = ctx->df.k11_mu0;
    d.k11_dT = ctx->df.k11_dT;
    d.k11_dmu = ctx->df.k11_dmu;
    d.k11 = ctx->d_k11.p;
    d.neos_ncol = ctx->df.neos_ncol;
    d.neos_tab = ctx->d_neos.p;
    d.neos_de = ctx->neos_de;
    return d;
}

int cell_prep(iss_ctx* ctx) {
    if (ctx->Nc == 0) return ISS_OK;
    HIPCHK(ctx->d_coef.ensure(6 * ctx->Nc));
    const int64_t nb = (ctx->Nc + 255) / 256;
    k_cell_prep<<<nb, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->Nc, deltaf_dev(ctx), ctx->d_coef.p);
    HIPCHK(hipGetLastError());
    return ISS_OK;
}

SpeciesDev species_dev(iss_ctx* ctx) {
    SpeciesDev s;
    const int Ns = ctx->Ns;
    s.mass = ctx->d_mass.p;
    s.g = ctx->d_ints.p;
    s.B = ctx->d_ints.p + Ns;
    s.S = ctx->d_ints.p + 2 * Ns;
    s.Q = ctx->d_ints.p + 3 * Ns;
    s.sign = ctx->d_ints.p + 4 * Ns;
    s.pid = ctx->d_ints.p + 5 * Ns;
    return s;
}

template <typename R, int MODE, bool RAW, bool EXACT>
int launch_yield_t(iss_ctx* ctx, const YieldArgs& a) {
    constexpr int TC = 16;
    const size_t lds = yield_lds_bytes(sizeof(R), ktables_for(MODE), TC, pad_species(a.Ns));
    auto kern = k_yield<R, MODE, RAW, TC, EXACT>;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
        return ISS_ERR_HIP;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kYieldThreads, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int64_t ntiles = (a.Nc + TC - 1) / TC;
    const int nblk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ntiles, int64_t(ctx->num_cu) * per_cu)));
    ctx->yield_grid = nblk;
    kern<<<nblk, kYieldThreads, lds, ctx->stream>>>(a);
    return hipGetLastError() == hipSuccess ? ISS_OK : ISS_ERR_HIP;
}

template <typename R, int MODE>
int launch_yield_m(iss_ctx* ctx, const YieldArgs& a, bool raw, bool exact) {
    if (exact) return raw ? launch_yield_t<R, MODE, true, true>(ctx, a) : launch_yield_t<R, MODE, false, true>(ctx, a);
    return raw ? launch_yield_t<R, MODE, true, false>(ctx, a) : launch_yield_t<R, MODE, false, false>(ctx, a);
}

template <typename R>
int launch_yield(iss_ctx* ctx, const YieldArgs& a, bool raw, bool exact) {
    switch (a.bulk) {
        case BULK_RTA: return launch_yield_m<R, BULK_RTA>(ctx, a, raw, exact);
        case BULK_CE: return launch_yield_m<R, BULK_CE>(ctx, a, raw, exact);
        case BULK_14MOM: return launch_yield_m<R, BULK_14MOM>(ctx, a, raw, exact);
        case BULK_22MOM: return launch_yield_m<R, BULK_22MOM>(ctx, a, raw, exact);
        default: return launch_yield_m<R, BULK_NONE>(ctx, a, raw, exact);
    }
}

int run_yields(iss_ctx* ctx, int precision, void* per_cell_host, double* totals_host) {
    if (ctx->Ns == 0) return set_err(ctx, ISS_ERR_STATE, "species not set");
    if (precision != 32 && precision != 64) return set_err(ctx, ISS_ERR_ARG, "precision must be 32 or 64");
    int rc = ensure_ktab(ctx);
    if (rc) return rc;
    const int64_t Nc = ctx->Nc;
    const int Ns = ctx->Ns;
    HIPCHK(ctx->d_cdf.ensure(std::max<int64_t>(1, Nc) * Ns));
    HIPCHK(ctx->d_ctot.ensure(std::max<int64_t>(1, Nc)));
    HIPCHK(ctx->d_stot.ensure(Ns));
    // partials sized for the largest grid the occupancy can give
    const int max_blk = ctx->num_cu * 8;
    HIPCHK(ctx->d_partial.ensure(static_cast<size_t>(max_blk) * Ns));
    const size_t esz = precision == 64 ? sizeof(double) : sizeof(float);
    const bool raw = per_cell_host != nullptr;
    if (raw) HIPCHK(ctx->d_raw.ensure(std::max<int64_t>(1, Nc) * Ns * esz));
    YieldArgs a;
    a.cells = ctx->d_cells.p;
    a.coef = ctx->d_coef.p;
    a.Nc = Nc;
    a.Ns = Ns;
    a.sp = species_dev(ctx);
    a.ktab = ctx->d_ktab.p;
    a.ktabf = ctx->d_ktabf.p;
    a.bulk = ctx->df.include_bulk ? ctx->bulk_mode : BULK_NONE;
    a.cdf = ctx->d_cdf.p;
    a.ctot = ctx->d_ctot.p;
    a.raw = raw ? static_cast<void*>(ctx->d_raw.p) : nullptr;
    a.partial = ctx->d_partial.p;
    HIPCHK(ctx->d_oob.ensure(1));
    a.oob = ctx->d_oob.p;
    HIPCHK(hipMemsetAsync(ctx->d_oob.p, 0, sizeof(int), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    int nblk = 1;
    if (Nc > 0) {
        int rc = (precision == 64) ? launch_yield<double>(ctx, a, raw, false) : launch_yield<float>(ctx, a, raw, false);
        if (rc) return set_err(ctx, rc, "yield kernel launch failed");
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        int oob = 0;
        HIPCHK(hipMemcpyAsync(&oob, ctx->d_oob.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (oob) {
            // a Bessel argument left the table: rerun with the exact fallback
            rc = (precision == 64) ? launch_yield<double>(ctx, a, raw, true) : launch_yield<float>(ctx, a, raw, true);
            if (rc) return set_err(ctx, rc, "yield kernel launch failed");
            HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        }
        nblk = ctx->yield_grid;
    } else {
        HIPCHK(hipMemsetAsync(ctx->d_partial.p, 0, sizeof(double) * Ns, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    }
    k_species_reduce<<<Ns, 256, 0, ctx->stream>>>(ctx->d_partial.p, nblk, Ns, ctx->d_stot.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    if (totals_host)
        HIPCHK(hipMemcpyAsync(totals_host, ctx->d_stot.p, sizeof(double) * Ns, hipMemcpyDeviceToHost, ctx->stream));
    if (raw && Nc > 0)
        HIPCHK(hipMemcpyAsync(per_cell_host, ctx->d_raw.p, esz * Nc * Ns, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms1 = 0.f, ms2 = 0.f;
    HIPCHK(hipEventElapsedTime(&ms1, ctx->ev[0], ctx->ev[1]));
    HIPCHK(hipEventElapsedTime(&ms2, ctx->ev[0], ctx->ev[2]));
    ctx->stats.ms_yield_kernel = ms1;
    ctx->stats.ms_yield = ms2;
    ctx->yields_valid = true;
    ctx->yield_precision = precision;
    return ISS_OK;
}

}  // namespace

// ================================================================ C-ABI

extern "C" {

int iss_ctx_create(int device, iss_ctx** out) {
    if (!out) return ISS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ISS_ERR_HIP;
    if (device < 0 || device >= ndev) return ISS_ERR_ARG;
    iss_ctx* ctx = new iss_ctx;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
        delete ctx;
        return ISS_ERR_HIP;
    }
    for (auto& e : ctx->ev) (void)hipEventCreate(&e);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
    if (ctx->d_stats.ensure(4) != hipSuccess) {
        delete ctx;
        return ISS_ERR_NOMEM;
    }
    *out = ctx;
    return ISS_OK;
}

int iss_ctx_destroy(iss_ctx* ctx) {
    if (!ctx) return ISS_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->d_mass.release(); ctx->d_ints.release(); ctx->d_k11.release(); ctx->d_neos.release();
    ctx->d_ktab.release(); ctx->d_ktabf.release(); ctx->d_ptab.release();
    ctx->d_cells.release(); ctx->d_tmp.release(); ctx->d_coef.release(); ctx->d_u32.release();
    ctx->d_i64.release(); ctx->d_scan.release(); ctx->d_eos.release(); ctx->d_cdf.release();
    ctx->d_ctot.release(); ctx->d_partial.release(); ctx->d_oob.release(); ctx->d_stot.release(); ctx->d_raw.release();
    ctx->d_counts.release(); ctx->d_evcount.release(); ctx->d_off.release(); ctx->d_evoff.release();
    ctx->d_tree0.release(); ctx->d_tree1.release(); ctx->d_out.release(); ctx->d_stats.release();
    for (auto& e : ctx->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ISS_OK;
}

const char* iss_last_error(const iss_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int iss_set_species(iss_ctx* ctx, const iss_species_desc* sp) {
    if (!ctx || !sp) return ISS_ERR_ARG;
    if (sp->n <= 0 || sp->n > kMaxSpecies)
        return set_err(ctx, ISS_ERR_RANGE, "number of species must be in [1, 512]");
    (void)hipSetDevice(ctx->device);
    const int Ns = sp->n;
    std::vector<int> ints(6 * Ns);
    for (int i = 0; i < Ns; i++) {
        if (sp->sign[i] < -1 || sp->sign[i] > 1) return set_err(ctx, ISS_ERR_ARG, "sign must be -1, 0 or 1");
        if (!(sp->mass[i] > 0.)) return set_err(ctx, ISS_ERR_ARG, "species mass must be positive");
        ints[i] = sp->degeneracy[i];
        ints[Ns + i] = sp->baryon[i];
        ints[2 * Ns + i] = sp->strange[i];
        ints[3 * Ns + i] = sp->charge[i];
        ints[4 * Ns + i] = sp->sign[i];
        ints[5 * Ns + i] = sp->pid[i];
    }
    HIPCHK(ctx->d_mass.ensure(Ns));
    HIPCHK(ctx->d_ints.ensure(6 * Ns));
    HIPCHK(hipMemcpyAsync(ctx->d_mass.p, sp->mass, sizeof(double) * Ns, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_ints.p, ints.data(), sizeof(int) * 6 * Ns, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->Ns = Ns;
    ctx->h_pid.assign(sp->pid, sp->pid + Ns);
    ctx->yields_valid = false;
    if (!ctx->tables_ready) {
        iss_deltaf_desc d{};
        d.fermion_cdf2_quirk = 1;
        int rc = iss_set_deltaf(ctx, &d);
        if (rc) return rc;
    }
    return ISS_OK;
}

int iss_set_deltaf(iss_ctx* ctx, const iss_deltaf_desc* d) {
    if (!ctx || !d) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (d->include_diffusion) return set_err(ctx, ISS_ERR_UNSUPPORTED, "baryon-diffusion delta f is not supported");
    int mode = BULK_NONE, neos = -1;
    const int k = d->bulk_kind;
    if (k == 21) neos = 1;
    if (k == 20) neos = 0;
    if (d->include_bulk) {
        if (k == 1) mode = BULK_RTA;
        else if (k == 11) mode = BULK_14MOM;
        else if (k == 20) mode = BULK_22MOM;
        else if (k == 21) mode = BULK_CE;
        else if (k == 0 || (k >= 2 && k <= 4)) mode = BULK_NONE;
        else return set_err(ctx, ISS_ERR_UNSUPPORTED, "unknown bulk_deltaf_kind");
    }
    if (d->include_bulk && k == 11) {
        if (!d->k11_c0 || !d->k11_c1 || !d->k11_c2 || d->k11_nT < 2 || d->k11_nmu < 2)
            return set_err(ctx, ISS_ERR_ARG, "kind 11 tables missing");
        const size_t len = static_cast<size_t>(d->k11_nT) * d->k11_nmu;
        HIPCHK(ctx->d_k11.ensure(3 * len));
        HIPCHK(hipMemcpyAsync(ctx->d_k11.p, d->k11_c0, len * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_k11.p + len, d->k11_c1, len * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->d_k11.p + 2 * len, d->k11_c2, len * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    if (neos >= 0) {
        const int nc = (neos == 1) ? 5 : 8;
        if (!d->neos || d->neos_ncol != nc) return set_err(ctx, ISS_ERR_ARG, "NEoS coefficient table missing");
        const size_t len = 200ull * 200ull * nc;
        HIPCHK(ctx->d_neos.ensure(len));
        HIPCHK(hipMemcpyAsync(ctx->d_neos.p, d->neos, len * 8, hipMemcpyHostToDevice, ctx->stream));
        ctx->neos_de = d->neos[200 * nc] - d->neos[0];
    }
    ctx->df = *d;
    ctx->df.k11_c0 = ctx->df.k11_c1 = ctx->df.k11_c2 = nullptr;
    ctx->df.neos = nullptr;
    ctx->bulk_mode = mode;
    ctx->neos = neos;
    int rc = build_ptables(ctx, d->fermion_cdf2_quirk ? 1 : 0);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->tables_ready = true;
    ctx->yields_valid = false;
    rc = cell_prep(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_set_surface_lab(iss_ctx* ctx, int64_t n, const float* cells34, const iss_eos_desc* eos) {
    if (!ctx || n < 0 || (n > 0 && !cells34)) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    int use_eos = 0;
    EosDev ed{0, 0, nullptr};
    if (eos && eos->rows && eos->n_rows > 0) {
        const int i = eos->iEOS;
        if (!(i == 9 || i == 91 || i == 12 || i == 14))
            return set_err(ctx, ISS_ERR_ARG, "EoS regulation needs iEOS 9, 91, 12 or 14");
        const int nBlen = (i == 12 || i == 14) ? 200 : 1;
        if (eos->n_rows < 3 * nBlen) return set_err(ctx, ISS_ERR_ARG, "EoS table too short");
        HIPCHK(ctx->d_eos.ensure(7 * eos->n_rows));
        HIPCHK(hipMemcpyAsync(ctx->d_eos.p, eos->rows, sizeof(double) * 7 * eos->n_rows,
                              hipMemcpyHostToDevice, ctx->stream));
        ed = EosDev{i, eos->n_rows, ctx->d_eos.p};
        use_eos = 1;
    }
    ctx->yields_valid = false;
    ctx->Nc = 0;
    if (n == 0) return ISS_OK;
    DBuf<float> lab;
    HIPCHK(lab.ensure(34 * n));
    HIPCHK(hipMemcpyAsync(lab.p, cells34, sizeof(float) * 34 * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx->d_tmp.ensure(kCellFields * n));
    HIPCHK(ctx->d_u32.ensure(n));
    HIPCHK(ctx->d_i64.ensure(n + 1));
    const int64_t nb = (n + 255) / 256;
    k_lab_to_lrf<<<nb, 256, 0, ctx->stream>>>(lab.p, n, ed, use_eos, ctx->d_tmp.p, ctx->d_u32.p);
    HIPCHK(hipGetLastError());
    int rc = launch_scan(ctx, ctx->d_u32.p, n, ctx->d_i64.p);
    if (rc) { lab.release(); return rc; }
    int64_t nk = 0;
    HIPCHK(hipMemcpyAsync(&nk, ctx->d_i64.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx->d_cells.ensure(kCellFields * std::max<int64_t>(nk, 1)));
    k_compact<<<nb, 256, 0, ctx->stream>>>(ctx->d_tmp.p, ctx->d_u32.p, ctx->d_i64.p, n, nk, ctx->d_cells.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    lab.release();
    ctx->Nc = nk;
    rc = cell_prep(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_set_surface_lrf(iss_ctx* ctx, int64_t n, const float* cells28) {
    if (!ctx || n < 0 || (n > 0 && !cells28)) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    ctx->yields_valid = false;
    ctx->Nc = 0;
    if (n == 0) return ISS_OK;
    HIPCHK(ctx->d_tmp.ensure(kCellFields * n));
    HIPCHK(ctx->d_cells.ensure(kCellFields * n));
    HIPCHK(hipMemcpyAsync(ctx->d_tmp.p, cells28, sizeof(float) * 28 * n, hipMemcpyHostToDevice, ctx->stream));
    k_aos_to_soa<<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_tmp.p, n, ctx->d_cells.p);
    HIPCHK(hipGetLastError());
    ctx->Nc = n;
    int rc = cell_prep(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int64_t iss_num_cells(const iss_ctx* ctx) { return ctx ? ctx->Nc : -1; }

int iss_get_surface_lrf(iss_ctx* ctx, float* cells28) {
    if (!ctx || (!cells28 && ctx->Nc > 0)) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->Nc == 0) return ISS_OK;
    HIPCHK(ctx->d_tmp.ensure(kCellFields * ctx->Nc));
    k_soa_to_aos<<<(ctx->Nc + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->Nc, ctx->d_tmp.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cells28, ctx->d_tmp.p, sizeof(float) * 28 * ctx->Nc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_compute_yields(iss_ctx* ctx, const iss_yield_opts* opts, double* species_totals, void* per_cell) {
    if (!ctx || !opts) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    return run_yields(ctx, opts->precision, per_cell, species_totals);
}

int iss_sample(iss_ctx* ctx, const iss_sample_opts* o, int64_t* n_hadrons) {
    if (!ctx || !o) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (o->n_events <= 0) return set_err(ctx, ISS_ERR_ARG, "n_events must be positive");
    if (o->sampling_model != 30)
        return set_err(ctx, ISS_ERR_UNSUPPORTED, "only dN_dy_sampling_model = 30 (Poisson) is supported");
    if (o->proposal != 0) return set_err(ctx, ISS_ERR_UNSUPPORTED, "learned proposal not available");
    if (o->max_trials <= 1) return set_err(ctx, ISS_ERR_ARG, "max_trials must exceed 1");
    if (!ctx->yields_valid) {
        int rc = run_yields(ctx, o->precision == 64 ? 64 : 32, nullptr, nullptr);
        if (rc) return rc;
    }
    const int64_t Nc = ctx->Nc;
    const int E = o->n_events;
    const bool binv = o->hydro_mode != 2;
    const double mult = static_cast<double>(E) * (binv ? (o->y_RB - o->y_LB) : 1.0);
    const uint32_t chunk = static_cast<uint32_t>(o->event0);
    HIPCHK(ctx->d_counts.ensure(std::max<int64_t>(Nc, 1)));
    HIPCHK(ctx->d_off.ensure(Nc + 1));
    HIPCHK(ctx->d_evcount.ensure(E));
    HIPCHK(ctx->d_evoff.ensure(E + 1));
    int D = 0;
    while ((1ll << D) < E) D++;
    HIPCHK(ctx->d_tree0.ensure(1ll << D));
    HIPCHK(ctx->d_tree1.ensure(1ll << D));
    HIPCHK(hipMemsetAsync(ctx->d_stats.p, 0, 4 * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if (Nc > 0)
        k_counts<<<(Nc + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_ctot.p, Nc, mult, o->seed, chunk,
                                                            ctx->d_counts.p);
    HIPCHK(hipGetLastError());
    int rc = launch_scan(ctx, ctx->d_counts.p, Nc, ctx->d_off.p);
    if (rc) return rc;
    // multinomial over events, root = K
    HIPCHK(hipMemcpyAsync(ctx->d_tree0.p, ctx->d_off.p + Nc, sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream));
    int64_t* cur = ctx->d_tree0.p;
    int64_t* nxt = ctx->d_tree1.p;
    for (int d = 0; d < D; d++) {
        const int64_t nodes = 1ll << d;
        k_multinomial_level<<<(nodes + 255) / 256, 256, 0, ctx->stream>>>(cur, nxt, d, E, o->seed, chunk);
        std::swap(cur, nxt);
    }
    k_multinomial_leaves<<<((1ll << D) + 255) / 256, 256, 0, ctx->stream>>>(cur, D, E, ctx->d_evcount.p);
    HIPCHK(hipGetLastError());
    rc = launch_scan(ctx, ctx->d_evcount.p, E, ctx->d_evoff.p);
    if (rc) return rc;
    int64_t K = 0;
    HIPCHK(hipMemcpyAsync(&K, ctx->d_off.p + Nc, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx->d_out.ensure(std::max<int64_t>(K, 1)));
    SampleArgs a;
    a.cells = ctx->d_cells.p;
    a.coef = ctx->d_coef.p;
    a.Nc = Nc;
    a.Ns = ctx->Ns;
    a.sp = species_dev(ctx);
    a.pt = ctx->pt;
    a.cdf = ctx->d_cdf.p;
    a.ctot = ctx->d_ctot.p;
    a.off = ctx->d_off.p;
    a.evoff = ctx->d_evoff.p;
    a.E = E;
    a.K = K;
    a.seed = o->seed;
    a.chunk = chunk;
    a.perm.init(static_cast<uint64_t>(std::max<int64_t>(K, 1)), o->seed ^ (0x5851F42D4C957F2Dull * (chunk + 1)));
    a.shear = ctx->df.include_shear;
    a.bulk = ctx->df.include_bulk ? ctx->bulk_mode : BULK_NONE;
    a.neos = ctx->neos;
    a.boost_invariant = binv ? 1 : 0;
    a.yLB = static_cast<float>(o->y_LB);
    a.yRB = static_cast<float>(o->y_RB);
    a.max_trials = o->max_trials;
    a.out = ctx->d_out.p;
    a.stats = ctx->d_stats.p;
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    if (K > 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sample, 256, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        const int64_t want = (K + 255) / 256;
        const int nblk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, int64_t(ctx->num_cu) * per_cu)));
        k_sample<<<nblk, 256, 0, ctx->stream>>>(a);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    unsigned long long st[4];
    HIPCHK(hipMemcpyAsync(st, ctx->d_stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms_all = 0.f, ms_k = 0.f;
    HIPCHK(hipEventElapsedTime(&ms_all, ctx->ev[3], ctx->ev[5]));
    HIPCHK(hipEventElapsedTime(&ms_k, ctx->ev[4], ctx->ev[5]));
    ctx->K = K;
    ctx->E = E;
    ctx->stats.n_hadrons = K;
    ctx->stats.n_trials = static_cast<int64_t>(st[0]);
    ctx->stats.n_inner = static_cast<int64_t>(st[1]);
    ctx->stats.n_failed = static_cast<int64_t>(st[2]);
    ctx->stats.n_range_errors = static_cast<int64_t>(st[3]);
    ctx->stats.ms_sample = ms_all;
    ctx->stats.ms_sample_kernel = ms_k;
    if (n_hadrons) *n_hadrons = K;
    if (st[3]) return set_err(ctx, ISS_ERR_RANGE, "m/T - mu/T outside the momentum-sampler tables");
    return ISS_OK;
}

int iss_get_event_counts(iss_ctx* ctx, int64_t* counts) {
    if (!ctx || !counts) return ISS_ERR_ARG;
    if (ctx->E == 0) return ISS_OK;
    (void)hipSetDevice(ctx->device);
    std::vector<int64_t> off(ctx->E + 1);
    HIPCHK(hipMemcpy(off.data(), ctx->d_evoff.p, sizeof(int64_t) * (ctx->E + 1), hipMemcpyDeviceToHost));
    for (int e = 0; e < ctx->E; e++) counts[e] = off[e + 1] - off[e];
    return ISS_OK;
}

int iss_get_hadrons(iss_ctx* ctx, iss_hadron* dst, int64_t capacity) {
    if (!ctx || (!dst && ctx->K > 0)) return ISS_ERR_ARG;
    if (capacity < ctx->K) return set_err(ctx, ISS_ERR_RANGE, "capacity smaller than the number of hadrons");
    (void)hipSetDevice(ctx->device);
    static_assert(sizeof(Hadron) == sizeof(iss_hadron), "hadron layout");
    if (ctx->K > 0) HIPCHK(hipMemcpy(dst, ctx->d_out.p, sizeof(iss_hadron) * ctx->K, hipMemcpyDeviceToHost));
    return ISS_OK;
}

int iss_get_stats(const iss_ctx* ctx, iss_stats* s) {
    if (!ctx || !s) return ISS_ERR_ARG;
    *s = ctx->stats;
    return ISS_OK;
}

}  // extern "C"
