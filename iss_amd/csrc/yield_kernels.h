// Yield pass: per-(cell, species) yields, FP64 species CDF rows and species totals.
// (one translation unit: included by iss_hip.hip)
#pragma once

#include "kernels.h"

namespace issdev {

// ============================================================== yield pass

template <typename R>
struct KLds {
    using T = R;
};

// LDS slot of K_order: K_2 first (every mode), then K_1, then K_3
__device__ __forceinline__ int kslot(int order) { return order == 2 ? 0 : (order == 1 ? 1 : 2); }

// K_n lookup: the reference's table and linear interpolation
// (src/FSSW.cpp:1646-1703), x = 0.5 + 0.05 i.  The fast variant
// (EXACT = false) reads only the first kKLds rows, staged in LDS (the yield
// pass), or only the global table (staged = false); an argument outside those
// rows raises *oob and the pass is rerun with EXACT = true, which reads the
// global table and calls the exact K_n outside [0.5, 399.95] as the reference
// calls GSL.  Keeping the two paths in different kernels keeps the fast path
// a plain LDS access.
template <bool EXACT, int NROWS = kKLds>
struct KLookup64 {
    const double* lds;   // [3][NROWS] staged rows (used when `staged`)
    const double* glob;  // [3][kKLen]
    int* oob;
    bool staged;
    __device__ __forceinline__ double operator()(int order, double arg) const {
        if (arg < kKx0 || arg > kKxmax - kKdx) {
            if constexpr (EXACT) return kn_exact(order, arg);
            *oob = 1;
            return 0.;
        }
        const int idx = static_cast<int>((arg - kKx0) / kKdx);
        const double fr = (arg - kKx0 - idx * kKdx) / kKdx;
        double k0, k1;
        if (!EXACT && staged) {
            if (idx + 1 >= NROWS) {
                *oob = 1;
                return 0.;
            }
            k0 = lds[kslot(order) * NROWS + idx];
            k1 = lds[kslot(order) * NROWS + idx + 1];
        } else {
            k0 = glob[(order - 1) * kKLen + idx];
            k1 = glob[(order - 1) * kKLen + idx + 1];
        }
        return (1. - fr) * k0 + fr * k1;
    }
    // K_2 and, for NK = 3 or 2, K_1 (and K_3) at one argument
    template <int NK>
    __device__ __forceinline__ void k123(double arg, double& K2, double& K1, double& K3) const {
        K2 = (*this)(2, arg);
        if (NK >= 2) K1 = (*this)(1, arg);
        if (NK >= 3) K3 = (*this)(3, arg);
    }
};

// K_n lookup, FP32: same grid, K_i + f (K_{i+1} - K_i); the fast variant
// reads FP32 rows from LDS, the exact one the global (K_i, K_{i+1} - K_i)
// table
template <bool EXACT>
struct KLookup32 {
    const float* lds;    // [3][kKLds]
    const float2* glob;  // [3][kKLen]
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
        if constexpr (EXACT) {
            const float2 v = glob[(order - 1) * kKLen + idx];
            return fmaf(fr, v.y, v.x);
        } else {
            if (idx + 1 >= kKLds) {
                *oob = 1;
                return 0.f;
            }
            const float k0 = lds[kslot(order) * kKLds + idx], k1 = lds[kslot(order) * kKLds + idx + 1];
            return fmaf(fr, k1 - k0, k0);
        }
    }
    // K_2 and, for NK = 3 or 2, K_1 (and K_3) at one argument: one index,
    // the LDS reads issued together, out-of-range arguments flagged without
    // a branch around the reads.  The table position t = (x - 0.5) / 0.05 is
    // formed from an FP64 argument: an FP32 x near 50 carries an absolute
    // error of a few 1e-6, which K_n(x) ~ exp(-x) would turn into a relative
    // error of the same size, a third of the FP32 budget.
    template <int NK>
    __device__ __forceinline__ void k123(double arg, float& K2, float& K1, float& K3) const {
        if constexpr (EXACT) {
            if (arg < kKx0 || arg > kKxmax - kKdx) {
                K2 = static_cast<float>(kn_exact(2, arg));
                if (NK >= 2) K1 = static_cast<float>(kn_exact(1, arg));
                if (NK >= 3) K3 = static_cast<float>(kn_exact(3, arg));
                return;
            }
            const double t = (arg - kKx0) * 20.0;
            const int idx = static_cast<int>(t);
            const float fr = static_cast<float>(t - idx);
            const float2 v2 = glob[kKLen + idx];
            K2 = fmaf(fr, v2.y, v2.x);
            if (NK >= 2) { const float2 v = glob[idx]; K1 = fmaf(fr, v.y, v.x); }
            if (NK >= 3) { const float2 v = glob[2 * kKLen + idx]; K3 = fmaf(fr, v.y, v.x); }
        } else {
            const double t = (arg - kKx0) * 20.0;
            int idx = static_cast<int>(t);
            const bool ok = (arg >= kKx0) & (arg <= kKxmax - kKdx) & (idx + 1 < kKLds);
            if (!ok) *oob = 1;
            idx = ok ? idx : 0;
            const float fr = static_cast<float>(t - idx);
            const float a2 = lds[idx], b2 = lds[idx + 1];
            float a1 = 0.f, b1 = 0.f, a3 = 0.f, b3 = 0.f;
            if (NK >= 2) { a1 = lds[kKLds + idx]; b1 = lds[kKLds + idx + 1]; }
            if (NK >= 3) { a3 = lds[2 * kKLds + idx]; b3 = lds[2 * kKLds + idx + 1]; }
            K2 = fmaf(fr, b2 - a2, a2);
            if (NK >= 2) K1 = fmaf(fr, b1 - a1, a1);
            if (NK >= 3) K3 = fmaf(fr, b3 - a3, a3);
        }
    }
};

// calculate_dN_analytic + the per-cell combination and clamp
// (src/FSSW.cpp:645-706, 719-848) for one (cell, species).  `ill` reports
// that the terms of N cancel (see kIllCond): rounding errors are amplified.

// exp(x) and 1/n: correctly rounded in FP64 (parity with the reference);
// the hardware exp2 / reciprocal in FP32 (within the 1e-5 FP32 budget).
__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ double frcp(double x) { return 1. / x; }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// E_2, ..., E_18 at x: the table's linear interpolation, exact outside it
// (src/FSSW.cpp:1700-1721)
__device__ __forceinline__ void en_lookup(const double* etab, double x, double* e) {
    if (x < kKx0 || x > kKxmax - kKdx) {
        e[0] = expint_en(2, x);
        for (int k = 1; k < kEnTerms; k++) e[k] = expint_en(2 * k + 2, x);
        return;
    }
    const int idx = static_cast<int>((x - kKx0) / kKdx);
    const double fr = (x - kKx0 - idx * kKdx) / kKdx;
    for (int k = 0; k < kEnTerms; k++)
        e[k] = (1. - fr) * etab[k * kKLen + idx] + fr * etab[k * kKLen + idx + 1];
}

// the diffusion series term I_1 of calculate_dN_analytic (src/FSSW.cpp:
// 785-809) at x = n m / T, before the factor -(m / T)^3
__device__ __forceinline__ double diffusion_i1(const double* etab, double x) {
    double e[kEnTerms];
    en_lookup(etab, x, e);
    double v = exp(-x) / x * (2. / (x * x) + 2. / x - 0.5) + 3. / 8. * e[0];
    double dfact = 1., fact = 2., p2 = 4.;
    for (int k = 3; k <= kEnTerms + 1; k++) {
        dfact *= (2 * k - 5);
        fact *= k;
        p2 *= 2.;
        v += 3. * dfact / p2 / fact * e[k - 2];
    }
    return v;
}

constexpr float kIllCond = 24.f;

// dnq: the diffusion correction of this (cell, species), computed by
// k_diffusion_dn (0 without diffusion delta f)
// md and betad: the mass and 1/T in FP64, which place the argument n m / T in
// the K table (the same product as the FP64 series forms)
template <typename R, int MODE, typename KL>
__device__ __forceinline__ R yield_one(R mass, R pref, R Bq, R Sq, R Qq, R sgn, R T, R beta,
                                       R muB, R muS, R muQ, R dsu, R Pi, const R* c,
                                       const KL& kl, bool* ill, R dnq, double md, double betad) {
    const R mu = Bq * muB + Sq * muS + Qq * muQ;
    const R lambda = fexp(beta * mu);
    const int nt = (mass < R(0.7) && T > R(0.05)) ? 10 : 1;
    const R mb = mass * beta;
    R r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    R th = 1, fug = lambda;
    auto term = [&](R rn, R inv) {
        const double arg = static_cast<double>(rn) * md * betad;
        R K2 = 0, K1 = 0, K3 = 0;
        kl.template k123<(MODE == BULK_NONE) ? 1 : ((MODE == BULK_RTA || MODE == BULK_CE) ? 2 : 3)>(arg, K2, K1, K3);
        // FP64: theta / n as the reference divides; FP32: times 1/n
        const R thn = (sizeof(R) == sizeof(double)) ? th / rn : th * inv;
        const R k2n = (sizeof(R) == sizeof(double)) ? 3 * K2 / rn : 3 * K2 * inv;
        r0 += thn * fug * K2;
        if (MODE == BULK_RTA || MODE == BULK_CE) {
            r1 += th * fug * (mb * K1 + k2n);
            r2 += th * fug * K1;
        } else if (MODE == BULK_14MOM || MODE == BULK_22MOM) {
            const R k3n = (sizeof(R) == sizeof(double)) ? 3 * K3 / rn : 3 * K3 * inv;
            r1 += th * fug * K2;
            r2 += th * fug * (mb * K1 + k2n);
            r3 += th * fug * (mb * K2 + k3n);
        }
        th *= -sgn;
        fug *= lambda;
    };
    if (nt == 1) {
        // one term (every species above 0.7 GeV): n = 1 folds away
        term(R(1), R(1));
    } else {
#pragma unroll 1
        for (int n = 1; n <= nt; n++) {
            const R rn = static_cast<R>(n);
            term(rn, (sizeof(R) == sizeof(double)) ? R(1) : frcp(rn));
        }
    }
    const R m2 = mass * mass;
    const R pd = pref * dsu;
    const R Neq = pd * (m2 * T * r0);
    R dN = 0, dabs = 0;  // dabs: the sum of the magnitudes of dN's terms
    if (MODE == BULK_RTA || MODE == BULK_CE) {
        const R b1 = m2 * T * r1, b2 = m2 * mass / R(3) * r2;
        dN = pd * (-Pi * c[0]) * (-c[1] * b1 + b2);
        if (ill) dabs = fabs(pd * Pi * c[0]) * (fabs(c[1] * b1) + fabs(b2));
    } else if (MODE == BULK_14MOM) {
        const R T2 = T * T;
        const R b1 = m2 * T * r1, b2 = m2 * T2 * r2, b3 = m2 * mass * T2 * r3;
        const R t1 = b1 * m2 * c[0], t2 = b2 * Bq * c[1], t3 = b3 * c[2];
        dN = pd * Pi * (t1 + t2 + t3);
        if (ill) dabs = fabs(pd * Pi) * (fabs(t1) + fabs(t2) + fabs(t3));
    } else if (MODE == BULK_22MOM) {
        const R T2 = T * T;
        const R b1 = m2 * T * r1, b2 = m2 * T2 * r2, b3 = m2 * mass * T2 * r3;
        const R t1 = b1 * m2 * c[2], t2 = b2 * (Bq * c[3] + Sq * c[4] + Qq * c[5]), t3 = b3 * (c[1] - c[2]);
        dN = pd * Pi * (t1 + t2 + t3);
        if (ill)
            dabs = fabs(pd * Pi) * (fabs(t1) + fabs(b2) * (fabs(Bq * c[3]) + fabs(Sq * c[4]) + fabs(Qq * c[5])) +
                                    fabs(b3) * (fabs(c[1]) + fabs(c[2])));
    }
    dN += dnq;
    const R N = Neq + dN;
    // ill-conditioned in the working precision: the terms of N cancel to a
    // small fraction of their magnitudes.  The relative error of an FP32 N is
    // about 2e-7 times (|N_eq| + sum |terms|) / |N| (14-moment bulk delta f
    // of baryons at mu_B ~ 0.45 GeV reaches a ratio of 70, where the three
    // bulk terms each exceed N_eq tenfold); beyond kIllCond the item is
    // recomputed in FP64 (0.8 % of the items of a 2M-cell surface with
    // mu_B up to 0.45 GeV).
    if (ill) *ill = (fabs(Neq) + dabs + fabs(dnq)) > R(kIllCond) * fabs(N);
    return N > R(0) ? N : R(0);
}

// FP64 yield of one (cell, species) from global memory only: the fallback
// for FP32 yields whose bulk correction nearly cancels N_eq ((|N_eq| + |dN|)
// > 4 |N|), so that every FP32 yield stays within 1e-5 of FP64 (rare).
template <int MODE, bool EXACT>
__device__ double yield_fp64_global(const YieldArgs& a, int64_t c, int s) {
    const int64_t Nc = a.Nc;
    const KLookup64<EXACT> k64{nullptr, a.ktab, a.oob, false};
    const double T = a.cells[F_T * Nc + c];
    double cd[6];
    for (int k = 0; k < 6; k++) cd[k] = a.coef[k * Nc + c];
    double Pi = a.cells[F_PI * Nc + c];
    if (MODE == BULK_RTA) Pi = Pi / kHbarC;
    const double pref = static_cast<double>(a.sp.g[s]) * (1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi));
    const double m = a.sp.mass[s];
    return yield_one<double, MODE>(m, pref, a.sp.B[s], a.sp.S[s], a.sp.Q[s], a.sp.sign[s], T,
                                   1. / T, (double)a.cells[F_MUB * Nc + c], (double)a.cells[F_MUS * Nc + c],
                                   (double)a.cells[F_MUQ * Nc + c], (double)a.cells[F_DA0 * Nc + c], Pi, cd,
                                   k64, (bool*)nullptr, a.dnq ? a.dnq[c * a.Ns + s] : 0., m, 1. / T);
}

constexpr int kYieldThreads = 256;

// an empty use of a register value: the compiler completes its load here
template <typename T>
__device__ __forceinline__ void settle(const T& x) {
    asm volatile("" ::"v"(x));
}
// species per CDF staging chunk
// species per CDF staging chunk: one 128-byte line of a CDF row in FP32 mode;
// half a line in FP64 mode, whose wider staging would halve the occupancy
__host__ __device__ constexpr int yield_chunk(int esz) { return esz == 8 ? 8 : 16; }

__host__ __device__ constexpr int ktables_for(int mode) {
    return mode == BULK_NONE ? 1 : ((mode == BULK_RTA || mode == BULK_CE) ? 2 : 3);
}

__host__ __device__ constexpr int pad_species(int ns) { return (ns + 3) & ~3; }

// LDS bytes of one k_yield workgroup for nsp (padded) species
__host__ __device__ constexpr size_t yield_lds_bytes(int esz, int ntab, int nsp) {
    return static_cast<size_t>(ntab) * kKLds * esz                                // K rows
           + static_cast<size_t>(nsp) * 8                                         // species sums
           + static_cast<size_t>(kYieldThreads) * (yield_chunk(esz) + 1) * (8 + esz);  // prefixes, yields
}

// Yield pass, one thread per cell (persistent over tiles of 256 cells).  The
// species loop is uniform across the wave: species constants come from scalar
// loads, the light species' 10-term series keeps whole waves together, and the
// cell's parameters stay in registers.  Per species:
//   the yield (FP32, or FP64 for precision 64), with ill-conditioned FP32
//   items ((|N_eq| + |dN|) > 4 |N|) recomputed in FP64;
//   the cell's running FP64 prefix, staged in LDS and stored coalesced as the
//   species CDF row, yield_chunk species at a time;
//   the yield, staged beside it: per chunk, a group of threads per species sum the
//   tile's column in FP64 (fixed order) into the workgroup's species sums.
template <typename R, int MODE, bool RAW, bool EXACT>
__global__ __launch_bounds__(kYieldThreads) void k_yield(YieldArgs a, const SpeciesK* __restrict__ spk) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = ktables_for(MODE);
    constexpr int SC = yield_chunk(sizeof(R));
    constexpr int G = kYieldThreads / SC;  // threads per species in the column sums
    using KT = typename KLds<R>::T;
    const int Ns = a.Ns;
    const int nsp = pad_species(Ns);
    static_assert(G * SC == kYieldThreads && G <= kWave, "one lane group per species of a chunk");
    KT* s_k = reinterpret_cast<KT*>(smem);
    double* s_acc = reinterpret_cast<double*>(s_k + NT * kKLds);
    double* s_stage = s_acc + nsp;
    R* s_y = reinterpret_cast<R*>(s_stage + kYieldThreads * (SC + 1));

    // K_2 first, then K_1, K_3 as the mode needs
    for (int i = threadIdx.x; i < NT * kKLds; i += blockDim.x) {
        const int t = i / kKLds, r = i % kKLds;
        const int o = (t == 0) ? 2 : (t == 1 ? 1 : 3);
        s_k[i] = static_cast<KT>(a.ktab[(o - 1) * kKLen + r]);
    }
    for (int i = threadIdx.x; i < nsp; i += blockDim.x) s_acc[i] = 0.;
    __syncthreads();

    const int64_t Nc = a.Nc;
    double* stage = s_stage + threadIdx.x * (SC + 1);
    R* ystage = s_y + threadIdx.x * (SC + 1);
    // column-sum role: species j of the chunk, cells sub, sub + 32, ...
    const int cj = threadIdx.x / G, csub = threadIdx.x % G;
    // store role: entry cj_store of rows cl_store, cl_store + 256 / SC, ...
    const int cj_store = threadIdx.x % SC, cl_store = threadIdx.x / SC;
    const int64_t ntiles = (Nc + kYieldThreads - 1) / kYieldThreads;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t c0 = tile * kYieldThreads;
        const int64_t c = c0 + threadIdx.x;
        const int nc = static_cast<int>(min<int64_t>(kYieldThreads, Nc - c0));
        const bool valid = c < Nc;
        R T = R(0.15), muB = 0, muS = 0, muQ = 0, dsu = 0, Pi = 0;
        R cf[6] = {0, 0, 0, 0, 0, 0};
        if (valid) {
            T = a.cells[F_T * Nc + c];
            muB = a.cells[F_MUB * Nc + c];
            muS = a.cells[F_MUS * Nc + c];
            muQ = a.cells[F_MUQ * Nc + c];
            dsu = a.cells[F_DA0 * Nc + c];
            if (MODE != BULK_NONE) {
                Pi = a.cells[F_PI * Nc + c];
                if (MODE == BULK_RTA) Pi = Pi / static_cast<R>(kHbarC);
#pragma unroll
                for (int k = 0; k < 6; k++) cf[k] = static_cast<R>(a.coef[k * Nc + c]);
            }
        }
        // the cell's loads complete here, once per tile: a wait placed inside
        // the species loop would also wait for the CDF stores issued since
        // (loads and stores share one counter)
        settle(T); settle(muB); settle(muS); settle(muQ); settle(dsu); settle(Pi);
#pragma unroll
        for (int k = 0; k < 6; k++) settle(cf[k]);
        const R beta = R(1) / T;
        const double betad = 1. / static_cast<double>(T);
        double run = 0.;
        for (int s0 = 0; s0 < Ns; s0 += SC) {
            const int sn = min(SC, Ns - s0);
            SpeciesK qn = spk[s0];
#pragma unroll 2
            for (int j = 0; j < sn; j++) {
                const int s = s0 + j;
                const SpeciesK q = qn;
                if (j + 1 < sn) qn = spk[s + 1];  // the next species' constants in flight
                R y = 0;
                if (valid && !(a.lcc && q.Q < 0.f)) {
                    const R dnq = a.dnq ? static_cast<R>(a.dnq[c * Ns + s]) : R(0);
                    if constexpr (sizeof(R) == sizeof(double)) {
                        const KLookup64<EXACT> kl{reinterpret_cast<const double*>(s_k), a.ktab, a.oob, true};
                        y = yield_one<R, MODE>(q.md, q.prefd, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu,
                                               Pi, cf, kl, (bool*)nullptr, dnq, q.md, betad);
                    } else {
                        const KLookup32<EXACT> kl{reinterpret_cast<const float*>(s_k), a.ktabf, a.oob};
                        bool ill = false;
                        y = yield_one<R, MODE>(q.m, q.pref, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu, Pi,
                                               cf, kl, (MODE == BULK_NONE && !a.dnq) ? (bool*)nullptr : &ill, dnq,
                                               q.md, betad);
                        if (ill) y = static_cast<R>(yield_fp64_global<MODE, EXACT>(a, c, s));
                    }
                    if (RAW) static_cast<R*>(a.raw)[c * Ns + s] = y;
                }
                if (a.ycol && valid) a.ycol[static_cast<int64_t>(s) * Nc + c] = static_cast<double>(y);
                run += static_cast<double>(y);
                stage[j] = run;
                ystage[j] = y;
            }
            __syncthreads();
            // the chunk's CDF entries: sn consecutive doubles per cell row
            // (thread t stores entry t % SC of rows t / SC, t / SC + 256 / SC, ...)
            if (cj_store < sn)
                for (int cl = cl_store; cl < nc; cl += kYieldThreads / SC)
                    a.cdf[(c0 + cl) * cdf_stride(Ns) + s0 + cj_store] = s_stage[cl * (SC + 1) + cj_store];
            // species sums of the chunk (the group of species j is uniform)
            if (cj < sn) {
                double v = 0.;
                for (int cl = csub; cl < nc; cl += G) v += static_cast<double>(s_y[cl * (SC + 1) + cj]);
#pragma unroll
                for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, G);
                if (csub == 0) s_acc[s0 + cj] += v;
            }
            __syncthreads();
        }
        if (valid) a.ctot[c] = run;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < Ns; s += blockDim.x) a.partial[static_cast<int64_t>(blockIdx.x) * Ns + s] = s_acc[s];
}

// FP64 per-species totals: one workgroup per species, fixed summation order
// an int flag cleared by a kernel (a memset node would not be safe in a
// captured graph)
__global__ void k_clear_flag(int* flag) {
    if (threadIdx.x == 0) *flag = 0;
}

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


// dsigma . q as the reference forms it from its float records
// (src/FSSW.cpp:627-629): a float expression, each product and sum rounded
// to float (a fused multiply-add would round differently, and the yield
// amplifies that where its terms cancel)
__device__ __forceinline__ float dot3_unfused(float a0, float b0, float a1, float b1, float a2, float b2) {
#pragma clang fp contract(off)
    return a0 * b0 + a1 * b1 + a2 * b2;
}

// The diffusion correction of every (cell, species) yield, FP64
// (src/FSSW.cpp:622-637, 696-702, 783-829):
//   dN_qmu = g / (2 pi^2 hbar c^3) (dsigma . q) / D_qmu
//            (-n_B/(e+P) m^2 T^2 S_1 - B T^3 / 3 S_2),
//   S_1 = sum_n theta_n / n lambda^n K_2(n m / T),
//   S_2 = sum_n n theta_n lambda^n I_1(n m / T) (-(m / T)^3),
// with the equilibrium series' truncation (10 terms below 0.7 GeV).
__global__ void k_diffusion_dn(YieldArgs a, const double* __restrict__ etab, double* __restrict__ dnq) {
    const int64_t Nc = a.Nc;
    const int Ns = a.Ns;
    const int64_t items = Nc * Ns;
    const KLookup64<true> kl{nullptr, a.ktab, a.oob, false};
    const double uf = 1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi);
    // dnq is cell-major, [Nc][Ns], as the yield rows
    for (int64_t it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; it < items;
         it += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = it / Ns;
        const int s = static_cast<int>(it - c * Ns);
        const double T = a.cells[F_T * Nc + c], beta = 1. / T;
        const float dqf = dot3_unfused(a.cells[F_QX * Nc + c], a.cells[F_DA1 * Nc + c], a.cells[F_QY * Nc + c],
                                       a.cells[F_DA2 * Nc + c], a.cells[F_QZ * Nc + c], a.cells[F_DA3 * Nc + c]);
        const double dq = static_cast<double>(dqf) * a.coef[6 * Nc + c];
        const double pq = a.coef[7 * Nc + c];
        const double m = a.sp.mass[s];
        const double B = a.sp.B[s], S = a.sp.S[s], Q = a.sp.Q[s], sg = a.sp.sign[s];
        const double mu = B * a.cells[F_MUB * Nc + c] + S * a.cells[F_MUS * Nc + c] + Q * a.cells[F_MUQ * Nc + c];
        const double lambda = exp(beta * mu);
        const int nt = (m < 0.7 && T > 0.05) ? 10 : 1;
        const double mb = m * beta;
        double s1 = 0., s2 = 0., th = 1., fug = lambda;
        for (int n = 1; n <= nt; n++) {
            const double arg = n * m * beta;
            s1 += th / n * fug * kl(2, arg);
            s2 += n * th * fug * (-(mb * mb * mb) * diffusion_i1(etab, arg));
            th *= -sg;
            fug *= lambda;
        }
        const double pref = static_cast<double>(a.sp.g[s]) * uf;
        dnq[it] = pref * dq * (-pq * (m * m * T * T * s1) - B * (T * T * T / 3.) * s2);
    }
}


// ============================================================ yield rows
// The yield pass of one surface sampled by Poisson splitting: every
// (cell, species) yield in the working precision, stored as cell rows of
// row_stride() entries (64-byte rows: 16 FP32 or 8 FP64 species per 64-byte
// segment, the padding zero), the FP64 per-cell totals and the workgroup's
// FP64 species partials.  The sampling pass builds each cell's FP64 species
// CDF from its row when it picks species (k_pick), so no CDF is stored.
//
// One cell per thread, tiles of 256 cells (persistent over tiles); the
// species loop is uniform across the workgroup, a 64-byte segment of species
// at a time:
//   * the species constants come from a padded table staged in LDS
//     (broadcast reads that wait in order with the K reads);
//   * FP32, one-term species (above 0.7 GeV): the fugacity is a product of
//     per-cell powers e^{beta mu q} picked by the species' charges, the K
//     table position t = 20 m / T - 10 is one FP64 fma, and the 14-moment
//     (or RTA / Chapman-Enskog / 22-moment) bulk terms are formed from
//     per-cell products, all with lambda factored out (the same sums as the
//     general series with n = 1);
//   * FP32, light species (the leading ones below 0.7 GeV): the same ten
//     terms as the general series, unrolled;
//   * diffusion: the per-item correction is added to either;
//   * FP64 mode, the exact-K rerun, 22-moment bulk delta f and
//     species lists outside these assumptions take the general series
//     (yield_one);
//   * items whose terms cancel beyond kIllCond are recomputed in FP64, 64
//     at a time from a per-wave queue (one item per lane);
//   * each thread stores its segment with 16-byte stores (whole 64-byte
//     sectors), adds the yields to its FP64 cell total in species order, and
//     the species' FP64 sums over the wave's 64 cells come from DPP
//     reductions, accumulated per wave in LDS (a fixed order).
constexpr int kRowThreads = 256;
constexpr int kKRows = 1024;     // K rows staged in LDS (arguments below 51.7)
constexpr int kLightMax = 16;

__host__ __device__ constexpr int row_chunk(int esz) { return esz == 8 ? 8 : 16; }
__host__ __device__ constexpr int row_stride(int ns, int esz) {
    return (ns + row_chunk(esz) - 1) / row_chunk(esz) * row_chunk(esz);
}

// K rows in LDS: FP64 values (the reference's interpolation), or FP32
// (K_i, K_{i+1} - K_i) pairs (one 8-byte read per order)
template <typename R>
struct RowK {
    using T = double;
};
template <>
struct RowK<float> {
    using T = float2;
};

template <bool EXACT>
struct KLookupRows32 {
    const float2* lds;   // [NT][kKRows], K_2, K_1, K_3
    const float2* glob;  // [3][kKLen], K_1, K_2, K_3
    int* oob;
    template <int NK>
    __device__ __forceinline__ void k123(double arg, float& K2, float& K1, float& K3) const {
        if constexpr (EXACT) {
            const KLookup32<true> g{nullptr, glob, oob};
            g.template k123<NK>(arg, K2, K1, K3);
        } else {
            at<NK>((arg - kKx0) * 20.0, K2, K1, K3);
        }
    }
    // the same at table position t = (x - 0.5) / 0.05
    template <int NK>
    __device__ __forceinline__ void at(double t, float& K2, float& K1, float& K3) const {
        if (!at_ok<NK>(t, K2, K1, K3)) *oob = 1;
    }
    // ... returning whether t lies in the staged rows (the caller flags it)
    template <int NK>
    __device__ __forceinline__ bool at_ok(double t, float& K2, float& K1, float& K3) const {
        int idx = static_cast<int>(t);
        const bool ok = (t >= 0.) & (idx + 1 < kKRows);
        idx = ok ? idx : 0;
        const float fr = static_cast<float>(t - static_cast<double>(idx));
        const float2 v2 = lds[idx];
        float2 v1 = make_float2(0.f, 0.f), v3 = make_float2(0.f, 0.f);
        if (NK >= 2) v1 = lds[kKRows + idx];
        if (NK >= 3) v3 = lds[2 * kKRows + idx];
        K2 = fmaf(fr, v2.y, v2.x);
        if (NK >= 2) K1 = fmaf(fr, v1.y, v1.x);
        if (NK >= 3) K3 = fmaf(fr, v3.y, v3.x);
        return ok;
    }
};

// the FP64 fallback (cold: ~0.5 % of the items)
template <int MODE, bool EXACT>
__device__ __forceinline__ float yield_fp64_cold(const YieldArgs& a, int64_t c, int s) {
    return static_cast<float>(yield_fp64_global<MODE, EXACT>(a, c, s));
}

// per-cell quantities of the FP32 one-term path
struct RowCellF {
    float T, beta, Pi, pdc;
    double betad;
    float eBp, eBm, eSp, eSm, eQp, eQm;  // e^{+beta mu}, e^{-beta mu} of the three charges
    float muB, muS, muQ, dsu;
    float k1, k2, k3, k4, k5;   // per-cell bulk products (MODE-dependent)
};

template <int MODE>
__device__ __forceinline__ void row_cell_f(const YieldArgs& a, int64_t c, bool valid, RowCellF& p, float cf[6]) {
    const int64_t Nc = a.Nc;
    double T = 0.15, muB = 0., muS = 0., muQ = 0., dsu = 0., Pi = 0.;
    double cd[6] = {0., 0., 0., 0., 0., 0.};
    if (valid) {
        T = a.cells[F_T * Nc + c];
        muB = a.cells[F_MUB * Nc + c];
        muS = a.cells[F_MUS * Nc + c];
        muQ = a.cells[F_MUQ * Nc + c];
        dsu = a.cells[F_DA0 * Nc + c];
        if (MODE != BULK_NONE) {
            Pi = a.cells[F_PI * Nc + c];
#pragma unroll
            for (int k = 0; k < 6; k++) cd[k] = a.coef[k * Nc + c];
        }
    }
    p.T = static_cast<float>(T);
    p.muB = static_cast<float>(muB);
    p.muS = static_cast<float>(muS);
    p.muQ = static_cast<float>(muQ);
    p.dsu = static_cast<float>(dsu);
    p.betad = 1. / T;
    p.beta = static_cast<float>(p.betad);
    // the general series divides the float Pi by hbar c in FP32 (RTA)
    float Pif = static_cast<float>(Pi);
    if (MODE == BULK_RTA) Pif = Pif / static_cast<float>(kHbarC);
    p.Pi = Pif;
    p.pdc = static_cast<float>(dsu) * static_cast<float>(1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi));
    const double xb = p.betad * muB, xs = p.betad * muS, xq = p.betad * muQ;
    p.eBp = static_cast<float>(exp(xb));
    p.eBm = static_cast<float>(exp(-xb));
    p.eSp = static_cast<float>(exp(xs));
    p.eSm = static_cast<float>(exp(-xs));
    p.eQp = static_cast<float>(exp(xq));
    p.eQm = static_cast<float>(exp(-xq));
#pragma unroll
    for (int k = 0; k < 6; k++) cf[k] = static_cast<float>(cd[k]);
    const double T2 = T * T;
    p.k1 = p.k2 = p.k3 = p.k4 = p.k5 = 0.f;
    if (MODE == BULK_14MOM) {
        p.k1 = static_cast<float>(T * cd[0]);        // m^4 T C0 K2
        p.k2 = static_cast<float>(T2 * cd[1]);       // m^2 B T^2 C1 (m b K1 + 3 K2)
        p.k3 = static_cast<float>(T2 * cd[2]);       // m^3 T^2 C2 (m b K2 + 3 K3)
    } else if (MODE == BULK_RTA || MODE == BULK_CE) {
        p.k1 = static_cast<float>(static_cast<double>(cf[0]) * static_cast<double>(cf[1]) * T);  // c0 c1 T m^2 B1
        p.k2 = cf[0];                                                                              // - c0 m^3 / 3 K1
    } else if (MODE == BULK_22MOM) {
        p.k1 = static_cast<float>(T * cd[2]);            // m^4 T c2 K2
        p.k2 = static_cast<float>(T2 * (cd[1] - cd[2])); // m^3 T^2 (c1 - c2) B2
        p.k3 = static_cast<float>(T2 * cd[3]);           // m^2 T^2 (B c3 + S c4 + Q c5) B1
        p.k4 = static_cast<float>(T2 * cd[4]);
        p.k5 = static_cast<float>(T2 * cd[5]);
    }
}

// FP32 yield of a one-term species (n = 1) from the padded species table:
// the general series' sums with lambda factored out, branch-free: the
// fugacity is the product of three per-cell factors read from the thread's
// LDS column (rows picked by the species' charges), K from the staged rows.
// Returns the unclamped yield; ill: the terms cancel beyond kIllCond; ok: the
// K argument lies in the staged rows.
template <int MODE>
__device__ __forceinline__ float row_yield_sp(const RowSp& q, const RowCellF& p, const float* fug,
                                              const KLookupRows32<false>& kl, bool& ill, bool& ok) {
    float K2 = 0.f, K1 = 0.f, K3 = 0.f;
    constexpr int NK = (MODE == BULK_NONE) ? 1 : ((MODE == BULK_RTA || MODE == BULK_CE) ? 2 : 3);
    ok = kl.template at_ok<NK>(fma(q.md20, p.betad, -10.0), K2, K1, K3);
    const float lam = fug[(q.frow & 0xFF) * kRowThreads] * fug[((q.frow >> 8) & 0xFF) * kRowThreads] *
                      fug[(q.frow >> 16) * kRowThreads];
    const float a0 = (q.m2 * p.T) * K2;
    float br = a0;
    ill = false;
    if (MODE == BULK_14MOM) {
        const float mb = q.m * p.beta;
        const float B1 = fmaf(mb, K1, 3.f * K2), B2 = fmaf(mb, K2, 3.f * K3);
        const float t1 = (q.m4 * p.k1) * K2, t2 = (q.m2B * p.k2) * B1, t3 = (q.m3 * p.k3) * B2;
        br = fmaf(p.Pi, t1 + t2 + t3, a0);
        ill = fmaf(fabsf(p.Pi), fabsf(t1) + fabsf(t2) + fabsf(t3), fabsf(a0)) > kIllCond * fabsf(br);
    } else if (MODE == BULK_RTA || MODE == BULK_CE) {
        const float mb = q.m * p.beta;
        const float B1 = fmaf(mb, K1, 3.f * K2);
        const float t1 = (q.m2 * p.k1) * B1, t2 = -(q.m3 * (1.f / 3.f) * p.k2) * K1;
        br = fmaf(p.Pi, t1 + t2, a0);
        ill = fmaf(fabsf(p.Pi), fabsf(t1) + fabsf(t2), fabsf(a0)) > kIllCond * fabsf(br);
    }
    return (q.g * p.pdc) * lam * br;
}

// FP32 yield of a light species (below 0.7 GeV): the general series' ten
// terms (one where T <= 0.05 GeV), unrolled so that their K reads are in
// flight together; lambda from the fugacity column, the constants from the
// padded table (sgn: +1 fermions, -1 bosons).  Returns the unclamped yield;
// ill and ok as row_yield_sp.
constexpr int kLightTerms = 10;
template <int MODE>
__device__ __forceinline__ float row_yield_light(const RowSp& q, const RowCellF& p, const float* fug,
                                                 const KLookupRows32<false>& kl, bool& ill, bool& ok) {
    constexpr int NK = (MODE == BULK_NONE) ? 1 : ((MODE == BULK_RTA || MODE == BULK_CE) ? 2 : 3);
    const float lam = fug[(q.frow & 0xFF) * kRowThreads] * fug[((q.frow >> 8) & 0xFF) * kRowThreads] *
                      fug[(q.frow >> 16) * kRowThreads];
    const float more = p.T > 0.05f ? 1.f : 0.f;  // weight of the terms beyond the first
    const float mb = q.m * p.beta;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, r3 = 0.f;
    float th = 1.f, fl = lam;
    ok = true;
#pragma unroll
    for (int n = 1; n <= kLightTerms; n++) {
        float K2 = 0.f, K1 = 0.f, K3 = 0.f;
        const bool okn = kl.template at_ok<NK>(fma(static_cast<double>(n) * q.md20, p.betad, -10.0), K2, K1, K3);
        const float w = n == 1 ? 1.f : more;
        ok = ok && (okn || w == 0.f);
        const float inv = 1.f / static_cast<float>(n);
        const float thn = th * inv, k2n = 3.f * K2 * inv, tf = w * (th * fl);
        r0 += w * (thn * fl * K2);
        if (MODE == BULK_RTA || MODE == BULK_CE) {
            r1 += tf * fmaf(mb, K1, k2n);
            r2 += tf * K1;
        } else if (MODE == BULK_14MOM) {
            r1 += tf * K2;
            r2 += tf * fmaf(mb, K1, k2n);
            r3 += tf * fmaf(mb, K2, 3.f * K3 * inv);
        }
        th *= -q.sgn;
        fl *= lam;
    }
    const float pd = q.g * p.pdc;
    const float Neq = pd * (q.m2 * p.T * r0);
    ill = false;
    if (MODE == BULK_RTA || MODE == BULK_CE) {
        const float t1 = (q.m2 * p.k1) * r1, t2 = -(q.m3 * (1.f / 3.f) * p.k2) * r2;
        const float N = fmaf(pd * p.Pi, t1 + t2, Neq);
        ill = fmaf(fabsf(pd * p.Pi), fabsf(t1) + fabsf(t2), fabsf(Neq)) > kIllCond * fabsf(N);
        return N;
    } else if (MODE == BULK_14MOM) {
        const float t1 = (q.m4 * p.k1) * r1, t2 = (q.m2B * p.k2) * r2, t3 = (q.m3 * p.k3) * r3;
        const float N = fmaf(pd * p.Pi, t1 + t2 + t3, Neq);
        ill = fmaf(fabsf(pd * p.Pi), fabsf(t1) + fabsf(t2) + fabsf(t3), fabsf(Neq)) > kIllCond * fabsf(N);
        return N;
    }
    return Neq;
}

// fugacity rows of the LDS column: e^{beta mu_B q}, q = -1..1; e^{beta mu_S q},
// q = -3..3; e^{beta mu_Q q}, q = -2..2
constexpr int kFugRows = 15;

template <typename R>
__host__ __device__ constexpr size_t yield_rows_lds_bytes(int ntab, int ns) {
    return static_cast<size_t>(ntab) * kKRows * sizeof(typename RowK<R>::T)               // K rows
           + static_cast<size_t>(kRowThreads / kWave) * row_stride(ns, sizeof(R)) * sizeof(double)   // wave sums
           + (sizeof(R) == sizeof(float)  // fugacities, the padded species table, the FP64 queues
                  ? static_cast<size_t>(kFugRows) * kRowThreads * sizeof(float) + row_stride(ns, 4) * sizeof(RowSp) +
                        static_cast<size_t>(kRowThreads) * (sizeof(double) + sizeof(int) + sizeof(float))
                  : 0);
}

template <typename R, int MODE, bool EXACT>
__global__ __launch_bounds__(kRowThreads) void k_yield_rows(YieldArgs a, const SpeciesK* __restrict__ spk,
                                                            const RowSp* __restrict__ rowsp) {
    // the exact-K pass after a first pass that stayed inside the staged rows:
    // nothing to redo (uniform over the grid; the flag only ever turns on)
    if (EXACT && a.gate && *a.gate == 0) return;
    extern __shared__ __align__(16) unsigned char smem[];
    using KT = typename RowK<R>::T;
    constexpr int NT = ktables_for(MODE);
    constexpr int SC = row_chunk(sizeof(R));
    constexpr bool F32 = sizeof(R) == sizeof(float);
    const int Ns = a.Ns, L = a.nlight, stride = row_stride(Ns, sizeof(R));
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    KT* s_k = reinterpret_cast<KT*>(smem);
    double* s_acc = reinterpret_cast<double*>(s_k + NT * kKRows);  // [waves][stride]
    float* s_fug = reinterpret_cast<float*>(s_acc + (kRowThreads / kWave) * stride);  // [kFugRows][threads]
    float* fug = s_fug + threadIdx.x;
    // the padded species table: broadcast LDS reads, in order with the K and
    // fugacity reads (scalar loads would share their wait counter out of order)
    RowSp* s_sp = reinterpret_cast<RowSp*>(s_fug + kFugRows * kRowThreads);
    // the wave's queue of items to recompute in FP64
    double* q_d = reinterpret_cast<double*>(s_sp + (F32 ? row_stride(Ns, 4) : 0)) + w * kWave;
    int* q_key = reinterpret_cast<int*>(q_d - w * kWave + kRowThreads) + w * kWave;
    float* q_old = reinterpret_cast<float*>(q_key - w * kWave + kRowThreads) + w * kWave;
    R* rows = static_cast<R*>(a.rows);
    if (F32 && rowsp) {
        const int n = row_stride(Ns, 4) * static_cast<int>(sizeof(RowSp) / sizeof(int));
        const int* src = reinterpret_cast<const int*>(rowsp);
        int* dstp = reinterpret_cast<int*>(s_sp);
        for (int i = threadIdx.x; i < n; i += blockDim.x) dstp[i] = src[i];
    }

    if (!EXACT) {
        for (int i = threadIdx.x; i < NT * kKRows; i += blockDim.x) {
            const int t = i / kKRows, r = i % kKRows;
            const int o = (t == 0) ? 2 : (t == 1 ? 1 : 3);
            if constexpr (F32) s_k[i] = a.ktabf[(o - 1) * kKLen + r];
            else s_k[i] = a.ktab[(o - 1) * kKLen + r];
        }
    }
    for (int i = threadIdx.x; i < (kRowThreads / kWave) * stride; i += blockDim.x) s_acc[i] = 0.;
    __syncthreads();
    double* wacc = s_acc + w * stride;

    const int64_t Nc = a.Nc;
    const int64_t ntiles = (Nc + kRowThreads - 1) / kRowThreads;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t c = tile * kRowThreads + threadIdx.x;
        const bool valid = c < Nc;
        // the cell's parameters for the general series (working precision)
        R T = R(0.15), muB = 0, muS = 0, muQ = 0, dsu = 0, Pi = 0;
        R cf[6] = {0, 0, 0, 0, 0, 0};
        RowCellF pf;
        float cff[6];
        if constexpr (F32) {
            row_cell_f<MODE>(a, c, valid, pf, cff);
            T = pf.T;
            muB = valid ? a.cells[F_MUB * Nc + c] : 0.f;
            muS = valid ? a.cells[F_MUS * Nc + c] : 0.f;
            muQ = valid ? a.cells[F_MUQ * Nc + c] : 0.f;
            dsu = valid ? a.cells[F_DA0 * Nc + c] : 0.f;
            Pi = pf.Pi;
#pragma unroll
            for (int k = 0; k < 6; k++) cf[k] = cff[k];
        } else if (valid) {
            T = a.cells[F_T * Nc + c];
            muB = a.cells[F_MUB * Nc + c];
            muS = a.cells[F_MUS * Nc + c];
            muQ = a.cells[F_MUQ * Nc + c];
            dsu = a.cells[F_DA0 * Nc + c];
            if (MODE != BULK_NONE) {
                Pi = a.cells[F_PI * Nc + c];
                if (MODE == BULK_RTA) Pi = Pi / static_cast<R>(kHbarC);
#pragma unroll
                for (int k = 0; k < 6; k++) cf[k] = a.coef[k * Nc + c];
            }
        }
        const R beta = R(1) / T;
        const double betad = 1. / static_cast<double>(T);
        double ctot = 0.;
        R* row = rows + c * stride;
        // the general series of species s (light species, FP64, the exact-K
        // rerun, lists whose light species are not all first)
        auto general = [&](int s) -> R {
            const SpeciesK q = spk[s];
            if (!valid || (a.lcc && q.Q < 0.f)) return R(0);
            const R dnq = a.dnq ? static_cast<R>(a.dnq[c * Ns + s]) : R(0);
            if constexpr (F32) {
                bool ill = false;
                const KLookupRows32<EXACT> kl{reinterpret_cast<const float2*>(s_k), a.ktabf, a.oob};
                R v = yield_one<float, MODE>(q.m, q.pref, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu, Pi, cf,
                                             kl, (MODE == BULK_NONE && !a.dnq) ? (bool*)nullptr : &ill, dnq, q.md,
                                             betad);
                if (ill) v = yield_fp64_cold<MODE, EXACT>(a, c, s);
                return v;
            } else {
                const KLookup64<EXACT, kKRows> kl{reinterpret_cast<const double*>(s_k), a.ktab, a.oob, true};
                return yield_one<double, MODE>(q.md, q.prefd, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu, Pi,
                                               cf, kl, (bool*)nullptr, dnq, q.md, betad);
            }
        };
        // every species on the general path, one at a time with word stores
        auto one_by_one = [&]() {
#pragma unroll 1
            for (int s = 0; s < stride; s++) {
                const R v = s < Ns ? general(s) : R(0);
                if (valid) row[s] = v;
                ctot += static_cast<double>(v);
                const double t = wave_sum_dpp(static_cast<double>(v));
                if (lane == 0) wacc[s] += t;
            }
        };
        // (the 22-moment bulk terms take the general series)
        bool fast = false;
        if constexpr (F32 && !EXACT && MODE != BULK_22MOM) fast = !a.all_general && a.rowsp;
        if constexpr (F32 && !EXACT && MODE != BULK_22MOM) if (fast) {
            {
                const double xb = pf.betad * static_cast<double>(muB), xs = pf.betad * static_cast<double>(muS),
                             xq = pf.betad * static_cast<double>(muQ);
#pragma unroll
                for (int q = -1; q <= 1; q++) fug[(q + 1) * kRowThreads] = static_cast<float>(exp(q * xb));
#pragma unroll
                for (int q = -3; q <= 3; q++) fug[(q + 6) * kRowThreads] = static_cast<float>(exp(q * xs));
#pragma unroll
                for (int q = -2; q <= 2; q++) fug[(q + 12) * kRowThreads] = static_cast<float>(exp(q * xq));
            }
            const KLookupRows32<false> kl1{reinterpret_cast<const float2*>(s_k), a.ktabf, a.oob};
            bool bad = false;  // a K argument beyond the staged rows
            // the diffusion correction of (c, s) joins the terms that may
            // cancel (zero weight: none)
            const double* dnq_row = (a.dnq && valid) ? a.dnq + c * Ns : nullptr;
            auto add_dnq = [&](int s, const RowSp& q, float& v, bool& ill) {
                const float dq = (dnq_row && s < Ns && q.g != 0.f) ? static_cast<float>(dnq_row[s]) : 0.f;
                const float N = v + dq;
                ill = ill || fabsf(v) + fabsf(dq) > kIllCond * fabsf(N);
                v = N;
            };
            // Items whose terms cancel beyond kIllCond (rare, and never of
            // zero weight) wait in the wave's queue: lane | species << 6 and
            // the FP32 value stored.  A flush recomputes up to 64 of them in
            // FP64, one per lane, rewrites their row entries, and adds the
            // differences to the cell totals and species sums in queue order.
            int qn = 0;  // queued items (uniform)
            auto flush = [&]() {
                wave_lds_sync();
                if (lane < qn) {
                    const int key = q_key[lane];
                    const int64_t cc = tile * kRowThreads + w * kWave + (key & (kWave - 1));
                    const int s = key >> 6;
                    const float v = yield_fp64_cold<MODE, EXACT>(a, cc, s);
                    // after the FP32 store of the same entry by this wave
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    reinterpret_cast<float*>(rows)[cc * stride + s] = v;
                    q_d[lane] = static_cast<double>(v) - static_cast<double>(q_old[lane]);
                }
                wave_lds_sync();
#pragma unroll 1
                for (int k = 0; k < qn; k++) {
                    const int key = q_key[k];
                    const double d = q_d[k];
                    if ((key & (kWave - 1)) == lane) ctot += d;
                    if (lane == 0) wacc[key >> 6] += d;
                }
                wave_lds_sync();
                qn = 0;
            };
            // the first segment: the light species (ten-term series), then
            // one-term species, one at a time with word stores
            auto first = [&](int s, float v, bool ill, bool ok, const RowSp& q) {
                if (a.dnq) add_dnq(s, q, v, ill);
                bad |= !ok;
                v = v > 0.f ? v : 0.f;
                if (valid) row[s] = v;
                ctot += static_cast<double>(v);
                const double t = wave_sum_dpp(static_cast<double>(v));
                if (lane == 0) wacc[s] += t;
                const bool mine = ill && valid && q.g != 0.f;
                const uint64_t m = __ballot(mine);
                if (m != 0u) {
                    const int n = __popcll(m);
                    if (qn + n > kWave) flush();
                    if (mine) {
                        const int pos =
                            qn + static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                                            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
                        q_key[pos] = lane | (s << 6);
                        q_old[pos] = v;
                    }
                    qn += n;
                }
            };
#pragma unroll 1
            for (int s = 0; s < L; s++) {
                const RowSp q = s_sp[s];
                bool ill = false, ok = true;
                const float v = row_yield_light<MODE>(q, pf, fug, kl1, ill, ok);
                first(s, v, ill, ok, q);
            }
#pragma unroll 1
            for (int s = L; s < SC; s++) {
                const RowSp q = s_sp[s];
                bool ill = false, ok = true;
                const float v = row_yield_sp<MODE>(q, pf, fug, kl1, ill, ok);
                first(s, v, ill, ok, q);
            }
            // the one-term species, a 64-byte segment at a time, from the
            // padded table (zero weight beyond Ns and for suppressed
            // species; lanes of cells beyond the surface have dsigma = 0)
            for (int s0 = SC; s0 < stride; s0 += SC) {
                float y[SC];
                uint32_t illm = 0u;  // items to recompute in FP64
#pragma unroll
                for (int j = 0; j < SC; j++) {
                    const RowSp q = s_sp[s0 + j];
                    bool ill = false, ok = true;
                    float v = row_yield_sp<MODE>(q, pf, fug, kl1, ill, ok);
                    if (a.dnq) add_dnq(s0 + j, q, v, ill);
                    v = v > 0.f ? v : 0.f;
                    illm |= (ill && q.g != 0.f) ? (1u << j) : 0u;
                    bad |= !ok;
                    y[j] = v;
                    ctot += static_cast<double>(v);
                }
                float* dst = reinterpret_cast<float*>(row) + s0;
                if (valid) {
#pragma unroll
                    for (int k = 0; k < SC; k += 4)
                        *reinterpret_cast<float4*>(dst + k) = make_float4(y[k], y[k + 1], y[k + 2], y[k + 3]);
                }
#pragma unroll
                for (int h = 0; h < SC; h += 8) {
                    double yd[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) yd[j] = static_cast<double>(y[h + j]);
                    const double t = wave_sum8(yd);
                    if ((lane & 7) == 0) wacc[s0 + h + wave_sum8_index(lane)] += t;
                }
                // the segment's ill items join the queue in (lane, species)
                // order, a queue's worth at a time
                if (!valid) illm = 0u;
                if (__any(illm != 0u)) {
                    const int k = __popc(illm);
                    const int incl = wave_inclusive_scan(k);
                    const int total = __shfl(incl, kWave - 1, kWave);
                    if (qn + total > kWave) flush();
                    for (int b = 0; b < total; b += kWave) {
                        if (b > 0) flush();
                        int pos = incl - k - b;  // this lane's first item, in this chunk
                        uint32_t m = illm;
                        while (m != 0u) {
                            const int j = __builtin_ctz(m);
                            m &= m - 1u;
                            if (pos >= 0 && pos < kWave) {
                                float old = y[0];
#pragma unroll
                                for (int t = 1; t < SC; t++) old = j == t ? y[t] : old;
                                q_key[qn + pos] = lane | ((s0 + j) << 6);
                                q_old[qn + pos] = old;
                            }
                            pos++;
                        }
                        qn += min(kWave, total - b);
                    }
                }
            }
            flush();
            if (__any(bad && valid) && lane == 0) *a.oob = 1;
        }
        if (!fast) one_by_one();
        if (valid && a.ctot) a.ctot[c] = ctot;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < Ns; s += blockDim.x) {
        double t = 0.;
        for (int k = 0; k < kRowThreads / kWave; k++) t += s_acc[k * stride + s];
        a.partial[static_cast<int64_t>(blockIdx.x) * Ns + s] = t;
    }
}

}  // namespace issdev
