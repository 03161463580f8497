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
template <bool EXACT>
struct KLookup64 {
    const double* lds;   // [3][kKLds] staged rows (used when `staged`)
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
            if (idx + 1 >= kKLds) {
                *oob = 1;
                return 0.;
            }
            k0 = lds[kslot(order) * kKLds + idx];
            k1 = lds[kslot(order) * kKLds + idx + 1];
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
    // a branch around the reads
    template <int NK>
    __device__ __forceinline__ void k123(float arg, float& K2, float& K1, float& K3) const {
        if constexpr (EXACT) {
            K2 = (*this)(2, arg);
            if (NK >= 2) K1 = (*this)(1, arg);
            if (NK >= 3) K3 = (*this)(3, arg);
        } else {
            const float t = (arg - 0.5f) * 20.f;
            int idx = static_cast<int>(t);
            const bool ok = (arg >= 0.5f) & (arg <= 399.95f) & (idx + 1 < kKLds);
            if (!ok) *oob = 1;
            idx = ok ? idx : 0;
            const float fr = t - static_cast<float>(idx);
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
// (|N_eq| + |dN_bulk|) > 4 |N|: rounding errors are amplified when the bulk
// correction nearly cancels the equilibrium yield.

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

// dnq: the diffusion correction of this (cell, species), computed by
// k_diffusion_dn (0 without diffusion delta f)
template <typename R, int MODE, typename KL>
__device__ __forceinline__ R yield_one(R mass, R pref, R Bq, R Sq, R Qq, R sgn, R T, R beta,
                                       R muB, R muS, R muQ, R dsu, R Pi, const R* c,
                                       const KL& kl, bool* ill, R dnq = R(0)) {
    const R mu = Bq * muB + Sq * muS + Qq * muQ;
    const R lambda = fexp(beta * mu);
    const int nt = (mass < R(0.7) && T > R(0.05)) ? 10 : 1;
    const R mb = mass * beta;
    R r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    R th = 1, fug = lambda;
    auto term = [&](R rn, R inv) {
        const R arg = rn * mass * beta;
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
    dN += dnq;
    const R N = Neq + dN;
    // ill-conditioned: the corrections nearly cancel N_eq
    if (ill) *ill = (fabs(Neq) + fabs(dN)) > R(4) * fabs(N);
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
    return yield_one<double, MODE>(a.sp.mass[s], pref, a.sp.B[s], a.sp.S[s], a.sp.Q[s], a.sp.sign[s], T,
                                   1. / T, (double)a.cells[F_MUB * Nc + c], (double)a.cells[F_MUS * Nc + c],
                                   (double)a.cells[F_MUQ * Nc + c], (double)a.cells[F_DA0 * Nc + c], Pi, cd,
                                   k64, (bool*)nullptr, a.dnq ? a.dnq[static_cast<int64_t>(s) * Nc + c] : 0.);
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
                    const R dnq = a.dnq ? static_cast<R>(a.dnq[static_cast<int64_t>(s) * Nc + c]) : R(0);
                    if constexpr (sizeof(R) == sizeof(double)) {
                        const KLookup64<EXACT> kl{reinterpret_cast<const double*>(s_k), a.ktab, a.oob, true};
                        y = yield_one<R, MODE>(q.md, q.prefd, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu,
                                               Pi, cf, kl, (bool*)nullptr, dnq);
                    } else {
                        const KLookup32<EXACT> kl{reinterpret_cast<const float*>(s_k), a.ktabf, a.oob};
                        bool ill = false;
                        y = yield_one<R, MODE>(q.m, q.pref, q.B, q.S, q.Q, q.sgn, T, beta, muB, muS, muQ, dsu, Pi,
                                               cf, kl, MODE == BULK_NONE ? (bool*)nullptr : &ill, dnq);
                        if (MODE != BULK_NONE && ill) y = static_cast<R>(yield_fp64_global<MODE, EXACT>(a, c, s));
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
    for (int64_t it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; it < items;
         it += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int s = static_cast<int>(it / Nc);
        const int64_t c = it - static_cast<int64_t>(s) * Nc;
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

}  // namespace issdev
