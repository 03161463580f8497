// Yield pass: per-(cell, species) yields, FP64 species CDF rows and species totals.
// (one translation unit: included by iss_hip.hip)
#pragma once

#include "kernels.h"

namespace issdev {

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
    dN += dnq;
    const R N = Neq + dN;
    // ill-conditioned: the corrections nearly cancel N_eq
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
                                   k64, (bool*)nullptr, a.dnq ? a.dnq[c * a.Ns + s] : 0.);
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
                const R dnq = a.dnq ? static_cast<R>(a.dnq[(c0 + cl) * Ns + s]) : R(0);
                const char4 q = s_q[s];
                if constexpr (sizeof(R) == sizeof(double)) {
                    const KLookup64<EXACT> kl{reinterpret_cast<const double*>(s_k), a.ktab, a.oob};
                    y = yield_one<R, MODE>(s_mass[s], s_pref[s], q.x, q.y, q.z, q.w, p[0], p[1], p[2],
                                           p[3], p[4], p[5], p[6], p + 7, kl, (bool*)nullptr, dnq);
                } else {
                    const KLookup32<EXACT> kl{reinterpret_cast<const float2*>(s_k), a.ktabf, a.oob};
                    bool ill = false;
                    y = yield_one<R, MODE>(s_mass[s], s_pref[s], q.x, q.y, q.z, q.w, p[0], p[1], p[2],
                                           p[3], p[4], p[5], p[6], p + 7, kl,
                                           MODE == BULK_NONE ? (bool*)nullptr : &ill, dnq);
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
    const KLookup64<true> kl{nullptr, a.ktab, a.oob};
    const double uf = 1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi);
    for (int64_t it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; it < items;
         it += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = it / Ns;
        const int s = static_cast<int>(it % Ns);
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
