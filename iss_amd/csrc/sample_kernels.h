// Sampling pass: per-cell Poisson counts, scans, the event multinomial and the momentum sampler.
// (one translation unit: included by iss_hip.hip)
#pragma once

#include "kernels.h"

namespace issdev {

// ============================================================ counting pass

enum StreamTag : uint32_t {
    TAG_COUNT = 1, TAG_MULTI = 2, TAG_SETUP = 3, TAG_TRIAL = 4,
    TAG_LEARN = 5, TAG_LEARN2 = 6, TAG_TRAIN = 7
};

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

// Up to kEventHistMax events the sizes come without the tree: each of the K
// slots draws its event uniformly (the high word of a 64-bit uniform times
// E, a bias below E / 2^64), and the counts are histograms.  The K draws are
// independent and uniform, so the counts are Multinomial(K, 1/E) as well.
// One block per kEventHistBlocks share of the slots counts into an LDS
// histogram and adds it to evcount (zeroed beforehand) with integer atomics,
// whose sums do not depend on their order.  root: K on the device.
constexpr int kEventHistMax = 8192, kEventHistBlocks = 256;
constexpr uint32_t TAG_EVHIST = 12;
__global__ __launch_bounds__(256) void k_event_hist(const int64_t* __restrict__ root, int E, uint64_t seed,
                                                    uint32_t chunk, uint32_t* __restrict__ evcount) {
    __shared__ uint32_t hist[kEventHistMax];
    for (int e = threadIdx.x; e < E; e += blockDim.x) hist[e] = 0u;
    __syncthreads();
    const int64_t K = *root;
    const int64_t npair = (K + 1) / 2;  // one Philox block gives two draws
    const int64_t per = (npair + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t p1 = min(npair, p0 + per);
    const uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
    const uint32_t tw = tag_word(TAG_EVHIST, chunk);
    const uint64_t Eu = static_cast<uint64_t>(E);
    for (int64_t q = p0 + threadIdx.x; q < p1; q += blockDim.x) {
        const u32x4 w = philox4x32(u32x4{static_cast<uint32_t>(q), static_cast<uint32_t>(q >> 32), tw, 0u}, k0, k1);
        const uint64_t u0 = (static_cast<uint64_t>(w.x) << 32) | w.y;
        atomicAdd(&hist[__umul64hi(u0, Eu)], 1u);
        if (2 * q + 1 < K) {
            const uint64_t u1 = (static_cast<uint64_t>(w.z) << 32) | w.w;
            atomicAdd(&hist[__umul64hi(u1, Eu)], 1u);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < E; e += blockDim.x)
        if (hist[e]) atomicAdd(&evcount[e], hist[e]);
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

// Boost a rest-frame momentum to the lab frame (src/FSSW.cpp:2001-2010) and
// write the hadron with its space-time position (src/FSSW.cpp:1969-1996).
// Boost-invariant surfaces draw y uniform in [y_LB, y_RB] and set
// eta_s = y - (y - eta_s) (src/FSSW.cpp:1024-1029).
__device__ __forceinline__ void emit_lab(const SampleArgs& a, uint32_t chunk, int64_t c, uint32_t jj, int pid, float fm,
                                         float px, float py, float pz, float p0, uint64_t outp) {
    const int64_t Nc = a.Nc;
    const float u0 = a.cells[F_U0 * Nc + c], u1 = a.cells[F_U1 * Nc + c],
                u2 = a.cells[F_U2 * Nc + c], u3 = a.cells[F_U3 * Nc + c];
    const float pdu = px * u1 + py * u2 + pz * u3;
    const float k1 = pdu / (u0 + 1.f) + p0;
    const float lx = px + k1 * u1, ly = py + k1 * u2, lz = pz + k1 * u3;
    const float mT = sqrtf(lx * lx + ly * ly + fm * fm);
    const float tau = a.cells[F_TAU * Nc + c], eta = a.cells[F_ETA * Nc + c];
    float eta_s, E, pzl;
    float t_par = 0.f, z_par = 0.f;
    const bool partner = (jj & kPartnerBit) != 0u;
    if (a.boost_invariant && partner) {
        // a local-charge-conservation partner keeps its parent's eta_s
        // (src/FSSW.cpp:1044-1046); the parent is the previous hadron, from
        // the same cell: eta_s = asinh(z / tau) keeps full precision at any
        // eta_s, where 0.5 log((t + z) / (t - z)) cancels (t - z rounds to 0
        // in FP32 beyond |eta_s| ~ 8)
        const Hadron hp = static_cast<const Hadron*>(a.out)[outp - 1];
        t_par = hp.t;
        z_par = hp.z;
        eta_s = asinhf(hp.z / tau);
        const float rap = eta_s + (asinhf(lz / mT) - eta);
        E = mT * coshf(rap);
        pzl = mT * sinhf(rap);
    } else if (a.boost_invariant) {
        Stream st(a.seed, static_cast<uint32_t>(c), jj, tag_word(TAG_SETUP, chunk) ^ 0x00800000u);
        const float rap = a.yLB + (a.yRB - a.yLB) * static_cast<float>(st.uniform());
        eta_s = rap - (asinhf(lz / mT) - eta);
        E = mT * coshf(rap);
        pzl = mT * sinhf(rap);
    } else {
        eta_s = eta;
        E = sqrtf(mT * mT + lz * lz);
        pzl = lz;
    }
    if (a.out_cap && outp >= a.out_cap) {  // batched rounds: beyond the reserved output
        atomicAdd(a.stats + 7, 1ull);
        return;
    }
    Hadron h;
    h.pid = pid;
    h.mass = fm;
    h.E = E; h.px = lx; h.py = ly; h.pz = pzl;
    h.t = tau * coshf(eta_s);
    h.x = a.cells[F_X * Nc + c];
    h.y = a.cells[F_Y * Nc + c];
    h.z = tau * sinhf(eta_s);
    if (a.boost_invariant && partner) {
        h.t = t_par;
        h.z = z_par;
    }
    static_cast<Hadron*>(a.out)[outp] = h;
}

// The per-cell part of the lab-frame record on 3+1D surfaces: the flow and
// the position (t, x, y, z) = (tau cosh eta, x, y, tau sinh eta), once per
// surface with the operations of emit_lab, so that a trial lane loads two
// 16-byte words when it takes a hadron instead of gathering eight fields
// (and evaluating cosh and sinh) at every acceptance.
__global__ void k_emit_cells(const float* __restrict__ cells, int64_t Nc, float4* __restrict__ ec) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < Nc;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float tau = cells[F_TAU * Nc + c], eta = cells[F_ETA * Nc + c];
        ec[2 * c] = make_float4(cells[F_U0 * Nc + c], cells[F_U1 * Nc + c], cells[F_U2 * Nc + c],
                                cells[F_U3 * Nc + c]);
        ec[2 * c + 1] = make_float4(tau * coshf(eta), cells[F_X * Nc + c], cells[F_Y * Nc + c], tau * sinhf(eta));
    }
}

// emit_lab on a 3+1D surface from the cell's words eu = u, ep = (t, x, y, z):
// the same arithmetic; the 40-byte record as five 8-byte stores
__device__ __forceinline__ void emit_cell_lab(const SampleArgs& a, float4 eu, float4 ep, int pid, float fm, float px,
                                              float py, float pz, float p0, uint64_t outp) {
    const float pdu = px * eu.y + py * eu.z + pz * eu.w;
    const float k1 = pdu / (eu.x + 1.f) + p0;
    const float lx = px + k1 * eu.y, ly = py + k1 * eu.z, lz = pz + k1 * eu.w;
    const float mT = sqrtf(lx * lx + ly * ly + fm * fm);
    const float E = sqrtf(mT * mT + lz * lz);
    if (a.out_cap && outp >= a.out_cap) {  // batched rounds: beyond the reserved output
        atomicAdd(a.stats + 7, 1ull);
        return;
    }
    uint2* dst = reinterpret_cast<uint2*>(static_cast<Hadron*>(a.out) + outp);
    dst[0] = make_uint2(static_cast<uint32_t>(pid), __float_as_uint(fm));
    dst[1] = make_uint2(__float_as_uint(E), __float_as_uint(lx));
    dst[2] = make_uint2(__float_as_uint(ly), __float_as_uint(lz));
    dst[3] = make_uint2(__float_as_uint(ep.x), __float_as_uint(ep.y));
    dst[4] = make_uint2(__float_as_uint(ep.z), __float_as_uint(ep.w));
}

// CDF_0, CDF_1, CDF_2 of one |p| table at E~ together: the per-term
// exponentials become powers of exp(m0 - E~) and exp(-m0), so one hadron's
// normalisation costs two exponentials and a logarithm.
// 1/n, 1/n^2 and 1/n^3 of the series terms n = 1..kMaxTrunc: folded
// constants are correctly rounded, so s / n^k (s = +-1, n^k exact) equals
// s * kInvPow[k-1][n-1] bit for bit, without FP64 divisions
constexpr int kMaxTrunc = 10;
__constant__ double kInvPow[3][kMaxTrunc] = {
    {1. / 1, 1. / 2, 1. / 3, 1. / 4, 1. / 5, 1. / 6, 1. / 7, 1. / 8, 1. / 9, 1. / 10},
    {1. / 1, 1. / 4, 1. / 9, 1. / 16, 1. / 25, 1. / 36, 1. / 49, 1. / 64, 1. / 81, 1. / 100},
    {1. / 1, 1. / 8, 1. / 27, 1. / 64, 1. / 125, 1. / 216, 1. / 343, 1. / 512, 1. / 729, 1. / 1000}};

__device__ void pcdf_all(const PTabDev& pt, int tb, double E, double out[3]) {
    const int fer = pt.fermion[tb], trunc = min(pt.trunc[tb], kMaxTrunc), quirk = pt.quirk;
    const double m = pt.m0[tb];
    const double r = exp(m - E), em1 = pt.em1[tb];
    double c0 = 0., c1 = 0., c2 = 0., sign = 1., em = 1., ee = r;
    for (int n = 0; n < trunc; n++) {
        const int n1 = n + 1;
        const double dn = static_cast<double>(n1);
        // (Fermi-Dirac: the integer quotient sign / n of the series, which is
        // sign for n = 1 and 0 beyond, as a select)
        if (trunc <= 5) c0 += (fer ? (n == 0 ? sign : 0.) : kInvPow[0][n]) * em * (1. - ee);
        const double s1 = fer ? sign : 1.;
        c1 += s1 * kInvPow[1][n] * em * ((m * dn + 1) - ee * (E * dn + 1));
        const double s2 = (fer && !quirk) ? sign : 1.;
        c2 += s2 * kInvPow[2][n] * em * ((m * dn * (m * dn + 2) + 2) - ee * (E * dn * (E * dn + 2) + 2));
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

// Per-hadron set-up record, written by k_setup, read once by k_trials.
struct alignas(16) HSetup {
    uint32_t cell;
    uint32_t j;       // index of the hadron within its cell (RNG counter)
    uint16_t species;
    uint8_t tb;       // |p| table
    uint8_t flags;    // 1: m/T - mu/T outside the table
    int32_t i0;       // first table row at or below m~ - mu~
    double mterm;     // proposal CDF at E~ = m~ - mu~
    uint64_t out;     // output slot (event-ordered, through the permutation)
};
static_assert(sizeof(HSetup) == 32, "HSetup layout");

// hadron-independent quantities of one |p| proposal, rebuilt from the record
struct PState {
    double mut, aa, x, mu, Td;
    float T, fm, fmu, fsg;
};

__device__ __forceinline__ void pstate_of(const SampleArgs& a, int64_t c, int s, PState& q) {
    const int64_t Nc = a.Nc;
    q.T = a.cells[F_T * Nc + c];
    const float muB = a.cells[F_MUB * Nc + c], muS = a.cells[F_MUS * Nc + c], muQ = a.cells[F_MUQ * Nc + c];
    const double mass = a.sp.mass[s];
    const float Bq = a.sp.B[s], Sq = a.sp.S[s], Qq = a.sp.Q[s];
    q.mu = fmin(mass, static_cast<double>(Bq * muB + Sq * muS + Qq * muQ));
    q.Td = q.T;
    const double mt = mass / q.Td;
    q.mut = q.mu / q.Td;
    q.x = mt - q.mut;
    q.aa = q.mut * q.mut - mt * mt / 2.;
    q.fm = static_cast<float>(mass);
    q.fmu = static_cast<float>(q.mu);
    q.fsg = static_cast<float>(a.sp.sign[s]);
}


// Species of every hadron slot of a pass, one wave per cell.  By Poisson
// splitting a hadron of cell c is of species s with probability
// N_{c,s} / N_c.  The cell's yield row (k_yield_rows, working precision)
// becomes its FP64 inclusive species prefix in LDS: lane l sums the G =
// ceil(Ns / 64) consecutive species from l G in order, one wave scan of the
// lanes' totals gives each lane its base, and a running maximum of the lanes'
// last values keeps the prefix non-decreasing across lanes (the scan's tree
// of roundings could otherwise step down where a lane's yields are zero).
// Within a lane the prefix of a zero-yield species equals its predecessor's.
// Slot j draws u from its stream (c, j, TAG_SETUP)
// and takes the first species whose prefix exceeds u times the cell total; a
// species of zero yield is never taken (the nearest species of positive
// yield replaces one that the roundings of the scan leave an interval of
// measure ~1e-16).  Writes each slot's species and cell; with PAIRS also the
// cell's pair bitmap and its number of pairs (the learned sampler's (cell,
// species) pairs, count_pairs).  Launch with ceil(Nc / kPickCells) workgroups
// of kPickCells waves and pick_lds_bytes(Ns, PAIRS) bytes of dynamic LDS (the
// prefix rows and hit counts, sized to the species table so that the
// latency-bound kernel keeps its full occupancy).
constexpr int kPickCells = 4;
constexpr int kPickWords = kMaxSpecies / 32;
constexpr int kPickPerLane = kMaxSpecies / kWave;
__host__ __device__ constexpr size_t pick_lds_bytes(int Ns, bool pairs) {
    return static_cast<size_t>(kPickCells) * Ns * (sizeof(double) + (pairs ? sizeof(uint32_t) : 0));
}

// The learned sampler's pairs of one cell from its slots' species, counted
// per species in LDS (`hits`) as the slots are drawn.  A pair is a species of
// at least `thr` slots; with scnt (routing, thr > 1) the slots of the other
// species are counted in scnt[c]: they take the plain sampler, whose
// per-hadron set-up costs less than the network and envelope of a pair of
// few hadrons.  Writes the pair bitmap (Wb words, gathered in the wave's
// LDS words `wbuf` and stored at once) and the number of pairs.
__device__ __forceinline__ void count_pairs(const uint32_t* hits, int Ns, int thr, int64_t c, int Wb,
                                            uint32_t* wbuf, uint32_t* __restrict__ gbits,
                                            uint32_t* __restrict__ cnt, uint32_t* __restrict__ scnt) {
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t np = 0, ns = 0;
    for (int s0 = 0; s0 < Ns; s0 += kWave) {
        const int s = s0 + lane;
        const uint32_t k = s < Ns ? hits[s] : 0u;
        const bool pair = k >= static_cast<uint32_t>(thr);
        if (k > 0 && !pair) ns += k;
        const unsigned long long m = __ballot(pair);
        if (lane == 0) {
            wbuf[s0 >> 5] = static_cast<uint32_t>(m);
            if ((s0 >> 5) + 1 < Wb) wbuf[(s0 >> 5) + 1] = static_cast<uint32_t>(m >> 32);
        }
        np += static_cast<uint32_t>(__popcll(m));
    }
    wave_lds_sync();
    if (lane < Wb) gbits[c * Wb + lane] = wbuf[lane];
    if (lane == 0) cnt[c] = np;
    if (scnt) {
        ns = wave_sum(ns);
        if (lane == 0) scnt[c] = ns;
    }
}

template <typename R, bool PAIRS>
__global__ __launch_bounds__(kWave * kPickCells) void k_pick(SampleArgs a, const R* __restrict__ rows, int stride,
                                                             uint16_t* __restrict__ sp16, uint32_t* __restrict__ cellidx,
                                                             uint32_t* __restrict__ gbits, uint32_t* __restrict__ cnt,
                                                             uint32_t* __restrict__ scnt, int thr) {
    extern __shared__ double s_pick[];  // [kPickCells][Ns] prefix, then (PAIRS) [kPickCells][Ns] hits
    __shared__ uint32_t s_pos[kPickCells][kPickWords];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kPickCells + w;
    if (c >= a.Nc) return;  // whole waves; no workgroup barrier below
    // (batched rounds: the round index lives on the device)
    const uint32_t chunk = a.chunk_dev ? *a.chunk_dev : a.chunk;
    const int Ns = a.Ns, Wb = (Ns + 31) / 32;
    const int64_t o0 = a.off[c];
    int64_t n = a.off[c + 1] - o0;
    // (batched rounds: only the slots with set-up records)
    if (a.setup_cap > 0) n = max<int64_t>(0, min<int64_t>(n, a.setup_cap - o0));
    if (n == 0) {
        if (PAIRS) {
            if (lane < Wb) gbits[c * Wb + lane] = 0u;
            if (lane == 0) cnt[c] = 0u;
            if (scnt && lane == 0) scnt[c] = 0u;
        }
        return;
    }
    double* pre = s_pick + w * Ns;
    uint32_t* pos = s_pos[w];
    uint32_t* hits = reinterpret_cast<uint32_t*>(s_pick + kPickCells * Ns) + w * Ns;
    if (PAIRS)
        for (int s = lane; s < Ns; s += kWave) hits[s] = 0u;
    const R* row = rows + c * stride;
    // the positive-yield bitmap, set bit by bit below
    if (lane < Wb) pos[lane] = 0u;
    wave_lds_sync();
    const int G = (Ns + kWave - 1) / kWave, sl = lane * G;
    // (the loads issued together: indices clamped into the row, values
    // beyond the lane's species zeroed)
    R yr[kPickPerLane];
#pragma unroll
    for (int k = 0; k < kPickPerLane; k++) yr[k] = row[min(sl + k, Ns - 1)];
    double v[kPickPerLane];
    double acc = 0.;
    uint32_t mb = 0u;  // bit k: species sl + k has a positive yield
#pragma unroll
    for (int k = 0; k < kPickPerLane; k++) {
        const double y = (k < G && sl + k < Ns) ? static_cast<double>(yr[k]) : 0.;
        acc += y;
        v[k] = acc;
        mb |= (y > 0.) ? (1u << k) : 0u;
    }
    // the lane's bits span at most two bitmap words (G <= 8); a word past
    // the species is only touched with bits of valid species
    const int b0 = sl & 31;
    atomicOr(&pos[sl >> 5], mb << b0);
    const uint32_t mhi = b0 ? (mb >> (32 - b0)) : 0u;
    if (mhi) atomicOr(&pos[(sl >> 5) + 1], mhi);
    // the earlier lanes' sum (DPP scan), and the running maximum of the
    // earlier lanes' last values
    double base = __shfl_up(wave_scan_dpp(acc), 1, kWave);
    if (lane == 0) base = 0.;
    double lmax = __shfl_up(wave_max_scan_dpp(base + acc), 1, kWave);
    if (lane == 0) lmax = 0.;
#pragma unroll
    for (int k = 0; k < kPickPerLane; k++) {
        const int s = sl + k;
        if (k < G && s < Ns) pre[s] = fmax(base + v[k], lmax);
    }
    wave_lds_sync();
    const double tot = pre[Ns - 1];
    auto positive = [&](int s) { return (pos[s >> 5] >> (s & 31)) & 1u; };
    for (int64_t r = lane; r < n; r += kWave) {
        Stream st(a.seed, static_cast<uint32_t>(c), static_cast<uint32_t>(r), tag_word(TAG_SETUP, chunk));
        int sp = upper_bound(pre, 0, Ns, st.uniform() * tot);
        if (sp >= Ns) sp = Ns - 1;
        if (!positive(sp)) {
            int t = sp + 1;
            while (t < Ns && !positive(t)) t++;
            if (t >= Ns) {
                t = sp - 1;
                while (t > 0 && !positive(t)) t--;
            }
            sp = max(t, 0);
        }
        sp16[o0 + r] = static_cast<uint16_t>(sp);
        cellidx[o0 + r] = static_cast<uint32_t>(c);
        if (PAIRS) atomicAdd(&hits[sp], 1u);
    }
    if (PAIRS) {
        wave_lds_sync();
        count_pairs(hits, Ns, thr, c, Wb, pos, gbits, cnt, scnt);  // (pos is free after the slots)
    }
}

// an empty record (out of trials, or outside the momentum tables); batched
// rounds count a slot beyond the reserved output instead of writing it
__device__ __forceinline__ void write_hole(const SampleArgs& a, uint64_t outp) {
    if (a.out_cap && outp >= a.out_cap) {
        atomicAdd(a.stats + 7, 1ull);
        return;
    }
    Hadron h{};
    static_cast<Hadron*>(a.out)[outp] = h;
}

// Batched rounds: the output slot of the round's hadron i (cell c), within
// its surface's block of this round, after the hadrons of the earlier rounds
__device__ __forceinline__ uint64_t batch_out(const SampleArgs& a, int64_t i, int64_t c) {
    const int k = upper_bound(a.surf_off + 1, 0, a.S, c);
    const int64_t h0 = a.off[a.surf_off[k]];
    return *a.out_base + static_cast<uint64_t>(h0) + a.perms[k](static_cast<uint64_t>(i - h0));
}

// the species-dependent part of a set-up record: the |p| table, its first
// row and the proposal's normalisation at E~ = m~ - mu~
__device__ __forceinline__ HSetup setup_record(const SampleArgs& a, int64_t c, uint32_t jj, int s) {
    PState q;
    pstate_of(a, c, s, q);
    const int sign = a.sp.sign[s];
    const int tb = ((sign == -1) ? 0 : 3) + ((q.x < 30.) ? 0 : (q.x < 50. ? 1 : 2));
    double cm[3];
    pcdf_all(a.pt, tb, q.x, cm);
    HSetup r;
    r.cell = static_cast<uint32_t>(c);
    r.j = jj;
    r.species = static_cast<uint16_t>(s);
    r.tb = static_cast<uint8_t>(tb);
    r.i0 = static_cast<int32_t>((q.x - a.pt.Et0[tb]) / a.pt.dEt[tb]);
    r.flags = (r.i0 < 0 || r.i0 >= a.pt.len[tb] - 1) ? 1 : 0;
    r.mterm = cm[2] + 2. * q.mut * cm[1] + q.aa * cm[0];
    r.out = 0;
    return r;
}

// Set-up of every hadron slot, one thread per slot (no divergence): the cell
// (offsets), the species (the cell's FP64 species CDF), the |p| table and its
// normalisation at E~ = m~ - mu~, and the output slot.
// With a slot list, record r describes slot list[r] for r < *nlist (the
// learned sampler's fallback hadrons); without one, record i describes slot
// i.  Grid-stride, one thread per record.
// BATCH: slots of batched surfaces (per-surface permutations, output base of
// the round); instantiated separately so that the one-surface kernel keeps
// its registers.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_setup(SampleArgs a, HSetup* __restrict__ rec,
                                               const int64_t* __restrict__ list,
                                               const unsigned long long* __restrict__ nlist) {
    // (the pass index of batched rounds is read here; the kernel argument
    // block itself is never written, which would copy it to scratch)
    const uint32_t chunk = a.chunk_dev ? *a.chunk_dev : a.chunk;
    int64_t n = list ? static_cast<int64_t>(*nlist) : (a.Kdev ? static_cast<int64_t>(*a.Kdev) : a.K);
    if (a.setup_cap > 0 && n > a.setup_cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.stats + 6, 1ull);
        n = a.setup_cap;
    }
    for (int64_t ri = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; ri < n;
         ri += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = list ? list[ri] : ri;
    const int64_t c = a.cellidx ? static_cast<int64_t>(a.cellidx[i]) : upper_bound(a.off + 1, int64_t(0), a.Nc, i);
    const uint32_t jj = static_cast<uint32_t>(i - a.off[c]);
    int s;
    if (a.slot_species) {
        s = a.slot_species[i];
    } else {
        Stream st(a.seed, static_cast<uint32_t>(c), jj, tag_word(TAG_SETUP, chunk));
        const double* row = a.cdf + c * cdf_stride(a.Ns);
        s = upper_bound(row, 0, a.Ns, st.uniform() * a.ctot[c]);
        if (s >= a.Ns) s = a.Ns - 1;
    }
    HSetup r = setup_record(a, c, jj, s);
    r.out = BATCH ? batch_out(a, i, c) : a.perm(static_cast<uint64_t>(i));
    rec[ri] = r;
    }
}

// Persistent trial kernel with per-lane refill.  Each wavefront owns a
// contiguous range of hadron slots; a lane that finishes its hadron
// (accepted, or out of trials) loads the next set-up record of the range, so
// the geometric tails of the rejection loops do not idle the other lanes.
// The momentum follows the reference's nested rejection
// (src/FSSW.cpp:1852-1966, src/MomentumSamplerBase.cpp:20-58) flattened into
// one loop: an iteration is one |p| proposal plus, when the |p| test passes,
// one outer trial.  Iteration `it` of hadron (c, j) uses Philox block
// (c, j, tag, it): word 0 is the 32-bit CDF draw, the other 96 bits give four
// 24-bit uniforms.  Streams depend on (c, j) only, so the output does not
// depend on the launch shape.
// idle lanes that trigger a refill, unless the autotuner chose otherwise
constexpr int kRefillMin = 16;
// The two |p| tables below t = 30 (Bose-Einstein 0, Fermi-Dirac 3), which
// nearly every hadron uses, are copied into LDS (e0, e1, e2 per row, 96 KB)
// by each workgroup, one workgroup of kTrialThreads per CU (4 waves per SIMD
// at the kernel's registers): the bracket search runs in LDS, without
// dependent global loads.  (The search over LDS rows g kCoarse narrows the
// bracket; with every row there, kCoarse = 1, it finds it.)  Other tables
// are searched in global memory.
constexpr int kTrialThreads = 1024;
constexpr int kCoarse = 1;
constexpr int kCoarseRows = (kPTabLen + kCoarse - 1) / kCoarse;

__device__ __forceinline__ void trials_body(const SampleArgs& a, const HSetup* __restrict__ rec) {
    __shared__ unsigned long long s_stats[4];
    __shared__ double s_ctab[2][kCoarseRows * 3];
    if (threadIdx.x < 4) s_stats[threadIdx.x] = 0ull;
    for (int i = threadIdx.x; i < 2 * kCoarseRows * 3; i += blockDim.x) {
        const int t = i / (kCoarseRows * 3), g = (i / 3) % kCoarseRows, k = i % 3;
        const int tbl = t == 0 ? 0 : 3, row = g * kCoarse;
        s_ctab[t][g * 3 + k] = row < a.pt.len[tbl] ? a.pt.C[(static_cast<int64_t>(tbl) * kPTabLen + row) * 4 + k] : 0.;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    const int64_t wid = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    const uint32_t chunk = a.chunk_dev ? *a.chunk_dev : a.chunk;
    int64_t Kn = a.Kdev ? static_cast<int64_t>(*a.Kdev) : a.K;
    if (a.setup_cap > 0 && Kn > a.setup_cap) Kn = a.setup_cap;
    const int64_t base = Kn / nwaves, rem = Kn % nwaves;
    const int64_t beg = wid * base + min(wid, rem);
    const int64_t end = beg + base + (wid < rem ? 1 : 0);
    const int64_t Nc = a.Nc;
    const PTabDev& pt = a.pt;
    const uint32_t key0 = static_cast<uint32_t>(a.seed), key1 = static_cast<uint32_t>(a.seed >> 32);
    const uint32_t tw = tag_word(TAG_TRIAL, chunk);
    const int max_it = a.max_trials * 64;

    int64_t next = beg;
    // per-lane counters (32 bits: a lane's share of one launch), summed in 64
    uint32_t n_trial = 0, n_inner = 0, n_fail = 0, n_range = 0;

    bool active = false;
    int64_t c = 0;
    uint64_t outp = 0;
    uint32_t jj = 0;
    int s = 0, tb = 0, i0 = 0, imax = 0, it = 0, trials = 1;
    double mterm = 0., cmax = 0., rmin0 = 0.;
    PState q{};
    float dsfac = 1.f, rdsfac = 1.f, rT = 1.f, dfpref = 0.f, d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f, Pi = 0.f;
    float Bq = 0.f, Sq = 0.f, Qq = 0.f;
    float pixx = 0.f, pixy = 0.f, pixz = 0.f, piyy = 0.f, piyz = 0.f;
    float cf0 = 0.f, cf1 = 0.f, cf2 = 0.f, cf3 = 0.f, cf4 = 0.f, cf5 = 0.f;
    float qx = 0.f, qy = 0.f, qz = 0.f, pqm = 0.f;  // diffusion: q / D_qmu, n_B / (e + P)
    float4 eu = make_float4(0.f, 0.f, 0.f, 0.f), ep = eu;  // the cell's flow and position (3+1D)
    int pid = 0;

    while (true) {
        const unsigned long long need = __ballot(!active);
        if (need && next < end && (__popcll(need) >= a.refill_min || need == ~0ull)) {
            const int64_t navail = end - next;
            const int rank = __popcll(need & ((1ull << lane) - 1));
            if (!active && rank < navail) {
                const HSetup r = rec[next + rank];
                c = r.cell;
                jj = r.j;
                s = r.species;
                tb = r.tb;
                i0 = r.i0;
                mterm = r.mterm;
                outp = r.out;
                pstate_of(a, c, s, q);
                if (a.ecell) {
                    eu = a.ecell[2 * c];
                    ep = a.ecell[2 * c + 1];
                    pid = a.sp.pid[s];
                }
                d0 = a.cells[F_DA0 * Nc + c]; d1 = a.cells[F_DA1 * Nc + c];
                d2 = a.cells[F_DA2 * Nc + c]; d3 = a.cells[F_DA3 * Nc + c];
                const float Ed = a.cells[F_E * Nc + c], Pd = a.cells[F_P * Nc + c];
                Bq = a.sp.B[s]; Sq = a.sp.S[s]; Qq = a.sp.Q[s];
                dsfac = fabsf(d0) + sqrtf(d1 * d1 + d2 * d2 + d3 * d3);
                rdsfac = 1.0f / dsfac;
                rT = 1.0f / q.T;
                dfpref = 1.0f / (2.0f * q.T * q.T * (Ed + Pd));
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
                if (a.diff) {
                    const float iq = static_cast<float>(a.coef[6 * Nc + c]);
                    qx = a.cells[F_QX * Nc + c] * iq;
                    qy = a.cells[F_QY * Nc + c] * iq;
                    qz = a.cells[F_QZ * Nc + c] * iq;
                    pqm = static_cast<float>(a.coef[7 * Nc + c]);
                }
                it = 0;
                trials = 1;
                imax = pt.len[tb] - 1;
                if (r.flags & 1) {
                    // m/T - mu/T outside the table: the reference exits
                    // (src/MomentumSamplerBase.cpp:33-43); leave a hole
                    n_range++;
                    write_hole(a, outp);
                } else {
                    const double* e0 = pt.C + (static_cast<int64_t>(tb) * kPTabLen + i0) * 4;
                    const double* em = pt.C + (static_cast<int64_t>(tb) * kPTabLen + imax) * 4;
                    rmin0 = e0[2] + 2. * q.mut * e0[1] + q.aa * e0[0] - mterm;
                    cmax = em[2] + 2. * q.mut * em[1] + q.aa * em[0] - mterm;
                    active = true;
                }
            }
            next += min<int64_t>(__popcll(need), navail);
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
        // search issues its three loads together
        int lo = i0, hi = imax;
        double rlo = rmin0, rhi = cmax;
        auto comb_of = [&](double e0, double e1, double e2) { return e2 + 2. * q.mut * e1 + q.aa * e0 - mterm; };
        auto comb = [&](int qq) {
            const double* e = C + 4 * qq;
            return comb_of(e[0], e[1], e[2]);
        };
        if (tb == 0 || tb == 3) {
            // LDS rows g kCoarse strictly inside (lo, hi): g in [ga, gb]; find
            // L < H with comb(L kCoarse) <= r < comb(H kCoarse), L = ga - 1
            // and H = gb + 1 standing for lo and hi
            const double* Cc = s_ctab[tb == 0 ? 0 : 1];
            auto combc = [&](int g) { return comb_of(Cc[3 * g], Cc[3 * g + 1], Cc[3 * g + 2]); };
            const int ga = lo / kCoarse + 1, gb = (hi - 1) / kCoarse;
            int L = ga - 1, H = gb + 1;
            double vl = rlo, vh = rhi;
            while (H - L > 4) {
                const int qd = (H - L) >> 2;
                const int m1 = L + qd, m2 = L + 2 * qd, m3 = L + 3 * qd;
                const double v1 = combc(m1), v2 = combc(m2), v3 = combc(m3);
                if (r < v2) {
                    if (r < v1) { H = m1; vh = v1; } else { L = m1; vl = v1; H = m2; vh = v2; }
                } else {
                    if (r < v3) { L = m2; vl = v2; H = m3; vh = v3; } else { L = m3; vl = v3; }
                }
            }
            while (H - L > 1) {
                const int mid = (H + L) >> 1;
                const double vm = combc(mid);
                if (r < vm) { H = mid; vh = vm; } else { L = mid; vl = vm; }
            }
            if (L >= ga) { lo = L * kCoarse; rlo = vl; }
            if (H <= gb) { hi = H * kCoarse; rhi = vh; }
        }
        while (hi - lo > 3) {
            const int qd = (hi - lo) >> 2;
            const int m1 = lo + qd, m2 = lo + 2 * qd, m3 = lo + 3 * qd;
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
        if (E0 < q.x) { E0 = q.x; rlo = 0.; }
        const double Ehi = Et0 + hi * dEt;
        const double Etl = E0 + (Ehi - E0) / fmax(1e-16, rhi - rlo) * (r - rlo);
        const float fm = q.fm;
        const float Ep = static_cast<float>(q.Td * Etl + q.mu);
        // (hardware square root and reciprocals: 1 ulp)
        const float pm = __builtin_amdgcn_sqrtf(fmaxf(Ep * Ep - fm * fm, 0.f));
        const float rEp = __builtin_amdgcn_rcpf(Ep);
        const float ratio = (pm * rEp) * __builtin_amdgcn_rcpf(1.f - fm * fm * (0.5f * rEp * rEp));
        bool done = false;
        if (u24(w.y) <= ratio) {
            n_trial++;
            const float T = q.T;
            const float phi = 2.f * static_cast<float>(kPi) * u24(w.z);
            const float cth = 2.f * u24(w.w) - 1.f;
            const float sth = __builtin_amdgcn_sqrtf(fmaxf(1.f - cth * cth, 0.f));
            float sphi, cphi;
            __sincosf(phi, &sphi, &cphi);
            const float pTc = pm * sth;
            const float px = pTc * cphi, py = pTc * sphi, pz = pm * cth;
            const float p0 = __builtin_amdgcn_sqrtf(fm * fm + pm * pm);
            const float rp0 = __builtin_amdgcn_rcpf(p0);
            const float pds = p0 * d0 + px * d1 + py * d2 + pz * d3;
            const float f0 = __builtin_amdgcn_rcpf(__expf((p0 - q.fmu) * rT) + q.fsg);
            const float omf = 1.f - q.fsg * f0;
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
            float dfq = 0.f;
            if (a.diff) {
                // (src/FSSW.cpp:1931-1937)
                dfq = omf * (pqm - Bq * rp0) * (-(px * qx + py * qy + pz * qz));
            }
            const float w1 = fminf(fmaxf(pds * rp0 * rdsfac, 0.f), 1.f);
            const float w2 = fminf(fmaxf((1.f + dfs + dfb + dfq) * 0.5f, 0.f), 1.f);
            const uint32_t u4 = (w.y & 0xFFu) | ((w.z & 0xFFu) << 8) | ((w.w & 0xFFu) << 16);
            const float ua = (static_cast<float>(u4) + 0.5f) * (1.0f / 16777216.0f);
            if (ua < w1 * w2) {
                if (a.ecell) emit_cell_lab(a, eu, ep, pid, fm, px, py, pz, p0, outp);
                else emit_lab(a, chunk, c, jj, a.sp.pid[s], fm, px, py, pz, p0, outp);
                done = true;
            } else {
                trials++;
            }
        }
        if (!done && (trials >= a.max_trials || it >= max_it)) {
            n_fail++;
            write_hole(a, outp);
            done = true;
        }
        if (done) active = false;
    }
    const unsigned long long t_trial = wave_sum(static_cast<unsigned long long>(n_trial));
    const unsigned long long t_inner = wave_sum(static_cast<unsigned long long>(n_inner));
    const unsigned long long t_fail = wave_sum(static_cast<unsigned long long>(n_fail));
    const unsigned long long t_range = wave_sum(static_cast<unsigned long long>(n_range));
    if (lane == 0) {
        atomicAdd(&s_stats[0], t_trial);
        atomicAdd(&s_stats[1], t_inner);
        atomicAdd(&s_stats[2], t_fail);
        atomicAdd(&s_stats[3], t_range);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_stats[threadIdx.x]) atomicAdd(&a.stats[threadIdx.x], s_stats[threadIdx.x]);
}

__global__ __launch_bounds__(kTrialThreads) void k_trials(SampleArgs a, const HSetup* __restrict__ rec) {
    trials_body(a, rec);
}

// The same trials for the hadrons a learned pass hands to the plain sampler
// (routed pairs and hadrons without an envelope): its own symbol, so that
// profiles and timings tell it from the plain pass's launch.
__global__ __launch_bounds__(kTrialThreads) void k_trials_routed(SampleArgs a, const HSetup* __restrict__ rec) {
    trials_body(a, rec);
}

// ============================================= local charge conservation
// FSSW's local_charge_conservation (src/FSSW.cpp:931-938, 1035-1048): the
// yields of negative species are zero, and every positive hadron is followed
// in its event by its antiparticle, sampled from the same cell with the
// negated charges (the antiparticle's species entry).  After the main pass:
//   k_lcc_flag    per slot, its species (the same draw as k_setup) and
//                 whether it is positive, by output position
//   scan          partner offsets g (exclusive) over output positions
//   k_lcc_place   hadron o moves to o + g[o]; a positive one leaves the slot
//                 after it to its partner, whose set-up record is built here
//   k_trials      samples the partners (j carries kPartnerBit: own streams)
//   k_lcc_events  event offsets e -> evoff[e] + g[evoff[e]]
__global__ void k_lcc_flag(SampleArgs a, uint32_t* __restrict__ pflag, int64_t* __restrict__ src) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.K) return;
    const int64_t c = a.cellidx ? static_cast<int64_t>(a.cellidx[i]) : upper_bound(a.off + 1, int64_t(0), a.Nc, i);
    const uint32_t jj = static_cast<uint32_t>(i - a.off[c]);
    int s;
    if (a.slot_species) {
        s = a.slot_species[i];
    } else {
        Stream st(a.seed, static_cast<uint32_t>(c), jj, tag_word(TAG_SETUP, a.chunk));
        const double* row = a.cdf + c * cdf_stride(a.Ns);
        s = upper_bound(row, 0, a.Ns, st.uniform() * a.ctot[c]);
        if (s >= a.Ns) s = a.Ns - 1;
    }
    const uint64_t o = a.perm(static_cast<uint64_t>(i));
    pflag[o] = (a.sp.Q[s] > 0) ? 1u : 0u;
    src[o] = (i << 10) | s;
}

// the same flags from the set-up records of the multiplicity models, whose
// output slot is the record index
__global__ void k_lcc_flag_rec(const HSetup* __restrict__ rec, int64_t K, const int* __restrict__ Q,
                               uint32_t* __restrict__ pflag, int64_t* __restrict__ src) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const int s = rec[i].species;
    pflag[i] = (Q[s] > 0) ? 1u : 0u;
    src[i] = (i << 10) | s;
}

// mrec: the set-up records the slots were sampled from (multiplicity
// models), or null for the Poisson pass (cell from the offsets)
__global__ void k_lcc_place(SampleArgs a, const uint32_t* __restrict__ pflag, const int64_t* __restrict__ g,
                            const int64_t* __restrict__ src, const Hadron* __restrict__ old_out,
                            Hadron* __restrict__ new_out, HSetup* __restrict__ prec,
                            const HSetup* __restrict__ mrec) {
    const int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= a.K) return;
    const int64_t n = o + g[o];
    new_out[n] = old_out[o];
    if (!pflag[o]) return;
    const int64_t i = src[o] >> 10;
    const int s = static_cast<int>(src[o] & 1023);
    int64_t c;
    uint32_t jj;
    if (mrec) {
        c = mrec[i].cell;
        jj = mrec[i].j | kPartnerBit;
    } else {
        c = upper_bound(a.off + 1, int64_t(0), a.Nc, i);
        jj = static_cast<uint32_t>(i - a.off[c]) | kPartnerBit;
    }
    HSetup r = setup_record(a, c, jj, a.anti[s]);
    r.out = static_cast<uint64_t>(n + 1);
    prec[g[o]] = r;
}

__global__ void k_lcc_events(const int64_t* __restrict__ evoff, int E, const int64_t* __restrict__ g,
                             int64_t* __restrict__ evoff_new) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    evoff_new[e] = evoff[e] + g[evoff[e]];
}

// ====================================== multiplicity models 1, 10 and 20
// FSSW's per-(species, event) multiplicity (src/FSSW.cpp:250-309) for the
// models other than Poisson, whose per-cell splitting the main pass uses:
// the reference's order species -> events -> hadrons -> cell pick
// (src/FSSW.cpp:920-979) becomes
//   k_model_counts   per (event, species): N from the model
//   scan             slots in (event, species) order: events are contiguous
//                    and, within an event, species follow the sampling order
//   k_model_fill     each slot's species
//   k_model_setup    each slot's cell from its species' FP64 cell prefix
//                    (the species-major yields, scanned along the cells) and
//                    its set-up record; the trial kernel follows
constexpr uint32_t TAG_MODEL = 10, TAG_MODEL_CELL = 11;

__global__ void k_model_counts(const double* __restrict__ stot, int Ns, int E, double mult, int model,
                               double para1, uint64_t seed, uint32_t chunk, uint32_t* __restrict__ cnt) {
    const int64_t it = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (it >= static_cast<int64_t>(E) * Ns) return;
    const int e = static_cast<int>(it / Ns), s = static_cast<int>(it % Ns);
    Stream st(seed, static_cast<uint32_t>(s), static_cast<uint32_t>(e), tag_word(TAG_MODEL, chunk));
    const double dN = stot[s] * mult;
    const double fl = floor(dN), frac = dN - fl;
    int64_t n = 0;
    if (model == 1) {
        n = static_cast<int64_t>(fl) + (st.uniform() < frac ? 1 : 0);
    } else if (model == 10) {
        const double k = para1 * frac, p = 1. / (1. + para1);
        n = static_cast<int64_t>(fl) + (k < 1e-15 ? 0 : negative_binomial(p, k, st));
    } else {  // 20
        const double k = para1 * dN, p = 1. / (1. + para1);
        n = k < 1e-15 ? static_cast<int64_t>(fl) : negative_binomial(p, k, st);
    }
    cnt[it] = static_cast<uint32_t>(min<int64_t>(n, 0xFFFFFFFFll));
}

// per (event, species) segment: its slots' species
__global__ void k_model_fill(const int64_t* __restrict__ segoff, int64_t nseg, int Ns, uint16_t* __restrict__ sp) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= nseg) return;
    const uint16_t s = static_cast<uint16_t>(g % Ns);
    for (int64_t i = segoff[g]; i < segoff[g + 1]; i++) sp[i] = s;
}

// event offsets: the first slot of each event's first species
__global__ void k_model_evoff(const int64_t* __restrict__ segoff, int E, int Ns, int64_t* __restrict__ evoff) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    evoff[e] = segoff[static_cast<int64_t>(e) * Ns];
}

// FP64 inclusive prefix of each species' row of cell yields ([Ns][Nc], in
// place): one workgroup per species, tiles in order (deterministic)
__global__ __launch_bounds__(256) void k_row_scan(double* __restrict__ rows, int64_t Nc) {
    __shared__ double wsum[256 / kWave];
    double* row = rows + static_cast<int64_t>(blockIdx.x) * Nc;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    double carry = 0.;
    for (int64_t t0 = 0; t0 < Nc; t0 += 256) {
        const int64_t c = t0 + threadIdx.x;
        const double v = c < Nc ? row[c] : 0.;
        const double incl = wave_inclusive_scan(v);
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        double pre = carry, tot = carry;
        for (int i = 0; i < 256 / kWave; i++) {
            if (i < w) pre += wsum[i];
            tot += wsum[i];
        }
        if (c < Nc) row[c] = pre + incl;
        carry = tot;
        __syncthreads();
    }
}

// per slot: the cell (the slot's species' cell prefix) and the set-up record;
// the slot index is the hadron's identity for its random streams
__global__ void k_model_setup(SampleArgs a, const uint16_t* __restrict__ sp, const double* __restrict__ ccum,
                              HSetup* __restrict__ rec) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.K) return;
    const int s = sp[i];
    const double* row = ccum + static_cast<int64_t>(s) * a.Nc;
    Stream st(a.seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), tag_word(TAG_MODEL_CELL, a.chunk));
    int64_t c = upper_bound(row, int64_t(0), a.Nc, st.uniform() * row[a.Nc - 1]);
    if (c >= a.Nc) c = a.Nc - 1;
    HSetup r = setup_record(a, c, static_cast<uint32_t>(i), s);
    r.out = static_cast<uint64_t>(i);
    rec[i] = r;
}

}  // namespace issdev
