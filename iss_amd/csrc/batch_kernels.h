// Batched surfaces: many freeze-out surfaces of uneven size, each with its
// own number of oversampled events, sampled together.  (one translation
// unit: included by iss_hip.hip)
//
// The surfaces are concatenated into one resident cell record (surface k owns
// cells [surf_off[k], surf_off[k+1])), so the yield pass is one launch over
// all of them.  Sampling runs in rounds: in round r surface k samples
// n_k(r) = clamp(E_k - r R, 0, R) events.  One round is
//   k_batch_counts        Poisson(n_k(r) N_c) hadrons per cell
//   scan                  round-local hadron offsets (all surfaces at once)
//   k_batch_roots         per surface: its hadron block, its permutation
//   k_batch_multi_*       per surface: Multinomial over its n_k(r) events
//   k_setup, k_trials     the one-surface kernels, reading the round, the
//                         hadron count and the output base from the device
//   k_batch_advance       output base += hadrons of the round, round += 1
// Every size the kernels need lives on the device, so the host captures one
// round into a hipGraph and replays it for all rounds (no host round trip,
// no memset inside the graph).  The output is round-major, then surface, then
// event: segment (r, k) starts at seg[2 (r S + k)] and holds seg[2 (r S + k) + 1]
// hadrons; evcount[(r S + k) R + e] is the size of its event e.
//
// The law per surface is the one-surface law: per-cell Poisson counts over
// independent events are split into rounds of independent events.  The random
// streams are keyed by the global cell index and the round.
#pragma once

#include "sample_kernels.h"

namespace issdev {

struct BatchDev {
    const int64_t* surf_off;        // [S + 1]
    const int32_t* events;          // [S] events of each surface
    int S;
    int R;                          // events per surface per round
    uint32_t* round;                // current round (device)
    unsigned long long* out_base;   // hadrons written by earlier rounds (device)
    double ymult;                   // y_RB - y_LB for boost-invariant surfaces, else 1
};

__device__ __forceinline__ int round_events(int E, uint32_t r, int R) {
    const int64_t rem = static_cast<int64_t>(E) - static_cast<int64_t>(r) * R;
    return rem <= 0 ? 0 : static_cast<int>(rem < R ? rem : R);
}

__global__ void k_batch_counts(BatchDev b, const double* __restrict__ ctot, int64_t N, uint64_t seed,
                               uint32_t* __restrict__ counts) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const uint32_t r = *b.round;
    const int k = upper_bound(b.surf_off + 1, 0, b.S, c);
    const int n = round_events(b.events[k], r, b.R);
    int64_t cnt = 0;
    if (n > 0) {
        Stream st(seed, static_cast<uint32_t>(c), static_cast<uint32_t>(c >> 32), tag_word(TAG_COUNT, r));
        cnt = poisson(static_cast<double>(n) * b.ymult * ctot[c], st);
    }
    counts[c] = static_cast<uint32_t>(min<int64_t>(cnt, 0xFFFFFFFFll));
}

// per surface: the root of its event tree (its hadrons this round), the keyed
// permutation of its block and the output segment of (round, surface)
__global__ void k_batch_roots(BatchDev b, const int64_t* __restrict__ off, uint64_t seed,
                              int64_t* __restrict__ tree, Permutation* __restrict__ perms,
                              int64_t* __restrict__ seg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= b.S) return;
    const uint32_t r = *b.round;
    const int64_t h0 = off[b.surf_off[k]], h1 = off[b.surf_off[k + 1]];
    const int64_t H = h1 - h0;
    tree[k] = H;
    Permutation p;
    p.init(static_cast<uint64_t>(H > 0 ? H : 1),
           seed ^ (0x5851F42D4C957F2Dull * (static_cast<uint64_t>(k) + 1)) ^ (static_cast<uint64_t>(r) << 40));
    perms[k] = p;
    const int64_t at = 2 * (static_cast<int64_t>(r) * b.S + k);
    seg[at] = static_cast<int64_t>(*b.out_base) + h0;
    seg[at + 1] = H;
}

// level d of every surface's binomial tree: node (k, j) covers the events
// [j n / 2^d, (j + 1) n / 2^d) of surface k's n = n_k(r)
__global__ void k_batch_multi_level(BatchDev b, const int64_t* __restrict__ parent, int64_t* __restrict__ child,
                                    int d, uint64_t seed) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (static_cast<int64_t>(b.S) << d)) return;
    const int k = static_cast<int>(t >> d);
    const int64_t j = t & ((1ll << d) - 1);
    const uint32_t r = *b.round;
    const int64_t E = round_events(b.events[k], r, b.R);
    const int64_t n = parent[t];
    const int64_t lo = (j * E) >> d, hi = ((j + 1) * E) >> d;
    const int64_t mid = ((2 * j + 1) * E) >> (d + 1);
    int64_t nl = 0;
    if (hi > lo && n > 0) {
        const double p = static_cast<double>(mid - lo) / static_cast<double>(hi - lo);
        Stream st(seed, static_cast<uint32_t>(t), static_cast<uint32_t>(d), tag_word(TAG_MULTI, r));
        nl = binomial(n, p, st);
    }
    child[2 * t] = nl;
    child[2 * t + 1] = n - nl;
}

// leaves (2^D >= R): event sizes of (round, surface)
__global__ void k_batch_multi_leaves(BatchDev b, const int64_t* __restrict__ leaf, int D,
                                     uint32_t* __restrict__ evcount) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (static_cast<int64_t>(b.S) << D)) return;
    const int k = static_cast<int>(t >> D);
    const int64_t j = t & ((1ll << D) - 1);
    const uint32_t r = *b.round;
    const int64_t E = round_events(b.events[k], r, b.R);
    const int64_t lo = (j * E) >> D, hi = ((j + 1) * E) >> D;
    if (hi - lo == 1)
        evcount[(static_cast<int64_t>(r) * b.S + k) * b.R + lo] = static_cast<uint32_t>(leaf[t]);
}

__global__ void k_batch_advance(BatchDev b, const int64_t* __restrict__ off, int64_t N) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *b.out_base += static_cast<unsigned long long>(off[N]);
        *b.round += 1u;
    }
}

// per-surface sums of the per-cell totals (one block per surface, FP64)
__global__ __launch_bounds__(256) void k_batch_lambda(const int64_t* __restrict__ surf_off,
                                                     const double* __restrict__ ctot, double* __restrict__ lam) {
    __shared__ double ws[256 / kWave];
    const int k = blockIdx.x;
    double s = 0.;
    for (int64_t c = surf_off[k] + threadIdx.x; c < surf_off[k + 1]; c += blockDim.x) s += ctot[c];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < 256 / kWave; w++) t += ws[w];
        lam[k] = t;
    }
}

}  // namespace issdev
