// Device helpers for the gfx950 Cooper-Frye sampler: counter-based RNG,
// exact Poisson / binomial variates, a keyed permutation, wave64 scans.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace issdev {

constexpr int kWave = 64;
constexpr double kHbarC = 0.197327053;  // GeV fm (src/data_struct.h:9)
constexpr double kPi = 3.14159265358979323846;

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
// as easy as 1, 2, 3"): 10 rounds of the published multipliers and Weyl keys.
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// uniform in (0, 1) with 24 random bits, exactly representable in float
__device__ __forceinline__ float u24(uint32_t x) {
    return (static_cast<float>(x >> 8) + 0.5f) * (1.0f / 16777216.0f);
}
// uniform in (0, 1) with 32 random bits, in double
__device__ __forceinline__ double u32d(uint32_t x) {
    return (static_cast<double>(x) + 0.5) * (1.0 / 4294967296.0);
}
// uniform in (0, 1) with 53 random bits from two words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    const uint64_t v = (static_cast<uint64_t>(a) << 21) ^ static_cast<uint64_t>(b >> 11);
    return (static_cast<double>(v & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}

// A stream of uniforms for one (key, c0, c1, c2) identity: draw n uses the
// Philox block (c0, c1, c2, n / 4).  Counter-based: any draw of any stream is
// reproducible from its identity alone.
struct Stream {
    uint32_t k0, k1, c0, c1, c2;
    uint32_t blk;
    u32x4 buf;
    int used;
    __device__ Stream(uint64_t seed, uint32_t a, uint32_t b, uint32_t c)
        : k0(static_cast<uint32_t>(seed)), k1(static_cast<uint32_t>(seed >> 32)),
          c0(a), c1(b), c2(c), blk(0), used(4) {}
    __device__ __forceinline__ uint32_t next_u32() {
        if (used == 4) {
            buf = philox4x32(u32x4{c0, c1, c2, blk++}, k0, k1);
            used = 0;
        }
        const uint32_t v = (used == 0) ? buf.x : (used == 1 ? buf.y : (used == 2 ? buf.z : buf.w));
        used++;
        return v;
    }
    __device__ __forceinline__ double uniform() {
        const uint32_t a = next_u32();
        const uint32_t b = next_u32();
        return u53(a, b);
    }
};

// --------------------------------------------------------------- variates
// Poisson(lambda), exact.  lambda < 10: inversion by sequential search;
// otherwise PTRS, the transformed rejection with squeeze of W. Hoermann,
// "The transformed rejection method for generating Poisson random
// variables", Insurance: Math. Econ. 12 (1993) 39-45.
__device__ inline int64_t poisson(double lam, Stream& s) {
    if (!(lam > 0.)) return 0;
    if (lam < 10.) {
        const double u = s.uniform();
        double p = exp(-lam), F = p;
        int64_t k = 0;
        while (u > F && k < 1000) {
            k++;
            p *= lam / static_cast<double>(k);
            F += p;
        }
        return k;
    }
    const double slam = sqrt(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.);
    for (int it = 0; it < 10000; it++) {
        const double U = s.uniform() - 0.5;
        const double V = s.uniform();
        const double us = 0.5 - fabs(U);
        const double kd = floor((2. * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return static_cast<int64_t>(kd);
        if (kd < 0. || (us < 0.013 && V > us)) continue;
        // past the squeeze (few draws): the exact test
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <=
            -lam + kd * log(lam) - lgamma(kd + 1.))
            return static_cast<int64_t>(kd);
    }
    return static_cast<int64_t>(floor(lam + 0.5));  // unreachable in practice
}

// standard normal by the polar method (Marsaglia and Bray 1964)
__device__ inline double std_normal(Stream& s) {
    for (int it = 0; it < 1000; it++) {
        const double u = 2. * s.uniform() - 1., v = 2. * s.uniform() - 1.;
        const double q = u * u + v * v;
        if (q > 0. && q < 1.) return u * sqrt(-2. * log(q) / q);
    }
    return 0.;
}

// Gamma(shape, 1): G. Marsaglia and W. W. Tsang, "A simple method for
// generating gamma variables", ACM TOMS 26 (2000) 363-372; shape < 1 through
// Gamma(shape + 1) U^(1/shape)
__device__ inline double gamma_variate(double shape, Stream& s) {
    if (!(shape > 0.)) return 0.;
    double boost = 1.;
    if (shape < 1.) {
        boost = pow(s.uniform(), 1. / shape);
        shape += 1.;
    }
    const double d = shape - 1. / 3., c = 1. / sqrt(9. * d);
    for (int it = 0; it < 10000; it++) {
        double x, v;
        do {
            x = std_normal(s);
            v = 1. + c * x;
        } while (v <= 0.);
        v = v * v * v;
        const double u = s.uniform();
        if (u < 1. - 0.0331 * x * x * x * x || log(u) < 0.5 * x * x + d * (1. - v + log(v))) return d * v * boost;
    }
    return d * boost;
}

// negative binomial with success probability p and n successes (mean
// n (1 - p) / p), as the Gamma-Poisson mixture Poisson(Gamma(n) (1 - p) / p)
__device__ inline int64_t negative_binomial(double p, double n, Stream& s) {
    return poisson(gamma_variate(n, s) * (1. - p) / p, s);
}

// Binomial(n, p), exact.  n p < 10 (after p -> min(p, 1-p)): inversion;
// otherwise BTRS of W. Hoermann, "The generation of binomial random variates",
// J. Statist. Comput. Simul. 46 (1993) 101-110.
__device__ inline int64_t binomial(int64_t n, double p, Stream& s) {
    if (n <= 0 || !(p > 0.)) return 0;
    if (p >= 1.) return n;
    const bool flip = p > 0.5;
    const double pp = flip ? 1. - p : p;
    const double q = 1. - pp;
    const double nd = static_cast<double>(n);
    int64_t k;
    if (nd * pp < 10.) {
        const double u = s.uniform();
        const double r = pp / q;
        double f = exp(nd * log(q)), F = f;
        k = 0;
        while (u > F && k < n) {
            f *= r * static_cast<double>(n - k) / static_cast<double>(k + 1);
            k++;
            F += f;
        }
    } else {
        const double spq = sqrt(nd * pp * q);
        const double b = 1.15 + 2.53 * spq;
        const double a = -0.0873 + 0.0248 * b + 0.01 * pp;
        const double c = nd * pp + 0.5;
        const double alpha = (2.83 + 5.1 / b) * spq;
        const double vr = 0.92 - 4.2 / b;
        const double lpq = log(pp / q);
        const double m = floor((nd + 1.) * pp);
        // log(m! (n - m)!), needed only past the squeeze (most draws stop
        // there), so computed on first use: two FP64 lgamma less per draw
        double h = 0.;
        bool have_h = false;
        k = static_cast<int64_t>(floor(nd * pp + 0.5));
        for (int it = 0; it < 10000; it++) {
            const double u = s.uniform() - 0.5;
            double v = s.uniform();
            const double us = 0.5 - fabs(u);
            const double kd = floor((2. * a / us + b) * u + c);
            if (kd < 0. || kd > nd) continue;
            if (us >= 0.07 && v <= vr) { k = static_cast<int64_t>(kd); break; }
            if (!have_h) {
                h = lgamma(m + 1.) + lgamma(nd - m + 1.);
                have_h = true;
            }
            v = log(v * alpha / (a / (us * us) + b));
            if (v <= h - lgamma(kd + 1.) - lgamma(nd - kd + 1.) + (kd - m) * lpq) {
                k = static_cast<int64_t>(kd);
                break;
            }
        }
    }
    return flip ? n - k : k;
}

// ------------------------------------------------------------ permutation
// A keyed pseudo-random permutation of [0, n), n < 2^32: a mixed-radix
// Feistel network on [0, a b), a = ceil(sqrt n), b = ceil(n / a) (both at
// least 2), six rounds, with cycle walking.  x = L b + R with L in [0, a),
// R in [0, b); a round maps (L, R) to (R, (L + F(R)) mod a) and swaps the
// radices, so after an even number of rounds the radices are (a, b) again.
// F(R) is a keyed 32-bit hash brought into [0, a) by a multiply-high.  Since
// a b - n < a, an output falls outside [0, n) with probability below
// 1 / sqrt(n): the walk almost never takes a second pass, so the lanes of a
// wave do not wait on each other's walks (with a power-of-two domain up to
// half of the outputs walk, and a wave runs several passes).
struct Permutation {
    uint64_t n;
    uint64_t cb;  // ceil(2^64 / b): x / b = mulhi(cb, x) for every x < 2^32
    uint32_t a, b;
    uint32_t key0, key1;
    __host__ __device__ void init(uint64_t n_in, uint64_t seed) {
        n = n_in;
        uint64_t s = 1;
        while (s * s < n_in) s++;
        a = static_cast<uint32_t>(s < 2 ? 2 : s);
        const uint64_t bb = (n_in + a - 1) / a;
        b = static_cast<uint32_t>(bb < 2 ? 2 : bb);
        cb = ~0ull / b + 1;
        key0 = static_cast<uint32_t>(seed * 0x9E3779B97F4A7C15ull >> 32);
        key1 = static_cast<uint32_t>(seed);
    }
    __device__ __forceinline__ static uint32_t mix(uint32_t x, uint32_t k) {
        x ^= k;
        x ^= x >> 16;
        x *= 0x7FEB352Du;
        x ^= x >> 15;
        x *= 0x846CA68Bu;
        x ^= x >> 16;
        return x;
    }
    __device__ __forceinline__ uint32_t once(uint32_t v) const {
        // (L, R) = (v / b, v mod b) by the multiply-high of cb
        const uint64_t t = static_cast<uint64_t>(static_cast<uint32_t>(cb)) * v;
        const uint64_t u = static_cast<uint64_t>(static_cast<uint32_t>(cb >> 32)) * v + (t >> 32);
        uint32_t L = static_cast<uint32_t>(u >> 32), R = v - L * b;
        uint32_t A = a, B = b;
#pragma unroll
        for (int r = 0; r < 6; r++) {
            const uint32_t f = mix(R, key0 + 0x632BE5ABu * r) ^ key1;
            const uint32_t add = static_cast<uint32_t>((static_cast<uint64_t>(f) * A) >> 32);  // in [0, A)
            uint32_t nr = L + add;
            nr = nr >= A ? nr - A : nr;
            L = R;
            R = nr;
            const uint32_t tmp = A;
            A = B;
            B = tmp;
        }
        return L * b + R;
    }
    __device__ __forceinline__ uint64_t operator()(uint64_t v) const {
        uint32_t x = once(static_cast<uint32_t>(v));
        while (x >= n) x = once(x);
        return x;
    }
};

// LDS traffic between lanes of one wave: wait for this wave's LDS operations
// and keep the compiler from moving memory accesses across
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// ------------------------------------------------------ DPP lane exchange
// Data-parallel-primitive moves between lanes (no LDS round trip):
// row_shr:n (0x110 + n) shifts within rows of 16 lanes; row_bcast:15 (0x142)
// and row_bcast:31 (0x143) pass lane 15 of each row, and lane 31, to the
// following rows.  Lanes without a source, and rows outside row_mask, take
// `old`.
constexpr int kDppRowShr = 0x110, kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(v), CTRL,
                                                             ROW_MASK, 0xf, false));
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double old, double v) {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = dpp_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(o), static_cast<uint32_t>(x));
    const uint32_t hi = dpp_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(o >> 32), static_cast<uint32_t>(x >> 32));
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// Inclusive sum over the lanes up to this one, FP64, in DPP steps: within
// rows of 16 (shifts 1, 2, 4, 8), then the rows' totals (rows 1 and 3 add
// row 0 and 2's last lane, rows 2 and 3 add lane 31).  Every lane's value
// comes from a fixed sequence of additions (deterministic).
__device__ __forceinline__ double wave_scan_dpp(double v) {
    v += dpp_f64<kDppRowShr + 1>(0., v);
    v += dpp_f64<kDppRowShr + 2>(0., v);
    v += dpp_f64<kDppRowShr + 4>(0., v);
    v += dpp_f64<kDppRowShr + 8>(0., v);
    v += dpp_f64<kDppRowBcast15, 0xa>(0., v);
    v += dpp_f64<kDppRowBcast31, 0xc>(0., v);
    return v;
}
// lane 63's value, returned uniformly
__device__ __forceinline__ double wave_last(double v) {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(x), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), 63);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
// the sum over all lanes, returned uniformly (from lane 63 of the scan)
__device__ __forceinline__ double wave_sum_dpp(double v) { return wave_last(wave_scan_dpp(v)); }
// Running maximum over the lanes up to this one, in the steps of
// wave_scan_dpp; for non-negative values (lanes without a source take 0)
__device__ __forceinline__ double wave_max_scan_dpp(double v) {
    v = fmax(v, dpp_f64<kDppRowShr + 1>(0., v));
    v = fmax(v, dpp_f64<kDppRowShr + 2>(0., v));
    v = fmax(v, dpp_f64<kDppRowShr + 4>(0., v));
    v = fmax(v, dpp_f64<kDppRowShr + 8>(0., v));
    v = fmax(v, dpp_f64<kDppRowBcast15, 0xa>(0., v));
    v = fmax(v, dpp_f64<kDppRowBcast31, 0xc>(0., v));
    return v;
}

// The 64-lane sums of 8 values at once (a transposed butterfly): lanes
// exchange half of their values with the partner lane and add, halving the
// values each step (lane l ^ 32 through v_permlane32_swap, l ^ 16 through
// v_permlane16_swap, l ^ 8 through DPP row_ror:8), then the group of 8 lanes
// adds its values pairwise (row_half_mirror, then the quad).  Lane l returns
// the sum of v[j], j = b3 + 2 b4 + 4 b5 (b_k bit k of l), bit-identical in
// every lane of its group of 8 (the same pairs in the same order); ~5
// instructions per value instead of ~21 for a DPP scan of each.
__device__ __forceinline__ double pair_add32(double a, double b) {
    const uint64_t x = __builtin_bit_cast(uint64_t, a), y = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane32_swap(static_cast<uint32_t>(x), static_cast<uint32_t>(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(static_cast<uint32_t>(x >> 32), static_cast<uint32_t>(y >> 32),
                                                     false, false);
    const double r0 = __builtin_bit_cast(double, (static_cast<uint64_t>(hi[0]) << 32) | lo[0]);
    const double r1 = __builtin_bit_cast(double, (static_cast<uint64_t>(hi[1]) << 32) | lo[1]);
    return r0 + r1;
}
__device__ __forceinline__ double pair_add16(double a, double b) {
    const uint64_t x = __builtin_bit_cast(uint64_t, a), y = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane16_swap(static_cast<uint32_t>(x), static_cast<uint32_t>(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(static_cast<uint32_t>(x >> 32), static_cast<uint32_t>(y >> 32),
                                                     false, false);
    const double r0 = __builtin_bit_cast(double, (static_cast<uint64_t>(hi[0]) << 32) | lo[0]);
    const double r1 = __builtin_bit_cast(double, (static_cast<uint64_t>(hi[1]) << 32) | lo[1]);
    return r0 + r1;
}
constexpr int kDppRowRor8 = 0x128, kDppRowHalfMirror = 0x141, kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E;

__device__ __forceinline__ double wave_sum8(const double (&v)[8]) {
    const int lane = threadIdx.x & (kWave - 1);
    double a[4], b[2];
#pragma unroll
    for (int k = 0; k < 4; k++) a[k] = pair_add32(v[k], v[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; k++) b[k] = pair_add16(a[k], a[k + 2]);
    const bool b3 = (lane & 8) != 0;
    const double send = b3 ? b[0] : b[1], keep = b3 ? b[1] : b[0];
    double d = keep + dpp_f64<kDppRowRor8>(0., send);
    d += dpp_f64<kDppRowHalfMirror>(0., d);
    d += dpp_f64<kDppQuadXor1>(0., d);
    d += dpp_f64<kDppQuadXor2>(0., d);
    return d;
}
__device__ __forceinline__ int wave_sum8_index(int lane) {
    return ((lane >> 3) & 1) | (((lane >> 4) & 1) << 1) | (((lane >> 5) & 1) << 2);
}

// ------------------------------------------------------------- wave scans
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T o = __shfl_up(v, d, kWave);
        if (lane >= d) v += o;
    }
    return v;
}

// running maximum over the lanes up to this one
template <typename T>
__device__ __forceinline__ T wave_inclusive_max(T v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T o = __shfl_up(v, d, kWave);
        if (lane >= d) v = v > o ? v : o;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// first index i in [lo, hi) with a[i] > x (a sorted ascending); hi if none
template <typename T, typename I>
__device__ __forceinline__ I upper_bound(const T* a, I lo, I hi, T x) {
    while (lo < hi) {
        const I mid = lo + (hi - lo) / 2;
        if (a[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

}  // namespace issdev
