// Resonance feed-down of the sampled hadrons: particle_decay::perform_decays
// (src/particle_decay.cpp:262-533) driven to stable particles as
// FSSW::perform_resonance_feed_down does (src/FSSW.cpp:1746-1779).
// (one translation unit: included by iss_hip.hip)
//
// Data-parallel over particles, one generation per pass:
//   k_decay_pick   per particle: table entry, channel (cumulative branching
//                  ratio), number of daughters; kind = final (stable or not in
//                  the table), decayed (2- or 3-body) or gone (other channels:
//                  no daughters, as in the reference)
//   scan           daughter offsets of the next generation
//   k_decay_emit   the daughters, with their parent's index
// After the last generation the final particles are placed deterministically
// without sorting: every particle's count of final descendants is gathered
// bottom-up from its (contiguous) daughters, a scan over the sampled hadrons
// gives each its block of the output (so events stay grouped), and each
// decayed particle hands its block to its daughters in order.
//
// Kinematics follow the reference at the pole mass (its Breit-Wigner draw is
// switched off): isotropic two-body decay, three-body energies by rejection on
// the Dalitz box then an Euler rotation, an exponential lifetime of mean
// E / (M Gamma) and the decay point on the mother's line; FP64 arithmetic,
// float records.  Random numbers: a Philox stream per (generation, particle).
#pragma once

#include "sample_kernels.h"

namespace issdev {

constexpr uint32_t TAG_DECAY = 8;
constexpr int kDecayMaxGen = 12;

struct DecayDev {
    int n;                    // table entries
    const int* key;           // [n] monvals, ascending
    const int* key_idx;       // [n] entry of key[i]
    const double* mass;       // [n]
    const double* width;      // [n]
    const int* stable;        // [n]
    const int* ch0;           // [n] first channel
    const int* nch;           // [n]
    const int* npart;         // [channels]
    const double* cumbr;      // [channels] cumulative branching ratio within the entry
    const int* dpid;          // [channels][5] daughter monvals
    const int* didx;          // [channels][5] daughter entries (-1: not in the table)
};

__device__ __forceinline__ int decay_entry(const DecayDev& t, int pid) {
    int lo = 0, hi = t.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t.key[mid] < pid) lo = mid + 1; else hi = mid;
    }
    return (lo < t.n && t.key[lo] == pid) ? t.key_idx[lo] : -1;
}

enum DecayKind : uint8_t { DK_FINAL = 0, DK_DECAYED = 1, DK_GONE = 2 };

__device__ __forceinline__ Stream decay_stream(uint64_t seed, int gen, int64_t i) {
    return Stream(seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), tag_word(TAG_DECAY, gen));
}

// kind, channel and daughter count of every particle of generation `gen`
// (a channel closed at the pole mass drops the mother and is counted in
// *nbelow: the reference exits there, src/particle_decay.cpp:351-357, 431-441)
__global__ void k_decay_pick(DecayDev t, const Hadron* __restrict__ P, int64_t n, int gen, uint64_t seed,
                             uint8_t* __restrict__ kind, int32_t* __restrict__ chan,
                             uint32_t* __restrict__ nchild, unsigned long long* __restrict__ nbelow) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = decay_entry(t, P[i].pid);
    uint8_t k = DK_FINAL;
    int c = -1;
    uint32_t nc = 0;
    if (e >= 0 && !t.stable[e] && t.nch[e] > 0) {
        Stream st = decay_stream(seed, gen, i);
        const double r = st.uniform();
        const int c0 = t.ch0[e], c1 = c0 + t.nch[e];
        c = c1 - 1;  // branching ratios summing below r: the last channel
        for (int j = c0; j < c1; j++)
            if (t.cumbr[j] > r) { c = j; break; }
        const int np = t.npart[c];
        if (np == 2 || np == 3) {
            float msum = 0.f;  // daughter masses as the float records hold them
            for (int j = 0; j < np; j++) {
                const int de = t.didx[5 * c + j];
                msum += static_cast<float>(de >= 0 ? t.mass[de] : 0.);
            }
            if (static_cast<double>(P[i].mass) < static_cast<double>(msum)) {
                k = DK_GONE;
                atomicAdd(nbelow, 1ull);
            } else {
                k = DK_DECAYED;
                nc = static_cast<uint32_t>(np);
            }
        } else {
            k = DK_GONE;
        }
    }
    kind[i] = k;
    chan[i] = c;
    nchild[i] = nc;
}

__device__ __forceinline__ void decay_boost(double vx, double vy, double vz, double E, double px, double py,
                                            double pz, Hadron& h) {
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double gamma = 1. / sqrt(1. - v2);
    const double vp = vx * px + vy * py + vz * pz;
    // (a mother exactly at rest: no boost, where the reference divides 0 by 0)
    const double k = v2 > 0. ? (gamma - 1.) * vp / v2 + gamma * E : 0.;
    h.E = static_cast<float>(gamma * (E + vp));
    h.px = static_cast<float>(px + k * vx);
    h.py = static_cast<float>(py + k * vy);
    h.pz = static_cast<float>(pz + k * vz);
}

// daughters of the decayed particles of generation `gen` into generation gen + 1
__global__ void k_decay_emit(DecayDev t, const Hadron* __restrict__ P, int64_t n, int gen, uint64_t seed,
                             const uint8_t* __restrict__ kind, const int32_t* __restrict__ chan,
                             const int64_t* __restrict__ off, Hadron* __restrict__ C, int64_t* __restrict__ parent,
                             unsigned long long* __restrict__ nfail) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || kind[i] != DK_DECAYED) return;
    const Hadron m = P[i];
    const int c = chan[i];
    const int e = decay_entry(t, m.pid);
    const int np = t.npart[c];
    Stream st = decay_stream(seed, gen, i);
    (void)st.uniform();  // the channel draw
    Hadron d[3];
    double ms[3];
    for (int k = 0; k < np; k++) {
        d[k].pid = t.dpid[5 * c + k];
        const int de = t.didx[5 * c + k];
        d[k].mass = static_cast<float>(de >= 0 ? t.mass[de] : 0.);
        ms[k] = d[k].mass;
    }
    const double Mp = m.mass, W = t.width[e];
    const double vx = static_cast<double>(m.px) / m.E, vy = static_cast<double>(m.py) / m.E,
                 vz = static_cast<double>(m.pz) / m.E;
    double life = 1e10;
    if (np == 2) {
        const double tt = Mp * Mp - ms[0] * ms[0] - ms[1] * ms[1];
        const double p = sqrt(fmax(tt * tt - 4 * ms[0] * ms[0] * ms[1] * ms[1], 0.)) / (2 * Mp);
        const double phi = st.uniform() * 2 * kPi;
        const double cth = 2. * (st.uniform() - 0.5);
        const double sth = sqrt(fmax(1. - cth * cth, 0.));
        double sphi, cphi;
        sincos(phi, &sphi, &cphi);
        const double px = p * sth * cphi, py = p * sth * sphi, pz = p * cth;
        decay_boost(vx, vy, vz, sqrt(p * p + ms[0] * ms[0]), px, py, pz, d[0]);
        decay_boost(vx, vy, vz, sqrt(p * p + ms[1] * ms[1]), -px, -py, -pz, d[1]);
        if (W > 1e-10) life = -(m.E / Mp / W) * log(st.uniform()) * 0.19733;
    } else {
        const double Q = Mp - ms[0] - ms[1] - ms[2];
        double E1 = 0., E2 = 0., p1 = 0., p2 = 0., c12 = 2.;
        int tries = 0;
        do {
            do {
                E1 = st.uniform() * Q + ms[0];
                E2 = st.uniform() * Q + ms[1];
            } while (E1 + E2 > Mp && ++tries < 100000);
            p1 = sqrt(fmax(E1 * E1 - ms[0] * ms[0], 0.));
            p2 = sqrt(fmax(E2 * E2 - ms[1] * ms[1], 0.));
            const double E3 = Mp - E1 - E2;
            c12 = (E3 * E3 - p1 * p1 - p2 * p2 - ms[2] * ms[2]) / (2. * p1 * p2);
        } while ((c12 < -1.0 || c12 > 1.0) && ++tries < 100000);
        if (tries >= 100000) {
            atomicAdd(nfail, 1ull);
            c12 = fmin(fmax(c12, -1.), 1.);
        }
        if (W > 1e-10) life = -(m.E / Mp / W) * log(st.uniform()) * 0.19733;
        const double t2x = p2 * sqrt(fmax(1. - c12 * c12, 0.)), t2z = p2 * c12;
        const double t3x = -t2x, t3z = -(p1 + t2z);
        const double phi = 2. * kPi * st.uniform();
        const double ksi = 2. * kPi * st.uniform();
        const double cth = 2. * st.uniform() - 1.0;
        double sp, cp, sk, ck;
        sincos(phi, &sp, &cp);
        sincos(ksi, &sk, &ck);
        const double sth = sqrt(fmax(1. - cth * cth, 0.));
        const double a = cp * cth * ck - sp * sk, b = -cp * cth * sk - sp * ck;
        const double q[3][3] = {{-p1 * sth * ck, p1 * sth * sk, p1 * cth},
                                {t2x * a - t2z * sth * ck, t2x * b + t2z * sth * sk, t2x * cp * sth + t2z * cth},
                                {t3x * a - t3z * sth * ck, t3x * b + t3z * sth * sk, t3x * cp * sth + t3z * cth}};
        for (int k = 0; k < 3; k++)
            decay_boost(vx, vy, vz, sqrt(ms[k] * ms[k] + q[k][0] * q[k][0] + q[k][1] * q[k][1] + q[k][2] * q[k][2]),
                        q[k][0], q[k][1], q[k][2], d[k]);
    }
    const int64_t o = off[i];
    for (int k = 0; k < np; k++) {
        d[k].t = static_cast<float>(m.t + life);
        d[k].x = static_cast<float>(m.x + static_cast<double>(m.px) / m.E * life);
        d[k].y = static_cast<float>(m.y + static_cast<double>(m.py) / m.E * life);
        d[k].z = static_cast<float>(m.z + static_cast<double>(m.pz) / m.E * life);
        C[o + k] = d[k];
        parent[o + k] = i;
    }
}

// final descendants of every particle of a generation, from its daughters'
// counts (the last generation: 1 for final, 0 otherwise)
__global__ void k_decay_size(const uint8_t* __restrict__ kind, const int64_t* __restrict__ off,
                             const uint32_t* __restrict__ nchild, const uint32_t* __restrict__ child_size,
                             int64_t n, uint32_t* __restrict__ size) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t s = 0;
    if (kind[i] == DK_FINAL) {
        s = 1;
    } else if (kind[i] == DK_DECAYED && child_size) {
        for (uint32_t k = 0; k < nchild[i]; k++) s += child_size[off[i] + k];
    }
    size[i] = s;
}

// positions of the daughters' blocks: each decayed particle's block, split in
// daughter order
__global__ void k_decay_place(const uint8_t* __restrict__ kind, const int64_t* __restrict__ off,
                              const uint32_t* __restrict__ nchild, const int64_t* __restrict__ pos,
                              const uint32_t* __restrict__ child_size, int64_t n, int64_t* __restrict__ child_pos) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || kind[i] != DK_DECAYED) return;
    int64_t p = pos[i];
    for (uint32_t k = 0; k < nchild[i]; k++) {
        child_pos[off[i] + k] = p;
        p += child_size[off[i] + k];
    }
}

__global__ void k_decay_write(const Hadron* __restrict__ P, const uint8_t* __restrict__ kind,
                              const int64_t* __restrict__ pos, int64_t n, Hadron* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || kind[i] != DK_FINAL) return;
    out[pos[i]] = P[i];
}

// new event offsets: the output position of each event's first sampled hadron
__global__ void k_decay_events(const int64_t* __restrict__ evoff, int E, const int64_t* __restrict__ pos0,
                               int64_t* __restrict__ evoff_new) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > E) return;
    evoff_new[e] = pos0[evoff[e]];
}

}  // namespace issdev
