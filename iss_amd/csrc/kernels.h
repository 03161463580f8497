// Device data layout and kernels of the gfx950 Cooper-Frye sampler.
//
// HBM layout (one context):
//   cells   float  [kCellFields][Nc]  rest-frame surface, field-major (SoA)
//   coef    double [8][Nc]            per-cell delta-f coefficients (6 bulk /
//                                     NEoS, 1 / D_qmu, n_B / (e + P))
//   cdf     double [Nc][cdf_stride]   per-cell inclusive species prefix of the
//                                     clamped yields (the species-selection CDF)
//   ctot    double [Nc]               per-cell total yield (last cdf entry)
//   ktab    double [3][kKLen]         K1, K2, K3 at x = 0.5 + 0.05 i
//   ktabf   float2 [3][kKLen]         (K_i, K_{i+1} - K_i) for the FP32 pass
//   ptab    double [6][3][2001]       |p|-sampler CDF_0/1/2 tables
#pragma once

#include "device.h"

namespace issdev {

// field index in the rest-frame cell record (FO_surf_LRF order,
// src/data_struct.h:66-76)
enum CellField {
    F_TAU = 0, F_X, F_Y, F_ETA,
    F_DA0, F_DA1, F_DA2, F_DA3,
    F_U0, F_U1, F_U2, F_U3,
    F_E, F_T, F_P, F_NB, F_MUB, F_MUS, F_MUQ, F_PI,
    F_PIXX, F_PIXY, F_PIXZ, F_PIYY, F_PIYZ,
    F_QX, F_QY, F_QZ,
    kCellFields
};

// special-function table (src/FSSW.cpp:1609-1615)
constexpr double kKx0 = 0.5, kKdx = 0.05, kKxmax = 400.;
constexpr int kKLen = 7991;
constexpr int kKLds = 1536;      // rows staged in LDS (args up to 77.25)

constexpr int kEnTerms = 9;     // E_2, E_4, ..., E_18 (src/FSSW.cpp:1611, 1637-1642)
constexpr int kCoefRows = 8;    // per-cell coefficients: 6 bulk / NEoS, 1 / D_qmu, n_B / (e + P)

constexpr int kMaxSpecies = 512;
// the species CDF's row stride: rows start on 128-byte boundaries, so that
// the yield pass stores whole 128-byte lines (16 species per staging chunk)
__host__ __device__ constexpr int cdf_stride(int ns) { return (ns + 15) & ~15; }
constexpr int kMaxChunks = kMaxSpecies / kWave;
constexpr int kPTabLen = 2001;

enum BulkMode { BULK_NONE = 0, BULK_RTA = 1, BULK_14MOM = 11, BULK_22MOM = 20, BULK_CE = 21 };

struct SpeciesDev {
    const double* mass;
    const int* g;
    const int* B;
    const int* S;
    const int* Q;
    const int* sign;
    const int* pid;
};

// species constants of the yield pass, read wave-uniformly (scalar loads):
// mass and g / (2 pi^2 (hbar c)^3) in FP32 and FP64, charges and statistics
struct SpeciesK {
    float m, pref, B, S, Q, sgn;
    double md, prefd;
    // the one-term yields of the row pass: m^2, m^3, m^4, degeneracy, the
    // charges as integers, 20 m (FP64: places m / T in the K table)
    float m2, m3, m4, g;
    int Bi, Si, Qi, pad;
    double md20;
};

// Species constants of the yield-row pass (padded to whole 64-byte segments
// of the row; padding and suppressed species have g = 0): 20 m in FP64, m,
// m^2, m^3, m^4, m^2 B, g, the LDS rows of the fugacity factors e^{beta mu q}
// of its baryon number, strangeness and charge, and the quantum-statistics
// sign (+1 fermions, -1 bosons)
struct RowSp {
    double md20;
    float m, m2, m3, m4, m2B, g;
    int frow;   // row of e^{beta mu_B B} | row of e^{beta mu_S S} << 8 | row of e^{beta mu_Q Q} << 16
    float sgn;
};
static_assert(sizeof(RowSp) == 40 && alignof(RowSp) == 8, "RowSp is copied to LDS as 10 words");

struct PTabDev {
    // per table: first E~, step, length, m0, truncation, fermion flag
    double Et0[6], dEt[6], m0[6];
    double ep[6], em1[6];  // exp(m0), exp(-m0)
    int len[6], trunc[6], fermion[6];
    int quirk;
    const double* C;  // [6][3][kPTabLen]
};

struct YieldArgs {
    const float* cells;
    const double* coef;
    int64_t Nc;
    int Ns;
    SpeciesDev sp;
    const double* ktab;
    const float2* ktabf;
    int bulk;        // BulkMode, BULK_NONE when bulk delta f is off
    double* cdf;
    double* ctot;
    void* raw;       // optional per-(cell, species) yields in working precision
    double* partial; // [gridDim.x][Ns]
    int* oob;        // set when a Bessel argument leaves the table
    const double* dnq;   // [Nc][Ns] diffusion corrections (k_diffusion_dn); null: none
    const SpeciesK* spk;  // [Ns] species constants of the yield pass
    int lcc;              // local charge conservation: negative species get no yield
    double* ycol;         // optional [Ns][Nc] FP64 yields, species-major (multiplicity models)
    void* rows;           // k_yield_rows: [Nc][row stride] yields in working precision, cell-major
    const struct RowSp* rowsp;  // k_yield_rows: species constants, padded (zero weight beyond Ns)
    int nlight;           // k_yield_rows: leading species below 0.7 GeV (the 10-term series)
    int all_general;      // k_yield_rows: a species below 0.7 GeV after the leading ones (at most 16),
                          // or a charge outside the fugacity rows: every species takes the general series
    const int* gate;      // k_yield_rows, exact-K pass: if set, the pass runs only where *gate != 0
                          // (the first pass's out-of-table flag, read on the device)
};

struct SampleArgs {
    const float* cells;
    const double* coef;
    int64_t Nc;
    int Ns;
    SpeciesDev sp;
    PTabDev pt;
    const double* cdf;
    const double* ctot;
    const int64_t* off;      // [Nc + 1] exclusive prefix of per-cell counts
    const int64_t* evoff;    // [E + 1]
    int E;
    int64_t K;               // hadrons in this pass
    const unsigned long long* Kdev;  // if set, the trial kernel's record count (device)
    uint64_t seed;
    uint32_t chunk;
    Permutation perm;
    int shear, bulk, neos;   // neos: -1 none, 1 CE, 0 22-moment
    int diff;                // diffusion delta f
    int boost_invariant;
    float yLB, yRB;
    int max_trials;
    int refill_min;          // idle lanes that trigger a refill in the trial kernels
    void* out;               // iss_hadron[K]
    unsigned long long* stats;  // trials, inner, failed, range errors
    // batched surfaces (all null / zero for one surface): the pass (round)
    // index and the hadrons written by earlier rounds live on the device, so
    // that one captured round can be replayed
    const uint32_t* chunk_dev;         // if set, overrides chunk
    const int64_t* surf_off;           // [S + 1] cell offsets of the surfaces
    int S;
    const Permutation* perms;          // [S] per-surface permutation of the round
    const unsigned long long* out_base;  // hadrons of earlier rounds
    uint64_t out_cap;                  // hadron slots in out (0: unchecked)
    int64_t setup_cap;                 // set-up records (0: unchecked)
    // local charge conservation: the antiparticle's species index ([Ns], -1
    // none); partner records carry kPartnerBit in j
    const int* anti;
    // the cell of every slot of the pass (k_pick); null: search the offsets
    const uint32_t* cellidx;
    // the species of every slot (k_pick); null: drawn from the FP64 species CDF rows
    const uint16_t* slot_species;
    // 3+1D surfaces: per cell (u0, u1, u2, u3) and (t, x, y, z) (k_emit_cells),
    // loaded when a trial lane takes a hadron; null: emit_lab gathers them
    const float4* ecell;
};

constexpr uint32_t kPartnerBit = 0x80000000u;

}  // namespace issdev
