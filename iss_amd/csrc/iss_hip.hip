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
//   k_spectra_rec + k_spectra smooth dN/(pT dpT dphi dy) on a (pT, phi) grid
//                             (+ k_spectra_reduce), host-side flows
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
#include <memory>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/iss_hip.h"
#include "kernels.h"

using namespace issdev;

#include "surface_kernels.h"
#include "yield_kernels.h"
#include "sample_kernels.h"
#include "learned_kernels.h"
#include "spectra_kernels.h"
#include "batch_kernels.h"
#include "decay_kernels.h"

// Scratch of the allocation-free yield entry point: a 256-byte header (the
// out-of-table flag in its first word), then the species partials of up to
// kNoAllocPartials (block, species) pairs: 4 MiB in all.  Each GPU holds its
// own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device.
constexpr int kNoAllocHeader = 32;         // doubles
constexpr int kNoAllocPartials = 1 << 19;  // doubles
constexpr int kNoAllocMaxSpecies = kNoAllocPartials / 256;  // keeps >= 256 blocks (one per CU)
static __device__ __align__(256) double scratch[kNoAllocHeader + kNoAllocPartials];

// =================================================================== host

namespace {

// the device address of `scratch`, looked up once per device (with that
// device current); null with the error text if the lookup failed
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];

double* scratch_address(int device, std::string* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = "device index beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device] && g_scratch_err[device].empty()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        else
            g_scratch_addr[device] = static_cast<double*>(p);
    }
    *err = g_scratch_err[device];
    return g_scratch_addr[device];
}

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
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    // grows with 1/8 headroom above a previous allocation, so that sizes
    // that vary from pass to pass (hadron counts) do not reallocate each
    // time a new maximum comes up
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        const size_t want = cap > 0 ? n + n / 8 : n;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host memory for the read-backs of counts, flags and totals: a copy
// from device memory into pageable memory goes through the runtime's staging
// buffer and costs tens of microseconds more than one into pinned memory.
// Slots: PIN_OOB int, PIN_K int64, PIN_STATS 16 x u64, PIN_AUX 2 x int64, then
// PIN_TOT [Ns] doubles.
struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
    PinBuf() = default;
    PinBuf(const PinBuf&) = delete;
    PinBuf& operator=(const PinBuf&) = delete;
    ~PinBuf() {
        if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
        if (e == hipSuccess) cap = n;
        return e;
    }
    template <typename T>
    T* at(size_t off) { return reinterpret_cast<T*>(static_cast<char*>(p) + off); }
};
constexpr size_t PIN_OOB = 0, PIN_K = 8, PIN_STATS = 64, PIN_AUX = 192, PIN_TOT = 256;
// device counters: [0..3] trials, |p| proposals, failed, out of table; [4..5]
// learned fallback and envelope violations; [6..7] batch overflow flags; [8]
// routed hadrons of batches; [9..12] the routed plain trials' [0..3]
constexpr int kStatRouted = 9, kStatSlots = 13;
// events per surface and round of a batch at most (one round covers up to
// this many events of every surface)
constexpr int kMaxEventsPerRound = 16384;

// Learned passes: (cell, species) pairs of fewer hadrons than the routing
// threshold take the plain sampler.  A pair's network and envelope cost
// about as much as one plain hadron's set-up and trials, and each learned
// hadron saves about 40 % of those, so pairs of fewer than about four
// hadrons are cheaper on the plain sampler; the autotuner times the choice.
constexpr int kLearnRouteDefault = 4;
constexpr int kLearnRouteMax = 64;
// every pair routed: the learned pass runs as the plain pass, without the
// pair bookkeeping (what the autotuner keeps where no pair pays)
constexpr int kLearnRouteAll = kLearnRouteMax + 1;
static_assert(kLearnRouteAll == ISS_LEARNED_ROUTE_ALL, "routing constant of the C ABI");

// one generation of the resonance feed-down: its particles and per-particle
// bookkeeping (kept by the context and reused across calls)
struct DecayGen {
    const Hadron* P = nullptr;
    int64_t n = 0;
    DBuf<Hadron> own;
    DBuf<uint8_t> kind;
    DBuf<int32_t> chan;
    DBuf<uint32_t> nchild, size;
    DBuf<int64_t> off, pos, parent;
    void release() {
        own.release(); kind.release(); chan.release(); nchild.release(); size.release();
        off.release(); pos.release(); parent.release();
    }
};

}  // namespace

struct iss_ctx {
    PinBuf h_pin;  // pinned read-back slots (PIN_*)
    int device = 0;
    double* scratch = nullptr;  // this device's `scratch` (allocation-free entry point)
    std::string scratch_err;
    hipStream_t stream = nullptr;
    hipEvent_t ev[11] = {};
    std::string err;
    int num_cu = 256;
    // species
    int Ns = 0;
    std::vector<int> h_pid, h_charge, h_anti;
    DBuf<int> d_anti;
    // local charge conservation: output positions' flags, partner offsets,
    // sources, partner set-up records, the second output list, new event offsets
    DBuf<uint32_t> d_lcc_flag;
    DBuf<int64_t> d_lcc_g, d_lcc_src, d_lcc_evoff;
    DBuf<HSetup> d_lcc_rec;
    DBuf<Hadron> d_out2;
    // resonance feed-down: generation buffers and the second output list
    std::vector<std::unique_ptr<DecayGen>> dgen;
    int yields_lcc = 0;
    DBuf<double> d_mass;
    DBuf<SpeciesK> d_spk;
    DBuf<int> d_ints;  // g, B, S, Q, sign, pid
    // delta f
    iss_deltaf_desc df{};
    int bulk_mode = BULK_NONE, neos = -1;
    DBuf<double> d_k11, d_neos, d_qmu, d_etab, d_dnq;
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
    // lab cells dropped for u.dsigma < 0 (kept for the smooth spectra)
    int64_t Nneg = 0;
    DBuf<float> d_cells_neg;
    DBuf<double> d_coef_neg;
    DBuf<uint32_t> d_u32;
    DBuf<int64_t> d_i64, d_scan;
    DBuf<double> d_eos;
    // yields: LAYOUT_ROWS (k_yield_rows, the Poisson pass of one surface) or
    // LAYOUT_CDF (k_yield, FP64 species CDF rows: batches, multiplicity models)
    bool yields_valid = false;
    int yield_precision = 0;
    int yield_layout = 0;
    int nlight = 0;  // leading species below 0.7 GeV (the 10-term series of the row pass)
    bool light_beyond = false;  // a species below 0.7 GeV after the first nlight ones
    DBuf<char> d_rows;
    DBuf<float4> d_ecell;      // [2 Nc]: per cell flow and 3+1D position (k_emit_cells)
    DBuf<RowSp> d_rowsp;       // [2][row_stride(Ns, 4)]: all species, then negative charges suppressed
    bool charges_outside = false;  // a species charge beyond the row pass's fugacity rows
    DBuf<double> d_cdf, d_ctot, d_partial, d_stot;
    DBuf<int> d_oob;
    DBuf<char> d_raw;
    // sampling
    DBuf<uint32_t> d_counts, d_evcount;
    DBuf<int64_t> d_off, d_evoff, d_tree0, d_tree1;
    DBuf<Hadron> d_out;
    DBuf<HSetup> d_setup;
    DBuf<uint32_t> d_cellidx;  // the cell of every slot of the pass
    // multiplicity models 1, 10, 20: species-major cell prefixes, per (event,
    // species) counts and offsets, slot species
    DBuf<double> d_ycol;
    DBuf<uint32_t> d_mcnt;
    DBuf<int64_t> d_msegoff;
    DBuf<uint16_t> d_msp;
    DBuf<unsigned long long> d_stats;
    int64_t K = 0;
    int E = 0;
    iss_stats stats{};
    int yield_grid = 1;
    // learned proposal
    DBuf<float> d_aux;
    DBuf<float> d_mlp_par, d_adam_m, d_adam_v, d_grad;
    DBuf<uint4> d_mlp_frag;
    bool mlp_ready = false;
    int adam_t = 0;
    DBuf<int64_t> d_fb_list;
    DBuf<LRec> d_lrec;
    // learned sampler's (cell, species) pairs: compact list, species
    // bitmaps, per-cell counts and first pairs, slot species and records;
    // per-cell counts and first positions of the slots routed to the plain
    // sampler
    DBuf<uint2> d_pkey;
    DBuf<uint32_t> d_lbits, d_lcnt, d_lscnt;
    DBuf<int64_t> d_pairoff, d_soff;
    // pairs of fewer hadrons than this take the plain sampler (1: none)
    int learn_route = kLearnRouteDefault;
    DBuf<uint16_t> d_sp16;
    DBuf<LSlot> d_lslot;
    DBuf<unsigned long long> d_fb_count;
    DBuf<float> d_tpart, d_tloss, d_tout;
    DBuf<uint32_t> d_tr_cell, d_ev_cell;
    DBuf<uint16_t> d_tr_spec, d_ev_spec;
    DBuf<int32_t> d_i32a, d_i32b;
    DBuf<float> d_f32a, d_f32b, d_f32c;
    int64_t n_train = 0;
    bool last_plain = false;  // d_setup / d_out hold the last plain pass
    // launch shape of the trial kernels per proposal (autotuner)
    int refill_min[2] = {kRefillMin, kRefillMin};
    int blocks_per_cu[2] = {0, 0};  // 0: occupancy limit
    // batched surfaces
    int S = 0;                       // surfaces of the resident record (0: one surface)
    std::vector<int64_t> h_surf_off;
    DBuf<int64_t> d_surf_off;
    DBuf<int32_t> d_bevents;
    DBuf<uint32_t> d_bround, d_bevcount;
    DBuf<unsigned long long> d_bbase;
    DBuf<Permutation> d_perms;
    DBuf<int64_t> d_seg;
    DBuf<double> d_lam;
    int batch_rounds = 0, batch_R = 0, batch_S = 0;
    std::vector<int32_t> h_bevents;
    int batch_R_tuned[2] = {16, 256};  // events per round, plain and learned proposal
    double batch_ms = 0., batch_ms_graph_launch = 0.;
    // resonance decay table
    int dec_n = 0;
    DBuf<int> d_dec_int;       // key, key_idx, stable, ch0, nch ([5][n]); npart, dpid[5], didx[5] ([11][nc])
    DBuf<double> d_dec_dbl;    // mass, width ([2][n]); cumulative br [nc]
    DecayDev dec{};
    int64_t dec_failed = 0;
    // smooth spectra
    DBuf<char> d_srec, d_sgrid;
    DBuf<double> d_spart, d_sout;
    DBuf<int32_t> d_sspec;
    double spectra_ms = 0.;
};

namespace {

int set_err(iss_ctx* ctx, int code, const std::string& m) {
    ctx->err = m;
    return code;
}

enum YieldLayout { LAYOUT_ROWS = 1, LAYOUT_CDF = 2 };
constexpr int kPartialBlocksPerCU = 8;  // yield-pass workgroups per CU at most (species partials)

// Forget the hadron list of the previous pass (single surface or batch).  A
// pass that fails part-way, or a batch round that overflowed its reserved
// output, must leave no count that points past what was written: the
// getters and iss_decay then see an empty list.
void invalidate_output(iss_ctx* ctx) {
    ctx->K = 0;
    ctx->E = 0;
    ctx->batch_rounds = 0;
    ctx->batch_S = 0;
    ctx->h_bevents.clear();
    ctx->last_plain = false;
}

// the context's hadron list is readable: its count fits the output buffer
int check_output(iss_ctx* ctx) {
    if (ctx->K < 0 || (ctx->K > 0 && (!ctx->d_out.p || static_cast<size_t>(ctx->K) > ctx->d_out.cap)))
        return set_err(ctx, ISS_ERR_STATE, "hadron count exceeds the output buffer");
    return ISS_OK;
}

// the learned pass's per-kernel figures, which only iss_sample fills
void clear_routed_stats(iss_ctx* ctx) {
    ctx->stats.ms_learned_trials_kernel = 0.;
    ctx->stats.ms_routed_setup_kernel = 0.;
    ctx->stats.ms_routed_trials_kernel = 0.;
    ctx->stats.n_routed_trials = 0;
    ctx->stats.n_routed_inner = 0;
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
    const int tr[6] = {10, 2, 1, 10, 2, 1};  // (at most kMaxTrunc terms on the device)
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
    d.include_diffusion = ctx->df.include_diffusion;
    d.qmu_tab = ctx->d_qmu.p;
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
    if (ctx->Nneg > 0) {
        HIPCHK(ctx->d_coef_neg.ensure(kCoefRows * ctx->Nneg));
        k_cell_prep<<<(ctx->Nneg + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_cells_neg.p, ctx->Nneg, deltaf_dev(ctx),
                                                                       ctx->d_coef_neg.p);
        HIPCHK(hipGetLastError());
    }
    if (ctx->Nc == 0) return ISS_OK;
    HIPCHK(ctx->d_coef.ensure(kCoefRows * ctx->Nc));
    const int64_t nb = (ctx->Nc + 255) / 256;
    k_cell_prep<<<nb, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->Nc, deltaf_dev(ctx), ctx->d_coef.p);
    HIPCHK(ctx->d_aux.ensure(static_cast<size_t>(kAuxFields) * ctx->Nc));
    k_cell_aux<<<nb, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->Nc, ctx->d_aux.p);
    HIPCHK(hipGetLastError());
    HIPCHK(ctx->d_ecell.ensure(2 * ctx->Nc));
    k_emit_cells<<<nb, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->Nc, ctx->d_ecell.p);
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
    const size_t lds = yield_lds_bytes(sizeof(R), ktables_for(MODE), pad_species(a.Ns));
    auto kern = k_yield<R, MODE, RAW, EXACT>;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
        return ISS_ERR_HIP;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kYieldThreads, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    per_cu = std::min(per_cu, kPartialBlocksPerCU);
    const int64_t ntiles = (a.Nc + kYieldThreads - 1) / kYieldThreads;
    const int nblk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(ntiles, int64_t(ctx->num_cu) * per_cu)));
    ctx->yield_grid = nblk;
    kern<<<nblk, kYieldThreads, lds, ctx->stream>>>(a, a.spk);
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

template <typename R, int MODE, bool EXACT>
int launch_rows_t(iss_ctx* ctx, const YieldArgs& a, int max_blocks, int force_blocks) {
    const size_t lds = yield_rows_lds_bytes<R>(ktables_for(MODE), ctx->Ns);
    auto kern = k_yield_rows<R, MODE, EXACT>;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
        return ISS_ERR_HIP;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kRowThreads, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    per_cu = std::min(per_cu, kPartialBlocksPerCU);  // the species partials hold num_cu * 8 blocks
    const int64_t ntiles = (a.Nc + kRowThreads - 1) / kRowThreads;
    // (a forced count: the same grid as an earlier pass, whose partials a
    // skipped pass leaves in place)
    const int nblk = force_blocks > 0 ? force_blocks
                                      : static_cast<int>(std::max<int64_t>(
                                            1, std::min<int64_t>({ntiles, int64_t(ctx->num_cu) * per_cu,
                                                                  int64_t(max_blocks)})));
    ctx->yield_grid = nblk;
    kern<<<nblk, kRowThreads, lds, ctx->stream>>>(a, a.spk, a.rowsp);
    return hipGetLastError() == hipSuccess ? ISS_OK : ISS_ERR_HIP;
}

template <typename R, int MODE>
int launch_rows_m(iss_ctx* ctx, const YieldArgs& a, bool exact, int max_blocks, int force_blocks) {
    return exact ? launch_rows_t<R, MODE, true>(ctx, a, max_blocks, force_blocks)
                 : launch_rows_t<R, MODE, false>(ctx, a, max_blocks, force_blocks);
}

template <typename R>
int launch_rows(iss_ctx* ctx, const YieldArgs& a, bool exact, int max_blocks = 1 << 30, int force_blocks = 0) {
    switch (a.bulk) {
        case BULK_RTA: return launch_rows_m<R, BULK_RTA>(ctx, a, exact, max_blocks, force_blocks);
        case BULK_CE: return launch_rows_m<R, BULK_CE>(ctx, a, exact, max_blocks, force_blocks);
        case BULK_14MOM: return launch_rows_m<R, BULK_14MOM>(ctx, a, exact, max_blocks, force_blocks);
        case BULK_22MOM: return launch_rows_m<R, BULK_22MOM>(ctx, a, exact, max_blocks, force_blocks);
        default: return launch_rows_m<R, BULK_NONE>(ctx, a, exact, max_blocks, force_blocks);
    }
}

// The yield-pass arguments every layout shares (the resident surface,
// species and tables); the output and scratch pointers are left null.
YieldArgs base_yield_args(iss_ctx* ctx, int lcc) {
    YieldArgs a{};
    const int Ns = ctx->Ns;
    a.cells = ctx->d_cells.p;
    a.coef = ctx->d_coef.p;
    a.Nc = ctx->Nc;
    a.Ns = Ns;
    a.sp = species_dev(ctx);
    a.ktab = ctx->d_ktab.p;
    a.ktabf = ctx->d_ktabf.p;
    a.bulk = ctx->df.include_bulk ? ctx->bulk_mode : BULK_NONE;
    a.spk = ctx->d_spk.p;
    a.lcc = lcc;
    a.nlight = ctx->nlight;
    a.rowsp = ctx->d_rowsp.p ? ctx->d_rowsp.p + (lcc ? row_stride(Ns, 4) : 0) : nullptr;
    a.all_general = (ctx->light_beyond || ctx->charges_outside) ? 1 : 0;
    return a;
}

// The yield pass.  LAYOUT_ROWS writes every (cell, species) yield in the
// working precision (d_rows, [Nc][Ns]); LAYOUT_CDF the FP64 species CDF row
// of every cell (d_cdf) and, with ycol, the FP64 species-major yields.
int run_yields(iss_ctx* ctx, int precision, void* per_cell_host, double* totals_host, int lcc = 0,
               double* ycol = nullptr, int layout = LAYOUT_ROWS) {
    if (ctx->Ns == 0) return set_err(ctx, ISS_ERR_STATE, "species not set");
    if (precision != 32 && precision != 64) return set_err(ctx, ISS_ERR_ARG, "precision must be 32 or 64");
    int rc = ensure_ktab(ctx);
    if (rc) return rc;
    if (layout != LAYOUT_ROWS && layout != LAYOUT_CDF) return set_err(ctx, ISS_ERR_ARG, "unknown yield layout");
    const bool rowsl = layout == LAYOUT_ROWS;
    if (rowsl && ycol) return set_err(ctx, ISS_ERR_ARG, "species-major yields come with the CDF layout");
    const int64_t Nc = ctx->Nc;
    const int Ns = ctx->Ns;
    const size_t esz = precision == 64 ? sizeof(double) : sizeof(float);
    const int stride = row_stride(Ns, static_cast<int>(esz));
    if (rowsl) HIPCHK(ctx->d_rows.ensure(std::max<int64_t>(1, Nc) * stride * esz));
    else HIPCHK(ctx->d_cdf.ensure(std::max<int64_t>(1, Nc) * cdf_stride(Ns)));
    HIPCHK(ctx->d_ctot.ensure(std::max<int64_t>(1, Nc)));
    HIPCHK(ctx->d_stot.ensure(Ns));
    // partials sized for the largest grid the launches use
    const int max_blk = ctx->num_cu * kPartialBlocksPerCU;
    HIPCHK(ctx->d_partial.ensure(static_cast<size_t>(max_blk) * Ns));
    // per-cell yields for the caller: the rows themselves, or a staging copy
    const bool raw = per_cell_host != nullptr && !rowsl;
    if (raw) HIPCHK(ctx->d_raw.ensure(std::max<int64_t>(1, Nc) * Ns * esz));
    YieldArgs a = base_yield_args(ctx, lcc);
    a.cdf = ctx->d_cdf.p;
    a.ctot = ctx->d_ctot.p;
    a.raw = raw ? static_cast<void*>(ctx->d_raw.p) : nullptr;
    a.partial = ctx->d_partial.p;
    a.ycol = ycol;
    a.rows = rowsl ? static_cast<void*>(ctx->d_rows.p) : nullptr;
    HIPCHK(ctx->d_oob.ensure(1));
    a.oob = ctx->d_oob.p;
    HIPCHK(hipMemsetAsync(ctx->d_oob.p, 0, sizeof(int), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    int nblk = 1;
    if (Nc > 0 && ctx->df.include_diffusion) {
        HIPCHK(ctx->d_dnq.ensure(static_cast<size_t>(Nc) * Ns));
        k_diffusion_dn<<<ctx->num_cu * 8, 256, 0, ctx->stream>>>(a, ctx->d_etab.p, ctx->d_dnq.p);
        HIPCHK(hipGetLastError());
        a.dnq = ctx->d_dnq.p;
    }
    // the pass, the species reduction and the read-backs, then one wait; a
    // Bessel argument outside the table (flag oob) reruns the pass with the
    // exact fallback
    HIPCHK(ctx->h_pin.ensure(PIN_TOT + sizeof(double) * Ns));
    int* h_oob = ctx->h_pin.at<int>(PIN_OOB);
    double* h_tot = ctx->h_pin.at<double>(PIN_TOT);
    for (int exact = 0; exact < 2; exact++) {
        if (Nc > 0) {
            int rc;
            // rows: at most 2^19 / Ns blocks, the exact-K rerun on the first
            // pass's grid (the grid fixes the order of the species sums; the
            // allocation-free entry point uses the same)
            const int max_rows_blk = kNoAllocPartials / Ns, force = exact ? nblk : 0;
            if (rowsl)
                rc = (precision == 64) ? launch_rows<double>(ctx, a, exact != 0, max_rows_blk, force)
                                       : launch_rows<float>(ctx, a, exact != 0, max_rows_blk, force);
            else
                rc = (precision == 64) ? launch_yield<double>(ctx, a, raw, exact != 0)
                                       : launch_yield<float>(ctx, a, raw, exact != 0);
            if (rc) return set_err(ctx, rc, "yield kernel launch failed");
            nblk = ctx->yield_grid;
        } else {
            HIPCHK(hipMemsetAsync(ctx->d_partial.p, 0, sizeof(double) * Ns, ctx->stream));
        }
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        k_species_reduce<<<Ns, 256, 0, ctx->stream>>>(ctx->d_partial.p, nblk, Ns, ctx->d_stot.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
        const bool check = Nc > 0 && !exact;
        if (check) HIPCHK(hipMemcpyAsync(h_oob, ctx->d_oob.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        if (totals_host)
            HIPCHK(hipMemcpyAsync(h_tot, ctx->d_stot.p, sizeof(double) * Ns, hipMemcpyDeviceToHost, ctx->stream));
        if (per_cell_host && Nc > 0) {
            if (rowsl)  // the rows without their padding
                HIPCHK(hipMemcpy2DAsync(per_cell_host, esz * Ns, ctx->d_rows.p, esz * stride, esz * Ns, Nc,
                                        hipMemcpyDeviceToHost, ctx->stream));
            else
                HIPCHK(hipMemcpyAsync(per_cell_host, ctx->d_raw.p, esz * Nc * Ns, hipMemcpyDeviceToHost, ctx->stream));
        }
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (totals_host) std::memcpy(totals_host, h_tot, sizeof(double) * Ns);
        if (!(check && *h_oob)) break;
    }
    float ms1 = 0.f, ms2 = 0.f;
    HIPCHK(hipEventElapsedTime(&ms1, ctx->ev[0], ctx->ev[1]));
    HIPCHK(hipEventElapsedTime(&ms2, ctx->ev[0], ctx->ev[2]));
    ctx->stats.ms_yield_kernel = ms1;
    ctx->stats.ms_yield = ms2;
    ctx->yields_valid = true;
    ctx->yield_precision = precision;
    ctx->yield_layout = layout;
    ctx->yields_lcc = lcc;
    return ISS_OK;
}


// surface, species and delta-f fields of SampleArgs (no sampling pass)
SampleArgs base_args(iss_ctx* ctx) {
    SampleArgs a{};
    a.cells = ctx->d_cells.p;
    a.coef = ctx->d_coef.p;
    a.Nc = ctx->Nc;
    a.Ns = ctx->Ns;
    a.sp = species_dev(ctx);
    a.pt = ctx->pt;
    a.shear = ctx->df.include_shear;
    a.bulk = ctx->df.include_bulk ? ctx->bulk_mode : BULK_NONE;
    a.neos = ctx->neos;
    a.diff = ctx->df.include_diffusion;
    a.max_trials = 5000;
    return a;
}

int mlp_pack(iss_ctx* ctx) {
    k_mlp_pack<<<(kNumFrags + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_mlp_par.p, ctx->d_mlp_frag.p);
    HIPCHK(hipGetLastError());
    return ISS_OK;
}

int mlp_alloc(iss_ctx* ctx) {
    HIPCHK(ctx->d_mlp_par.ensure(kNumParams));
    HIPCHK(ctx->d_adam_m.ensure(kNumParams));
    HIPCHK(ctx->d_adam_v.ensure(kNumParams));
    HIPCHK(ctx->d_grad.ensure(kNumParams));
    HIPCHK(ctx->d_mlp_frag.ensure(kNumFrags));
    return ISS_OK;
}

int mlp_upload(iss_ctx* ctx, const float* params) {
    int rc = mlp_alloc(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->d_mlp_par.p, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_adam_m.p, 0, sizeof(float) * kNumParams, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_adam_v.p, 0, sizeof(float) * kNumParams, ctx->stream));
    rc = mlp_pack(ctx);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->adam_t = 0;
    ctx->mlp_ready = true;
    return ISS_OK;
}

int check_samples(iss_ctx* ctx, int64_t n, const int32_t* cell, const int32_t* species) {
    if (n < 0 || (n > 0 && (!cell || !species))) return set_err(ctx, ISS_ERR_ARG, "bad sample arrays");
    if (ctx->Nc <= 0 || ctx->Ns <= 0 || !ctx->tables_ready)
        return set_err(ctx, ISS_ERR_STATE, "surface, species and delta f must be set");
    for (int64_t i = 0; i < n; i++)
        if (cell[i] < 0 || cell[i] >= ctx->Nc || species[i] < 0 || species[i] >= ctx->Ns)
            return set_err(ctx, ISS_ERR_ARG, "cell or species index out of range");
    return ISS_OK;
}

// one training step (or a gradient evaluation) over nb * 64 samples
int mlp_step(iss_ctx* ctx, TrainArgs& t, int nb, float lr, float* grad_out) {
    HIPCHK(ctx->d_tpart.ensure(static_cast<size_t>(nb) * kNumParams));
    HIPCHK(ctx->d_tloss.ensure(2 * static_cast<size_t>(nb)));
    t.partial = ctx->d_tpart.p;
    t.loss_part = ctx->d_tloss.p;
    HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_train),
                               hipFuncAttributeMaxDynamicSharedMemorySize, kTrainLds));
    k_mlp_train<<<nb, 64, kTrainLds, ctx->stream>>>(t);
    HIPCHK(hipGetLastError());
    if (lr > 0.f) ctx->adam_t++;
    k_mlp_reduce_adam<<<(kNumParams + 255) / 256, 256, 0, ctx->stream>>>(
        ctx->d_tpart.p, nb, ctx->d_tloss.p, ctx->d_mlp_par.p, ctx->d_adam_m.p, ctx->d_adam_v.p, grad_out, lr,
        ctx->adam_t);
    HIPCHK(hipGetLastError());
    if (lr > 0.f) return mlp_pack(ctx);
    return ISS_OK;
}

double mean_loss(iss_ctx* ctx, int nb) {
    std::vector<float> h(2 * static_cast<size_t>(nb));
    if (hipMemcpyAsync(h.data(), ctx->d_tloss.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return 0.;
    double ls = 0., nv = 0.;
    for (int b = 0; b < nb; b++) { ls += h[b]; nv += h[nb + b]; }
    return nv > 0. ? ls / nv : 0.;
}

template <typename R>
int run_spectra(iss_ctx* ctx, const iss_spectra_opts* o, const int32_t* species, int ns, const double* pT, int npT,
                const double* phi, int nphi, const double* eta, const double* w, int ny, double* out) {
    const int64_t Nk = ctx->Nc, Nd = o->all_cells ? ctx->Nneg : 0, N = Nk + Nd;
    const int64_t npl = static_cast<int64_t>(npT) * nphi;
    const int64_t npts = npl * (o->mode == 1 ? ny : 1);
    const int64_t nout = ns * npts;
    if (npts > (int64_t{1} << 30)) return set_err(ctx, ISS_ERR_ARG, "spectra grid too large");
    const int64_t wps = (npts + 64 * kSpecPts - 1) / (64 * kSpecPts);
    if (ns * wps > 65535) return set_err(ctx, ISS_ERR_ARG, "spectra: species x grid points exceed one launch");
    ctx->spectra_ms = 0.;
    if (N == 0) {
        std::fill(out, out + nout, 0.);
        return ISS_OK;
    }
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    // per-cell records: the sampled cells, then the dropped ones
    HIPCHK(ctx->d_srec.ensure(sizeof(double) * kSpecRec * N));
    double* rec = reinterpret_cast<double*>(ctx->d_srec.p);
    const int shear = ctx->df.include_shear;
    const int bulk = ctx->df.include_bulk ? ctx->bulk_mode : BULK_NONE;
    if (Nk > 0)
        k_spectra_rec<<<(Nk + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_cells.p, ctx->d_coef.p, Nk, shear, bulk,
                                                                   ctx->neos, ctx->df.include_diffusion, rec);
    if (Nd > 0)
        k_spectra_rec<<<(Nd + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_cells_neg.p, ctx->d_coef_neg.p, Nd, shear,
                                                                   bulk, ctx->neos, ctx->df.include_diffusion,
                                                                   rec + Nk * kSpecRec);
    HIPCHK(hipGetLastError());
    // grids: pT, cos phi, sin phi, cosh, sinh, weight of the y - eta_s nodes
    // (mode 0: y - eta_s = -eta[k]) or of the rapidities (mode 1)
    std::vector<double> g(static_cast<size_t>(npT) + 2 * nphi + 3 * ny);
    double* gp = g.data();
    for (int i = 0; i < npT; i++) gp[i] = pT[i];
    for (int j = 0; j < nphi; j++) {
        gp[npT + j] = std::cos(phi[j]);
        gp[npT + nphi + j] = std::sin(phi[j]);
    }
    const int oy = npT + 2 * nphi;
    for (int k = 0; k < ny; k++) {
        const double y = o->mode == 1 ? eta[k] : -eta[k];
        gp[oy + k] = std::cosh(y);
        gp[oy + ny + k] = std::sinh(y);
        gp[oy + 2 * ny + k] = o->mode == 1 ? 1. : w[k];
    }
    HIPCHK(ctx->d_sgrid.ensure(sizeof(double) * g.size()));
    HIPCHK(ctx->d_sspec.ensure(ns));
    HIPCHK(hipMemcpyAsync(ctx->d_sgrid.p, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_sspec.p, species, sizeof(int32_t) * ns, hipMemcpyHostToDevice, ctx->stream));
    const double* dg = reinterpret_cast<const double*>(ctx->d_sgrid.p);
    // cell chunks: about 32 waves per CU over all (chunk, point-wave) pairs
    const int64_t pw = ns * wps;
    const int64_t target = static_cast<int64_t>(ctx->num_cu) * 32;
    int64_t nch = std::max<int64_t>(1, std::min<int64_t>((target + pw - 1) / pw, (N + 7) / 8));
    const int64_t cpc = (N + nch - 1) / nch;
    nch = (N + cpc - 1) / cpc;
    HIPCHK(ctx->d_spart.ensure(nch * nout));
    HIPCHK(ctx->d_sout.ensure(nout));
    SpectraArgs a{};
    a.rec = rec;
    a.N = N;
    a.cells_per_chunk = cpc;
    a.sp = species_dev(ctx);
    a.species = ctx->d_sspec.p;
    a.ns = ns;
    a.pT = dg;
    a.npT = npT;
    a.cphi = dg + npT;
    a.sphi = dg + npT + nphi;
    a.nphi = nphi;
    a.ych = dg + oy;
    a.ysh = dg + oy + ny;
    a.yw = dg + oy + 2 * ny;
    a.ny = ny;
    a.mode = o->mode;
    a.positive_only = o->use_pos_dN_only != 0;
    a.restrict_ratio = o->restrict_deltaf ? o->deltaf_max_ratio : 0.;
    a.waves_per_species = static_cast<int>(wps);
    a.partial = ctx->d_spart.p;
    k_spectra<R><<<dim3(static_cast<unsigned>(nch), static_cast<unsigned>(pw)), 64, 0, ctx->stream>>>(a);
    HIPCHK(hipGetLastError());
    k_spectra_reduce<<<(nout + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_spart.p, nch, nout, ctx->d_sout.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->d_sout.p, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->spectra_ms = ms;
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
    ctx->scratch = scratch_address(device, &ctx->scratch_err);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
    if (ctx->d_stats.ensure(kStatSlots) != hipSuccess || ctx->d_fb_count.ensure(1) != hipSuccess) {
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
    ctx->d_mass.release(); ctx->d_spk.release(); ctx->d_ints.release(); ctx->d_k11.release(); ctx->d_neos.release();
    ctx->d_ktab.release(); ctx->d_ktabf.release(); ctx->d_ptab.release();
    ctx->d_cells.release(); ctx->d_tmp.release(); ctx->d_coef.release(); ctx->d_u32.release();
    ctx->d_i64.release(); ctx->d_scan.release(); ctx->d_eos.release(); ctx->d_cdf.release();
    ctx->d_ctot.release(); ctx->d_partial.release(); ctx->d_oob.release(); ctx->d_stot.release(); ctx->d_raw.release();
    ctx->d_counts.release(); ctx->d_evcount.release(); ctx->d_off.release(); ctx->d_evoff.release();
    ctx->d_tree0.release(); ctx->d_tree1.release(); ctx->d_out.release(); ctx->d_setup.release(); ctx->d_stats.release();
    ctx->d_cellidx.release(); ctx->d_ecell.release();
    ctx->d_ycol.release(); ctx->d_mcnt.release(); ctx->d_msegoff.release(); ctx->d_msp.release();
    ctx->d_anti.release(); ctx->d_lcc_flag.release(); ctx->d_lcc_g.release(); ctx->d_lcc_src.release();
    ctx->d_lcc_evoff.release(); ctx->d_lcc_rec.release(); ctx->d_out2.release();
    ctx->d_aux.release(); ctx->d_mlp_par.release(); ctx->d_adam_m.release(); ctx->d_adam_v.release();
    ctx->d_grad.release(); ctx->d_mlp_frag.release(); ctx->d_fb_list.release(); ctx->d_fb_count.release(); ctx->d_lrec.release();
    ctx->d_pkey.release(); ctx->d_lbits.release(); ctx->d_lcnt.release(); ctx->d_pairoff.release();
    ctx->d_lscnt.release(); ctx->d_soff.release();
    ctx->d_sp16.release(); ctx->d_lslot.release();
    ctx->d_tpart.release(); ctx->d_tloss.release(); ctx->d_tout.release(); ctx->d_tr_cell.release();
    ctx->d_ev_cell.release(); ctx->d_tr_spec.release(); ctx->d_ev_spec.release(); ctx->d_i32a.release();
    ctx->d_i32b.release(); ctx->d_f32a.release(); ctx->d_f32b.release(); ctx->d_f32c.release();
    ctx->d_cells_neg.release(); ctx->d_coef_neg.release(); ctx->d_srec.release(); ctx->d_sgrid.release();
    ctx->d_spart.release(); ctx->d_sout.release(); ctx->d_sspec.release();
    ctx->d_surf_off.release(); ctx->d_bevents.release(); ctx->d_bround.release(); ctx->d_bevcount.release();
    ctx->d_bbase.release(); ctx->d_perms.release(); ctx->d_seg.release(); ctx->d_lam.release();
    for (auto& g : ctx->dgen) g->release();
    ctx->dgen.clear();
    ctx->d_dec_int.release(); ctx->d_dec_dbl.release(); ctx->d_qmu.release(); ctx->d_etab.release(); ctx->d_dnq.release();
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
    // yield-pass constants: pref = g / (2 pi^2 (hbar c)^3), formed in each precision as the FP64 fallback
    // and the FP32 series use it
    std::vector<SpeciesK> spk(Ns);
    const double uf = 1.0 / (kHbarC * kHbarC * kHbarC) / (2. * kPi * kPi);
    for (int i = 0; i < Ns; i++) {
        SpeciesK& k = spk[i];
        k.m = static_cast<float>(sp->mass[i]);
        k.pref = static_cast<float>(sp->degeneracy[i]) * static_cast<float>(uf);
        k.B = static_cast<float>(sp->baryon[i]);
        k.S = static_cast<float>(sp->strange[i]);
        k.Q = static_cast<float>(sp->charge[i]);
        k.sgn = static_cast<float>(sp->sign[i]);
        k.md = sp->mass[i];
        k.prefd = static_cast<double>(sp->degeneracy[i]) * uf;
        k.m2 = static_cast<float>(sp->mass[i] * sp->mass[i]);
        k.m3 = static_cast<float>(sp->mass[i] * sp->mass[i] * sp->mass[i]);
        k.m4 = static_cast<float>(sp->mass[i] * sp->mass[i] * sp->mass[i] * sp->mass[i]);
        k.g = static_cast<float>(sp->degeneracy[i]);
        k.Bi = sp->baryon[i];
        k.Si = sp->strange[i];
        k.Qi = sp->charge[i];
        k.pad = 0;
        k.md20 = 20. * sp->mass[i];
    }
    HIPCHK(ctx->d_mass.ensure(Ns));
    HIPCHK(ctx->d_ints.ensure(6 * Ns));
    HIPCHK(ctx->d_spk.ensure(Ns));
    HIPCHK(hipMemcpyAsync(ctx->d_mass.p, sp->mass, sizeof(double) * Ns, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_ints.p, ints.data(), sizeof(int) * 6 * Ns, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_spk.p, spk.data(), sizeof(SpeciesK) * Ns, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->Ns = Ns;
    // the leading species below 0.7 GeV take the 10-term series (the row
    // pass gives them lanes of their own; lighter species further down the
    // list, if any, are still summed correctly, by the one-term lanes' loop)
    ctx->nlight = 0;
    while (ctx->nlight < std::min(Ns, kLightMax) && sp->mass[ctx->nlight] < 0.7) ctx->nlight++;
    ctx->light_beyond = false;
    for (int i = ctx->nlight; i < Ns; i++)
        if (sp->mass[i] < 0.7) ctx->light_beyond = true;
    // the row pass's padded one-term table: all species, then the same with
    // the negative species suppressed (local charge conservation)
    {
        const int st = row_stride(Ns, 4);
        std::vector<RowSp> rs(2 * static_cast<size_t>(st));
        ctx->charges_outside = false;
        for (int v = 0; v < 2; v++)
            for (int i = 0; i < st; i++) {
                RowSp& r = rs[static_cast<size_t>(v) * st + i];
                const double m = i < Ns ? sp->mass[i] : 1.0;
                const int B = i < Ns ? sp->baryon[i] : 0, S = i < Ns ? sp->strange[i] : 0,
                          Q = i < Ns ? sp->charge[i] : 0;
                r.md20 = 20. * m;
                r.m = static_cast<float>(m);
                r.m2 = static_cast<float>(m * m);
                r.m3 = static_cast<float>(m * m * m);
                r.m4 = static_cast<float>(m * m * m * m);
                r.m2B = static_cast<float>(m * m * B);
                const bool off = i >= Ns || (v == 1 && Q < 0);
                r.g = off ? 0.f : static_cast<float>(sp->degeneracy[i]);
                if (B < -1 || B > 1 || S < -3 || S > 3 || Q < -2 || Q > 2) {
                    ctx->charges_outside = true;
                    r.frow = 1 | (6 << 8) | (12 << 16);
                } else {
                    r.frow = (B + 1) | ((S + 6) << 8) | ((Q + 12) << 16);
                }
                r.sgn = i < Ns ? static_cast<float>(sp->sign[i]) : 0.f;
            }
        HIPCHK(ctx->d_rowsp.ensure(rs.size()));
        HIPCHK(hipMemcpy(ctx->d_rowsp.p, rs.data(), sizeof(RowSp) * rs.size(), hipMemcpyHostToDevice));
    }
    ctx->h_pid.assign(sp->pid, sp->pid + Ns);
    ctx->h_charge.assign(sp->charge, sp->charge + Ns);
    // antiparticle entries (local charge conservation)
    ctx->h_anti.assign(Ns, -1);
    for (int i = 0; i < Ns; i++)
        for (int j = 0; j < Ns; j++)
            if (sp->pid[j] == -sp->pid[i]) {
                ctx->h_anti[i] = j;
                break;
            }
    HIPCHK(ctx->d_anti.ensure(Ns));
    HIPCHK(hipMemcpy(ctx->d_anti.p, ctx->h_anti.data(), sizeof(int) * Ns, hipMemcpyHostToDevice));
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
    if (d->include_diffusion) {
        if (!d->qmu_table) return set_err(ctx, ISS_ERR_ARG, "diffusion delta f needs its coefficient table");
        HIPCHK(ctx->d_qmu.ensure(150 * 100));
        HIPCHK(hipMemcpyAsync(ctx->d_qmu.p, d->qmu_table, sizeof(double) * 150 * 100, hipMemcpyHostToDevice,
                              ctx->stream));
        if (!ctx->d_etab.p) {
            HIPCHK(ctx->d_etab.ensure(static_cast<size_t>(kEnTerms) * kKLen));
            k_expint_table<<<(kKLen + 127) / 128, 128, 0, ctx->stream>>>(ctx->d_etab.p);
            HIPCHK(hipGetLastError());
        }
    }
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
    ctx->df.qmu_table = nullptr;
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
    ctx->Nneg = 0;
    ctx->S = 0;
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
    if (n > nk) {
        HIPCHK(ctx->d_cells_neg.ensure(kCellFields * (n - nk)));
        k_compact_dropped<<<nb, 256, 0, ctx->stream>>>(ctx->d_tmp.p, ctx->d_u32.p, ctx->d_i64.p, n, n - nk,
                                                       ctx->d_cells_neg.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    lab.release();
    ctx->Nc = nk;
    ctx->Nneg = n - nk;
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
    ctx->Nneg = 0;
    ctx->S = 0;
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
    // (one surface and batched surfaces alike are sampled from the yield rows)
    return run_yields(ctx, opts->precision, per_cell, species_totals, 0, nullptr, LAYOUT_ROWS);
}

int iss_compute_yields_noalloc(iss_ctx* ctx, const iss_yield_opts* opts, double* species_totals, void* per_cell,
                               int64_t per_cell_capacity) {
    if (!ctx || !opts || !species_totals || !per_cell) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (!ctx->scratch) return set_err(ctx, ISS_ERR_HIP, "no scratch on this device: " + ctx->scratch_err);
    if (ctx->Ns == 0) return set_err(ctx, ISS_ERR_STATE, "species not set");
    const int precision = opts->precision;
    if (precision != 32 && precision != 64) return set_err(ctx, ISS_ERR_ARG, "precision must be 32 or 64");
    if (ctx->ktab_bulk != 1)
        return set_err(ctx, ISS_ERR_STATE,
                       "Bessel tables not built: the first iss_compute_yields or iss_sample of the context builds them");
    if (ctx->df.include_diffusion)
        return set_err(ctx, ISS_ERR_UNSUPPORTED,
                       "diffusion needs a per-(cell, species) buffer: use iss_compute_yields");
    const int Ns = ctx->Ns;
    const int64_t Nc = ctx->Nc;
    if (Ns > kNoAllocMaxSpecies)
        return set_err(ctx, ISS_ERR_RANGE,
                       "more than " + std::to_string(kNoAllocMaxSpecies) + " species: beyond the static scratch");
    const int esz = precision == 64 ? 8 : 4;
    const int stride = row_stride(Ns, esz);
    if (per_cell_capacity < Nc * stride)
        return set_err(ctx, ISS_ERR_RANGE,
                       "per_cell holds fewer than n_cells x " + std::to_string(stride) + " entries");
    if (reinterpret_cast<uintptr_t>(per_cell) % 16 != 0)
        return set_err(ctx, ISS_ERR_ARG, "per_cell must be 16-byte aligned");
    int* flag = reinterpret_cast<int*>(ctx->scratch);
    double* partial = ctx->scratch + kNoAllocHeader;
    YieldArgs a = base_yield_args(ctx, 0);
    a.rows = per_cell;
    a.partial = partial;
    a.oob = flag;
    k_clear_flag<<<1, 64, 0, ctx->stream>>>(flag);
    HIPCHK(hipGetLastError());
    int nblk = 0;
    if (Nc > 0) {
        // the first pass, then the exact-K pass on the same grid, which
        // returns at once on the device unless the first pass set the flag
        for (int exact = 0; exact < 2; exact++) {
            if (exact) a.gate = flag;
            const int rc = precision == 64 ? launch_rows<double>(ctx, a, exact != 0, kNoAllocPartials / Ns, nblk)
                                           : launch_rows<float>(ctx, a, exact != 0, kNoAllocPartials / Ns, nblk);
            if (rc) return set_err(ctx, rc, "yield kernel launch failed");
            nblk = ctx->yield_grid;
        }
    }
    k_species_reduce<<<Ns, 256, 0, ctx->stream>>>(partial, nblk, Ns, species_totals);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

// local charge conservation needs every positive species' antiparticle
static int lcc_check(iss_ctx* ctx) {
    for (int s = 0; s < ctx->Ns; s++)
        if (ctx->h_charge[s] > 0 && ctx->h_anti[s] < 0)
            return set_err(ctx, ISS_ERR_ARG,
                           "local charge conservation: the antiparticle of pid " + std::to_string(ctx->h_pid[s]) +
                               " is not in the species table");
    return ISS_OK;
}

// partners of the positive hadrons (src/FSSW.cpp:1035-1048): each positive
// hadron of the pass gets its antiparticle, from the same cell, in the slot
// after it.  a: the pass's arguments (a.out = d_out, a.K hadrons); mrec: the
// multiplicity models' set-up records, or null for the Poisson pass.
static int lcc_partners(iss_ctx* ctx, const SampleArgs& a, int E, const HSetup* mrec, int per_cu_plain,
                        int64_t* Kio) {
    const int64_t K = *Kio;
    HIPCHK(ctx->d_lcc_flag.ensure(K));
    HIPCHK(ctx->d_lcc_g.ensure(K + 1));
    HIPCHK(ctx->d_lcc_src.ensure(K));
    HIPCHK(ctx->d_lcc_evoff.ensure(E + 1));
    if (mrec)
        k_lcc_flag_rec<<<(K + 255) / 256, 256, 0, ctx->stream>>>(mrec, K, a.sp.Q, ctx->d_lcc_flag.p,
                                                                 ctx->d_lcc_src.p);
    else
        k_lcc_flag<<<(K + 255) / 256, 256, 0, ctx->stream>>>(a, ctx->d_lcc_flag.p, ctx->d_lcc_src.p);
    HIPCHK(hipGetLastError());
    int rc = launch_scan(ctx, ctx->d_lcc_flag.p, K, ctx->d_lcc_g.p);
    if (rc) return rc;
    int64_t P = 0;
    HIPCHK(hipMemcpyAsync(&P, ctx->d_lcc_g.p + K, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (K + P >= (1ll << 31)) return set_err(ctx, ISS_ERR_RANGE, "more than 2^31 hadrons in one pass");
    HIPCHK(ctx->d_out2.ensure(K + P));
    HIPCHK(ctx->d_lcc_rec.ensure(std::max<int64_t>(P, 1)));
    k_lcc_place<<<(K + 255) / 256, 256, 0, ctx->stream>>>(a, ctx->d_lcc_flag.p, ctx->d_lcc_g.p, ctx->d_lcc_src.p,
                                                          ctx->d_out.p, ctx->d_out2.p, ctx->d_lcc_rec.p, mrec);
    HIPCHK(hipGetLastError());
    if (P > 0) {
        SampleArgs ap = a;
        ap.out = ctx->d_out2.p;
        ap.K = P;
        ap.Kdev = nullptr;
        const int nblk = static_cast<int>(
            std::max<int64_t>(1, std::min<int64_t>((P + kTrialThreads - 1) / kTrialThreads, int64_t(ctx->num_cu) * per_cu_plain)));
        k_trials<<<nblk, kTrialThreads, 0, ctx->stream>>>(ap, ctx->d_lcc_rec.p);
        HIPCHK(hipGetLastError());
    }
    k_lcc_events<<<(E + 1 + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_evoff.p, E, ctx->d_lcc_g.p, ctx->d_lcc_evoff.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->d_evoff.p, ctx->d_lcc_evoff.p, sizeof(int64_t) * (E + 1), hipMemcpyDeviceToDevice,
                          ctx->stream));
    std::swap(ctx->d_out.p, ctx->d_out2.p);
    std::swap(ctx->d_out.cap, ctx->d_out2.cap);
    *Kio = K + P;
    ctx->last_plain = false;  // output positions no longer match the set-up records
    return ISS_OK;
}

// dN_dy_sampling_model 1, 10, 20 (src/FSSW.cpp:250-309): per (event,
// species) multiplicities, then each hadron's cell from its species' cell
// prefix (src/FSSW.cpp:920-979), then the plain momentum sampler
static int sample_model(iss_ctx* ctx, const iss_sample_opts* o, int64_t* n_hadrons) {
    const int model = o->sampling_model;
    if (model != 1 && model != 10 && model != 20)
        return set_err(ctx, ISS_ERR_UNSUPPORTED, "dN_dy_sampling_model must be 1, 10, 20 or 30");
    if (o->proposal != 0)
        return set_err(ctx, ISS_ERR_UNSUPPORTED, "the learned proposal samples with dN_dy_sampling_model = 30");
    const int lcc = o->local_charge_conservation ? 1 : 0;
    if (lcc) {
        int rc = lcc_check(ctx);
        if (rc) return rc;
    }
    if ((model == 10 || model == 20) && !(o->sampling_para1 > 0.))
        return set_err(ctx, ISS_ERR_ARG, "dN_dy_sampling_para1 must be positive for the negative binomial models");
    invalidate_output(ctx);
    if (o->max_trials <= 1) return set_err(ctx, ISS_ERR_ARG, "max_trials must exceed 1");
    const int64_t Nc = ctx->Nc;
    const int Ns = ctx->Ns;
    const int E = o->n_events;
    // yields with their species-major copy, scanned along the cells
    HIPCHK(ctx->d_ycol.ensure(std::max<int64_t>(Nc, 1) * Ns));
    int rc = run_yields(ctx, o->precision == 64 ? 64 : 32, nullptr, nullptr, lcc, ctx->d_ycol.p, LAYOUT_CDF);
    if (rc) return rc;
    const bool binv = o->hydro_mode != 2;
    const double mult = binv ? (o->y_RB - o->y_LB) : 1.0;
    const uint32_t chunk = static_cast<uint32_t>(o->event0);
    const int64_t nseg = static_cast<int64_t>(E) * Ns;
    HIPCHK(ctx->d_mcnt.ensure(nseg));
    HIPCHK(ctx->d_msegoff.ensure(nseg + 1));
    HIPCHK(ctx->d_evoff.ensure(E + 1));
    HIPCHK(hipMemsetAsync(ctx->d_stats.p, 0, kStatSlots * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if (Nc > 0) k_row_scan<<<Ns, 256, 0, ctx->stream>>>(ctx->d_ycol.p, Nc);
    k_model_counts<<<(nseg + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_stot.p, Ns, E, mult, model, o->sampling_para1,
                                                                o->seed, chunk, ctx->d_mcnt.p);
    HIPCHK(hipGetLastError());
    rc = launch_scan(ctx, ctx->d_mcnt.p, nseg, ctx->d_msegoff.p);
    if (rc) return rc;
    k_model_evoff<<<(E + 1 + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_msegoff.p, E, Ns, ctx->d_evoff.p);
    int64_t K = 0;
    HIPCHK(hipMemcpyAsync(&K, ctx->d_msegoff.p + nseg, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (K >= (1ll << 31)) return set_err(ctx, ISS_ERR_RANGE, "more than 2^31 hadrons in one pass");
    if (K > 0 && Nc == 0) return set_err(ctx, ISS_ERR_STATE, "hadrons without cells");
    HIPCHK(ctx->d_msp.ensure(std::max<int64_t>(K, 1)));
    HIPCHK(ctx->d_out.ensure(std::max<int64_t>(K, 1)));
    HIPCHK(ctx->d_setup.ensure(std::max<int64_t>(K, 1)));
    k_model_fill<<<(nseg + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_msegoff.p, nseg, Ns, ctx->d_msp.p);
    SampleArgs a = base_args(ctx);
    a.evoff = ctx->d_evoff.p;
    a.E = E;
    a.K = K;
    a.seed = o->seed;
    a.chunk = chunk;
    a.boost_invariant = binv ? 1 : 0;
    a.ecell = binv ? nullptr : ctx->d_ecell.p;  // 3+1D: the cells' flow and position words
    a.yLB = static_cast<float>(o->y_LB);
    a.yRB = static_cast<float>(o->y_RB);
    a.max_trials = o->max_trials;
    a.refill_min = ctx->refill_min[0];
    a.out = ctx->d_out.p;
    a.stats = ctx->d_stats.p;
    a.anti = ctx->d_anti.p;
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    if (K > 0) {
        k_model_setup<<<(K + 255) / 256, 256, 0, ctx->stream>>>(a, ctx->d_msp.p, ctx->d_ycol.p, ctx->d_setup.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trials, kTrialThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (ctx->blocks_per_cu[0] > 0) per_cu = std::min(ctx->blocks_per_cu[0], per_cu);
    if (K > 0) {
        const int nblk = static_cast<int>(std::max<int64_t>(
            1, std::min<int64_t>((K + kTrialThreads - 1) / kTrialThreads, int64_t(ctx->num_cu) * per_cu)));
        k_trials<<<nblk, kTrialThreads, 0, ctx->stream>>>(a, ctx->d_setup.p);
        HIPCHK(hipGetLastError());
    }
    ctx->last_plain = true;  // the records' output slots index d_out
    if (lcc && K > 0) {
        rc = lcc_partners(ctx, a, E, ctx->d_setup.p, per_cu, &K);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    unsigned long long st[6];
    HIPCHK(hipMemcpyAsync(st, ctx->d_stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms_all = 0.f, ms_k = 0.f, ms_su = 0.f, ms_tr = 0.f;
    HIPCHK(hipEventElapsedTime(&ms_all, ctx->ev[3], ctx->ev[5]));
    HIPCHK(hipEventElapsedTime(&ms_k, ctx->ev[4], ctx->ev[5]));
    HIPCHK(hipEventElapsedTime(&ms_su, ctx->ev[4], ctx->ev[6]));
    HIPCHK(hipEventElapsedTime(&ms_tr, ctx->ev[6], ctx->ev[5]));
    ctx->stats.ms_setup_kernel = ms_su;
    ctx->stats.ms_trials_kernel = ms_tr;
    ctx->stats.ms_sample = ms_all;
    ctx->stats.ms_sample_kernel = ms_k;
    ctx->K = K;
    ctx->E = E;
    ctx->stats.n_hadrons = K;
    ctx->stats.n_trials = static_cast<int64_t>(st[0]);
    ctx->stats.n_inner = static_cast<int64_t>(st[1]);
    ctx->stats.n_failed = static_cast<int64_t>(st[2]);
    ctx->stats.n_range_errors = static_cast<int64_t>(st[3]);
    ctx->stats.n_fallback = 0;
    ctx->stats.n_violations = 0;
    ctx->stats.n_routed = 0;
    clear_routed_stats(ctx);
    if (n_hadrons) *n_hadrons = K;
    if (st[3]) return set_err(ctx, ISS_ERR_RANGE, "m/T - mu/T outside the momentum-sampler tables");
    return ISS_OK;
}

int iss_sample(iss_ctx* ctx, const iss_sample_opts* o, int64_t* n_hadrons) {
    if (!ctx || !o) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (o->n_events <= 0) return set_err(ctx, ISS_ERR_ARG, "n_events must be positive");
    if (o->sampling_model != 30) return sample_model(ctx, o, n_hadrons);
    if (o->proposal != 0 && o->proposal != 1) return set_err(ctx, ISS_ERR_ARG, "proposal must be 0 or 1");
    if (o->proposal == 1 && !ctx->mlp_ready)
        return set_err(ctx, ISS_ERR_STATE, "learned proposal: no network parameters (init, set or train)");
    if (o->max_trials <= 1) return set_err(ctx, ISS_ERR_ARG, "max_trials must exceed 1");
    invalidate_output(ctx);
    const int lcc = o->local_charge_conservation ? 1 : 0;
    if (lcc) {
        int rc = lcc_check(ctx);
        if (rc) return rc;
    }
    if (!ctx->yields_valid || ctx->yields_lcc != lcc || ctx->yield_layout != LAYOUT_ROWS) {
        int rc = run_yields(ctx, o->precision == 64 ? 64 : 32, nullptr, nullptr, lcc, nullptr, LAYOUT_ROWS);
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
    HIPCHK(hipMemsetAsync(ctx->d_stats.p, 0, kStatSlots * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if (Nc > 0)
        k_counts<<<(Nc + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_ctot.p, Nc, mult, o->seed, chunk,
                                                            ctx->d_counts.p);
    HIPCHK(hipGetLastError());
    int rc = launch_scan(ctx, ctx->d_counts.p, Nc, ctx->d_off.p);
    if (rc) return rc;
    // multinomial over events, root = K: per-slot uniform events and
    // histograms up to kEventHistMax events, else the binomial tree, one
    // launch per level
    if (E <= kEventHistMax) {
        HIPCHK(hipMemsetAsync(ctx->d_evcount.p, 0, sizeof(uint32_t) * E, ctx->stream));
        k_event_hist<<<kEventHistBlocks, 256, 0, ctx->stream>>>(ctx->d_off.p + Nc, E, o->seed, chunk,
                                                                ctx->d_evcount.p);
    } else {
        HIPCHK(hipMemcpyAsync(ctx->d_tree0.p, ctx->d_off.p + Nc, sizeof(int64_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
        int64_t* cur = ctx->d_tree0.p;
        int64_t* nxt = ctx->d_tree1.p;
        for (int d = 0; d < D; d++) {
            const int64_t nodes = 1ll << d;
            k_multinomial_level<<<(nodes + 255) / 256, 256, 0, ctx->stream>>>(cur, nxt, d, E, o->seed, chunk);
            std::swap(cur, nxt);
        }
        k_multinomial_leaves<<<((1ll << D) + 255) / 256, 256, 0, ctx->stream>>>(cur, D, E, ctx->d_evcount.p);
    }
    HIPCHK(hipGetLastError());
    rc = launch_scan(ctx, ctx->d_evcount.p, E, ctx->d_evoff.p);
    if (rc) return rc;
    HIPCHK(ctx->h_pin.ensure(PIN_TOT + sizeof(double) * ctx->Ns));
    HIPCHK(hipMemcpyAsync(ctx->h_pin.at<int64_t>(PIN_K), ctx->d_off.p + Nc, sizeof(int64_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    int64_t K = *ctx->h_pin.at<int64_t>(PIN_K);
    HIPCHK(ctx->d_out.ensure(std::max<int64_t>(K, 1)));
    SampleArgs a{};
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
    a.diff = ctx->df.include_diffusion;
    a.boost_invariant = binv ? 1 : 0;
    a.ecell = binv ? nullptr : ctx->d_ecell.p;  // 3+1D: the cells' flow and position words
    a.yLB = static_cast<float>(o->y_LB);
    a.yRB = static_cast<float>(o->y_RB);
    a.max_trials = o->max_trials;
    // (the learned proposal with every pair routed runs the plain pass)
    const bool learned = o->proposal == 1 && ctx->learn_route < kLearnRouteAll;
    a.refill_min = ctx->refill_min[learned ? 1 : 0];
    a.out = ctx->d_out.p;
    a.stats = ctx->d_stats.p;
    a.Kdev = nullptr;
    a.anti = ctx->d_anti.p;
    // each slot's cell and species from its cell's yield row; the learned
    // pass also gets each cell's pair bitmap and pair count
    const bool route = learned && ctx->learn_route > 1;
    const int Wb = (ctx->Ns + 31) / 32;
    HIPCHK(ctx->d_cellidx.ensure(std::max<int64_t>(K, 1)));
    HIPCHK(ctx->d_sp16.ensure(std::max<int64_t>(K, 1)));
    if (learned) {
        HIPCHK(ctx->d_lbits.ensure(std::max<int64_t>(Nc, 1) * Wb));
        HIPCHK(ctx->d_lcnt.ensure(std::max<int64_t>(Nc, 1)));
        HIPCHK(ctx->d_lscnt.ensure(std::max<int64_t>(Nc, 1)));
        HIPCHK(ctx->d_soff.ensure(Nc + 1));
    }
    a.cellidx = ctx->d_cellidx.p;
    a.slot_species = ctx->d_sp16.p;
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    if (K > 0) {
        const int pblk = static_cast<int>((Nc + kPickCells - 1) / kPickCells);
        const bool f64 = ctx->yield_precision == 64;
        const void* rows = ctx->d_rows.p;
        uint32_t* gb = learned ? ctx->d_lbits.p : nullptr;
        uint32_t* gc = learned ? ctx->d_lcnt.p : nullptr;
        uint32_t* gs = route ? ctx->d_lscnt.p : nullptr;
        if (f64 && learned)
            k_pick<double, true><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, true), ctx->stream>>>(a, static_cast<const double*>(rows), row_stride(ctx->Ns, 8), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, ctx->learn_route);
        else if (f64)
            k_pick<double, false><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, false), ctx->stream>>>(a, static_cast<const double*>(rows), row_stride(ctx->Ns, 8), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, ctx->learn_route);
        else if (learned)
            k_pick<float, true><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, true), ctx->stream>>>(a, static_cast<const float*>(rows), row_stride(ctx->Ns, 4), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, ctx->learn_route);
        else
            k_pick<float, false><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, false), ctx->stream>>>(a, static_cast<const float*>(rows), row_stride(ctx->Ns, 4), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, ctx->learn_route);
        HIPCHK(hipGetLastError());
    }
    int per_cu_trials = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_trials, k_trials, kTrialThreads, 0) != hipSuccess ||
        per_cu_trials < 1)
        per_cu_trials = 1;
    const int per_cu_plain = ctx->blocks_per_cu[0] > 0 ? std::min(ctx->blocks_per_cu[0], per_cu_trials) : per_cu_trials;
    HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
    if (!learned) {
        HIPCHK(ctx->d_setup.ensure(std::max<int64_t>(K, 1)));
        if (K > 0) {
            HIPCHK(hipEventRecord(ctx->ev[7], ctx->stream));
            k_setup<false><<<(K + 255) / 256, 256, 0, ctx->stream>>>(a, ctx->d_setup.p, nullptr, nullptr);
            HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
            const int64_t want = (K + kTrialThreads - 1) / kTrialThreads;
            const int nblk = static_cast<int>(
                std::max<int64_t>(1, std::min<int64_t>(want, int64_t(ctx->num_cu) * per_cu_plain)));
            k_trials<<<nblk, kTrialThreads, 0, ctx->stream>>>(a, ctx->d_setup.p);
        }
        ctx->last_plain = true;
    } else {
        ctx->last_plain = false;
        HIPCHK(ctx->d_fb_list.ensure(std::max<int64_t>(K, 1)));
        HIPCHK(ctx->d_setup.ensure(std::max<int64_t>(K, 1)));
        HIPCHK(hipMemsetAsync(ctx->d_fb_count.p, 0, sizeof(unsigned long long), ctx->stream));
        if (K > 0) {
            LearnArgs L{};
            L.a = a;
            L.aux = ctx->d_aux.p;
            L.W = MlpDev{ctx->d_mlp_frag.p, ctx->d_mlp_par.p};
            L.fb_list = ctx->d_fb_list.p;
            L.fb_count = ctx->d_fb_count.p;
            L.stats2 = ctx->d_stats.p + 4;
            // the pass's (cell, species) pairs (k_pick counted them per
            // cell): one network evaluation and band envelope per pair
            HIPCHK(ctx->d_pkey.ensure(K));
            HIPCHK(ctx->d_lslot.ensure(K));
            HIPCHK(ctx->d_pairoff.ensure(Nc + 1));
            const int pblk = static_cast<int>((Nc + kPairCellsPerBlock - 1) / kPairCellsPerBlock);
            rc = launch_scan(ctx, ctx->d_lcnt.p, Nc, ctx->d_pairoff.p);
            if (rc) return rc;
            // routing: the slots of single-hadron pairs head the plain list
            if (route) {
                rc = launch_scan(ctx, ctx->d_lscnt.p, Nc, ctx->d_soff.p);
                if (rc) return rc;
            }
            const int64_t* soff = route ? ctx->d_soff.p : nullptr;
            k_lpair_fill<<<pblk, 64 * kPairCellsPerBlock, 0, ctx->stream>>>(a, Wb, ctx->d_lbits.p, ctx->d_pairoff.p,
                                                                            ctx->d_sp16.p, ctx->d_lslot.p,
                                                                            ctx->d_pkey.p, soff, ctx->d_fb_list.p);
            k_plain_count<<<1, 64, 0, ctx->stream>>>(route ? soff + Nc : nullptr, ctx->d_fb_count.p, nullptr);
            HIPCHK(hipGetLastError());
            L.nsingle = route ? soff + Nc : nullptr;
            // the pair count stays on the device: the set-up grid covers the
            // bound P <= K and the waves beyond P return at once
            L.pkey = ctx->d_pkey.p;
            L.P = ctx->d_pairoff.p + Nc;
            L.slot = ctx->d_lslot.p;
            HIPCHK(ctx->d_lrec.ensure(std::max<int64_t>(K, 1)));
            const int lds = kMlpBytes * (kSetupThreads / kWave);
            const int64_t nwave = (K + 63) / 64;
            HIPCHK(hipEventRecord(ctx->ev[7], ctx->stream));
            k_learn_setup<<<(nwave + kSetupThreads / 64 - 1) / (kSetupThreads / 64), kSetupThreads, lds, ctx->stream>>>(
                L, ctx->d_lrec.p);
            HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_learn_trials, 256, 0) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            if (ctx->blocks_per_cu[1] > 0) per_cu = std::min(ctx->blocks_per_cu[1], per_cu);
            const int64_t want = (K + 255) / 256;
            const int nblk = static_cast<int>(
                std::max<int64_t>(1, std::min<int64_t>(want, int64_t(ctx->num_cu) * per_cu)));
            k_learn_trials<<<nblk, 256, 0, ctx->stream>>>(L, ctx->d_lrec.p);
            HIPCHK(hipEventRecord(ctx->ev[8], ctx->stream));
            // routed hadrons and hadrons without an envelope: the plain
            // sampler (its own launch shape and counters), sized on the device
            const int64_t fb_want = route ? (K + 255) / 256 : ctx->num_cu * 4;
            k_setup<false><<<static_cast<int>(std::max<int64_t>(1, fb_want)), 256, 0, ctx->stream>>>(
                a, ctx->d_setup.p, ctx->d_fb_list.p, ctx->d_fb_count.p);
            HIPCHK(hipEventRecord(ctx->ev[9], ctx->stream));
            SampleArgs af = a;
            af.Kdev = ctx->d_fb_count.p;
            af.refill_min = ctx->refill_min[0];
            af.stats = ctx->d_stats.p + kStatRouted;
            k_trials_routed<<<ctx->num_cu * per_cu_plain, kTrialThreads, 0, ctx->stream>>>(af, ctx->d_setup.p);
            HIPCHK(hipEventRecord(ctx->ev[10], ctx->stream));
        }
    }
    HIPCHK(hipGetLastError());
    if (lcc && K > 0) {
        rc = lcc_partners(ctx, a, E, nullptr, per_cu_plain, &K);
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    unsigned long long* h_st = ctx->h_pin.at<unsigned long long>(PIN_STATS);
    HIPCHK(hipMemcpyAsync(h_st, ctx->d_stats.p, kStatSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          ctx->stream));
    int64_t npairs = 0, nrouted = 0;
    if (learned && K > 0)
        HIPCHK(hipMemcpyAsync(ctx->h_pin.at<int64_t>(PIN_AUX), ctx->d_pairoff.p + Nc, sizeof(int64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
    if (route && K > 0)
        HIPCHK(hipMemcpyAsync(ctx->h_pin.at<int64_t>(PIN_AUX + 8), ctx->d_soff.p + Nc, sizeof(int64_t),
                              hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    unsigned long long st[kStatSlots];
    std::memcpy(st, h_st, sizeof(st));
    if (learned && K > 0) npairs = *ctx->h_pin.at<int64_t>(PIN_AUX);
    if (o->proposal == 1 && !learned) nrouted = K;
    if (route && K > 0) nrouted = *ctx->h_pin.at<int64_t>(PIN_AUX + 8);
    float ms_all = 0.f, ms_k = 0.f;
    HIPCHK(hipEventElapsedTime(&ms_all, ctx->ev[3], ctx->ev[5]));
    HIPCHK(hipEventElapsedTime(&ms_k, ctx->ev[4], ctx->ev[5]));
    float ms_su = 0.f, ms_tr = 0.f;
    // the set-up kernel alone (the species pick and the learned pass's pair
    // passes run before it)
    HIPCHK(hipEventElapsedTime(&ms_su, K > 0 ? ctx->ev[7] : ctx->ev[4], K > 0 ? ctx->ev[6] : ctx->ev[4]));
    HIPCHK(hipEventElapsedTime(&ms_tr, ctx->ev[6], ctx->ev[5]));
    ctx->stats.ms_setup_kernel = ms_su;
    ctx->stats.ms_trials_kernel = ms_tr;
    // learned pass: its trial kernel, and the plain set-up and trials of the
    // routed and envelope-less hadrons, apart
    float ms_lt = 0.f, ms_rs = 0.f, ms_rt = 0.f;
    if (learned && K > 0) {
        HIPCHK(hipEventElapsedTime(&ms_lt, ctx->ev[6], ctx->ev[8]));
        HIPCHK(hipEventElapsedTime(&ms_rs, ctx->ev[8], ctx->ev[9]));
        HIPCHK(hipEventElapsedTime(&ms_rt, ctx->ev[9], ctx->ev[10]));
    }
    ctx->stats.ms_learned_trials_kernel = ms_lt;
    ctx->stats.ms_routed_setup_kernel = ms_rs;
    ctx->stats.ms_routed_trials_kernel = ms_rt;
    ctx->stats.n_routed_trials = static_cast<int64_t>(st[kStatRouted + 0]);
    ctx->stats.n_routed_inner = static_cast<int64_t>(st[kStatRouted + 1]);
    ctx->K = K;
    ctx->E = E;
    ctx->stats.n_hadrons = K;
    ctx->stats.n_trials = static_cast<int64_t>(st[0] + st[kStatRouted + 0]);
    ctx->stats.n_inner = static_cast<int64_t>(st[1] + st[kStatRouted + 1]);
    ctx->stats.n_failed = static_cast<int64_t>(st[2] + st[kStatRouted + 2]);
    ctx->stats.n_range_errors = static_cast<int64_t>(st[3] + st[kStatRouted + 3]);
    ctx->stats.ms_sample = ms_all;
    ctx->stats.ms_sample_kernel = ms_k;
    ctx->stats.n_fallback = static_cast<int64_t>(st[4]);
    ctx->stats.n_violations = static_cast<int64_t>(st[5]);
    if (o->proposal == 1) ctx->stats.n_pairs = npairs;
    ctx->stats.n_routed = nrouted;
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
    if (int rc = check_output(ctx)) return rc;
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

int iss_autotune(iss_ctx* ctx, const iss_tune_opts* o, iss_tune_report* rep) {
    if (!ctx || !o || (o->proposal != 0 && o->proposal != 1) || o->n_events <= 0 || o->repeats <= 0)
        return ISS_ERR_ARG;
    if (o->proposal == 1 && !ctx->mlp_ready) return set_err(ctx, ISS_ERR_STATE, "no network parameters");
    const int p = o->proposal;
    iss_sample_opts so{};
    so.seed = o->seed;
    so.event0 = 0;
    so.n_events = o->n_events;
    so.proposal = p;
    so.sampling_model = 30;
    so.hydro_mode = 2;
    so.y_LB = -5.;
    so.y_RB = 5.;
    so.max_trials = 5000;
    so.precision = ctx->yields_valid ? ctx->yield_precision : 32;
    const int refill[] = {8, 16, 24, 32, 48};
    const int bpc[] = {0, 3, 2};  // occupancy limit, then fewer resident blocks
    auto timed = [&](int rf, int bp, double* ms) -> int {
        ctx->refill_min[p] = rf;
        ctx->blocks_per_cu[p] = bp;
        double best = 1e30;
        for (int r = 0; r < o->repeats; r++) {
            int64_t n = 0;
            int rc = iss_sample(ctx, &so, &n);
            if (rc) return rc;
            best = std::min(best, ctx->stats.ms_trials_kernel);
        }
        *ms = best;
        return ISS_OK;
    };
    double ms_def = 0.;
    int rc = timed(kRefillMin, 0, &ms_def);
    if (rc) return rc;
    double best = ms_def;
    int best_rf = kRefillMin, best_bp = 0, tried = 1;
    for (int bp : bpc)
        for (int rf : refill) {
            if (rf == kRefillMin && bp == 0) continue;
            double ms = 0.;
            rc = timed(rf, bp, &ms);
            if (rc) return rc;
            tried++;
            if (ms < best) { best = ms; best_rf = rf; best_bp = bp; }
        }
    ctx->refill_min[p] = best_rf;
    ctx->blocks_per_cu[p] = best_bp;
    // learned proposal: the routing threshold, on the whole sampling pass
    // (pair bookkeeping, set-up and trials of both samplers)
    int best_route = ctx->learn_route;
    if (p == 1) {
        const int routes[] = {1, 2, 3, 4, 6, 8, 12, 16, kLearnRouteAll};
        double best_pass = 1e30;
        for (int rt : routes) {
            ctx->learn_route = rt;
            double ms = 1e30;
            for (int r = 0; r < o->repeats; r++) {
                int64_t n = 0;
                rc = iss_sample(ctx, &so, &n);
                if (rc) return rc;
                ms = std::min(ms, ctx->stats.ms_sample_kernel);
            }
            tried++;
            if (ms < best_pass) { best_pass = ms; best_route = rt; }
        }
        ctx->learn_route = best_route;
    }
    if (rep) {
        rep->refill_min = best_rf;
        rep->blocks_per_cu = best_bp;
        rep->ms_best = best;
        rep->ms_default = ms_def;
        rep->n_tried = tried;
        rep->learned_min_hadrons = p == 1 ? best_route : 0;
    }
    return ISS_OK;
}

int iss_mlp_num_params(void) { return kNumParams; }

int iss_set_learned_routing(iss_ctx* ctx, int32_t min_hadrons) {
    if (!ctx) return ISS_ERR_ARG;
    if (min_hadrons < 1 || min_hadrons > kLearnRouteAll)
        return set_err(ctx, ISS_ERR_ARG, "learned routing: min_hadrons must be in [1, " +
                                             std::to_string(kLearnRouteAll) + "]");
    ctx->learn_route = min_hadrons;
    return ISS_OK;
}

int iss_get_learned_routing(const iss_ctx* ctx) { return ctx ? ctx->learn_route : ISS_ERR_ARG; }

int iss_mlp_init(iss_ctx* ctx, uint64_t seed) {
    if (!ctx) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    std::vector<float> par(kNumParams, 0.f);
    if (seed != 0) {
        // Xavier-uniform from a 64-bit LCG stream
        uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1ull;
        auto uni = [&st]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return static_cast<double>(st >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int fan_out, int fan_in, int used_in) {
            const double lim = std::sqrt(6.0 / (fan_in + fan_out));
            for (int o = 0; o < fan_out; o++)
                for (int i = 0; i < fan_in; i++)
                    par[off + o * fan_in + i] = i < used_in ? static_cast<float>((2. * uni() - 1.) * lim) : 0.f;
        };
        fill(P_W1, kHid, kIn, 14);
        fill(P_W2, kHid, kHid, kHid);
        fill(P_W3, kOut, kHid, kHid);
    }
    return mlp_upload(ctx, par.data());
}

int iss_mlp_set_params(iss_ctx* ctx, const float* params, int n) {
    if (!ctx || !params || n != kNumParams) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    return mlp_upload(ctx, params);
}

int iss_mlp_get_params(iss_ctx* ctx, float* params, int n) {
    if (!ctx || !params || n != kNumParams) return ISS_ERR_ARG;
    if (!ctx->mlp_ready) return set_err(ctx, ISS_ERR_STATE, "no network parameters");
    (void)hipSetDevice(ctx->device);
    HIPCHK(hipMemcpyAsync(params, ctx->d_mlp_par.p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_mlp_collect(iss_ctx* ctx, int64_t max_samples, int64_t* n_samples) {
    if (!ctx || max_samples <= 0) return ISS_ERR_ARG;
    if (!ctx->last_plain) return set_err(ctx, ISS_ERR_STATE, "the last iss_sample was not a plain pass");
    (void)hipSetDevice(ctx->device);
    const int64_t K = ctx->K;
    HIPCHK(ctx->d_u32.ensure(std::max<int64_t>(K, 1)));
    HIPCHK(ctx->d_i64.ensure(K + 1));
    if (K > 0) {
        k_collect_flag<<<(K + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_setup.p, K, ctx->d_out.p, ctx->d_u32.p);
        HIPCHK(hipGetLastError());
    }
    int rc = launch_scan(ctx, ctx->d_u32.p, K, ctx->d_i64.p);
    if (rc) return rc;
    int64_t nacc = 0;
    HIPCHK(hipMemcpyAsync(&nacc, ctx->d_i64.p + K, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int64_t n = std::min(nacc, max_samples);
    HIPCHK(ctx->d_tr_cell.ensure(std::max<int64_t>(n, 1)));
    HIPCHK(ctx->d_tr_spec.ensure(std::max<int64_t>(n, 1)));
    if (K > 0) {
        k_collect_write<<<(K + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_setup.p, K, ctx->d_u32.p, ctx->d_i64.p,
                                                                  max_samples, ctx->d_tr_cell.p, ctx->d_tr_spec.p);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->n_train = n;
    if (n_samples) *n_samples = n;
    return ISS_OK;
}

int iss_mlp_train(iss_ctx* ctx, const iss_train_opts* o, iss_train_report* rep) {
    if (!ctx || !o || o->batch <= 0 || o->batch % 64 != 0 || o->steps < 0 || !(o->lr > 0.f)) return ISS_ERR_ARG;
    if (!ctx->mlp_ready) return set_err(ctx, ISS_ERR_STATE, "no network parameters (iss_mlp_init)");
    if (ctx->n_train <= 0) return set_err(ctx, ISS_ERR_STATE, "empty training set (iss_mlp_collect)");
    if (ctx->Nc <= 0 || !ctx->tables_ready) return set_err(ctx, ISS_ERR_STATE, "surface and tables must be set");
    (void)hipSetDevice(ctx->device);
    const int nb = static_cast<int>(o->batch / 64);
    TrainArgs t{};
    t.a = base_args(ctx);
    t.aux = ctx->d_aux.p;
    t.W = MlpDev{ctx->d_mlp_frag.p, ctx->d_mlp_par.p};
    t.cell = ctx->d_tr_cell.p;
    t.species = ctx->d_tr_spec.p;
    t.n = ctx->n_train;
    t.pick = nullptr;
    t.direct = 0;
    t.npick = o->batch;
    t.seed = o->seed;
    double first = 0., last = 0.;
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int step = 0; step < o->steps; step++) {
        t.step = static_cast<uint32_t>(step);
        int rc = mlp_step(ctx, t, nb, o->lr, nullptr);
        if (rc) return rc;
        if (step == 0) first = mean_loss(ctx, nb);
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    if (o->steps > 0) last = mean_loss(ctx, nb);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (rep) {
        rep->n_samples = ctx->n_train;
        rep->steps = o->steps;
        rep->loss_first = first;
        rep->loss_last = last;
        rep->ms = ms;
    }
    return ISS_OK;
}

int iss_mlp_eval(iss_ctx* ctx, int64_t n, const int32_t* cell, const int32_t* species, float* loss, float* dz,
                 float* inputs, float* grad) {
    if (!ctx) return ISS_ERR_ARG;
    if (!ctx->mlp_ready) return set_err(ctx, ISS_ERR_STATE, "no network parameters");
    int rc = check_samples(ctx, n, cell, species);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    if (n == 0) return ISS_OK;
    std::vector<uint32_t> hc(n);
    std::vector<uint16_t> hs(n);
    for (int64_t i = 0; i < n; i++) { hc[i] = static_cast<uint32_t>(cell[i]); hs[i] = static_cast<uint16_t>(species[i]); }
    HIPCHK(ctx->d_ev_cell.ensure(n));
    HIPCHK(ctx->d_ev_spec.ensure(n));
    HIPCHK(ctx->d_f32a.ensure(n));
    HIPCHK(ctx->d_f32b.ensure(n * kOut));
    HIPCHK(ctx->d_f32c.ensure(n * kIn));
    HIPCHK(hipMemcpyAsync(ctx->d_ev_cell.p, hc.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_ev_spec.p, hs.data(), sizeof(uint16_t) * n, hipMemcpyHostToDevice, ctx->stream));
    TrainArgs t{};
    t.a = base_args(ctx);
    t.aux = ctx->d_aux.p;
    t.W = MlpDev{ctx->d_mlp_frag.p, ctx->d_mlp_par.p};
    t.cell = ctx->d_ev_cell.p;
    t.species = ctx->d_ev_spec.p;
    t.n = n;
    t.pick = nullptr;
    t.direct = 1;
    t.npick = n;
    t.out_loss = ctx->d_f32a.p;
    t.out_dz = ctx->d_f32b.p;
    t.out_x = ctx->d_f32c.p;
    const int nb = static_cast<int>((n + 63) / 64);
    rc = mlp_step(ctx, t, nb, 0.f, ctx->d_grad.p);
    if (rc) return rc;
    if (loss) HIPCHK(hipMemcpyAsync(loss, ctx->d_f32a.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (dz) HIPCHK(hipMemcpyAsync(dz, ctx->d_f32b.p, sizeof(float) * n * kOut, hipMemcpyDeviceToHost, ctx->stream));
    if (inputs)
        HIPCHK(hipMemcpyAsync(inputs, ctx->d_f32c.p, sizeof(float) * n * kIn, hipMemcpyDeviceToHost, ctx->stream));
    if (grad) HIPCHK(hipMemcpyAsync(grad, ctx->d_grad.p, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_band_diag(iss_ctx* ctx, int64_t n, const int32_t* cell, const int32_t* species, const float* rho,
                  float* tables, float* params) {
    if (!ctx || (n > 0 && (!rho || !tables))) return ISS_ERR_ARG;
    int rc = check_samples(ctx, n, cell, species);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    if (n == 0) return ISS_OK;
    HIPCHK(ctx->d_i32a.ensure(n));
    HIPCHK(ctx->d_i32b.ensure(n));
    HIPCHK(ctx->d_f32a.ensure(n * kOut));
    HIPCHK(ctx->d_f32b.ensure(n * kBandDiag));
    HIPCHK(ctx->d_f32c.ensure(n * 24));
    HIPCHK(hipMemcpyAsync(ctx->d_i32a.p, cell, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_i32b.p, species, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_f32a.p, rho, sizeof(float) * n * kOut, hipMemcpyHostToDevice, ctx->stream));
    k_band_diag<<<(n + 255) / 256, 256, 0, ctx->stream>>>(base_args(ctx), ctx->d_aux.p, ctx->d_i32a.p, ctx->d_i32b.p,
                                                         ctx->d_f32a.p, n, ctx->d_f32b.p, ctx->d_f32c.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(tables, ctx->d_f32b.p, sizeof(float) * n * kBandDiag, hipMemcpyDeviceToHost, ctx->stream));
    if (params) HIPCHK(hipMemcpyAsync(params, ctx->d_f32c.p, sizeof(float) * n * 24, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return ISS_OK;
}

int iss_spectra(iss_ctx* ctx, const iss_spectra_opts* o, const int32_t* species, int32_t n_species,
                const double* pT, int32_t n_pT, const double* phi, int32_t n_phi, const double* eta,
                const double* eta_weight, int32_t n_eta, double* out) {
    if (!ctx || !o || !species || !pT || !phi || !eta || !out) return ISS_ERR_ARG;
    if (n_species <= 0 || n_pT <= 0 || n_phi <= 0 || n_eta <= 0)
        return set_err(ctx, ISS_ERR_ARG, "spectra: empty species list or grid");
    if (o->mode != 0 && o->mode != 1) return set_err(ctx, ISS_ERR_ARG, "spectra: mode must be 0 or 1");
    if (o->mode == 0 && !eta_weight) return set_err(ctx, ISS_ERR_ARG, "spectra: mode 0 needs eta weights");
    if (o->precision != 32 && o->precision != 64) return set_err(ctx, ISS_ERR_ARG, "spectra: precision 32 or 64");
    if (o->restrict_deltaf && !(o->deltaf_max_ratio > 0.))
        return set_err(ctx, ISS_ERR_ARG, "spectra: deltaf_max_ratio must be positive");
    if (ctx->Ns <= 0 || !ctx->tables_ready) return set_err(ctx, ISS_ERR_STATE, "species and delta f must be set");
    for (int i = 0; i < n_species; i++)
        if (species[i] < 0 || species[i] >= ctx->Ns) return set_err(ctx, ISS_ERR_ARG, "spectra: species index out of range");
    for (int i = 0; i < n_pT; i++)
        if (!(pT[i] >= 0.) || !std::isfinite(pT[i])) return set_err(ctx, ISS_ERR_ARG, "spectra: pT must be finite, >= 0");
    for (int k = 0; k < n_eta; k++)
        if (!std::isfinite(eta[k]) || std::fabs(eta[k]) > 20.) return set_err(ctx, ISS_ERR_ARG, "spectra: |eta| <= 20");
    (void)hipSetDevice(ctx->device);
    if (o->precision == 64)
        return run_spectra<double>(ctx, o, species, n_species, pT, n_pT, phi, n_phi, eta, eta_weight, n_eta, out);
    return run_spectra<float>(ctx, o, species, n_species, pT, n_pT, phi, n_phi, eta, eta_weight, n_eta, out);
}

double iss_spectra_ms(const iss_ctx* ctx) { return ctx ? ctx->spectra_ms : 0.; }

// calculate_flows (src/emissionfunction.cpp:875-1023): Fourier moments of one
// spectrum in phi per pT (differential) and weighted by pT dpT (integrated)
int iss_flows(const double* dN, const double* pT, const double* pT_weight, int32_t n_pT, const double* phi,
              const double* phi_weight, int32_t n_phi, double mass, int32_t to_order, double* vn_diff,
              double* vn_inte) {
    if (!dN || !pT || !pT_weight || !phi || !phi_weight || !vn_diff || !vn_inte) return ISS_ERR_ARG;
    if (n_pT <= 0 || n_phi <= 0 || to_order < 1) return ISS_ERR_ARG;
    const int nc = 3 + 3 * to_order;
    std::vector<double> norm(n_pT, 0.), num(static_cast<size_t>(n_pT) * to_order * 2, 0.);
    for (int i = 0; i < n_pT; i++) {
        for (int j = 0; j < n_phi; j++) {
            const double v = dN[static_cast<size_t>(i) * n_phi + j] * phi_weight[j];
            norm[i] += v;
            for (int n = 1; n <= to_order; n++) {
                num[(static_cast<size_t>(i) * to_order + n - 1) * 2] += v * std::cos(n * phi[j]);
                num[(static_cast<size_t>(i) * to_order + n - 1) * 2 + 1] += v * std::sin(n * phi[j]);
            }
        }
        norm[i] += 1e-30;
        double* row = vn_diff + static_cast<size_t>(i) * nc;
        row[0] = pT[i];
        row[1] = std::sqrt(mass * mass + pT[i] * pT[i]) - mass;
        row[2] = norm[i] / (2. * kPi);
        for (int n = 0; n < to_order; n++) {
            const double re = num[(static_cast<size_t>(i) * to_order + n) * 2];
            const double im = num[(static_cast<size_t>(i) * to_order + n) * 2 + 1];
            row[3 + 3 * n] = re / norm[i];
            row[4 + 3 * n] = im / norm[i];
            row[5 + 3 * n] = std::sqrt(re * re + im * im) / norm[i];
        }
    }
    double ni = 0.;
    std::vector<double> vi(static_cast<size_t>(to_order) * 2, 0.);
    for (int i = 0; i < n_pT; i++) {
        const double wt = pT[i] * pT_weight[i];
        ni += norm[i] * wt;
        for (int n = 0; n < to_order; n++) {
            vi[2 * n] += num[(static_cast<size_t>(i) * to_order + n) * 2] * wt;
            vi[2 * n + 1] += num[(static_cast<size_t>(i) * to_order + n) * 2 + 1] * wt;
        }
    }
    const double row0[6] = {0., ni, 0., 1., 0., 1.};
    for (int c = 0; c < 6; c++) vn_inte[c] = row0[c];
    for (int n = 0; n < to_order; n++) {
        double* row = vn_inte + 6 * (n + 1);
        const double re = vi[2 * n], im = vi[2 * n + 1];
        row[0] = n + 1;
        row[1] = re;
        row[2] = im;
        row[3] = re / ni;
        row[4] = im / ni;
        row[5] = std::sqrt(re * re + im * im) / ni;
    }
    return ISS_OK;
}

// ------------------------------------------------------------ batched surfaces

int iss_set_surfaces_lrf(iss_ctx* ctx, int32_t n_surfaces, const int64_t* n_cells, const float* cells28) {
    if (!ctx || n_surfaces <= 0 || !n_cells) return ISS_ERR_ARG;
    std::vector<int64_t> off(n_surfaces + 1, 0);
    for (int k = 0; k < n_surfaces; k++) {
        if (n_cells[k] < 0) return set_err(ctx, ISS_ERR_ARG, "negative cell count");
        off[k + 1] = off[k] + n_cells[k];
    }
    if (off[n_surfaces] > (int64_t{1} << 32) - 2) return set_err(ctx, ISS_ERR_RANGE, "more than 2^32 cells");
    int rc = iss_set_surface_lrf(ctx, off[n_surfaces], cells28);
    if (rc) return rc;
    HIPCHK(ctx->d_surf_off.ensure(n_surfaces + 1));
    HIPCHK(hipMemcpy(ctx->d_surf_off.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
    ctx->h_surf_off = off;
    ctx->S = n_surfaces;
    return ISS_OK;
}

int iss_num_surfaces(const iss_ctx* ctx) { return ctx ? ctx->S : -1; }

int iss_sample_batch(iss_ctx* ctx, const iss_batch_opts* o, const int32_t* n_events, int64_t* n_hadrons) {
    if (!ctx || !o || !n_events) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->S <= 0) return set_err(ctx, ISS_ERR_STATE, "no batched surfaces (iss_set_surfaces_lrf)");
    if (o->max_trials <= 1) return set_err(ctx, ISS_ERR_ARG, "max_trials must exceed 1");
    if (o->proposal != 0 && o->proposal != 1) return set_err(ctx, ISS_ERR_ARG, "proposal must be 0 or 1");
    if (o->proposal == 1 && !ctx->mlp_ready) return set_err(ctx, ISS_ERR_STATE, "no network parameters");
    // (the learned proposal with every pair routed runs the plain rounds)
    const bool learned = o->proposal == 1 && ctx->learn_route < kLearnRouteAll;
    const int S = ctx->S;
    int maxE = 0;
    for (int k = 0; k < S; k++) {
        if (n_events[k] < 0) return set_err(ctx, ISS_ERR_ARG, "negative event count");
        maxE = std::max(maxE, n_events[k]);
    }
    const int R = o->events_per_round > 0 ? o->events_per_round : ctx->batch_R_tuned[o->proposal == 1 ? 1 : 0];
    if (R > kMaxEventsPerRound)
        return set_err(ctx, ISS_ERR_ARG, "events_per_round must be at most " + std::to_string(kMaxEventsPerRound));
    invalidate_output(ctx);
    if (!ctx->yields_valid || ctx->yields_lcc || ctx->yield_layout != LAYOUT_ROWS ||
        (o->precision != ctx->yield_precision && (o->precision == 32 || o->precision == 64))) {
        int rc = run_yields(ctx, o->precision == 64 ? 64 : 32, nullptr, nullptr, 0, nullptr, LAYOUT_ROWS);
        if (rc) return rc;
    }
    const int64_t N = ctx->Nc;
    const int rounds = (maxE + R - 1) / R;
    const bool binv = o->hydro_mode != 2;
    const double ymult = binv ? (o->y_RB - o->y_LB) : 1.0;
    // expected hadrons per round -> reserved output and set-up records
    HIPCHK(ctx->d_lam.ensure(S));
    HIPCHK(ctx->d_bevents.ensure(S));
    HIPCHK(hipMemcpyAsync(ctx->d_bevents.p, n_events, sizeof(int32_t) * S, hipMemcpyHostToDevice, ctx->stream));
    k_batch_lambda<<<S, 256, 0, ctx->stream>>>(ctx->d_surf_off.p, ctx->d_ctot.p, ctx->d_lam.p);
    HIPCHK(hipGetLastError());
    std::vector<double> lam(S);
    HIPCHK(hipMemcpyAsync(lam.data(), ctx->d_lam.p, sizeof(double) * S, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double total = 0., maxr = 0.;
    for (int r = 0; r < rounds; r++) {
        double kr = 0.;
        for (int k = 0; k < S; k++) {
            const int64_t rem = static_cast<int64_t>(n_events[k]) - static_cast<int64_t>(r) * R;
            const int n = rem <= 0 ? 0 : static_cast<int>(std::min<int64_t>(rem, R));
            kr += n * ymult * lam[k];
        }
        total += kr;
        maxr = std::max(maxr, kr);
    }
    const int64_t out_cap = static_cast<int64_t>(total + 10. * std::sqrt(total) + 4096.);
    const int64_t setup_cap = static_cast<int64_t>(maxr + 10. * std::sqrt(maxr) + 4096.);
    // (a round's slots are placed by 32-bit permutations)
    if (setup_cap >= (1ll << 31))
        return set_err(ctx, ISS_ERR_RANGE, "more than 2^31 hadrons in one batched round: lower events_per_round");
    int D = 0;
    while ((1ll << D) < R) D++;
    HIPCHK(ctx->d_out.ensure(out_cap));
    HIPCHK(ctx->d_setup.ensure(setup_cap));
    HIPCHK(ctx->d_counts.ensure(std::max<int64_t>(N, 1)));
    HIPCHK(ctx->d_off.ensure(N + 1));
    HIPCHK(ctx->d_perms.ensure(S));
    HIPCHK(ctx->d_seg.ensure(2 * static_cast<size_t>(std::max(rounds, 1)) * S));
    HIPCHK(ctx->d_bevcount.ensure(static_cast<size_t>(std::max(rounds, 1)) * S * R));
    HIPCHK(ctx->d_tree0.ensure(static_cast<size_t>(S) << D));
    HIPCHK(ctx->d_tree1.ensure(static_cast<size_t>(S) << D));
    HIPCHK(ctx->d_scan.ensure(std::max<int64_t>(1, (N + kScanTile - 1) / kScanTile) + 1));
    HIPCHK(ctx->d_bround.ensure(1));
    HIPCHK(ctx->d_bbase.ensure(1));
    HIPCHK(hipMemsetAsync(ctx->d_bround.p, 0, sizeof(uint32_t), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_bbase.p, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_stats.p, 0, kStatSlots * sizeof(unsigned long long), ctx->stream));
    BatchDev b{ctx->d_surf_off.p, ctx->d_bevents.p, S, R, ctx->d_bround.p, ctx->d_bbase.p, ymult};
    SampleArgs a = base_args(ctx);
    a.cdf = ctx->d_cdf.p;
    a.ctot = ctx->d_ctot.p;
    a.off = ctx->d_off.p;
    a.K = 0;
    a.Kdev = reinterpret_cast<const unsigned long long*>(ctx->d_off.p + N);
    a.seed = o->seed;
    a.chunk = 0;
    a.chunk_dev = ctx->d_bround.p;
    a.surf_off = ctx->d_surf_off.p;
    a.S = S;
    a.perms = ctx->d_perms.p;
    a.out_base = ctx->d_bbase.p;
    a.out_cap = static_cast<uint64_t>(out_cap);
    a.setup_cap = setup_cap;
    HIPCHK(ctx->d_cellidx.ensure(std::max<int64_t>(setup_cap, 1)));
    HIPCHK(ctx->d_sp16.ensure(std::max<int64_t>(setup_cap, 1)));
    a.cellidx = ctx->d_cellidx.p;
    a.slot_species = ctx->d_sp16.p;  // (every slot's species from the round's pick)
    a.boost_invariant = binv ? 1 : 0;
    a.ecell = binv ? nullptr : ctx->d_ecell.p;  // 3+1D: the cells' flow and position words
    a.yLB = static_cast<float>(o->y_LB);
    a.yRB = static_cast<float>(o->y_RB);
    a.max_trials = o->max_trials;
    a.refill_min = ctx->refill_min[learned ? 1 : 0];
    a.out = ctx->d_out.p;
    a.stats = ctx->d_stats.p;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trials, kTrialThreads, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int per_cu_fb = per_cu;  // (the learned rounds' plain fallback)
    if (ctx->blocks_per_cu[0] > 0) per_cu = std::min(ctx->blocks_per_cu[0], per_cu);
    const int trial_blocks = ctx->num_cu * per_cu;
    const int setup_blocks = ctx->num_cu * 8;
    const int64_t ntiles = std::max<int64_t>(1, (N + kScanTile - 1) / kScanTile);
    // learned proposal: the round's (cell, species) pairs, one network
    // evaluation and band envelope per pair, the learned trials, and the
    // plain sampler for the hadrons without an envelope; all bounded by the
    // round's set-up records
    LearnArgs L{};
    int learn_blocks = 0, lsetup_blocks = 0;
    const int Wb = (ctx->Ns + 31) / 32;
    const bool route = learned && ctx->learn_route > 1;
    if (learned) {
        const int64_t cap = std::max<int64_t>(setup_cap, 1);
        HIPCHK(ctx->d_lbits.ensure(std::max<int64_t>(N, 1) * Wb));
        HIPCHK(ctx->d_lcnt.ensure(std::max<int64_t>(N, 1)));
        HIPCHK(ctx->d_lscnt.ensure(std::max<int64_t>(N, 1)));
        HIPCHK(ctx->d_pairoff.ensure(N + 1));
        HIPCHK(ctx->d_soff.ensure(N + 1));
        HIPCHK(ctx->d_pkey.ensure(cap));
        HIPCHK(ctx->d_lslot.ensure(cap));
        HIPCHK(ctx->d_lrec.ensure(cap));
        HIPCHK(ctx->d_fb_list.ensure(cap));
        L.a = a;
        L.aux = ctx->d_aux.p;
        L.W = MlpDev{ctx->d_mlp_frag.p, ctx->d_mlp_par.p};
        L.fb_list = ctx->d_fb_list.p;
        L.fb_count = ctx->d_fb_count.p;
        L.stats2 = ctx->d_stats.p + 4;
        L.pkey = ctx->d_pkey.p;
        L.P = ctx->d_pairoff.p + N;
        L.slot = ctx->d_lslot.p;
        L.nsingle = route ? ctx->d_soff.p + N : nullptr;
        int pl = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&pl, k_learn_trials, 256, 0) != hipSuccess || pl < 1) pl = 1;
        if (ctx->blocks_per_cu[1] > 0) pl = std::min(ctx->blocks_per_cu[1], pl);
        learn_blocks = ctx->num_cu * pl;
        lsetup_blocks = static_cast<int>((cap + kSetupThreads - 1) / kSetupThreads);
    }
    SampleArgs af = a;  // fallback: the listed slots, counted on the device
    af.Kdev = ctx->d_fb_count.p;
    af.slot_species = ctx->d_sp16.p;
    af.refill_min = ctx->refill_min[0];
    // one round: only kernel launches, every size read on the device
    auto round = [&]() {
        if (N > 0) k_batch_counts<<<(N + 255) / 256, 256, 0, ctx->stream>>>(b, ctx->d_ctot.p, N, o->seed, ctx->d_counts.p);
        k_scan_tiles<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_counts.p, N, ctx->d_scan.p);
        k_scan_top<<<1, kScanBlock, 0, ctx->stream>>>(ctx->d_scan.p, ntiles);
        k_scan_apply<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_counts.p, N, ctx->d_scan.p, ctx->d_off.p);
        k_batch_roots<<<(S + 63) / 64, 64, 0, ctx->stream>>>(b, ctx->d_off.p, o->seed, ctx->d_tree0.p, ctx->d_perms.p,
                                                             ctx->d_seg.p);
        int64_t* cur = ctx->d_tree0.p;
        int64_t* nxt = ctx->d_tree1.p;
        for (int d = 0; d < D; d++) {
            const int64_t nodes = static_cast<int64_t>(S) << d;
            k_batch_multi_level<<<(nodes + 255) / 256, 256, 0, ctx->stream>>>(b, cur, nxt, d, o->seed);
            std::swap(cur, nxt);
        }
        k_batch_multi_leaves<<<((static_cast<int64_t>(S) << D) + 255) / 256, 256, 0, ctx->stream>>>(
            b, cur, D, ctx->d_bevcount.p);
        // every slot's cell and species from its cell's yield row (the FP64
        // prefix in LDS, as one surface's pass); the learned rounds' pairs
        if (N > 0) {
            const int pblk = static_cast<int>((N + kPickCells - 1) / kPickCells);
            const bool f64 = ctx->yield_precision == 64;
            const void* rows = ctx->d_rows.p;
            uint32_t* gb = learned ? ctx->d_lbits.p : nullptr;
            uint32_t* gc = learned ? ctx->d_lcnt.p : nullptr;
            uint32_t* gs = route ? ctx->d_lscnt.p : nullptr;
            const int thr = ctx->learn_route;
            if (f64 && learned)
                k_pick<double, true><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, true), ctx->stream>>>(
                    a, static_cast<const double*>(rows), row_stride(ctx->Ns, 8), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, thr);
            else if (f64)
                k_pick<double, false><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, false), ctx->stream>>>(
                    a, static_cast<const double*>(rows), row_stride(ctx->Ns, 8), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, thr);
            else if (learned)
                k_pick<float, true><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, true), ctx->stream>>>(
                    a, static_cast<const float*>(rows), row_stride(ctx->Ns, 4), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, thr);
            else
                k_pick<float, false><<<pblk, kWave * kPickCells, pick_lds_bytes(ctx->Ns, false), ctx->stream>>>(
                    a, static_cast<const float*>(rows), row_stride(ctx->Ns, 4), ctx->d_sp16.p, ctx->d_cellidx.p, gb, gc, gs, thr);
        }
        if (!learned) {
            k_setup<true><<<setup_blocks, 256, 0, ctx->stream>>>(a, ctx->d_setup.p, nullptr, nullptr);
            k_trials<<<trial_blocks, kTrialThreads, 0, ctx->stream>>>(a, ctx->d_setup.p);
        } else {
            k_scan_tiles<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_lcnt.p, N, ctx->d_scan.p);
            k_scan_top<<<1, kScanBlock, 0, ctx->stream>>>(ctx->d_scan.p, ntiles);
            k_scan_apply<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_lcnt.p, N, ctx->d_scan.p, ctx->d_pairoff.p);
            // routing: the slots of single-hadron pairs head the plain list
            if (route) {
                k_scan_tiles<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_lscnt.p, N, ctx->d_scan.p);
                k_scan_top<<<1, kScanBlock, 0, ctx->stream>>>(ctx->d_scan.p, ntiles);
                k_scan_apply<<<ntiles, kScanBlock, 0, ctx->stream>>>(ctx->d_lscnt.p, N, ctx->d_scan.p, ctx->d_soff.p);
            }
            const int64_t* soff = route ? ctx->d_soff.p : nullptr;
            const int fblk = static_cast<int>(std::max<int64_t>(1, (N + kPairCellsPerBlock - 1) / kPairCellsPerBlock));
            k_lpair_fill<<<fblk, 64 * kPairCellsPerBlock, 0, ctx->stream>>>(a, Wb, ctx->d_lbits.p, ctx->d_pairoff.p,
                                                                            ctx->d_sp16.p, ctx->d_lslot.p,
                                                                            ctx->d_pkey.p, soff, ctx->d_fb_list.p);
            k_plain_count<<<1, 64, 0, ctx->stream>>>(route ? soff + N : nullptr, ctx->d_fb_count.p,
                                                     ctx->d_stats.p + 8);
            k_learn_setup<<<lsetup_blocks, kSetupThreads, kMlpBytes * (kSetupThreads / kWave), ctx->stream>>>(
                L, ctx->d_lrec.p);
            k_learn_trials<<<learn_blocks, 256, 0, ctx->stream>>>(L, ctx->d_lrec.p);
            k_setup<true><<<setup_blocks, 256, 0, ctx->stream>>>(a, ctx->d_setup.p, ctx->d_fb_list.p,
                                                                 ctx->d_fb_count.p);
            k_trials_routed<<<ctx->num_cu * (route ? per_cu : per_cu_fb), kTrialThreads, 0, ctx->stream>>>(af, ctx->d_setup.p);
        }
        k_batch_advance<<<1, 64, 0, ctx->stream>>>(b, ctx->d_off.p, N);
    };
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    if (rounds > 0) {
        if (o->use_graph) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            round();
            const hipError_t le = hipGetLastError();
            const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
            if (le != hipSuccess || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return set_err(ctx, ISS_ERR_HIP, std::string("round capture: ") +
                                                     hipGetErrorString(le != hipSuccess ? le : ce));
            }
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return set_err(ctx, ISS_ERR_HIP, std::string("round graph: ") + hipGetErrorString(ie));
            }
            for (int r = 0; r < rounds && ie == hipSuccess; r++) ie = hipGraphLaunch(exec, ctx->stream);
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return set_err(ctx, ISS_ERR_HIP, std::string("round replay: ") + hipGetErrorString(ie));
        } else {
            for (int r = 0; r < rounds; r++) round();
            HIPCHK(hipGetLastError());
        }
    }
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    unsigned long long st[9];
    unsigned long long K = 0;
    HIPCHK(hipMemcpyAsync(st, ctx->d_stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&K, ctx->d_bbase.p, sizeof(K), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    ctx->batch_ms = ms;
    ctx->batch_rounds = rounds;
    ctx->batch_R = R;
    ctx->batch_S = S;
    ctx->h_bevents.assign(n_events, n_events + S);
    ctx->K = static_cast<int64_t>(K);
    ctx->E = 0;
    ctx->last_plain = false;
    ctx->stats.n_hadrons = ctx->K;
    ctx->stats.n_trials = static_cast<int64_t>(st[0]);
    ctx->stats.n_inner = static_cast<int64_t>(st[1]);
    ctx->stats.n_failed = static_cast<int64_t>(st[2]);
    ctx->stats.n_range_errors = static_cast<int64_t>(st[3]);
    ctx->stats.n_fallback = learned ? static_cast<int64_t>(st[4]) : 0;
    ctx->stats.n_violations = learned ? static_cast<int64_t>(st[5]) : 0;
    ctx->stats.n_routed = learned ? static_cast<int64_t>(st[8]) : (o->proposal == 1 ? ctx->K : 0);
    clear_routed_stats(ctx);
    ctx->stats.ms_sample = ms;
    ctx->stats.ms_sample_kernel = ms;
    if (st[6] || st[7] || ctx->K > out_cap) {
        // slots past the reserved output or set-up records were never
        // written: no list is left for the getters or iss_decay
        invalidate_output(ctx);
        if (n_hadrons) *n_hadrons = 0;
        return set_err(ctx, ISS_ERR_RANGE, "batched sampling: a round exceeded the reserved output (10 sigma)");
    }
    if (n_hadrons) *n_hadrons = ctx->K;
    if (st[3]) return set_err(ctx, ISS_ERR_RANGE, "m/T - mu/T outside the momentum-sampler tables");
    return ISS_OK;
}

int iss_batch_rounds(const iss_ctx* ctx) { return ctx ? ctx->batch_rounds : -1; }
int iss_batch_events_per_round(const iss_ctx* ctx) { return ctx ? ctx->batch_R : -1; }
double iss_batch_ms(const iss_ctx* ctx) { return ctx ? ctx->batch_ms : 0.; }

int iss_get_batch_layout(iss_ctx* ctx, int64_t* segments, int64_t* event_sizes) {
    if (!ctx) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    const int S = ctx->batch_S, R = ctx->batch_R, rounds = ctx->batch_rounds;
    if (rounds == 0) return ISS_OK;
    const size_t nseg = static_cast<size_t>(rounds) * S;
    if (segments)
        HIPCHK(hipMemcpy(segments, ctx->d_seg.p, sizeof(int64_t) * 2 * nseg, hipMemcpyDeviceToHost));
    if (event_sizes) {
        std::vector<uint32_t> ev(nseg * R);
        HIPCHK(hipMemcpy(ev.data(), ctx->d_bevcount.p, sizeof(uint32_t) * ev.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < rounds; r++)
            for (int k = 0; k < S; k++) {
                const int64_t rem = static_cast<int64_t>(ctx->h_bevents[k]) - static_cast<int64_t>(r) * R;
                const int n = rem <= 0 ? 0 : static_cast<int>(std::min<int64_t>(rem, R));
                for (int e = 0; e < R; e++) {
                    const size_t i = (static_cast<size_t>(r) * S + k) * R + e;
                    event_sizes[i] = e < n ? static_cast<int64_t>(ev[i]) : 0;
                }
            }
    }
    return ISS_OK;
}

int iss_autotune_batch(iss_ctx* ctx, const iss_batch_opts* o, const int32_t* n_events, iss_tune_report* rep) {
    if (!ctx || !o || !n_events) return ISS_ERR_ARG;
    // the same candidates for both proposals (the learned one evaluates its
    // network and envelopes once per (cell, species) pair of a round, so
    // longer rounds share them among more hadrons)
    const bool learned = o->proposal == 1;
    const int cand[] = {4, 16, 64, 256, 1024, 4096, kMaxEventsPerRound};
    const int ncand = 7, R_def = learned ? 256 : 16;
    iss_batch_opts bo = *o;
    double best = 1e30, ms_def = 0.;
    int best_R = R_def, tried = 0;
    for (int ci = 0; ci < ncand; ci++) {
        const int R = cand[ci];
        bo.events_per_round = R;
        double t = 1e30;
        for (int rep_i = 0; rep_i < 2; rep_i++) {
            int64_t n = 0;
            int rc = iss_sample_batch(ctx, &bo, n_events, &n);
            if (rc) return rc;
            t = std::min(t, ctx->batch_ms);
        }
        tried++;
        if (R == R_def) ms_def = t;
        if (t < best) { best = t; best_R = R; }
    }
    ctx->batch_R_tuned[learned ? 1 : 0] = best_R;
    // learned proposal: the routing threshold at the chosen round size
    int best_route = ctx->learn_route;
    if (learned) {
        const int routes[] = {1, 2, 3, 4, 6, 8, 12, 16, kLearnRouteAll};
        bo.events_per_round = best_R;
        for (int rt : routes) {
            ctx->learn_route = rt;
            double t = 1e30;
            for (int rep_i = 0; rep_i < 2; rep_i++) {
                int64_t n = 0;
                int rc = iss_sample_batch(ctx, &bo, n_events, &n);
                if (rc) return rc;
                t = std::min(t, ctx->batch_ms);
            }
            tried++;
            if (t < best) { best = t; best_route = rt; }
        }
        ctx->learn_route = best_route;
    }
    if (rep) {
        rep->refill_min = best_R;   // events per round
        rep->blocks_per_cu = ctx->blocks_per_cu[0];
        rep->ms_best = best;
        rep->ms_default = ms_def;
        rep->n_tried = tried;
        rep->learned_min_hadrons = learned ? best_route : 0;
    }
    return ISS_OK;
}

// ------------------------------------------------------------ resonance decays

int iss_set_decay_table(iss_ctx* ctx, const iss_decay_table* t) {
    if (!ctx || !t || t->n <= 0 || t->n_channels_total < 0) return ISS_ERR_ARG;
    if (!t->monval || !t->mass || !t->width || !t->stable || !t->first_channel || !t->n_channels ||
        (t->n_channels_total > 0 && (!t->npart || !t->branching_ratio || !t->daughters)))
        return ISS_ERR_ARG;
    const int n = t->n, nc = t->n_channels_total;
    for (int i = 0; i < n; i++)
        if (t->first_channel[i] < 0 || t->n_channels[i] < 0 || t->first_channel[i] + t->n_channels[i] > nc)
            return set_err(ctx, ISS_ERR_ARG, "decay table: channel range out of bounds");
    // ascending monval keys; for a repeated monval the first entry wins, as
    // the reference's linear searches do
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t->monval[a] < t->monval[b]; });
    std::vector<int> key, kidx;
    for (int i : order)
        if (key.empty() || key.back() != t->monval[i]) {
            key.push_back(t->monval[i]);
            kidx.push_back(i);
        }
    const int nk = static_cast<int>(key.size());
    auto first_of = [&](int mv) {
        const auto it = std::lower_bound(key.begin(), key.end(), mv);
        return (it != key.end() && *it == mv) ? kidx[it - key.begin()] : -1;
    };
    std::vector<int> ints(5 * static_cast<size_t>(n) + 11 * static_cast<size_t>(std::max(nc, 1)), 0);
    std::vector<double> dbl(2 * static_cast<size_t>(n) + std::max(nc, 1), 0.);
    for (int i = 0; i < nk; i++) {
        ints[i] = key[i];
        ints[n + i] = kidx[i];
    }
    for (int i = 0; i < n; i++) {
        ints[2 * n + i] = t->stable[i];
        ints[3 * n + i] = t->first_channel[i];
        ints[4 * n + i] = t->n_channels[i];
        dbl[i] = t->mass[i];
        dbl[n + i] = t->width[i];
        double cum = 0.;
        for (int j = t->first_channel[i]; j < t->first_channel[i] + t->n_channels[i]; j++) {
            cum += t->branching_ratio[j];
            dbl[2 * n + j] = cum;
        }
    }
    int* ci = ints.data() + 5 * n;
    for (int j = 0; j < nc; j++) {
        ci[j] = t->npart[j];
        for (int k = 0; k < 5; k++) {
            const int mv = t->daughters[5 * j + k];
            ci[nc + 5 * j + k] = mv;
            ci[6 * nc + 5 * j + k] = mv != 0 ? first_of(mv) : -1;
        }
    }
    (void)hipSetDevice(ctx->device);
    HIPCHK(ctx->d_dec_int.ensure(ints.size()));
    HIPCHK(ctx->d_dec_dbl.ensure(dbl.size()));
    HIPCHK(hipMemcpy(ctx->d_dec_int.p, ints.data(), sizeof(int) * ints.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_dec_dbl.p, dbl.data(), sizeof(double) * dbl.size(), hipMemcpyHostToDevice));
    DecayDev& D = ctx->dec;
    const int* I = ctx->d_dec_int.p;
    const double* Dd = ctx->d_dec_dbl.p;
    D.n = nk;
    D.key = I;
    D.key_idx = I + n;
    D.stable = I + 2 * n;
    D.ch0 = I + 3 * n;
    D.nch = I + 4 * n;
    D.npart = I + 5 * n;
    D.dpid = I + 5 * n + nc;
    D.didx = I + 5 * n + 6 * nc;
    D.mass = Dd;
    D.width = Dd + n;
    D.cumbr = Dd + 2 * n;
    ctx->dec_n = n;
    return ISS_OK;
}


// The output of a batch as one list of events: blocks (round, surface) in
// order, each holding its events contiguously (the layout of
// iss_get_batch_layout).  Writes the event offsets to d_evoff, the number of
// events to *E and each block's first event to first[] (rounds x S + 1).
static int batch_events(iss_ctx* ctx, int* E, std::vector<int64_t>& first) {
    const int S = ctx->batch_S, R = ctx->batch_R, rounds = ctx->batch_rounds;
    const size_t nseg = static_cast<size_t>(rounds) * S;
    std::vector<int64_t> seg(2 * nseg);
    std::vector<uint32_t> ev(nseg * R);
    HIPCHK(hipMemcpy(seg.data(), ctx->d_seg.p, sizeof(int64_t) * seg.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ev.data(), ctx->d_bevcount.p, sizeof(uint32_t) * ev.size(), hipMemcpyDeviceToHost));
    std::vector<int64_t> off;
    first.assign(nseg + 1, 0);
    int64_t o = 0;
    for (int r = 0; r < rounds; r++)
        for (int k = 0; k < S; k++) {
            const size_t b = static_cast<size_t>(r) * S + k;
            first[b] = static_cast<int64_t>(off.size());
            if (seg[2 * b] != o) return set_err(ctx, ISS_ERR_STATE, "batch output blocks are not contiguous");
            const int64_t rem = static_cast<int64_t>(ctx->h_bevents[k]) - static_cast<int64_t>(r) * R;
            const int n = rem <= 0 ? 0 : static_cast<int>(std::min<int64_t>(rem, R));
            for (int e = 0; e < n; e++) {
                off.push_back(o);
                o += ev[b * R + e];
            }
            if (o != seg[2 * b] + seg[2 * b + 1]) return set_err(ctx, ISS_ERR_STATE, "batch event sizes disagree");
        }
    first[nseg] = static_cast<int64_t>(off.size());
    if (o != ctx->K) return set_err(ctx, ISS_ERR_STATE, "batch layout does not cover the hadron list");
    off.push_back(o);
    if (off.size() - 1 > static_cast<size_t>(INT32_MAX)) return set_err(ctx, ISS_ERR_RANGE, "too many events");
    *E = static_cast<int>(off.size() - 1);
    HIPCHK(ctx->d_evoff.ensure(off.size()));
    HIPCHK(hipMemcpy(ctx->d_evoff.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
    return ISS_OK;
}

// the batch layout (segments, event sizes) from the event offsets in d_evoff
// after the feed-down, block by block as batch_events laid them out
static int batch_relayout(iss_ctx* ctx, int E, const std::vector<int64_t>& first) {
    const int S = ctx->batch_S, R = ctx->batch_R, rounds = ctx->batch_rounds;
    const size_t nseg = static_cast<size_t>(rounds) * S;
    std::vector<int64_t> off(static_cast<size_t>(E) + 1), seg(2 * nseg);
    std::vector<uint32_t> ev(nseg * R, 0u);
    HIPCHK(hipMemcpy(off.data(), ctx->d_evoff.p, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost));
    for (size_t b = 0; b < nseg; b++) {
        const int64_t e0 = first[b], e1 = first[b + 1];
        seg[2 * b] = off[e0];
        seg[2 * b + 1] = off[e1] - off[e0];
        for (int64_t e = e0; e < e1; e++) ev[b * R + (e - e0)] = static_cast<uint32_t>(off[e + 1] - off[e]);
    }
    HIPCHK(hipMemcpy(ctx->d_seg.p, seg.data(), sizeof(int64_t) * seg.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_bevcount.p, ev.data(), sizeof(uint32_t) * ev.size(), hipMemcpyHostToDevice));
    return ISS_OK;
}

int iss_decay(iss_ctx* ctx, uint64_t seed, int64_t* n_hadrons) {
    if (!ctx) return ISS_ERR_ARG;
    if (ctx->dec_n == 0) return set_err(ctx, ISS_ERR_STATE, "no decay table (iss_set_decay_table)");
    if (int rc = check_output(ctx)) return rc;
    (void)hipSetDevice(ctx->device);
    HIPCHK(ctx->h_pin.ensure(PIN_TOT + sizeof(double) * std::max(ctx->Ns, 1)));
    // a batch's output decays as one list of its events, then gets its
    // layout back
    const bool batch = ctx->E <= 0 && ctx->K > 0 && ctx->batch_rounds > 0;
    if (ctx->E <= 0 && ctx->K > 0 && !batch) return set_err(ctx, ISS_ERR_STATE, "hadron list without events");
    std::vector<int64_t> bfirst;
    int E = ctx->E;
    if (batch) {
        int rc = batch_events(ctx, &E, bfirst);
        if (rc) return rc;
    }
    // generation buffers live in the context and only grow
    auto& gens = ctx->dgen;
    if (gens.empty()) gens.emplace_back(new DecayGen);
    int ngen = 1;
    gens[0]->P = ctx->d_out.p;
    gens[0]->n = ctx->K;
    HIPCHK(ctx->d_stats.ensure(8));
    HIPCHK(hipMemsetAsync(ctx->d_stats.p + 6, 0, 2 * sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int g = 0;; g++) {
        DecayGen& G = *gens[g];
        const int64_t n = G.n;
        HIPCHK(G.kind.ensure(n));
        HIPCHK(G.chan.ensure(n));
        HIPCHK(G.nchild.ensure(n));
        HIPCHK(G.size.ensure(n));
        HIPCHK(G.off.ensure(n + 1));
        HIPCHK(G.pos.ensure(n + 1));
        if (n > 0)
            k_decay_pick<<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->dec, G.P, n, g, seed, G.kind.p, G.chan.p,
                                                                   G.nchild.p, ctx->d_stats.p + 6);
        HIPCHK(hipGetLastError());
        int rc = launch_scan(ctx, G.nchild.p, n, G.off.p);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(ctx->h_pin.at<int64_t>(PIN_K), G.off.p + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        const int64_t nn = *ctx->h_pin.at<int64_t>(PIN_K);
        if (nn == 0) break;
        if (g + 1 >= kDecayMaxGen) return set_err(ctx, ISS_ERR_RANGE, "decay chains deeper than 12 generations");
        if (static_cast<int>(gens.size()) < g + 2) gens.emplace_back(new DecayGen);
        ngen = g + 2;
        DecayGen& H = *gens[g + 1];
        HIPCHK(H.own.ensure(nn));
        HIPCHK(H.parent.ensure(nn));
        H.P = H.own.p;
        H.n = nn;
        k_decay_emit<<<(n + 255) / 256, 256, 0, ctx->stream>>>(ctx->dec, G.P, n, g, seed, G.kind.p, G.chan.p,
                                                               G.off.p, H.own.p, H.parent.p, ctx->d_stats.p + 7);
        HIPCHK(hipGetLastError());
    }
    // final descendants, bottom-up
    const int last = ngen - 1;
    for (int g = last; g >= 0; g--) {
        DecayGen& G = *gens[g];
        if (G.n == 0) continue;
        k_decay_size<<<(G.n + 255) / 256, 256, 0, ctx->stream>>>(G.kind.p, G.off.p, G.nchild.p,
                                                                 g < last ? gens[g + 1]->size.p : nullptr, G.n,
                                                                 G.size.p);
    }
    HIPCHK(hipGetLastError());
    int rc = launch_scan(ctx, gens[0]->size.p, gens[0]->n, gens[0]->pos.p);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->h_pin.at<int64_t>(PIN_K), gens[0]->pos.p + gens[0]->n, sizeof(int64_t),
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int64_t F = *ctx->h_pin.at<int64_t>(PIN_K);
    for (int g = 0; g < last; g++) {
        DecayGen& G = *gens[g];
        if (G.n == 0) continue;
        k_decay_place<<<(G.n + 255) / 256, 256, 0, ctx->stream>>>(G.kind.p, G.off.p, G.nchild.p, G.pos.p,
                                                                  gens[g + 1]->size.p, G.n, gens[g + 1]->pos.p);
    }
    DBuf<Hadron>& out = ctx->d_out2;
    HIPCHK(out.ensure(std::max<int64_t>(F, 1)));
    for (int g = 0; g <= last; g++) {
        DecayGen& G = *gens[g];
        if (G.n == 0) continue;
        k_decay_write<<<(G.n + 255) / 256, 256, 0, ctx->stream>>>(G.P, G.kind.p, G.pos.p, G.n, out.p);
    }
    if (E > 0) {
        HIPCHK(ctx->d_tree0.ensure(E + 1));
        k_decay_events<<<(E + 1 + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_evoff.p, E, gens[0]->pos.p,
                                                                      ctx->d_tree0.p);
        HIPCHK(hipMemcpyAsync(ctx->d_evoff.p, ctx->d_tree0.p, sizeof(int64_t) * (E + 1), hipMemcpyDeviceToDevice,
                              ctx->stream));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
    unsigned long long* h_nfb = ctx->h_pin.at<unsigned long long>(PIN_STATS);
    HIPCHK(hipMemcpyAsync(h_nfb, ctx->d_stats.p + 6, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const unsigned long long nbelow = h_nfb[0], nf = h_nfb[1];
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    std::swap(ctx->d_out.p, out.p);
    std::swap(ctx->d_out.cap, out.cap);
    ctx->K = F;
    if (batch) {
        rc = batch_relayout(ctx, E, bfirst);
        if (rc) return rc;
    }
    ctx->stats.n_hadrons = F;
    ctx->stats.ms_decay = ms;
    ctx->stats.n_decay_generations = last + 1;
    ctx->dec_failed = static_cast<int64_t>(nf);
    ctx->last_plain = false;
    if (n_hadrons) *n_hadrons = F;
    if (nf) return set_err(ctx, ISS_ERR_RANGE, "three-body decay found no kinematically allowed point");
    if (nbelow)
        return set_err(ctx, ISS_ERR_RANGE, std::to_string(nbelow) +
                                               " decays with the pole mass below the daughters' masses "
                                               "(the reference exits; the mothers were dropped)");
    return ISS_OK;
}

int iss_set_hadrons(iss_ctx* ctx, int32_t n_events, const int64_t* counts, const iss_hadron* hadrons) {
    if (!ctx || n_events < 0 || (n_events > 0 && !counts)) return ISS_ERR_ARG;
    std::vector<int64_t> off(n_events + 1, 0);
    for (int e = 0; e < n_events; e++) {
        if (counts[e] < 0) return set_err(ctx, ISS_ERR_ARG, "negative event size");
        off[e + 1] = off[e] + counts[e];
    }
    const int64_t K = off[n_events];
    if (K > 0 && !hadrons) return ISS_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    invalidate_output(ctx);
    HIPCHK(ctx->d_out.ensure(std::max<int64_t>(K, 1)));
    HIPCHK(ctx->d_evoff.ensure(n_events + 1));
    if (K > 0) HIPCHK(hipMemcpy(ctx->d_out.p, hadrons, sizeof(Hadron) * K, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_evoff.p, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice));
    ctx->K = K;
    ctx->E = n_events;
    ctx->stats.n_hadrons = K;
    ctx->last_plain = false;
    return ISS_OK;
}

}  // extern "C"
