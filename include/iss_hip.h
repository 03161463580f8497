/*
 * iss_hip.h — C-ABI of the MI355X (gfx950) Cooper-Frye sampler.
 *
 * The reference (chunshen1987/iSS) has no C-ABI: its FSSW sampler is a C++
 * class driven by `iSS::generate_samples` (src/iSS.cpp:130-165).  Each entry
 * point below replaces one step of that class; the comment names the
 * reference code it stands in for.  Plain pointers and sizes only; all host
 * buffers are caller-owned, device memory is owned by the context.  Every
 * call is synchronous at return (one internal HIP stream per context) and a
 * context must be used by one host thread at a time, as the reference's single
 * shared RNG makes an `iSS` instance single-threaded (src/iSS.h:27).
 *
 * Status: 0 = OK, negative = error; `iss_last_error` gives the message.  No
 * entry point calls exit(); the C++ drop-in layer maps errors to the
 * reference's message-then-exit behaviour only where the reference does.
 */
#ifndef ISS_HIP_H_
#define ISS_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISS_OK 0
#define ISS_ERR_ARG (-1)
#define ISS_ERR_HIP (-2)
#define ISS_ERR_STATE (-3)
#define ISS_ERR_UNSUPPORTED (-4)
#define ISS_ERR_NOMEM (-5)
#define ISS_ERR_RANGE (-6)

typedef struct iss_ctx iss_ctx;

/* iSS_Hadron, same field order and size (src/data_struct.h:79-84). */
typedef struct {
    int32_t pid;
    float mass;
    float E, px, py, pz;
    float t, x, y, z;
} iss_hadron;

/* Species in sampling order (the caller sorts by mass as
 * src/FSSW.cpp:152-162 does).  sign: -1 Bose-Einstein, +1 Fermi-Dirac,
 * 0 Boltzmann (src/readindata.cpp:1103-1113). */
typedef struct {
    int32_t n;
    const int32_t* pid;
    const double* mass;
    const int32_t* degeneracy;
    const int32_t* baryon;
    const int32_t* strange;
    const int32_t* charge;
    const int32_t* sign;
} iss_species_desc;

/* delta-f switches and coefficient tables (src/FSSW.cpp:72-83, 175-195).
 * bulk_kind: 1 RTA polynomial, 11 OSU 14-moment (T, muB) table, 20 WSU
 * 22-moment NEoS table, 21 WSU Chapman-Enskog NEoS table; kinds 0, 2, 3, 4
 * contribute nothing on the FSSW path (src/FSSW.cpp:663-692, 1916-1923). */
typedef struct {
    int32_t include_shear;
    int32_t include_bulk;
    int32_t bulk_kind;
    int32_t include_diffusion;     /* baryon-diffusion delta f (needs qmu_table) */
    /* kind 11: grid of nT x nmu, values [iT * nmu + imu] */
    int32_t k11_nT, k11_nmu;
    double k11_T0, k11_mu0, k11_dT, k11_dmu;
    const double* k11_c0;
    const double* k11_c1;
    const double* k11_c2;
    /* kinds 20/21: 200 x 200 rows of neos_ncol (5 CE, 8 22-moment) */
    int32_t neos_ncol;
    const double* neos;
    /* keep FermionMomentumSampler::CDF_2's missing alternating sign
     * (src/FermionMomentumSampler.cpp:80-93); 1 = as the reference */
    int32_t fermion_cdf2_quirk;
    /* diffusion: Coefficients_RTA_diffusion.dat, 150 x 100 values [iT][imu] on
     * T = 0.05 + 0.001 iT, mu_B = 0.007892 imu (src/FSSW.cpp:1546-1566) */
    const double* qmu_table;
} iss_deltaf_desc;

/* Hadron-resonance-gas EoS used to regulate T, mu, P from (e, n_B)
 * (src/readindata.cpp:768-801, 1249-1309).  rows: n_rows x 7
 * {e, nB, P, T, muB, muS, muQ}; iEOS 9/91: 1-D in e, 12/14: 200 nB per e. */
typedef struct {
    int32_t iEOS;
    int64_t n_rows;
    const double* rows;
} iss_eos_desc;

typedef struct {
    int32_t precision;       /* 32 or 64 */
    int32_t hydro_mode;      /* 2: 3+1D; otherwise boost invariant */
    double y_LB, y_RB;       /* rapidity window (src/FSSW.cpp:950-958) */
} iss_yield_opts;

typedef struct {
    uint64_t seed;
    int64_t event0;          /* index of the first event (counter-based RNG) */
    int32_t n_events;
    int32_t proposal;        /* 0 plain rejection, 1 learned mixture */
    int32_t sampling_model;  /* dN_dy_sampling_model (src/FSSW.cpp:250-309): 30 Poisson, 1 integer part
                                plus one with the fraction's probability, 10 and 20 negative binomial */
    int32_t hydro_mode;
    double y_LB, y_RB;
    int32_t max_trials;      /* maximum_impatience (src/FSSW.cpp:1872) */
    int32_t precision;       /* yield precision used if yields are stale */
    /* local_charge_conservation (src/FSSW.cpp:931-938, 1035-1048): species of
     * negative charge are not sampled; every positive hadron is followed, in
     * its event, by its antiparticle from the same cell (same eta_s).  Needs
     * every positive species' antiparticle in the species table. */
    int32_t local_charge_conservation;
    /* dN_dy_sampling_para1 of the multiplicity models 10 and 20 (src/FSSW.cpp:275-292) */
    double sampling_para1;
} iss_sample_opts;

typedef struct {
    int64_t n_hadrons;
    int64_t n_trials;        /* outer rejection trials */
    int64_t n_inner;         /* |p| proposals */
    int64_t n_failed;        /* slots that exhausted max_trials */
    int64_t n_range_errors;  /* m/T - mu/T outside the |p| tables */
    double ms_yield;         /* device time of the last yield pass */
    double ms_sample;        /* device time of the last sampling pass */
    double ms_sample_kernel; /* device time of the last sampling kernel alone */
    double ms_yield_kernel;  /* device time of the last yield kernel alone */
    int64_t n_fallback;      /* learned proposal: hadrons handed to the plain sampler */
    int64_t n_violations;    /* learned proposal: draws above the envelope (0 when exact) */
    double ms_setup_kernel;  /* device time of the last set-up kernel (plain: per hadron;
                                learned: the network and envelopes, per (cell, species) pair) */
    double ms_trials_kernel; /* device time of the last trial kernel(s), fallback included */
    double ms_decay;         /* device time of the last iss_decay */
    int64_t n_decay_generations;
    int64_t n_pairs;         /* learned proposal: (cell, species) pairs of the last pass */
    int64_t n_routed;        /* learned proposal: hadrons routed to the plain sampler
                                (iss_set_learned_routing) */
    /* learned iss_sample: device time of the learned trial kernel alone, and
     * of the plain set-up and trials of the routed and envelope-less hadrons
     * (ms_trials_kernel holds all three); the plain trials' outer trials and
     * |p| proposals (included in n_trials, n_inner) */
    double ms_learned_trials_kernel;
    double ms_routed_setup_kernel;
    double ms_routed_trials_kernel;
    int64_t n_routed_trials;
    int64_t n_routed_inner;
} iss_stats;

/* ---- lifecycle (the reference: iSS ctor/dtor, src/iSS.cpp:14-30) ---- */
int iss_ctx_create(int device, iss_ctx** out);
int iss_ctx_destroy(iss_ctx* ctx);
const char* iss_last_error(const iss_ctx* ctx);

/* ---- tables: FSSW ctor (src/FSSW.cpp:43-201) ---- */
int iss_set_species(iss_ctx* ctx, const iss_species_desc* species);
int iss_set_deltaf(iss_ctx* ctx, const iss_deltaf_desc* deltaf);

/* ---- surfaces ----
 * Lab-frame MUSIC cells, 34 floats each in FO_surf order (src/data_struct.h:53-63):
 * regulated (src/readindata.cpp:768-842, EoS optional) and boosted to the fluid
 * rest frame on the device (src/iSS.cpp:170-293); cells with u.dsigma < 0 are
 * dropped (src/iSS.cpp:226). */
int iss_set_surface_lab(iss_ctx* ctx, int64_t n, const float* cells34, const iss_eos_desc* eos);
/* Cells already in the rest frame, 28 floats each in FO_surf_LRF order
 * (src/data_struct.h:66-76). */
int iss_set_surface_lrf(iss_ctx* ctx, int64_t n, const float* cells28);
int64_t iss_num_cells(const iss_ctx* ctx);
int iss_get_surface_lrf(iss_ctx* ctx, float* cells28);

/* ---- yields: calculate_dN_dxtdy_for_one_particle_species for all species
 * (src/FSSW.cpp:565-715, 719-848).  species_totals [n_species] FP64 sums of
 * the clamped per-cell yields; per_cell [n_cells][n_species] in the working
 * precision (float or double), may be NULL. ---- */
int iss_compute_yields(iss_ctx* ctx, const iss_yield_opts* opts, double* species_totals,
                       void* per_cell);

/* ---- the same yields with no allocation (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops).  Both outputs are DEVICE pointers:
 * species_totals [n_species] FP64; per_cell [n_cells][stride] in the working
 * precision, stride = n_species rounded up to a multiple of 16 (FP32) or
 * 8 (FP64), entries past n_species zero, 16-byte aligned, holding at least
 * per_cell_capacity elements.  Values equal iss_compute_yields' bit for bit.
 * Scratch (species partials, the out-of-table flag) is a static 4 MiB array
 * per GPU: at most 2048 species; calls that overlap on one GPU race on it.
 * Needs the Bessel tables (built by the context's first iss_compute_yields
 * or iss_sample); diffusion delta f is refused (ISS_ERR_UNSUPPORTED).  The
 * context's own yields and sampling state are left unchanged. ---- */
int iss_compute_yields_noalloc(iss_ctx* ctx, const iss_yield_opts* opts, double* species_totals,
                               void* per_cell, int64_t per_cell_capacity);

/* ---- sampling: sample_using_dN_dxtdy_4all_particles_conventional
 * (src/FSSW.cpp:873-1071) with sample_momemtum_from_a_fluid_cell,
 * the boost and add_one_sampled_particle (src/FSSW.cpp:1852-2010). ---- */
int iss_sample(iss_ctx* ctx, const iss_sample_opts* opts, int64_t* n_hadrons);
/* hadrons per event of the last iss_sample (FSSW::get_number_of_particles) */
int iss_get_event_counts(iss_ctx* ctx, int64_t* counts);
/* all hadrons, grouped by event in event order (FSSW::Hadron_list) */
int iss_get_hadrons(iss_ctx* ctx, iss_hadron* dst, int64_t capacity);
int iss_get_stats(const iss_ctx* ctx, iss_stats* stats);

/* ---- launch autotuner (no counterpart in the reference).  Times short
 * sampling passes of the current surface for each candidate launch shape of
 * the trial kernel (idle lanes before a refill, i.e. the draws a lane takes
 * at once; resident blocks per CU, i.e. the hadrons per wavefront) and keeps
 * the fastest for that proposal.  The samples do not depend on the launch
 * shape (counter-based random numbers). ---- */
typedef struct {
    int32_t proposal;        /* 0 plain, 1 learned (needs network parameters) */
    int32_t n_events;        /* events per timed pass */
    uint64_t seed;
    int32_t repeats;         /* passes per candidate (the minimum counts) */
} iss_tune_opts;

typedef struct {
    int32_t refill_min;      /* chosen */
    int32_t blocks_per_cu;   /* chosen; 0 = occupancy limit */
    double ms_best;          /* trial-kernel time of the chosen shape */
    double ms_default;       /* trial-kernel time of the default shape */
    int32_t n_tried;
    int32_t learned_min_hadrons;  /* learned proposal: the routing threshold chosen on whole
                                     sampling passes (iss_set_learned_routing); 0 for plain */
} iss_tune_report;

int iss_autotune(iss_ctx* ctx, const iss_tune_opts* opts, iss_tune_report* report);

/* ---- learned proposal (no counterpart in the reference; the reference's
 * rejection sampler, src/FSSW.cpp:1852-1966, is the plain proposal).
 * A 14 -> 64 -> 64 -> 8 network (BF16, FP32 accumulation) shapes the
 * angular part of a banded envelope of the exact rest-frame target; accepted
 * momenta follow the same distribution as the plain sampler's.  Parameters
 * are FP32, row-major [out][in]: W1 64x16, b1 64, W2 64x64, b2 64, W3 8x64,
 * b3 8 (iss_mlp_num_params() values).  iss_sample with proposal = 1 needs
 * parameters (init, set or trained). ---- */
int iss_mlp_num_params(void);
/* Xavier-uniform weights from seed, zero biases; seed 0 gives all zeros */
int iss_mlp_init(iss_ctx* ctx, uint64_t seed);
int iss_mlp_set_params(iss_ctx* ctx, const float* params, int n);
int iss_mlp_get_params(iss_ctx* ctx, float* params, int n);
/* Routing of the learned passes (iss_sample, iss_sample_batch): a (cell,
 * species) pair with fewer than min_hadrons hadrons in the pass takes the
 * plain sampler, whose per-hadron set-up and trials cost less than one
 * network evaluation and envelope shared by few hadrons.  1: every pair
 * takes the learned proposal; default 4; 1..64, or ISS_LEARNED_ROUTE_ALL:
 * every pair routed, the pass runs as the plain one without the pair
 * bookkeeping.  iss_stats.n_routed counts the routed hadrons; iss_autotune
 * and iss_autotune_batch (proposal 1) pick the threshold on timed passes.
 * The output law is the same for every threshold. */
#define ISS_LEARNED_ROUTE_ALL 65
int iss_set_learned_routing(iss_ctx* ctx, int32_t min_hadrons);
int iss_get_learned_routing(const iss_ctx* ctx);

typedef struct {
    int64_t batch;           /* samples per step, a multiple of 64 */
    int32_t steps;
    float lr;                /* Adam step size */
    uint64_t seed;           /* minibatch draws */
} iss_train_opts;

typedef struct {
    int64_t n_samples;       /* training-set size */
    int32_t steps;
    double loss_first;       /* mean log envelope mass, first and last step */
    double loss_last;
    double ms;               /* device time of all steps */
} iss_train_report;

/* Keep the (cell, species) of the hadrons accepted by the last plain
 * iss_sample (at most max_samples) as the training set. */
int iss_mlp_collect(iss_ctx* ctx, int64_t max_samples, int64_t* n_samples);
/* Adam on the mean log envelope mass over minibatches of the training set. */
int iss_mlp_train(iss_ctx* ctx, const iss_train_opts* opts, iss_train_report* report);
/* One forward and backward pass on explicit samples: per-sample log
 * envelope mass, its gradient at the 8 outputs, the 16 network inputs
 * (after BF16 rounding) and the gradient of the summed loss with respect to
 * every parameter.  Any output pointer may be NULL. */
int iss_mlp_eval(iss_ctx* ctx, int64_t n, const int32_t* cell, const int32_t* species, float* loss,
                 float* dz, float* inputs, float* grad);
/* Envelope tables for explicit (cell, species, rho[8]): 47 floats per
 * sample (ok, log M, c[9], s[9], kappa[9], beta[9], Q[8] / 2^16, last band)
 * and the 24 hadron parameters they were built from. */
int iss_band_diag(iss_ctx* ctx, int64_t n, const int32_t* cell, const int32_t* species, const float* rho,
                  float* tables, float* params);

/* ---- smooth spectra: EmissionFunctionArray::calculate_dN_pTdpTdphidy
 * (src/emissionfunction.cpp:625-844) for several species at once, from the
 * resident surface with the context's delta f.  dN / (pT dpT dphi dy) in
 * GeV^-2 on the grid pT[n_pT] (GeV) x phi[n_phi].
 *   mode 0 (the reference): each cell is integrated over the eta_s table,
 *          the particle at rapidity y = eta_cell - eta[k] with weight w[k]
 *          (src/emissionfunction.cpp:232-238, 756-757); out[s][i_pT][i_phi].
 *   mode 1: no eta_s integration; eta[] are particle rapidities, w unused;
 *          out[s][i_y][i_pT][i_phi] (rapidity-resolved 3+1D spectra).
 * species: indices into the context's species table. ---- */
typedef struct {
    int32_t precision;        /* 64: FP64 throughout; 32: FP32 integrand, FP64 cell sums */
    int32_t mode;             /* 0 or 1, above */
    int32_t all_cells;        /* 1: also the lab cells dropped for u.dsigma < 0 (the
                                 reference's smooth path uses every cell); 0: sampled cells */
    int32_t use_pos_dN_only;  /* drop negative contributions (:804-806) */
    int32_t restrict_deltaf;  /* rescale to |delta f| <= deltaf_max_ratio (:791-799) */
    double deltaf_max_ratio;
} iss_spectra_opts;

int iss_spectra(iss_ctx* ctx, const iss_spectra_opts* opts, const int32_t* species, int32_t n_species,
                const double* pT, int32_t n_pT, const double* phi, int32_t n_phi, const double* eta,
                const double* eta_weight, int32_t n_eta, double* out);
/* device time of the last iss_spectra, ms (record + integral + reduction) */
double iss_spectra_ms(const iss_ctx* ctx);

/* ---- batched surfaces (no counterpart in the reference, which runs one
 * surface per iSS instance, src/iSS.cpp:130-168).  Many rest-frame surfaces
 * of uneven size are held as one record; surface k owns the next n_cells[k]
 * cells.  iss_sample_batch samples n_events[k] events of every surface k in
 * rounds of at most events_per_round events per surface; one round is a
 * fixed sequence of kernels (counts, scan, per-surface event multinomial,
 * set-up, trials) captured once in a hipGraph and replayed, every size read
 * on the device.  Each surface's output follows the one-surface law.  The
 * hadron list (iss_get_hadrons) is round-major, then surface, then event:
 * iss_get_batch_layout gives segments[(r S + k) 2 + {0, 1}] = (first hadron,
 * count) of round r, surface k, and event_sizes[(r S + k) R + e] the size of
 * its e-th event (event r R + e of surface k; 0 beyond the surface's
 * events).  iss_compute_yields and iss_sample treat the record as one
 * surface. ---- */
typedef struct {
    uint64_t seed;
    int32_t hydro_mode;          /* 1 boost invariant (y in [y_LB, y_RB]), 2 3+1D */
    double y_LB, y_RB;
    int32_t events_per_round;    /* R (at most 16384); 0: the autotuned value (default 16) */
    int32_t max_trials;
    int32_t precision;           /* yields: 64 or 32 */
    int32_t use_graph;           /* 1: one captured round replayed; 0: launched per round */
    int32_t proposal;            /* 0 plain rejection, 1 learned (needs network parameters); the
                                    species of every slot are the same under both */
} iss_batch_opts;

int iss_set_surfaces_lrf(iss_ctx* ctx, int32_t n_surfaces, const int64_t* n_cells, const float* cells28);
int iss_num_surfaces(const iss_ctx* ctx);
int iss_sample_batch(iss_ctx* ctx, const iss_batch_opts* opts, const int32_t* n_events, int64_t* n_hadrons);
int iss_batch_rounds(const iss_ctx* ctx);
int iss_batch_events_per_round(const iss_ctx* ctx);
double iss_batch_ms(const iss_ctx* ctx);   /* device time of the last batch (all rounds) */
int iss_get_batch_layout(iss_ctx* ctx, int64_t* segments, int64_t* event_sizes);
/* times whole batches for R in {4, 16, 64, 256, 1024, 4096, 16384}, keeps the
 * fastest as the default events_per_round of that proposal (the learned one
 * also times its routing thresholds at that R); report: refill_min = R
 * chosen, ms_* batch times, learned_min_hadrons */
int iss_autotune_batch(iss_ctx* ctx, const iss_batch_opts* opts, const int32_t* n_events, iss_tune_report* report);

/* ---- resonance decays: particle_decay (src/particle_decay.cpp:33-533) driven
 * to stable particles as FSSW::perform_resonance_feed_down
 * (src/FSSW.cpp:1746-1779).  The table is the PDG list with its channels as
 * particle_decay reads it (anti-baryon entries with conjugated daughters);
 * entry i owns channels [first_channel[i], first_channel[i] + n_channels[i]),
 * daughters[5 j .. 5 j + 4] the monvals of channel j (0 = none).  A particle
 * whose first channel has one daughter is stable.  iss_decay replaces the
 * hadron list of the last iss_sample by its stable descendants, event by
 * event (iss_get_event_counts / iss_get_hadrons read the result): two- and
 * three-body channels at the pole mass, other channel sizes leave no
 * daughters, as in the reference.  After iss_sample_batch it decays the
 * batch's events the same way and rewrites the batch layout
 * (iss_get_batch_layout) for the decayed list. ---- */
typedef struct {
    int32_t n;
    const int32_t* monval;
    const double* mass;
    const double* width;
    const int32_t* stable;
    const int32_t* first_channel;
    const int32_t* n_channels;
    int32_t n_channels_total;
    const int32_t* npart;
    const double* branching_ratio;
    const int32_t* daughters;
} iss_decay_table;

int iss_set_decay_table(iss_ctx* ctx, const iss_decay_table* table);
int iss_decay(iss_ctx* ctx, uint64_t seed, int64_t* n_hadrons);
/* an external hadron list (n_events events of counts[e] hadrons, grouped by
 * event) as the context's list, e.g. to decay it (no counterpart: the
 * reference decays its own Hadron_list) */
int iss_set_hadrons(iss_ctx* ctx, int32_t n_events, const int64_t* counts, const iss_hadron* hadrons);

/* ---- flows: EmissionFunctionArray::calculate_flows
 * (src/emissionfunction.cpp:875-1023) on one spectrum dN[n_pT][n_phi].
 * vn_diff [n_pT][3 + 3 to_order]: pT, mT - m, dN/(pT dpT dy) / (2 pi)... as the
 *   reference's thermal_*_vndata rows (pT, mT - m, norm / 2pi, then per order
 *   re, im, |v_n|); vn_inte [to_order + 1][6]: order, numerator re, im, v_n re,
 *   im, |v_n|, row 0 holding the normalisation.  Host arithmetic, no context. ---- */
int iss_flows(const double* dN, const double* pT, const double* pT_weight, int32_t n_pT, const double* phi,
              const double* phi_weight, int32_t n_phi, double mass, int32_t to_order, double* vn_diff,
              double* vn_inte);

#ifdef __cplusplus
}
#endif

#endif /* ISS_HIP_H_ */
