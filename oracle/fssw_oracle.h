// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement, in FP64 with std::mt19937, of the iSS "FSSW" Cooper-Frye
// sampling path (chunshen1987/iSS, MC_sampling = 4).  It is the checker that the
// HIP product is checked against.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product never links, imports or
// calls it.
//
// Each function names the reference lines it follows.  Parity status per part:
//   * species list, mass sort ........ pinned vs oracle/_ref (reference reader)
//   * surface reader + EoS regulation  pinned vs oracle/_ref (reference reader)
//   * K1/K2/K3 table ................. pinned vs scipy.special.kn golden fixture
//   * yields ......................... pinned by closed forms (Boltzmann K2
//                                      formula, BE/FD quadrature, truncation rule)
//   * LRF transform .................. pinned by Lorentz invariants
//   * |p| sampler .................... pinned draw-for-draw vs oracle/_ref
//                                      (reference MomentumSampler* sources)
//   * cell pick (inverse CDF) ........ pinned draw-for-draw vs oracle/_ref
//   * Poisson multiplicity ........... law pinned (mean/var/KS vs scipy);
//                                      draw-level parity UNPINNED (GSL absent)
//   * full sampler ................... pinned by the reference's T^{mu nu}
//                                      reconstruction test (1e-3 GeV/fm^3)
//   * smooth spectra ................. pinned by closed forms (Boltzmann and
//                                      quantum-statistics K1 series, boosted
//                                      cell, shear cos 2phi term, bulk K0/K2)
//   * flows .......................... pinned by Bessel I_n ratios and by
//                                      pure Fourier inputs
//   * resonance decays ............... pinned draw-for-draw vs oracle/_ref
//                                      (the reference's particle_decay sources)
//                                      and by conservation laws
#pragma once

#include <array>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace iss_oracle {

constexpr double kHbarC = 0.197327053;  // src/data_struct.h:9

// particle_info (src/data_struct.h:34-50), decay channels omitted.
struct Particle {
    int monval = 0;
    std::string name;
    double mass = 0., width = 0.;
    int gspin = 0, baryon = 0, strange = 0, charm = 0, bottom = 0;
    int gisospin = 0, charge = 0, decays = 0, stable = 0;
    int sign = 0;
};

// FO_surf (src/data_struct.h:53-63) without the PCE vector.
struct Surf {
    float tau, xpt, ypt, eta;
    float da0, da1, da2, da3;
    float u0, u1, u2, u3;
    float Edec, Tdec, Pdec;
    float Bn, muB, muS, muQ;
    float pi00, pi01, pi02, pi03, pi11, pi12, pi13, pi22, pi23, pi33;
    float bulkPi;
    float qmu0, qmu1, qmu2, qmu3;
};

// FO_surf_LRF (src/data_struct.h:66-76): 28 floats, in this order.
struct SurfLRF {
    float tau, xpt, ypt, eta;
    float da[4];    // da_mu_LRF
    float u[4];     // u_tz
    float Edec, Tdec, Pdec;
    float Bn, muB, muS, muQ;
    float bulkPi;
    float pixx, pixy, pixz, piyy, piyz;
    float qx, qy, qz;
};
static_assert(sizeof(SurfLRF) == 28 * sizeof(float), "SurfLRF layout");

// iSS_Hadron (src/data_struct.h:79-84).
struct Hadron {
    int pid;
    float mass;
    float E, px, py, pz;
    float t, x, y, z;
};

struct Config {
    int hydro_mode = 2;              // src/FSSW.cpp:67
    int afterburner = 1;             // 0 PDG, 1 UrQMD, 2 SMASH
    int quantum_statistics = 1;      // src/readindata.cpp:37-41
    int include_shear = 0;           // src/FSSW.cpp:72
    int include_bulk = 0;            // src/FSSW.cpp:73
    int bulk_kind = 1;               // src/FSSW.cpp:74
    int include_diffusion = 0;       // src/FSSW.cpp:75 (not restated: rejected)
    int sampling_model = 30;         // src/FSSW.cpp:885
    double sampling_para1 = 0.16;    // src/FSSW.cpp:886
    double y_LB = -5., y_RB = 5.;    // src/FSSW.cpp:950-951
    // Reproduce FermionMomentumSampler::CDF_2's missing alternating sign
    // (src/FermionMomentumSampler.cpp:80-93).  1 = as the reference does.
    int fermion_cdf2_quirk = 1;
};

// ---------------------------------------------------------------- readers
std::vector<Particle> read_resonances(const std::string& file, bool quantum_statistics);
std::vector<int> read_chosen(const std::string& file);
std::vector<int> sampling_order(const std::vector<Particle>& pdg,
                                const std::vector<int>& chosen_monvals);

// MUSIC surface (src/readindata.cpp:395-546 boost invariant, 626-765 3+1D).
std::vector<Surf> read_surface_music(const std::string& file, bool binary,
                                     bool boost_invariant, int turn_on_bulk,
                                     int turn_on_rhob, int turn_on_diff);

struct HRGEoS {
    int iEOS = 0;
    std::vector<std::array<double, 7>> rows;  // {e, nB, P, T, muB, muS, muQ}
    bool load(const std::string& file, int iEOS_music);
    int lookup(double ed, double nB, double out[5]) const;  // {P,T,muB,muS,muQ}
};
void regulate_surface(std::vector<Surf>& cells, const HRGEoS* eos);
std::vector<SurfLRF> transform_to_lrf(const std::vector<Surf>& cells);
void surface_tmunu(const std::vector<Surf>& cells, double T[16], double Q[3]);

// ------------------------------------------------------ special functions
double bessel_kn(int n, double x);  // own quadrature; pinned vs scipy
double expint_en(int n, double x);  // own quadrature; pinned vs scipy

struct SpecialFunctions {
    static constexpr double x_min = 0.5, x_max = 400., dx = 0.05;
    int length = 0;
    std::vector<double> K1, K2, K3;
    std::vector<double> En;   // [9][length]: E_2, E_4, ..., E_18 (diffusion only)
    void init(bool with_bulk, bool with_diffusion = false);
    double K(int order, double arg) const;
    void E(double arg, double out[9]) const;   // get_special_function_En (src/FSSW.cpp:1706-1726)
};

// ------------------------------------------------------ delta-f tables
struct DeltafTables {
    // kind 0: T[1/fm], B0, D0, E0 (src/FSSW.cpp:1088-1097)
    std::vector<double> k0_T, k0_B0, k0_D0, k0_E0;
    // kind 11: OSU 14-moment (src/FSSW.cpp:1215-1301); c0 loaded correctly
    int k11_nT = 0, k11_nmu = 0;
    double k11_T0 = 0, k11_mu0 = 0, k11_dT = 0, k11_dmu = 0;
    std::vector<double> k11_c0, k11_c1, k11_c2;   // [iT*nmu + imu]
    // kinds 20/21: NEoS BQS table, 200 x 200 rows of 5 (CE) or 8 (22-mom)
    int neos_ncol = 0;
    std::vector<double> neos;                       // [row*ncol + col]
    double neos_de = 0.;
    // diffusion: Coefficients_RTA_diffusion.dat, [iT * 100 + imu] (src/FSSW.cpp:1546-1566)
    std::vector<double> qmu;
    bool load_qmu(const std::string& file);
    double qmu_coeff(double T, double muB) const;   // get_deltaf_qmu_coeff (:1569-1606)
    bool load_kind0(const std::string& file);
    bool load_kind11(const std::string& dir);
    bool load_neos(const std::string& file, int ncol);
};

// per-cell delta-f coefficients (src/FSSW.cpp:1074-1212, 1379-1543)
void bulk_coefficients_T(int kind, const DeltafTables& tb, double Tdec, double c[3]);
void bulk_coefficients_14mom(const DeltafTables& tb, double Tdec, double muB, double c[3]);
void neos_coefficients(const DeltafTables& tb, double Edec, double nB, double* c);

// -------------------------------------------------------- |p| samplers
// MomentumSamplerBase/Boson/Fermion (src/MomentumSamplerBase.cpp:12-93,
// src/BosonMomentumSampler.cpp:7-87, src/FermionMomentumSampler.cpp:8-93).
struct PSamplerTable {
    bool fermion = false;
    double m0 = 0.;
    int trunc = 1;
    bool cdf2_quirk = true;
    std::vector<double> Et, C0, C1, C2;
    double cache_key = 0., c0m = 0., c1m = 0., c2m = 0.;
    void build(bool is_fermion, double m0_in, int trunc_in, bool quirk);
    double cdf(int k, double Etilde) const;
};

struct PSampler {
    PSamplerTable tb[6];  // boson 0,1,2 then fermion 0,1,2
    void build(bool quirk);
    // MomentumSamplerShell::Sample_a_momentum (src/MomentumSamplerShell.cpp:24-48)
    double sample(double m, double T, double mu, int sign, std::mt19937& g,
                  std::uniform_real_distribution<double>& U);
};

// ---------------------------------------------------------------- FSSW
class FSSWOracle {
 public:
    Config cfg;
    std::vector<Particle> particles;     // full PDG list (with anti-baryons)
    std::vector<int> order;              // chosen species, sorted by mass
    std::vector<SurfLRF> cells;
    SpecialFunctions sf;
    DeltafTables dft;
    PSampler psampler;
    std::vector<std::vector<Hadron>> events;
    long n_failed = 0;

    // table_dir holds the reference's iSS_tables layout.
    bool setup(const std::string& table_dir, const Config& c, std::string* err);
    int neos_kind() const;

    // calculate_dN_analytic (src/FSSW.cpp:719-848), results[0..3]
    void dN_analytic(const Particle& p, double mu, double T, double res[4]) const;
    // calculate_dN_dxtdy_for_one_particle_species (src/FSSW.cpp:565-715)
    void yields_one_species(int real_idx, std::vector<double>& out) const;

    // sample_using_dN_dxtdy_4all_particles_conventional (src/FSSW.cpp:873-1071)
    // with the per-(event, species) multiplicities drawn by std::poisson.
    void sample(uint32_t seed, int nev);

    int sample_from_cell(double mass, int sign, int baryon, int strange, int charge,
                         const SurfLRF& s, const double* coeff, std::mt19937& g,
                         std::uniform_real_distribution<double>& U,
                         double& pT, double& phi, double& y_minus_eta);
};

// construct_Tmunu_from_particle_samples (src/iSS.cpp:296-363)
void tmunu_from_samples(const std::vector<std::vector<Hadron>>& events,
                        const std::vector<Particle>& particles, double volume,
                        double T[16], double Q[3]);

// ------------------------------------------------ smooth spectra and flows
struct SpectraOpts {
    int mode = 0;                // 0: eta_s table integral, 1: fixed rapidities
    int use_pos_dN_only = 0;
    int restrict_deltaf = 1;
    double deltaf_max_ratio = 1.0;
};
// dN/(pT dpT dphi dy), out[(plane * npT + i) * nphi + j]
void smooth_spectra(const FSSWOracle& o, const std::vector<Surf>& lab, int real_idx, const SpectraOpts& so,
                    const std::vector<double>& pT, const std::vector<double>& phi,
                    const std::vector<double>& eta, const std::vector<double>& w,
                    std::vector<double>& out);
void flows(const std::vector<double>& dN, const std::vector<double>& pT, const std::vector<double>& pTw,
           const std::vector<double>& phi, const std::vector<double>& phiw, double mass, int to_order,
           std::vector<double>& diff, std::vector<double>& inte);

// ---------------------------------------------------------- resonance decays
struct DecayChannel {
    int npart = 0;
    double br = 0.;
    int part[5] = {0, 0, 0, 0, 0};
};
struct DecayParticle {
    int monval = 0;
    double mass = 0., width = 0.;
    int baryon = 0, strange = 0, charge = 0, stable = 0;
    std::vector<DecayChannel> ch;
};
class Decayer {
 public:
    std::vector<DecayParticle> table;
    // decays below threshold (pole mass under the daughters' masses): the
    // reference exits there (src/particle_decay.cpp:351-357, 431-441); here
    // the mother is dropped and counted
    mutable long n_below = 0;
    // particle_decay::read_resonances_list (src/particle_decay.cpp:33-170)
    bool load(const std::string& pdg_file);
    const DecayParticle* find(int monval) const;
    // particle_decay::perform_decays (src/particle_decay.cpp:262-320)
    void decay(const Hadron& m, std::vector<Hadron>& out, std::mt19937& g,
               std::uniform_real_distribution<double>& U) const;
    // FSSW::perform_resonance_feed_down (src/FSSW.cpp:1746-1779) on one event
    void feed_down(std::vector<Hadron>& event, std::mt19937& g,
                   std::uniform_real_distribution<double>& U) const;
};

}  // namespace iss_oracle
