// The iSS drop-in class over the gfx950 C-ABI.  Host work here is file I/O,
// table set-up and the per-event hadron lists; the surface transform, yields,
// multiplicities and momentum sampling run in the HIP library.
#include "iSS.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <random>

#include "iss_hip.h"
#include "surface_io.h"

namespace {
void info(const std::string& m) { std::cout << "[Info] " << m << std::endl; }
void warn(const std::string& m) { std::cout << "[Warning] " << m << std::endl; }
void error(const std::string& m) { std::cout << "[Error] " << m << std::endl; }
}  // namespace

iSS::iSS(std::string path, std::string table_path, std::string particle_table_path,
         std::string inputfile, std::string surface_filename)
    : path_(path), table_path_(table_path), particle_table_path_(particle_table_path),
      surface_filename_(surface_filename) {
    paraRdr_ptr = new ParameterReader;
    paraRdr_ptr->readFromFile(inputfile);
}

iSS::~iSS() {
    clear();
    release_hadrons();
    if (ctx_) iss_ctx_destroy(ctx_);
    delete paraRdr_ptr;
}

void iSS::release_hadrons() {
    for (auto* e : hadron_list_) delete e;
    hadron_list_.clear();
}

void iSS::clear() {
    FOsurf_array_.clear();
    FOsurf_LRF_array_.clear();
    FOsurf_Tmunu_.clear();
    for (auto& p : particle_)
        for (auto* d : p.decay_channels) delete d;
    particle_.clear();
}

int iSS::ensure_context() {
    if (ctx_) return 0;
    const int dev = static_cast<int>(paraRdr_ptr->getVal("gpu_device", 0));
    const int rc = iss_ctx_create(dev, &ctx_);
    if (rc != ISS_OK) {
        error("no usable GPU for the sampler (iss_ctx_create status " + std::to_string(rc) + ")");
        ctx_ = nullptr;
        return rc;
    }
    return 0;
}

// src/iSS.cpp:41-56
int iSS::shell() {
    int status = read_in_FO_surface();
    if (status != 0) {
        error("Some errors happened in reading in the hyper-surface");
        std::exit(-1);
    }
    set_random_seed();
    status = generate_samples();
    if (status != 0) {
        error("Some errors happened in generating particle samples");
        std::exit(-1);
    }
    return 0;
}

// src/iSS.cpp:86-113 with the reader of src/readindata.cpp:25-111: music_input
// switches, the afterburner (EOS 9 / 91 force UrQMD / SMASH), the surface,
// EoS regulation for EOS 9, 91, 12, 14 and the PDG list.
int iSS::read_in_FO_surface() {
    const int mode = static_cast<int>(paraRdr_ptr->getVal("hydro_mode"));
    if (mode != 1 && mode != 2) {
        error("hydro_mode " + std::to_string(mode) + " is not supported (MUSIC modes 1 and 2 are)");
        return -1;
    }
    auto mi = iss_io::read_music_input(path_, static_cast<int>(paraRdr_ptr->getVal("turn_on_bulk")),
                                       static_cast<int>(paraRdr_ptr->getVal("turn_on_rhob")),
                                       static_cast<int>(paraRdr_ptr->getVal("turn_on_diff")));
    if (!mi.found) {
        error("read_FOdata: can not find configuration file " + path_ + "/music_input");
        return -1;
    }
    iEOS_ = mi.iEOS;
    const int ab = static_cast<int>(paraRdr_ptr->getVal("afterburner_type"));
    afterburner_type_ = (ab == 1) ? AfterburnerType::UrQMD
                                  : (ab == 2 ? AfterburnerType::SMASH : AfterburnerType::PDG_Decay);
    if (iEOS_ == 9) afterburner_type_ = AfterburnerType::UrQMD;
    if (iEOS_ == 91) afterburner_type_ = AfterburnerType::SMASH;

    std::string err;
    std::vector<FO_surf> cells;
    if (!iss_io::read_surface(path_ + "/" + surface_filename_, mode, mi, cells, &err)) {
        error(err);
        return -1;
    }
    info("total number of cells: " + std::to_string(cells.size()));
    if (cells.empty()) {
        warn("No freeze-out fluid cell, exit now ...");
        std::exit(1);
    }
    std::vector<double> eos;
    const bool use_eos = (iEOS_ == 9 || iEOS_ == 91 || iEOS_ == 12 || iEOS_ == 14);
    if (use_eos) {
        std::string f = table_path_ + "/EOS_tables/";
        f += (iEOS_ == 9 || iEOS_ == 91) ? "HRGEOS_PST-" : (iEOS_ == 12 ? "HRGNEOS_B-" : "HRGNEOS_BQS-");
        f += (afterburner_type_ == AfterburnerType::SMASH) ? "SMASH.dat" : "urqmd_v3.3+.dat";
        if (!iss_io::read_hrg_eos(f, iEOS_, eos)) {
            error("Can not found EOS file: " + f);
            return -1;
        }
    }
    const bool qs = paraRdr_ptr->getVal("quantum_statistics") == 1;
    std::string pdg = particle_table_path_;
    pdg += (afterburner_type_ == AfterburnerType::SMASH) ? "/pdg-SMASH.dat"
           : (afterburner_type_ == AfterburnerType::UrQMD ? "/pdg-urqmd_v3.3+.dat" : "/pdg-s95pv1.dat");
    for (auto& p : particle_)
        for (auto* d : p.decay_channels) delete d;
    particle_ = iss_io::read_pdg(pdg, qs);
    if (particle_.empty()) {
        error("Can not found pdg file: " + pdg);
        return -1;
    }
    info("total number of particle species: " + std::to_string(particle_.size()));

    // the GPU regulates and boosts the lab-frame cells (34 floats each)
    if (ensure_context()) return -1;
    std::vector<float> lab(34 * cells.size());
    for (size_t i = 0; i < cells.size(); i++) {
        const FO_surf& s = cells[i];
        const float v[34] = {s.tau, s.xpt, s.ypt, s.eta, s.da0, s.da1, s.da2, s.da3, s.u0, s.u1, s.u2,
                             s.u3, s.Edec, s.Tdec, s.Pdec, s.Bn, s.muB, s.muS, s.muQ, s.pi00, s.pi01,
                             s.pi02, s.pi03, s.pi11, s.pi12, s.pi13, s.pi22, s.pi23, s.pi33, s.bulkPi,
                             s.qmu0, s.qmu1, s.qmu2, s.qmu3};
        std::copy(v, v + 34, lab.begin() + 34 * i);
    }
    iss_eos_desc ed{iEOS_, static_cast<int64_t>(eos.size() / 7), eos.data()};
    int rc = iss_set_surface_lab(ctx_, static_cast<int64_t>(cells.size()), lab.data(), use_eos ? &ed : nullptr);
    if (rc != ISS_OK) {
        error(std::string("iss_set_surface_lab: ") + iss_last_error(ctx_));
        return -1;
    }
    // host copy of the regulated cells for T^{mu nu} (src/iSS.cpp:378-445)
    iss_io::regulate(cells, use_eos ? &eos : nullptr, iEOS_);
    computeFOSurfTmunu(cells);
    FOsurf_array_ = cells;
    transform_to_local_rest_frame(FOsurf_array_, FOsurf_LRF_array_);
    info(" -- Read in data finished!");
    return 0;
}

// The rest-frame cells as the GPU computed them (src/iSS.cpp:170-293).
void iSS::transform_to_local_rest_frame(std::vector<FO_surf>&, std::vector<FO_surf_LRF>& out) {
    out.clear();
    if (!ctx_) return;
    const int64_t n = iss_num_cells(ctx_);
    std::vector<float> buf(28 * static_cast<size_t>(n));
    if (iss_get_surface_lrf(ctx_, buf.data()) != ISS_OK) return;
    out.resize(n);
    for (int64_t i = 0; i < n; i++) {
        const float* v = &buf[28 * i];
        FO_surf_LRF& r = out[i];
        r.tau = v[0]; r.xpt = v[1]; r.ypt = v[2]; r.eta = v[3];
        for (int k = 0; k < 4; k++) { r.da_mu_LRF[k] = v[4 + k]; r.u_tz[k] = v[8 + k]; }
        r.Edec = v[12]; r.Tdec = v[13]; r.Pdec = v[14]; r.Bn = v[15];
        r.muB = v[16]; r.muS = v[17]; r.muQ = v[18]; r.bulkPi = v[19];
        r.piLRF_xx = v[20]; r.piLRF_xy = v[21]; r.piLRF_xz = v[22]; r.piLRF_yy = v[23]; r.piLRF_yz = v[24];
        r.qmuLRF_x = v[25]; r.qmuLRF_y = v[26]; r.qmuLRF_z = v[27];
    }
}

// src/iSS.cpp:116-127: randomSeed < 0 draws the seed from the OS
void iSS::set_random_seed() { set_random_seed(static_cast<int>(paraRdr_ptr->getVal("randomSeed"))); }

void iSS::set_random_seed(int seed) {
    randomSeed_ = seed;
    if (seed == -1) {
        std::random_device rd;
        randomSeed_ = rd();
    }
    seeded_ = true;
}

int iSS::upload_species_and_tables(std::string* err) {
    // chosen particles, sorted by mass (src/iSS.cpp:133-142, src/FSSW.cpp:116-162)
    std::string chosen = particle_table_path_;
    chosen += (afterburner_type_ == AfterburnerType::SMASH) ? "/chosen_particles_SMASH.dat"
              : (afterburner_type_ == AfterburnerType::UrQMD ? "/chosen_particles_urqmd_v3.3+.dat"
                                                             : "/chosen_particles_s95p-v1.dat");
    std::vector<int> missing;
    const auto order = iss_io::sampling_order(particle_, iss_io::read_chosen(chosen), &missing);
    if (!missing.empty()) warn("not all chosen particles are in the pdg particle list!");
    const int n = static_cast<int>(order.size());
    std::vector<int> pid(n), g(n), B(n), S(n), Q(n), sg(n);
    std::vector<double> m(n);
    for (int i = 0; i < n; i++) {
        const particle_info& p = particle_[order[i]];
        pid[i] = p.monval; m[i] = p.mass; g[i] = p.gspin; B[i] = p.baryon;
        S[i] = p.strange; Q[i] = p.charge; sg[i] = p.sign;
    }
    species_pid_ = pid;
    species_index_ = order;
    iss_species_desc sd{n, pid.data(), m.data(), g.data(), B.data(), S.data(), Q.data(), sg.data()};
    int rc = iss_set_species(ctx_, &sd);
    if (rc != ISS_OK) {
        *err = iss_last_error(ctx_);
        return rc;
    }
    // delta-f switches and tables (src/FSSW.cpp:72-83, 175-195)
    iss_deltaf_desc df{};
    df.include_shear = static_cast<int>(paraRdr_ptr->getVal("include_deltaf_shear"));
    df.include_bulk = static_cast<int>(paraRdr_ptr->getVal("include_deltaf_bulk"));
    df.bulk_kind = static_cast<int>(paraRdr_ptr->getVal("bulk_deltaf_kind"));
    df.include_diffusion = static_cast<int>(paraRdr_ptr->getVal("include_deltaf_diffusion"));
    df.fermion_cdf2_quirk = static_cast<int>(paraRdr_ptr->getVal("fermion_cdf2_reference_quirk", 1));
    const bool smash = afterburner_type_ == AfterburnerType::SMASH;
    iss_io::Kind11 k11;
    std::vector<double> neos, qmu;
    if (df.include_diffusion == 1) {
        const std::string f = table_path_ + "/deltaf_tables/Coefficients_RTA_diffusion.dat";
        if (!iss_io::read_qmu_table(f, qmu)) {
            *err = "Can not found file: " + f;
            return ISS_ERR_ARG;
        }
        df.qmu_table = qmu.data();
    }
    if (df.include_bulk == 1 && df.bulk_kind == 11) {
        const std::string d = table_path_ + "/deltaf_tables" + (smash ? "/smash_box" : "/urqmd");
        if (!iss_io::read_kind11(d, k11)) {
            *err = "Can not found file: " + d + "/c0.dat";
            return ISS_ERR_ARG;
        }
        df.k11_nT = k11.nT; df.k11_nmu = k11.nmu; df.k11_T0 = k11.T0; df.k11_mu0 = k11.mu0;
        df.k11_dT = k11.dT; df.k11_dmu = k11.dmu;
        df.k11_c0 = k11.c0.data(); df.k11_c1 = k11.c1.data(); df.k11_c2 = k11.c2.data();
    }
    if (df.bulk_kind == 20 || df.bulk_kind == 21) {
        const int nc = (df.bulk_kind == 21) ? 5 : 8;
        const std::string f = table_path_ + "/deltaf_tables" + (smash ? "/smash" : "/urqmd") +
                              (nc == 5 ? "/NEoSBQS_CE_deltafCoeff.dat" : "/NEoSBQS_22mom_deltafCoeff.dat");
        if (!iss_io::read_neos(f, nc, neos)) {
            *err = "Can not found file: " + f;
            return ISS_ERR_ARG;
        }
        df.neos_ncol = nc;
        df.neos = neos.data();
    }
    rc = iss_set_deltaf(ctx_, &df);
    if (rc != ISS_OK) *err = iss_last_error(ctx_);
    return rc;
}

// src/iSS.cpp:130-165 with FSSW::shell (src/FSSW.cpp:344-361): yields,
// multiplicities and momenta on the GPU, per-event hadron lists, writers.
int iSS::generate_samples() {
    const int mc = static_cast<int>(paraRdr_ptr->getVal("MC_sampling"));
    if (mc != 4 && mc != 0) {
        error("only MC_sampling = 4 (FSSW) and 0 (smooth spectra only) are provided by this sampler");
        return -1;
    }
    if (ensure_context()) return -1;
    if (!seeded_) set_random_seed();
    std::string err;
    int rc = upload_species_and_tables(&err);
    if (rc != ISS_OK) {
        error(err);
        return rc;
    }
    // EmissionFunctionArray::shell (src/emissionfunction.cpp:2542-2561)
    if (static_cast<int>(paraRdr_ptr->getVal("calculate_vn", 0)) == 1) {
        rc = calculate_spectra_and_flows();
        if (rc != 0) return rc;
    }
    if (mc == 0) return 0;
    const int hydro_mode = static_cast<int>(paraRdr_ptr->getVal("hydro_mode"));
    const double yLB = paraRdr_ptr->getVal("y_LB"), yRB = paraRdr_ptr->getVal("y_RB");
    const int precision = static_cast<int>(paraRdr_ptr->getVal("yield_precision", 64));
    iss_yield_opts yo{precision, hydro_mode, yLB, yRB};
    species_dNdy_.assign(species_pid_.size(), 0.);
    rc = iss_compute_yields(ctx_, &yo, species_dNdy_.data(), nullptr);
    if (rc != ISS_OK) {
        error(std::string("iss_compute_yields: ") + iss_last_error(ctx_));
        return rc;
    }
    int nev = static_cast<int>(paraRdr_ptr->getVal("number_of_repeated_sampling"));
    if (static_cast<int>(paraRdr_ptr->getVal("sample_upto_desired_particle_number")) == 1) {
        // compute_number_of_sampling_needed (src/FSSW.cpp:851-870): thermal pi+
        double pion = 0.;
        for (size_t i = 0; i < species_pid_.size(); i++)
            if (species_pid_[i] == 211) pion = species_dNdy_[i];
        const double need = paraRdr_ptr->getVal("number_of_particles_needed");
        int n = static_cast<int>(need / (6. * pion));
        if (hydro_mode == 2) n *= 10;
        const int mx = static_cast<int>(paraRdr_ptr->getVal("maximum_sampling_events"));
        nev = std::max(1, std::min(mx, n));
    }
    iss_sample_opts so{};
    so.seed = static_cast<uint64_t>(randomSeed_);
    so.event0 = 0;
    so.n_events = nev;
    so.proposal = 0;
    so.sampling_model = static_cast<int>(paraRdr_ptr->getVal("dN_dy_sampling_model"));
    so.sampling_para1 = paraRdr_ptr->getVal("dN_dy_sampling_para1", 0.16);
    so.hydro_mode = hydro_mode;
    so.y_LB = yLB;
    so.y_RB = yRB;
    so.max_trials = 5000;
    so.precision = precision;
    // FSSW's local_charge_conservation (src/FSSW.cpp:85, 931-938, 1035-1048)
    so.local_charge_conservation = static_cast<int>(paraRdr_ptr->getVal("local_charge_conservation", 0)) == 1;
    release_hadrons();
    const bool decays = static_cast<int>(paraRdr_ptr->getVal("perform_decays", 0)) == 1;
    if (decays) {
        // the decay table as particle_decay reads it (src/particle_decay.cpp:33-170)
        std::vector<int32_t> mv, st, c0, nc, np, dd;
        std::vector<double> ms, wd, br;
        for (const auto& p : particle_) {
            mv.push_back(p.monval);
            ms.push_back(p.mass);
            wd.push_back(p.width);
            st.push_back(p.stable);
            c0.push_back(static_cast<int32_t>(np.size()));
            nc.push_back(static_cast<int32_t>(p.decay_channels.size()));
            for (const auto* d : p.decay_channels) {
                np.push_back(d->decay_Npart);
                br.push_back(d->branching_ratio);
                for (int k = 0; k < 5; k++) dd.push_back(d->decay_part[k]);
            }
        }
        iss_decay_table dt{static_cast<int32_t>(mv.size()), mv.data(), ms.data(), wd.data(), st.data(), c0.data(),
                           nc.data(), static_cast<int32_t>(np.size()), np.data(), br.data(), dd.data()};
        if (iss_set_decay_table(ctx_, &dt) != ISS_OK) {
            error(std::string("iss_set_decay_table: ") + iss_last_error(ctx_));
            return -1;
        }
    }
    // one sampling pass over events [event0, event0 + n): appended to the lists
    // (train: keep the pass's accepted hadrons as the network's training set,
    // before any decays replace them)
    int64_t ntrain = 0;
    auto run_pass = [&](int event0, int n, int proposal, bool train) -> int {
        so.event0 = event0;
        so.n_events = n;
        so.proposal = proposal;
        int64_t nh = 0;
        int r = iss_sample(ctx_, &so, &nh);
        if (r != ISS_OK) {
            error(std::string("iss_sample: ") + iss_last_error(ctx_));
            return r;
        }
        if (train && iss_mlp_collect(ctx_, 10000000, &ntrain) != ISS_OK) {
            error(std::string("learned proposal training set: ") + iss_last_error(ctx_));
            return -1;
        }
        if (decays) {
            // FSSW::shell -> perform_resonance_feed_down (src/FSSW.cpp:344-361, 1746-1779)
            r = iss_decay(ctx_, so.seed ^ 0xD3CA7ull ^ (static_cast<uint64_t>(event0) << 20), &nh);
            if (r != ISS_OK) {
                error(std::string("iss_decay: ") + iss_last_error(ctx_));
                return r;
            }
        }
        std::vector<int64_t> counts(n);
        std::vector<iss_hadron> all(static_cast<size_t>(nh));
        if (iss_get_event_counts(ctx_, counts.data()) != ISS_OK || iss_get_hadrons(ctx_, all.data(), nh) != ISS_OK) {
            error(std::string("hadron read-back: ") + iss_last_error(ctx_));
            return -1;
        }
        size_t k = 0;
        for (int e = 0; e < n; e++) {
            auto* v = new std::vector<iSS_Hadron>();
            v->reserve(counts[e]);
            for (int64_t i = 0; i < counts[e]; i++, k++) {
                const iss_hadron& h = all[k];
                if (h.pid == 0) continue;  // slot that ran out of trials
                v->push_back(iSS_Hadron{h.pid, h.mass, h.E, h.px, h.py, h.pz, h.t, h.x, h.y, h.z});
            }
            hadron_list_.push_back(v);
        }
        return ISS_OK;
    };
    int proposal = static_cast<int>(paraRdr_ptr->getVal("sampling_proposal", 0));
    if (proposal == 1 && so.local_charge_conservation) {
        warn("sampling_proposal = 1 with local_charge_conservation = 1: the plain proposal is used");
        proposal = 0;
    }
    if (proposal == 1 && so.sampling_model != 30) {
        warn("sampling_proposal = 1 with dN_dy_sampling_model != 30: the plain proposal is used");
        proposal = 0;
    }
    if (proposal == 1) {
        // learned proposal: the first events (about 2e6 hadrons) come from the
        // plain sampler; the network is trained on their accepted hadrons and
        // samples the remaining events
        double per_event = 0.;
        for (double v : species_dNdy_) per_event += v;
        if (hydro_mode != 2) per_event *= (yRB - yLB);
        const int n1 = std::max(1, std::min(nev, static_cast<int>(std::ceil(2.0e6 / std::max(per_event, 1e-9)))));
        rc = run_pass(0, n1, 0, n1 < nev);
        if (rc != ISS_OK) return rc;
        if (n1 < nev) {
            iss_train_opts to{8192, 300, 3e-3f, static_cast<uint64_t>(randomSeed_) + 17u};
            iss_train_report tr{};
            if (iss_mlp_init(ctx_, 1) != ISS_OK || iss_mlp_train(ctx_, &to, &tr) != ISS_OK) {
                error(std::string("learned proposal training: ") + iss_last_error(ctx_));
                return -1;
            }
            info("learned proposal trained on " + std::to_string(ntrain) + " hadrons, log envelope mass " +
                 std::to_string(tr.loss_first) + " -> " + std::to_string(tr.loss_last));
            rc = run_pass(n1, nev - n1, 1, false);
            if (rc != ISS_OK) return rc;
        }
    } else {
        rc = run_pass(0, nev, 0, false);
        if (rc != ISS_OK) return rc;
    }
    int64_t nh = 0;
    for (auto* v : hadron_list_) nh += static_cast<int64_t>(v->size());
    iss_stats st{};
    iss_get_stats(ctx_, &st);
    if (st.n_failed > 0)
        warn(std::to_string(st.n_failed) + " hadrons exhausted the maximum number of trials");
    info("sampled " + std::to_string(nh) + " hadrons in " + std::to_string(nev) + " events");

    if (static_cast<int>(paraRdr_ptr->getVal("include_spectators", 0)) != 0)
        warn("include_spectators is not supported by this sampler");
    // writers (src/FSSW.cpp:354-360)
    if (static_cast<int>(paraRdr_ptr->getVal("use_OSCAR_format")) == 1) {
        if (!iss_io::write_oscar("OSCAR.DAT", table_path_ + "/OSCAR_header.txt", hadron_list_)) {
            error("combine_samples_to_OSCAR error: OSCAR header file not found.");
            std::exit(-1);
        }
    } else if (static_cast<int>(paraRdr_ptr->getVal("use_gzip_format")) == 1) {
        iss_io::write_gzip("particle_samples.gz", hadron_list_);
    } else if (static_cast<int>(paraRdr_ptr->getVal("use_binary_format")) == 1) {
        iss_io::write_binary("particle_samples.bin", hadron_list_);
    }
    return 0;
}

namespace {
// two-column bin table (src/Table.cpp:180-...): node, weight
bool read_bin_table(const std::string& file, std::vector<double>& x, std::vector<double>& w) {
    x.clear();
    w.clear();
    std::ifstream f(file);
    if (!f.good()) return false;
    double a, b;
    while (f >> a >> b) {
        x.push_back(a);
        w.push_back(b);
    }
    return !x.empty();
}

// Table::printTable (src/Table.cpp:166-176): rows of "%15.8e  " fields
void print_rows(std::ostream& os, const double* v, int rows, int cols, bool transpose) {
    for (int r = 0; r < rows; r++) {
        for (int c = 0; c < cols; c++)
            os << std::scientific << std::setw(15) << std::setprecision(8)
               << (transpose ? v[static_cast<size_t>(c) * rows + r] : v[static_cast<size_t>(r) * cols + c]) << "  ";
        os << std::endl;
    }
}
}  // namespace

// EmissionFunctionArray::calculate_dN_pTdpTdphidy_and_flows_4all
// (src/emissionfunction.cpp:1129-1218) and its historic-format variant
// (:1031-1120): spectra of all chosen species in one GPU pass over every lab
// cell, flows per species on the host, the reference's files and layout.
int iSS::calculate_spectra_and_flows() {
    std::vector<double> pT, pTw, phi, phiw, eta, etaw;
    const std::string d = table_path_ + "/bin_tables/";
    if (!read_bin_table(d + "pT_gauss_table.dat", pT, pTw) || !read_bin_table(d + "phi_gauss_table.dat", phi, phiw) ||
        !read_bin_table(d + "eta_uni_table.dat", eta, etaw)) {
        error("Table::loadTableFromFile error: the bin tables in " + d + " cannot be opened.");
        return -1;
    }
    const int to_order = static_cast<int>(paraRdr_ptr->getVal("calculate_vn_to_order", 9));
    if (static_cast<int>(paraRdr_ptr->getVal("calculate_dN_dphi", 0)) == 1)
        warn("calculate_dN_dphi = 1 is not supported by this sampler; dN/dphi is not written");
    const int ns = static_cast<int>(species_pid_.size());
    const int npT = static_cast<int>(pT.size()), nphi = static_cast<int>(phi.size());
    std::vector<int32_t> idx(ns);
    for (int i = 0; i < ns; i++) idx[i] = i;
    iss_spectra_opts so{};
    so.precision = static_cast<int>(paraRdr_ptr->getVal("spectra_precision", 64));
    so.mode = 0;
    so.all_cells = 1;
    so.use_pos_dN_only = static_cast<int>(paraRdr_ptr->getVal("use_pos_dN_only", 0));
    so.restrict_deltaf = static_cast<int>(paraRdr_ptr->getVal("restrict_deltaf", 1));
    so.deltaf_max_ratio = paraRdr_ptr->getVal("deltaf_max_ratio", 1.0);
    std::vector<double> dN(static_cast<size_t>(ns) * npT * nphi);
    int rc = iss_spectra(ctx_, &so, idx.data(), ns, pT.data(), npT, phi.data(), nphi, eta.data(), etaw.data(),
                         static_cast<int>(eta.size()), dN.data());
    if (rc != ISS_OK) {
        error(std::string("iss_spectra: ") + iss_last_error(ctx_));
        return rc;
    }
    info("smooth spectra of " + std::to_string(ns) + " species on " + std::to_string(npT) + " x " +
         std::to_string(nphi) + " (pT, phi) x " + std::to_string(eta.size()) + " eta_s: " +
         std::to_string(iss_spectra_ms(ctx_)) + " ms on the GPU");
    // species of each particle-list entry (-1: not chosen)
    std::vector<int> of_particle(particle_.size(), -1);
    for (int i = 0; i < ns; i++)
        if (of_particle[species_index_[i]] < 0) of_particle[species_index_[i]] = i;
    const int nc = 3 + 3 * to_order;
    std::vector<double> diff(static_cast<size_t>(npT) * nc), inte(6 * static_cast<size_t>(to_order + 1));
    std::vector<double> zero(static_cast<size_t>(npT) * nphi, 0.);
    auto flows_of = [&](const double* spec, double mass) {
        return iss_flows(spec, pT.data(), pTw.data(), npT, phi.data(), phiw.data(), nphi, mass, to_order,
                         diff.data(), inte.data());
    };
    const std::string dn_file = path_ + "/dN_pTdpTdphidy.dat";
    std::remove(dn_file.c_str());
    std::ofstream odn(dn_file);
    // dN_pTdpTdphidy tables: one row per phi, one column per pT
    if (static_cast<int>(paraRdr_ptr->getVal("use_historic_flow_output_format", 0)) == 1) {
        const std::string f1 = path_ + "/v2data.dat", f2 = path_ + "/v2data-inte.dat";
        std::ofstream o1(f1), o2(f2);
        for (size_t n = 0; n < particle_.size(); n++) {
            const int s = of_particle[n];
            const double* spec = s >= 0 ? &dN[static_cast<size_t>(s) * npT * nphi] : zero.data();
            print_rows(odn, spec, nphi, npT, true);
            o1 << "# Output for particle: " << particle_[n].name << std::endl;
            o1 << "#                 " << particle_[n].monval << std::endl;
            o2 << "# For: " << particle_[n].name << std::endl;
            if (flows_of(spec, particle_[n].mass) != ISS_OK) return -1;
            print_rows(o1, diff.data(), npT, nc, false);
            print_rows(o2, inte.data(), to_order + 1, 6, false);
        }
        return 0;
    }
    for (int s = 0; s < ns; s++) {
        const particle_info& p = particle_[species_index_[s]];
        if (flows_of(&dN[static_cast<size_t>(s) * npT * nphi], p.mass) != ISS_OK) return -1;
        std::ofstream o1(path_ + "/thermal_" + std::to_string(p.monval) + "_vndata.dat");
        print_rows(o1, diff.data(), npT, nc, false);
        std::ofstream o2(path_ + "/thermal_" + std::to_string(p.monval) + "_integrated__vndata.dat");
        print_rows(o2, inte.data(), to_order + 1, 6, false);
    }
    for (size_t n = 0; n < particle_.size(); n++) {
        const int s = of_particle[n];
        print_rows(odn, s >= 0 ? &dN[static_cast<size_t>(s) * npT * nphi] : zero.data(), nphi, npT, true);
    }
    return 0;
}

// src/iSS.cpp:378-445 (FOsurf_Q_[0] is the last cell's n_B, as there)
void iSS::computeFOSurfTmunu(std::vector<FO_surf>& cells) {
    FOsurf_Tmunu_.assign(16, 0.f);
    FOsurf_Q_.assign(3, 0.f);
    const float g[4] = {1.f, -1.f, -1.f, -1.f};
    for (const auto& s : cells) {
        const float ch = std::cosh(s.eta), sh = std::sinh(s.eta);
        const float u[4] = {s.u0 * ch + s.u3 * sh, s.u1, s.u2, s.u3 * ch + s.u0 * sh};
        float P[4][4];
        P[0][0] = s.pi00 * ch * ch + 2.f * s.pi03 * ch * sh + s.pi33 * sh * sh;
        P[0][1] = s.pi01 * ch + s.pi13 * sh;
        P[0][2] = s.pi02 * ch + s.pi23 * sh;
        P[0][3] = s.pi00 * ch * sh + s.pi03 * (ch * ch + sh * sh) + s.pi33 * sh * ch;
        P[1][1] = s.pi11;
        P[1][2] = s.pi12;
        P[1][3] = s.pi01 * sh + s.pi13 * ch;
        P[2][2] = s.pi22;
        P[2][3] = s.pi02 * sh + s.pi23 * ch;
        P[3][3] = s.pi00 * sh * sh + 2.f * s.pi03 * sh * ch + s.pi33 * ch * ch;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < i; j++) P[i][j] = P[j][i];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                FOsurf_Tmunu_[4 * i + j] += s.Edec * u[i] * u[j] -
                                            (s.Pdec + s.bulkPi) * ((i == j ? g[i] : 0.f) - u[i] * u[j]) +
                                            P[i][j];
        FOsurf_Q_[0] = s.Bn;
        FOsurf_Q_[2] = 0.4f * s.Bn;
    }
}

void iSS::getParticleQuantumNumbers(long monval, std::array<int, 3>& Q) {
    for (const auto& p : particle_)
        if (p.monval == monval) {
            Q = {p.baryon, p.strange, p.charge};
            return;
        }
}

// src/iSS.cpp:296-363: T^{mu nu} and n_B of the samples vs the surface
void iSS::construct_Tmunu_from_particle_samples() {
    if (FOsurf_LRF_array_.empty()) return;
    const double volume = FOsurf_LRF_array_[0].da_mu_LRF[0] / FOsurf_LRF_array_[0].u_tz[0];
    double T[4][4] = {{0}};
    double q[3] = {0, 0, 0};
    long cached = 0;
    std::array<int, 3> Q{0, 0, 0};
    const int nev = get_number_of_sampled_events();
    for (int e = 0; e < nev; e++)
        for (const auto& h : *hadron_list_[e]) {
            if (h.pid != cached) {
                getParticleQuantumNumbers(h.pid, Q);
                cached = h.pid;
            }
            for (int k = 0; k < 3; k++) q[k] += Q[k];
            const double p[4] = {h.E, h.px, h.py, h.pz};
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) T[i][j] += p[i] * p[j] / p[0];
        }
    std::ofstream out("checkReconstructedTmunu.dat");
    out << "# Tmunu_FOcell[GeV/fm^3]  Tmunu_Particles[GeV/fm^3]  diff" << std::endl;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            const double t = T[i][j] / (nev * volume);
            out << std::scientific << std::setprecision(8) << FOsurf_Tmunu_[4 * i + j] << "  " << t
                << "  " << FOsurf_Tmunu_[4 * i + j] - t << std::endl;
        }
    const double nb = q[0] / (nev * volume);
    out << std::scientific << std::setprecision(8) << FOsurf_Q_[0] << "  " << nb << "  "
        << FOsurf_Q_[0] - nb << std::endl;
}

// src/iSS.cpp:59-83
void iSS::perform_checks() {
    info("Performing checks for the samples ...");
    construct_Tmunu_from_particle_samples();
    const int nb = 100;
    const double x0 = 0., x1 = 5., w = (x1 - x0) / (nb - 1);
    for (int which : {211, 2212}) {
        std::vector<double> y(nb, 0.), y2(nb, 0.), xs(nb, 0.), cur(nb, 0.);
        std::vector<long> cnt(nb, 0);
        const int nev = get_number_of_sampled_events();
        for (int e = 0; e < nev; e++) {
            std::fill(cur.begin(), cur.end(), 0.);
            for (const auto& h : *hadron_list_[e]) {
                if (h.pid != which) continue;
                const double pT = std::sqrt(h.px * h.px + h.py * h.py);
                const int i = static_cast<int>((pT - x0) / w);
                if (i >= 0 && i < nb) {
                    xs[i] += pT;
                    cur[i] += 1.;
                    cnt[i]++;
                }
            }
            for (int i = 0; i < nb; i++) {
                y[i] += cur[i];
                y2[i] += cur[i] * cur[i];
            }
        }
        std::ofstream of(which == 211 ? "check_211_spectra.dat" : "check_2212_spectra.dat");
        of << "# x  y  y_err  bin_counts" << std::endl;
        for (int i = 0; i < nb; i++) {
            const double x = cnt[i] ? xs[i] / cnt[i] : x0 + (i + 0.5) * w;
            const double err = std::sqrt(std::max(0., y2[i] / nev - y[i] * y[i] / (1.0 * nev * nev))) /
                               std::sqrt(static_cast<double>(nev));
            of << std::scientific << std::setprecision(6) << x << "  " << y[i] / nev << "  " << err << "  "
               << cnt[i] << std::endl;
        }
    }
}
