// ORACLE CHECKER — TEST INFRASTRUCTURE ONLY.
// A thin extern "C" driver over the reference's own GSL-free sources
// (compiled where they lie under the reference tree by ref/Makefile), used to
// pin the oracle's restatement draw-for-draw:
//   * MomentumSamplerShell (src/MomentumSamplerShell.cpp:6-48) + Random
//   * RandomVariable1DArray (src/RandomVariable1DArray.cpp:25-74)
//   * read_FOdata reader + regulation + PDG list (src/readindata.cpp)
//   * particle_decay (src/particle_decay.cpp) driven as
//     FSSW::perform_resonance_feed_down does (src/FSSW.cpp:1746-1779)
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "MomentumSamplerShell.h"
#include "ParameterReader.h"
#include "particle_decay.h"
#include "Random.h"
#include "RandomVariable1DArray.h"
#include "readindata.h"

extern "C" {

int ref_psample(long seed, double m, double T, double mu, int sign, int n, double* out) {
    auto rng = std::make_shared<RandomUtil::Random>(seed);
    MomentumSamplerShell shell(rng);
    for (int i = 0; i < n; i++) out[i] = shell.Sample_a_momentum(m, T, mu, sign);
    return 0;
}

void ref_cdf_pick(const double* vals, long n, long seed, long ndraw, long* out) {
    std::vector<double> v(vals, vals + n);
    auto rng = std::make_shared<RandomUtil::Random>(seed);
    RandomVariable1DArray rv(&v, rng, 0);
    for (long k = 0; k < ndraw; k++) out[k] = rv.rand();
}

// Reads <work_dir>/<surface> with the reference reader (music_input in
// work_dir decides EoS / binary / bulk / rhob), regulates it and reads the PDG
// list.  Cells are returned as 34 floats each in FO_surf field order.
static std::vector<FO_surf> g_cells;
static std::vector<particle_info> g_parts;

long ref_read_surface(const char* work_dir, const char* table_dir, const char* surface,
                      int hydro_mode, int afterburner, int qs, int bulk, int rhob, int diff) {
    ParameterReader pr;
    pr.setVal("hydro_mode", hydro_mode);
    pr.setVal("turn_on_bulk", bulk);
    pr.setVal("turn_on_rhob", rhob);
    pr.setVal("turn_on_diff", diff);
    pr.setVal("quantum_statistics", qs);
    pr.setVal("afterburner_type", afterburner);
    read_FOdata rd(&pr, work_dir, table_dir, table_dir);
    g_cells.clear();
    g_parts.clear();
    rd.read_in_freeze_out_data(g_cells, surface);
    rd.read_in_chemical_potentials(g_cells, g_parts);
    return static_cast<long>(g_cells.size());
}

void ref_get_cells(float* out34) {
    for (size_t i = 0; i < g_cells.size(); i++) {
        const FO_surf& s = g_cells[i];
        const float v[34] = {s.tau, s.xpt, s.ypt, s.eta, s.da0, s.da1, s.da2, s.da3,
                             s.u0, s.u1, s.u2, s.u3, s.Edec, s.Tdec, s.Pdec, s.Bn,
                             s.muB, s.muS, s.muQ, s.pi00, s.pi01, s.pi02, s.pi03, s.pi11,
                             s.pi12, s.pi13, s.pi22, s.pi23, s.pi33, s.bulkPi, s.qmu0,
                             s.qmu1, s.qmu2, s.qmu3};
        std::memcpy(out34 + 34 * i, v, sizeof(v));
    }
}

long ref_num_particles() { return static_cast<long>(g_parts.size()); }

void ref_particles(int* pid, double* mass, int* g, int* B, int* S, int* Q, int* sign) {
    for (size_t i = 0; i < g_parts.size(); i++) {
        pid[i] = g_parts[i].monval; mass[i] = g_parts[i].mass; g[i] = g_parts[i].gspin;
        B[i] = g_parts[i].baryon; S[i] = g_parts[i].strange; Q[i] = g_parts[i].charge;
        sign[i] = g_parts[i].sign;
    }
}

// resonance feed-down of events (counts[nev] hadrons each, iSS_Hadron
// records) with the reference's decayer and one generator for all events
long ref_decay(const char* table_dir, int afterburner, long seed, long nev, const long* counts,
               const void* in, long cap, long* counts_out, void* out) {
    auto rng = std::make_shared<RandomUtil::Random>(seed);
    const AfterburnerType ab = afterburner == 2 ? AfterburnerType::SMASH
                               : (afterburner == 1 ? AfterburnerType::UrQMD : AfterburnerType::PDG_Decay);
    particle_decay dec(rng, ab, table_dir);
    const iSS_Hadron* h = static_cast<const iSS_Hadron*>(in);
    iSS_Hadron* o = static_cast<iSS_Hadron*>(out);
    long k = 0, w = 0;
    for (long e = 0; e < nev; e++) {
        std::vector<iSS_Hadron> work(h + k, h + k + counts[e]);
        k += counts[e];
        long n = 0;
        for (size_t i = 0; i < work.size(); i++) {
            std::vector<iSS_Hadron> ds;
            iSS_Hadron m = work[i];
            dec.perform_decays(&m, &ds);
            for (auto& d : ds) {
                if (dec.check_particle_stable(&d) == 1) {
                    if (w < cap) o[w] = d;
                    w++;
                    n++;
                } else {
                    work.push_back(d);
                }
            }
        }
        counts_out[e] = n;
    }
    return w;
}

}  // extern "C"
