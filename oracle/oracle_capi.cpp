// ORACLE — TEST INFRASTRUCTURE ONLY.  extern "C" surface of the CPU oracle for
// the tests and the cpu_baseline leg of bench.py (loaded with ctypes).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

#include "fssw_oracle.h"

using namespace iss_oracle;

namespace {
struct Handle {
    FSSWOracle o;
    std::vector<Surf> lab;
    std::string err;
};
}  // namespace

extern "C" {

void* orc_create(const char* table_dir, int afterburner, int qs, int shear, int bulk,
                 int bulk_kind, int hydro_mode, int model, double para1, double yLB,
                 double yRB, int fd_quirk, int diffusion) {
    Handle* h = new Handle;
    Config c;
    c.afterburner = afterburner;
    c.quantum_statistics = qs;
    c.include_shear = shear;
    c.include_bulk = bulk;
    c.bulk_kind = bulk_kind;
    c.hydro_mode = hydro_mode;
    c.sampling_model = model;
    c.sampling_para1 = para1;
    c.y_LB = yLB;
    c.y_RB = yRB;
    c.fermion_cdf2_quirk = fd_quirk;
    c.include_diffusion = diffusion;
    if (!h->o.setup(table_dir, c, &h->err)) {
        delete h;
        return nullptr;
    }
    return h;
}

void orc_destroy(void* p) { delete static_cast<Handle*>(p); }

int orc_num_species(void* p) { return static_cast<int>(static_cast<Handle*>(p)->o.order.size()); }

int orc_num_particles(void* p) { return static_cast<int>(static_cast<Handle*>(p)->o.particles.size()); }

// species in sampling order (mass sorted)
void orc_species(void* p, int* pid, double* mass, int* g, int* B, int* S, int* Q, int* sign) {
    auto& o = static_cast<Handle*>(p)->o;
    for (size_t i = 0; i < o.order.size(); i++) {
        const Particle& q = o.particles[o.order[i]];
        pid[i] = q.monval; mass[i] = q.mass; g[i] = q.gspin;
        B[i] = q.baryon; S[i] = q.strange; Q[i] = q.charge; sign[i] = q.sign;
    }
}

// full PDG list (with anti-baryons) in file order
void orc_particles(void* p, int* pid, double* mass, int* g, int* B, int* S, int* Q, int* sign) {
    auto& o = static_cast<Handle*>(p)->o;
    for (size_t i = 0; i < o.particles.size(); i++) {
        const Particle& q = o.particles[i];
        pid[i] = q.monval; mass[i] = q.mass; g[i] = q.gspin;
        B[i] = q.baryon; S[i] = q.strange; Q[i] = q.charge; sign[i] = q.sign;
    }
}

void orc_set_cells_lrf(void* p, long n, const float* aos28) {
    auto& o = static_cast<Handle*>(p)->o;
    o.cells.resize(n);
    std::memcpy(o.cells.data(), aos28, sizeof(SurfLRF) * n);
}

long orc_num_cells(void* p) { return static_cast<long>(static_cast<Handle*>(p)->o.cells.size()); }

void orc_get_cells_lrf(void* p, float* out) {
    auto& o = static_cast<Handle*>(p)->o;
    std::memcpy(out, o.cells.data(), sizeof(SurfLRF) * o.cells.size());
}

// read + regulate + LRF transform; eos_file may be empty (no EoS regulation)
long orc_read_surface(void* p, const char* path, int binary, int boost_inv, int bulk, int rhob,
                      int diff, const char* eos_file, int iEOS) {
    Handle* h = static_cast<Handle*>(p);
    h->lab = read_surface_music(path, binary != 0, boost_inv != 0, bulk, rhob, diff);
    HRGEoS eos;
    const HRGEoS* pe = nullptr;
    if (eos_file && eos_file[0]) {
        if (!eos.load(eos_file, iEOS)) return -1;
        pe = &eos;
    }
    regulate_surface(h->lab, pe);
    h->o.cells = transform_to_lrf(h->lab);
    return static_cast<long>(h->o.cells.size());
}

// lab cells after regulation, FO_surf order (34 floats each)
void orc_get_cells_lab(void* p, float* out) {
    Handle* h = static_cast<Handle*>(p);
    std::memcpy(out, h->lab.data(), sizeof(Surf) * h->lab.size());
}

void orc_surface_tmunu(void* p, double* T16, double* Q3) {
    surface_tmunu(static_cast<Handle*>(p)->lab, T16, Q3);
}

// stand-alone LRF transform of n lab cells (34 floats each); returns kept count
long orc_transform(long n, const float* lab34, float* out28, long* kept_index) {
    std::vector<Surf> in(n);
    std::memcpy(in.data(), lab34, sizeof(Surf) * n);
    long k = 0;
    for (long i = 0; i < n; i++) {
        std::vector<Surf> one(1, in[i]);
        auto r = transform_to_lrf(one);
        if (!r.empty()) {
            std::memcpy(out28 + 28 * k, &r[0], sizeof(SurfLRF));
            if (kept_index) kept_index[k] = i;
            k++;
        }
    }
    return k;
}

void orc_regulate(long n, float* lab34) {
    std::vector<Surf> in(n);
    std::memcpy(in.data(), lab34, sizeof(Surf) * n);
    regulate_surface(in, nullptr);
    std::memcpy(lab34, in.data(), sizeof(Surf) * n);
}

// per-(cell, species) yields, species in sampling order: out[c*Ns + s]
void orc_yields(void* p, double* out) {
    auto& o = static_cast<Handle*>(p)->o;
    const size_t Ns = o.order.size(), Nc = o.cells.size();
    std::vector<double> col;
    for (size_t s = 0; s < Ns; s++) {
        o.yields_one_species(o.order[s], col);
        for (size_t c = 0; c < Nc; c++) out[c * Ns + s] = col[c];
    }
}

// yields for a contiguous cell sub-range [c0, c1), all species: used to time
// a bounded sample of the yield pass on the host
double orc_time_yields(void* p, long c0, long c1, int reps) {
    auto& o = static_cast<Handle*>(p)->o;
    std::vector<SurfLRF> all = o.cells;
    o.cells.assign(all.begin() + c0, all.begin() + c1);
    std::vector<double> col;
    double sink = 0.;
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++)
        for (size_t s = 0; s < o.order.size(); s++) {
            o.yields_one_species(o.order[s], col);
            sink += col[0];
        }
    auto t1 = std::chrono::steady_clock::now();
    o.cells = all;
    (void)sink;
    return std::chrono::duration<double>(t1 - t0).count();
}

long orc_sample(void* p, unsigned seed, int nev) {
    auto& o = static_cast<Handle*>(p)->o;
    o.sample(seed, nev);
    long tot = 0;
    for (auto& e : o.events) tot += static_cast<long>(e.size());
    return tot;
}

// The same sampler on `nthreads` host threads (the CPU baseline on all
// cores): thread t samples its share of the nev events with its own copy of
// the oracle and its own seeds, as separate runs of the single-threaded
// reference would.  Returns the total number of hadrons; the handle's own
// event list is left untouched.
long orc_sample_mt(void* p, unsigned seed, int nev, int nthreads) {
    const FSSWOracle& o = static_cast<Handle*>(p)->o;
    nthreads = std::max(1, std::min(nthreads, nev));
    std::vector<FSSWOracle> copies(nthreads, o);
    std::vector<long> tot(nthreads, 0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const int n_t = nev / nthreads + (t < nev % nthreads ? 1 : 0);
        th.emplace_back([&, t, n_t]() {
            copies[t].sample(seed + 7919u * static_cast<unsigned>(t), n_t);
            for (auto& e : copies[t].events) tot[t] += static_cast<long>(e.size());
        });
    }
    for (auto& x : th) x.join();
    long all = 0;
    for (long v : tot) all += v;
    return all;
}

long orc_num_failed(void* p) { return static_cast<Handle*>(p)->o.n_failed; }

void orc_event_counts(void* p, long* counts) {
    auto& o = static_cast<Handle*>(p)->o;
    for (size_t e = 0; e < o.events.size(); e++) counts[e] = static_cast<long>(o.events[e].size());
}

void orc_hadrons(void* p, void* out) {
    auto& o = static_cast<Handle*>(p)->o;
    Hadron* d = static_cast<Hadron*>(out);
    for (auto& e : o.events)
        for (auto& h : e) *d++ = h;
}

void orc_tmunu_samples(void* p, double volume, double* T16, double* Q3) {
    auto& o = static_cast<Handle*>(p)->o;
    tmunu_from_samples(o.events, o.particles, volume, T16, Q3);
}

double orc_bessel_kn(int n, double x) { return bessel_kn(n, x); }
double orc_expint(int n, double x) { return expint_en(n, x); }

double orc_kn_table(void* p, int order, double x) { return static_cast<Handle*>(p)->o.sf.K(order, x); }

// n draws of the |p| sampler with a fresh std::mt19937(seed)
int orc_psample(void* p, unsigned seed, double m, double T, double mu, int sign, int n,
                double* out) {
    auto& o = static_cast<Handle*>(p)->o;
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0., 1.);
    for (int i = 0; i < n; i++) {
        out[i] = o.psampler.sample(m, T, mu, sign, g, U);
        if (out[i] < 0) return -1;
    }
    return 0;
}

// inverse-CDF cell picks with a fresh std::mt19937(seed)
void orc_cdf_pick(const double* vals, long n, unsigned seed, long ndraw, long* out) {
    std::vector<double> cdf(n + 1, 0.);
    double sum = 0.;
    for (long l = 0; l < n; l++) {
        sum += std::max(0., vals[l]);
        cdf[l + 1] = sum;
    }
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0., 1.);
    for (long k = 0; k < ndraw; k++) {
        const double r = cdf[0] + (sum - cdf[0] - 1e-15) * U(g);
        long lo = 0, hi = n, mid = (lo + hi) / 2;
        while (hi - lo > 1) {
            if (cdf[mid] < r) lo = mid; else hi = mid;
            mid = (lo + hi) / 2;
        }
        out[k] = lo;
    }
}

// neos / 14-moment coefficient lookups, for table checks
void orc_neos_coeff(void* p, double e, double nB, double* out) {
    neos_coefficients(static_cast<Handle*>(p)->o.dft, e, nB, out);
}
void orc_k11_coeff(void* p, double T, double muB, double* out) {
    bulk_coefficients_14mom(static_cast<Handle*>(p)->o.dft, T, muB, out);
}
void orc_k1_coeff(void* p, double T, double* out) {
    bulk_coefficients_T(1, static_cast<Handle*>(p)->o.dft, T, out);
}

// lab cells (34 floats each), regulated without an EoS as the surface reader
// does, and their rest-frame copy
long orc_set_cells_lab(void* p, long n, const float* lab34) {
    Handle* h = static_cast<Handle*>(p);
    h->lab.resize(n);
    std::memcpy(h->lab.data(), lab34, sizeof(Surf) * n);
    regulate_surface(h->lab, nullptr);
    h->o.cells = transform_to_lrf(h->lab);
    return static_cast<long>(h->o.cells.size());
}

// smooth spectra of species (sampling-order indices), out[s][plane][pT][phi]
void orc_spectra(void* p, const int* species, int ns, int mode, int pos_only, int restrict_df, double ratio,
                 const double* pT, int npT, const double* phi, int nphi, const double* eta, const double* w,
                 int ny, double* out) {
    Handle* h = static_cast<Handle*>(p);
    SpectraOpts so;
    so.mode = mode;
    so.use_pos_dN_only = pos_only;
    so.restrict_deltaf = restrict_df;
    so.deltaf_max_ratio = ratio;
    std::vector<double> vp(pT, pT + npT), vf(phi, phi + nphi), ve(eta, eta + ny), vw;
    if (w) vw.assign(w, w + ny);
    else vw.assign(ny, 1.);
    std::vector<double> r;
    size_t off = 0;
    for (int s = 0; s < ns; s++) {
        smooth_spectra(h->o, h->lab, h->o.order[species[s]], so, vp, vf, ve, vw, r);
        std::memcpy(out + off, r.data(), sizeof(double) * r.size());
        off += r.size();
    }
}

void orc_flows(const double* dN, const double* pT, const double* pTw, int npT, const double* phi,
               const double* phiw, int nphi, double mass, int to_order, double* diff, double* inte) {
    std::vector<double> d(dN, dN + static_cast<size_t>(npT) * nphi), a(pT, pT + npT), aw(pTw, pTw + npT),
        f(phi, phi + nphi), fw(phiw, phiw + nphi), o1, o2;
    flows(d, a, aw, f, fw, mass, to_order, o1, o2);
    std::memcpy(diff, o1.data(), sizeof(double) * o1.size());
    std::memcpy(inte, o2.data(), sizeof(double) * o2.size());
}

// resonance feed-down of events (hadrons grouped by event, counts[nev]) with
// one mt19937 stream over all events, as the reference's shared generator
static Decayer* g_decayer = nullptr;
static std::string g_decayer_file;

long orc_decay(const char* pdg_file, unsigned seed, long nev, const long* counts, const void* in, long cap,
               long* counts_out, void* out) {
    if (!g_decayer || g_decayer_file != pdg_file) {
        delete g_decayer;
        g_decayer = new Decayer;
        if (!g_decayer->load(pdg_file)) return -1;
        g_decayer_file = pdg_file;
    }
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0., 1.);
    g_decayer->n_below = 0;
    const Hadron* h = static_cast<const Hadron*>(in);
    Hadron* o = static_cast<Hadron*>(out);
    long k = 0, w = 0;
    for (long e = 0; e < nev; e++) {
        std::vector<Hadron> ev(h + k, h + k + counts[e]);
        k += counts[e];
        g_decayer->feed_down(ev, g, U);
        counts_out[e] = static_cast<long>(ev.size());
        for (const auto& x : ev) {
            if (w < cap) o[w] = x;
            w++;
        }
    }
    return w;
}

long orc_decay_below() { return g_decayer ? g_decayer->n_below : 0; }

// the decay table as the decayer holds it: per entry monval, mass, width,
// stable, baryon, strange, charge, first channel; per channel npart, br, 5 daughters
long orc_decay_table(const char* pdg_file, int* ints, double* dbls, int* chan_ints, double* chan_br, long cap_p,
                     long cap_c, long* n_chan) {
    Decayer d;
    if (!d.load(pdg_file)) return -1;
    long nc = 0;
    for (size_t i = 0; i < d.table.size(); i++) {
        const auto& p = d.table[i];
        if (static_cast<long>(i) < cap_p) {
            int* q = ints + 7 * i;
            q[0] = p.monval; q[1] = p.stable; q[2] = p.baryon; q[3] = p.strange; q[4] = p.charge;
            q[5] = static_cast<int>(nc); q[6] = static_cast<int>(p.ch.size());
            dbls[2 * i] = p.mass;
            dbls[2 * i + 1] = p.width;
        }
        for (const auto& c : p.ch) {
            if (nc < cap_c) {
                int* q = chan_ints + 6 * nc;
                q[0] = c.npart;
                for (int k = 0; k < 5; k++) q[1 + k] = c.part[k];
                chan_br[nc] = c.br;
            }
            nc++;
        }
    }
    *n_chan = nc;
    return static_cast<long>(d.table.size());
}

}  // extern "C"
