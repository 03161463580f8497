// ORACLE — TEST INFRASTRUCTURE ONLY (see fssw_oracle.h for the rules and the
// parity status of each part).  CPU restatement of chunshen1987/iSS's FSSW path.
#include "fssw_oracle.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>

namespace iss_oracle {

// ===================================================================== readers

// read_FOdata::read_resonances_list (src/readindata.cpp:971-1116): one entry per
// line, its decay channels on the following lines; anti-baryons are appended
// right after their baryon; sign = -1 (BE) for mesons, +1 (FD) for baryons,
// 0 without quantum statistics (src/readindata.cpp:1103-1113).
std::vector<Particle> read_resonances(const std::string& file, bool qs) {
    std::vector<Particle> out;
    std::ifstream f(file);
    if (!f.good()) return out;
    while (true) {
        Particle p;
        if (!(f >> p.monval)) break;
        f >> p.name >> p.mass >> p.width >> p.gspin >> p.baryon >> p.strange >>
            p.charm >> p.bottom >> p.gisospin >> p.charge >> p.decays;
        int first_npart = 0;
        for (int j = 0; j < p.decays; j++) {
            int dummy, npart;
            double br;
            int parts[5];
            f >> dummy >> npart >> br >> parts[0] >> parts[1] >> parts[2] >> parts[3] >>
                parts[4];
            std::string rest;
            std::getline(f, rest);
            if (j == 0) first_npart = npart;
        }
        if (!f) break;
        p.stable = (first_npart == 1) ? 1 : 0;
        out.push_back(p);
        if (p.baryon > 0) {
            Particle a = p;
            a.monval = -p.monval;
            a.name = "Anti-" + p.name;
            a.baryon = -p.baryon;
            a.strange = -p.strange;
            a.charm = -p.charm;
            a.bottom = -p.bottom;
            a.charge = -p.charge;
            out.push_back(a);
        }
    }
    for (auto& p : out) p.sign = qs ? (p.baryon == 0 ? -1 : 1) : 0;
    return out;
}

// Table::loadTableFromFile of the chosen-particle list; column 1 is the
// Monte-Carlo id (src/iSS.cpp:133-142, src/FSSW.cpp:116-131).
std::vector<int> read_chosen(const std::string& file) {
    std::vector<int> out;
    std::ifstream f(file);
    double v;
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        if (ss >> v) out.push_back(static_cast<int>(v));
    }
    return out;
}

// FSSW ctor: chosen monval -> first matching PDG index, then the bubble sort
// by mass with strict '>' (stable) (src/FSSW.cpp:116-162).
std::vector<int> sampling_order(const std::vector<Particle>& pdg,
                                const std::vector<int>& chosen) {
    std::vector<int> idx;
    for (int mv : chosen) {
        for (size_t n = 0; n < pdg.size(); n++) {
            if (pdg[n].monval == mv) {
                idx.push_back(static_cast<int>(n));
                break;
            }
        }
    }
    const int N = static_cast<int>(idx.size());
    for (int m = 0; m < N; m++)
        for (int n = 0; n < N - m - 1; n++)
            if (pdg[idx[n]].mass > pdg[idx[n + 1]].mass) std::swap(idx[n], idx[n + 1]);
    return idx;
}

// read_FOsurfdat_MUSIC (src/readindata.cpp:626-765) and its boost-invariant
// twin (src/readindata.cpp:395-546): 34 floats per cell in binary, or the text
// columns with optional bulk / rhob / diffusion fields; energies, temperature,
// chemical potentials and pi^{mu nu} times hbarC; cells with T <= 0.01 GeV are
// dropped (src/readindata.cpp:751-761).
std::vector<Surf> read_surface_music(const std::string& file, bool binary, bool binv,
                                     int bulk, int rhob, int diff) {
    std::vector<Surf> out;
    std::ifstream f(file, binary ? std::ios::binary : std::ios::in);
    if (!f.good()) return out;
    const double hc = kHbarC;
    while (!f.eof()) {
        Surf s{};
        if (binary) {
            float a[34];
            for (int i = 0; i < 34; i++) {
                float t = 0.f;
                f.read(reinterpret_cast<char*>(&t), sizeof(float));
                a[i] = t;
            }
            s.tau = a[0]; s.xpt = a[1]; s.ypt = a[2]; s.eta = binv ? 0.f : a[3];
            s.da0 = a[4]; s.da1 = a[5]; s.da2 = a[6]; s.da3 = binv ? 0.f : a[7];
            s.u0 = a[8]; s.u1 = a[9]; s.u2 = a[10]; s.u3 = a[11];
            s.Edec = a[12] * hc; s.Tdec = a[13] * hc;
            s.muB = a[14] * hc; s.muS = a[15] * hc; s.muQ = a[16] * hc;
            s.Pdec = a[17] * s.Tdec - s.Edec;  // float arithmetic as in the reference
            float* pi = &s.pi00;
            for (int k = 0; k < 10; k++) pi[k] = a[18 + k] * hc;
            s.bulkPi = a[28] * hc;
            s.Bn = a[29];
            s.qmu0 = a[30]; s.qmu1 = a[31]; s.qmu2 = a[32]; s.qmu3 = a[33];
        } else {
            double d;
            if (binv) {
                // line based (src/readindata.cpp:468-535)
                std::string line;
                std::getline(f, line, '\n');
                std::stringstream ss(line);
                double tt, xx, yy, ee;
                ss >> tt >> xx >> yy >> ee;
                s.tau = tt; s.xpt = xx; s.ypt = yy; s.eta = 0.f;
                ss >> s.da0 >> s.da1 >> s.da2 >> s.da3;
                s.da3 = 0.f;
                ss >> s.u0 >> s.u1 >> s.u2 >> s.u3;
                ss >> d; s.Edec = d * hc;
                ss >> d; s.Tdec = d * hc;
                ss >> d; s.muB = d * hc;
                ss >> d; s.muS = d * hc;
                ss >> d; s.muQ = d * hc;
                ss >> d; s.Pdec = d * s.Tdec - s.Edec;
                float* pi = &s.pi00;
                for (int k = 0; k < 10; k++) { ss >> d; pi[k] = d * hc; }
                if (bulk == 1) { ss >> d; s.bulkPi = d * hc; } else s.bulkPi = 0.f;
                if (rhob == 1) { ss >> d; s.Bn = d; } else s.Bn = 0.f;
                if (diff == 1) { ss >> s.qmu0 >> s.qmu1 >> s.qmu2 >> s.qmu3; }
                else { s.qmu0 = s.qmu1 = s.qmu2 = s.qmu3 = 0.f; }
            } else {
                f >> s.tau >> s.xpt >> s.ypt >> s.eta;
                f >> s.da0 >> s.da1 >> s.da2 >> s.da3;
                f >> s.u0 >> s.u1 >> s.u2 >> s.u3;
                f >> d; s.Edec = d * hc;
                f >> d; s.Tdec = d * hc;
                f >> d; s.muB = d * hc;
                f >> d; s.muS = d * hc;
                f >> d; s.muQ = d * hc;
                f >> d; s.Pdec = d * s.Tdec - s.Edec;
                float* pi = &s.pi00;
                for (int k = 0; k < 10; k++) { f >> d; pi[k] = d * hc; }
                if (bulk == 1) { f >> d; s.bulkPi = d * hc; } else s.bulkPi = 0.f;
                if (rhob == 1) { f >> d; s.Bn = d; } else s.Bn = 0.f;
                if (diff == 1) { f >> s.qmu0 >> s.qmu1 >> s.qmu2 >> s.qmu3; }
                else { s.qmu0 = s.qmu1 = s.qmu2 = s.qmu3 = 0.f; }
            }
        }
        if (!f.eof() && s.Tdec > 0.01) out.push_back(s);
    }
    return out;
}

// read_FOdata::read_in_HRG_EOS (src/readindata.cpp:913-968)
bool HRGEoS::load(const std::string& file, int iEOS_music) {
    iEOS = iEOS_music;
    rows.clear();
    std::ifstream f(file);
    if (!f.good()) return false;
    std::string line;
    std::getline(f, line);
    std::getline(f, line);
    while (!f.eof()) {
        std::stringstream ss(line);
        std::array<double, 7> it{};
        if (iEOS == 9 || iEOS == 91) {
            double tmp, e, P, T;
            ss >> e >> P >> tmp >> T;
            it[0] = e; it[2] = P; it[3] = T;
        } else if (iEOS == 12) {
            for (int i = 0; i < 5; i++) ss >> it[i];
        } else if (iEOS == 14) {
            for (int i = 0; i < 7; i++) ss >> it[i];
        }
        rows.push_back(it);
        std::getline(f, line);
    }
    return true;
}

// read_FOdata::getValuesFromHRGEOS (src/readindata.cpp:1249-1309)
int HRGEoS::lookup(double ed, double nB, double out[5]) const {
    int nBlen = (iEOS == 12 || iEOS == 14) ? 200 : 1;
    const double de = rows[nBlen][0] - rows[0][0];
    const double e0 = rows[0][0];
    int e_idx = static_cast<int>((ed - e0) / de);
    if (e_idx < 0 || e_idx >= static_cast<int>(rows.size() / nBlen) - 2) return -1;
    const int i1 = e_idx * nBlen, i2 = (e_idx + 1) * nBlen;
    const double e_frac = (ed - rows[i1][0]) / de;
    double f1 = 0, f2 = 0;
    int n1 = 0, n2 = 0;
    if (nBlen > 1) {
        const double d1 = rows[i1 + 1][1], d2 = rows[i2 + 1][1];
        n1 = std::min(nBlen - 2, static_cast<int>(nB / d1));
        n2 = std::min(nBlen - 2, static_cast<int>(nB / d2));
        f1 = std::min(1., (nB - rows[i1 + n1][1]) / d1);
        f2 = std::min(1., (nB - rows[i2 + n2][1]) / d2);
    }
    auto col = [&](int c) {
        const double a = rows[i1 + n1][c] * (1. - f1) + rows[i1 + n1 + 1][c] * f1;
        const double b = rows[i2 + n2][c] * (1. - f2) + rows[i2 + n2 + 1][c] * f2;
        return a * (1 - e_frac) + b * e_frac;
    };
    out[0] = col(2);
    out[1] = col(3);
    out[2] = out[3] = out[4] = 0.;
    if (nBlen > 1)
        for (int c = 4; c < 7; c++) out[c - 2] = col(c);
    return 0;
}

// read_FOdata::regulate_surface_cells + regulate_Wmunu
// (src/readindata.cpp:768-842, 1216-1246)
void regulate_surface(std::vector<Surf>& cells, const HRGEoS* eos) {
    for (auto& s : cells) {
        if (eos) {
            double v[5];
            if (eos->lookup(s.Edec, s.Bn, v) == 0) {
                s.Tdec = v[1]; s.muB = v[2]; s.muS = v[3]; s.muQ = v[4];
                s.Pdec = v[0];
            }
        }
        s.u0 = std::sqrt(1. + s.u1 * s.u1 + s.u2 * s.u2 + s.u3 * s.u3);
        s.qmu0 = (s.u1 * s.qmu1 + s.u2 * s.qmu2 + s.u3 * s.qmu3) / s.u0;
        const double u[4] = {s.u0, s.u1, s.u2, s.u3};
        const double W[4][4] = {{s.pi00, s.pi01, s.pi02, s.pi03},
                                {s.pi01, s.pi11, s.pi12, s.pi13},
                                {s.pi02, s.pi12, s.pi22, s.pi23},
                                {s.pi03, s.pi13, s.pi23, s.pi33}};
        const double g[4] = {-1, 1, 1, 1};
        double udp[4], ul[4];
        for (int i = 0; i < 4; i++) {
            udp[i] = -u[0] * W[0][i] + u[1] * W[1][i] + u[2] * W[2][i] + u[3] * W[3][i];
            ul[i] = g[i] * u[i];
        }
        const double tr = -W[0][0] + W[1][1] + W[2][2] + W[3][3];
        double upu = 0.;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) upu += ul[i] * W[i][j] * ul[j];
        double R[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const double gij = (i == j) ? g[i] : 0.;
                R[i][j] = W[i][j] + u[i] * udp[j] + u[j] * udp[i] + u[i] * u[j] * upu -
                          1. / 3. * (gij + u[i] * u[j]) * (tr + upu);
            }
        s.pi00 = R[0][0]; s.pi01 = R[0][1]; s.pi02 = R[0][2]; s.pi03 = R[0][3];
        s.pi11 = R[1][1]; s.pi12 = R[1][2]; s.pi13 = R[1][3];
        s.pi22 = R[2][2]; s.pi23 = R[2][3]; s.pi33 = R[3][3];
    }
}

// pi^{mu nu} rotated from Milne (tau, eta) to (t, z) components
// (src/iSS.cpp:248-274 and 396-419); float arithmetic as in the reference.
static void pi_tz_of(const Surf& s, float ch, float sh, float P[4][4]) {
    P[0][0] = s.pi00 * ch * ch + 2. * s.pi03 * ch * sh + s.pi33 * sh * sh;
    P[0][1] = s.pi01 * ch + s.pi13 * sh;
    P[0][2] = s.pi02 * ch + s.pi23 * sh;
    P[0][3] = s.pi00 * ch * sh + s.pi03 * (ch * ch + sh * sh) + s.pi33 * sh * ch;
    P[1][0] = P[0][1];
    P[1][1] = s.pi11;
    P[1][2] = s.pi12;
    P[1][3] = s.pi01 * sh + s.pi13 * ch;
    P[2][0] = P[0][2];
    P[2][1] = s.pi12;
    P[2][2] = s.pi22;
    P[2][3] = s.pi02 * sh + s.pi23 * ch;
    P[3][0] = P[0][3];
    P[3][1] = P[1][3];
    P[3][2] = P[2][3];
    P[3][3] = s.pi00 * sh * sh + 2. * s.pi03 * sh * ch + s.pi33 * ch * ch;
}

// iSS::transform_to_local_rest_frame (src/iSS.cpp:170-293): boost dsigma,
// q^mu and pi^{mu nu} into the fluid rest frame; cells with u.dsigma < 0 are
// dropped (src/iSS.cpp:226).
std::vector<SurfLRF> transform_to_lrf(const std::vector<Surf>& cells) {
    std::vector<SurfLRF> out;
    for (const auto& s : cells) {
        SurfLRF r{};
        r.tau = s.tau; r.xpt = s.xpt; r.ypt = s.ypt; r.eta = s.eta;
        r.Edec = s.Edec; r.Tdec = s.Tdec; r.Pdec = s.Pdec; r.Bn = s.Bn;
        r.muB = s.muB; r.muS = s.muS; r.muQ = s.muQ; r.bulkPi = s.bulkPi;
        const float ch = std::cosh(s.eta), sh = std::sinh(s.eta);
        const float ut = s.u0 * ch + s.u3 * sh;
        const float uz = s.u3 * ch + s.u0 * sh;
        const float ux = s.u1, uy = s.u2;
        r.u[0] = ut; r.u[1] = ux; r.u[2] = uy; r.u[3] = uz;
        const double L[4][4] = {
            {ut, -ux, -uy, -uz},
            {-ux, 1. + ux * ux / (ut + 1.), ux * uy / (ut + 1.), ux * uz / (ut + 1.)},
            {-uy, ux * uy / (ut + 1.), 1. + uy * uy / (ut + 1.), uy * uz / (ut + 1.)},
            {-uz, ux * uz / (ut + 1.), uy * uz / (ut + 1.), 1. + uz * uz / (ut + 1.)}};
        const float dup[4] = {s.tau * s.da0 * ch - s.da3 * sh, -s.tau * s.da1,
                              -s.tau * s.da2, -s.da3 * ch + s.tau * s.da0 * sh};
        float dl[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < 4; i++) {
            for (int j = 0; j < 4; j++) dl[i] += L[i][j] * dup[j];
            r.da[i] = (i == 0) ? dl[i] : -dl[i];
        }
        if (r.da[0] < 0) continue;
        const float qtz[4] = {s.qmu0 * ch + s.qmu3 * sh, s.qmu1, s.qmu2,
                              s.qmu3 * ch + s.qmu0 * sh};
        float ql[4] = {0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) ql[i] += L[i][j] * qtz[j];
        r.qx = ql[1]; r.qy = ql[2]; r.qz = ql[3];
        float P[4][4], PL[4][4];
        pi_tz_of(s, ch, sh, P);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                PL[i][j] = 0.f;
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++) PL[i][j] += L[i][a] * P[a][b] * L[b][j];
            }
        r.pixx = PL[1][1]; r.pixy = PL[1][2]; r.pixz = PL[1][3];
        r.piyy = PL[2][2]; r.piyz = PL[2][3];
        out.push_back(r);
    }
    return out;
}

// iSS::computeFOSurfTmunu (src/iSS.cpp:378-445).  Q[0] is the LAST cell's
// n_B, not a sum, as in the reference (src/iSS.cpp:431).
void surface_tmunu(const std::vector<Surf>& cells, double T[16], double Q[3]) {
    float acc[16] = {0};
    float q[3] = {0, 0, 0};
    const float g[4] = {1, -1, -1, -1};
    for (const auto& s : cells) {
        const float ch = std::cosh(s.eta), sh = std::sinh(s.eta);
        float u[4] = {s.u0 * ch + s.u3 * sh, s.u1, s.u2, s.u3 * ch + s.u0 * sh};
        float P[4][4];
        pi_tz_of(s, ch, sh, P);
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const float gij = (i == j) ? g[i] : 0.f;
                const float t = s.Edec * u[i] * u[j] -
                                (s.Pdec + s.bulkPi) * (gij - u[i] * u[j]) + P[i][j];
                acc[4 * i + j] += t;
            }
        q[0] = s.Bn;
        q[2] = 0.4f * s.Bn;
    }
    for (int i = 0; i < 16; i++) T[i] = acc[i];
    for (int i = 0; i < 3; i++) Q[i] = q[i];
}

// ========================================================= special functions

// K_n(x) = int_0^inf exp(-x cosh t) cosh(n t) dt, trapezoid rule (spectrally
// accurate for this analytic, doubly-exponentially decaying integrand).  The
// reference fills its table with GSL's gsl_sf_bessel_K1 / gsl_sf_bessel_Kn
// (src/FSSW.cpp:1626-1633); GSL is absent here, so the table is pinned by a
// scipy.special.kn fixture instead.
double bessel_kn(int n, double x) {
    const double h = std::min(0.1, 0.25 / std::sqrt(x));
    const double tpeak = std::asinh(n / x);
    double sum = 0.5;  // t = 0 term, exp(-x (cosh 0 - 1)) cosh 0 = 1
    for (int k = 1; k < 100000; k++) {
        const double t = k * h;
        const double f = std::exp(-x * (std::cosh(t) - 1.)) * std::cosh(n * t);
        sum += f;
        if (t > tpeak && f < 1e-18 * sum) break;
    }
    return h * sum * std::exp(-x);
}

// FSSW::initialize_special_function_arrays (src/FSSW.cpp:1609-1643); the K1
// and K3 columns are zero unless bulk delta f is on (src/FSSW.cpp:1625-1631).
void SpecialFunctions::init(bool with_bulk, bool with_diffusion) {
    length = static_cast<int>((x_max - x_min) / dx) + 1;
    K1.assign(length, 0.);
    K2.assign(length, 0.);
    K3.assign(length, 0.);
    for (int i = 0; i < length; i++) {
        const double x = x_min + i * dx;
        if (with_bulk) {
            K1[i] = bessel_kn(1, x);
            K3[i] = bessel_kn(3, x);
        }
        K2[i] = bessel_kn(2, x);
    }
    En.clear();
    if (with_diffusion) {
        En.assign(9 * static_cast<size_t>(length), 0.);
        for (int i = 0; i < length; i++) {
            const double x = x_min + i * dx;
            En[i] = expint_en(2, x);
            for (int k = 1; k < 9; k++) En[k * length + i] = expint_en(2 * k + 2, x);
        }
    }
}

void SpecialFunctions::E(double arg, double out[9]) const {
    if (arg < x_min || arg > x_max - dx || En.empty()) {
        out[0] = expint_en(2, arg);
        for (int k = 1; k < 9; k++) out[k] = expint_en(2 * k + 2, arg);
        return;
    }
    const int idx = static_cast<int>((arg - x_min) / dx);
    const double fr = (arg - x_min - idx * dx) / dx;
    for (int k = 0; k < 9; k++) out[k] = (1. - fr) * En[k * length + idx] + fr * En[k * length + idx + 1];
}

// E_n(x) = exp(-x) / x int_0^inf exp(-u) (1 + u/x)^-n du, u = e^v, trapezoid
// in v on [-40, 6] (the integrand decays like e^v and exp(-e^v) at the ends)
double expint_en(int n, double x) {
    const double h = 0.05;
    double sum = 0.;
    for (double v = -40.; v <= 6.; v += h) {
        const double u = std::exp(v);
        sum += std::exp(-u + v - n * std::log1p(u / x));
    }
    return std::exp(-x) / x * sum * h;
}

// get_special_function_K1/K2/K3 (src/FSSW.cpp:1646-1703)
double SpecialFunctions::K(int order, double arg) const {
    if (arg < x_min || arg > x_max - dx) return bessel_kn(order, arg);
    const std::vector<double>& t = (order == 1) ? K1 : (order == 2 ? K2 : K3);
    const int idx = static_cast<int>((arg - x_min) / dx);
    const double fr = (arg - x_min - idx * dx) / dx;
    return (1. - fr) * t[idx] + fr * t[idx + 1];
}

// ============================================================ delta-f tables

// load_bulk_deltaf_14mom_table (src/FSSW.cpp:1215-1301).  The reference never
// skips c0.dat's header line, so every c0 read fails; here c0 is loaded like
// c1 and c2 (documented divergence: the reference's c0 is uninitialised).
bool DeltafTables::load_kind11(const std::string& dir) {
    std::ifstream c[3] = {std::ifstream(dir + "/c0.dat"), std::ifstream(dir + "/c1.dat"),
                          std::ifstream(dir + "/c2.dat")};
    for (auto& s : c)
        if (!s.good()) return false;
    c[0] >> k11_nT >> k11_nmu;
    std::string dummy;
    std::getline(c[0], dummy);
    std::getline(c[0], dummy);
    for (int i = 0; i < 3; i++) {
        std::getline(c[1], dummy);
        std::getline(c[2], dummy);
    }
    k11_c0.assign(k11_nT * k11_nmu, 0.);
    k11_c1.assign(k11_nT * k11_nmu, 0.);
    k11_c2.assign(k11_nT * k11_nmu, 0.);
    for (int j = 0; j < k11_nmu; j++)
        for (int i = 0; i < k11_nT; i++) {
            double t1, t2;
            c[0] >> t1 >> t2 >> k11_c0[i * k11_nmu + j];
            c[1] >> t1 >> t2 >> k11_c1[i * k11_nmu + j];
            c[2] >> t1 >> t2 >> k11_c2[i * k11_nmu + j];
            if (j == 0 && i == 0) { k11_T0 = t1; k11_mu0 = t2; }
            if (j == 0 && i == 1) k11_dT = t1 - k11_T0;
            if (j == 1 && i == 0) k11_dmu = t2 - k11_mu0;
        }
    return true;
}

// load_CE_deltaf_NEOSBQS_table / load_22mom_deltaf_NEOSBQS_table
// (src/FSSW.cpp:1304-1376): 200 x 200 rows after one header line.
// rows (T, mu_B, D) with mu_B outer and T inner
bool DeltafTables::load_qmu(const std::string& file) {
    std::ifstream f(file);
    if (!f.good()) return false;
    qmu.assign(150 * 100, 0.);
    for (int j = 0; j < 100; j++)
        for (int i = 0; i < 150; i++) {
            double a, b, c;
            if (!(f >> a >> b >> c)) return false;
            qmu[i * 100 + j] = c;
        }
    return true;
}

double DeltafTables::qmu_coeff(double T, double muB) const {
    const double T0 = 0.05, mu0 = 0.0, dT = 0.001, dmu = 0.007892;
    const int iT = static_cast<int>((T - T0) / dT);
    const int imu = static_cast<int>((muB - mu0) / dmu);
    const double xf = (T - T0) / dT - iT, yf = (muB - mu0) / dmu - imu;
    if (imu > 100 - 2 || iT > 150 - 2 || imu < 0 || iT < 0) return 1e30;
    const double f1 = qmu[iT * 100 + imu], f2 = qmu[iT * 100 + imu + 1];
    const double f3 = qmu[(iT + 1) * 100 + imu + 1], f4 = qmu[(iT + 1) * 100 + imu];
    return f1 * (1. - xf) * (1. - yf) + f2 * (1. - xf) * yf + f3 * xf * yf + f4 * xf * (1. - yf);
}

bool DeltafTables::load_neos(const std::string& file, int ncol) {
    std::ifstream f(file);
    if (!f.good()) return false;
    std::string dummy;
    std::getline(f, dummy);
    neos_ncol = ncol;
    neos.assign(200 * 200 * ncol, 0.);
    for (int r = 0; r < 200 * 200; r++)
        for (int k = 0; k < ncol; k++) f >> neos[r * ncol + k];
    neos_de = neos[200 * ncol] - neos[0];
    return true;
}

bool DeltafTables::load_kind0(const std::string& file) {
    std::ifstream f(file);
    if (!f.good()) return false;
    double a, b, c, d;
    while (f >> a >> b >> c >> d) {
        k0_T.push_back(a); k0_B0.push_back(b); k0_D0.push_back(c); k0_E0.push_back(d);
    }
    return true;
}

// getbulkvisCoefficients(T) (src/FSSW.cpp:1074-1212).  Only kind 1 reaches
// the FSSW yields and weights: kinds 0, 2, 3, 4 have no yield term
// (src/FSSW.cpp:663-692) and the sampler zeroes Pi for them
// (src/FSSW.cpp:1916-1923), so their coefficients never matter.
void bulk_coefficients_T(int kind, const DeltafTables&, double Tdec, double c[3]) {
    c[0] = c[1] = c[2] = 0.;
    if (kind != 1) return;
    const double x = Tdec / kHbarC;
    // polynomial fit coefficients of the relaxation-time bulk delta f
    static const double a0[11] = {642096.624265727, -8163329.49562861, 47162768.4292073,
                                  -162590040.002683, 369637951.096896, -578181331.809836,
                                  629434830.225675, -470493661.096657, 230936465.421,
                                  -67175218.4629078, 8789472.32652964};
    static const double a1[11] = {1.18171174036192, -17.6740645873717, 136.298469057177,
                                  -635.999435106846, 1918.77100633321, -3836.32258307711,
                                  5136.35746882372, -4566.22991441914, 2593.45375240886,
                                  -853.908199724349, 124.260460450113};
    double pw[11];
    pw[0] = 1.;
    pw[1] = x;
    for (int i = 2; i < 11; i++) pw[i] = pw[i - 1] * x;
    double s0 = a0[0], s1 = a1[0];
    for (int i = 1; i < 11; i++) {
        s0 += a0[i] * pw[i];
        s1 += a1[i] * pw[i];
    }
    c[0] = s0;
    c[1] = s1;
}

// getbulkvisCoefficients(T, muB) (src/FSSW.cpp:1379-1430): bilinear in
// (T, muB), clamped indices, returns (c0-c2)/T^4, c1/T^3, (4c2-c0)/T^4.
void bulk_coefficients_14mom(const DeltafTables& tb, double T, double muB, double c[3]) {
    const int iT = static_cast<int>((T - tb.k11_T0) / tb.k11_dT);
    const int im = static_cast<int>((muB - tb.k11_mu0) / tb.k11_dmu);
    const double xf = (T - tb.k11_T0) / tb.k11_dT - iT;
    const double yf = (muB - tb.k11_mu0) / tb.k11_dmu - im;
    const int T1 = std::max(0, std::min(tb.k11_nT - 1, iT));
    const int m1 = std::max(0, std::min(tb.k11_nmu - 1, im));
    const int T2 = std::max(0, std::min(tb.k11_nT - 1, iT + 1));
    const int m2 = std::max(0, std::min(tb.k11_nmu - 1, im + 1));
    auto bil = [&](const std::vector<double>& t) {
        const int n = tb.k11_nmu;
        return t[T1 * n + m1] * (1. - xf) * (1. - yf) + t[T1 * n + m2] * (1. - xf) * yf +
               t[T2 * n + m2] * xf * yf + t[T2 * n + m1] * xf * (1. - yf);
    };
    const double c0 = bil(tb.k11_c0), c1 = bil(tb.k11_c1), c2 = bil(tb.k11_c2);
    const double T3 = T * T * T, T4 = T3 * T;
    c[0] = (c0 - c2) / T4;
    c[1] = c1 / T3;
    c[2] = (4. * c2 - c0) / T4;
}

// getCENEOSBQSCoefficients / get22momNEOSBQSCoefficients
// (src/FSSW.cpp:1433-1543): bilinear in (e, n_B) on a 200 x 200 table whose
// n_B spacing differs per e row.  CE returns {1/zeta_hat, 1/3 - c2_hat, eta_hat};
// 22-moment returns its six columns.
void neos_coefficients(const DeltafTables& tb, double Edec, double nB, double* c) {
    const int nc = tb.neos_ncol;
    auto at = [&](int row, int col) { return tb.neos[row * nc + col]; };
    int ie = static_cast<int>((Edec - at(0, 0)) / tb.neos_de);
    ie = std::max(0, std::min(200 - 2, ie));
    const int r1 = ie * 200, r2 = (ie + 1) * 200;
    const double ef = (Edec - at(r1, 0)) / tb.neos_de;
    const double d1 = at(r1 + 1, 1), d2 = at(r2 + 1, 1);
    int n1 = static_cast<int>(nB / d1), n2 = static_cast<int>(nB / d2);
    n1 = std::min(200 - 2, n1);
    n2 = std::min(200 - 2, n2);
    const double f1 = std::min(1., (nB - at(r1 + n1, 1)) / d1);
    const double f2 = std::min(1., (nB - at(r2 + n2, 1)) / d2);
    double ic[6];
    for (int i = 2; i < nc; i++) {
        const double a = at(r1 + n1, i) * (1 - f1) + at(r1 + n1 + 1, i) * f1;
        const double b = at(r2 + n2, i) * (1 - f2) + at(r2 + n2 + 1, i) * f2;
        ic[i - 2] = a * (1. - ef) + b * ef;
    }
    if (nc == 5) {
        c[0] = 1. / ic[1];
        c[1] = 1. / 3. - ic[0];
        c[2] = ic[2];
    } else {
        for (int i = 0; i < 6; i++) c[i] = ic[i];
    }
}

// ============================================================ |p| samplers

// CDF_0/1/2 of BosonMomentumSampler (src/BosonMomentumSampler.cpp:36-87) and
// FermionMomentumSampler (src/FermionMomentumSampler.cpp:35-93): integrals of
// E^k f(E) from m0, times e^{m0}.  The fermion CDF_2 lacks the alternating
// sign in the reference; `cdf2_quirk` keeps that.
static double cdf_series(bool fermion, bool quirk, int k, int trunc, double E, double m) {
    double res = 0.;
    if (k == 0 && trunc > 5) {
        if (fermion) return -std::exp(m) * std::log((1. + std::exp(-E)) / (1. + std::exp(-m)));
        return std::exp(m) * std::log((1. - std::exp(-E)) / (1. - std::exp(-m)));
    }
    double sign = 1.;
    for (int n = 0; n < trunc; n++) {
        const int n1 = n + 1;
        const double s = fermion ? ((k == 2 && quirk) ? 1. : sign) : 1.;
        const double em = std::exp(-m * n), ee = std::exp((m - E) * n1);
        if (k == 0) {
            // the fermion CDF_0 series divides int sign by int n1
            // (src/FermionMomentumSampler.cpp:49-53): only n1 = 1 survives
            const double c0 = fermion ? static_cast<double>(static_cast<int>(sign) / n1)
                                      : 1. / n1;
            res += c0 * em * (1. - ee);
        } else if (k == 1) {
            res += s / (n1 * n1) * em * ((m * n1 + 1) - ee * (E * n1 + 1));
        } else {
            res += s / (double(n1) * n1 * n1) * em *
                   ((m * n1 * (m * n1 + 2) + 2) - ee * (E * n1 * (E * n1 + 2) + 2));
        }
        sign = -sign;
    }
    return res;
}

double PSamplerTable::cdf(int k, double Et) const {
    return cdf_series(fermion, cdf2_quirk, k, trunc, Et, m0);
}

// Boson: E~ in [m0 + 0.05, m0 + 50.05] step 0.05; fermion: [m0, m0 + 40] step
// 0.02 (src/BosonMomentumSampler.cpp:7-31, src/FermionMomentumSampler.cpp:8-31).
void PSamplerTable::build(bool is_fermion, double m0_in, int trunc_in, bool quirk) {
    fermion = is_fermion;
    m0 = m0_in;
    trunc = trunc_in;
    cdf2_quirk = quirk;
    const double Emin = fermion ? m0 : m0 + 0.05;
    const double Emax = Emin + (fermion ? 40. : 50.);
    const double dE = fermion ? 0.02 : 0.05;
    const int np = static_cast<int>((Emax - Emin) / dE + 1);
    Et.resize(np); C0.resize(np); C1.resize(np); C2.resize(np);
    for (int i = 0; i < np; i++) {
        const double e = Emin + i * dE;
        Et[i] = e;
        C0[i] = cdf(0, e);
        C1[i] = cdf(1, e);
        C2[i] = cdf(2, e);
    }
    // the constructors call update_cache(1.0) first (src/BosonMomentumSampler.cpp:15)
    cache_key = 1.0;
    c0m = cdf(0, 1.0); c1m = cdf(1, 1.0); c2m = cdf(2, 1.0);
}

// MomentumSamplerShell ctor (src/MomentumSamplerShell.cpp:6-21)
void PSampler::build(bool quirk) {
    tb[0].build(false, 0.05, 10, quirk);
    tb[1].build(false, 30., 2, quirk);
    tb[2].build(false, 50., 1, quirk);
    tb[3].build(true, 0., 10, quirk);
    tb[4].build(true, 30., 2, quirk);
    tb[5].build(true, 50., 1, quirk);
}

// MomentumSamplerBase::Sample_a_momentum + inverse_CDF
// (src/MomentumSamplerBase.cpp:20-93).  Returns -1 where the reference exits
// (m~ - mu~ outside the table, src/MomentumSamplerBase.cpp:33-43).
static double sample_table(PSamplerTable& t, double m, double T, double mu, std::mt19937& g,
                           std::uniform_real_distribution<double>& U) {
    const double eps = 1e-16, tol = 1e-8;
    T = std::max(eps, T);
    const double mt = m / T, mut = mu / T;
    if (std::abs((mt - mut) - t.cache_key) > tol) {
        t.cache_key = mt - mut;
        t.c0m = t.cdf(0, t.cache_key);
        t.c1m = t.cdf(1, t.cache_key);
        t.c2m = t.cdf(2, t.cache_key);
    }
    const double a = mut * mut - mt * mt / 2.;
    const double mterm = t.c2m + 2. * mut * t.c1m + a * t.c0m;
    const int imax = static_cast<int>(t.Et.size()) - 1;
    auto comb = [&](int i) { return t.C2[i] + 2. * mut * t.C1[i] + a * t.C0[i] - mterm; };
    const double cmax = comb(imax);
    const int i0 = static_cast<int>((mt - mut - t.Et[0]) / (t.Et[1] - t.Et[0]));
    if (i0 < 0 || i0 >= imax) return -1.;
    double p = -1., E = -1., ratio = 1.;
    do {
        const double r = U(g) * cmax;
        int lo = i0, hi = imax;
        double rlo = comb(lo), rhi = cmax;
        while (hi - lo > 1) {
            const int mid = static_cast<int>((hi + lo) / 2 + 0.1);
            const double rm = comb(mid);
            if (r < rm) { hi = mid; rhi = rm; } else { lo = mid; rlo = rm; }
        }
        double E0 = t.Et[lo];
        if (E0 < mt - mut) { E0 = mt - mut; rlo = 0.; }
        const double Et = E0 + (t.Et[hi] - E0) / std::max(eps, rhi - rlo) * (r - rlo);
        E = T * Et + mu;
        p = std::sqrt(E * E - m * m);
        ratio = (p / E) / (1. - m * m / (2. * E * E));
    } while (U(g) > ratio);
    return p;
}

double PSampler::sample(double m, double T, double mu, int sign, std::mt19937& g,
                        std::uniform_real_distribution<double>& U) {
    const double m0t = m / T - mu / T;
    const int base = (sign == -1) ? 0 : 3;
    const int which = (m0t < 30.) ? 0 : (m0t < 50. ? 1 : 2);
    return sample_table(tb[base + which], m, T, mu, g, U);
}

// ===================================================================== FSSW

int FSSWOracle::neos_kind() const {
    // src/FSSW.cpp:77-83
    if (cfg.bulk_kind == 21) return 1;
    if (cfg.bulk_kind == 20) return 0;
    return -1;
}

bool FSSWOracle::setup(const std::string& dir, const Config& c, std::string* err) {
    cfg = c;
    std::string pdg, chosen, sub;
    if (cfg.afterburner == 2) {
        pdg = "/pdg-SMASH.dat"; chosen = "/chosen_particles_SMASH.dat"; sub = "/smash";
    } else if (cfg.afterburner == 1) {
        pdg = "/pdg-urqmd_v3.3+.dat"; chosen = "/chosen_particles_urqmd_v3.3+.dat"; sub = "/urqmd";
    } else {
        pdg = "/pdg-s95pv1.dat"; chosen = "/chosen_particles_s95p-v1.dat"; sub = "/urqmd";
    }
    particles = read_resonances(dir + pdg, cfg.quantum_statistics == 1);
    if (particles.empty()) { if (err) *err = "pdg table missing"; return false; }
    order = sampling_order(particles, read_chosen(dir + chosen));
    if (cfg.include_diffusion &&
        !dft.load_qmu(dir + "/deltaf_tables/Coefficients_RTA_diffusion.dat")) {
        if (err) *err = "diffusion table missing";
        return false;
    }
    if (cfg.include_bulk && cfg.bulk_kind == 11) {
        // kind 11 reads /smash_box for SMASH (src/FSSW.cpp:1216-1219)
        const std::string d = dir + "/deltaf_tables" + (cfg.afterburner == 2 ? "/smash_box" : "/urqmd");
        if (!dft.load_kind11(d)) { if (err) *err = "kind 11 tables missing"; return false; }
    }
    if (neos_kind() == 1) {
        if (!dft.load_neos(dir + "/deltaf_tables" + sub + "/NEoSBQS_CE_deltafCoeff.dat", 5)) {
            if (err) *err = "CE table missing";
            return false;
        }
    } else if (neos_kind() == 0) {
        if (!dft.load_neos(dir + "/deltaf_tables" + sub + "/NEoSBQS_22mom_deltafCoeff.dat", 8)) {
            if (err) *err = "22-moment table missing";
            return false;
        }
    }
    sf.init(cfg.include_bulk == 1, cfg.include_diffusion == 1);
    psampler.build(cfg.fermion_cdf2_quirk == 1);
    return true;
}

void FSSWOracle::dN_analytic(const Particle& p, double mu, double T, double res[4]) const {
    double Neq = 0., b1 = 0., b2 = 0., b3 = 0.;
    const int sign = p.sign;
    const double mass = p.mass, beta = 1. / T;
    const double lambda = std::exp(beta * mu);
    const int trunc = (mass < 0.7 && T > 0.05) ? 10 : 1;
    const int kind = cfg.bulk_kind;
    for (int n = 1; n <= trunc; n++) {
        const double arg = n * mass * beta;
        const double theta = std::pow(-sign, n - 1);
        const double fug = std::pow(lambda, n);
        const double K2 = sf.K(2, arg);
        Neq += theta / n * fug * K2;
        if (cfg.include_bulk == 1) {
            const double K1 = sf.K(1, arg);
            if (kind == 1 || kind == 21) {
                b1 += theta * fug * (mass * beta * K1 + 3 * K2 / n);
                b2 += theta * fug * K1;
            } else if (kind == 11 || kind == 20) {
                const double K3 = sf.K(3, arg);
                b1 += theta * fug * K2;
                b2 += theta * fug * (mass * beta * K1 + 3 * K2 / n);
                b3 += theta * fug * (mass * beta * K2 + 3 * K3 / n);
            }
        }
    }
    Neq *= mass * mass * T;
    if (cfg.include_bulk == 1) {
        if (kind == 1 || kind == 21) {
            b1 = mass * mass / beta * b1;
            b2 = mass * mass * mass / 3. * b2;
            b3 = 0.;
        } else if (kind == 11 || kind == 20) {
            b1 = mass * mass / beta * b1;
            b2 = mass * mass / (beta * beta) * b2;
            b3 = mass * mass * mass / (beta * beta) * b3;
        }
    } else {
        b1 = b2 = b3 = 0.;
    }
    res[0] = Neq; res[1] = b1; res[2] = b2; res[3] = b3;
}

void FSSWOracle::yields_one_species(int ri, std::vector<double>& out) const {
    out.assign(cells.size(), 0.);
    const double uf = 1.0 / std::pow(kHbarC, 3);
    const Particle& p = particles[ri];
    const int kind = cfg.bulk_kind;
    for (size_t l = 0; l < cells.size(); l++) {
        const SurfLRF& s = cells[l];
        const double T = s.Tdec, muB = s.muB;
        const double dsu = s.da[0];
        double c[6] = {0, 0, 0, 0, 0, 0};
        double Pi = 0.;
        if (neos_kind() >= 0) neos_coefficients(dft, s.Edec, s.Bn, c);
        if (cfg.include_bulk == 1) {
            if (kind == 21 || kind == 20) {
                Pi = s.bulkPi;
            } else if (kind == 11) {
                Pi = s.bulkPi;
                bulk_coefficients_14mom(dft, T, muB, c);
            } else {
                Pi = (kind == 0) ? s.bulkPi : s.bulkPi / kHbarC;
                bulk_coefficients_T(kind, dft, T, c);
            }
        }
        const double mu = p.baryon * s.muB + p.strange * s.muS + p.charge * s.muQ;
        const double pref = p.gspin / (2. * M_PI * M_PI);
        double r[4];
        dN_analytic(p, mu, T, r);
        // diffusion (src/FSSW.cpp:622-637, 696-702, 783-829)
        double dNq = 0.;
        if (cfg.include_diffusion == 1) {
            // a float dot product of the float records, stored in double (:627-629)
            const double dsq = s.qx * s.da[1] + s.qy * s.da[2] + s.qz * s.da[3];
            const double coeff = dft.qmu_coeff(T, muB);
            // double locals (:633-635)
            const double rho_B = s.Bn, Edec = s.Edec, Pdec = s.Pdec;
            const double pq = rho_B / (Edec + Pdec);
            const int nt = (p.mass < 0.7 && T > 0.05) ? 10 : 1;
            const double beta = 1. / T, lambda = std::exp(beta * mu), mb = p.mass * beta;
            double t1 = 0., t2 = 0.;
            for (int n = 1; n <= nt; n++) {
                const double arg = n * p.mass * beta;
                const double theta = std::pow(-p.sign, n - 1), fug = std::pow(lambda, n);
                t1 += theta / n * fug * sf.K(2, arg);
                double e[9];
                sf.E(arg, e);
                double I1 = std::exp(-arg) / arg * (2. / (arg * arg) + 2. / arg - 1. / 2.) + 3. / 8. * e[0];
                double df2 = 1., fact = 2., p2k = 4.;
                for (int k = 3; k <= 10; k++) {
                    df2 *= (2 * k - 5);
                    fact *= k;
                    p2k *= 2;
                    I1 += 3. * df2 / p2k / fact * e[k - 2];
                }
                t2 += n * theta * fug * (-(mb * mb * mb) * I1);
            }
            t1 = p.mass * p.mass * T * T * t1;
            t2 = T * T * T / 3. * t2;
            dNq = uf * pref * dsq / coeff * (-pq * t1 - p.baryon * t2);
        }
        const double Neq = uf * pref * dsu * r[0];
        double dNb = 0.;
        if (cfg.include_bulk == 1) {
            if (kind == 1 || kind == 21) {
                dNb = uf * pref * dsu * (-Pi * c[0]) * (-c[1] * r[1] + r[2]);
            } else if (kind == 11) {
                dNb = uf * pref * dsu * Pi *
                      (r[1] * p.mass * p.mass * c[0] + r[2] * p.baryon * c[1] + r[3] * c[2]);
            } else if (kind == 20) {
                dNb = uf * pref * dsu * Pi *
                      (r[1] * p.mass * p.mass * c[2] +
                       r[2] * (p.baryon * c[3] + p.strange * c[4] + p.charge * c[5]) +
                       r[3] * (c[1] - c[2]));
            }
        }
        out[l] = std::max(0., Neq + dNb + dNq);
    }
}

// FSSW::sample_momemtum_from_a_fluid_cell (src/FSSW.cpp:1852-1966) with
// get_deltaf_bulk (src/FSSW.cpp:1795-1849) and the boost
// (src/FSSW.cpp:2001-2010).  Draw order per trial: |p| (2 per inner try),
// phi, cos(theta), accept.
int FSSWOracle::sample_from_cell(double mass, int sign, int baryon, int strange, int charge,
                                 const SurfLRF& s, const double* c, std::mt19937& g,
                                 std::uniform_real_distribution<double>& U, double& pT,
                                 double& phi, double& ymeta) {
    const double T = s.Tdec;
    const double mu = std::min(mass, static_cast<double>(baryon * s.muB + strange * s.muS +
                                                         charge * s.muQ));
    const double dfpref = 1.0 / (2.0 * T * T * (s.Edec + s.Pdec));
    const double dsfac = std::abs(s.da[0]) +
                         std::sqrt(s.da[1] * s.da[1] + s.da[2] * s.da[2] + s.da[3] * s.da[3]);
    const int kind = cfg.bulk_kind, nk = neos_kind();
    // diffusion (src/FSSW.cpp:1006-1008, 1865, 1931-1937)
    const double qcoeff = cfg.include_diffusion == 1 ? dft.qmu_coeff(s.Tdec, s.muB) : 1.;
    const double prefactor_qmu = s.Bn / (s.Edec + s.Pdec);
    for (int tries = 1; tries < 5000; tries++) {
        const double pmag = psampler.sample(mass, T, mu, sign, g, U);
        if (pmag < 0.) return -1;
        const double phic = 2 * M_PI * U(g);
        const double cth = 2. * U(g) - 1.;
        const double sth = std::sqrt(1. - cth * cth);
        const double pTc = pmag * sth;
        const double px = pTc * std::cos(phic), py = pTc * std::sin(phic);
        const double p0 = std::sqrt(mass * mass + pmag * pmag);
        const double pz = pmag * cth;
        const double pds = p0 * s.da[0] + px * s.da[1] + py * s.da[2] + pz * s.da[3];
        const double f0 = 1. / (std::exp((p0 - mu) / T) + sign);
        double dfs = 0.;
        if (cfg.include_shear == 1) {
            const double W = px * px * s.pixx + 2. * px * py * s.pixy + 2. * px * pz * s.pixz +
                             py * py * s.piyy + 2. * py * pz * s.piyz +
                             pz * pz * (-s.pixx - s.piyy);
            if (nk == 1) dfs = (1. - sign * f0) * W / (2. * c[2]) / (p0 * T);
            else if (nk == 0) dfs = (1. - sign * f0) * W * c[0];
            else dfs = (1. - sign * f0) * W * dfpref;
        }
        double dfb = 0.;
        if (cfg.include_bulk == 1) {
            double Pi = 0.;
            if (kind == 11 || kind == 21 || kind == 20) Pi = s.bulkPi;
            else if (kind == 1) Pi = s.bulkPi / kHbarC;
            const double Eo = p0 / T, mo = mass / T;
            if (kind == 1 || kind == 21) {
                dfb = -1.0 * (1. - sign * f0) * c[0] * (mo * mo / (3. * Eo) - c[1] * Eo) * Pi;
            } else if (kind == 11) {
                dfb = (1. - sign * f0) * Pi * (c[0] * mass * mass + c[1] * baryon * p0 + c[2] * p0 * p0);
            } else if (kind == 20) {
                dfb = (1. - sign * f0) * Pi *
                      (mass * mass * c[2] + p0 * (baryon * c[3] + strange * c[4] + charge * c[5]) +
                       p0 * p0 * (c[1] - c[2]));
            }
        }
        double dfq = 0.;
        if (cfg.include_diffusion == 1) {
            const double qf = -px * s.qx - py * s.qy - pz * s.qz;
            dfq = (1. - sign * f0) * (prefactor_qmu - baryon / p0) * qf / qcoeff;
        }
        double f1 = pds / p0 / dsfac;
        double f2 = (1. + dfs + dfb + dfq) / 2.;
        f1 = std::max(0., std::min(1., f1));
        f2 = std::max(0., std::min(1., f2));
        if (U(g) < f1 * f2) {
            const float pL[4] = {static_cast<float>(p0), static_cast<float>(px),
                                 static_cast<float>(py), static_cast<float>(pz)};
            const float u[4] = {s.u[0], s.u[1], s.u[2], s.u[3]};
            double pdu = 0.;
            for (int i = 1; i < 4; i++) pdu += pL[i] * u[i];
            float pl[4];
            pl[0] = pL[0] * u[0] + pdu;
            for (int i = 1; i < 4; i++) pl[i] = pL[i] + (pdu / (u[0] + 1) + pL[0]) * u[i];
            pT = std::sqrt(pl[1] * pl[1] + pl[2] * pl[2]);
            phi = std::atan2(pl[2], pl[1]);
            const double mT = std::sqrt(pT * pT + mass * mass);
            const double y = std::asinh(pl[3] / mT);
            ymeta = y - s.eta;
            return 1;
        }
    }
    return 0;
}

void FSSWOracle::sample(uint32_t seed, int nev) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> U(0., 1.);
    std::mt19937 gp(seed);  // the reference's separate GSL stream (src/FSSW.cpp:169-172)
    events.assign(nev, {});
    n_failed = 0;
    const int kind = cfg.bulk_kind, nk = neos_kind();
    std::vector<double> dn;
    std::vector<double> cdf;
    for (int ri : order) {
        const Particle& p = particles[ri];
        yields_one_species(ri, dn);
        // RandomVariable1DArray (src/RandomVariable1DArray.cpp:25-52)
        cdf.assign(dn.size() + 1, 0.);
        double sum = 0.;
        for (size_t l = 0; l < dn.size(); l++) {
            sum += std::max(0., dn[l]);
            cdf[l + 1] = sum;
        }
        const double dN = (cfg.hydro_mode != 2) ? (cfg.y_RB - cfg.y_LB) * sum : sum;
        for (int ev = 0; ev < nev; ev++) {
            // determine_number_to_sample (src/FSSW.cpp:250-309)
            long N = 0;
            const long iN = static_cast<long>(dN);
            const double frac = dN - iN;
            if (cfg.sampling_model == 30) {
                if (dN >= 1e-15) N = std::poisson_distribution<long>(dN)(gp);
            } else if (cfg.sampling_model == 1) {
                N = iN + (U(g) < frac ? 1 : 0);
            } else if (cfg.sampling_model == 10 || cfg.sampling_model == 20) {
                // negative binomial as a gamma-Poisson mixture
                const double k = cfg.sampling_para1 * (cfg.sampling_model == 10 ? frac : dN);
                const double pp = 1.0 / (1.0 + cfg.sampling_para1);
                long extra = 0;
                if (k >= 1e-15) {
                    const double lam = std::gamma_distribution<double>(k, (1. - pp) / pp)(gp);
                    extra = lam > 0 ? std::poisson_distribution<long>(lam)(gp) : 0;
                }
                N = (cfg.sampling_model == 10 ? iN : 0) + extra;
            }
            long k = 1;
            int guard = 0;
            while (k <= N) {
                // RandomVariable1DArray::rand + binarySearch
                // (src/RandomVariable1DArray.cpp:63-67, src/arsenal.cpp:644-678)
                const double r = cdf[0] + (sum - cdf[0] - 1e-15) * U(g);
                long lo = 0, hi = static_cast<long>(cdf.size()) - 1;
                long mid = (lo + hi) / 2;
                while (hi - lo > 1) {
                    if (cdf[mid] < r) lo = mid; else hi = mid;
                    mid = (lo + hi) / 2;
                }
                const SurfLRF& s = cells[lo];
                double c[6] = {0, 0, 0, 0, 0, 0};
                if (nk >= 0) neos_coefficients(dft, s.Edec, s.Bn, c);
                if (cfg.include_bulk == 1 && nk == -1) {
                    if (kind == 11) bulk_coefficients_14mom(dft, s.Tdec, s.muB, c);
                    else bulk_coefficients_T(kind, dft, s.Tdec, c);
                }
                double pT, phi, ym;
                const int st = sample_from_cell(p.mass, p.sign, p.baryon, p.strange, p.charge, s,
                                                c, g, U, pT, phi, ym);
                if (st < 0) return;  // reference exits (src/MomentumSamplerBase.cpp:42)
                if (st == 0) {
                    n_failed++;
                    if (++guard > 100000) return;
                    continue;
                }
                k++;
                double eta_s = s.eta;
                if (cfg.hydro_mode != 2) {
                    const double rap = cfg.y_LB + (cfg.y_RB - cfg.y_LB) * U(g);
                    eta_s = rap - ym;
                }
                // add_one_sampled_particle (src/FSSW.cpp:1969-1996)
                const double y = ym + eta_s;
                const double px = pT * std::cos(phi), py = pT * std::sin(phi);
                const double mT = std::sqrt(p.mass * p.mass + pT * pT);
                Hadron h;
                h.pid = p.monval;
                h.mass = p.mass;
                h.E = mT * std::cosh(y);
                h.px = px;
                h.py = py;
                h.pz = mT * std::sinh(y);
                h.t = s.tau * std::cosh(eta_s);
                h.x = s.xpt;
                h.y = s.ypt;
                h.z = s.tau * std::sinh(eta_s);
                events[ev].push_back(h);
            }
        }
    }
}

void tmunu_from_samples(const std::vector<std::vector<Hadron>>& events,
                        const std::vector<Particle>& particles, double volume, double T[16],
                        double Q[3]) {
    for (int i = 0; i < 16; i++) T[i] = 0.;
    Q[0] = Q[1] = Q[2] = 0.;
    for (const auto& ev : events)
        for (const auto& h : ev) {
            for (const auto& p : particles)
                if (p.monval == h.pid) {
                    Q[0] += p.baryon; Q[1] += p.strange; Q[2] += p.charge;
                    break;
                }
            const double pm[4] = {h.E, h.px, h.py, h.pz};
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) T[4 * i + j] += pm[i] * pm[j] / pm[0];
        }
    const double nv = events.size() * volume;
    for (int i = 0; i < 16; i++) T[i] /= nv;
    for (int i = 0; i < 3; i++) Q[i] /= nv;
}

// EmissionFunctionArray::calculate_dN_pTdpTdphidy (src/emissionfunction.cpp:
// 625-844), evaluated as the reference does on the lab-frame Milne cells:
// p.u = pt u^tau - px u^x - py u^y - pz tau u^eta, p.dsigma tau =
// (pt da0 + px da1 + py da2 + pz da3 / tau) tau, and the shear factor
// p_mu p_nu pi^{mu nu} from the ten lab components (:759-776).  The
// delta-f coefficients are FSSW's (the same choice as yields_one_species and
// sample_from_cell), so that for shear and bulk kind 1 this is the
// reference's smooth path and otherwise its smooth counterpart of the sampler.
// mode 0: pt = mT cosh(-eta_k), pz = mT sinh(-eta_k), weight w_k (:232-238,
// 756-760); mode 1: y - eta_s = y_i - eta_cell, weight 1.
void smooth_spectra(const FSSWOracle& o, const std::vector<Surf>& lab, int ri, const SpectraOpts& so,
                    const std::vector<double>& pT, const std::vector<double>& phi,
                    const std::vector<double>& eta, const std::vector<double>& w,
                    std::vector<double>& out) {
    const Particle& p = o.particles[ri];
    const double mass = p.mass, sign = p.sign, degen = p.gspin;
    const double prefactor = 1.0 / (8.0 * (M_PI * M_PI * M_PI)) / kHbarC / kHbarC / kHbarC;
    const int kind = o.cfg.bulk_kind, nk = o.neos_kind();
    const size_t npT = pT.size(), nphi = phi.size(), ny = eta.size();
    const size_t nplanes = so.mode == 1 ? ny : 1;
    out.assign(nplanes * npT * nphi, 0.);
    for (size_t iy = 0; iy < nplanes; iy++)
        for (size_t i = 0; i < npT; i++) {
            const double mT = std::sqrt(mass * mass + pT[i] * pT[i]);
            for (size_t j = 0; j < nphi; j++) {
                const double px = pT[i] * std::cos(phi[j]), py = pT[i] * std::sin(phi[j]);
                double acc = 0.;
                for (const Surf& s : lab) {
                    const double T = s.Tdec, tau = s.tau;
                    const double mu = p.baryon * s.muB + p.strange * s.muS + p.charge * s.muQ;
                    double c[6] = {0, 0, 0, 0, 0, 0};
                    double Pi = 0.;
                    if (nk >= 0) neos_coefficients(o.dft, s.Edec, s.Bn, c);
                    if (o.cfg.include_bulk == 1) {
                        if (kind == 21 || kind == 20) {
                            Pi = s.bulkPi;
                        } else if (kind == 11) {
                            Pi = s.bulkPi;
                            bulk_coefficients_14mom(o.dft, T, s.muB, c);
                        } else {
                            Pi = (kind == 0) ? s.bulkPi : s.bulkPi / kHbarC;
                            bulk_coefficients_T(kind, o.dft, T, c);
                        }
                    }
                    const double Edec = s.Edec, Pdec = s.Pdec;  // double locals (:687-689)
                    const double dfpref = 1.0 / (2.0 * T * T * (Edec + Pdec));
                    // diffusion (src/emissionfunction.cpp:737-754, 777-786)
                    double qcoeff = 1., prefactor_qmu = 0.;
                    if (o.cfg.include_diffusion == 1) {
                        qcoeff = o.dft.qmu_coeff(T, s.muB);
                        prefactor_qmu = s.Bn / (Edec + Pdec);
                    }
                    const size_t nkk = so.mode == 1 ? 1 : ny;
                    double cell = 0.;
                    for (size_t k = 0; k < nkk; k++) {
                        const double yme = so.mode == 1 ? eta[iy] - s.eta : -eta[k];
                        const double wk = so.mode == 1 ? 1. : w[k];
                        const double pt = mT * std::cosh(yme), pz = mT * std::sinh(yme);
                        const double pdotu = pt * s.u0 - px * s.u1 - py * s.u2 - pz * s.u3;
                        const double f0 = 1. / (std::exp((pdotu - mu) / T) + sign);
                        const double pdsigma = pt * s.da0 + px * s.da1 + py * s.da2 + pz * s.da3 / tau;
                        double dfs = 0.;
                        if (o.cfg.include_shear == 1) {
                            const double W = pt * pt * s.pi00 - 2.0 * pt * px * s.pi01 - 2.0 * pt * py * s.pi02 -
                                             2.0 * pt * pz * s.pi03 + px * px * s.pi11 + 2.0 * px * py * s.pi12 +
                                             2.0 * px * pz * s.pi13 + py * py * s.pi22 + 2.0 * py * pz * s.pi23 +
                                             pz * pz * s.pi33;
                            if (nk == 1) dfs = (1. - sign * f0) * W / (2. * c[2]) / (pdotu * T);
                            else if (nk == 0) dfs = (1. - sign * f0) * W * c[0];
                            else dfs = (1. - sign * f0) * W * dfpref;
                        }
                        double dfb = 0.;
                        if (o.cfg.include_bulk == 1) {
                            const double Eo = pdotu / T, mo = mass / T;
                            if (kind == 1 || kind == 21) {
                                dfb = -1.0 * (1. - sign * f0) * c[0] * (mo * mo / (3. * Eo) - c[1] * Eo) * Pi;
                            } else if (kind == 11) {
                                dfb = (1. - sign * f0) * Pi *
                                      (c[0] * mass * mass + c[1] * p.baryon * pdotu + c[2] * pdotu * pdotu);
                            } else if (kind == 20) {
                                dfb = (1. - sign * f0) * Pi *
                                      (mass * mass * c[2] +
                                       pdotu * (p.baryon * c[3] + p.strange * c[4] + p.charge * c[5]) +
                                       pdotu * pdotu * (c[1] - c[2]));
                            }
                        }
                        double dfq = 0.;
                        if (o.cfg.include_diffusion == 1) {
                            const double qf = pt * s.qmu0 - px * s.qmu1 - py * s.qmu2 - pz * s.qmu3;
                            dfq = (1. - sign * f0) * (prefactor_qmu - p.baryon / pdotu) * qf / qcoeff;
                        }
                        double resize = 1.0;
                        if (so.restrict_deltaf == 1)
                            resize = std::min(1., so.deltaf_max_ratio / (std::abs(dfs + dfb + dfq) + 1e-10));
                        const double result =
                            prefactor * degen * f0 * pdsigma * tau * (1. + (dfs + dfb + dfq) * resize);
                        if (so.use_pos_dN_only && result < 0.) continue;
                        cell += result * wk;
                    }
                    acc += cell;
                }
                out[(iy * npT + i) * nphi + j] = acc;
            }
        }
}

// EmissionFunctionArray::calculate_flows (src/emissionfunction.cpp:875-1023)
void flows(const std::vector<double>& dN, const std::vector<double>& pT, const std::vector<double>& pTw,
           const std::vector<double>& phi, const std::vector<double>& phiw, double mass, int to_order,
           std::vector<double>& diff, std::vector<double>& inte) {
    const size_t npT = pT.size(), nphi = phi.size();
    const size_t nc = 3 + 3 * to_order;
    diff.assign(npT * nc, 0.);
    inte.assign(6 * (to_order + 1), 0.);
    std::vector<double> norm(npT, 0.), re(npT * to_order, 0.), im(npT * to_order, 0.);
    for (size_t i = 0; i < npT; i++) {
        for (size_t j = 0; j < nphi; j++) {
            const double dn = dN[i * nphi + j];
            norm[i] += dn * phiw[j];
            for (int n = 1; n <= to_order; n++) {
                re[i * to_order + n - 1] += dn * phiw[j] * std::cos(n * phi[j]);
                im[i * to_order + n - 1] += dn * phiw[j] * std::sin(n * phi[j]);
            }
        }
        norm[i] = norm[i] + 1e-30;
        diff[i * nc + 0] = pT[i];
        diff[i * nc + 1] = std::sqrt(mass * mass + pT[i] * pT[i]) - mass;
        diff[i * nc + 2] = norm[i] / (2.0 * M_PI);
        for (int t = 0; t < to_order; t++) {
            const double a = re[i * to_order + t], b = im[i * to_order + t];
            diff[i * nc + 3 + 3 * t] = a / norm[i];
            diff[i * nc + 4 + 3 * t] = b / norm[i];
            diff[i * nc + 5 + 3 * t] = std::sqrt(a * a + b * b) / norm[i];
        }
    }
    double ni = 0.;
    std::vector<double> vr(to_order, 0.), vi(to_order, 0.);
    for (size_t i = 0; i < npT; i++) {
        ni += norm[i] * pT[i] * pTw[i];
        for (int t = 0; t < to_order; t++) {
            vr[t] += re[i * to_order + t] * pT[i] * pTw[i];
            vi[t] += im[i * to_order + t] * pT[i] * pTw[i];
        }
    }
    inte[0] = 0; inte[1] = ni; inte[2] = 0; inte[3] = 1; inte[4] = 0; inte[5] = 1;
    for (int t = 0; t < to_order; t++) {
        double* r = &inte[6 * (t + 1)];
        r[0] = 1 + t;
        r[1] = vr[t];
        r[2] = vi[t];
        r[3] = vr[t] / ni;
        r[4] = vi[t] / ni;
        r[5] = std::sqrt(vr[t] * vr[t] + vi[t] * vi[t]) / ni;
    }
}

// ---------------------------------------------------------- resonance decays

// One entry per particle with all its channels; an anti-baryon entry follows
// every baryon with B = 1, its daughters conjugated unless neutral in B, Q
// and S (src/particle_decay.cpp:33-170).  A particle is stable when its first
// channel has one daughter.
bool Decayer::load(const std::string& file) {
    table.clear();
    std::ifstream f(file);
    if (!f.good()) return false;
    while (true) {
        DecayParticle p;
        std::string name;
        int gspin, charm, bottom, giso, ndec;
        if (!(f >> p.monval)) break;
        f >> name >> p.mass >> p.width >> gspin >> p.baryon >> p.strange >> charm >> bottom >> giso >>
            p.charge >> ndec;
        for (int j = 0; j < ndec; j++) {
            DecayChannel c;
            int dummy;
            f >> dummy >> c.npart >> c.br >> c.part[0] >> c.part[1] >> c.part[2] >> c.part[3] >> c.part[4];
            p.ch.push_back(c);
        }
        if (!f) break;
        p.stable = (!p.ch.empty() && p.ch[0].npart == 1) ? 1 : 0;
        table.push_back(p);
        if (p.baryon == 1) {
            DecayParticle a = p;
            a.monval = -p.monval;
            a.baryon = -p.baryon;
            a.strange = -p.strange;
            a.charge = -p.charge;
            for (auto& c : a.ch)
                for (int k = 0; k < 5; k++) {
                    const int d = c.part[k];
                    if (d == 0) continue;
                    const DecayParticle* dp = nullptr;
                    for (const auto& q : table)
                        if (q.monval == d) { dp = &q; break; }
                    // a daughter missing from the list is conjugated
                    const bool neutral = dp && dp->baryon == 0 && dp->charge == 0 && dp->strange == 0;
                    c.part[k] = neutral ? d : -d;
                }
            table.push_back(a);
        }
    }
    return !table.empty();
}

const DecayParticle* Decayer::find(int monval) const {
    for (const auto& p : table)
        if (p.monval == monval) return &p;
    return nullptr;
}

namespace {
// boost a rest-frame momentum by the mother's velocity (src/particle_decay.cpp:376-394)
void boost_to_lab(double vx, double vy, double vz, double E, double px, double py, double pz, Hadron& h) {
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double gamma = 1. / std::sqrt(1. - v2);
    const double gm1 = gamma - 1.;
    const double vp = vx * px + vy * py + vz * pz;
    const double k = v2 > 0. ? gm1 * vp / v2 + gamma * E : 0.;  // at rest: no boost (the reference: 0/0)
    h.E = static_cast<float>(gamma * (E + vp));
    h.px = static_cast<float>(px + k * vx);
    h.py = static_cast<float>(py + k * vy);
    h.pz = static_cast<float>(pz + k * vz);
}
}  // namespace

// Channel by cumulative branching ratio, then the two-body
// (src/particle_decay.cpp:345-418) or three-body (:422-533) decay at the pole
// mass (the Breit-Wigner draw is off there), isotropic in the rest frame, an
// exponential lifetime of mean E / (M Gamma) (hbar c = 0.19733 GeV fm), the
// decay point on the mother's straight line.  Channels of any other size
// leave no daughters, as in the reference.
void Decayer::decay(const Hadron& m, std::vector<Hadron>& out, std::mt19937& g,
                    std::uniform_real_distribution<double>& U) const {
    const DecayParticle* M = find(m.pid);
    if (!M || M->stable) {
        out.push_back(m);
        return;
    }
    const double r = U(g);
    double cum = 0.;
    const DecayChannel* c = nullptr;
    for (const auto& ch : M->ch) {
        cum += ch.br;
        if (cum > r) { c = &ch; break; }
    }
    if (!c) c = &M->ch.back();   // branching ratios summing below r (the reference dereferences null)
    const double Mp = m.mass, W = M->width;
    const double vx = m.px / m.E, vy = m.py / m.E, vz = m.pz / m.E;
    if (c->npart == 2) {
        Hadron d1{}, d2{};
        d1.pid = c->part[0];
        d2.pid = c->part[1];
        d1.mass = static_cast<float>(find(d1.pid) ? find(d1.pid)->mass : 0.);
        d2.mass = static_cast<float>(find(d2.pid) ? find(d2.pid)->mass : 0.);
        const double m1 = d1.mass, m2 = d2.mass;
        if (Mp < m1 + m2) {
            n_below++;
            return;
        }
        const double t = Mp * Mp - m1 * m1 - m2 * m2;
        const double p = std::sqrt(t * t - 4 * m1 * m1 * m2 * m2) / (2 * Mp);
        const double phi = U(g) * 2 * M_PI;
        const double cth = 2. * (U(g) - 0.5);
        const double sth = std::sqrt(1. - cth * cth);
        const double px = p * sth * std::cos(phi), py = p * sth * std::sin(phi), pz = p * cth;
        boost_to_lab(vx, vy, vz, std::sqrt(p * p + m1 * m1), px, py, pz, d1);
        boost_to_lab(vx, vy, vz, std::sqrt(p * p + m2 * m2), -px, -py, -pz, d2);
        double life = 1e10;
        if (W > 1e-10) life = -(m.E / Mp / W) * std::log(U(g)) * 0.19733;
        for (Hadron* d : {&d1, &d2}) {
            d->t = static_cast<float>(m.t + life);
            d->x = static_cast<float>(m.x + m.px / m.E * life);
            d->y = static_cast<float>(m.y + m.py / m.E * life);
            d->z = static_cast<float>(m.z + m.pz / m.E * life);
        }
        out.push_back(d1);
        out.push_back(d2);
    } else if (c->npart == 3) {
        Hadron d[3]{};
        double ms[3];
        for (int k = 0; k < 3; k++) {
            d[k].pid = c->part[k];
            d[k].mass = static_cast<float>(find(d[k].pid) ? find(d[k].pid)->mass : 0.);
            ms[k] = d[k].mass;
        }
        if (Mp < ms[0] + ms[1] + ms[2]) {
            n_below++;
            return;
        }
        double E1, E2, E3, p1, p2, c12;
        do {
            do {
                E1 = U(g) * (Mp - ms[0] - ms[1] - ms[2]) + ms[0];
                E2 = U(g) * (Mp - ms[0] - ms[1] - ms[2]) + ms[1];
            } while (E1 + E2 > Mp);
            p1 = std::sqrt(E1 * E1 - ms[0] * ms[0]);
            p2 = std::sqrt(E2 * E2 - ms[1] * ms[1]);
            E3 = Mp - E1 - E2;
            c12 = (E3 * E3 - p1 * p1 - p2 * p2 - ms[2] * ms[2]) / (2. * p1 * p2);
        } while (c12 < -1.0 || c12 > 1.0);
        double life = 1e10;
        if (W > 1e-10) life = -(m.E / Mp / W) * std::log(U(g)) * 0.19733;
        // daughters 2 and 3 in the plane of daughter 1 (along z), then an
        // Euler rotation (phi, ksi, theta)
        const double t2x = p2 * std::sqrt(1. - c12 * c12), t2z = p2 * c12;
        const double t3x = -t2x, t3z = -(p1 + t2z);
        const double phi = 2. * M_PI * U(g);
        const double ksi = 2. * M_PI * U(g);
        const double cth = 2. * U(g) - 1.0;
        const double sp = std::sin(phi), cp = std::cos(phi), sk = std::sin(ksi), ck = std::cos(ksi);
        const double sth = std::sqrt(1. - cth * cth);
        const double q[3][3] = {
            {-p1 * sth * ck, p1 * sth * sk, p1 * cth},
            {t2x * (cp * cth * ck - sp * sk) - t2z * sth * ck, t2x * (-cp * cth * sk - sp * ck) + t2z * sth * sk,
             t2x * (cp * sth) + t2z * cth},
            {t3x * (cp * cth * ck - sp * sk) - t3z * sth * ck, t3x * (-cp * cth * sk - sp * ck) + t3z * (sth * sk),
             t3x * cp * sth + t3z * cth}};
        for (int k = 0; k < 3; k++) {
            const double Ek = std::sqrt(ms[k] * ms[k] + q[k][0] * q[k][0] + q[k][1] * q[k][1] + q[k][2] * q[k][2]);
            boost_to_lab(vx, vy, vz, Ek, q[k][0], q[k][1], q[k][2], d[k]);
            d[k].t = static_cast<float>(m.t + life);
            d[k].x = static_cast<float>(m.x + m.px / m.E * life);
            d[k].y = static_cast<float>(m.y + m.py / m.E * life);
            d[k].z = static_cast<float>(m.z + m.pz / m.E * life);
        }
        for (int k = 0; k < 3; k++) out.push_back(d[k]);
    }
}

// Decays every particle of the event, appending unstable daughters to the
// work list and stable ones to the event, in the reference's order.
void Decayer::feed_down(std::vector<Hadron>& event, std::mt19937& g,
                        std::uniform_real_distribution<double>& U) const {
    std::vector<Hadron> work(event);
    event.clear();
    std::vector<Hadron> ds;
    for (size_t i = 0; i < work.size(); i++) {
        ds.clear();
        const Hadron m = work[i];
        decay(m, ds, g, U);
        for (const auto& d : ds) {
            const DecayParticle* q = find(d.pid);
            // particles missing from the list stay final (the reference would
            // dereference a null entry for them)
            if (!q || q->stable == 1) event.push_back(d);
            else work.push_back(d);
        }
    }
}

}  // namespace iss_oracle
