#include "surface_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

namespace iss_io {

using iSS_data::hbarC;

MusicInput read_music_input(const std::string& work_dir, int bulk, int rhob, int diff) {
    MusicInput m;
    m.bulk = bulk;
    m.rhob = rhob;
    m.diff = diff;
    std::ifstream f(work_dir + "/music_input");
    if (!f.is_open()) return m;
    m.found = true;
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string key;
        ss >> key;
        if (key == "Include_Bulk_Visc_Yes_1_No_0") ss >> m.bulk;
        else if (key == "Include_Rhob_Yes_1_No_0") ss >> m.rhob;
        else if (key == "turn_on_baryon_diffusion") ss >> m.diff;
        else if (key == "freeze_surface_in_binary") {
            int b = 1;
            ss >> b;
            m.binary = (b == 1);
        } else if (key == "EOS_to_use") ss >> m.iEOS;
    }
    return m;
}

namespace {
// one record of 34 floats in FO_surf order -> FO_surf with hbarC factors
void from_array(const float* a, bool binv, FO_surf& s) {
    s.tau = a[0]; s.xpt = a[1]; s.ypt = a[2]; s.eta = binv ? 0.f : a[3];
    s.da0 = a[4]; s.da1 = a[5]; s.da2 = a[6]; s.da3 = binv ? 0.f : a[7];
    s.u0 = a[8]; s.u1 = a[9]; s.u2 = a[10]; s.u3 = a[11];
    s.Edec = a[12] * hbarC; s.Tdec = a[13] * hbarC;
    s.muB = a[14] * hbarC; s.muS = a[15] * hbarC; s.muQ = a[16] * hbarC;
    s.Pdec = a[17] * s.Tdec - s.Edec;
    s.pi00 = a[18] * hbarC; s.pi01 = a[19] * hbarC; s.pi02 = a[20] * hbarC;
    s.pi03 = a[21] * hbarC; s.pi11 = a[22] * hbarC; s.pi12 = a[23] * hbarC;
    s.pi13 = a[24] * hbarC; s.pi22 = a[25] * hbarC; s.pi23 = a[26] * hbarC;
    s.pi33 = a[27] * hbarC;
    s.bulkPi = a[28] * hbarC;
    s.Bn = a[29];
    s.qmu0 = a[30]; s.qmu1 = a[31]; s.qmu2 = a[32]; s.qmu3 = a[33];
}
}  // namespace

bool read_surface(const std::string& file, int hydro_mode, const MusicInput& in,
                  std::vector<FO_surf>& out, std::string* err) {
    out.clear();
    const bool binv = (hydro_mode == 1);
    std::ifstream f(file, in.binary ? std::ios::binary : std::ios::in);
    if (!f.good()) {
        if (err) *err = "Surface file is not found! " + file;
        return false;
    }
    if (in.binary) {
        float a[34];
        while (f.read(reinterpret_cast<char*>(a), sizeof(a))) {
            FO_surf s{};
            from_array(a, binv, s);
            if (s.Tdec > 0.01) out.push_back(s);
        }
        return true;
    }
    // text: whitespace separated fields, optional bulk / rhob / diffusion
    const int nf = 28 + (in.bulk == 1) + (in.rhob == 1) + 4 * (in.diff == 1);
    std::vector<double> v(nf);
    std::string line;
    std::stringstream all;
    all << f.rdbuf();
    while (true) {
        int k = 0;
        for (; k < nf; k++)
            if (!(all >> v[k])) break;
        if (k < nf) break;
        float a[34] = {0};
        for (int i = 0; i < 28; i++) a[i] = static_cast<float>(v[i]);
        // text column 17 is (e+P)/T: P = (e+P)/T * T - e
        int j = 28;
        if (in.bulk == 1) a[28] = static_cast<float>(v[j++]);
        if (in.rhob == 1) a[29] = static_cast<float>(v[j++]);
        if (in.diff == 1)
            for (int q = 0; q < 4; q++) a[30 + q] = static_cast<float>(v[j++]);
        FO_surf s{};
        s.tau = a[0]; s.xpt = a[1]; s.ypt = a[2]; s.eta = binv ? 0.f : a[3];
        s.da0 = a[4]; s.da1 = a[5]; s.da2 = a[6]; s.da3 = binv ? 0.f : a[7];
        s.u0 = a[8]; s.u1 = a[9]; s.u2 = a[10]; s.u3 = a[11];
        s.Edec = v[12] * hbarC; s.Tdec = v[13] * hbarC;
        s.muB = v[14] * hbarC; s.muS = v[15] * hbarC; s.muQ = v[16] * hbarC;
        s.Pdec = v[17] * s.Tdec - s.Edec;
        float* pi = &s.pi00;
        for (int q = 0; q < 10; q++) pi[q] = v[18 + q] * hbarC;
        s.bulkPi = (in.bulk == 1) ? a[28] * hbarC : 0.f;
        s.Bn = (in.rhob == 1) ? a[29] : 0.f;
        s.qmu0 = a[30]; s.qmu1 = a[31]; s.qmu2 = a[32]; s.qmu3 = a[33];
        if (s.Tdec > 0.01) out.push_back(s);
    }
    return true;
}

std::vector<particle_info> read_pdg(const std::string& file, bool qs) {
    std::vector<particle_info> out;
    std::ifstream f(file);
    if (!f.good()) return out;
    while (true) {
        particle_info p;
        if (!(f >> p.monval)) break;
        f >> p.name >> p.mass >> p.width >> p.gspin >> p.baryon >> p.strange >> p.charm >>
            p.bottom >> p.gisospin >> p.charge >> p.decays;
        for (int j = 0; j < p.decays; j++) {
            auto* d = new decay_channel_info;
            int dummy;
            f >> dummy >> d->decay_Npart >> d->branching_ratio >> d->decay_part[0] >>
                d->decay_part[1] >> d->decay_part[2] >> d->decay_part[3] >> d->decay_part[4];
            std::string rest;
            std::getline(f, rest);
            p.decay_channels.push_back(d);
        }
        if (!f) break;
        p.stable = (!p.decay_channels.empty() && p.decay_channels[0]->decay_Npart == 1) ? 1 : 0;
        out.push_back(p);
        if (p.baryon > 0) {
            // anti-baryon; decay products flip unless neutral and non-strange
            particle_info a = p;
            a.monval = -p.monval;
            a.name = "Anti-" + p.name;
            a.baryon = -p.baryon; a.strange = -p.strange; a.charm = -p.charm;
            a.bottom = -p.bottom; a.charge = -p.charge;
            a.decay_channels.clear();
            for (auto* d : p.decay_channels) {
                auto* e = new decay_channel_info(*d);
                for (int k = 0; k < 5; k++) {
                    const int mv = d->decay_part[k];
                    if (mv == 0) continue;
                    const particle_info* q = nullptr;
                    for (const auto& r : out)
                        if (r.monval == mv) { q = &r; break; }
                    const bool neutral = q && q->baryon == 0 && q->charge == 0 && q->strange == 0;
                    e->decay_part[k] = neutral ? mv : -mv;
                }
                a.decay_channels.push_back(e);
            }
            out.push_back(a);
        }
    }
    for (auto& p : out) p.sign = qs ? (p.baryon == 0 ? -1 : 1) : 0;
    return out;
}

std::vector<int> read_chosen(const std::string& file) {
    std::vector<int> out;
    std::ifstream f(file);
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        double v;
        if (ss >> v) out.push_back(static_cast<int>(v));
    }
    return out;
}

std::vector<int> sampling_order(const std::vector<particle_info>& pdg, const std::vector<int>& chosen,
                                std::vector<int>* missing) {
    std::vector<int> idx;
    for (int mv : chosen) {
        bool hit = false;
        for (size_t n = 0; n < pdg.size(); n++)
            if (pdg[n].monval == mv) {
                idx.push_back(static_cast<int>(n));
                hit = true;
                break;
            }
        if (!hit && missing) missing->push_back(mv);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return pdg[a].mass < pdg[b].mass; });
    return idx;
}

bool read_hrg_eos(const std::string& file, int iEOS, std::vector<double>& rows) {
    rows.clear();
    std::ifstream f(file);
    if (!f.good()) return false;
    std::string line;
    std::getline(f, line);
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        double it[7] = {0, 0, 0, 0, 0, 0, 0};
        if (iEOS == 9 || iEOS == 91) {
            double e, P, s, T;
            if (!(ss >> e >> P >> s >> T)) continue;
            it[0] = e; it[2] = P; it[3] = T;
        } else {
            const int n = (iEOS == 12) ? 5 : 7;
            int k = 0;
            for (; k < n; k++)
                if (!(ss >> it[k])) break;
            if (k < n) continue;
        }
        rows.insert(rows.end(), it, it + 7);
    }
    return !rows.empty();
}

int hrg_lookup(const std::vector<double>& rows, int iEOS, double ed, double nB, double out[5]) {
    const int nBlen = (iEOS == 12 || iEOS == 14) ? 200 : 1;
    const long n = static_cast<long>(rows.size() / 7);
    auto R = [&](long r, int c) { return rows[r * 7 + c]; };
    const double de = R(nBlen, 0) - R(0, 0);
    const int ie = static_cast<int>((ed - R(0, 0)) / de);
    if (ie < 0 || ie >= static_cast<int>(n / nBlen) - 2) return -1;
    const long i1 = static_cast<long>(ie) * nBlen, i2 = static_cast<long>(ie + 1) * nBlen;
    const double ef = (ed - R(i1, 0)) / de;
    double f1 = 0, f2 = 0;
    int n1 = 0, n2 = 0;
    if (nBlen > 1) {
        const double d1 = R(i1 + 1, 1), d2 = R(i2 + 1, 1);
        n1 = std::min(nBlen - 2, static_cast<int>(nB / d1));
        n2 = std::min(nBlen - 2, static_cast<int>(nB / d2));
        f1 = std::min(1., (nB - R(i1 + n1, 1)) / d1);
        f2 = std::min(1., (nB - R(i2 + n2, 1)) / d2);
    }
    auto col = [&](int c) {
        const double a = R(i1 + n1, c) * (1. - f1) + R(i1 + n1 + 1, c) * f1;
        const double b = R(i2 + n2, c) * (1. - f2) + R(i2 + n2 + 1, c) * f2;
        return a * (1. - ef) + b * ef;
    };
    out[0] = col(2);
    out[1] = col(3);
    out[2] = out[3] = out[4] = 0.;
    if (nBlen > 1)
        for (int c = 4; c < 7; c++) out[c - 2] = col(c);
    return 0;
}

void regulate(std::vector<FO_surf>& cells, const std::vector<double>* eos, int iEOS) {
    for (auto& s : cells) {
        if (eos) {
            double v[5];
            if (hrg_lookup(*eos, iEOS, s.Edec, s.Bn, v) == 0) {
                s.Tdec = v[1]; s.muB = v[2]; s.muS = v[3]; s.muQ = v[4]; s.Pdec = v[0];
            }
        }
        s.u0 = std::sqrt(1. + s.u1 * s.u1 + s.u2 * s.u2 + s.u3 * s.u3);
        s.qmu0 = (s.u1 * s.qmu1 + s.u2 * s.qmu2 + s.u3 * s.qmu3) / s.u0;
        const double u[4] = {s.u0, s.u1, s.u2, s.u3};
        const double W[4][4] = {{s.pi00, s.pi01, s.pi02, s.pi03}, {s.pi01, s.pi11, s.pi12, s.pi13},
                                {s.pi02, s.pi12, s.pi22, s.pi23}, {s.pi03, s.pi13, s.pi23, s.pi33}};
        const double g[4] = {-1, 1, 1, 1};
        double udp[4];
        for (int i = 0; i < 4; i++)
            udp[i] = -u[0] * W[0][i] + u[1] * W[1][i] + u[2] * W[2][i] + u[3] * W[3][i];
        const double tr = -W[0][0] + W[1][1] + W[2][2] + W[3][3];
        double upu = 0.;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) upu += g[i] * u[i] * W[i][j] * g[j] * u[j];
        double Rg[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                Rg[i][j] = W[i][j] + u[i] * udp[j] + u[j] * udp[i] + u[i] * u[j] * upu -
                           1. / 3. * ((i == j ? g[i] : 0.) + u[i] * u[j]) * (tr + upu);
        s.pi00 = Rg[0][0]; s.pi01 = Rg[0][1]; s.pi02 = Rg[0][2]; s.pi03 = Rg[0][3];
        s.pi11 = Rg[1][1]; s.pi12 = Rg[1][2]; s.pi13 = Rg[1][3];
        s.pi22 = Rg[2][2]; s.pi23 = Rg[2][3]; s.pi33 = Rg[3][3];
    }
}

bool read_kind11(const std::string& dir, Kind11& t) {
    std::ifstream f[3] = {std::ifstream(dir + "/c0.dat"), std::ifstream(dir + "/c1.dat"),
                          std::ifstream(dir + "/c2.dat")};
    for (auto& s : f)
        if (!s.good()) return false;
    std::vector<double>* dst[3] = {&t.c0, &t.c1, &t.c2};
    for (int k = 0; k < 3; k++) {
        int nT = 0, nmu = 0;
        f[k] >> nT >> nmu;
        std::string line;
        std::getline(f[k], line);
        std::getline(f[k], line);  // column header
        t.nT = nT;
        t.nmu = nmu;
        dst[k]->assign(static_cast<size_t>(nT) * nmu, 0.);
        for (int j = 0; j < nmu; j++)
            for (int i = 0; i < nT; i++) {
                double T, mu, v;
                f[k] >> T >> mu >> v;
                (*dst[k])[static_cast<size_t>(i) * nmu + j] = v;
                if (k == 1) {
                    if (i == 0 && j == 0) { t.T0 = T; t.mu0 = mu; }
                    if (i == 1 && j == 0) t.dT = T - t.T0;
                    if (i == 0 && j == 1) t.dmu = mu - t.mu0;
                }
            }
    }
    return t.nT > 1 && t.nmu > 1;
}

bool read_neos(const std::string& file, int ncol, std::vector<double>& rows) {
    std::ifstream f(file);
    if (!f.good()) return false;
    std::string line;
    std::getline(f, line);
    rows.assign(200ull * 200ull * ncol, 0.);
    for (auto& v : rows)
        if (!(f >> v)) return false;
    return true;
}

bool read_qmu_table(const std::string& file, std::vector<double>& D) {
    std::ifstream f(file);
    if (!f.good()) return false;
    D.assign(150 * 100, 0.);
    for (int imu = 0; imu < 100; imu++)
        for (int iT = 0; iT < 150; iT++) {
            double T, mu;
            if (!(f >> T >> mu >> D[iT * 100 + imu])) return false;
        }
    return true;
}

bool write_binary(const std::string& file, const std::vector<std::vector<iSS_Hadron>*>& ev) {
    FILE* o = std::fopen(file.c_str(), "wb");
    if (!o) return false;
    for (const auto* e : ev) {
        const int n = static_cast<int>(e->size());
        std::fwrite(&n, sizeof(int), 1, o);
        for (const auto& h : *e) {
            std::fwrite(&h.pid, sizeof(int), 1, o);
            const float a[9] = {h.mass, h.t, h.x, h.y, h.z, h.E, h.px, h.py, h.pz};
            std::fwrite(a, sizeof(float), 9, o);
        }
    }
    std::fclose(o);
    return true;
}

bool write_oscar(const std::string& file, const std::string& header_file,
                 const std::vector<std::vector<iSS_Hadron>*>& ev) {
    std::ifstream hdr(header_file);
    if (!hdr.is_open()) return false;
    std::ofstream o(file);
    std::string line;
    while (std::getline(hdr, line)) o << line << "\n";
    char buf[512];
    for (size_t iev = 0; iev < ev.size(); iev++) {
        const auto& e = *ev[iev];
        if (e.empty()) continue;
        std::snprintf(buf, sizeof(buf), "%10zu  %10zu  %8g  %8g\n", iev, e.size(), 0.0, 0.0);
        o << buf;
        for (size_t i = 0; i < e.size(); i++) {
            const auto& h = e[i];
            std::snprintf(buf, sizeof(buf),
                          "%10zu  %10d  %24.16e  %24.16e  %24.16e  %24.16e  %24.16e  %24.16e  %24.16e  "
                          "%24.16e  %24.16e\n",
                          i + 1, h.pid, (double)h.px, (double)h.py, (double)h.pz, (double)h.E,
                          (double)h.mass, (double)h.x, (double)h.y, (double)h.z, (double)h.t);
            o << buf;
        }
    }
    return true;
}

bool write_gzip(const std::string& file, const std::vector<std::vector<iSS_Hadron>*>& ev) {
    gzFile g = gzopen(file.c_str(), "wb");
    if (!g) return false;
    for (const auto* e : ev) {
        gzprintf(g, "%d \n", static_cast<int>(e->size()));
        for (const auto& h : *e) {
            gzprintf(g, "%d ", h.pid);
            gzprintf(g, "%.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e\n", (double)h.mass, (double)h.t,
                     (double)h.x, (double)h.y, (double)h.z, (double)h.E, (double)h.px, (double)h.py,
                     (double)h.pz);
        }
    }
    gzclose(g);
    return true;
}

}  // namespace iss_io
