// Test probe (built by tests/test_surface_formats.py with g++, no GPU): the
// drop-in's MUSIC surface reader and host regulation
// (iss_amd/csrc/host/surface_io.cpp) behind a C entry point, so that the
// text, 34-float binary, 3+1D and boost-invariant paths can be compared with
// each other, with the oracle and with the reference's own reader on the CPU.
#include <cstring>
#include <string>
#include <vector>

#include "surface_io.h"

extern "C" long probe_read(const char* work_dir, const char* surface, int hydro_mode, int bulk, int rhob, int diff,
                           const char* eos_file, int iEOS, float* out34, long cap) {
    const iss_io::MusicInput mi = iss_io::read_music_input(work_dir, bulk, rhob, diff);
    if (!mi.found) return -1;
    std::vector<FO_surf> cells;
    std::string err;
    if (!iss_io::read_surface(std::string(work_dir) + "/" + surface, hydro_mode, mi, cells, &err)) return -2;
    std::vector<double> eos;
    if (eos_file && eos_file[0]) {
        if (!iss_io::read_hrg_eos(eos_file, iEOS, eos)) return -3;
        iss_io::regulate(cells, &eos, iEOS);
    } else {
        iss_io::regulate(cells, nullptr, 0);
    }
    const long n = static_cast<long>(cells.size());
    for (long i = 0; i < n && i < cap; i++) {
        const FO_surf& s = cells[i];
        const float v[34] = {s.tau, s.xpt, s.ypt, s.eta, s.da0, s.da1, s.da2, s.da3, s.u0, s.u1, s.u2,
                             s.u3, s.Edec, s.Tdec, s.Pdec, s.Bn, s.muB, s.muS, s.muQ, s.pi00, s.pi01,
                             s.pi02, s.pi03, s.pi11, s.pi12, s.pi13, s.pi22, s.pi23, s.pi33, s.bulkPi,
                             s.qmu0, s.qmu1, s.qmu2, s.qmu3};
        std::memcpy(out34 + 34 * i, v, sizeof(v));
    }
    return n;
}
