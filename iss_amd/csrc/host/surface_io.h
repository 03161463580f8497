// File formats of the iSS drop-in: MUSIC freeze-out surfaces and their
// music_input side file, the PDG / chosen-particle lists, the HRG EoS and
// delta-f coefficient tables, and the three hadron-list writers.
#ifndef ISS_AMD_SURFACE_IO_H_
#define ISS_AMD_SURFACE_IO_H_

#include <array>
#include <string>
#include <vector>

#include "data_struct.h"

namespace iss_io {

// <work_dir>/music_input switches (src/readindata.cpp:44-94)
struct MusicInput {
    bool found = false;
    int bulk = 0, rhob = 0, diff = 0;
    bool binary = true;
    int iEOS = 0;
};
MusicInput read_music_input(const std::string& work_dir, int bulk, int rhob, int diff);

// MUSIC surfaces: boost-invariant (mode 1, src/readindata.cpp:395-546) and
// 3+1D (mode 2, src/readindata.cpp:626-765); text or 34-float binary records.
bool read_surface(const std::string& file, int hydro_mode, const MusicInput& in,
                  std::vector<FO_surf>& out, std::string* err);

// PDG list with anti-baryons and statistics sign (src/readindata.cpp:971-1116)
std::vector<particle_info> read_pdg(const std::string& file, bool quantum_statistics);
std::vector<int> read_chosen(const std::string& file);
// chosen Monte-Carlo ids -> PDG indices, stable-sorted by mass (src/FSSW.cpp:116-162)
std::vector<int> sampling_order(const std::vector<particle_info>& pdg, const std::vector<int>& chosen,
                                std::vector<int>* missing);

// HRG EoS rows {e, nB, P, T, muB, muS, muQ} (src/readindata.cpp:913-968)
bool read_hrg_eos(const std::string& file, int iEOS, std::vector<double>& rows7);
int hrg_lookup(const std::vector<double>& rows7, int iEOS, double ed, double nB, double out[5]);
// regulation (src/readindata.cpp:768-842, 1216-1246), host copy for diagnostics
void regulate(std::vector<FO_surf>& cells, const std::vector<double>* eos_rows, int iEOS);

// delta-f coefficient tables
struct Kind11 {
    int nT = 0, nmu = 0;
    double T0 = 0, mu0 = 0, dT = 0, dmu = 0;
    std::vector<double> c0, c1, c2;  // [iT * nmu + imu]
};
bool read_kind11(const std::string& dir, Kind11& t);
bool read_neos(const std::string& file, int ncol, std::vector<double>& rows);
// Coefficients_RTA_diffusion.dat (src/FSSW.cpp:1546-1566): 100 mu_B blocks of
// 150 T rows (T, mu_B, D); returned as D[iT * 100 + imu]
bool read_qmu_table(const std::string& file, std::vector<double>& D);

// writers (src/FSSW.cpp:365-561)
bool write_binary(const std::string& file, const std::vector<std::vector<iSS_Hadron>*>& events);
bool write_oscar(const std::string& file, const std::string& header_file,
                 const std::vector<std::vector<iSS_Hadron>*>& events);
bool write_gzip(const std::string& file, const std::vector<std::vector<iSS_Hadron>*>& events);

}  // namespace iss_io

#endif  // ISS_AMD_SURFACE_IO_H_
