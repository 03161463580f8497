// Data types of the iSS drop-in surface.  Field order and types match the
// reference's structs (src/data_struct.h:8-84) so that code written against
// iSS compiles unchanged and iSS_Hadron stays binary compatible (40 bytes).
#ifndef ISS_AMD_DATA_STRUCT_H_
#define ISS_AMD_DATA_STRUCT_H_

#include <array>
#include <string>
#include <vector>

namespace iSS_data {
const double hbarC = 0.197327053;  // GeV fm
typedef std::array<float, 4> Vec4;
const int AMOUNT_OF_OUTPUT = 0;
}  // namespace iSS_data

enum class AfterburnerType { UrQMD, SMASH, JAM, PDG_Decay };

typedef struct {
    int decay_Npart;
    double branching_ratio;
    int decay_part[5];
} decay_channel_info;

typedef struct {
    int monval;
    std::string name;
    double mass;
    double width;
    int gspin;
    int baryon;
    int strange;
    int charm;
    int bottom;
    int gisospin;
    int charge;
    int decays;
    int stable;
    std::vector<decay_channel_info*> decay_channels;
    int sign;
} particle_info;

typedef struct {
    float tau, xpt, ypt, eta;
    float da0, da1, da2, da3;
    float u0, u1, u2, u3;
    float Edec, Tdec, Pdec;
    float Bn, muB, muS, muQ;
    float pi00, pi01, pi02, pi03, pi11, pi12, pi13, pi22, pi23, pi33;
    float bulkPi;
    float qmu0, qmu1, qmu2, qmu3;
    std::vector<float> particle_mu_PCE;
} FO_surf;

typedef struct {
    float tau, xpt, ypt, eta;
    iSS_data::Vec4 da_mu_LRF;
    iSS_data::Vec4 u_tz;
    float Edec, Tdec, Pdec;
    float Bn, muB, muS, muQ;
    float bulkPi;
    float piLRF_xx, piLRF_xy, piLRF_xz, piLRF_yy, piLRF_yz;
    float qmuLRF_x, qmuLRF_y, qmuLRF_z;
    std::vector<float> particle_mu_PCE;
} FO_surf_LRF;

struct iSS_Hadron {
    int pid;
    float mass;
    float E, px, py, pz;
    float t, x, y, z;
};

static_assert(sizeof(iSS_Hadron) == 40, "iSS_Hadron must stay 40 bytes");

#endif  // ISS_AMD_DATA_STRUCT_H_
