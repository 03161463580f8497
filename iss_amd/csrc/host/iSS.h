// The iSS drop-in class (src/iSS.h:16-102): same constructor, parameter
// access, surface read-in, sampling and hadron-list getters, with the FSSW
// sampling path (MC_sampling = 4) running on the GPU through the C-ABI of
// include/iss_hip.h.  Returned hadron-list pointers stay valid until the
// next generate_samples(), clear() or the destructor.  One instance per host
// thread, as in the reference.
#ifndef ISS_AMD_ISS_H_
#define ISS_AMD_ISS_H_

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "ParameterReader.h"
#include "data_struct.h"

struct iss_ctx;

class iSS {
 private:
    const std::string path_;
    const std::string table_path_;
    const std::string particle_table_path_;
    const std::string surface_filename_;

    std::vector<FO_surf> FOsurf_array_;          // regulated lab-frame cells
    std::vector<FO_surf_LRF> FOsurf_LRF_array_;  // rest-frame cells (from the GPU)
    std::vector<float> FOsurf_Tmunu_;
    std::vector<float> FOsurf_Q_;
    std::vector<particle_info> particle_;
    AfterburnerType afterburner_type_ = AfterburnerType::UrQMD;
    int iEOS_ = 0;
    int flag_PCE_ = 0;
    long randomSeed_ = 0;
    bool seeded_ = false;

    iss_ctx* ctx_ = nullptr;
    std::vector<std::vector<iSS_Hadron>*> hadron_list_;
    std::vector<double> species_dNdy_;  // FP64 per-species totals, sampling order
    std::vector<int> species_pid_;
    std::vector<int> species_index_;    // sampling order -> index in particle_

    int ensure_context();
    int upload_species_and_tables(std::string* err);
    void release_hadrons();

 public:
    iSS(std::string path, std::string table_path = "iSS_tables",
        std::string particle_table_path = "iSS_tables",
        std::string inputfile = "iSS_parameters.dat",
        std::string surface_filename = "surface.dat");
    ~iSS();
    iSS(const iSS&) = delete;
    iSS& operator=(const iSS&) = delete;

    ParameterReader* paraRdr_ptr;

    void set_random_seed();
    void set_random_seed(int randomSeed_in);

    void perform_checks();
    void construct_Tmunu_from_particle_samples();

    int shell();
    int read_in_FO_surface();
    int generate_samples();
    // smooth dN/(pT dpT dphi dy) and v_n of every chosen species on the
    // bin_tables grids, written as the reference's calculate_vn output
    int calculate_spectra_and_flows();

    int get_number_of_sampled_events() { return static_cast<int>(hadron_list_.size()); }
    int get_number_of_particles(int iev) { return static_cast<int>(hadron_list_[iev]->size()); }
    iSS_Hadron get_hadron(int iev, int ipart) { return (*hadron_list_[iev])[ipart]; }
    std::vector<iSS_Hadron>* get_hadron_list_iev(const int iev) { return hadron_list_.at(iev); }

    void clear();

    void transform_to_local_rest_frame(std::vector<FO_surf>& FOsurf_ptr,
                                       std::vector<FO_surf_LRF>& FOsurf_LRF_ptr);
    void computeFOSurfTmunu(std::vector<FO_surf>& FOsurf_ptr);
    void getParticleQuantumNumbers(long monval, std::array<int, 3>& Qarr);

    // per-species FP64 dN/dy of the last yield pass, in sampling (mass) order
    const std::vector<double>& get_species_dNdy() const { return species_dNdy_; }
    const std::vector<int>& get_species_pid() const { return species_pid_; }
};

#endif  // ISS_AMD_ISS_H_
