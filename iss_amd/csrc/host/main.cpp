// iSS.e: the reference's command line (src/main.cpp:22-74) over the gfx950
// sampler: iSS.e [parameter file] [work dir] [surface file] [key=value ...]
#include <iostream>
#include <string>

#include "iSS.h"

int main(int argc, char** argv) {
    const std::string table_path = "iSS_tables";
    const std::string particle_table_path = "iSS_tables";
    std::string path = "results";
    std::string input_file = "iSS_parameters.dat";
    std::string surface_filename = "surface.dat";
    if (argc > 1 && std::string(argv[1]).find('=') == std::string::npos) input_file = argv[1];
    if (argc > 2 && std::string(argv[2]).find('=') == std::string::npos) path = argv[2];
    if (argc > 3 && std::string(argv[3]).find('=') == std::string::npos) surface_filename = argv[3];
    std::cout << "input file : " << input_file << std::endl
              << "work folder path : " << path << std::endl
              << "surface filename : " << surface_filename << std::endl;
    iSS sampler(path, table_path, particle_table_path, input_file, surface_filename);
    sampler.paraRdr_ptr->readFromArguments(argc, argv);
    sampler.paraRdr_ptr->echo();
    const int status = sampler.shell();
    if (static_cast<int>(sampler.paraRdr_ptr->getVal("perform_checks", 0)) == 1) sampler.perform_checks();
    if (status == 0) std::cout << "Program executed normally." << std::endl;
    return status;
}
