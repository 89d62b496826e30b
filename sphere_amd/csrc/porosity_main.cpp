// porosity_main.cpp -- the `porosity` tool (src/porosity.cpp:22-99): for
// each input .bin, one "z<TAB>porosity" line per horizontal slice on stdout,
// computed by DEM::porosity on the device.
//   -h/--help, -V/--version, -v/--verbose, -s/--slices N (default 10) as in
//   the reference; --exact uses exact slab volumes instead of the
//   reference's rule.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/sphere_dem.hpp"

int main(int argc, char** argv) {
    int verbose = 0, nfiles = 0, slices = 10, rule = SPH_POROSITY_REFERENCE;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << argv[0] << ": porosity of horizontal slices of a sphere input\n"
                      << "usage: " << argv[0] << " [OPTION[S]]... [FILE1 ...]\n"
                      << "  -h, --help        print this help\n"
                      << "  -V, --version     print version\n"
                      << "  -v, --verbose     name each input file\n"
                      << "  -s, --slices N    number of horizontal slices (default 10)\n"
                      << "      --exact       exact sphere volumes in each slice\n"
                      << "Writes one 'z<TAB>porosity' line per slice to stdout." << std::endl;
            return 0;
        }
        if (a == "-V" || a == "--version") {
            std::cout << "porosity (gfx950 DEM), C-ABI " << SPH_ABI_VERSION << std::endl;
            return 0;
        }
        if (a == "-v" || a == "--verbose") { verbose = 1; continue; }
        if (a == "--exact") { rule = SPH_POROSITY_EXACT; continue; }
        if (a == "-s" || a == "--slices") {
            if (i + 1 >= argc) {
                std::cerr << argv[0] << ": " << a << " needs a value" << std::endl;
                return 1;
            }
            slices = std::atoi(argv[++i]);
            if (slices < 1) {
                std::cerr << "Error: The number of slices must be a positive, real number (was " << slices << ")"
                          << std::endl;
                return 1;
            }
            continue;
        }
        ++nfiles;
        if (verbose) std::cout << argv[0] << ": processing input file: " << a << std::endl;
        try {
            sph::DEM dem(a, verbose, 0);
            dem.porosity(slices, rule);
        } catch (const sph::Error& e) {
            std::cerr << e.what() << std::endl;
            return 1;
        }
    }
    if (nfiles == 0) {
        std::cerr << argv[0] << ": missing input binary file\n"
                  << "See `" << argv[0] << " --help` for more information" << std::endl;
        return 1;
    }
    return 0;
}
