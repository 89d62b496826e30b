// forcechains_main.cpp -- the `forcechains` tool (src/forcechains.cpp:22-109):
// for each input .bin, the overlapping pairs with their normal force
// f_n = -k_n delta_n as a gnuplot script or a table on stdout
// (DEM::forcechains).
//   -h/--help, -V/--version, -v/--verbose, -f/--format F (interactive, png,
//   epslatex, epslatex-color, txt), -lc/--lower-cutoff V,
//   -uc/--upper-cutoff V, -2d as in the reference.
#include <cstdlib>
#include <iostream>
#include <string>

#include "../../include/sphere_dem.hpp"

int main(int argc, char** argv) {
    int verbose = 0, nfiles = 0, threedim = 1;
    double lower = 0.0, upper = 1.0e9;
    std::string format = "interactive";
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << argv[0] << ": force chains of a sphere input\n"
                      << "usage: " << argv[0] << " [OPTION[S]]... [FILE1 ...] > outputfile\n"
                      << "  -h, --help                 print this help\n"
                      << "  -V, --version              print version\n"
                      << "  -v, --verbose              name each input file\n"
                      << "  -lc, --lower-cutoff V      only contacts with a normal force above V\n"
                      << "  -uc, --upper-cutoff V      top of the colour range\n"
                      << "  -f, --format F             interactive (default), png, epslatex,\n"
                      << "                             epslatex-color (gnuplot scripts) or txt\n"
                      << "  -2d                        x and z coordinates only" << std::endl;
            return 0;
        }
        if (a == "-V" || a == "--version") {
            std::cout << "forcechains (gfx950 DEM), C-ABI " << SPH_ABI_VERSION << std::endl;
            return 0;
        }
        if (a == "-v" || a == "--verbose") { verbose = 1; continue; }
        if (a == "-2d") { threedim = 0; continue; }
        if (a == "-f" || a == "--format" || a == "-lc" || a == "--lower-cutoff" || a == "-uc" ||
            a == "--upper-cutoff") {
            if (i + 1 >= argc) {
                std::cerr << argv[0] << ": " << a << " needs a value" << std::endl;
                return 1;
            }
            const char* v = argv[++i];
            if (a == "-f" || a == "--format") format = v;
            else if (a == "-lc" || a == "--lower-cutoff") lower = std::atof(v);
            else upper = std::atof(v);
            continue;
        }
        ++nfiles;
        if (verbose) std::cout << argv[0] << ": processing input file: " << a << std::endl;
        try {
            sph::DEM dem(a, verbose, 0);
            dem.forcechains(format, threedim, lower, upper);
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
