// sphere_main.cpp -- command-line front end of the DEM time step.
//
// Mirrors the reference's `sphere [options] input.bin ...` for the
// time-step path (src/main.cpp:27-182): -h/--help, -V/--version,
// -q/--quiet, -n/--dry, then each input is run with DEM::startTime, writing
// output/<sid>.outputNNNNN.bin and output/<sid>.status.dat.  Rendering and
// the fluid solvers of the reference's other flags are outside this path.
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>

#include "../../include/sphere_dem.hpp"

int main(int argc, char** argv) {
    int verbose = 1, dry = 0, nfiles = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::printf("usage: %s [OPTION[S]]... [FILE1 ...]\n"
                        "  -h, --help     print this help\n"
                        "  -V, --version  print version\n"
                        "  -q, --quiet    suppress status messages\n"
                        "  -n, --dry      read and check the input, report values, do not run\n",
                        argv[0]);
            return 0;
        }
        if (a == "-V" || a == "--version") {
            std::printf("sphere (gfx950 DEM time step), C-ABI %d\n", SPH_ABI_VERSION);
            return 0;
        }
        if (a == "-q" || a == "--quiet") { verbose = 0; continue; }
        if (a == "-n" || a == "--dry") { dry = 1; continue; }
        if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "%s: unsupported option '%s'\n", argv[0], a.c_str());
            return 2;
        }
        ++nfiles;
        try {
            sph::DEM dem(a, verbose, 1, dry);
            if (dry) {
                std::printf("%s: input is valid\n", a.c_str());
                continue;
            }
            if (verbose) dem.reportValues();
            mkdir("output", 0755);
            dem.startTime("output");
        } catch (const sph::Error& e) {
            std::fprintf(stderr, "%s: %s (status %d)\n", a.c_str(), e.what(), e.code());
            return 1;
        }
    }
    if (nfiles == 0) {
        std::fprintf(stderr, "%s: no input file given (see --help)\n", argv[0]);
        return 1;
    }
    return 0;
}
