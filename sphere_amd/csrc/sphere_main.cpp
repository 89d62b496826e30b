// sphere_main.cpp -- command-line front end of the DEM time step.
//
// Mirrors the reference's `sphere [options] input.bin ...`
// (src/main.cpp:27-182): -h/--help, -V/--version, -q/--quiet, -n/--dry,
// -dc/--dont-check, then each input is run with DEM::startTime, writing
// output/<sid>.outputNNNNN.bin and output/<sid>.status.dat; or, with
// -r/--render or -m/--method <method> [<maxval> [-l <cutoff>]], ray traced
// into img_out/<sid>.ppm.  The fluid solvers' flags are outside this path.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>

#include "../../include/sphere_dem.hpp"

int main(int argc, char** argv) {
    int verbose = 1, dry = 0, nfiles = 0, check_vals = 1, render = 0, method = 0;
    float max_val = 0.0f, lower_cutoff = 0.0f;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::printf("usage: %s [OPTION[S]]... [FILE1 ...]\n"
                        "  -h, --help     print this help\n"
                        "  -V, --version  print version\n"
                        "  -q, --quiet    suppress status messages\n"
                        "  -n, --dry      read and check the input, report values, do not run\n"
                        "  -r, --render   render the inputs to img_out/<sid>.ppm instead of running them\n"
                        "  -dc, --dont-check  do not check the input values\n"
                        "  -m, --method <method> <maxval> [-l <cutoff>]\n"
                        "                 colour by: normal (no maxval), pres, vel, angvel, xdisp, angpos;\n"
                        "                 with -l, spheres at or below the cutoff are not drawn\n",
                        argv[0]);
            return 0;
        }
        if (a == "-V" || a == "--version") {
            std::printf("sphere (gfx950 DEM time step), C-ABI %d\n", SPH_ABI_VERSION);
            return 0;
        }
        if (a == "-q" || a == "--quiet") { verbose = 0; continue; }
        if (a == "-n" || a == "--dry") { dry = 1; continue; }
        if (a == "-r" || a == "--render") { render = 1; check_vals = 0; continue; }
        if (a == "-dc" || a == "--dont-check") { check_vals = 0; continue; }
        if (a == "-m" || a == "--method") {
            static const char* names[] = {"normal", "pres", "vel", "angvel", "xdisp", "angpos"};
            render = 1;
            method = -1;
            for (int k = 0; k < 6 && i + 1 < argc; ++k)
                if (std::strcmp(argv[i + 1], names[k]) == 0) method = k;
            if (method < 0) {
                std::fprintf(stderr, "Visualization method not understood. See `%s --help` for more information.\n",
                             argv[0]);
                return 1;
            }
            ++i;
            if (method != 0) {
                if (i + 1 >= argc) {
                    std::fprintf(stderr, "%s: -m %s needs a maximum value\n", argv[0], names[method]);
                    return 1;
                }
                max_val = (float)std::atof(argv[++i]);
                if (i + 2 < argc && std::strcmp(argv[i + 1], "-l") == 0) {
                    lower_cutoff = (float)std::atof(argv[i + 2]);
                    i += 2;
                }
            }
            continue;
        }
        if (!a.empty() && a[0] == '-') {
            std::fprintf(stderr, "%s: unsupported option '%s'\n", argv[0], a.c_str());
            return 2;
        }
        ++nfiles;
        try {
            sph::DEM dem(a, verbose, check_vals, dry);
            if (dry) {
                std::printf("%s: input is valid\n", a.c_str());
                continue;
            }
            if (render) {
                mkdir("img_out", 0755);
                dem.render(method, max_val, lower_cutoff);
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
