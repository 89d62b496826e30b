// sphere_dem.hpp -- C++ DEM class over the C-ABI of sphere_dem.h.
//
// Header-only drop-in for the reference's DEM class (src/sphere.h:248-321):
// the same constructor arguments and public members for the time-step path.
// Failures throw sph::Error carrying the C-ABI status and message where the
// reference calls exit() (src/utility.cu:9-28, src/sphere.cpp:104-246).
// Link with libsphere_dem.so.
#ifndef SPHERE_DEM_HPP_
#define SPHERE_DEM_HPP_

#include <algorithm>
#include <array>
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sphere_dem.h"

namespace sph {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class DEM {
public:
    // DEM::DEM (src/sphere.cpp:16-77).  dry=1 reads and checks only (no GPU)
    // and then throws nothing; initCuda/transferConstMem are accepted for
    // source compatibility (device setup always happens); darcyflow != 0 is
    // outside this path.
    explicit DEM(const std::string& inputbin, int verbosity = 1, int checkVals = 1, int dry = 0,
                 int initCuda = 1, int transferConstMem = 1, int darcyflow = 0)
        : inputbin_(inputbin), verbose_(verbosity) {
        (void)initCuda;
        (void)transferConstMem;
        if (darcyflow) throw Error(SPH_E_UNSUPPORTED, "Darcy flow is outside this path");
        const size_t dot = inputbin.rfind('.'), slash = inputbin.rfind('/');
        const size_t b = slash == std::string::npos ? 0 : slash + 1;
        sid_ = inputbin.substr(b, dot == std::string::npos || dot < b ? std::string::npos : dot - b);
        if (dry) {
            std::vector<char> buf = read_file(inputbin);
            char err[512] = {0};
            check(sph_dry_run(buf.data(), buf.size(), checkVals, nullptr, 0, nullptr, err, sizeof(err)), err);
            dry_ = true;
            return;
        }
        sph_t* h = nullptr;
        const int rc = sph_open(inputbin.c_str(), verbosity, checkVals, &h);
        if (rc) throw Error(rc, sph_last_error(nullptr));
        h_ = h;
    }
    ~DEM() { sph_close(h_); }
    DEM(const DEM&) = delete;
    DEM& operator=(const DEM&) = delete;

    // DEM::readbin (src/file_io.cpp:29-301): replaces the simulation, keeping
    // the handle's settings; on a read or check error the state is unchanged.
    void readbin(const char* target) { check(sph_readbin(handle(), target)); }
    // DEM::writebin (src/file_io.cpp:304-490)
    void writebin(const char* target) { check(sph_writebin(handle(), target)); }
    // DEM::checkValues (src/sphere.cpp:104-246)
    void checkValues() { check(sph_check_values(handle())); }
    // DEM::reportValues (src/sphere.cpp:249-341), abridged
    void reportValues() {
        sph_info_t i;
        check(sph_info(handle(), &i));
        std::printf("  np = %llu, nb0 = %llu, nw = %llu, grid = %u x %u x %u, periodic = %d\n"
                    "  dt = %g s, current = %g s, total = %g s, file_dt = %g s\n",
                    (unsigned long long)i.np, (unsigned long long)i.nb0, (unsigned long long)i.nw, i.num[0],
                    i.num[1], i.num[2], i.periodic, i.dt, i.t_current, i.t_total, i.file_dt);
    }
    // DEM::startTime (src/device.cu:576-1299): output/<sid>.outputNNNNN.bin
    // every file_dt and output/<sid>.status.dat.
    void startTime(const std::string& outdir = "output") { check(sph_run(handle(), outdir.c_str(), sid_.c_str())); }

    // Beyond the reference: n steps without file output, field access.
    void step(uint64_t n) { check(sph_step(handle(), n)); }
    template <class T> std::vector<T> get(int field) {
        std::vector<T> out(sph_field_bytes(handle(), field) / sizeof(T));
        check(sph_get(handle(), field, out.data(), out.size() * sizeof(T)));
        return out;
    }
    template <class T> void set(int field, const std::vector<T>& v) {
        check(sph_set(handle(), field, v.data(), v.size() * sizeof(T)));
    }
    // DEM::porosity (src/sphere.cpp:363-483): one "z<TAB>porosity" line per
    // slice on `out`; rule SPH_POROSITY_EXACT for exact slab volumes.
    void porosity(const int z_slices = 10, int rule = SPH_POROSITY_REFERENCE, std::ostream& out = std::cout) {
        std::vector<double> z(z_slices > 0 ? z_slices : 0), p(z.size());
        check(sph_porosity(handle(), z_slices, rule, z.data(), p.data()));
        for (size_t k = 0; k < z.size(); ++k) out << z[k] << '\t' << p[k] << '\n';
    }

    // DEM::findOverlaps (src/sphere.cpp:553-604)
    void findOverlaps(std::vector<std::vector<unsigned int>>& ij, std::vector<double>& delta_n_ij) {
        uint64_t n = 0;
        check(sph_find_overlaps(handle(), nullptr, nullptr, 0, &n));
        std::vector<uint32_t> pairs(2 * n);
        std::vector<double> dn(n);
        if (n) check(sph_find_overlaps(handle(), pairs.data(), dn.data(), n, &n));
        for (uint64_t k = 0; k < n; ++k) {
            ij.push_back({pairs[2 * k], pairs[2 * k + 1]});
            delta_n_ij.push_back(dn[k]);
        }
    }

    // DEM::minPos / DEM::maxPos (src/sphere.cpp:486-551): componentwise
    // extremes of the positions, starting from 0 as the reference does.
    std::array<double, 3> minPos() { return extreme(true); }
    std::array<double, 3> maxPos() { return extreme(false); }

    // DEM::forcechains (src/sphere.cpp:606-760): the contacts with
    // f_n = -k_n delta_n >= lower_cutoff, as a table (format "txt") or as a
    // gnuplot script (any other format; "png", "epslatex" and
    // "epslatex-color" add a terminal).  In the 3-D gnuplot arrows the
    // second point's y and z are separated by a comma, which the reference
    // omits (src/sphere.cpp:733).
    void forcechains(const std::string format = "interactive", const int threedim = 1,
                     const double lower_cutoff = 0.0, const double upper_cutoff = 1.0e9,
                     std::ostream& out = std::cout) {
        std::vector<std::vector<unsigned int>> ij;
        std::vector<double> dn;
        findOverlaps(ij, dn);
        const std::array<double, 3> lo = minPos(), hi = maxPos();
        const std::vector<double> x = get<double>(SPH_X), vel = get<double>(SPH_VEL);
        sph_params_t prm;
        check(sph_get_params(handle(), &prm));
        const double f_n_max = dn.empty() ? 0.0 : -prm.k_n * *std::min_element(dn.begin(), dn.end());
        const bool txt = format == "txt", d3 = threedim == 1;
        if (txt) {
            out << "x_1, [m]\t" << (d3 ? "y_1, [m]\t" : "") << "z_1, [m]\t"
                << "x_2, [m]\t" << (d3 ? "y_2, [m]\t" : "") << "z_2, [m]\t" << "||f_n||, [N]" << std::endl;
        } else {
            std::string name = sid_;
            std::replace(name.begin(), name.end(), '.', '-');
            out << "#!/usr/bin/env gnuplot\n"
                << "# This Gnuplot script is automatically generated using\n"
                << "# the forcechain utility in sphere. For more information,\n"
                << "# see https://github.com/anders-dc/sphere\n"
                << "set size ratio -1\n";
            if (format == "png")
                out << "set term pngcairo size 30 cm,20 cm\n" << "set out '" << name << "-fc.png'\n";
            else if (format == "epslatex")
                out << "set term epslatex size 8.6 cm, 5.6 cm\n" << "set out '" << name << "-fc.tex'\n";
            else if (format == "epslatex-color")
                out << "set term epslatex color size 8.6 cm, 5.6 cm\n" << "set out 'plots/" << name << "-fc.tex'\n";
            out << "set xlabel '\\sffamily $x_1$, [m]'\n";
            if (d3)
                out << "set ylabel '\\sffamily $x_2$, [m]'\n" << "set zlabel '\\sffamily $x_3$, [m]' offset 2\n";
            else
                out << "set ylabel '\\sffamily $x_3$, [m]' offset 2\n";
            out << "set cblabel '\\sffamily $||\\boldsymbol{f}_n||$, [N]'\n"
                << "set xyplane at " << lo[2] << '\n' << "set pm3d\n" << "set view 90.0,0.0\n"
                << "set palette defined ( 1 '#000fff', 2 '#0090ff', 3 '#0fffee', 4 '#90ff70', 5 '#ffee00', "
                   "6 '#ff7000', 7 '#ee0000', 8 '#7f0000')\n"
                << "set cbrange [" << lower_cutoff << ':' << upper_cutoff << "]\n" << std::endl;
        }
        for (size_t n = 0; n < ij.size(); ++n) {
            const size_t i = ij[n][0], j = ij[n][1];
            const double f_n = -prm.k_n * dn[n];
            if (f_n < lower_cutoff) continue;
            const double* a = &x[4 * i];
            const double* b = &x[4 * j];
            if (txt) {
                out << a[0] << '\t';
                if (d3) out << a[1] << '\t';
                out << a[2] << '\t' << b[0] << '\t';
                if (d3) out << b[1] << '\t';
                out << b[2] << '\t' << f_n << std::endl;
            } else if (f_n > lower_cutoff && (vel[4 * i + 3] + vel[4 * j + 3]) == 0.0) {
                // contacts between free particles only (fixvel sum zero)
                out << "set arrow " << n + 1 << " from " << a[0] << ',';
                if (d3) out << a[1] << ',';
                out << a[2] << " to " << b[0] << ',';
                if (d3) out << b[1] << ',';
                out << b[2] << " nohead " << "lw " << f_n / f_n_max * 8.0 << " lc palette cb " << f_n << std::endl;
            }
        }
        if (!txt) {
            out << (d3 ? "splot " : "plot ") << '[' << lo[0] << ':' << hi[0] << "] ";
            if (d3) out << '[' << lo[1] << ':' << hi[1] << "] ";
            out << '[' << lo[2] << ':' << hi[2] << "] " << "NaN notitle" << std::endl;
        }
    }

    // DEM::render (src/raytracer.cuh:374-553): writes img_out/<sid>.ppm (the
    // reference's DEM::writePPM format, src/file_io.cpp:493-516).
    // focalLength is accepted and unused, as in the reference's cameraInit.
    void render(const int method = 1, const float maxval = 1.0e3f, const float lower_cutoff = 0.0f,
                const float focalLength = 1.0f, const unsigned int img_width = 800,
                const unsigned int img_height = 800, const std::string& outdir = "img_out") {
        (void)focalLength;
        std::vector<unsigned char> rgb(3 * (size_t)img_width * img_height);
        check(sph_render(handle(), method, maxval, lower_cutoff, img_width, img_height, 1, rgb.data()));
        const std::string target = outdir + "/" + sid_ + ".ppm";
        FILE* f = std::fopen(target.c_str(), "wb");
        if (!f) throw Error(SPH_E_IO, "Could not create output PPM file '" + target + "'");
        std::fprintf(f, "P6 %u %u 255\n", img_width, img_height);
        const size_t wrote = std::fwrite(rgb.data(), 1, rgb.size(), f);
        std::fclose(f);
        if (wrote != rgb.size()) throw Error(SPH_E_IO, "short write to '" + target + "'");
        if (verbose_ == 1) std::printf("  Saving image: %s\n", target.c_str());
    }

    sph_t* handle() const {
        if (!h_) throw Error(SPH_E_STATE, dry_ ? "dry-run DEM has no device state" : "DEM is closed");
        return h_;
    }
    const std::string& sid() const { return sid_; }

private:
    void check(int rc, const char* msg = nullptr) {
        if (rc) throw Error(rc, msg ? msg : sph_last_error(h_));
    }
    std::array<double, 3> extreme(bool lower) {
        const std::vector<double> x = get<double>(SPH_X);
        std::array<double, 3> e = {0.0, 0.0, 0.0};
        for (size_t i = 0; i < x.size(); i += 4)
            for (int a = 0; a < 3; ++a) e[a] = lower ? std::min(e[a], x[i + a]) : std::max(e[a], x[i + a]);
        return e;
    }
    static std::vector<char> read_file(const std::string& path) {
        std::vector<char> buf;
        if (FILE* f = std::fopen(path.c_str(), "rb")) {
            char tmp[1 << 16];
            size_t n;
            while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
            std::fclose(f);
            return buf;
        }
        throw Error(SPH_E_IO, "could not read input binary file '" + path + "'");
    }

    std::string inputbin_, sid_;
    int verbose_ = 1;
    bool dry_ = false;
    sph_t* h_ = nullptr;
};

}  // namespace sph

#endif
