// sphere_dem.hpp -- C++ DEM class over the C-ABI of sphere_dem.h.
//
// Header-only drop-in for the reference's DEM class (src/sphere.h:248-321):
// the same constructor arguments and public members for the time-step path.
// Failures throw sph::Error carrying the C-ABI status and message where the
// reference calls exit() (src/utility.cu:9-28, src/sphere.cpp:104-246).
// Link with libsphere_dem.so.
#ifndef SPHERE_DEM_HPP_
#define SPHERE_DEM_HPP_

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
    sph_t* handle() const {
        if (!h_) throw Error(SPH_E_STATE, dry_ ? "dry-run DEM has no device state" : "DEM is closed");
        return h_;
    }
    const std::string& sid() const { return sid_; }

private:
    void check(int rc, const char* msg = nullptr) {
        if (rc) throw Error(rc, msg ? msg : sph_last_error(h_));
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
