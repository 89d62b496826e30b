// binio.cpp -- .bin reader/writer and input validation (host C++).
#include "binio.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "../../include/sphere_dem.h"

namespace sph {

namespace {

class Cursor {
public:
    Cursor(const void* data, size_t bytes) : p_(static_cast<const char*>(data)), n_(bytes) {}
    template <class T> bool get(T* dst, size_t count = 1) {
        const size_t sz = sizeof(T) * count;
        if (sz > n_ - off_) { ok_ = false; return false; }
        if (sz) std::memcpy(dst, p_ + off_, sz);
        off_ += sz;
        return true;
    }
    template <class T> bool getv(std::vector<T>& v, size_t count) {
        v.resize(count);
        return get(v.data(), count);
    }
    bool ok() const { return ok_; }
    size_t left() const { return n_ - off_; }
    size_t offset() const { return off_; }

private:
    const char* p_;
    size_t n_;
    size_t off_ = 0;
    bool ok_ = true;
};

template <class T> void put(std::vector<char>& out, const T* src, size_t count = 1) {
    const size_t sz = sizeof(T) * count;
    const size_t at = out.size();
    out.resize(at + sz);
    if (sz) std::memcpy(out.data() + at, src, sz);
}

}  // namespace

int parse_bin(const void* data, size_t bytes, HostState& st, std::string& err) {
    Cursor c(data, bytes);
    auto truncated = [&](const char* what) {
        std::ostringstream os;
        os << "truncated .bin while reading " << what << " at byte " << c.offset();
        err = os.str();
        return SPH_E_IO;
    };
    if (!c.get(&st.nd) || !c.get(&st.np)) return truncated("header");
    if (st.nd != 3) {
        err = "dimensionality mismatch: the file is " + std::to_string(st.nd) + "-D, only 3-D is supported";
        return SPH_E_INVALID;
    }
    if (!c.get(&st.dt) || !c.get(&st.t_cur) || !c.get(&st.t_total) || !c.get(&st.file_dt) ||
        !c.get(&st.step_count))
        return truncated("time");
    if (!c.get(st.origo, 3) || !c.get(st.L, 3) || !c.get(st.num, 3) || !c.get(&st.periodic))
        return truncated("grid");
    const size_t n = st.np;
    if ((double)n * 216.0 > (double)c.left()) return truncated("particles");
    if (!c.getv(st.x, 4 * n) || !c.getv(st.xysum, 2 * n) || !c.getv(st.vel, 4 * n) ||
        !c.getv(st.force, 3 * n) || !c.getv(st.angpos, 3 * n) || !c.getv(st.angvel, 3 * n) ||
        !c.getv(st.torque, 3 * n) || !c.getv(st.es_dot, n) || !c.getv(st.es, n) ||
        !c.getv(st.ev_dot, n) || !c.getv(st.ev, n) || !c.getv(st.p, n))
        return truncated("particles");
    double pr[14];
    if (!c.get(st.g, 3) || !c.get(pr, 14) || !c.get(&st.contactmodel) || !c.get(&st.kappa) ||
        !c.get(&st.db) || !c.get(&st.V_b))
        return truncated("params");
    st.k_n = pr[0]; st.k_t = pr[1]; st.k_r = pr[2]; st.gamma_n = pr[3]; st.gamma_t = pr[4];
    st.gamma_r = pr[5]; st.mu_s = pr[6]; st.mu_d = pr[7]; st.mu_r = pr[8]; st.gamma_wn = pr[9];
    st.gamma_wt = pr[10]; st.mu_ws = pr[11]; st.mu_wd = pr[12]; st.rho = pr[13];
    if (!c.get(&st.nw)) return truncated("walls");
    if (st.nw > 6) {
        err = "nw=" + std::to_string(st.nw) + " exceeds MAXWALLS (6)";
        return SPH_E_INVALID;
    }
    if (!c.getv(st.wmode, st.nw) || !c.getv(st.wnx, 4 * (size_t)st.nw) ||
        !c.getv(st.wmvfd, 4 * (size_t)st.nw) || !c.get(&st.devs_A) || !c.get(&st.devs_f))
        return truncated("walls");
    if (!c.get(&st.lambda_bar) || !c.get(&st.nb0) || !c.get(&st.sigma_b) || !c.get(&st.tau_b))
        return truncated("bond params");
    const size_t nb = st.nb0;
    if ((double)nb * 72.0 > (double)c.left()) return truncated("bonds");
    if (!c.getv(st.bonds, 2 * nb) || !c.getv(st.bonds_delta_n, nb) || !c.getv(st.bonds_delta_t, 3 * nb) ||
        !c.getv(st.bonds_omega_n, nb) || !c.getv(st.bonds_omega_t, 3 * nb))
        return truncated("bonds");
    if (!c.get(&st.nu)) return truncated("fluid");
    if (st.nu > 0.0) {
        err = "fluid coupling (nu > 0) is outside this path";
        return SPH_E_UNSUPPORTED;
    }
    return SPH_OK;
}

int read_bin_file(const char* path, HostState& st, std::string& err) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = std::string("could not read input binary file '") + path + "'";
        return SPH_E_IO;
    }
    std::vector<char> buf;
    char tmp[1 << 16];
    size_t got;
    while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
        err = std::string("read error on '") + path + "'";
        return SPH_E_IO;
    }
    return parse_bin(buf.data(), buf.size(), st, err);
}

void serialize_bin(const HostState& st, std::vector<char>& out) {
    out.clear();
    const size_t n = st.np, nb = st.nb0;
    out.reserve(112 + 216 * n + 164 + 20 + 68 * (size_t)st.nw + 28 + 72 * nb + 8);
    put(out, &st.nd); put(out, &st.np);
    put(out, &st.dt); put(out, &st.t_cur); put(out, &st.t_total); put(out, &st.file_dt);
    put(out, &st.step_count);
    put(out, st.origo, 3); put(out, st.L, 3); put(out, st.num, 3); put(out, &st.periodic);
    put(out, st.x.data(), 4 * n); put(out, st.xysum.data(), 2 * n); put(out, st.vel.data(), 4 * n);
    put(out, st.force.data(), 3 * n); put(out, st.angpos.data(), 3 * n);
    put(out, st.angvel.data(), 3 * n); put(out, st.torque.data(), 3 * n);
    put(out, st.es_dot.data(), n); put(out, st.es.data(), n); put(out, st.ev_dot.data(), n);
    put(out, st.ev.data(), n); put(out, st.p.data(), n);
    put(out, st.g, 3);
    const double pr[14] = {st.k_n, st.k_t, st.k_r, st.gamma_n, st.gamma_t, st.gamma_r, st.mu_s,
                           st.mu_d, st.mu_r, st.gamma_wn, st.gamma_wt, st.mu_ws, st.mu_wd, st.rho};
    put(out, pr, 14);
    put(out, &st.contactmodel); put(out, &st.kappa); put(out, &st.db); put(out, &st.V_b);
    put(out, &st.nw);
    put(out, st.wmode.data(), st.nw); put(out, st.wnx.data(), 4 * (size_t)st.nw);
    put(out, st.wmvfd.data(), 4 * (size_t)st.nw);
    put(out, &st.devs_A); put(out, &st.devs_f);
    put(out, &st.lambda_bar); put(out, &st.nb0); put(out, &st.sigma_b); put(out, &st.tau_b);
    put(out, st.bonds.data(), 2 * nb); put(out, st.bonds_delta_n.data(), nb);
    put(out, st.bonds_delta_t.data(), 3 * nb); put(out, st.bonds_omega_n.data(), nb);
    put(out, st.bonds_omega_t.data(), 3 * nb);
    put(out, &st.nu);
}

int write_bin_file(const char* path, const HostState& st, std::string& err) {
    std::vector<char> buf;
    serialize_bin(st, buf);
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        err = std::string("could not create output binary file '") + path + "'";
        return SPH_E_IO;
    }
    const size_t w = std::fwrite(buf.data(), 1, buf.size(), f);
    const int rc = std::fclose(f);
    if (w != buf.size() || rc != 0) {
        err = std::string("short write on '") + path + "'";
        return SPH_E_IO;
    }
    return SPH_OK;
}

int check_values(const HostState& st, std::string& err) {
    std::ostringstream os;
    auto fail = [&]() { err = os.str(); return SPH_E_INVALID; };
    if (st.nd != 3) { os << "Error: nd = " << st.nd << ", ND = 3"; return fail(); }
    if (st.np < 1) { os << "Error: np = " << st.np; return fail(); }
    if (st.t_cur > st.t_total || st.t_cur < 0.0) {
        os << "Error: time.current = " << st.t_cur << " s, time.total = " << st.t_total << " s";
        return fail();
    }
    if (st.origo[0] < 0.0 || st.origo[1] < 0.0 || st.origo[2] < 0.0) {
        os << "Error: Negative values in grid.origo are known to cause problems.";
        return fail();
    }
    if (st.L[0] <= 0.0 || st.L[1] <= 0.0 || st.L[2] <= 0.0) {
        os << "Error: grid.L = [" << st.L[0] << ", " << st.L[1] << ", " << st.L[2] << "] m";
        return fail();
    }
    if (st.num[0] == 0 || st.num[1] == 0 || st.num[2] == 0) {
        os << "Error: grid.num = [" << st.num[0] << ", " << st.num[1] << ", " << st.num[2] << "]";
        return fail();
    }
    if (st.periodic == 2 && st.num[0] < 3) {
        os << "Error: When 1st dimension boundaries are periodic, there must be at least 3 cells in that dimension.";
        return fail();
    }
    if (st.periodic == 1 && (st.num[0] < 3 || st.num[1] < 3)) {
        os << "Error: When 1st and 2nd dimension boundaries are periodic, there must be at least 3 cells in each of those dimensions.";
        return fail();
    }
    for (uint32_t i = 0; i < st.np; ++i) {
        const double* x = &st.x[4 * (size_t)i];
        if (!(x[3] > 0.0)) { os << "Error: Particle " << i << " has a radius of " << x[3] << " m."; return fail(); }
        if (x[0] < st.origo[0] || x[1] < st.origo[1] || x[2] < st.origo[2] || x[0] > st.L[0] ||
            x[1] > st.L[1] || x[2] > st.L[2]) {
            os << "Error: Particle " << i << " is outside of the computational grid: [" << x[0] << ", "
               << x[1] << ", " << x[2] << "]";
            return fail();
        }
    }
    if (st.nw > 0) {
        double z_max = 0.0;
        for (uint32_t i = 0; i < st.np; ++i) z_max = st.x[4 * (size_t)i + 2] > z_max ? st.x[4 * (size_t)i + 2] : z_max;
        if (st.wnx[3] < z_max) { os << "Error: One or more particles have centres above the upper, dynamic wall"; return fail(); }
    }
    if (!(st.k_n > 0.0)) { os << "Error: k_n = " << st.k_n << " N/m"; return fail(); }
    if (!(st.rho > 0.0)) { os << "Error: rho = " << st.rho << " kg/m3"; return fail(); }
    return SPH_OK;
}

}  // namespace sph

extern "C" int sph_dry_run(const void* bin, size_t bytes, int check_vals, void* out, size_t out_cap,
                           size_t* out_len, char* err, size_t err_len) {
    sph::HostState st;
    std::string msg;
    int rc = sph::parse_bin(bin, bytes, st, msg);
    if (!rc && check_vals == 1) rc = sph::check_values(st, msg);
    if (!rc && out_len) {
        std::vector<char> buf;
        sph::serialize_bin(st, buf);
        *out_len = buf.size();
        if (out) {
            if (out_cap < buf.size()) {
                msg = "output buffer too small";
                rc = SPH_E_INVALID;
            } else {
                std::memcpy(out, buf.data(), buf.size());
            }
        }
    }
    if (rc && err && err_len) std::snprintf(err, err_len, "%s", msg.c_str());
    return rc;
}
