// binio.h -- host-side state of one simulation and its .bin codec.
#ifndef SPHERE_BINIO_H_
#define SPHERE_BINIO_H_

#include <stdint.h>

#include <string>
#include <vector>

namespace sph {

// Host mirror of the reference's DEM members (src/sphere.h:12-127) in the
// file's own array layout.
struct HostState {
    uint32_t nd = 3, np = 0;
    double dt = 0, t_cur = 0, t_total = 0, file_dt = 0;
    uint32_t step_count = 0;
    double origo[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    uint32_t num[3] = {0, 0, 0};
    int32_t periodic = 0;
    std::vector<double> x;       // np*4: x, y, z, r
    std::vector<double> xysum;   // np*2
    std::vector<double> vel;     // np*4: vx, vy, vz, fixvel
    std::vector<double> force, angpos, angvel, torque;  // np*3 each
    std::vector<double> es_dot, es, ev_dot, ev, p;      // np each
    double g[3] = {0, 0, 0};
    double k_n = 0, k_t = 0, k_r = 0, gamma_n = 0, gamma_t = 0, gamma_r = 0;
    double mu_s = 0, mu_d = 0, mu_r = 0, gamma_wn = 0, gamma_wt = 0, mu_ws = 0, mu_wd = 0, rho = 0;
    uint32_t contactmodel = 2;
    double kappa = 0, db = 0, V_b = 0;
    uint32_t nw = 0;
    std::vector<int32_t> wmode;  // nw
    std::vector<double> wnx;     // nw*4
    std::vector<double> wmvfd;   // nw*4
    double devs_A = 0, devs_f = 0;
    double lambda_bar = 1, sigma_b = 0, tau_b = 0;
    uint32_t nb0 = 0;
    std::vector<uint32_t> bonds;        // nb0*2
    std::vector<double> bonds_delta_n;  // nb0
    std::vector<double> bonds_delta_t;  // nb0*3
    std::vector<double> bonds_omega_n;  // nb0
    std::vector<double> bonds_omega_t;  // nb0*3
    double nu = 0;
};

// readbin (src/file_io.cpp:29-301); returns 0 or SPH_E_IO / SPH_E_UNSUPPORTED
// with a message.  Bond arrays are sized nb0 (the reference sizes them np).
int parse_bin(const void* data, size_t bytes, HostState& st, std::string& err);
int read_bin_file(const char* path, HostState& st, std::string& err);
// writebin (src/file_io.cpp:304-490)
void serialize_bin(const HostState& st, std::vector<char>& out);
int write_bin_file(const char* path, const HostState& st, std::string& err);
// checkValues (src/sphere.cpp:104-246); returns 0 or SPH_E_INVALID.
int check_values(const HostState& st, std::string& err);

}  // namespace sph

#endif
