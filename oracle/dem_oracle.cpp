// dem_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see dem_oracle.h).
//
// Serial restatement of the reference's DEM step.  Every function names the
// reference lines it restates; expressions keep the reference's evaluation
// order (left to right, no contraction: build with -ffp-contract=off) so the
// HIP kernels can be held to bit equality.  Where the reference is undefined
// the definition adopted by both sides is stated at the function.
//
// Deliberate definitions (the reference is undefined or racy there):
//  * vector lengths use IEEE double sqrt; the reference's length() narrows to
//    float through sqrtf (src/vector_arithmetic.h:317-324);
//  * periodic positions are wrapped with the integrate rule
//    (src/integration.cuh:80-95) before the cell id is computed, and the cell
//    coordinate is clamped to [0, num-1]; a particle more than one cell off
//    the grid raises ORC_F_OUT_OF_GRID (src/sorting.cuh:9-14 converts a
//    negative floor to unsigned);
//  * a contact that finds neither its own slot nor an empty one, or more than
//    NC overlaps, raises ORC_F_CONTACT_OVERFLOW and ends the particle's search
//    (src/contactsearch.cuh:252-284 writes the partner into slot nc, the
//    count of overlaps found so far: over another contact when nc < NC, one
//    past the particle's slots when nc = NC);
//  * delta_t starts at zero and a tangential history that stays untouched
//    keeps its projected value (src/contactmodels.cuh:357,487-491 leave it
//    uninitialised);
//  * each wall sums its own per-particle forces (src/integration.cuh:250-252
//    sums wall 0 for every wall);
//  * bond forces are added in ascending bond index (src/cohesion.cuh:220-226
//    races), a broken bond's first index becomes max(nb0, np);
//  * extension (no reference counterpart): fixvel >= 2 also holds the z
//    velocity (acc_z = 0), a velocity-controlled particle layer for tension
//    tests; fixvel in (0, 2) is the reference's fixed horizontal velocity
//    (src/integration.cuh:65-77), which the reference's inputs use (1);
//  * bondsLinear reads the positions before the periodic wrap, as the
//    reference does (it runs before integrate); running it first in the step
//    is otherwise equivalent because no stage before it in the reference
//    writes x, vel or angvel.

#include "dem_oracle.h"

// Planted mistakes for the pin tests' own check (tests/test_oracle_mutants.py
// builds the oracle with -DORC_MUTANT=k and expects the closed-form pins to
// fail); 0, the default, is the restatement itself.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace {

const double PI = 3.14159265358979;  // src/constants.h:14
const uint32_t EMPTY = 0xffffffffu;

struct D4 { double x, y, z, w; };
struct D2 { double x, y; };
struct V3 { double x, y, z; };

inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
// src/vector_arithmetic.h:304-307, 317-320 (double sqrt, see header note)
inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len3(V3 a) { return std::sqrt(dot3(a, a)); }
// src/vector_arithmetic.h:412-415
inline V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 smul(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }   // Float*Float3
inline V3 muls(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }   // Float3*Float
inline V3 divs(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 neg3(V3 a) { return v3(-a.x, -a.y, -a.z); }

}  // namespace

struct orc_state {
    uint32_t np = 0;
    int nc = 32;
    double dt = 0, t_cur = 0, t_total = 0, file_dt = 0;
    uint32_t step_count = 0;
    double origo[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    uint32_t num[3] = {0, 0, 0};
    int periodic = 0;
    std::vector<D4> x, vel;
    std::vector<D2> xysum;
    std::vector<V3> force, angpos, angvel, torque, vel0, angvel0;
    std::vector<double> es_dot, es, ev_dot, ev, p;
    double g[3] = {0, 0, 0};
    double k_n = 0, k_t = 0, k_r = 0, gamma_n = 0, gamma_t = 0, gamma_r = 0;
    double mu_s = 0, mu_d = 0, mu_r = 0, gamma_wn = 0, gamma_wt = 0, mu_ws = 0, mu_wd = 0, rho = 0;
    uint32_t contactmodel = 2;
    double kappa = 0, db = 0, V_b = 0;
    uint32_t nw = 0;
    int wmode[6] = {0, 0, 0, 0, 0, 0};
    D4 wnx[6], wmvfd[6];
    double wvel0[6] = {0, 0, 0, 0, 0, 0};
    double devs_A = 0, devs_f = 0;
    double lambda_bar = 1, sigma_b = 0, tau_b = 0;
    uint32_t nb0 = 0;
    std::vector<uint32_t> bonds;  // 2*nb0
    std::vector<D4> bdelta, bomega;
    double nu = 0;
    // contact bookkeeping (src/device.cu:264-267, 295-297)
    std::vector<uint32_t> contacts;
    std::vector<D4> distmod;
    std::vector<V3> delta_t;
    // per-step scratch
    std::vector<uint32_t> key, idx, key_sorted, cellStart, cellEnd;
    std::vector<D4> x_sorted, vel_sorted;
    std::vector<V3> angvel_sorted;
    std::vector<double> wall_pp, wall_partial;
    std::vector<V3> bond_f, bond_t;
    std::vector<uint8_t> bond_live;
    uint32_t flags = 0;

    uint64_t K() const { return (uint64_t)num[0] * num[1] * num[2]; }
    uint32_t blocks() const { return (np + 255) / 256; }
};

// ---------------------------------------------------------------- .bin I/O
namespace {

struct Reader {
    std::vector<char> buf;
    size_t off = 0;
    bool ok = true;
    explicit Reader(FILE* f) {
        char tmp[1 << 16];
        size_t got;
        while ((got = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
    }
    template <class T> void get(T* p, size_t n = 1) {
        if (!ok) return;
        const size_t sz = sizeof(T) * n;
        if (sz > buf.size() - off) { ok = false; return; }
        std::memcpy(p, buf.data() + off, sz);
        off += sz;
    }
};

struct Writer {
    FILE* f;
    bool ok = true;
    template <class T> void put(const T* p, size_t n = 1) {
        if (!ok) return;
        if (fwrite(p, sizeof(T), n, f) != n) ok = false;
    }
};

void seterr(char* err, size_t errlen, const std::string& m) {
    if (err && errlen) {
        std::snprintf(err, errlen, "%s", m.c_str());
    }
}

}  // namespace

// Reads the layout of src/file_io.cpp:45-237 (bond arrays sized nb0, not np:
// src/file_io.cpp:218-219 allocates np).
extern "C" orc_t* orc_open(const char* path, int nc, char* err, size_t errlen) {
    FILE* f = std::fopen(path, "rb");
    if (!f) { seterr(err, errlen, std::string("cannot open ") + path); return nullptr; }
    orc_t* s = new orc_t();
    s->nc = nc > 0 ? nc : 32;
    Reader r(f);
    uint32_t nd = 0;
    r.get(&nd); r.get(&s->np);
    if (!r.ok || nd != 3) { std::fclose(f); delete s; seterr(err, errlen, "bad header"); return nullptr; }
    r.get(&s->dt); r.get(&s->t_cur); r.get(&s->t_total); r.get(&s->file_dt); r.get(&s->step_count);
    r.get(s->origo, 3); r.get(s->L, 3); r.get(s->num, 3); r.get(&s->periodic);
    const uint32_t n = s->np;
    s->x.resize(n); s->vel.resize(n); s->xysum.resize(n);
    s->force.resize(n); s->angpos.resize(n); s->angvel.resize(n); s->torque.resize(n);
    s->es_dot.resize(n); s->es.resize(n); s->ev_dot.resize(n); s->ev.resize(n); s->p.resize(n);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->x[i].x, 4);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->xysum[i].x, 2);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->vel[i].x, 4);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->force[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->angpos[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->angvel[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) r.get(&s->torque[i].x, 3);
    r.get(s->es_dot.data(), n); r.get(s->es.data(), n); r.get(s->ev_dot.data(), n);
    r.get(s->ev.data(), n); r.get(s->p.data(), n);
    r.get(s->g, 3);
    r.get(&s->k_n); r.get(&s->k_t); r.get(&s->k_r); r.get(&s->gamma_n); r.get(&s->gamma_t);
    r.get(&s->gamma_r); r.get(&s->mu_s); r.get(&s->mu_d); r.get(&s->mu_r); r.get(&s->gamma_wn);
    r.get(&s->gamma_wt); r.get(&s->mu_ws); r.get(&s->mu_wd); r.get(&s->rho);
    r.get(&s->contactmodel); r.get(&s->kappa); r.get(&s->db); r.get(&s->V_b);
    r.get(&s->nw);
    if (!r.ok || s->nw > 6) { std::fclose(f); delete s; seterr(err, errlen, "bad wall block"); return nullptr; }
    r.get(s->wmode, s->nw);
    for (uint32_t i = 0; i < s->nw; ++i) r.get(&s->wnx[i].x, 4);
    for (uint32_t i = 0; i < s->nw; ++i) r.get(&s->wmvfd[i].x, 4);
    r.get(&s->devs_A); r.get(&s->devs_f);
    r.get(&s->lambda_bar); r.get(&s->nb0); r.get(&s->sigma_b); r.get(&s->tau_b);
    const uint32_t nb = s->nb0;
    if (!r.ok) { std::fclose(f); delete s; seterr(err, errlen, "truncated"); return nullptr; }
    s->bonds.resize(2 * (size_t)nb); s->bdelta.resize(nb); s->bomega.resize(nb);
    for (uint32_t i = 0; i < nb; ++i) r.get(&s->bonds[2 * i], 2);
    for (uint32_t i = 0; i < nb; ++i) r.get(&s->bdelta[i].w);
    for (uint32_t i = 0; i < nb; ++i) r.get(&s->bdelta[i].x, 3);
    for (uint32_t i = 0; i < nb; ++i) r.get(&s->bomega[i].w);
    for (uint32_t i = 0; i < nb; ++i) r.get(&s->bomega[i].x, 3);
    r.get(&s->nu);
    std::fclose(f);
    if (!r.ok) { delete s; seterr(err, errlen, "truncated"); return nullptr; }
    if (s->nu > 0.0) { delete s; seterr(err, errlen, "fluid (nu>0) unsupported"); return nullptr; }
    if (s->contactmodel < 1 || s->contactmodel > 3) {
        delete s;
        seterr(err, errlen, "only contactmodels 1, 2 and 3");
        return nullptr;
    }
    // src/device.cu:413,423: vel0 <- vel, angvel0 <- angvel; walls vel0 <- mvfd.y
    s->vel0.resize(n); s->angvel0.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        s->vel0[i] = v3(s->vel[i].x, s->vel[i].y, s->vel[i].z);
        s->angvel0[i] = s->angvel[i];
    }
    for (uint32_t w = 0; w < s->nw; ++w) s->wvel0[w] = s->wmvfd[w].y;
    // src/device.cu:295-297: every slot holds the sentinel np
    const size_t ns = (size_t)n * s->nc;
    s->contacts.assign(ns, n);
    s->distmod.assign(ns, D4{0, 0, 0, 0});
    s->delta_t.assign(ns, v3(0, 0, 0));
    s->key.assign(n, 0); s->idx.assign(n, 0); s->key_sorted.assign(n, 0);
    s->cellStart.assign(s->K(), EMPTY); s->cellEnd.assign(s->K(), EMPTY);
    s->x_sorted.resize(n); s->vel_sorted.resize(n); s->angvel_sorted.resize(n);
    s->wall_pp.assign((size_t)s->nw * n, 0.0);
    s->wall_partial.assign((size_t)s->nw * s->blocks(), 0.0);
    s->bond_f.assign(nb, v3(0, 0, 0)); s->bond_t.assign(nb, v3(0, 0, 0));
    s->bond_live.assign(nb, 0);
    return s;
}

extern "C" void orc_close(orc_t* s) { delete s; }

double det_exp(double x);
// The exponential the capillary force uses, for its own accuracy check.
extern "C" double orc_exp(double x) { return det_exp(x); }

// Writes the layout of src/file_io.cpp:319-452.
extern "C" int orc_writebin(orc_t* s, const char* path) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    Writer w{f};
    const uint32_t nd = 3, n = s->np;
    w.put(&nd); w.put(&n);
    w.put(&s->dt); w.put(&s->t_cur); w.put(&s->t_total); w.put(&s->file_dt); w.put(&s->step_count);
    w.put(s->origo, 3); w.put(s->L, 3); w.put(s->num, 3); w.put(&s->periodic);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->x[i].x, 4);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->xysum[i].x, 2);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->vel[i].x, 4);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->force[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->angpos[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->angvel[i].x, 3);
    for (uint32_t i = 0; i < n; ++i) w.put(&s->torque[i].x, 3);
    w.put(s->es_dot.data(), n); w.put(s->es.data(), n); w.put(s->ev_dot.data(), n);
    w.put(s->ev.data(), n); w.put(s->p.data(), n);
    w.put(s->g, 3);
    const double pr[14] = {s->k_n, s->k_t, s->k_r, s->gamma_n, s->gamma_t, s->gamma_r, s->mu_s,
                           s->mu_d, s->mu_r, s->gamma_wn, s->gamma_wt, s->mu_ws, s->mu_wd, s->rho};
    w.put(pr, 14);
    w.put(&s->contactmodel); w.put(&s->kappa); w.put(&s->db); w.put(&s->V_b);
    w.put(&s->nw);
    w.put(s->wmode, s->nw);
    for (uint32_t i = 0; i < s->nw; ++i) w.put(&s->wnx[i].x, 4);
    for (uint32_t i = 0; i < s->nw; ++i) w.put(&s->wmvfd[i].x, 4);
    w.put(&s->devs_A); w.put(&s->devs_f);
    w.put(&s->lambda_bar); w.put(&s->nb0); w.put(&s->sigma_b); w.put(&s->tau_b);
    for (uint32_t i = 0; i < s->nb0; ++i) w.put(&s->bonds[2 * i], 2);
    for (uint32_t i = 0; i < s->nb0; ++i) w.put(&s->bdelta[i].w);
    for (uint32_t i = 0; i < s->nb0; ++i) w.put(&s->bdelta[i].x, 3);
    for (uint32_t i = 0; i < s->nb0; ++i) w.put(&s->bomega[i].w);
    for (uint32_t i = 0; i < s->nb0; ++i) w.put(&s->bomega[i].x, 3);
    w.put(&s->nu);
    std::fclose(f);
    return w.ok ? 0 : -1;
}

// ------------------------------------------------------------ cell hashing
namespace {

// Periodic wrap of src/integration.cuh:80-95, applied before hashing.
void wrap_position(const orc_t* s, D4& x) {
    if (s->periodic == 1) {
        if (x.x < s->origo[0]) x.x += s->L[0];
        if (x.x > s->L[0]) x.x -= s->L[0];
        if (x.y < s->origo[1]) x.y += s->L[1];
        if (x.y > s->L[1]) x.y -= s->L[1];
    } else if (s->periodic == 2) {
        if (x.x < s->origo[0]) x.x += s->L[0];
        if (x.x > s->L[0]) x.x -= s->L[0];
    }
}

// src/sorting.cuh:12-14: floor((x - origo)/(L/num)), clamped to [0, num-1].
int cell_coord(orc_t* s, double v, int d) {
    const double f = std::floor((v - s->origo[d]) / (s->L[d] / (double)s->num[d]));
    if (f != f) { s->flags |= ORC_F_NAN; return 0; }
    if (f < -1.0 || f > (double)s->num[d]) s->flags |= ORC_F_OUT_OF_GRID;
    if (f < 0.0) return 0;
    if (f > (double)(s->num[d] - 1)) return (int)(s->num[d] - 1);
    return (int)f;
}

// src/sorting.cuh:17
uint32_t cell_id(const orc_t* s, int ix, int iy, int iz) {
    return ((uint32_t)iz * s->num[1]) * s->num[0] + (uint32_t)iy * s->num[0] + (uint32_t)ix;
}

}  // namespace

// src/sorting.cuh:24-43 (calcParticleCellID), with the wrap written back.
static void stage_cellid(orc_t* s) {
    for (uint32_t i = 0; i < s->np; ++i) {
        wrap_position(s, s->x[i]);
        const D4 x = s->x[i];
        const int ix = cell_coord(s, x.x, 0), iy = cell_coord(s, x.y, 1), iz = cell_coord(s, x.z, 2);
        s->key[i] = cell_id(s, ix, iy, iz);
        s->idx[i] = i;
    }
}

// src/device.cu:763-765 thrust::sort_by_key, defined as a stable sort.
static void stage_sort(orc_t* s) {
    std::vector<uint32_t> order(s->np);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return s->key[a] < s->key[b]; });
    for (uint32_t i = 0; i < s->np; ++i) {
        s->idx[i] = order[i];
        s->key_sorted[i] = s->key[order[i]];
    }
}

// src/device.cu:776-777 memset + src/sorting.cuh:49-121 reorderArrays.
// cellEnd of an empty cell is defined as 0xffffffff (the reference leaves it stale).
static void stage_reorder(orc_t* s) {
    std::fill(s->cellStart.begin(), s->cellStart.end(), EMPTY);
    std::fill(s->cellEnd.begin(), s->cellEnd.end(), EMPTY);
    for (uint32_t i = 0; i < s->np; ++i) {
        const uint32_t c = s->key_sorted[i];
        if (i == 0 || c != s->key_sorted[i - 1]) {
            s->cellStart[c] = i;
            if (i > 0) s->cellEnd[s->key_sorted[i - 1]] = i;
        }
        if (i == s->np - 1) s->cellEnd[c] = i + 1;
        const uint32_t o = s->idx[i];
        s->x_sorted[i] = s->x[o];
        s->vel_sorted[i] = s->vel[o];
        s->angvel_sorted[i] = s->angvel[o];
    }
}

// ------------------------------------------------------------- topology
namespace {

// src/contactsearch.cuh:12-74 findDistMod
int find_dist_mod(const orc_t* s, int& tx, int& ty, int tz, V3& dm) {
    const int nx = (int)s->num[0], ny = (int)s->num[1], nz = (int)s->num[2];
    if (s->periodic == 1) {
        if (tx < 0) { tx = nx - 1; dm.x += s->L[0]; dm.y += 0.0; dm.z += 0.0; }
        if (tx == nx) { tx = 0; dm.x -= s->L[0]; dm.y -= 0.0; dm.z -= 0.0; }
        if (ty < 0) { ty = ny - 1; dm.x += 0.0; dm.y += s->L[1]; dm.z += 0.0; }
        if (ty == ny) { ty = 0; dm.x -= 0.0; dm.y -= s->L[1]; dm.z -= 0.0; }
    } else if (s->periodic == 2) {
        if (tx < 0) { tx = nx - 1; dm.x += s->L[0]; dm.y += 0.0; dm.z += 0.0; }
        if (tx == nx) { tx = 0; dm.x -= s->L[0]; dm.y -= 0.0; dm.z -= 0.0; }
        if (ty < 0 || ty == ny) return -1;
    } else {
        if (tx < 0 || tx == nx) return -1;
        if (ty < 0 || ty == ny) return -1;
    }
    if (tz < 0 || tz == nz) return -1;
    return 0;
}

}  // namespace

// src/contactsearch.cuh:181-307 findContactsInCell + :314-355 topology
// (called for contact models 2 and 3 only, src/device.cu:809).
static void stage_topology(orc_t* s) {
    if (s->contactmodel == 1) return;
    const uint32_t NC = (uint32_t)s->nc;
    const uint32_t n = s->np;
    for (uint32_t a = 0; a < n; ++a) {
        const D4 xa = s->x_sorted[a];
        const double ra = xa.w;
        const uint32_t ao = s->idx[a];
        uint32_t* slots = &s->contacts[(size_t)ao * NC];
        D4* dms = &s->distmod[(size_t)ao * NC];
        uint32_t nc = 0;
        bool stop = false;
        const int gx = cell_coord(s, xa.x, 0), gy = cell_coord(s, xa.y, 1), gz = cell_coord(s, xa.z, 2);
        for (int dz = -1; dz < 2 && !stop; ++dz)
            for (int dy = -1; dy < 2 && !stop; ++dy)
                for (int dx = -1; dx < 2 && !stop; ++dx) {
                    int tx = gx + dx, ty = gy + dy;
                    const int tz = gz + dz;
                    V3 dm = v3(0.0, 0.0, 0.0);
                    if (find_dist_mod(s, tx, ty, tz, dm) == -1) continue;
                    const uint32_t c = (uint32_t)tx + (uint32_t)ty * s->num[0] +
                                       (s->num[0] * s->num[1]) * (uint32_t)tz;
                    const uint32_t st = s->cellStart[c];
                    if (st == EMPTY) continue;
                    const uint32_t en = s->cellEnd[c];
                    for (uint32_t b = st; b < en && !stop; ++b) {
                        if (b == a) continue;
                        const D4 xb = s->x_sorted[b];
                        const double rb = xb.w;
                        const uint32_t bo = s->idx[b];
                        V3 xab = v3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                        xab = add3(xab, dm);
                        const double len = len3(xab);
                        const double delta = len - (ra + rb);
                        if (delta < 0.0) {
                            // src/contactsearch.cuh:252-265: last empty slot, or the
                            // slot already holding b (first match ends the scan)
                            uint32_t cpos = nc;
                            bool placed = false;
                            for (uint32_t i = 0; i < NC; ++i) {
                                const uint32_t ci = slots[i];
                                if (ci == n) { cpos = i; placed = true; }
                                else if (ci == bo) { cpos = i; placed = true; break; }
                            }
                            if (!placed) { s->flags |= ORC_F_CONTACT_OVERFLOW; stop = true; break; }
                            slots[cpos] = bo;
                            dms[cpos] = D4{dm.x, dm.y, dm.z, rb};
                            ++nc;
                            if (nc > NC) { s->flags |= ORC_F_CONTACT_OVERFLOW; stop = true; }
                        }
                    }
                }
    }
}

// -------------------------------------------------------- contact models
namespace {

struct Acc { V3 F, T; double es_dot, ev_dot, p; };

// src/contactmodels.cuh:283-493 contactLinear.
void contact_linear(orc_t* s, Acc& acc, const D4& vel_a, V3 angvel_a, const D4& vel_b4,
                    V3 angvel_b, double radius_a, double radius_b, V3 x, double x_length,
                    double delta, V3& delta_t_io) {
    const V3 delta_t0_uncor = delta_t_io;
    const V3 n = divs(x, x_length);
    const V3 vel_linear = v3(vel_a.x - vel_b4.x, vel_a.y - vel_b4.y, vel_a.z - vel_b4.z);
    V3 vel = add3(add3(vel_linear, smul(radius_a + delta / 2.0, cross3(n, angvel_a))),
                  smul(radius_b + delta / 2.0, cross3(n, angvel_b)));
    const double vel_n = -dot3(vel_linear, n);
    const V3 vel_t = sub3(vel, muls(n, dot3(n, vel)));
    const double vel_t_length = len3(vel_t);
    const V3 delta_t0 = sub3(delta_t0_uncor, muls(n, dot3(n, delta_t0_uncor)));
    const double delta_t0_length = len3(delta_t0);
    V3 delta_t = delta_t0;  // definition for the untouched case (header note)
    V3 f_n = smul(-s->k_n * delta - s->gamma_n * vel_n, n);
    acc.ev_dot += s->gamma_n * vel_n * vel_n;
    if (dot3(f_n, n) < 0.0) f_n = v3(0.0, 0.0, 0.0);
    const double f_n_length = len3(f_n);
    const V3 f_c = smul(-s->kappa * std::sqrt(radius_a * radius_b), n);
    V3 f_t = v3(0.0, 0.0, 0.0);
    if (delta_t0_length > 0.0 || vel_t_length > 0.0) {
        delta_t = add3(delta_t0, muls(vel_t, s->dt));
        const V3 f_t_elast = smul(-s->k_t, delta_t);
        const V3 f_t_visc = smul(-s->gamma_t, vel_t);
        f_t = add3(f_t_elast, f_t_visc);
        const double f_t_length = len3(f_t);
        const double f_t_limit = s->mu_s * len3(sub3(f_n, f_c));
        delta_t = add3(delta_t0, muls(vel_t, s->dt));
        if (f_t_length > f_t_limit) {
            const V3 t = divs(f_t, len3(f_t));
            f_t = smul(f_t_limit, t);
            delta_t = smul(-1.0 / s->k_t,
                           add3(smul(s->mu_d * len3(sub3(f_n, f_c)), t), smul(s->gamma_t, vel_t)));
            acc.es_dot += len3(muls(smul(len3(f_t), vel_t), s->dt)) / s->dt;
        }
    }
    acc.F = add3(acc.F, add3(add3(f_n, f_t), f_c));
    acc.T = add3(acc.T, cross3(smul(-(radius_a + delta * 0.5), n), f_t));
    acc.p += f_n_length / (4.0 * PI * radius_a * radius_a);
    delta_t_io = delta_t;
}

// src/contactmodels.cuh:499-666 contactHertz, contact model 3, restated for a
// positive overlap xi = -delta.  The reference passes the (negative) overlap
// delta to powf(delta, 3/2), powf(delta, 1/4) and powf(delta, 1/2)
// (src/contactsearch.cuh:517-529), which gives NaN forces; here each power
// is taken of xi with the sign the linear law's terms carry: the elastic
// normal term -k_n delta becomes k_n xi^(3/2), and the viscous and tangential
// coefficients gamma_n xi^(1/4), k_t xi^(1/2), gamma_t xi^(1/4).  The powers
// are evaluated as xi sqrt(xi), sqrt(sqrt(xi)) and sqrt(xi) (correctly
// rounded on both sides).  The rolling resistance (:642-653) opposes the
// relative spin with the smaller of gamma_r r_a |w_ab| and mu_r r_a |f_n|.
void contact_hertz(orc_t* s, Acc& acc, const D4& vel_a, V3 angvel_a, const D4& vel_b4, V3 angvel_b,
                   double radius_a, double radius_b, V3 x, double x_length, double delta, V3& delta_t_io) {
    const double xi = -delta;
    const double sq = std::sqrt(xi);
    const double xi32 = ORC_MUTANT == 10 ? xi : xi * sq;
    const double xi14 = std::sqrt(sq), xi12 = ORC_MUTANT == 14 ? 1.0 : sq;
    const V3 delta_t0_uncor = delta_t_io;
    const V3 n = divs(x, x_length);
    const V3 vel_linear = v3(vel_a.x - vel_b4.x, vel_a.y - vel_b4.y, vel_a.z - vel_b4.z);
    const V3 vel = add3(add3(vel_linear, smul(radius_a + delta / 2.0, cross3(n, angvel_a))),
                        smul(radius_b + delta / 2.0, cross3(n, angvel_b)));
    const V3 angvel_ab = sub3(angvel_a, angvel_b);
    const double angvel_ab_length = len3(angvel_ab);
    const double vel_n = dot3(vel_linear, n);
    const V3 vel_t = sub3(vel, muls(n, dot3(vel, n)));
    const double vel_t_length = len3(vel_t);
    const V3 delta_t0 = sub3(delta_t0_uncor, muls(n, dot3(delta_t0_uncor, n)));
    const double delta_t0_length = len3(delta_t0);
    V3 delta_t = delta_t0;  // the reference leaves it unset when neither branch runs
    V3 f_n = smul(s->k_n * xi32 - s->gamma_n * (ORC_MUTANT == 13 ? 1.0 : xi14) * vel_n, n);
    acc.ev_dot += s->gamma_n * vel_n * vel_n;
    if (dot3(f_n, n) < 0.0) f_n = v3(0.0, 0.0, 0.0);
    const double f_n_length = len3(f_n);
    const V3 f_c = smul(-s->kappa * std::sqrt(radius_a * radius_b), n);
    V3 f_t = v3(0.0, 0.0, 0.0), t_res = v3(0.0, 0.0, 0.0);
    if (delta_t0_length > 0.0 || vel_t_length > 0.0) {
        const V3 f_t_elast = smul(-s->k_t * xi12, delta_t0);
        const V3 f_t_visc = smul(-s->gamma_t * xi14, vel_t);
        const double f_t_limit = vel_t_length > (double)0.001f ? s->mu_d * len3(sub3(f_n, f_c))
                                                               : s->mu_s * len3(sub3(f_n, f_c));
        f_t = add3(f_t_elast, f_t_visc);
        const double f_t_length = len3(f_t);
        if (f_t_length > f_t_limit) {
            f_t = muls(f_t, f_t_limit / f_t_length);
            delta_t = divs(add3(f_t, smul(s->gamma_t * xi14, vel_t)), -s->k_t * xi12);
            acc.es_dot += len3(sub3(delta_t0, delta_t)) * s->k_t / s->dt;
        } else {
            delta_t = add3(delta_t0, muls(vel_t, s->dt));
        }
    }
    if (angvel_ab_length > 0.0) {
        const double cap = ORC_MUTANT == 12 ? s->gamma_r * radius_a * angvel_ab_length
                                            : std::fmin(s->gamma_r * radius_a * angvel_ab_length,
                                                        s->mu_r * radius_a * f_n_length);
        t_res = divs(smul((ORC_MUTANT == 11 ? 1.0 : -1.0) * cap, angvel_ab), angvel_ab_length);
    }
    acc.F = add3(acc.F, add3(add3(f_n, f_t), f_c));
    acc.T = add3(acc.T, add3(smul(-(radius_a + delta / 2.0), cross3(n, f_t)), t_res));
    acc.p += f_n_length / (4.0 * PI * radius_a * radius_a);
    delta_t_io = delta_t;
}

// src/contactmodels.cuh:10-109 contactLinear_wall.
double contact_linear_wall(orc_t* s, Acc& acc, const D4& linvel4, V3 angvel, double radius_a,
                           V3 n, double delta, double wvel) {
    const V3 linvel = v3(linvel4.x, linvel4.y, linvel4.z);
    // operator+(Float3, Float) adds the scalar wall velocity to every component
    const V3 vel0 = add3(linvel, smul(radius_a, cross3(n, angvel)));
    V3 vel = v3(vel0.x + wvel, vel0.y + wvel, vel0.z + wvel);
    if (ORC_MUTANT == 4) vel = v3(vel0.x, vel0.y, vel0.z + wvel);
    const double vel_n = dot3(vel, n);
    const V3 vel_t = sub3(vel, muls(n, dot3(vel, n)));
    const double vel_t_length = len3(vel_t);
    V3 f_n = smul(-s->k_n * delta - s->gamma_wn * vel_n, n);
    if (dot3(f_n, n) < 0.0) f_n = v3(0.0, 0.0, 0.0);
    const double f_n_length = len3(f_n);
    V3 f_t = v3(0.0, 0.0, 0.0);
    const V3 T_res = v3(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0) {
        const double f_t_visc = s->gamma_wt * vel_t_length;
        double f_t_limit;
        const bool dynamic = ORC_MUTANT == 2 ? vel_t_length < (double)0.001f : vel_t_length > (double)0.001f;
        if (dynamic) f_t_limit = s->mu_wd * f_n_length;
        else f_t_limit = s->mu_ws * f_n_length;
        if (f_t_visc < f_t_limit) f_t = divs(smul((ORC_MUTANT == 1 ? 1.0 : -1.0) * f_t_visc, vel_t), vel_t_length);
        else f_t = divs(smul((ORC_MUTANT == 5 ? 1.0 : -1.0) * f_t_limit, vel_t), vel_t_length);
    }
    acc.F = add3(acc.F, add3(f_n, f_t));
    acc.T = add3(acc.T, add3(smul(ORC_MUTANT == 3 ? radius_a : -radius_a, cross3(n, f_t)), T_res));
    acc.p += f_n_length / (4.0 * PI * radius_a * radius_a);
    return dot3(f_n, n);
}

}  // namespace

// The reference's expf of a double argument, restated as a double-precision
// exponential of fixed steps: Cody-Waite reduction x = k ln2 + r with
// |r| <= ln2/2 (ln2 split into a 32-bit-exact high part and a low part), the
// Taylor series of e^r to degree 13 by Horner's rule (truncation error below
// 1e-17 relative on that range), and the scaling by 2^k.
double det_exp(double x) {
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double k = std::nearbyint(x * inv_ln2);
    const double r = (x - k * ln2_hi) - k * ln2_lo;
    double q = 1.0 / 6227020800.0;                      // 1/13!
    const double c[13] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
                          1.0 / 40320.0, 1.0 / 5040.0, 1.0 / 720.0, 1.0 / 120.0, 1.0 / 24.0,
                          1.0 / 6.0, 0.5, 1.0, 1.0};  // 1/12! .. 1/0!
    for (int i = 0; i < 13; ++i) q = q * r + c[i];
    return std::ldexp(q, (int)k);
}

namespace {

// src/contactmodels.cuh:114-253 contactLinearViscous: no spring history, no
// torque (commented out in the reference), viscous tangential force capped
// by Coulomb friction.
void contact_viscous(orc_t* s, Acc& acc, const D4& vel_a, V3 angvel_a, const D4& vel_b, V3 angvel_b,
                     double radius_a, double radius_b, V3 x, double x_length, double delta) {
    const V3 n = divs(x, x_length);
    const V3 vel_linear = v3(vel_a.x - vel_b.x, vel_a.y - vel_b.y, vel_a.z - vel_b.z);
    const V3 vel = add3(add3(vel_linear, smul(radius_a + delta / 2.0, cross3(n, angvel_a))),
                        smul(radius_b + delta / 2.0, cross3(n, angvel_b)));
    const double vel_n = dot3(vel_linear, n);
    const V3 vel_t = sub3(vel, muls(n, dot3(vel, n)));
    const double vel_t_length = len3(vel_t);
    V3 f_n = smul(-s->k_n * delta - s->gamma_n * vel_n, n);
    if (dot3(f_n, n) < 0.0) f_n = v3(0.0, 0.0, 0.0);
    const double f_n_length = len3(f_n);
    const V3 f_c = smul(-s->kappa * std::sqrt(radius_a * radius_b), n);
    V3 f_t = v3(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0) {
        const double f_t_visc = s->gamma_t * vel_t_length;
        const double f_t_limit = vel_t_length > (double)0.001f ? s->mu_d * len3(sub3(f_n, f_c))
                                                               : s->mu_s * len3(sub3(f_n, f_c));
        if (f_t_visc < f_t_limit) f_t = divs(smul(-f_t_visc, vel_t), vel_t_length);
        else f_t = divs(smul(-f_t_limit, vel_t), vel_t_length);
    }
    acc.F = add3(acc.F, add3(add3(f_n, f_t), f_c));
    acc.p += f_n_length / (4.0 * PI * radius_a * radius_a);
}

// src/cohesion.cuh:314-350 capillaryCohesion_exp (as written: the decay
// length lambda = 0.9 h sqrt(V_b / R_har) has h = -sqrt(r) < 0).
void capillary(orc_t* s, Acc& acc, double radius_a, double radius_b, double delta, V3 x, double x_length) {
    const V3 n = divs(x, x_length);
    const double r = (radius_a / radius_b) > (radius_b / radius_a) ? radius_a / radius_b : radius_b / radius_a;
    const double h = -std::sqrt(r);
    const double R_har = (2.0 * radius_a * radius_b) / (radius_a + radius_b);
    const double R_geo = std::sqrt(radius_a * radius_b);
    const double lambda = (double)0.9f * h * std::sqrt(s->V_b / R_har);
    acc.F = add3(acc.F, smul(-s->kappa * R_geo * det_exp(-delta / lambda), n));
}

// src/contactsearch.cuh:76-173 findAndProcessContactsInCell over the 27
// cells (:548-576): contact model 1 finds and processes contacts at once.
void interact_viscous(orc_t* s, uint32_t a, Acc& acc) {
    const D4 xa = s->x_sorted[a];
    const double ra = xa.w;
    const int gx = cell_coord(s, xa.x, 0), gy = cell_coord(s, xa.y, 1), gz = cell_coord(s, xa.z, 2);
    for (int dz = -1; dz < 2; ++dz)
        for (int dy = -1; dy < 2; ++dy)
            for (int dx = -1; dx < 2; ++dx) {
                int tx = gx + dx, ty = gy + dy;
                const int tz = gz + dz;
                V3 dm = v3(0.0, 0.0, 0.0);
                if (find_dist_mod(s, tx, ty, tz, dm) == -1) continue;
                const uint32_t c = (uint32_t)tx + (uint32_t)ty * s->num[0] + (s->num[0] * s->num[1]) * (uint32_t)tz;
                const uint32_t st = s->cellStart[c];
                if (st == EMPTY) continue;
                const uint32_t en = s->cellEnd[c];
                for (uint32_t b = st; b < en; ++b) {
                    if (b == a) continue;
                    const D4 xb = s->x_sorted[b];
                    const double rb = xb.w;
                    V3 xab = v3(xa.x - xb.x, xa.y - xb.y, xa.z - xb.z);
                    xab = add3(xab, dm);
                    const double len = len3(xab);
                    const double delta = len - (ra + rb);
                    if (delta < 0.0)
                        contact_viscous(s, acc, s->vel_sorted[a], s->angvel_sorted[a], s->vel_sorted[b],
                                        s->angvel_sorted[b], ra, rb, xab, len, delta);
                    else if (delta < s->db)
                        capillary(s, acc, ra, rb, delta, xab, len);
                }
            }
}

}  // namespace

// src/contactsearch.cuh:369-689 interact (contact models 1 and 2).
static void stage_interact(orc_t* s) {
    const uint32_t NC = (uint32_t)s->nc;
    const uint32_t n = s->np;
    const double Lx = s->L[0], Ly = s->L[1], Lz = s->L[2];
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t ao = s->idx[a];
        const D4 xa = s->x_sorted[a];
        const double ra = xa.w;
        D4 wnx[5] = {{0.0, 0.0, -1.0, Lz}, {-1.0, 0.0, 0.0, Lx}, {1.0, 0.0, 0.0, 0.0},
                     {0.0, -1.0, 0.0, Ly}, {0.0, 1.0, 0.0, 0.0}};
        D4 wmv[5] = {{0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}};
        for (uint32_t w = 0; w < s->nw && w < 5; ++w) { wnx[w] = s->wnx[w]; wmv[w] = s->wmvfd[w]; }
        Acc acc;
        acc.F = v3(0.0, 0.0, 0.0); acc.T = v3(0.0, 0.0, 0.0);
        acc.es_dot = 0.0; acc.ev_dot = 0.0; acc.p = 0.0;
        const D4 vel_a = s->vel_sorted[a];
        const V3 angvel_a = s->angvel_sorted[a];
        if (s->contactmodel == 1) interact_viscous(s, a, acc);
        for (uint32_t i = 0; i < NC && s->contactmodel != 1; ++i) {
            const size_t mp = (size_t)ao * NC + i;
            const uint32_t bo = s->contacts[mp];
            if (bo != n) {
                const D4 dm = s->distmod[mp];
                const D4 xb = s->x[bo];
                const double rb = xb.w;
                const V3 xab = v3(xa.x - xb.x + dm.x, xa.y - xb.y + dm.y, xa.z - xb.z + dm.z);
                const double len = len3(xab);
                const double delta_n = len - (ra + rb);
                if (delta_n < 0.0 && s->contactmodel == 3) {
                    contact_hertz(s, acc, vel_a, angvel_a, s->vel[bo], s->angvel[bo], ra, rb, xab, len,
                                  delta_n, s->delta_t[mp]);
                } else if (delta_n < 0.0) {
                    contact_linear(s, acc, vel_a, angvel_a, s->vel[bo], s->angvel[bo], ra, rb, xab,
                                   len, delta_n, s->delta_t[mp]);
                } else {
                    s->contacts[mp] = n;
                    s->delta_t[mp] = v3(0.0, 0.0, 0.0);
                }
            } else {
                s->delta_t[mp] = v3(0.0, 0.0, 0.0);
            }
        }
        double wf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        double dw;
        // upper wall (idx 0)
        dw = wnx[0].w - (xa.z + ra);
        if (dw < 0.0)
            wf[0] = contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(wnx[0].x, wnx[0].y, wnx[0].z), dw, wmv[0].y);
        // lower wall at origo z
        dw = xa.z - ra - s->origo[2];
        if (dw < 0.0)
            (void)contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(0.0, 0.0, 1.0), dw, 0.0);
        if (s->periodic == 0) {
            dw = wnx[1].w - (xa.x + ra);
            if (dw < 0.0)
                wf[1] = contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(wnx[1].x, wnx[1].y, wnx[1].z), dw, wmv[1].y);
            dw = xa.x - ra - wnx[2].w;
            if (dw < 0.0)
                wf[2] = contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(wnx[2].x, wnx[2].y, wnx[2].z), dw, wmv[2].y);
        }
        if (s->periodic == 0 || s->periodic == 2) {
            dw = wnx[3].w - (xa.y + ra);
            if (dw < 0.0)
                wf[3] = contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(wnx[3].x, wnx[3].y, wnx[3].z), dw, wmv[3].y);
            dw = xa.y - ra - wnx[4].w;
            if (dw < 0.0)
                wf[4] = contact_linear_wall(s, acc, vel_a, angvel_a, ra, v3(wnx[4].x, wnx[4].y, wnx[4].z), dw, wmv[4].y);
        }
        s->force[ao] = acc.F;
        s->torque[ao] = acc.T;
        s->es_dot[ao] = acc.es_dot;
        s->ev_dot[ao] = acc.ev_dot;
        s->es[ao] += acc.es_dot * s->dt;
        s->ev[ao] += acc.ev_dot * s->dt;
        s->p[ao] = acc.p;
        for (uint32_t w = 0; w < s->nw && w < 5; ++w) s->wall_pp[(size_t)w * n + ao] = wf[w];
    }
}

// -------------------------------------------------------------- bonds
// src/cohesion.cuh:8-233 bondsLinear (forces kept per bond, applied in
// ascending bond index by stage_bond_apply).
static void stage_bonds(orc_t* s) {
    const uint32_t n = s->np;
    const uint32_t broken = s->nb0 >= n ? s->nb0 : n;
    const double dt = s->dt;
    for (uint32_t k = 0; k < s->nb0; ++k) {
        s->bond_live[k] = 0;
        const uint32_t bi = s->bonds[2 * k], bj = s->bonds[2 * k + 1];
        if (!(bi < n)) continue;
        const D4 d0 = s->bdelta[k], o0 = s->bomega[k];
        V3 delta0_t = v3(d0.x, d0.y, d0.z);
        const double delta0_n = d0.w;
        V3 omega0_t = v3(o0.x, o0.y, o0.z);
        const double omega0_n = o0.w;
        const D4 x_i = s->x[bi], x_j = s->x[bj];
        const D4 vel_i = s->vel[bi], vel_j = s->vel[bj];
        const V3 angvel_i = s->angvel[bi], angvel_j = s->angvel[bj];
        const double R_bar = s->lambda_bar * std::fmin(x_i.w, x_j.w);
        const double A = PI * R_bar * R_bar;
        const double I = 0.25 * PI * R_bar * R_bar * R_bar * R_bar;
        const double J = I * 2.0;
        const V3 x = v3(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
        const double x_length = len3(x);
        const double overlap = std::fmin(0.0, x_length - (x_i.w + x_j.w));
        const V3 nrm = divs(x, x_length);
        delta0_t = sub3(delta0_t, muls(nrm, dot3(nrm, delta0_t)));
        const V3 ddelta = muls(add3(add3(v3(vel_i.x - vel_j.x, vel_i.y - vel_j.y, vel_i.z - vel_j.z),
                                         smul(x_i.w + overlap / 2.0, cross3(nrm, angvel_i))),
                                    smul(x_j.w + overlap / 2.0, cross3(nrm, angvel_j))),
                               dt);
        const double ddelta_n = -dot3(ddelta, nrm);
        const double delta_n = delta0_n + ddelta_n;
        const V3 ddelta_t = sub3(ddelta, muls(nrm, dot3(nrm, ddelta)));
        const V3 delta_t = add3(delta0_t, ddelta_t);
        const V3 f_n = smul(s->k_n * A * delta_n + s->gamma_n * ddelta_n / dt, nrm);
        const V3 f_t = sub3(smul(-s->k_t * A, delta_t), divs(smul(s->gamma_t, ddelta_t), dt));
        const V3 f = add3(f_n, f_t);
        const V3 mn = neg3(nrm);
        omega0_t = sub3(omega0_t, muls(mn, dot3(omega0_t, mn)));
        const V3 domega = muls(v3(angvel_j.x - angvel_i.x, angvel_j.y - angvel_i.y, angvel_j.z - angvel_i.z), dt);
        const double domega_n = dot3(domega, mn);
        const double omega_n = omega0_n + domega_n;
        const V3 domega_t = sub3(domega, smul(domega_n, mn));
        const V3 omega_t = add3(omega0_t, domega_t);
        const V3 t_n = smul(-s->k_t * J * omega_n, mn);
        const V3 t_t = smul(-s->k_n * I, omega_t);
        const V3 t_j = add3(t_n, t_t);
        const double sigma_max = len3(f_n) / A + len3(t_t) * R_bar / I;
        const double tau_max = len3(f_t) / A + len3(t_n) * R_bar / J;
        if (sigma_max >= s->sigma_b || tau_max >= s->tau_b) {
            s->bonds[2 * k] = broken;
            continue;
        }
        s->bdelta[k] = D4{delta_t.x, delta_t.y, delta_t.z, delta_n};
        s->bomega[k] = D4{omega_t.x, omega_t.y, omega_t.z, omega_n};
        s->bond_f[k] = f;
        s->bond_t[k] = t_j;
        s->bond_live[k] = 1;
    }
}

// src/cohesion.cuh:220-226 in ascending bond index.
static void stage_bond_apply(orc_t* s) {
    for (uint32_t k = 0; k < s->nb0; ++k) {
        if (!s->bond_live[k]) continue;
        const uint32_t bi = s->bonds[2 * k], bj = s->bonds[2 * k + 1];
        const V3 f = s->bond_f[k], t = s->bond_t[k];
        s->force[bi] = v3(s->force[bi].x + f.x, s->force[bi].y + f.y, s->force[bi].z + f.z);
        s->force[bj] = v3(s->force[bj].x - f.x, s->force[bj].y - f.y, s->force[bj].z - f.z);
        s->torque[bi] = v3(s->torque[bi].x - t.x, s->torque[bi].y - t.y, s->torque[bi].z - t.z);
        s->torque[bj] = v3(s->torque[bj].x + t.x, s->torque[bj].y + t.y, s->torque[bj].z + t.z);
    }
}

// ------------------------------------------------------------ integrate
// src/integration.cuh:9-187.
static void stage_integrate(orc_t* s) {
    const double dt = s->dt;
    for (uint32_t i = 0; i < s->np; ++i) {
        const uint32_t o = s->idx[i];
        const V3 force = s->force[o], torque = s->torque[o];
        V3 angpos = s->angpos[o], vel0 = s->vel0[o], angvel0 = s->angvel0[o];
        D4 x = s->x_sorted[i];
        D4 vel = s->vel_sorted[i];
        V3 angvel = s->angvel_sorted[i];
        const double radius = x.w;
        D2 xysum = D2{0.0, 0.0};
        const double m = 4.0 / 3.0 * PI * radius * radius * radius * s->rho;
        V3 acc = v3(force.x / m, force.y / m, force.z / m);
        V3 angacc = v3(torque.x * 1.0 / (2.0 / 5.0 * m * radius * radius),
                       torque.y * 1.0 / (2.0 / 5.0 * m * radius * radius),
                       torque.z * 1.0 / (2.0 / 5.0 * m * radius * radius));
        acc.x += s->g[0]; acc.y += s->g[1]; acc.z += s->g[2];
        if (vel.w > 0.0 && ORC_MUTANT != 6) {
            acc.x = 0.0; acc.y = 0.0;
            if (ORC_MUTANT != 7) acc.z -= s->g[2];
            if (vel.w >= ORC_FIXVEL_HOLD_Z && ORC_MUTANT != 8) acc.z = 0.0;  // extension (header note)
            if (ORC_MUTANT != 9) angacc = v3(0.0, 0.0, 0.0);
        }
        wrap_position(s, x);
        vel0.x += acc.x * dt; vel0.y += acc.y * dt; vel0.z += acc.z * dt;
        angvel0.x += angacc.x * dt; angvel0.y += angacc.y * dt; angvel0.z += angacc.z * dt;
        x.x += vel0.x * dt; x.y += vel0.y * dt; x.z += vel0.z * dt;
        angpos.x += angvel0.x * dt; angpos.y += angvel0.y * dt; angpos.z += angvel0.z * dt;
        vel.x = vel0.x + 0.5 * acc.x * dt;
        vel.y = vel0.y + 0.5 * acc.y * dt;
        vel.z = vel0.z + 0.5 * acc.z * dt;
        angvel.x = angvel0.x + 0.5 * angacc.x * dt;
        angvel.y = angvel0.y + 0.5 * angacc.y * dt;
        angvel.z = angvel0.z + 0.5 * angacc.z * dt;
        xysum.x += vel.x * dt;
        xysum.y += vel.y * dt;
        s->xysum[o].x += xysum.x;
        s->xysum[o].y += xysum.y;
        s->angvel[o] = angvel;
        s->angvel0[o] = angvel0;
        s->vel[o] = vel;
        s->vel0[o] = vel0;
        s->angpos[o] = angpos;
        s->x[o] = x;
    }
}

// src/integration.cuh:191-221 summation: one block of 256 threads per 256
// original indices, tree halving 128..1, one partial per block; per wall.
static void stage_summation(orc_t* s) {
    const uint32_t nb = s->blocks();
    double cache[256];
    for (uint32_t w = 0; w < s->nw; ++w) {
        const double* in = &s->wall_pp[(size_t)w * s->np];
        for (uint32_t b = 0; b < nb; ++b) {
            for (uint32_t t = 0; t < 256; ++t) {
                const uint32_t i = b * 256 + t;
                double temp = 0.0;
                if (i < s->np) temp += in[i];
                cache[t] = temp;
            }
            for (uint32_t h = 128; h != 0; h /= 2)
                for (uint32_t t = 0; t < h; ++t) cache[t] += cache[t + h];
            s->wall_partial[(size_t)w * nb + b] = cache[0];
        }
    }
}

// src/integration.cuh:224-295 integrateWalls.
static void stage_walls(orc_t* s) {
    const uint32_t nb = s->blocks();
    const double dt = s->dt;
    for (uint32_t w = 0; w < s->nw; ++w) {
        D4 w_nx = s->wnx[w], w_mvfd = s->wmvfd[w];
        const int wmode = s->wmode[w];
        double vel0 = s->wvel0[w];
        if (wmode == 0) continue;
        w_mvfd.z = 0.0;
        for (uint32_t i = 0; i < nb; ++i) w_mvfd.z += s->wall_partial[(size_t)w * nb + i];
        const double sigma_0 = w_mvfd.w + s->devs_A * std::sin(2.0 * 3.141596654 * s->devs_f * s->t_cur);
        const double N = -sigma_0 * s->L[0] * s->L[1];
        double acc = (w_mvfd.z + N) / w_mvfd.x;
        if (wmode == 2) acc = 0.0;
        vel0 += acc * dt;
        w_nx.w += vel0 * dt;
        w_mvfd.y = vel0 + 0.5 * acc * dt;
        s->wnx[w] = w_nx;
        s->wmvfd[w] = w_mvfd;
        s->wvel0[w] = vel0;
    }
}

extern "C" int orc_stage(orc_t* s, int stage) {
    switch (stage) {
        case ORC_STAGE_BONDS: stage_bonds(s); break;
        case ORC_STAGE_CELLID: stage_cellid(s); break;
        case ORC_STAGE_SORT: stage_sort(s); break;
        case ORC_STAGE_REORDER: stage_reorder(s); break;
        case ORC_STAGE_TOPOLOGY: stage_topology(s); break;
        case ORC_STAGE_INTERACT: stage_interact(s); break;
        case ORC_STAGE_BOND_APPLY: stage_bond_apply(s); break;
        case ORC_STAGE_INTEGRATE: stage_integrate(s); break;
        case ORC_STAGE_SUMMATION: stage_summation(s); break;
        case ORC_STAGE_WALLS: stage_walls(s); break;
        case ORC_STAGE_TIME: s->t_cur += s->dt; break;
        default: return -1;
    }
    return (int)s->flags;
}

// One step in the order of src/device.cu:747-1111 (bondsLinear first, see
// header note).
extern "C" int orc_step(orc_t* s, uint64_t nsteps) {
    for (uint64_t k = 0; k < nsteps; ++k) {
        for (int st = 0; st < ORC_NSTAGES; ++st) orc_stage(s, st);
        if (s->flags) break;
    }
    return (int)s->flags;
}

// ------------------------------------------------------------ field access
namespace {

template <class T> size_t vbytes(const std::vector<T>& v) { return v.size() * sizeof(T); }

}  // namespace

extern "C" size_t orc_field_bytes(const orc_t* s, int field) {
    const size_t n = s->np, nc = (size_t)s->nc, nb = s->nb0, nw = s->nw;
    switch (field) {
        case ORC_X: case ORC_VEL: case ORC_X_SORTED: return n * 32;
        case ORC_ANGVEL: case ORC_FORCE: case ORC_TORQUE: case ORC_ANGPOS: case ORC_VEL0:
        case ORC_ANGVEL0: return n * 24;
        case ORC_XYSUM: return n * 16;
        case ORC_ES_DOT: case ORC_ES: case ORC_EV_DOT: case ORC_EV: case ORC_P: return n * 8;
        case ORC_CONTACTS: return n * nc * 4;
        case ORC_DISTMOD: return n * nc * 32;
        case ORC_DELTA_T: return n * nc * 24;
        case ORC_CELL_KEYS: case ORC_SORT_IDX: return n * 4;
        case ORC_CELL_START: case ORC_CELL_END: return s->K() * 4;
        case ORC_BONDS: return nb * 8;
        case ORC_BONDS_DELTA: case ORC_BONDS_OMEGA: return nb * 32;
        case ORC_WALLS_NX: case ORC_WALLS_MVFD: return nw * 32;
        case ORC_WALLS_VEL0: return nw * 8;
        case ORC_WALL_PP: return nw * n * 8;
        case ORC_TIME: return 32;
        default: return 0;
    }
}

namespace {

void* field_ptr(orc_t* s, int field) {
    switch (field) {
        case ORC_X: return s->x.data();
        case ORC_VEL: return s->vel.data();
        case ORC_X_SORTED: return s->x_sorted.data();
        case ORC_ANGVEL: return s->angvel.data();
        case ORC_FORCE: return s->force.data();
        case ORC_TORQUE: return s->torque.data();
        case ORC_ANGPOS: return s->angpos.data();
        case ORC_VEL0: return s->vel0.data();
        case ORC_ANGVEL0: return s->angvel0.data();
        case ORC_XYSUM: return s->xysum.data();
        case ORC_ES_DOT: return s->es_dot.data();
        case ORC_ES: return s->es.data();
        case ORC_EV_DOT: return s->ev_dot.data();
        case ORC_EV: return s->ev.data();
        case ORC_P: return s->p.data();
        case ORC_CONTACTS: return s->contacts.data();
        case ORC_DISTMOD: return s->distmod.data();
        case ORC_DELTA_T: return s->delta_t.data();
        case ORC_CELL_KEYS: return s->key.data();
        case ORC_SORT_IDX: return s->idx.data();
        case ORC_CELL_START: return s->cellStart.data();
        case ORC_CELL_END: return s->cellEnd.data();
        case ORC_BONDS: return s->bonds.data();
        case ORC_BONDS_DELTA: return s->bdelta.data();
        case ORC_BONDS_OMEGA: return s->bomega.data();
        case ORC_WALLS_NX: return s->wnx;
        case ORC_WALLS_MVFD: return s->wmvfd;
        case ORC_WALLS_VEL0: return s->wvel0;
        case ORC_WALL_PP: return s->wall_pp.data();
        default: return nullptr;
    }
}

}  // namespace

extern "C" int orc_get(const orc_t* cs, int field, void* dst, size_t bytes) {
    orc_t* s = const_cast<orc_t*>(cs);
    if (field < 0 || field >= ORC_NFIELDS) return -1;
    const size_t need = orc_field_bytes(s, field);
    if (bytes < need) return -2;
    if (field == ORC_TIME) {
        const double t[4] = {s->dt, s->t_cur, s->t_total, s->file_dt};
        std::memcpy(dst, t, 32);
        return 0;
    }
    void* src = field_ptr(s, field);
    if (need) std::memcpy(dst, src, need);
    return 0;
}

extern "C" int orc_set(orc_t* s, int field, const void* src, size_t bytes) {
    const size_t need = orc_field_bytes(s, field);
    if (bytes != need) return -2;
    if (field == ORC_TIME) {
        double t[4];
        std::memcpy(t, src, 32);
        s->dt = t[0]; s->t_cur = t[1]; s->t_total = t[2]; s->file_dt = t[3];
        return 0;
    }
    void* dst = field_ptr(s, field);
    if (!dst) return -1;
    if (need) std::memcpy(dst, src, need);
    return 0;
}

extern "C" void orc_dims(const orc_t* s, uint64_t* out5) {
    out5[0] = s->np; out5[1] = s->nb0; out5[2] = s->nw; out5[3] = (uint64_t)s->nc; out5[4] = s->K();
}

extern "C" void orc_get_params(const orc_t* s, double* o) {
    o[0] = s->k_n; o[1] = s->k_t; o[2] = s->gamma_n; o[3] = s->gamma_t; o[4] = s->mu_s; o[5] = s->mu_d;
}

extern "C" void orc_set_params(orc_t* s, const double* p) {
    s->k_n = p[0]; s->k_t = p[1]; s->gamma_n = p[2]; s->gamma_t = p[3]; s->mu_s = p[4]; s->mu_d = p[5];
}

// ------------------------------------------------------------ host analysis
namespace {

// sphericalCap, src/sphere.cpp:343-347: volume of a cap of height hc.
double cap_volume(double hc, double r) { return M_PI * hc * hc / 3.0 * (3.0 * r - hc); }

// Volume of the sphere (centre zc, radius r) between the planes z = lo and
// z = hi, integrated exactly: pi (r^2 t - t^3 / 3) from t1 to t2.
double slab_volume(double zc, double r, double lo, double hi) {
    const double t1 = std::max(lo - zc, -r), t2 = std::min(hi - zc, r);
    if (!(t2 > t1)) return 0.0;
    return M_PI * (r * r * (t2 - t1) - (t2 * t2 * t2 - t1 * t1 * t1) / 3.0);
}

}  // namespace

// DEM::porosity, src/sphere.cpp:363-483.  rule 0 restates the reference's
// branches as they are (a sphere whose centre lies inside the slice but which
// crosses one of its planes removes V_sphere + cap, src/sphere.cpp:430-432
// and :445-447); rule 1 removes the exact volume of the sphere inside the
// slice.  Slices start at z = 0 (src/sphere.cpp:401) and have the thickness
// (top - origo_z) / slices, top being wall 0's position when there are
// walls, else L_z (src/sphere.cpp:367-374).  rule 0 refuses slices no
// thicker than twice the radius that DEM::r_max returns, which is the last
// particle's (src/sphere.cpp:350-360 and :379-383).  Returns 0, or 1 for
// that refusal, 2 for slices < 1.
extern "C" int orc_porosity(const orc_t* s, int slices, int rule, double* z_pos, double* porosity) {
    if (slices < 1) return 2;
    const double top = s->nw > 0 ? s->wnx[0].w : s->L[2];
    const double h_slice = (top - s->origo[2]) / (double)slices;
    const double r_last = s->np ? s->x[s->np - 1].w : 0.0;
    if (rule == 0 && h_slice <= r_last * 2.0) return 1;
    const double V_slice = h_slice * s->L[0] * s->L[1];
    for (int iz = 0; iz < slices; ++iz) {
        double V_void = V_slice;
        const double lo = iz * h_slice, hi = lo + h_slice;
        for (uint32_t i = 0; i < s->np; ++i) {
            const double zc = s->x[i].z, r = s->x[i].w;
            if (rule != 0) {
                V_void -= slab_volume(zc, r, lo, hi);
                continue;
            }
            const double zl = zc - r, zh = zc + r;
            const double Vs = 4.0 / 3.0 * M_PI * r * r * r;
            if (lo <= zl && zh <= hi) {
                V_void -= Vs;
            } else {
                if (lo >= zc && lo <= zh)
                    V_void -= cap_volume(zh - lo, r);
                else if (lo < zc && lo > zl)
                    V_void -= Vs + cap_volume(lo - zl, r);
                if (hi <= zc && hi >= zl)
                    V_void -= cap_volume(hi - zl, r);
                else if (hi > zc && hi < zh)
                    V_void -= Vs + cap_volume(zh - hi, r);
            }
        }
        z_pos[iz] = lo + 0.5 * h_slice;
        porosity[iz] = V_void / V_slice;
    }
    return 0;
}

// DEM::findOverlaps, src/sphere.cpp:553-604: every pair i < j (original
// indices, i major) whose centre distance is below r_i + r_j, with no
// periodic images; delta_n = |x_j - x_i| - (r_i + r_j) < 0.  Writes at most
// `capacity` pairs and returns the number found.
extern "C" uint64_t orc_find_overlaps(const orc_t* s, uint32_t* ij, double* delta_n, uint64_t capacity) {
    uint64_t n = 0;
    for (uint32_t i = 0; i < s->np; ++i) {
        const D4 a = s->x[i];
        for (uint32_t j = i + 1; j < s->np; ++j) {
            const D4 b = s->x[j];
            const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            const double delta = std::sqrt(dx * dx + dy * dy + dz * dz) - (a.w + b.w);
            if (delta < 0.0) {
                if (n < capacity) {
                    ij[2 * n] = i;
                    ij[2 * n + 1] = j;
                    delta_n[n] = delta;
                }
                ++n;
            }
        }
    }
    return n;
}

// ---------------------------------------------------------------- raytracer
// DEM::render (src/raytracer.cuh:14-553): the camera of cameraInit
// (:281-333) and render (:374-420), the rays of rayInitPerspective (:27-51)
// and the per-pixel loops of rayIntersectSpheres (:55-137) and
// rayIntersectSpheresColormap (:140-278), all in FP32 as there.  The vector
// helpers follow helper_math's host definitions (a third-party header the
// reference includes, src/darcy.cuh:11): component-wise float3 arithmetic,
// length = sqrtf(dot), normalize = v * (1 / sqrtf(dot)), division by a
// scalar per component.
namespace {

struct F3 {
    float x, y, z;
};
F3 f3(float x, float y, float z) { return {x, y, z}; }
F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float flength(F3 a) { return std::sqrt(fdot(a, a)); }
F3 fnormalize(F3 a) { return a * (1.0f / std::sqrt(fdot(a, a))); }
F3 fcross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// colorbar.h:7-22
float cb_red(float r) { return std::fmin(1.0f, 0.209f * r * r * r - 2.49f * r * r + 3.0f * r + 0.0109f); }
float cb_green(float r) { return std::fmin(1.0f, -2.44f * r * r + 2.15f * r + 0.369f); }
float cb_blue(float r) { return std::fmin(1.0f, -2.21f * r * r + 1.61f * r + 0.573f); }

double vnorm(const V3& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

}  // namespace

// rgb: width * height * 3 bytes in the reference's pixel order.  Returns 0,
// or 2 for an unknown method or an empty image.
extern "C" int orc_render(const orc_t* s, int method, float max_val, float lower_cutoff, uint32_t width,
                          uint32_t height, uint8_t* rgb) {
    if (method < 0 || method > 5 || width == 0 || height == 0) return 2;
    const uint32_t np = s->np;
    // camera (render, src/raytracer.cuh:409-418; cameraInit, :281-313)
    const float imgw = (float)(s->L[0] * 1.35f);
    double mx[3] = {0.0, 0.0, 0.0};  // DEM::maxPos, src/sphere.cpp:520-551
    for (uint32_t i = 0; i < np; ++i) {
        mx[0] = std::max(mx[0], s->x[i].x);
        mx[1] = std::max(mx[1], s->x[i].y);
        mx[2] = std::max(mx[2], s->x[i].z);
    }
    const F3 lookat = f3((float)mx[0], (float)mx[1], (float)mx[2]) / 2.0f;
    const F3 eye = f3((float)(s->L[0] * 2.3f), (float)(s->L[1] * -5.0f), (float)(s->L[2] * 1.3f));
    const float hw_ratio = (float)(height / width);
    const float pl = -0.5f * imgw, pr = 0.5f * imgw, pb = -0.5f * imgw * hw_ratio, pt = 0.5f * imgw * hw_ratio;
    const F3 view = eye - lookat;
    const F3 up = f3(0.0f, 0.0f, 1.0f);
    const F3 w = -view / flength(view);
    const F3 u = fcross(up, w) / flength(fcross(up, w));
    const F3 v = fcross(w, u);
    const float dfoc = flength(view) * 0.8f;
    const F3 light = fnormalize(-1.0f * eye * f3(1.0f, 0.2f, 0.6f));
    // per-particle colour values (render, :445-503), in FP64 as there
    std::vector<double> lin(np, 0.0);
    for (uint32_t i = 0; i < np; ++i) {
        switch (method) {
            case 1: lin[i] = s->p[i]; break;
            case 2: lin[i] = std::sqrt(s->vel[i].x * s->vel[i].x + s->vel[i].y * s->vel[i].y + s->vel[i].z * s->vel[i].z); break;
            case 3: lin[i] = vnorm(s->angvel[i]); break;
            case 4: lin[i] = s->xysum[i].x; break;
            case 5: lin[i] = vnorm(s->angpos[i]); break;
            default: break;
        }
    }
    const uint64_t pixels = (uint64_t)width * height;
    for (uint64_t m = 0; m < pixels; ++m) {
        const uint32_t mempos = (uint32_t)m;
        const uint32_t i = mempos % width;
        const uint32_t j = (uint32_t)(int)std::floor((float)mempos / width) % width;
        const float p_u = pl + (pr - pl) * (i + 0.5f) / width;
        const float p_v = pb + (pt - pb) * (j + 0.5f) / height;
        const F3 e = eye;
        const F3 d = -dfoc * w + p_u * u + p_v * v;
        float tdist = 1e10f;
        F3 n = f3(0, 0, 0);
        uint32_t hit = 0;
        for (uint32_t k = 0; k < np; ++k) {
            const F3 c = f3((float)s->x[k].x, (float)s->x[k].y, (float)s->x[k].z);
            const float R = (float)s->x[k].w;
            const float Delta = (2.0f * fdot(d, (e - c))) * (2.0f * fdot(d, (e - c))) -
                                4.0f * fdot(d, d) * (fdot((e - c), (e - c)) - R * R);
            if (method != 0 && lower_cutoff > 0.0) {
                const float val = (float)lin[k];
                const float fixvel = (float)s->vel[k].w;
                if (!(Delta > 0.0f && val > lower_cutoff && fixvel == 0.f)) continue;
            } else if (!(Delta > 0.0f)) {
                continue;
            }
            const float t_minus = ((fdot(-d, (e - c)) -
                                    std::sqrt(fdot(d, (e - c)) * fdot(d, (e - c)) -
                                              fdot(d, d) * (fdot((e - c), (e - c)) - R * R))) /
                                   fdot(d, d));
            if (std::fabs(t_minus) < tdist) {
                const F3 p = e + t_minus * d;
                tdist = std::fabs(t_minus);
                n = fnormalize(2.0f * (p - c));
                hit = k;
            }
        }
        uint8_t* px = rgb + 3 * m;
        px[0] = px[1] = px[2] = 255;  // imageInit, :14-24
        if (!(tdist < 1e10f)) continue;
        const float dotprod = std::fmax(0.0f, fdot(n, light));
        const float I_d = 40.0f, k_d = 5.0f, k_a = 10.0f, I_a = 5.0f;
        if (method == 0) {
            px[0] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.48f);
            px[1] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.41f);
            px[2] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.27f);
            continue;
        }
        const float fixvel = (float)s->vel[hit].w;
        float ratio = (float)(lin[hit] / max_val);
        if (ratio < 0.01f) ratio = 0.01f;
        if (ratio > 0.99f) ratio = 0.99f;
        float redv = cb_red(ratio), greenv = cb_green(ratio), bluev = cb_blue(ratio);
        if (fixvel > 0.f) redv = greenv = bluev = 0.5f;
        px[0] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * redv);
        px[1] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * greenv);
        px[2] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * bluev);
    }
    return 0;
}
