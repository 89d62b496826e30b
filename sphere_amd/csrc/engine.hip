// engine.hip -- host engine behind the C-ABI of include/sphere_dem.h.
//
// Owns the device buffers of one simulation, enqueues the step kernels on
// its own stream (captured once into a HIP graph and replayed), converts
// between the device layout and the reference's array layout, and
// reproduces DEM::startTime's output cadence (src/device.cu:576-1299).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "binio.h"
#include "dem_internal.h"
#include "engine_internal.h"

using sph::DemDev;
using sph::HostState;
using sph::StepHooks;


namespace {

thread_local std::string g_open_error;

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return SPH_E_HIP;                                                      \
        }                                                                          \
    } while (0)

template <class T> int dalloc(sph_dem* h, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " bytes): " + hipGetErrorString(e);
        return SPH_E_HIP;
    }
    h->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return SPH_OK;
}

#define ALLOC(ptr, count)                              \
    do {                                               \
        int rc_ = dalloc(h, &(ptr), (count));          \
        if (rc_) return rc_;                           \
    } while (0)

void destroy_graph(sph_dem* h) {
    for (int l = 0; l < 2; ++l)
        for (int k = 0; k < 2; ++k) {
            if (h->gexec[l][k]) { (void)hipGraphExecDestroy(h->gexec[l][k]); h->gexec[l][k] = nullptr; }
            if (h->graph[l][k]) { (void)hipGraphDestroy(h->graph[l][k]); h->graph[l][k] = nullptr; }
        }
}

sph::OwnSet own_set(const sph::PState& p) {
    return {p.x, p.vel, p.angvel, p.pr0, p.pr1, p.pr2, p.id, p.cmask, p.lidx, p.vf, p.wf, p.cslot, p.cdt, p.cdt32};
}

// the buffer set holding the current private records
const sph::PState& own(const sph_dem* h) { return h->own_cur ? h->d.a : h->d.b; }

// A range of particles: original indices [base, base + n), which are also
// the storage positions they occupy -- all of a handle's particles, or one
// ensemble member's (the sort keeps each member inside its own range).
struct Range {
    size_t base, n;
};
Range whole(const sph_dem* h) { return {0, h->hs.np}; }

// Margin of the search's FP32 distance prefilter: every coordinate is below
// M = max |origo| + 2 L in magnitude, so converting, subtracting and shifting
// by a periodic image errs by at most 3.5 ulp(M) per axis and 6.1 ulp(M) in
// the distance; 16 ulp(M) plus 1e-6 of the largest radius sum also covers
// the FP32 rounding of the squares.
// The systems of a handle: its own, or an ensemble's members.
std::vector<const HostState*> systems_of(const sph_dem* h) {
    std::vector<const HostState*> v;
    if (h->members.empty()) v.push_back(&h->hs);
    for (const HostState& m : h->members) v.push_back(&m);
    return v;
}

double max_radius(const HostState& s) {
    double rmax = 0.0;
    for (uint32_t i = 0; i < s.np; ++i) rmax = std::max(rmax, s.x[4 * (size_t)i + 3]);
    return rmax;
}

void update_search_eps(sph_dem* h) {
    double M = 0.0, rmax = 0.0;
    for (const HostState* m : systems_of(h)) {
        for (int a = 0; a < 3; ++a) M = std::max(M, std::fabs(m->origo[a]) + 2.0 * std::fabs(m->L[a]));
        rmax = std::max(rmax, max_radius(*m));
    }
    const double ulp = std::ldexp(1.0, std::ilogb(std::max(M, 1e-30)) - 23);
    h->d.search_eps = (float)(16.0 * ulp + 2e-6 * rmax);
    h->d.rmax_f = (float)(rmax * (1.0 + 1e-6));
    // neighbour-list reach beyond r_a + r_b: a tenth of the smallest cell
    // (at least 2 r_max / 10), so lists hold the touching and nearly touching
    // spheres; a rebuild is due once a sphere has moved 0.45 skin.  Two
    // spheres more than a cell apart at the build (outside its 27-cell scan)
    // can then meet before the next one if the cell leaves less room than
    // 0.9 skin beyond the largest pair: the build then also scans the cells
    // two away (list_shell), or, where a periodic axis has fewer than 5 cells,
    // the rebuild distance shrinks to half that room less the prefilter
    // margin (0 for a cell no wider than the largest pair: rebuild on any
    // motion).  (An ensemble takes the smallest cell and least room of its
    // members.)
    double cmin = HUGE_VAL, room = HUGE_VAL;
    bool shell_ok = true;
    for (const HostState* m : systems_of(h)) {
        double c = m->L[0] / m->num[0];
        for (int a = 1; a < 3; ++a) c = std::min(c, m->L[a] / m->num[a]);
        cmin = std::min(cmin, c);
        room = std::min(room, c - 2.0 * (h->members.empty() ? rmax : max_radius(*m)));
        if ((m->periodic == 1 || m->periodic == 2) && m->num[0] < 5) shell_ok = false;
        if (m->periodic == 1 && m->num[1] < 5) shell_ok = false;
    }
    h->d.list_skin = (float)(0.1 * cmin);
    // a sphere's FP32 position lies in its cell's box widened by the FP32
    // rounding (below search_eps) and the cell-id rounding (relative 1e-12)
    h->d.cell_margin = (float)(1e-3 * cmin + 4.0 * (double)h->d.search_eps);
    const double eps = (double)h->d.search_eps, skin = (double)h->d.list_skin;
    const bool tight = room - 2.0 * eps < 0.9 * skin;
    h->d.list_shell = tight && shell_ok ? 1u : 0u;
    const double lim = tight && !shell_ok ? std::min(0.45 * skin, 0.5 * room - eps) : 0.45 * skin;
    h->d.list_lim = lim > 0.0 ? (float)lim : 0.0f;
}

// Box, grid, parameters and walls of system s as the kernels read them; the
// caller places it (cbase, pbase, wbase, bbase).
sph::Sys make_sys(const HostState& s) {
    sph::Sys S{};
    for (int a = 0; a < 3; ++a) {
        S.num[a] = s.num[a]; S.origo[a] = s.origo[a]; S.L[a] = s.L[a];
        S.cell_len[a] = s.L[a] / (double)s.num[a];
        S.g[a] = s.g[a];
    }
    S.div_x = sph::make_fastdiv(s.num[0]);
    S.div_xy = sph::make_fastdiv(s.num[0] * s.num[1]);
    S.periodic = s.periodic;
    S.Lf[0] = (float)s.L[0];
    S.Lf[1] = (float)s.L[1];
    S.K = s.num[0] * s.num[1] * s.num[2];
    S.np = s.np;
    S.nw = s.nw;
    S.nblocks = (s.np + 255) / 256;
    S.dt = s.dt;
    S.k_n = s.k_n; S.k_t = s.k_t; S.gamma_n = s.gamma_n; S.gamma_t = s.gamma_t;
    S.mu_s = s.mu_s; S.mu_d = s.mu_d; S.gamma_wn = s.gamma_wn; S.gamma_wt = s.gamma_wt;
    S.mu_ws = s.mu_ws; S.mu_wd = s.mu_wd; S.rho = s.rho; S.kappa = s.kappa;
    S.gamma_r = s.gamma_r; S.mu_r = s.mu_r;
    S.devs_A = s.devs_A; S.devs_f = s.devs_f;
    S.contactmodel = s.contactmodel;
    S.db = s.db; S.V_b = s.V_b;
    return S;
}

void fill_params(sph_dem* h) {
    const HostState& s = h->hs;
    DemDev& d = h->d;
    d.nb0 = s.nb0;
    d.bond_broken = s.nb0 >= s.np ? s.nb0 : s.np;
    d.lambda_bar = s.lambda_bar; d.sigma_b = s.sigma_b; d.tau_b = s.tau_b;
    if (h->members.empty()) {
        d.sys = make_sys(s);
        d.np = s.np; d.nw = s.nw; d.K = d.sys.K; d.nwall_blocks = d.sys.nblocks;
        d.systems = nullptr; d.sys_of = nullptr; d.sys_of_block = nullptr; d.nsys = 1;
    } else {
        // members back to back: indices, cells, wall entries (6 each) and
        // wall-sum blocks; the device copy of the table is made at upload
        h->sys_host.clear();
        uint32_t pb = 0, cb = 0, bb = 0, nw = 0;
        for (size_t k = 0; k < h->members.size(); ++k) {
            sph::Sys S = make_sys(h->members[k]);
            S.pbase = pb; S.cbase = cb; S.wbase = (uint32_t)(6 * k); S.bbase = bb;
            pb += S.np; cb += S.K; bb += S.nblocks; nw = std::max(nw, S.nw);
            h->sys_host.push_back(S);
        }
        d.sys = h->sys_host[0];
        d.np = pb; d.nw = nw; d.K = cb; d.nwall_blocks = bb;
        d.nsys = (uint32_t)h->members.size();
    }
    d.scan_blocks = (d.K + sph::SCAN_TILE - 1) / sph::SCAN_TILE;
    if (!d.tpb_force) d.tpb_force = sph::BLOCK;
    if (!d.tpb_search) d.tpb_search = sph::BLOCK;
}

// Contact-search tile: tile_x x tile_y cells of one z plane per block.  Aim
// for about one block of home particles with the staged neighbourhood
// (3 (tile_y+2)(tile_x+2) cells) inside the LDS capacity, using the mean
// occupancy of the non-empty cells.
void choose_tile(sph_dem* h, uint32_t want_x = 0, uint32_t want_y = 0) {
    const HostState& s = h->hs;
    DemDev& d = h->d;
    if (!h->members.empty()) {
        d.tile_x = d.tile_y = d.tiles_x = d.tiles_y = 1;
        if (d.search_mode == 0) d.search_mode = 2;
        return;
    }
    const uint32_t nx = s.num[0], ny = s.num[1];
    uint32_t bx = 1, by = 1;
    if (want_x && want_y) {
        bx = want_x; by = want_y;
    } else {
        std::vector<uint32_t> cnt(d.K, 0);
        uint64_t nonempty = 0;
        for (uint32_t i = 0; i < s.np; ++i) {
            uint32_t c[3];
            for (int a = 0; a < 3; ++a) {
                double f = std::floor((s.x[4 * (size_t)i + a] - s.origo[a]) / d.sys.cell_len[a]);
                if (!(f >= 0.0)) f = 0.0;
                if (f > (double)(s.num[a] - 1)) f = (double)(s.num[a] - 1);
                c[a] = (uint32_t)f;
            }
            uint32_t& k = cnt[(c[2] * s.num[1] + c[1]) * s.num[0] + c[0]];
            nonempty += k == 0;
            ++k;
        }
        const double lam = nonempty ? (double)s.np / (double)nonempty : 1.0;
        double best = -1.0;
        for (uint32_t ty = 1; ty <= ny && ty <= 16; ++ty)
            for (uint32_t tx = 1; tx <= nx && tx <= 64; ++tx) {
                const double cells = 3.0 * (tx + 2) * (ty + 2);
                if (cells > sph::TOPO_MAX_CELLS) continue;
                const double staged = cells * lam, home = tx * ty * lam;
                if (staged > 0.8 * sph::TOPO_CAP) continue;
                const double passes = std::ceil(home / sph::BLOCK);
                const double util = home / (passes * sph::BLOCK);
                const double score = util * home / staged;
                if (score > best) { best = score; bx = tx; by = ty; }
            }
    }
    bx = bx < 1 ? 1 : (bx > nx ? nx : bx);
    by = by < 1 ? 1 : (by > ny ? ny : by);
    while (3u * (bx + 2) * (by + 2) > (uint32_t)sph::TOPO_MAX_CELLS) {
        if (bx > by) --bx; else --by;
    }
    d.tile_x = bx;
    d.tile_y = by;
    if (d.search_mode == 0) d.search_mode = 2;
    d.tiles_x = (nx + bx - 1) / bx;
    d.tiles_y = (ny + by - 1) / by;
}

int check_bond_pairs(sph_dem* h, const uint32_t* bonds, size_t nb, uint32_t np);

int validate_system(sph_dem* h, const HostState& s) {
    if (s.contactmodel < 1 || s.contactmodel > 3) {
        h->err = "contactmodel " + std::to_string(s.contactmodel) + " is not a model of the reference";
        return SPH_E_INVALID;
    }
    if (s.np == 0 || s.np > sph::MAX_PARTICLES) {
        h->err = "np=" + std::to_string(s.np) + " outside [1, 2^28-1]";
        return SPH_E_INVALID;
    }
    if (s.num[0] == 0 || s.num[1] == 0 || s.num[2] == 0 || !(s.L[0] > 0) || !(s.L[1] > 0) || !(s.L[2] > 0)) {
        h->err = "grid has a zero-sized axis";
        return SPH_E_INVALID;
    }
    const double K = (double)s.num[0] * s.num[1] * s.num[2];
    if (K >= 4294967295.0) {
        h->err = "grid has too many cells for 32-bit cell ids";
        return SPH_E_INVALID;
    }
    if ((s.periodic == 1 || s.periodic == 2) && s.num[0] < 3) {
        h->err = "periodic x needs at least 3 cells";
        return SPH_E_INVALID;
    }
    if (s.periodic == 1 && s.num[1] < 3) {
        h->err = "periodic y needs at least 3 cells";
        return SPH_E_INVALID;
    }
    return check_bond_pairs(h, s.bonds.data(), s.nb0, s.np);
}

int validate_for_device(sph_dem* h) {
    if (h->members.empty()) return validate_system(h, h->hs);
    double K = 0.0, np = 0.0;
    for (size_t k = 0; k < h->members.size(); ++k) {
        const HostState& m = h->members[k];
        int rc;
        if ((rc = validate_system(h, m))) {
            h->err = "member " + std::to_string(k) + ": " + h->err;
            return rc;
        }
        if (m.nb0) {
            h->err = "member " + std::to_string(k) + " has bonds; ensembles step unbonded systems";
            return SPH_E_UNSUPPORTED;
        }
        if (m.contactmodel != 2) {
            h->err = "member " + std::to_string(k) + ": ensembles step contact model 2";
            return SPH_E_UNSUPPORTED;
        }
        K += (double)m.num[0] * m.num[1] * m.num[2];
        np += m.np;
    }
    if (K >= 4294967295.0 || np > (double)sph::MAX_PARTICLES) {
        h->err = "ensemble has too many cells or particles for 32-bit cell ids and 28-bit indices";
        return SPH_E_INVALID;
    }
    return SPH_OK;
}

// Host-side view of the device state in original particle order.
struct Snapshot {
    std::vector<uint32_t> id;       // storage position -> original index
    std::vector<double> x, vel, angvel, pr0, pr1, pr2;  // 4 per particle, original order
};

// Per-particle bond lists in ascending bond index, first-index side first
// (entry = k << 1 | side), over the intact bonds of `bonds` (pairs, first <
// np), replacing the handle's lists.
int upload_bond_lists(sph_dem* h, const std::vector<uint32_t>& bonds) {
    DemDev& d = h->d;
    const size_t n = h->hs.np, nb = h->hs.nb0;
    std::vector<uint32_t> off(n + 1, 0), ent;
    for (size_t k = 0; k < nb; ++k) {
        const uint32_t bi = bonds[2 * k], bj = bonds[2 * k + 1];
        if (!(bi < h->hs.np)) continue;
        ++off[bi + 1];
        ++off[bj + 1];
    }
    for (size_t i = 0; i < n; ++i) off[i + 1] += off[i];
    ent.assign(off[n] ? off[n] : 1, 0);
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (size_t k = 0; k < nb; ++k) {
        const uint32_t bi = bonds[2 * k], bj = bonds[2 * k + 1];
        if (!(bi < h->hs.np)) continue;
        ent[cur[bi]++] = (uint32_t)(k << 1);
        ent[cur[bj]++] = (uint32_t)(k << 1) | 1u;
    }
    if (d.bcsr_ent) {
        auto it = std::find(h->allocs.begin(), h->allocs.end(), (void*)d.bcsr_ent);
        if (it != h->allocs.end()) h->allocs.erase(it);
        (void)hipFree(d.bcsr_ent);
        d.bcsr_ent = nullptr;
    }
    if (dalloc(h, &d.bcsr_ent, ent.size())) return SPH_E_HIP;
    HIP_TRY(hipMemcpy(d.bcsr_off, off.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.bcsr_ent, ent.data(), ent.size() * 4, hipMemcpyHostToDevice));
    return SPH_OK;
}

// The bond check of validate_system on a pair list: an intact bond (first
// index < np) must name a second particle < np.
int check_bond_pairs(sph_dem* h, const uint32_t* bonds, size_t nb, uint32_t np) {
    for (size_t k = 0; k < nb; ++k) {
        if (bonds[2 * k] < np && bonds[2 * k + 1] >= np) {
            h->err = "bond " + std::to_string(k) + " names particle " + std::to_string(bonds[2 * k + 1]) + " >= np";
            return SPH_E_INVALID;
        }
    }
    return SPH_OK;
}

int upload_all(sph_dem* h) {
    HostState& s = h->hs;
    DemDev& d = h->d;
    const size_t n = s.np;
    // storage starts in original order: b holds the state, id and pos_of are the identity
    h->own_cur = 0;
    // src/device.cu:413,423: vel0 <- vel, angvel0 <- angvel; packed records
    // pr0 = (vel0, angvel0.x), pr1 = (angvel0.yz, angpos.xy),
    // pr2 = (angpos.z, xysum, es), angvel.w = ev
    std::vector<double> w4(4 * n), r0(4 * n), r1(4 * n), r2(4 * n), f4(4 * n), t4(4 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            w4[4 * i + a] = s.angvel[3 * i + a];
            f4[4 * i + a] = s.force[3 * i + a];
            t4[4 * i + a] = s.torque[3 * i + a];
            f4[4 * i + 3] = s.es_dot[i];
            t4[4 * i + 3] = s.ev_dot[i];
            r0[4 * i + a] = s.vel[4 * i + a];
        }
        w4[4 * i + 3] = s.ev[i];
        r0[4 * i + 3] = s.angvel[3 * i + 0];
        r1[4 * i + 0] = s.angvel[3 * i + 1];
        r1[4 * i + 1] = s.angvel[3 * i + 2];
        r1[4 * i + 2] = s.angpos[3 * i + 0];
        r1[4 * i + 3] = s.angpos[3 * i + 1];
        r2[4 * i + 0] = s.angpos[3 * i + 2];
        r2[4 * i + 1] = s.xysum[2 * i + 0];
        r2[4 * i + 2] = s.xysum[2 * i + 1];
        r2[4 * i + 3] = s.es[i];
    }
    std::vector<uint32_t> iota(n);
    for (size_t i = 0; i < n; ++i) iota[i] = (uint32_t)i;
    HIP_TRY(hipMemcpy(own(h).x, s.x.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(own(h).vel, s.vel.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(own(h).angvel, w4.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(own(h).pr0, r0.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(own(h).pr1, r1.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(own(h).pr2, r2.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.b.id, iota.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.a.id, iota.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.sid, iota.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.pos_of, iota.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d.b.cmask, 0, n * 4));
    HIP_TRY(hipMemset(d.a.cmask, 0, n * 4));
    HIP_TRY(hipMemcpy(d.force, f4.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.torque, t4.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.pres, s.p.data(), n * 8, hipMemcpyHostToDevice));
    for (const sph::PState* p : {&d.a, &d.b}) {
        HIP_TRY(hipMemset(p->cslot, 0, n * sph::NC * 4));
        HIP_TRY(hipMemset(p->cdt, 0, n * sph::NC * 3 * 8));
        if (p->cdt32) HIP_TRY(hipMemset(p->cdt32, 0, n * sph::NC * 3 * 4));
    }
    HIP_TRY(hipMemset(d.flags, 0, 8));
    HIP_TRY(hipMemset(d.key, 0, n * 4));
    HIP_TRY(hipMemset(d.perm, 0, n * 4));
    HIP_TRY(hipMemset(d.off, 0, ((size_t)d.K + 1) * 4));
    HIP_TRY(hipMemset(d.count, 0, (size_t)d.K * 4));
    HIP_TRY(hipMemset(d.hcnt, 0, n * 4));
    HIP_TRY(hipMemset(d.ncnt, 0, n * 4));
    HIP_TRY(hipMemset(d.nhint, 0xff, n * sph::NLIST));
    HIP_TRY(hipMemset(d.a.lidx, 0, n * 4));
    HIP_TRY(hipMemset(d.b.lidx, 0, n * 4));
    {
        const uint32_t ns[6] = {0u, 1u, 0u, 0u, 0u, 0u};  // neighbour lists: rebuild at the first step
        HIP_TRY(hipMemcpy(d.nstate, ns, 24, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(d.a.x, s.x.data(), n * 32, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d.perm, 0, n * 4));
    // walls, 6 entries per system: entries past its nw keep the defaults of
    // src/contactsearch.cuh:416-427
    const std::vector<const HostState*> sys = systems_of(h);
    const size_t ns = sys.size();
    std::vector<double> wnx(24 * ns), wmv(24 * ns, 0.0), wv0(6 * ns, 0.0), tcur(ns);
    std::vector<int> wmode(6 * ns, 0);
    for (size_t k = 0; k < ns; ++k) {
        const HostState& m = *sys[k];
        const double def[24] = {0.0, 0.0, -1.0, m.L[2], -1.0, 0.0, 0.0, m.L[0], 1.0, 0.0, 0.0, 0.0,
                                0.0, -1.0, 0.0, m.L[1], 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        std::copy(def, def + 24, wnx.begin() + 24 * k);
        for (uint32_t w = 0; w < m.nw; ++w) {
            for (int a = 0; a < 4; ++a) {
                wnx[24 * k + 4 * w + a] = m.wnx[4 * w + a];
                wmv[24 * k + 4 * w + a] = m.wmvfd[4 * w + a];
            }
            wv0[6 * k + w] = m.wmvfd[4 * w + 1];
            wmode[6 * k + w] = m.wmode[w];
        }
        tcur[k] = m.t_cur;
    }
    HIP_TRY(hipMemcpy(d.wnx, wnx.data(), wnx.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.wmvfd, wmv.data(), wmv.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.wvel0, wv0.data(), wv0.size() * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.wmode, wmode.data(), wmode.size() * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.t_cur, tcur.data(), ns * 8, hipMemcpyHostToDevice));
    if (d.nw) HIP_TRY(hipMemset(d.wall_pp, 0, (size_t)d.nw * n * 8));
    if (!h->members.empty()) {
        // system table, system of each index and of each wall-sum block
        std::vector<uint16_t> of(n), ofb(d.nwall_blocks);
        for (size_t k = 0; k < ns; ++k) {
            const sph::Sys& S = h->sys_host[k];
            std::fill(of.begin() + S.pbase, of.begin() + S.pbase + S.np, (uint16_t)k);
            std::fill(ofb.begin() + S.bbase, ofb.begin() + S.bbase + S.nblocks, (uint16_t)k);
        }
        HIP_TRY(hipMemcpy(const_cast<sph::Sys*>(d.systems), h->sys_host.data(), ns * sizeof(sph::Sys),
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(const_cast<uint16_t*>(d.sys_of), of.data(), n * 2, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(const_cast<uint16_t*>(d.sys_of_block), ofb.data(), ofb.size() * 2, hipMemcpyHostToDevice));
    }
    // bonds
    const size_t nb = s.nb0;
    if (nb) {
        std::vector<double> bd(4 * nb), bo(4 * nb);
        for (size_t k = 0; k < nb; ++k) {
            for (int a = 0; a < 3; ++a) {
                bd[4 * k + a] = s.bonds_delta_t[3 * k + a];
                bo[4 * k + a] = s.bonds_omega_t[3 * k + a];
            }
            bd[4 * k + 3] = s.bonds_delta_n[k];
            bo[4 * k + 3] = s.bonds_omega_n[k];
        }
        HIP_TRY(hipMemcpy(d.bonds, s.bonds.data(), nb * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.bdelta, bd.data(), nb * 32, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.bomega, bo.data(), nb * 32, hipMemcpyHostToDevice));
        int rc;
        if ((rc = upload_bond_lists(h, s.bonds))) return rc;
    }
    return SPH_OK;
}

// Copies a storage-order double4 array into original order (w kept), for the
// particles of range r (ids relative to r.base, as fetch_ids returns them).
int fetch4(sph_dem* h, const double4* src, const std::vector<uint32_t>& id, std::vector<double>& out, Range r) {
    const size_t n = r.n;
    std::vector<double> t(4 * n);
    HIP_TRY(hipMemcpy(t.data(), src + r.base, n * 32, hipMemcpyDeviceToHost));
    out.resize(4 * n);
    for (size_t q = 0; q < n; ++q)
        for (int a = 0; a < 4; ++a) out[4 * (size_t)id[q] + a] = t[4 * q + a];
    return SPH_OK;
}

int fetch4(sph_dem* h, const double4* src, const std::vector<uint32_t>& id, std::vector<double>& out) {
    return fetch4(h, src, id, out, whole(h));
}

// Original indices of the records at the storage positions of range r,
// relative to r.base.
int fetch_ids(sph_dem* h, std::vector<uint32_t>& id, Range r) {
    id.resize(r.n);
    HIP_TRY(hipMemcpy(id.data(), own(h).id + r.base, r.n * 4, hipMemcpyDeviceToHost));
    for (auto& v : id) v -= (uint32_t)r.base;
    return SPH_OK;
}

int fetch_ids(sph_dem* h, std::vector<uint32_t>& id) { return fetch_ids(h, id, whole(h)); }

int alloc_pstate(sph_dem* h, sph::PState& p, size_t n) {
    ALLOC(p.x, n); ALLOC(p.vel, n); ALLOC(p.angvel, n); ALLOC(p.pr0, n); ALLOC(p.pr1, n);
    ALLOC(p.pr2, n); ALLOC(p.id, n); ALLOC(p.cmask, n); ALLOC(p.lidx, n); ALLOC(p.cslot, n * sph::NC); ALLOC(p.cdt, n * sph::NC * 3);
    p.cdt32 = nullptr;
    return SPH_OK;
}

int allocate(sph_dem* h) {
    DemDev& d = h->d;
    const size_t n = h->hs.np;
    int rc;
    if ((rc = alloc_pstate(h, d.a, n)) || (rc = alloc_pstate(h, d.b, n))) return rc;
    ALLOC(d.pos_of, n);
    ALLOC(d.force, n); ALLOC(d.torque, n); ALLOC(d.pres, n);
    ALLOC(d.key, n); ALLOC(d.rank, n); ALLOC(d.perm, n);
    ALLOC(d.tmask, n); ALLOC(d.tfresh, n); ALLOC(d.skey, n); ALLOC(d.xf, n);
    ALLOC(d.hcnt, n); ALLOC(d.hits, n * sph::NC); ALLOC(d.sid, n);
    ALLOC(d.nlist, n * sph::NLIST); ALLOC(d.nlistg, n * sph::NLIST); ALLOC(d.nhint, n * sph::NLIST); ALLOC(d.nlo, n * sph::NLIST); ALLOC(d.ldisp, n); ALLOC(d.ncnt, n); ALLOC(d.xref, n); ALLOC(d.nstate, 6); ALLOC(d.linv, n); ALLOC(d.slidx, n);
    ALLOC(d.count, d.K); ALLOC(d.off, (size_t)d.K + 1); ALLOC(d.scan_bsum, d.scan_blocks);
    const size_t ns = h->members.empty() ? 1 : h->members.size();
    ALLOC(d.wnx, 6 * ns); ALLOC(d.wmvfd, 6 * ns); ALLOC(d.wvel0, 6 * ns); ALLOC(d.wmode, 6 * ns);
    ALLOC(d.wall_pp, (size_t)d.nw * n); ALLOC(d.wall_partial, (size_t)d.nw * d.nwall_blocks);
    ALLOC(d.t_cur, ns);
    if (!h->members.empty()) {
        sph::Sys* tbl;
        uint16_t *of, *ofb;
        ALLOC(tbl, ns); ALLOC(of, n); ALLOC(ofb, d.nwall_blocks);
        d.systems = tbl; d.sys_of = of; d.sys_of_block = ofb;
    }
    const size_t nb = h->hs.nb0;
    ALLOC(d.bonds, nb); ALLOC(d.bdelta, nb); ALLOC(d.bomega, nb); ALLOC(d.bond_f, nb); ALLOC(d.bond_t, nb);
    ALLOC(d.bcsr_off, n + 1);
    d.bcsr_ent = nullptr;
    ALLOC(d.flags, 2);
    d.flags_next = d.flags + 1;
    return SPH_OK;
}

void free_all(sph_dem* h) {
    destroy_graph(h);
    if (h->events_made)
        for (auto& e : h->ev) (void)hipEventDestroy(e);
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->join_ev) (void)hipEventDestroy(h->join_ev);
    h->fork_ev = h->join_ev = nullptr;
    if (h->side) (void)hipStreamDestroy(h->side);
    h->side = nullptr;
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->stream = nullptr;
}

}  // namespace

// The wall update only feeds the next step's force kernel: with `lead` the
// step starts with the previous step's wall update on the side stream, joined
// before its own force kernel; with `tail` it ends with its own wall update,
// otherwise that update is left pending for the next step or the caller.
void sph::enqueue_step(sph_dem* h, bool timed, int parity, bool lead, bool tail, const StepHooks* hooks) {
    DemDev d = h->d;
    d.ro = own_set(parity ? h->d.a : h->d.b);
    d.wo = own_set(parity ? h->d.b : h->d.a);
    hipStream_t s = h->stream;
    if (lead) {
        (void)hipEventRecord(h->fork_ev, s);
        (void)hipStreamWaitEvent(h->side, h->fork_ev, 0);
        sph::launch_walls(d, h->side);
        (void)hipEventRecord(h->join_ev, h->side);
    }
    if (timed) (void)hipEventRecord(h->ev[0], s);
    sph::launch_bonds(d, s);
    if (timed) (void)hipEventRecord(h->ev[1], s);
    // cell ids, per-cell counts and ranks come from the previous step's
    // force kernel (or prepare_keys)
    if (timed) (void)hipEventRecord(h->ev[2], s);
    sph::launch_scan(d, s);
    sph::launch_scatter(d, s);
    sph::launch_cellsort(d, s);
    if (timed) (void)hipEventRecord(h->ev[3], s);
    sph::launch_reorder(d, s);
    if (timed) (void)hipEventRecord(h->ev[4], s);
    sph::launch_topology(d, s);
    if (timed) (void)hipEventRecord(h->ev[5], s);
    if (hooks && hooks->before_force) hooks->before_force(d, s);
    if (lead) (void)hipStreamWaitEvent(s, h->join_ev, 0);
    sph::launch_forces(d, h->precision, s);
    if (hooks && hooks->after_force) hooks->after_force(d, s);
    if (timed) (void)hipEventRecord(h->ev[6], s);
    if (tail) sph::launch_walls(d, s);
    if (timed) (void)hipEventRecord(h->ev[7], s);
}

namespace {

// Cell ids of the stored positions (after upload or sph_set of positions):
// what a step's force kernel leaves for the next step.
int prepare_keys(sph_dem* h) {
    DemDev d = h->d;
    d.ro = own_set(own(h));
    sph::launch_fill_u32(h->d.flags_next, 1, 0u, h->stream);
    sph::launch_fill_u32(h->d.count, h->d.K, 0u, h->stream);
    sph::launch_cellid(d, h->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SPH_OK;
}

// Neighbour lists (search mode 5) no longer describe the state: rebuild them
// at the next step.
void request_list_rebuild(sph_dem* h) { sph::launch_fill_u32(h->d.nstate + 1, 1, 1u, h->stream); }

int ensure_graph(sph_dem* h) {
    for (int l = 0; l < 2; ++l)
        for (int k = 0; k < 2; ++k) {
            if (h->gexec[l][k]) continue;
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            sph::enqueue_step(h, false, k, l == 1, false);
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamEndCapture(h->stream, &g));
            h->graph[l][k] = g;
            HIP_TRY(hipGraphInstantiate(&h->gexec[l][k], g, nullptr, nullptr, 0));
        }
    return SPH_OK;
}

int check_flags(sph_dem* h) {
    uint32_t f = 0;
    HIP_TRY(hipMemcpyAsync(&f, h->d.flags, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (!f) return SPH_OK;
    HIP_TRY(hipMemsetAsync(h->d.flags, 0, 4, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (f & sph::FLAG_NAN) { h->err = "non-finite particle position"; return SPH_E_NAN; }
    if (f & sph::FLAG_OUT_OF_GRID) {
        h->err = "one or more particles moved more than one cell outside the grid; is the time step too large?";
        return SPH_E_OUT_OF_GRID;
    }
    h->err = "a particle has more than " + std::to_string(sph::NC) + " contacts";
    return SPH_E_CONTACT_OVERFLOW;
}

int ensure_mixed_buffers(sph_dem* h) {
    if (h->d.a.cdt32) return SPH_OK;
    const size_t n = h->hs.np;
    ALLOC(h->d.a.cdt32, n * sph::NC * 3); ALLOC(h->d.b.cdt32, n * sph::NC * 3);
    HIP_TRY(hipMemset(h->d.a.cdt32, 0, n * sph::NC * 3 * 4));
    HIP_TRY(hipMemset(h->d.b.cdt32, 0, n * sph::NC * 3 * 4));
    ALLOC(h->d.a.vf, n); ALLOC(h->d.a.wf, n); ALLOC(h->d.b.vf, n); ALLOC(h->d.b.wf, n);
    return SPH_OK;
}

// Mixed mode: the FP32 velocity copies of the current records, after any
// change of the state outside a step (steps in mixed mode keep them).
void refresh_f32(sph_dem* h) {
    if (h->precision != SPH_PREC_MIXED || !h->d.a.vf) return;
    DemDev d = h->d;
    d.ro = own_set(own(h));
    sph::launch_f32_copies(d, h->stream);
}

// Downloads every particle array into the host mirror (for writebin).
int download_all(sph_dem* h) {
    HostState& s = h->hs;
    DemDev& d = h->d;
    const size_t n = s.np;
    HIP_TRY(hipStreamSynchronize(h->stream));
    int rc;
    std::vector<uint32_t> id;
    if ((rc = fetch_ids(h, id))) return rc;
    std::vector<double> t;
    if ((rc = fetch4(h, own(h).x, id, t))) return rc;
    s.x = t;
    if ((rc = fetch4(h, own(h).vel, id, t))) return rc;
    s.vel = t;
    auto unpad = [&](std::vector<double>& dst) {
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) dst[3 * i + a] = t[4 * i + a];
    };
    if ((rc = fetch4(h, d.force, id, t))) return rc;
    unpad(s.force);
    for (size_t i = 0; i < n; ++i) s.es_dot[i] = t[4 * i + 3];
    if ((rc = fetch4(h, own(h).angvel, id, t))) return rc;
    unpad(s.angvel);
    for (size_t i = 0; i < n; ++i) s.ev[i] = t[4 * i + 3];
    if ((rc = fetch4(h, d.torque, id, t))) return rc;
    unpad(s.torque);
    for (size_t i = 0; i < n; ++i) s.ev_dot[i] = t[4 * i + 3];
    std::vector<double> t1, t2;
    if ((rc = fetch4(h, own(h).pr1, id, t1)) || (rc = fetch4(h, own(h).pr2, id, t2))) return rc;
    for (size_t i = 0; i < n; ++i) {
        s.angpos[3 * i + 0] = t1[4 * i + 2];
        s.angpos[3 * i + 1] = t1[4 * i + 3];
        s.angpos[3 * i + 2] = t2[4 * i + 0];
        s.xysum[2 * i] = t2[4 * i + 1]; s.xysum[2 * i + 1] = t2[4 * i + 2];
        s.es[i] = t2[4 * i + 3];
    }
    {
        std::vector<double> pv(n);
        HIP_TRY(hipMemcpy(pv.data(), d.pres, n * 8, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) s.p[id[q]] = pv[q];
    }
    if (s.nw) {
        double wnx[24], wmv[24];
        HIP_TRY(hipMemcpy(wnx, d.wnx, sizeof(wnx), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(wmv, d.wmvfd, sizeof(wmv), hipMemcpyDeviceToHost));
        for (uint32_t w = 0; w < s.nw; ++w)
            for (int a = 0; a < 4; ++a) { s.wnx[4 * w + a] = wnx[4 * w + a]; s.wmvfd[4 * w + a] = wmv[4 * w + a]; }
    }
    const size_t nb = s.nb0;
    if (nb) {
        std::vector<double> bd(4 * nb), bo(4 * nb);
        HIP_TRY(hipMemcpy(s.bonds.data(), d.bonds, nb * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bd.data(), d.bdelta, nb * 32, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bo.data(), d.bomega, nb * 32, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nb; ++k) {
            for (int a = 0; a < 3; ++a) {
                s.bonds_delta_t[3 * k + a] = bd[4 * k + a];
                s.bonds_omega_t[3 * k + a] = bo[4 * k + a];
            }
            s.bonds_delta_n[k] = bd[4 * k + 3];
            s.bonds_omega_n[k] = bo[4 * k + 3];
        }
    }
    return SPH_OK;
}

int open_common(sph_dem* h, int verbosity, int check_vals) {
    h->verbosity = verbosity;
    h->check_vals = check_vals;
    int rc;
    if (check_vals == 1 && (rc = sph::check_values(h->hs, h->err))) return rc;
    if ((rc = validate_for_device(h))) return rc;
    fill_params(h);
    update_search_eps(h);
    choose_tile(h);
    if (h->hs.contactmodel == 1 && h->members.empty()) {  // no contact search to choose
        h->search_auto = false;
        h->d.search_mode = 2;
    }
    HIP_TRY(hipGetDevice(&h->device));
    HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&h->join_ev, hipEventDisableTiming));
    if ((rc = allocate(h))) return rc;
    if ((rc = upload_all(h))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    if ((rc = prepare_keys(h))) return rc;
    if (verbosity > 0) {
        std::printf("sphere_dem: np=%u nb0=%u nw=%u grid=%ux%ux%u periodic=%d dt=%g\n", h->hs.np, h->hs.nb0,
                    h->hs.nw, h->hs.num[0], h->hs.num[1], h->hs.num[2], h->hs.periodic, h->hs.dt);
    }
    return SPH_OK;
}

int finish_open(sph_dem* h, int rc, sph_t** out) {
    if (rc) {
        g_open_error = h->err;
        free_all(h);
        delete h;
        *out = nullptr;
        return rc;
    }
    *out = h;
    return SPH_OK;
}

int make_events(sph_dem* h) {
    if (h->events_made) return SPH_OK;
    for (auto& e : h->ev) HIP_TRY(hipEventCreate(&e));
    h->events_made = true;
    return SPH_OK;
}

// Times the contact search of four steps with each of the row (2),
// half-shell (4) and neighbour-list (5) kernels and keeps the fastest by the
// mean of the last three (the first warms up, or builds the lists), then the
// force kernel with and without the partners' velocity gathers batched
// (three steps each), each when its choice is automatic; the steps are
// ordinary steps of the run (all kernels give the same bits).
constexpr uint64_t TUNE_SEARCH_STEPS = 12, TUNE_FORCE_STEPS = 6, TUNE_BLOCK_STEPS = 9;

uint64_t tune_steps(const sph_dem* h) {
    return (h->search_auto ? TUNE_SEARCH_STEPS : 0) + (h->force_auto ? TUNE_FORCE_STEPS : 0) +
           (h->blocks_auto ? TUNE_BLOCK_STEPS : 0);
}

// One ordinary step, timed per stage: ms[0] search, ms[1] force.
int timed_step(sph_dem* h, float ms[2]) {
    sph::enqueue_step(h, true, h->own_cur, false, true);
    h->own_cur ^= 1;
    h->hs.t_cur += h->hs.dt;
    ++h->steps_done;
    HIP_TRY(hipEventSynchronize(h->ev[7]));
    HIP_TRY(hipEventElapsedTime(&ms[0], h->ev[4], h->ev[5]));
    HIP_TRY(hipEventElapsedTime(&ms[1], h->ev[5], h->ev[6]));
    return SPH_OK;
}

// Threads per block tried for the search and force kernels.
constexpr uint32_t TUNE_BLOCKS[3] = {64, 128, 256};

}  // namespace

int sph::tune_search(sph_dem* h, uint64_t* ran) {
    int rc;
    if ((rc = make_events(h))) return rc;
    const int modes[3] = {2, 4, 5};
    float mean[3] = {0.0f, 0.0f, 0.0f};
    uint64_t steps = 0;
    for (int k = 0; k < 3 && h->search_auto; ++k) {
        h->d.search_mode = modes[k];
        if (modes[k] == 5) request_list_rebuild(h);
        for (int rep = 0; rep < 4; ++rep) {
            float ms[2];
            if ((rc = timed_step(h, ms))) return rc;
            ++steps;
            if (rep > 0) mean[k] += ms[0] / 3.0f;
        }
    }
    if (h->search_auto) {
        int best = 0;
        for (int k = 1; k < 3; ++k)
            if (mean[k] < mean[best]) best = k;
        h->d.search_mode = modes[best];
        if (modes[best] == 5) request_list_rebuild(h);
    }
    float fv[2] = {0.0f, 0.0f};
    for (int v = 0; v < 2 && h->force_auto; ++v) {
        h->d.force_vw = (uint32_t)v;
        for (int rep = 0; rep < 3; ++rep) {
            float ms[2];
            if ((rc = timed_step(h, ms))) return rc;
            ++steps;
            if (rep > 0) fv[v] += ms[1] / 2.0f;
        }
    }
    if (h->force_auto) h->d.force_vw = fv[1] < fv[0] ? 1u : 0u;
    if (h->blocks_auto) {
        // threads per block of the chosen search kernel and of k_force, on
        // the same steps (mean of the last two of three per size)
        float tb[3][2] = {};
        for (int k = 0; k < 3; ++k) {
            h->d.tpb_search = h->d.tpb_force = TUNE_BLOCKS[k];
            for (int rep = 0; rep < 3; ++rep) {
                float ms[2];
                if ((rc = timed_step(h, ms))) return rc;
                ++steps;
                if (rep > 0) { tb[k][0] += ms[0] / 2.0f; tb[k][1] += ms[1] / 2.0f; }
            }
        }
        int bs = 0, bf = 0;
        for (int k = 1; k < 3; ++k) {
            if (tb[k][0] < tb[bs][0]) bs = k;
            if (tb[k][1] < tb[bf][1]) bf = k;
        }
        h->d.tpb_search = TUNE_BLOCKS[bs];
        h->d.tpb_force = TUNE_BLOCKS[bf];
    }
    if (ran) *ran = steps;
    destroy_graph(h);
    h->tuned = true;
    h->steps_since_tune = 0;
    if (h->verbosity > 1)
        std::printf("sphere_dem: contact search %.4f (rows) %.4f (half shell) %.4f (lists) ms: mode %d; "
                    "force %.4f / %.4f ms: velocity gathers %s\n", mean[0], mean[1], mean[2], h->d.search_mode,
                    fv[0], fv[1], h->d.force_vw ? "batched" : "per slot");
    return check_flags(h);
}

namespace {

struct Region {
    void* p;
    size_t bytes;
};

// `own_cur`: which buffer set holds the state (the slot words and springs
// are kept for that set only: the other set's are written by the next step
// before any read).
std::vector<Region> state_regions(sph_dem* h, int own_cur) {
    const DemDev& d = h->d;
    const sph::PState& cur = own_cur ? d.a : d.b;
    const size_t n = h->hs.np, nc = sph::NC;
    const size_t ns = h->members.empty() ? 1 : h->members.size();
    std::vector<Region> r;
    for (const sph::PState* p : {&d.a, &d.b}) {
        r.push_back({p->x, n * 32});
        r.push_back({p->vel, n * 32});
        r.push_back({p->angvel, n * 32});
        r.push_back({p->pr0, n * 32});
        r.push_back({p->pr1, n * 32});
        r.push_back({p->pr2, n * 32});
        r.push_back({p->id, n * 4});
        r.push_back({p->cmask, n * 4});
    }
    r.push_back({d.pos_of, n * 4});
    r.push_back({d.key, n * 4});  // the next step's keys with their counts and ranks
    r.push_back({d.count, (size_t)d.K * 4});
    r.push_back({d.rank, n * 4});
    r.push_back({cur.cslot, n * nc * 4});
    r.push_back({cur.cdt, n * nc * 3 * 8});
    if (cur.cdt32) r.push_back({cur.cdt32, n * nc * 3 * 4});
    r.push_back({d.wnx, 6 * ns * 32});
    r.push_back({d.wmvfd, 6 * ns * 32});
    r.push_back({d.wvel0, 6 * ns * 8});
    r.push_back({d.t_cur, ns * 8});
    r.push_back({d.flags, 8});
    if (d.nstate) r.push_back({d.nstate, 24});  // list-build requests and count: builds on a copy do not count
    r.push_back({d.force, n * 32});
    r.push_back({d.torque, n * 32});
    r.push_back({d.pres, n * 8});
    if (d.nw) r.push_back({d.wall_pp, (size_t)d.nw * n * 8});
    if (h->hs.nb0) {
        const size_t nb = h->hs.nb0;
        r.push_back({d.bonds, nb * 8});
        r.push_back({d.bdelta, nb * 32});
        r.push_back({d.bomega, nb * 32});
    }
    return r;
}

}  // namespace

sph::Checkpoint::~Checkpoint() {
    for (void* p : bufs) (void)hipFree(p);
}

int sph::save_state(sph_dem* h, Checkpoint& c) {
    const auto regs = state_regions(h, h->own_cur);
    if (c.bufs.size() != regs.size()) {
        for (void* p : c.bufs) (void)hipFree(p);
        c.bufs.clear();
        for (const Region& r : regs) {
            void* q = nullptr;
            if (hipMalloc(&q, r.bytes) != hipSuccess) {
                h->err = "state copy: hipMalloc of " + std::to_string(r.bytes) + " bytes failed";
                return SPH_E_HIP;
            }
            c.bufs.push_back(q);
        }
    }
    for (size_t k = 0; k < regs.size(); ++k)
        HIP_TRY(hipMemcpyAsync(c.bufs[k], regs[k].p, regs[k].bytes, hipMemcpyDeviceToDevice, h->stream));
    c.own_cur = h->own_cur;
    c.t_cur = h->hs.t_cur;
    c.steps_done = h->steps_done;
    return SPH_OK;
}

int sph::restore_state(sph_dem* h, const Checkpoint& c) {
    const auto regs = state_regions(h, c.own_cur);
    if (c.bufs.size() != regs.size()) {
        h->err = "state copy: the checkpoint does not match the handle";
        return SPH_E_STATE;
    }
    for (size_t k = 0; k < regs.size(); ++k)
        HIP_TRY(hipMemcpyAsync(regs[k].p, c.bufs[k], regs[k].bytes, hipMemcpyDeviceToDevice, h->stream));
    h->own_cur = c.own_cur;
    h->hs.t_cur = c.t_cur;
    h->steps_done = c.steps_done;
    // neighbour lists (search mode 5) describe another state: rebuild
    request_list_rebuild(h);
    refresh_f32(h);
    return SPH_OK;
}

int sph::autotune(sph_dem* h) {
    if (!tune_steps(h) || (h->tuned && h->steps_since_tune < 4096)) return SPH_OK;
    Checkpoint c;
    int rc;
    if ((rc = save_state(h, c))) return rc;
    // a simulation error inside the trial steps (the state blowing up within
    // them) is the trial's, not the caller's: the kernels chosen so far are
    // kept and the restored state is stepped as asked
    rc = tune_search(h);
    if (rc == SPH_E_HIP) return rc;
    h->tuned = true;
    h->steps_since_tune = 0;
    destroy_graph(h);
    if ((rc = restore_state(h, c))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return SPH_OK;
}

namespace {

int run_steps(sph_dem* h, uint64_t n) {
    if (n == 0) return SPH_OK;
    int rc;
    const uint64_t chunk = 512;
    if (h->profiling) {
        if ((rc = make_events(h))) return rc;
        sph_timing_t t{};
        for (uint64_t i = 0; i < n; ++i) {
            sph::enqueue_step(h, true, h->own_cur, false, true);
            h->own_cur ^= 1;
            HIP_TRY(hipEventSynchronize(h->ev[7]));
            float ms[7];
            for (int k = 0; k < 7; ++k) HIP_TRY(hipEventElapsedTime(&ms[k], h->ev[k], h->ev[k + 1]));
            t.bonds_ms += ms[0]; t.cellid_ms += ms[1]; t.sort_ms += ms[2]; t.reorder_ms += ms[3];
            t.search_ms += ms[4]; t.force_ms += ms[5]; t.walls_ms += ms[6];
            h->hs.t_cur += h->hs.dt;
            ++h->steps_done;
        }
        t.contact_ms = t.search_ms + t.force_ms;
        t.total_ms = t.bonds_ms + t.cellid_ms + t.sort_ms + t.reorder_ms + t.contact_ms + t.walls_ms;
        t.steps = n;
        t.tile_x = h->d.tile_x;
        t.tile_y = h->d.tile_y;
        t.search_mode = h->d.search_mode;
        t.search_auto = h->search_auto ? 1 : 0;
        t.force_variant = h->d.force_vw ? 2 : 1;
        h->timing = t;
        return check_flags(h);
    }
    if ((rc = sph::autotune(h))) return rc;
    h->steps_since_tune += n;
    if (h->use_graph && (rc = ensure_graph(h))) return rc;
    // step i > 0 runs step i-1's wall update beside its cell list and search;
    // the last step's wall update is enqueued before returning, also on error
    bool pending = false;
    for (uint64_t i = 0; i < n; ++i) {
        const int lead = i > 0 ? 1 : 0;
        if (h->use_graph) HIP_TRY(hipGraphLaunch(h->gexec[lead][h->own_cur], h->stream));
        else sph::enqueue_step(h, false, h->own_cur, lead == 1, false);
        pending = true;
        h->own_cur ^= 1;
        h->hs.t_cur += h->hs.dt;
        ++h->steps_done;
        if ((i + 1) % chunk == 0 && i + 1 < n) {
            if ((rc = check_flags(h))) {
                sph::launch_walls(h->d, h->stream);
                (void)hipStreamSynchronize(h->stream);
                return rc;
            }
        }
    }
    if (pending) sph::launch_walls(h->d, h->stream);
    HIP_TRY(hipGetLastError());
    return check_flags(h);
}

bool field_is_known(int f) { return f >= 0 && f < SPH_NFIELDS; }

// A field component kept as component c of a per-particle double4 array.
struct Comp {
    double4* arr;
    int c;
};

// Fields packed into the private records and the spare lane of angvel.
std::vector<Comp> packed_comps(const sph::PState& p, int f) {
    switch (f) {
        case SPH_VEL0: return {{p.pr0, 0}, {p.pr0, 1}, {p.pr0, 2}};
        case SPH_ANGVEL0: return {{p.pr0, 3}, {p.pr1, 0}, {p.pr1, 1}};
        case SPH_ANGPOS: return {{p.pr1, 2}, {p.pr1, 3}, {p.pr2, 0}};
        case SPH_XYSUM: return {{p.pr2, 1}, {p.pr2, 2}};
        case SPH_ES: return {{p.pr2, 3}};
        case SPH_EV: return {{p.angvel, 3}};
        default: return {};
    }
}

// storage-order arrays -> original-order values, comps.size() per particle,
// for the particles of range r
int get_comps(sph_dem* h, const std::vector<Comp>& comps, double* out, Range r) {
    const size_t n = r.n, nc = comps.size();
    std::vector<uint32_t> id;
    int rc;
    if ((rc = fetch_ids(h, id, r))) return rc;
    std::vector<double> t(4 * n);
    for (size_t a = 0; a < nc; ++a) {
        if (a == 0 || comps[a].arr != comps[a - 1].arr)
            HIP_TRY(hipMemcpy(t.data(), comps[a].arr + r.base, n * 32, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) out[nc * id[q] + a] = t[4 * q + comps[a].c];
    }
    return SPH_OK;
}

// original-order values -> the given components, the rest of each record kept
int set_comps(sph_dem* h, const std::vector<Comp>& comps, const double* in) {
    const size_t n = h->hs.np, nc = comps.size();
    std::vector<uint32_t> id;
    int rc;
    if ((rc = fetch_ids(h, id))) return rc;
    std::vector<double> t(4 * n);
    for (size_t a = 0; a < nc; ++a) {
        const bool first = a == 0 || comps[a].arr != comps[a - 1].arr;
        const bool last = a + 1 == nc || comps[a + 1].arr != comps[a].arr;
        if (first) HIP_TRY(hipMemcpy(t.data(), comps[a].arr, n * 32, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) t[4 * q + comps[a].c] = in[nc * id[q] + a];
        if (last) HIP_TRY(hipMemcpy(comps[a].arr, t.data(), n * 32, hipMemcpyHostToDevice));
    }
    return SPH_OK;
}


// Per-particle fields of the particles of range r, in their original order
// relative to r.base (contact partners too; an empty slot reads r.n).
// Returns 1 for a field that is not per particle.
int get_particle_field(sph_dem* h, int f, void* dst, Range r) {
    const DemDev& d = h->d;
    const size_t n = r.n, nc = sph::NC;
    double* dd = static_cast<double*>(dst);
    uint32_t* du = static_cast<uint32_t*>(dst);
    int rc;
    std::vector<uint32_t> id;
    std::vector<double> t;
    // per-particle double4 record in storage order -> comps [c0, c0+nc_) in original order
    auto gather = [&](const double4* src, int c0, int ncomp) -> int {
        int q;
        if ((q = fetch_ids(h, id, r)) || (q = fetch4(h, src, id, t, r))) return q;
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < ncomp; ++a) dd[(size_t)ncomp * i + a] = t[4 * i + c0 + a];
        return SPH_OK;
    };
    switch (f) {
        case SPH_X: return gather(own(h).x, 0, 4);
        case SPH_VEL: return gather(own(h).vel, 0, 4);
        case SPH_ANGVEL: return gather(own(h).angvel, 0, 3);
        case SPH_FORCE: return gather(d.force, 0, 3);
        case SPH_TORQUE: return gather(d.torque, 0, 3);
        case SPH_ANGPOS: case SPH_VEL0: case SPH_ANGVEL0: case SPH_XYSUM: case SPH_ES: case SPH_EV:
            return get_comps(h, packed_comps(own(h), f), dd, r);
        case SPH_ES_DOT: return gather(d.force, 3, 1);
        case SPH_EV_DOT: return gather(d.torque, 3, 1);
        case SPH_P: {
            std::vector<double> pv(n);
            if ((rc = fetch_ids(h, id, r))) return rc;
            HIP_TRY(hipMemcpy(pv.data(), d.pres + r.base, n * 8, hipMemcpyDeviceToHost));
            for (size_t q = 0; q < n; ++q) dd[id[q]] = pv[q];
            return SPH_OK;
        }
        case SPH_CONTACTS: case SPH_DISTMOD: case SPH_DELTA_T: {
            std::vector<uint32_t> smask(n), mask(n), slot(n * nc);
            if ((rc = fetch_ids(h, id, r))) return rc;
            HIP_TRY(hipMemcpy(smask.data(), own(h).cmask + r.base, n * 4, hipMemcpyDeviceToHost));
            for (size_t q = 0; q < n; ++q) mask[id[q]] = smask[q];
            {  // slot k of storage position q at k * np + q in the current set
                std::vector<uint32_t> t(n * nc);
                HIP_TRY(hipMemcpy2D(t.data(), n * 4, own(h).cslot + r.base, (size_t)d.np * 4, n * 4, nc,
                                    hipMemcpyDeviceToHost));
                for (size_t q = 0; q < n; ++q)
                    for (size_t k = 0; k < nc; ++k) slot[id[q] * nc + k] = t[k * n + q];
            }
            auto occ = [&](size_t i, size_t k) { return ((mask[i] >> k) & 1u) != 0; };
            if (f == SPH_CONTACTS) {
                for (size_t i = 0; i < n; ++i)
                    for (size_t k = 0; k < nc; ++k)
                        du[i * nc + k] = occ(i, k) ? (slot[i * nc + k] & sph::SLOT_IDX_MASK) - (uint32_t)r.base
                                                   : (uint32_t)n;
                return SPH_OK;
            }
            if (f == SPH_DISTMOD) {
                std::vector<double> x;
                if ((rc = fetch4(h, own(h).x, id, x, r))) return rc;
                for (size_t i = 0; i < n; ++i)
                    for (size_t k = 0; k < nc; ++k) {
                        double* o4 = dd + 4 * (i * nc + k);
                        if (!occ(i, k)) { o4[0] = o4[1] = o4[2] = o4[3] = 0.0; continue; }
                        const uint32_t w = slot[i * nc + k], code = w >> 28;
                        const size_t b = (w & sph::SLOT_IDX_MASK) - r.base;
                        const uint32_t cx = code & 3u, cy = (code >> 2) & 3u;
                        o4[0] = cx == 1u ? h->hs.L[0] : (cx == 2u ? -h->hs.L[0] : 0.0);
                        o4[1] = cy == 1u ? h->hs.L[1] : (cy == 2u ? -h->hs.L[1] : 0.0);
                        o4[2] = 0.0;
                        o4[3] = x[4 * b + 3];
                    }
                return SPH_OK;
            }
            // springs: slot k of storage position q at 3 * (k * np + q), like the slot words
            const size_t np = d.np;
            std::vector<double> t(n * nc * 3);
            if (h->precision == SPH_PREC_MIXED && own(h).cdt32) {
                std::vector<float> tf(n * nc * 3);
                HIP_TRY(hipMemcpy2D(tf.data(), n * 12, own(h).cdt32 + 3 * r.base, np * 12, n * 12, nc,
                                    hipMemcpyDeviceToHost));
                for (size_t i = 0; i < tf.size(); ++i) t[i] = (double)tf[i];
            } else {
                HIP_TRY(hipMemcpy2D(t.data(), n * 24, own(h).cdt + 3 * r.base, np * 24, n * 24, nc,
                                    hipMemcpyDeviceToHost));
            }
            for (size_t q = 0; q < n; ++q) {
                const size_t i = id[q];
                for (size_t k = 0; k < nc; ++k)
                    for (int a = 0; a < 3; ++a)
                        dd[3 * (i * nc + k) + a] = occ(i, k) ? t[3 * (k * n + q) + a] : 0.0;
            }
            return SPH_OK;
        }
        default: return 1;
    }
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

int sph_open_mem(const void* bin, size_t bytes, int verbosity, int check_vals, sph_t** out) {
    if (!out) return SPH_E_INVALID;
    sph_dem* h = new sph_dem();
    int rc = sph::parse_bin(bin, bytes, h->hs, h->err);
    if (!rc) rc = open_common(h, verbosity, check_vals);
    return finish_open(h, rc, out);
}

int sph_open(const char* path, int verbosity, int check_vals, sph_t** out) {
    if (!out || !path) return SPH_E_INVALID;
    sph_dem* h = new sph_dem();
    int rc = sph::read_bin_file(path, h->hs, h->err);
    if (!rc) rc = open_common(h, verbosity, check_vals);
    return finish_open(h, rc, out);
}

int sph_readbin(sph_t* h, const char* path) {
    if (!h || !path) return SPH_E_INVALID;
    if (!h->members.empty()) {
        h->err = "sph_readbin: not on an ensemble";
        return SPH_E_STATE;
    }
    // parse and check on a scratch handle first: errors leave h unchanged
    auto* t = new sph_dem();
    int rc = sph::read_bin_file(path, t->hs, t->err);
    if (!rc && h->check_vals == 1) rc = sph::check_values(t->hs, t->err);
    if (!rc) rc = validate_for_device(t);
    if (rc) {
        h->err = t->err;
        delete t;
        return rc;
    }
    // the new simulation is opened completely on the scratch handle, with
    // this handle's settings, and swapped in only once it is ready: any
    // error (an allocation that fails for a larger file, say) leaves h as it
    // was
    t->d.search_mode = h->d.search_mode;
    t->d.tpb_force = h->d.tpb_force;
    t->d.tpb_search = h->d.tpb_search;
    t->d.force_vw = h->d.force_vw;
    t->force_auto = h->force_auto;
    rc = open_common(t, h->verbosity, h->check_vals);
    if (!rc) {
        t->use_graph = h->use_graph;
        t->search_auto = h->search_auto && t->hs.contactmodel != 1;
        t->profiling = h->profiling;
        if (t->d.search_mode == 5) request_list_rebuild(t);
        if (h->precision != SPH_PREC_FP64) rc = sph_set_precision(t, h->precision);
    }
    if (rc) {
        h->err = t->err;
        free_all(t);
        delete t;
        return rc;
    }
    h->adj.reset();
    std::swap(*h, *t);
    free_all(t);
    delete t;
    return SPH_OK;
}

void sph_close(sph_t* h) {
    if (!h) return;
    free_all(h);
    delete h;
}

const char* sph_last_error(const sph_t* h) { return h ? h->err.c_str() : g_open_error.c_str(); }

int sph_step(sph_t* h, uint64_t n) {
    if (!h) return SPH_E_INVALID;
    return run_steps(h, n);
}

int sph_writebin(sph_t* h, const char* path) {
    if (!h || !path) return SPH_E_INVALID;
    int rc = download_all(h);
    if (rc) return rc;
    return sph::write_bin_file(path, h->hs, h->err);
}

int sph_check_values(sph_t* h) {
    if (!h) return SPH_E_INVALID;
    int rc = download_all(h);
    if (rc) return rc;
    return sph::check_values(h->hs, h->err);
}

int sph_run(sph_t* h, const char* outdir, const char* sid) {
    if (!h || !outdir || !sid) return SPH_E_INVALID;
    HostState& s = h->hs;
    const std::string base = std::string(outdir) + "/" + sid;
    auto write_status = [&]() -> int {
        const std::string p = base + ".status.dat";
        FILE* f = std::fopen(p.c_str(), "w");
        if (!f) { h->err = "cannot write " + p; return SPH_E_IO; }
        std::fprintf(f, "%2.4e %2.4e %d\n", s.t_cur, 100.0 * s.t_cur / s.t_total, (int)s.step_count);
        std::fclose(f);
        return SPH_OK;
    };
    int rc = sph_writebin(h, (base + ".output00000.bin").c_str());
    if (rc) return rc;
    if ((rc = write_status())) return rc;
    // src/device.cu:594,1109-1138: float output clock, double simulation time
    float filetimeclock = 0.0f;
    while (s.t_cur <= s.t_total) {
        // count the steps until the next output or the end of the run
        uint64_t k = 0;
        double t = s.t_cur;
        float ftc = filetimeclock;
        bool out = false;
        do {
            t += s.dt;
            ftc += s.dt;
            ++k;
            if (ftc >= s.file_dt) { out = true; break; }
        } while (t <= s.t_total && k < 100000);
        if ((rc = run_steps(h, k))) return rc;
        filetimeclock = ftc;
        if (out) {
            s.step_count += 1;
            char name[64];
            std::snprintf(name, sizeof(name), ".output%05u.bin", s.step_count);
            if ((rc = sph_writebin(h, (base + name).c_str()))) return rc;
            if ((rc = write_status())) return rc;
            filetimeclock = 0.0f;
        }
    }
    return SPH_OK;
}

size_t sph_field_bytes(const sph_t* h, int f) {
    if (!h || !field_is_known(f)) return 0;
    const size_t n = h->hs.np, nb = h->hs.nb0, nw = h->hs.nw, nc = sph::NC;
    switch (f) {
        case SPH_X: case SPH_VEL: case SPH_X_SORTED: return n * 32;
        case SPH_ANGVEL: case SPH_FORCE: case SPH_TORQUE: case SPH_ANGPOS: case SPH_VEL0:
        case SPH_ANGVEL0: return n * 24;
        case SPH_XYSUM: return n * 16;
        case SPH_ES_DOT: case SPH_ES: case SPH_EV_DOT: case SPH_EV: case SPH_P: return n * 8;
        case SPH_CONTACTS: return n * nc * 4;
        case SPH_DISTMOD: return n * nc * 32;
        case SPH_DELTA_T: return n * nc * 24;
        case SPH_CELL_KEYS: case SPH_SORT_IDX: return n * 4;
        case SPH_CELL_START: case SPH_CELL_END: return (size_t)h->d.K * 4;
        case SPH_BONDS: return nb * 8;
        case SPH_BONDS_DELTA: case SPH_BONDS_OMEGA: return nb * 32;
        case SPH_WALLS_NX: case SPH_WALLS_MVFD: return nw * 32;
        case SPH_WALLS_VEL0: return nw * 8;
        case SPH_WALL_PP: return nw * n * 8;
        case SPH_TIME: return 32;
        default: return 0;
    }
}

int sph_get(sph_t* h, int f, void* dst, size_t bytes) {
    if (!h || !field_is_known(f)) return SPH_E_INVALID;
    const size_t need = sph_field_bytes(h, f);
    if (bytes != need) {
        h->err = "sph_get: field " + std::to_string(f) + " needs " + std::to_string(need) + " bytes";
        return SPH_E_INVALID;
    }
    if (need == 0) return SPH_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    int rc = get_particle_field(h, f, dst, whole(h));
    if (rc <= 0) return rc;
    const DemDev& d = h->d;
    const size_t n = h->hs.np, K = d.K;
    double* dd = static_cast<double*>(dst);
    uint32_t* du = static_cast<uint32_t*>(dst);
    auto copy = [&](const void* src) -> int {
        HIP_TRY(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
        return SPH_OK;
    };
    switch (f) {
        case SPH_X_SORTED: {
            // the last step's input positions (the other set) in its sorted
            // order, wrapped as the step wrapped them
            const sph::PState& in = h->own_cur ? d.b : d.a;
            std::vector<double> xs(4 * n);
            std::vector<uint32_t> perm(n);
            HIP_TRY(hipMemcpy(xs.data(), in.x, n * 32, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(perm.data(), d.perm, n * 4, hipMemcpyDeviceToHost));
            const HostState& hs = h->hs;
            for (size_t q = 0; q < n; ++q) {
                double v[4];
                std::memcpy(v, &xs[4 * (size_t)perm[q]], 32);
                if (hs.periodic == 1 || hs.periodic == 2) {
                    if (v[0] < hs.origo[0]) v[0] += hs.L[0];
                    if (v[0] > hs.L[0]) v[0] -= hs.L[0];
                }
                if (hs.periodic == 1) {
                    if (v[1] < hs.origo[1]) v[1] += hs.L[1];
                    if (v[1] > hs.L[1]) v[1] -= hs.L[1];
                }
                std::memcpy(dd + 4 * q, v, 32);
            }
            return SPH_OK;
        }
        case SPH_SORT_IDX: return copy(d.sid);
        case SPH_CELL_KEYS: case SPH_CELL_START: case SPH_CELL_END: {
            std::vector<uint32_t> off(K + 1);
            HIP_TRY(hipMemcpy(off.data(), d.off, (K + 1) * 4, hipMemcpyDeviceToHost));
            if (f == SPH_CELL_KEYS) {
                std::vector<uint32_t> sid(n);
                HIP_TRY(hipMemcpy(sid.data(), d.sid, n * 4, hipMemcpyDeviceToHost));
                for (size_t c = 0; c < K; ++c)
                    for (uint32_t q = off[c]; q < off[c + 1]; ++q) du[sid[q]] = (uint32_t)c;
                return SPH_OK;
            }
            for (size_t c = 0; c < K; ++c) {
                const bool occ = off[c + 1] > off[c];
                du[c] = occ ? (f == SPH_CELL_START ? off[c] : off[c + 1]) : sph::EMPTY_CELL;
            }
            return SPH_OK;
        }
        case SPH_BONDS: return copy(d.bonds);
        case SPH_BONDS_DELTA: return copy(d.bdelta);
        case SPH_BONDS_OMEGA: return copy(d.bomega);
        case SPH_WALLS_NX: return copy(d.wnx);
        case SPH_WALLS_MVFD: return copy(d.wmvfd);
        case SPH_WALLS_VEL0: return copy(d.wvel0);
        case SPH_WALL_PP: return copy(d.wall_pp);
        case SPH_TIME: {
            const double tm[4] = {h->hs.dt, h->hs.t_cur, h->hs.t_total, h->hs.file_dt};
            std::memcpy(dst, tm, 32);
            return SPH_OK;
        }
        default: break;
    }
    h->err = "sph_get: unknown field";
    return SPH_E_INVALID;
}

}  // extern "C"

namespace {
int set_field(sph_t* h, int f, const void* src, size_t bytes);
}  // namespace

extern "C" {

int sph_set(sph_t* h, int f, const void* src, size_t bytes) {
    const int rc = set_field(h, f, src, bytes);
    if (rc == SPH_OK) refresh_f32(h);
    return rc;
}

}  // extern "C"

namespace {

int set_field(sph_t* h, int f, const void* src, size_t bytes) {
    if (!h || !field_is_known(f)) return SPH_E_INVALID;
    const size_t need = sph_field_bytes(h, f);
    if (bytes != need) {
        h->err = "sph_set: field " + std::to_string(f) + " needs " + std::to_string(need) + " bytes";
        return SPH_E_INVALID;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    DemDev& d = h->d;
    const size_t n = h->hs.np;
    const double* sd = static_cast<const double*>(src);
    auto copy = [&](void* dptr) -> int {
        HIP_TRY(hipMemcpy(dptr, src, need, hipMemcpyHostToDevice));
        return SPH_OK;
    };
    // original-order input -> comps [c0, c0+ncomp) of a storage-order double4 record
    auto scatter = [&](double4* dptr, int c0, int ncomp) -> int {
        std::vector<uint32_t> id;
        int r;
        if ((r = fetch_ids(h, id))) return r;
        std::vector<double> t(4 * n);
        HIP_TRY(hipMemcpy(t.data(), dptr, n * 32, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q)
            for (int a = 0; a < ncomp; ++a) t[4 * q + c0 + a] = sd[(size_t)ncomp * id[q] + a];
        HIP_TRY(hipMemcpy(dptr, t.data(), n * 32, hipMemcpyHostToDevice));
        return SPH_OK;
    };
    switch (f) {
        case SPH_X: {
            const int r = scatter(own(h).x, 0, 4);
            if (r) return r;
            std::memcpy(h->hs.x.data(), sd, n * 32);  // radii for the search margin
            update_search_eps(h);
            destroy_graph(h);
            request_list_rebuild(h);
            return prepare_keys(h);
        }
        case SPH_VEL: return scatter(own(h).vel, 0, 4);
        case SPH_ANGVEL: return scatter(own(h).angvel, 0, 3);
        case SPH_FORCE: return scatter(d.force, 0, 3);
        case SPH_TORQUE: return scatter(d.torque, 0, 3);
        case SPH_ANGPOS: case SPH_VEL0: case SPH_ANGVEL0: case SPH_XYSUM: case SPH_ES: case SPH_EV:
            return set_comps(h, packed_comps(own(h), f), sd);
        case SPH_ES_DOT: return scatter(d.force, 3, 1);
        case SPH_EV_DOT: return scatter(d.torque, 3, 1);
        case SPH_P: {
            std::vector<uint32_t> id;
            int r;
            if ((r = fetch_ids(h, id))) return r;
            std::vector<double> pv(n);
            for (size_t q = 0; q < n; ++q) pv[q] = sd[id[q]];
            HIP_TRY(hipMemcpy(d.pres, pv.data(), n * 8, hipMemcpyHostToDevice));
            return SPH_OK;
        }
        case SPH_BONDS: {
            // new pairs: checked like a .bin's, then the per-particle bond
            // lists the force kernel gathers through are rebuilt from them
            const uint32_t* pairs = static_cast<const uint32_t*>(src);
            int rc;
            if ((rc = check_bond_pairs(h, pairs, h->hs.nb0, h->hs.np))) return rc;
            std::vector<uint32_t> b(pairs, pairs + 2 * (size_t)h->hs.nb0);
            if ((rc = copy(d.bonds))) return rc;
            h->hs.bonds = b;
            return upload_bond_lists(h, b);
        }
        case SPH_BONDS_DELTA: return copy(d.bdelta);
        case SPH_BONDS_OMEGA: return copy(d.bomega);
        case SPH_WALLS_NX: return copy(d.wnx);
        case SPH_WALLS_MVFD: return copy(d.wmvfd);
        case SPH_WALLS_VEL0: return copy(d.wvel0);
        case SPH_TIME: {
            double tm[4];
            std::memcpy(tm, src, 32);
            h->hs.dt = tm[0]; h->hs.t_cur = tm[1]; h->hs.t_total = tm[2]; h->hs.file_dt = tm[3];
            HIP_TRY(hipMemcpy(d.t_cur, &h->hs.t_cur, 8, hipMemcpyHostToDevice));
            if (d.sys.dt != h->hs.dt) { d.sys.dt = h->hs.dt; destroy_graph(h); }
            return SPH_OK;
        }
        default: break;
    }
    h->err = "sph_set: field " + std::to_string(f) + " is read-only";
    return SPH_E_INVALID;
}

}  // namespace

extern "C" {

int sph_info(const sph_t* h, sph_info_t* out) {
    if (!h || !out) return SPH_E_INVALID;
    out->np = h->hs.np; out->nb0 = h->hs.nb0; out->nw = h->hs.nw; out->nc = sph::NC;
    out->ncells = h->d.K;
    for (int a = 0; a < 3; ++a) out->num[a] = h->hs.num[a];
    out->periodic = h->hs.periodic;
    out->dt = h->hs.dt; out->t_current = h->hs.t_cur; out->t_total = h->hs.t_total; out->file_dt = h->hs.file_dt;
    out->steps_done = h->steps_done;
    out->precision = h->precision;
    return SPH_OK;
}

int sph_get_params(const sph_t* h, sph_params_t* out) {
    if (!h || !out) return SPH_E_INVALID;
    out->k_n = h->hs.k_n; out->k_t = h->hs.k_t; out->gamma_n = h->hs.gamma_n;
    out->gamma_t = h->hs.gamma_t; out->mu_s = h->hs.mu_s; out->mu_d = h->hs.mu_d;
    return SPH_OK;
}

int sph_set_params(sph_t* h, const sph_params_t* in) {
    if (!h || !in) return SPH_E_INVALID;
    h->hs.k_n = in->k_n; h->hs.k_t = in->k_t; h->hs.gamma_n = in->gamma_n;
    h->hs.gamma_t = in->gamma_t; h->hs.mu_s = in->mu_s; h->hs.mu_d = in->mu_d;
    fill_params(h);
    destroy_graph(h);
    return SPH_OK;
}

int sph_set_precision(sph_t* h, int precision) {
    if (!h || (precision != SPH_PREC_FP64 && precision != SPH_PREC_MIXED)) return SPH_E_INVALID;
    if (precision == h->precision) return SPH_OK;
    if (precision == SPH_PREC_MIXED && h->hs.contactmodel == 3) {
        h->err = "sph_set_precision: contact model 3 runs in FP64 only";
        return SPH_E_UNSUPPORTED;
    }
    int rc;
    const size_t cnt = (size_t)h->hs.np * sph::NC * 3;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (precision == SPH_PREC_MIXED && h->steps_done == 0) {
        // no step yet: every spring is zero, as the fresh FP32 buffer is
        if ((rc = ensure_mixed_buffers(h))) return rc;
    } else if (precision == SPH_PREC_MIXED) {
        if ((rc = ensure_mixed_buffers(h))) return rc;
        std::vector<double> a(cnt);
        std::vector<float> b(cnt);
        HIP_TRY(hipMemcpy(a.data(), own(h).cdt, cnt * 8, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cnt; ++i) b[i] = (float)a[i];
        HIP_TRY(hipMemcpy(own(h).cdt32, b.data(), cnt * 4, hipMemcpyHostToDevice));
    } else {
        std::vector<double> a(cnt);
        std::vector<float> b(cnt);
        HIP_TRY(hipMemcpy(b.data(), own(h).cdt32, cnt * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < cnt; ++i) a[i] = (double)b[i];
        HIP_TRY(hipMemcpy(own(h).cdt, a.data(), cnt * 8, hipMemcpyHostToDevice));
    }
    h->precision = precision;
    refresh_f32(h);
    destroy_graph(h);
    return SPH_OK;
}

int sph_set_profiling(sph_t* h, int on) {
    if (!h) return SPH_E_INVALID;
    h->profiling = on ? 1 : 0;
    return SPH_OK;
}

int sph_get_timing(const sph_t* h, sph_timing_t* out) {
    if (!h || !out) return SPH_E_INVALID;
    *out = h->timing;
    out->list_builds = 0;
    if (h->d.nstate) {
        if (hipMemcpyAsync(&out->list_builds, h->d.nstate + 3, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess)
            return SPH_E_HIP;
    }
    return SPH_OK;
}

int sph_set_tile(sph_t* h, uint32_t tile_x, uint32_t tile_y) {
    if (!h) return SPH_E_INVALID;
    choose_tile(h, tile_x, tile_y);
    destroy_graph(h);
    return SPH_OK;
}

int sph_set_search(sph_t* h, int mode) {
    if (!h || mode < 0 || mode > 5) return SPH_E_INVALID;
    h->search_auto = mode == 0;
    h->tuned = false;
    h->d.search_mode = mode == 0 ? 2 : mode;
    if (mode == 5) request_list_rebuild(h);
    destroy_graph(h);
    return SPH_OK;
}

int sph_set_force_kernel(sph_t* h, int variant) {
    if (!h || variant < 0 || variant > 2) return SPH_E_INVALID;
    h->force_auto = variant == 0;
    h->tuned = false;
    h->d.force_vw = variant == 2 ? 1u : 0u;
    destroy_graph(h);
    return SPH_OK;
}

int sph_set_graph(sph_t* h, int on) {
    if (!h) return SPH_E_INVALID;
    h->use_graph = on != 0;
    if (!h->use_graph) destroy_graph(h);
    return SPH_OK;
}

int sph_bandwidth_probe(uint64_t bytes, int reps, double* gbs) {
    if (bytes < 64 || reps < 1 || !gbs) return SPH_E_INVALID;
    const uint64_t n = bytes / 32;
    double4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPH_OK;
    if (hipMalloc(&a, n * 32) != hipSuccess || hipMalloc(&b, n * 32) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemset(a, 0, n * 32) != hipSuccess) {
        rc = SPH_E_HIP;
    } else {
        sph::launch_copy_probe(a, b, n, st);  // warm up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) sph::launch_copy_probe(a, b, n, st);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = SPH_E_HIP;
        else *gbs = 2.0 * (double)(n * 32) * reps / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int sph_flops_probe(int fp64, uint32_t blocks, uint32_t iters, int reps, double* tflops) {
    if (blocks == 0 || iters == 0 || reps < 1 || !tflops) return SPH_E_INVALID;
    const size_t elem = fp64 ? 8 : 4;
    void* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPH_OK;
    if (hipMalloc(&out, (size_t)blocks * sph::BLOCK * elem) != hipSuccess ||
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = SPH_E_HIP;
    } else {
        sph::launch_flops_probe(out, blocks, iters, fp64 != 0, st);  // warm up
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < reps; ++r) sph::launch_flops_probe(out, blocks, iters, fp64 != 0, st);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = SPH_E_HIP;
        else *tflops = 16.0 * iters * (double)blocks * sph::BLOCK * reps / (ms * 1e-3) / 1e12;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}

}  // extern "C"

// ----------------------------------------------------------------- ensembles

struct sph_ens {
    sph_dem* h = nullptr;
    double cell_factor = -1.0;
    double tuned_ms[4] = {0.0, 0.0, 0.0, 0.0};
    std::string err;
};

namespace {

thread_local std::string g_ens_error;
constexpr double CELL_FACTORS[4] = {1.0, 1.25, 1.5, 2.0};

// Cells of factor x 2 r_max on every axis (at least one cell, three on a
// periodic axis).
void regrid(HostState& m, double factor) {
    const double edge = factor * 2.0 * max_radius(m);
    for (int a = 0; a < 3; ++a) {
        double c = edge > 0.0 ? std::floor(m.L[a] / edge) : 1.0;
        const bool per = (a == 0 && (m.periodic == 1 || m.periodic == 2)) || (a == 1 && m.periodic == 1);
        c = std::max(c, per ? 3.0 : 1.0);
        m.num[a] = (uint32_t)std::min(c, 4.0e9);
    }
}

// The members' particles back to back, in member order.
void concatenate(const std::vector<HostState>& ms, HostState& out) {
    out = HostState{};
    const HostState& f = ms[0];
    out.dt = f.dt; out.t_cur = f.t_cur; out.t_total = f.t_total; out.file_dt = f.file_dt;
    for (int a = 0; a < 3; ++a) { out.origo[a] = f.origo[a]; out.L[a] = f.L[a]; out.num[a] = f.num[a]; out.g[a] = f.g[a]; }
    out.periodic = f.periodic;
    out.k_n = f.k_n; out.k_t = f.k_t; out.gamma_n = f.gamma_n; out.gamma_t = f.gamma_t;
    out.mu_s = f.mu_s; out.mu_d = f.mu_d; out.gamma_wn = f.gamma_wn; out.gamma_wt = f.gamma_wt;
    out.mu_ws = f.mu_ws; out.mu_wd = f.mu_wd; out.rho = f.rho; out.kappa = f.kappa;
    for (const HostState& m : ms) {
        out.np += m.np;
        out.nw = std::max(out.nw, m.nw);
        auto app = [](std::vector<double>& dst, const std::vector<double>& src) {
            dst.insert(dst.end(), src.begin(), src.end());
        };
        app(out.x, m.x); app(out.xysum, m.xysum); app(out.vel, m.vel); app(out.force, m.force);
        app(out.angpos, m.angpos); app(out.angvel, m.angvel); app(out.torque, m.torque);
        app(out.es_dot, m.es_dot); app(out.es, m.es); app(out.ev_dot, m.ev_dot); app(out.ev, m.ev); app(out.p, m.p);
    }
}

void drop_particles(HostState& m) {
    for (std::vector<double>* v : {&m.x, &m.xysum, &m.vel, &m.force, &m.angpos, &m.angvel, &m.torque, &m.es_dot,
                                   &m.es, &m.ev_dot, &m.ev, &m.p}) {
        v->clear();
        v->shrink_to_fit();
    }
}

// Opens an ensemble on `ms` (already on their final grids).
int ens_open(std::vector<HostState> ms, int verbosity, sph_ens** out, std::string& err) {
    *out = nullptr;
    auto* h = new sph_dem();
    h->members = std::move(ms);
    for (size_t k = 0; k < h->members.size(); ++k) {
        std::string e;
        int rc = sph::check_values(h->members[k], e);
        if (rc) {
            err = "member " + std::to_string(k) + ": " + e;
            delete h;
            return rc;
        }
    }
    concatenate(h->members, h->hs);
    h->blocks_auto = true;
    int rc = open_common(h, verbosity, 0);
    if (rc) {
        err = h->err;
        free_all(h);
        delete h;
        return rc;
    }
    for (HostState& m : h->members) drop_particles(m);
    drop_particles(h->hs);
    auto* e = new sph_ens();
    e->h = h;
    *out = e;
    return SPH_OK;
}

// ms per step of an ensemble on `ms` regridded by `factor`: the autotuner's
// steps (search kernel and block sizes) and then ten graph-replayed steps.
int trial_ms(const std::vector<HostState>& ms, double factor, double& ms_step, std::string& err) {
    std::vector<HostState> t = ms;
    for (HostState& m : t) regrid(m, factor);
    sph_ens* e = nullptr;
    int rc = ens_open(std::move(t), 0, &e, err);
    if (rc) return rc;
    sph_dem* h = e->h;
    rc = run_steps(h, 4);  // tunes the kernels on a copy, then warms up
    hipEvent_t a = nullptr, b = nullptr;
    if (!rc && (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)) rc = SPH_E_HIP;
    float ms10 = 0.0f;
    if (!rc) {
        (void)hipEventRecord(a, h->stream);
        rc = run_steps(h, 10);
        (void)hipEventRecord(b, h->stream);
        if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms10, a, b) != hipSuccess))
            rc = SPH_E_HIP;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) err = h->err;
    ms_step = ms10 / 10.0;
    sph_ens_close(e);
    return rc;
}

int ens_open_states(std::vector<HostState> ms, double cell_factor, int verbosity, sph_ens_t** out) {
    *out = nullptr;
    std::string err;
    if (ms.empty() || ms.size() > 65535) {
        g_ens_error = "an ensemble has 1 to 65535 members";
        return SPH_E_INVALID;
    }
    if (cell_factor > 0.0 && cell_factor < 1.0) {
        g_ens_error = "cell_factor below 1 would make cells narrower than the largest pair";
        return SPH_E_INVALID;
    }
    double tuned[4] = {0.0, 0.0, 0.0, 0.0};
    if (cell_factor == 0.0) {
        int best = -1;
        for (int k = 0; k < 4; ++k) {
            double t = 0.0;
            const int rc = trial_ms(ms, CELL_FACTORS[k], t, err);
            if (rc) continue;  // e.g. too few cells on a periodic axis at this size
            tuned[k] = t;
            if (best < 0 || t < tuned[best]) best = k;
        }
        if (best < 0) {
            g_ens_error = "autotuner: no cell factor gives a valid ensemble: " + err;
            return SPH_E_INVALID;
        }
        cell_factor = CELL_FACTORS[best];
        if (verbosity > 0)
            std::printf("sphere_dem: ensemble cell factor %.2f (ms/step %.4f %.4f %.4f %.4f)\n", cell_factor,
                        tuned[0], tuned[1], tuned[2], tuned[3]);
    }
    if (cell_factor >= 1.0)
        for (HostState& m : ms) regrid(m, cell_factor);
    const int rc = ens_open(std::move(ms), verbosity, out, err);
    if (rc) {
        g_ens_error = err;
        return rc;
    }
    (*out)->cell_factor = cell_factor;
    std::copy(tuned, tuned + 4, (*out)->tuned_ms);
    return SPH_OK;
}

}  // namespace

extern "C" {

int sph_ens_open_mem(const void* const* images, const size_t* bytes, uint32_t n, double cell_factor, int verbosity,
                     sph_ens_t** out) {
    if (!out) return SPH_E_INVALID;
    *out = nullptr;
    if (!images || !bytes) return SPH_E_INVALID;
    std::vector<HostState> ms(n);
    for (uint32_t k = 0; k < n; ++k) {
        std::string e;
        const int rc = sph::parse_bin(images[k], bytes[k], ms[k], e);
        if (rc) {
            g_ens_error = "member " + std::to_string(k) + ": " + e;
            return rc;
        }
    }
    return ens_open_states(std::move(ms), cell_factor, verbosity, out);
}

int sph_ens_open(const char* const* paths, uint32_t n, double cell_factor, int verbosity, sph_ens_t** out) {
    if (!out) return SPH_E_INVALID;
    *out = nullptr;
    if (!paths) return SPH_E_INVALID;
    std::vector<HostState> ms(n);
    for (uint32_t k = 0; k < n; ++k) {
        std::string e;
        const int rc = sph::read_bin_file(paths[k], ms[k], e);
        if (rc) {
            g_ens_error = "member " + std::to_string(k) + ": " + e;
            return rc;
        }
    }
    return ens_open_states(std::move(ms), cell_factor, verbosity, out);
}

void sph_ens_close(sph_ens_t* e) {
    if (!e) return;
    if (e->h) {
        sph_close(e->h);
        e->h = nullptr;
    }
    delete e;
}

const char* sph_ens_last_error(const sph_ens_t* e) { return e ? e->h->err.c_str() : g_ens_error.c_str(); }

int sph_ens_step(sph_ens_t* e, uint64_t nsteps) {
    if (!e) return SPH_E_INVALID;
    return run_steps(e->h, nsteps);
}

size_t sph_ens_field_bytes(const sph_ens_t* e, uint32_t k, int f) {
    if (!e || k >= e->h->sys_host.size() || !field_is_known(f)) return 0;
    const sph::Sys& S = e->h->sys_host[k];
    const size_t n = S.np, nw = S.nw, nc = sph::NC;
    switch (f) {
        case SPH_X: case SPH_VEL: return n * 32;
        case SPH_ANGVEL: case SPH_FORCE: case SPH_TORQUE: case SPH_ANGPOS: case SPH_VEL0:
        case SPH_ANGVEL0: return n * 24;
        case SPH_XYSUM: return n * 16;
        case SPH_ES_DOT: case SPH_ES: case SPH_EV_DOT: case SPH_EV: case SPH_P: return n * 8;
        case SPH_CONTACTS: return n * nc * 4;
        case SPH_DELTA_T: return n * nc * 24;
        case SPH_WALLS_NX: case SPH_WALLS_MVFD: return nw * 32;
        case SPH_WALLS_VEL0: return nw * 8;
        case SPH_WALL_PP: return nw * n * 8;
        case SPH_TIME: return 32;
        default: return 0;
    }
}

int sph_ens_get(sph_ens_t* e, uint32_t k, int f, void* dst, size_t bytes) {
    if (!e || !dst) return SPH_E_INVALID;
    sph_dem* h = e->h;
    if (k >= h->sys_host.size()) {
        h->err = "sph_ens_get: no member " + std::to_string(k);
        return SPH_E_INVALID;
    }
    const size_t need = sph_ens_field_bytes(e, k, f);
    if (need == 0) {
        h->err = "sph_ens_get: field " + std::to_string(f) + " is not available per member";
        return SPH_E_UNSUPPORTED;
    }
    if (bytes != need) {
        h->err = "sph_ens_get: field " + std::to_string(f) + " needs " + std::to_string(need) + " bytes";
        return SPH_E_INVALID;
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    const sph::Sys& S = h->sys_host[k];
    const size_t n = S.np, pb = S.pbase, N = h->d.np;
    char* out = static_cast<char*>(dst);
    switch (f) {
        case SPH_WALLS_NX: HIP_TRY(hipMemcpy(dst, h->d.wnx + S.wbase, need, hipMemcpyDeviceToHost)); return SPH_OK;
        case SPH_WALLS_MVFD: HIP_TRY(hipMemcpy(dst, h->d.wmvfd + S.wbase, need, hipMemcpyDeviceToHost)); return SPH_OK;
        case SPH_WALLS_VEL0: HIP_TRY(hipMemcpy(dst, h->d.wvel0 + S.wbase, need, hipMemcpyDeviceToHost)); return SPH_OK;
        case SPH_WALL_PP:
            for (uint32_t w = 0; w < S.nw; ++w)
                HIP_TRY(hipMemcpy(out + (size_t)w * n * 8, h->d.wall_pp + (size_t)w * N + pb, n * 8,
                                  hipMemcpyDeviceToHost));
            return SPH_OK;
        case SPH_TIME: {
            const HostState& m = h->members[k];
            double tm[4] = {m.dt, 0.0, m.t_total, m.file_dt};
            HIP_TRY(hipMemcpy(&tm[1], h->d.t_cur + k, 8, hipMemcpyDeviceToHost));
            std::memcpy(dst, tm, 32);
            return SPH_OK;
        }
        default: break;
    }
    // per-particle fields: the member's own storage range (its particles'
    // original indices and sorted positions are [pbase, pbase + np)), its
    // contact partners in its own numbering
    const int rc = get_particle_field(h, f, out, Range{pb, n});
    if (rc > 0) {
        h->err = "sph_ens_get: field " + std::to_string(f) + " is not available per member";
        return SPH_E_UNSUPPORTED;
    }
    return rc;
}

int sph_ens_member_info(const sph_ens_t* e, uint32_t k, sph_info_t* out) {
    if (!e || !out || k >= e->h->sys_host.size()) return SPH_E_INVALID;
    const sph_dem* h = e->h;
    const sph::Sys& S = h->sys_host[k];
    const HostState& m = h->members[k];
    out->np = S.np; out->nb0 = 0; out->nw = S.nw; out->nc = sph::NC; out->ncells = S.K;
    for (int a = 0; a < 3; ++a) out->num[a] = S.num[a];
    out->periodic = S.periodic;
    out->dt = m.dt;
    out->t_current = m.t_cur;  // time.current of the member, accumulated on the device
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(&out->t_current, h->d.t_cur + k, 8, hipMemcpyDeviceToHost) != hipSuccess)
        return SPH_E_HIP;
    out->t_total = m.t_total; out->file_dt = m.file_dt;
    out->steps_done = h->steps_done;
    out->precision = h->precision;
    return SPH_OK;
}

int sph_ens_info(const sph_ens_t* e, sph_ens_info_t* out) {
    if (!e || !out) return SPH_E_INVALID;
    const sph_dem* h = e->h;
    out->members = (uint32_t)h->sys_host.size();
    out->np = h->d.np; out->ncells = h->d.K; out->steps_done = h->steps_done;
    out->cell_factor = e->cell_factor;
    for (int k = 0; k < 4; ++k) out->tuned_ms[k] = e->tuned_ms[k];
    out->search_mode = h->d.search_mode;
    out->force_block = h->d.tpb_force; out->search_block = h->d.tpb_search;
    out->precision = h->precision;
    return SPH_OK;
}

int sph_ens_set_precision(sph_ens_t* e, int precision) { return e ? sph_set_precision(e->h, precision) : SPH_E_INVALID; }

int sph_ens_set_search(sph_ens_t* e, int mode) {
    if (!e) return SPH_E_INVALID;
    if (mode != 0 && mode != 2 && mode != 4 && mode != 5) {
        e->h->err = "ensembles use search mode 0, 2, 4 or 5";
        return SPH_E_INVALID;
    }
    e->h->blocks_auto = mode == 0;
    return sph_set_search(e->h, mode);
}

int sph_ens_set_blocks(sph_ens_t* e, uint32_t force_block, uint32_t search_block) {
    if (!e) return SPH_E_INVALID;
    auto ok = [](uint32_t b) { return b >= 64 && b <= (uint32_t)sph::BLOCK && b % 64 == 0; };
    if (!ok(force_block) || !ok(search_block)) {
        e->h->err = "threads per block: a multiple of 64 up to 256";
        return SPH_E_INVALID;
    }
    e->h->blocks_auto = false;
    e->h->d.tpb_force = force_block;
    e->h->d.tpb_search = search_block;
    destroy_graph(e->h);
    return SPH_OK;
}

int sph_ens_set_graph(sph_ens_t* e, int on) { return e ? sph_set_graph(e->h, on) : SPH_E_INVALID; }
int sph_ens_set_profiling(sph_ens_t* e, int on) { return e ? sph_set_profiling(e->h, on) : SPH_E_INVALID; }
int sph_ens_get_timing(const sph_ens_t* e, sph_timing_t* out) { return e ? sph_get_timing(e->h, out) : SPH_E_INVALID; }

}  // extern "C"
