// kernels.hip -- gfx950 kernels of one DEM time step.
//
// Step order (one HIP stream, replayed as a HIP graph):
//   bonds      one thread per bond: parallel-bond force, torque, breakage
//              (src/cohesion.cuh:8-233); forces go to per-bond records, no
//              scatter
//   (cell ids  the previous step's force kernel computed them with the
//              per-cell counts, src/sorting.cuh:7-43)
//   scan       exclusive scan of the per-cell counts -> cell offsets
//   scatter    counting-sort placement of the particles
//   cellsort   one thread per cell: order each cell's members by original
//              index, which makes the counting sort the stable sort of
//              src/device.cu:763-765
//   reorder    gather every per-particle record into the new sorted order
//              (src/sorting.cuh:49-121 gathers x, vel, angvel only)
//   contact    one thread per sorted particle: 27-cell search and slot
//              update (src/contactsearch.cuh:181-355), contact and wall
//              forces (:369-689, src/contactmodels.cuh:10-109,283-493),
//              bond forces gathered in ascending bond index, leapfrog
//              (src/integration.cuh:9-187)
//   walls      per-wall tree sum of per-particle wall forces in the
//              reference's block order, wall leapfrog
//              (src/integration.cuh:191-295)
//
// Arithmetic keeps the reference's expression order and is built with
// -ffp-contract=off, so the FP64 path reproduces the CPU oracle bit for bit.
// Vector lengths use IEEE double sqrt (the reference narrows to float).

#include <hip/hip_runtime.h>

#include "dem_internal.h"
#include "vec3.h"

namespace sph {

namespace {

constexpr double PI_REF = 3.14159265358979;   // src/constants.h:14
constexpr double PI_WALL = 3.141596654;       // src/integration.cuh:258


// Workgroups are dispatched round-robin over the 8 XCDs; remap so each XCD
// walks one contiguous range of sorted particles and its L2 holds that
// range's neighbourhood.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    const uint32_t q = nb >> 3, r = nb & 7u, xcd = b & 7u, k = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + k;
}

// The system of index / sorted position i: the handle's only one, or its
// entry of the ensemble table.
template <bool B>
__device__ __forceinline__ const Sys& sys_at(const DemDev& d, uint32_t i) {
    if constexpr (B) return d.systems[d.sys_of[i]];
    else return d.sys;
}

// ------------------------------------------------------------ cell hashing
// Periodic wrap with the rule of src/integration.cuh:80-95.
__device__ __forceinline__ bool wrap_xy(const Sys& S, double& x, double& y) {
    bool changed = false;
    if (S.periodic == 1 || S.periodic == 2) {
        if (x < S.origo[0]) { x += S.L[0]; changed = true; }
        if (x > S.L[0]) { x -= S.L[0]; changed = true; }
    }
    if (S.periodic == 1) {
        if (y < S.origo[1]) { y += S.L[1]; changed = true; }
        if (y > S.L[1]) { y -= S.L[1]; changed = true; }
    }
    return changed;
}

// Start-of-step position of sorted particle q: its stored record at the
// previous position perm[q], wrapped as the step's integrate wraps it.
__device__ __forceinline__ double4 spos(const DemDev& d, const Sys& S, uint32_t q) {
    double4 x = d.ro.x[d.perm[q]];
    (void)wrap_xy(S, x.x, x.y);
    return x;
}

// The same for non-inlined helpers, which take this small view by value
// instead of the DemDev.
struct PosView {
    const double4* x;
    const uint32_t* perm;
    double ox, oy, Lx, Ly;
    int periodic;
};

__device__ __forceinline__ PosView pos_view(const DemDev& d, const Sys& S) {
    PosView v;
    v.x = d.ro.x; v.perm = d.perm;
    v.ox = S.origo[0]; v.oy = S.origo[1]; v.Lx = S.L[0]; v.Ly = S.L[1];
    v.periodic = S.periodic;
    return v;
}

__device__ __forceinline__ double4 pv_pos(const PosView& v, uint32_t q) {
    double4 x = v.x[v.perm[q]];
    if (v.periodic == 1 || v.periodic == 2) {
        if (x.x < v.ox) x.x += v.Lx;
        if (x.x > v.Lx) x.x -= v.Lx;
    }
    if (v.periodic == 1) {
        if (x.y < v.oy) x.y += v.Ly;
        if (x.y > v.Ly) x.y -= v.Ly;
    }
    return x;
}

// floor((v - origo)/(L/num)) (src/sorting.cuh:12-14), clamped to the grid.
__device__ __forceinline__ int cell_coord(const Sys& S, double v, int a, uint32_t& err) {
    const double f = floor((v - S.origo[a]) / S.cell_len[a]);
    if (f != f) { err |= FLAG_NAN; return 0; }
    if (f < -1.0 || f > (double)S.num[a]) err |= FLAG_OUT_OF_GRID;
    if (f < 0.0) return 0;
    if (f > (double)(S.num[a] - 1)) return (int)(S.num[a] - 1);
    return (int)f;
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Cell id of a position at the start of a step: the periodic wrap the step's
// integrate applies to it (src/integration.cuh:80-95) on a copy, then
// src/sorting.cuh:7-19.  The position itself is wrapped when it is gathered
// into the sorted order.
__device__ __forceinline__ uint32_t cell_key(const Sys& S, double x, double y, double z, uint32_t& err) {
    (void)wrap_xy(S, x, y);
    const int ix = cell_coord(S, x, 0, err);
    const int iy = cell_coord(S, y, 1, err);
    const int iz = cell_coord(S, z, 2, err);
    return S.cbase + (((uint32_t)iz * S.num[1]) * S.num[0] + (uint32_t)iy * S.num[0] + (uint32_t)ix);
}

// Per-cell counts and each particle's rank among its cell's members (in any
// order: the cell sort orders members by original index afterwards), taken
// where the keys are computed (k_force for the next step, k_cellid after an
// upload).  Storage is the previous step's sorted order, so a wavefront's
// keys come in runs; one atomic per run.  Every lane of the wavefront below
// the grid's end calls it (the active lanes are a prefix).
__device__ __forceinline__ void count_key(const DemDev& d, uint32_t i, uint32_t key) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0u || prev != key;
    const uint64_t heads = __ballot(head);
    const uint64_t acts = __ballot(true);
    const uint64_t upto = (2ull << lane) - 1ull;  // lanes 0..lane (all 64 for lane 63)
    const uint32_t start = 63u - (uint32_t)__clzll((long long)(heads & upto));
    uint32_t base = 0;
    if (head) {
        const uint64_t above = heads & ~upto;
        const uint32_t end = above ? (uint32_t)__ffsll((long long)above) - 1u : 64u - (uint32_t)__clzll((long long)acts);
        base = atomicAdd(&d.count[key], end - lane);
    }
    base = __shfl(base, (int)start, 64);
    d.rank[i] = base + (lane - start);
}

// Cell ids of the stored positions, for a state that did not come out of a
// step (after open or a position update); steps compute the next step's ids
// in the force kernel.
template <bool B>
__global__ __launch_bounds__(BLOCK) void k_cellid(DemDev d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.np) return;
    const Sys& S = sys_at<B>(d, i);
    const double4 x = d.ro.x[i];
    uint32_t err = 0;
    const uint32_t key = cell_key(S, x.x, x.y, x.z, err);
    d.key[i] = key;
    if (err) atomicOr(d.flags_next, err);
    count_key(d, i, key);
}

// -------------------------------------------------------------------- scan
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds, uint32_t& total) {
    // wave-level inclusive scan with shuffles, then across the 4 waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        const uint32_t ws = lds[w];
        if (w < wave) woff += ws;
        tot += ws;
    }
    __syncthreads();
    total = tot;
    return woff + inc - v;
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(DemDev d) {
    __shared__ uint32_t lds[BLOCK / 64];
    // the cell-id errors belong to this step; so does a neighbour-list
    // rebuild requested by the last step's force kernel (or a failed build)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t f = *d.flags_next;
        if (f) {
            atomicOr(d.flags, f);
            *d.flags_next = 0u;
        }
        const uint32_t build = d.nstate[1] | d.nstate[2];
        d.nstate[0] = build;
        if (build && d.search_mode == 5) d.nstate[3] += 1u;  // list builds so far
        d.nstate[1] = 0u;
        d.nstate[2] = 0u;
        d.nstate[4] = d.nstate[5];  // a sphere crossed a periodic face in the last step
        d.nstate[5] = 0u;
    }
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_PER_THREAD; ++j)
        if (base + j < d.K) sum += d.count[base + j];
    uint32_t total;
    (void)block_exclusive_scan(sum, lds, total);
    if (threadIdx.x == 0) d.scan_bsum[blockIdx.x] = total;
}

// off[c] = first sorted position of cell c, off[K] = np.  Each block adds
// the totals of the blocks before it (integer sums: any order).
__global__ __launch_bounds__(BLOCK) void k_scan_down(DemDev d) {
    __shared__ uint32_t lds[BLOCK / 64];
    uint32_t before = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += BLOCK) before += d.scan_bsum[b];
    uint32_t prefix;
    (void)block_exclusive_scan(before, lds, prefix);
    const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_PER_THREAD;
    uint32_t c[SCAN_PER_THREAD];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        c[j] = base + j < d.K ? d.count[base + j] : 0u;
        sum += c[j];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(sum, lds, total) + prefix;
#pragma unroll
    for (int j = 0; j < SCAN_PER_THREAD; ++j) {
        if (base + j < d.K) {
            d.off[base + j] = off;
            d.count[base + j] = 0u;  // zero for the next step's counts
        }
        off += c[j];
        if (base + j + 1 == d.K) d.off[d.K] = off;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scatter(DemDev d) {
    const uint32_t i = xcd_block() * blockDim.x + threadIdx.x;
    if (i >= d.np) return;
    const uint32_t q = d.off[d.key[i]] + d.rank[i];
    if (q < d.np) d.perm[q] = i;  // (counts and keys agree by construction; never write past perm)
}

// Order one cell's members by original index: odd-even transposition network
// on (id, position) pairs in registers for up to W members, insertion sort in
// memory beyond that.
template <int W>
__device__ __forceinline__ void sort_small(const DemDev& d, uint32_t* p, uint32_t n) {
    uint64_t v[W];
#pragma unroll
    for (int j = 0; j < W; ++j)
        v[j] = (uint32_t)j < n ? ((uint64_t)d.ro.id[p[j]] << 32) | p[j] : ~0ull;
#pragma unroll
    for (int r = 0; r < W; ++r) {
#pragma unroll
        for (int j = (r & 1); j + 1 < W; j += 2) {
            const uint64_t a = v[j], b = v[j + 1];
            v[j] = a < b ? a : b;
            v[j + 1] = a < b ? b : a;
        }
    }
#pragma unroll
    for (int j = 0; j < W; ++j)
        if ((uint32_t)j < n) p[j] = (uint32_t)v[j];
}

__global__ __launch_bounds__(BLOCK) void k_cellsort(DemDev d) {
    const uint32_t c = xcd_block() * blockDim.x + threadIdx.x;
    if (c >= d.K) return;
    const uint32_t st = d.off[c];
    const uint32_t n = d.off[c + 1] - st;
    if (n < 2) return;
    uint32_t* p = d.perm + st;
    if (n <= 4) sort_small<4>(d, p, n);
    else if (n <= 8) sort_small<8>(d, p, n);
    else if (n <= 16) sort_small<16>(d, p, n);
    else {
        for (uint32_t i = 1; i < n; ++i) {
            const uint32_t v = p[i];
            const uint32_t key = d.ro.id[v];
            uint32_t j = i;
            while (j > 0 && d.ro.id[p[j - 1]] > key) { p[j] = p[j - 1]; --j; }
            p[j] = v;
        }
    }
}

template <bool B>
__global__ __launch_bounds__(BLOCK) void k_reorder(DemDev d) {
    const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
    if (q >= d.np) return;
    const Sys& S = sys_at<B>(d, q);
    const uint32_t p = d.perm[q];
    double4 x = d.ro.x[p];
    (void)wrap_xy(S, x.x, x.y);
    d.xf[q] = make_float4((float)x.x, (float)x.y, (float)x.z, (float)x.w);
    const uint32_t id = d.ro.id[p];
    d.sid[q] = id;
    d.skey[q] = d.key[p];
    if (d.search_mode == 5) {
        const uint32_t li = d.ro.lidx[p];
        d.linv[li] = q;
        d.slidx[q] = li;
    }
}

// ------------------------------------------------------------------- bonds
// src/cohesion.cuh:8-233.  The force and torque of an intact bond go to
// bond_f/bond_t; the contact kernel adds them per particle in ascending bond
// index (the reference adds them with racing += at :220-226).
__global__ __launch_bounds__(BLOCK) void k_bonds(DemDev d) {
    const uint32_t k = xcd_block() * blockDim.x + threadIdx.x;
    if (k >= d.nb0) return;
    const Sys& S = d.sys;
    const uint2 bond = d.bonds[k];
    if (!(bond.x < d.np)) return;
    const double dt = S.dt;
    const double4 d0 = d.bdelta[k], o0 = d.bomega[k];
    V3<double> delta0_t = mk(d0.x, d0.y, d0.z);
    const double delta0_n = d0.w;
    V3<double> omega0_t = mk(o0.x, o0.y, o0.z);
    const double omega0_n = o0.w;
    const uint32_t pi = d.pos_of[bond.x], pj = d.pos_of[bond.y];
    const double4 x_i = d.ro.x[pi], x_j = d.ro.x[pj];
    const double4 vel_i = d.ro.vel[pi], vel_j = d.ro.vel[pj];
    const V3<double> angvel_i = xyz(d.ro.angvel[pi]), angvel_j = xyz(d.ro.angvel[pj]);
    const double R_bar = d.lambda_bar * fmin(x_i.w, x_j.w);
    const double A = PI_REF * R_bar * R_bar;
    const double I = 0.25 * PI_REF * R_bar * R_bar * R_bar * R_bar;
    const double J = I * 2.0;
    const V3<double> x = mk(x_i.x - x_j.x, x_i.y - x_j.y, x_i.z - x_j.z);
    const double x_length = vlen(x);
    const double overlap = fmin(0.0, x_length - (x_i.w + x_j.w));
    const V3<double> n = x / x_length;
    delta0_t = delta0_t - (n * dot(n, delta0_t));
    const V3<double> ddelta = (mk(vel_i.x - vel_j.x, vel_i.y - vel_j.y, vel_i.z - vel_j.z)
                               + (x_i.w + overlap / 2.0) * cross(n, angvel_i)
                               + (x_j.w + overlap / 2.0) * cross(n, angvel_j)) * dt;
    const double ddelta_n = -dot(ddelta, n);
    const double delta_n = delta0_n + ddelta_n;
    const V3<double> ddelta_t = ddelta - n * dot(n, ddelta);
    const V3<double> delta_t = delta0_t + ddelta_t;
    const V3<double> f_n = (S.k_n * A * delta_n + S.gamma_n * ddelta_n / dt) * n;
    const V3<double> f_t = -S.k_t * A * delta_t - S.gamma_t * ddelta_t / dt;
    const V3<double> f = f_n + f_t;
    const V3<double> mn = -n;
    omega0_t = omega0_t - mn * dot(omega0_t, mn);
    const V3<double> domega = mk(angvel_j.x - angvel_i.x, angvel_j.y - angvel_i.y, angvel_j.z - angvel_i.z) * dt;
    const double domega_n = dot(domega, mn);
    const double omega_n = omega0_n + domega_n;
    const V3<double> domega_t = domega - domega_n * mn;
    const V3<double> omega_t = omega0_t + domega_t;
    const V3<double> t_n = -S.k_t * J * omega_n * mn;
    const V3<double> t_t = -S.k_n * I * omega_t;
    const V3<double> t_j = t_n + t_t;
    const double sigma_max = vlen(f_n) / A + vlen(t_t) * R_bar / I;
    const double tau_max = vlen(f_t) / A + vlen(t_n) * R_bar / J;
    if (sigma_max >= d.sigma_b || tau_max >= d.tau_b) {
        d.bonds[k].x = d.bond_broken;
        return;
    }
    d.bdelta[k] = make_double4(delta_t.x, delta_t.y, delta_t.z, delta_n);
    d.bomega[k] = make_double4(omega_t.x, omega_t.y, omega_t.z, omega_n);
    d.bond_f[k] = make_double4(f.x, f.y, f.z, 0.0);
    d.bond_t[k] = make_double4(t_j.x, t_j.y, t_j.z, 0.0);
}

// --------------------------------------------------------- contact models
struct Acc {
    V3<double> F, T;
    double es_dot, ev_dot, p;
};

// src/contactmodels.cuh:283-493 in arithmetic T (double: the reference's
// FP64 path; float: the mixed mode's pair force).  Sums go to FP64.
template <class T>
__device__ __forceinline__ void contact_linear(const Sys& S, Acc& acc, V3<T> vel_a, V3<T> angvel_a,
                                               V3<T> vel_b, V3<T> angvel_b, T radius_a, T radius_b,
                                               V3<T> x, T x_length, T delta, V3<T>& delta_t_io) {
    const T k_n = (T)S.k_n, k_t = (T)S.k_t, gamma_n = (T)S.gamma_n, gamma_t = (T)S.gamma_t;
    const T mu_s = (T)S.mu_s, mu_d = (T)S.mu_d, kappa = (T)S.kappa, dt = (T)S.dt;
    const V3<T> delta_t0_uncor = delta_t_io;
    const V3<T> n = x / x_length;
    const V3<T> vel_linear = vel_a - vel_b;
    const V3<T> vel = vel_linear + (radius_a + delta / (T)2.0) * cross(n, angvel_a)
                      + (radius_b + delta / (T)2.0) * cross(n, angvel_b);
    const T vel_n = -dot(vel_linear, n);
    const V3<T> vel_t = vel - n * dot(n, vel);
    const T vel_t_length = vlen(vel_t);
    const V3<T> delta_t0 = delta_t0_uncor - (n * dot(n, delta_t0_uncor));
    const T delta_t0_length = vlen(delta_t0);
    V3<T> delta_t = delta_t0;
    V3<T> f_n = (-k_n * delta - gamma_n * vel_n) * n;
    acc.ev_dot += (double)(gamma_n * vel_n * vel_n);
    if (dot(f_n, n) < (T)0.0) f_n = mk<T>(0, 0, 0);
    const T f_n_length = vlen(f_n);
    const V3<T> f_c = (-kappa * (T)sqrt(radius_a * radius_b)) * n;
    V3<T> f_t = mk<T>(0, 0, 0);
    if (delta_t0_length > (T)0.0 || vel_t_length > (T)0.0) {
        delta_t = delta_t0 + vel_t * dt;
        const V3<T> f_t_elast = -k_t * delta_t;
        const V3<T> f_t_visc = -gamma_t * vel_t;
        f_t = f_t_elast + f_t_visc;
        const T f_t_length = vlen(f_t);
        const T f_t_limit = mu_s * vlen(f_n - f_c);
        delta_t = delta_t0 + vel_t * dt;
        if (f_t_length > f_t_limit) {
            const V3<T> t = f_t / vlen(f_t);
            f_t = f_t_limit * t;
            delta_t = (T)-1.0 / k_t * (mu_d * vlen(f_n - f_c) * t + gamma_t * vel_t);
            acc.es_dot += (double)(vlen(vlen(f_t) * vel_t * dt) / dt);
        }
    }
    const V3<T> fsum = f_n + f_t + f_c;
    acc.F = acc.F + mk<double>(fsum.x, fsum.y, fsum.z);
    const V3<T> tq = cross(-(radius_a + delta * (T)0.5) * n, f_t);
    acc.T = acc.T + mk<double>(tq.x, tq.y, tq.z);
    acc.p += (double)(f_n_length / ((T)4.0 * (T)PI_REF * radius_a * radius_a));
    delta_t_io = delta_t;
}

// src/contactmodels.cuh:499-666 contactHertz (contact model 3), FP64, for a
// positive overlap xi = -delta.  The reference raises the negative overlap
// itself to the powers 3/2, 1/4 and 1/2 (powf of a negative number: NaN
// forces); each power is taken here of xi, with the sign the linear law's
// terms carry (elastic normal term k_n xi^(3/2) repulsive, positive viscous
// and tangential coefficients), as xi sqrt(xi), sqrt(sqrt(xi)), sqrt(xi).
// The rolling resistance (:642-653) opposes the relative spin with the
// smaller of gamma_r r_a |w_ab| and mu_r r_a |f_n|.
__device__ __forceinline__ void contact_hertz(const Sys& S, Acc& acc, V3<double> vel_a, V3<double> angvel_a,
                                              V3<double> vel_b, V3<double> angvel_b, double radius_a,
                                              double radius_b, V3<double> x, double x_length, double delta,
                                              V3<double>& delta_t_io) {
    const double xi = -delta;
    const double sq = sqrt(xi);
    const double xi32 = xi * sq, xi14 = sqrt(sq), xi12 = sq;
    const V3<double> delta_t0_uncor = delta_t_io;
    const V3<double> n = x / x_length;
    const V3<double> vel_linear = vel_a - vel_b;
    const V3<double> vel = vel_linear + (radius_a + delta / 2.0) * cross(n, angvel_a)
                           + (radius_b + delta / 2.0) * cross(n, angvel_b);
    const V3<double> angvel_ab = angvel_a - angvel_b;
    const double angvel_ab_length = vlen(angvel_ab);
    const double vel_n = dot(vel_linear, n);
    const V3<double> vel_t = vel - n * dot(vel, n);
    const double vel_t_length = vlen(vel_t);
    const V3<double> delta_t0 = delta_t0_uncor - n * dot(delta_t0_uncor, n);
    const double delta_t0_length = vlen(delta_t0);
    V3<double> delta_t = delta_t0;
    V3<double> f_n = (S.k_n * xi32 - S.gamma_n * xi14 * vel_n) * n;
    acc.ev_dot += S.gamma_n * vel_n * vel_n;
    if (dot(f_n, n) < 0.0) f_n = mk(0.0, 0.0, 0.0);
    const double f_n_length = vlen(f_n);
    const V3<double> f_c = (-S.kappa * sqrt(radius_a * radius_b)) * n;
    V3<double> f_t = mk(0.0, 0.0, 0.0), t_res = mk(0.0, 0.0, 0.0);
    if (delta_t0_length > 0.0 || vel_t_length > 0.0) {
        const V3<double> f_t_elast = (-S.k_t * xi12) * delta_t0;
        const V3<double> f_t_visc = (-S.gamma_t * xi14) * vel_t;
        const double f_t_limit = vel_t_length > (double)0.001f ? S.mu_d * vlen(f_n - f_c) : S.mu_s * vlen(f_n - f_c);
        f_t = f_t_elast + f_t_visc;
        const double f_t_length = vlen(f_t);
        if (f_t_length > f_t_limit) {
            f_t = f_t * (f_t_limit / f_t_length);
            delta_t = (f_t + (S.gamma_t * xi14) * vel_t) / (-S.k_t * xi12);
            acc.es_dot += vlen(delta_t0 - delta_t) * S.k_t / S.dt;
        } else {
            delta_t = delta_t0 + vel_t * S.dt;
        }
    }
    if (angvel_ab_length > 0.0)
        t_res = ((-1.0 * fmin(S.gamma_r * radius_a * angvel_ab_length, S.mu_r * radius_a * f_n_length)) * angvel_ab) /
                angvel_ab_length;
    acc.F = acc.F + ((f_n + f_t) + f_c);
    acc.T = acc.T + ((-(radius_a + delta / 2.0)) * cross(n, f_t) + t_res);
    acc.p += f_n_length / (4.0 * PI_REF * radius_a * radius_a);
    delta_t_io = delta_t;
}

// src/contactmodels.cuh:10-109 contactLinear_wall (FP64 in both modes).
__device__ __forceinline__ double contact_wall(const Sys& S, Acc& acc, V3<double> linvel, V3<double> angvel,
                                               double radius_a, V3<double> n, double delta, double wvel) {
    const V3<double> v0 = linvel + radius_a * cross(n, angvel);
    // operator+(Float3, Float) adds the scalar wall velocity to every component
    const V3<double> vel = mk(v0.x + wvel, v0.y + wvel, v0.z + wvel);
    const double vel_n = dot(vel, n);
    const V3<double> vel_t = vel - n * dot(vel, n);
    const double vel_t_length = vlen(vel_t);
    V3<double> f_n = (-S.k_n * delta - S.gamma_wn * vel_n) * n;
    if (dot(f_n, n) < 0.0) f_n = mk(0.0, 0.0, 0.0);
    const double f_n_length = vlen(f_n);
    V3<double> f_t = mk(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0) {
        const double f_t_visc = S.gamma_wt * vel_t_length;
        const double f_t_limit = vel_t_length > (double)0.001f ? S.mu_wd * f_n_length : S.mu_ws * f_n_length;
        if (f_t_visc < f_t_limit) f_t = (-1.0 * f_t_visc) * vel_t / vel_t_length;
        else f_t = (-1.0 * f_t_limit) * vel_t / vel_t_length;
    }
    acc.F = acc.F + (f_n + f_t);
    const V3<double> tq = (-radius_a) * cross(n, f_t);
    acc.T = acc.T + mk(tq.x + 0.0, tq.y + 0.0, tq.z + 0.0);
    acc.p += f_n_length / (4.0 * PI_REF * radius_a * radius_a);
    return dot(f_n, n);
}

// The exponential of the capillary bond (the reference calls expf on a
// double): fixed steps, the same as the oracle's -- Cody-Waite reduction
// x = k ln2 + r with |r| <= ln2/2, the Taylor series of e^r to degree 13 by
// Horner's rule, scaling by 2^k -- so both sides get the same bits.
__device__ __forceinline__ double det_exp(double x) {
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double k = rint(x * inv_ln2);
    const double r = (x - k * ln2_hi) - k * ln2_lo;
    double q = 1.0 / 6227020800.0;  // 1/13!
    q = q * r + 1.0 / 479001600.0;
    q = q * r + 1.0 / 39916800.0;
    q = q * r + 1.0 / 3628800.0;
    q = q * r + 1.0 / 362880.0;
    q = q * r + 1.0 / 40320.0;
    q = q * r + 1.0 / 5040.0;
    q = q * r + 1.0 / 720.0;
    q = q * r + 1.0 / 120.0;
    q = q * r + 1.0 / 24.0;
    q = q * r + 1.0 / 6.0;
    q = q * r + 0.5;
    q = q * r + 1.0;
    q = q * r + 1.0;
    return ldexp(q, (int)k);
}

// src/contactmodels.cuh:114-253 contactLinearViscous (contact model 1, FP64
// in both precision modes): no spring history, no torque (commented out in
// the reference), viscous tangential force capped by Coulomb friction.
__device__ __forceinline__ void contact_viscous(const Sys& S, Acc& acc, V3<double> vel_a, V3<double> angvel_a,
                                                V3<double> vel_b, V3<double> angvel_b, double radius_a,
                                                double radius_b, V3<double> x, double x_length, double delta) {
    const V3<double> n = x / x_length;
    const V3<double> vel_linear = vel_a - vel_b;
    const V3<double> vel = vel_linear + (radius_a + delta / 2.0) * cross(n, angvel_a)
                           + (radius_b + delta / 2.0) * cross(n, angvel_b);
    const double vel_n = dot(vel_linear, n);
    const V3<double> vel_t = vel - n * dot(vel, n);
    const double vel_t_length = vlen(vel_t);
    V3<double> f_n = (-S.k_n * delta - S.gamma_n * vel_n) * n;
    if (dot(f_n, n) < 0.0) f_n = mk(0.0, 0.0, 0.0);
    const double f_n_length = vlen(f_n);
    const V3<double> f_c = (-S.kappa * sqrt(radius_a * radius_b)) * n;
    V3<double> f_t = mk(0.0, 0.0, 0.0);
    if (vel_t_length > 0.0) {
        const double f_t_visc = S.gamma_t * vel_t_length;
        const double f_t_limit = vel_t_length > (double)0.001f ? S.mu_d * vlen(f_n - f_c) : S.mu_s * vlen(f_n - f_c);
        if (f_t_visc < f_t_limit) f_t = ((-f_t_visc) * vel_t) / vel_t_length;
        else f_t = ((-f_t_limit) * vel_t) / vel_t_length;
    }
    acc.F = acc.F + (f_n + f_t + f_c);
    acc.p += f_n_length / (4.0 * PI_REF * radius_a * radius_a);
}

// src/cohesion.cuh:314-350 capillaryCohesion_exp, as written (the decay
// length has h = -sqrt(r) < 0).
__device__ __forceinline__ void capillary(const Sys& S, Acc& acc, double radius_a, double radius_b, double delta,
                                          V3<double> x, double x_length) {
    const V3<double> n = x / x_length;
    const double r = (radius_a / radius_b) > (radius_b / radius_a) ? radius_a / radius_b : radius_b / radius_a;
    const double h = -sqrt(r);
    const double R_har = (2.0 * radius_a * radius_b) / (radius_a + radius_b);
    const double R_geo = sqrt(radius_a * radius_b);
    const double lambda = (double)0.9f * h * sqrt(S.V_b / R_har);
    acc.F = acc.F + (-S.kappa * R_geo * det_exp(-delta / lambda)) * n;
}

__device__ __forceinline__ void decode_image(const Sys& S, uint32_t code, double& dx, double& dy) {
    const uint32_t cx = code & 3u, cy = (code >> 2) & 3u;
    dx = cx == 1u ? S.L[0] : (cx == 2u ? -S.L[0] : 0.0);
    dy = cy == 1u ? S.L[1] : (cy == 2u ? -S.L[1] : 0.0);
}

// A partner's velocity records as the pair force of arithmetic T reads
// them: the FP64 records, or (mixed mode) their FP32 copies, which the
// previous step's force kernel wrote next to them -- the same values the
// pair force would round them to, in half the bytes.
template <class T> struct PartnerVel;
template <> struct PartnerVel<double> {
    using type = double4;
    static __device__ __forceinline__ double4 vel(const DemDev& d, uint32_t p) { return d.ro.vel[p]; }
    static __device__ __forceinline__ double4 angvel(const DemDev& d, uint32_t p) { return d.ro.angvel[p]; }
    static __device__ __forceinline__ double4 zero() { return make_double4(0.0, 0.0, 0.0, 0.0); }
    static __device__ __forceinline__ V3<double> v3(double4 a) { return xyz(a); }
};
template <> struct PartnerVel<float> {
    using type = float4;
    static __device__ __forceinline__ float4 vel(const DemDev& d, uint32_t p) { return d.ro.vf[p]; }
    static __device__ __forceinline__ float4 angvel(const DemDev& d, uint32_t p) { return d.ro.wf[p]; }
    static __device__ __forceinline__ float4 zero() { return make_float4(0.f, 0.f, 0.f, 0.f); }
    static __device__ __forceinline__ V3<float> v3(float4 a) { return mk<float>(a.x, a.y, a.z); }
};

// The spring of contact slot k: read at the particle's storage position in
// the step's input set, written at its sorted position in the output set.
template <class T> struct SlotStore;
template <> struct SlotStore<double> {
    static __device__ __forceinline__ V3<double> load(const DemDev& d, int k, uint32_t q) {
        const double* p = d.ro.cdt + 3 * ((size_t)k * d.np + q);
        return mk(p[0], p[1], p[2]);
    }
    static __device__ __forceinline__ void store(const DemDev& d, int k, uint32_t q, V3<double> v) {
        double* p = d.wo.cdt + 3 * ((size_t)k * d.np + q);
        p[0] = v.x; p[1] = v.y; p[2] = v.z;
    }
};
template <> struct SlotStore<float> {
    static __device__ __forceinline__ V3<float> load(const DemDev& d, int k, uint32_t q) {
        const float* p = d.ro.cdt32 + 3 * ((size_t)k * d.np + q);
        return mk(p[0], p[1], p[2]);
    }
    static __device__ __forceinline__ void store(const DemDev& d, int k, uint32_t q, V3<float> v) {
        float* p = d.wo.cdt32 + 3 * ((size_t)k * d.np + q);
        p[0] = v.x; p[1] = v.y; p[2] = v.z;
    }
};

// ----------------------------------------------------------- contact search
// One block per tile of tile_x x tile_y cells in one z plane.  The positions
// of every particle in the 3 x (tile_y+2) x (tile_x+2) cells around the tile
// are staged once into LDS (each staged cell is a contiguous run of the
// sorted arrays); each thread then walks its particle's 27 cells from LDS in
// the reference's z, y, x order and, within a cell, in sorted order
// (src/contactsearch.cuh:181-355).  A tile whose neighbourhood exceeds the
// LDS capacity reads the same candidates from HBM instead.
struct TopoShared {
    double4 pos[TOPO_CAP];
    uint32_t cstart[TOPO_MAX_CELLS + 1];  // LDS offset of each staged cell
    uint32_t cglob[TOPO_MAX_CELLS];       // first sorted index of each staged cell
    uint8_t ccode[TOPO_MAX_CELLS];        // periodic image code of each staged cell
    uint32_t scan[BLOCK / 64];
};

// The slot words of one particle in the step's input set: word k of the
// particle stored at q is cslot[k * np + q] (slot-major, like the springs).
// A search updates the row in place (each particle's row is its own); the
// force kernel carries the words of the contacts that last into the output
// set at the particle's sorted position.
struct SlotRow {
    uint32_t* p;
    size_t ld;
    __device__ __forceinline__ uint32_t& operator[](int k) const { return p[(size_t)k * ld]; }
};
__device__ __forceinline__ SlotRow slot_row(const DemDev& d, uint32_t s) { return {d.ro.cslot + d.perm[s], d.np}; }

// Slot bookkeeping of one confirmed overlap (src/contactsearch.cuh:252-284):
// the slot already holding b (lowest), else the highest empty slot.
struct SlotState {
    uint32_t mask, fresh, nc, err;
    bool stop;
};

__device__ __noinline__ void slot_insert(SlotRow slot, SlotState& st, uint32_t bo, uint32_t code) {
    int cpos = -1;
    for (uint32_t m = st.mask; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        if ((slot[i] & SLOT_IDX_MASK) == bo) { cpos = i; break; }
    }
    if (cpos < 0) {
        const uint32_t empty = ~st.mask;
        if (empty == 0u) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; return; }
        cpos = 31 - __clz(empty);
        st.fresh |= 1u << cpos;
    }
    slot[cpos] = bo | (code << 28);
    st.mask |= 1u << cpos;
    if (++st.nc > (uint32_t)NC) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; }
}

template <bool FROM_LDS>
__device__ __forceinline__ double4 staged_pos(const DemDev& d, const TopoShared& sh, uint32_t lb, uint32_t sb) {
    if constexpr (FROM_LDS) return sh.pos[lb];
    else return spos(d, d.sys, sb);
}

// Walks the 27 cells of every home particle of the tile: 9 rows (z, y order),
// each row the contiguous run of cells x-1, x, x+1, within a cell in sorted
// order -- the candidate order of src/contactsearch.cuh:340-353.  Candidates
// come from LDS (FROM_LDS) or from HBM.  The hot loop only rejects pairs whose
// squared distance exceeds (ra+rb)^2 by more than rounding; the exact test
// sqrt(d2) - (ra+rb) < 0 runs on the few that remain.
template <bool FROM_LDS>
__device__ __forceinline__ void topo_search(const DemDev& d, const TopoShared& sh, int W, int H, int nxh,
                                            int nyh) {
    uint32_t home_total = 0;
    for (int hr = 0; hr < nyh; ++hr) {
        const int e0 = (1 * H + hr + 1) * W + 1;
        home_total += sh.cstart[e0 + nxh] - sh.cstart[e0];
    }
    for (uint32_t h = threadIdx.x; h < home_total; h += BLOCK) {
        int hr = 0, e0 = 0;
        uint32_t li = 0, rem = h;
        for (; hr < nyh; ++hr) {
            e0 = (1 * H + hr + 1) * W + 1;
            const uint32_t n_row = sh.cstart[e0 + nxh] - sh.cstart[e0];
            if (rem < n_row) { li = sh.cstart[e0] + rem; break; }
            rem -= n_row;
        }
        int eh = e0;
        while (sh.cstart[eh + 1] <= li) ++eh;
        const int colh = eh - (1 * H + hr + 1) * W;  // 1..nxh
        const uint32_t s = sh.cglob[eh] + (li - sh.cstart[eh]);
        const double4 xa = staged_pos<FROM_LDS>(d, sh, li, s);
        const double ra = xa.w;
        const SlotRow slot = slot_row(d, s);
        SlotState st;
        st.mask = d.ro.cmask[d.perm[s]];
        st.fresh = 0; st.nc = 0; st.err = 0; st.stop = false;
        for (int lay = 0; lay < 3 && !st.stop; ++lay) {
            for (int dy = 0; dy < 3 && !st.stop; ++dy) {
                const int e = (lay * H + hr + dy) * W + colh - 1;
                const uint32_t b0 = sh.cstart[e], b3 = sh.cstart[e + 3];
                if (b0 == b3) continue;
                const uint32_t codes = (uint32_t)sh.ccode[e] | ((uint32_t)sh.ccode[e + 1] << 8) |
                                       ((uint32_t)sh.ccode[e + 2] << 16);
                for (int c = 0; c < 3 && !st.stop; ++c) {
                    const uint32_t lo = c == 0 ? b0 : sh.cstart[e + c];
                    const uint32_t hi = c == 2 ? b3 : sh.cstart[e + c + 1];
                    if (lo == hi) continue;
                    const uint32_t code = (codes >> (8 * c)) & 0xffu;
                    double dmx, dmy;
                    decode_image(d.sys, code, dmx, dmy);
                    const uint32_t gb = sh.cglob[e + c] - lo;
                    for (uint32_t lb = lo; lb < hi; ++lb) {
                        const double4 xb = staged_pos<FROM_LDS>(d, sh, lb, gb + lb);
                        const double ex = (xa.x - xb.x) + dmx;
                        const double ey = (xa.y - xb.y) + dmy;
                        const double ez = xa.z - xb.z;
                        // (ez + 0.0)^2 == ez^2, so d2 equals dot(xab, xab) of the exact test
                        const double d2 = ex * ex + ey * ey + ez * ez;
                        const double rsum = ra + xb.w;
                        if (__builtin_expect(d2 <= rsum * rsum * 1.0000001, 0) && lb != li) {
                            // exact test of src/contactsearch.cuh:242-248 (double sqrt)
                            if (sqrt(d2) - rsum < 0.0) {
                                slot_insert(slot, st, d.sid[gb + lb], code);
                                if (st.stop) break;
                            }
                        }
                    }
                }
            }
        }
        d.tmask[s] = st.mask;
        d.tfresh[s] = st.fresh;
        if (st.err) atomicOr(d.flags, st.err);
    }
}

__global__ __launch_bounds__(BLOCK) void k_topo(DemDev d) {
    __shared__ TopoShared sh;
    const Sys& S = d.sys;
    const uint32_t bt = xcd_block();
    const uint32_t txi = bt % d.tiles_x, rest = bt / d.tiles_x;
    const uint32_t tyi = rest % d.tiles_y, tz = rest / d.tiles_y;
    const int nx = (int)S.num[0], ny = (int)S.num[1], nz = (int)S.num[2];
    const int x0 = (int)(txi * d.tile_x), y0 = (int)(tyi * d.tile_y);
    const int nxh = min((int)d.tile_x, nx - x0), nyh = min((int)d.tile_y, ny - y0);
    const int W = nxh + 2, H = nyh + 2;
    const int ncell = 3 * H * W;  // <= TOPO_MAX_CELLS (host-checked)

    // 1. staged cells: global run, image code, count -> LDS offsets
    uint32_t cnt2[2] = {0u, 0u};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int e = 2 * (int)threadIdx.x + k;
        if (e >= ncell) continue;
        const int col = e % W, rr = e / W, row = rr % H, lay = rr / H;
        int cx = x0 - 1 + col, cy = y0 - 1 + row;
        const int cz = (int)tz - 1 + lay;
        uint32_t code = 0, g0 = 0, cnt = 0;
        bool ok = cz >= 0 && cz < nz;
        if (ok && (cy < 0 || cy >= ny)) {
            if (S.periodic == 1) { code |= (cy < 0 ? 1u : 2u) << 2; cy = cy < 0 ? ny - 1 : 0; }
            else ok = false;
        }
        if (ok && (cx < 0 || cx >= nx)) {
            if (S.periodic != 0) { code |= cx < 0 ? 1u : 2u; cx = cx < 0 ? nx - 1 : 0; }
            else ok = false;
        }
        if (ok) {
            const uint32_t cid = (uint32_t)cx + (uint32_t)cy * S.num[0] + (S.num[0] * S.num[1]) * (uint32_t)cz;
            g0 = d.off[cid];
            cnt = d.off[cid + 1] - g0;
        }
        sh.cglob[e] = g0;
        sh.ccode[e] = (uint8_t)code;
        cnt2[k] = cnt;
    }
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(cnt2[0] + cnt2[1], sh.scan, total);
    if (2 * (int)threadIdx.x < ncell) sh.cstart[2 * threadIdx.x] = ex;
    if (2 * (int)threadIdx.x + 1 < ncell) sh.cstart[2 * threadIdx.x + 1] = ex + cnt2[0];
    if (threadIdx.x == 0) sh.cstart[ncell] = total;
    __syncthreads();
    const bool use_lds = total <= (uint32_t)TOPO_CAP;

    // 2. stage positions
    if (use_lds) {
        for (uint32_t i = threadIdx.x; i < total; i += BLOCK) {
            int lo = 0, hi = ncell - 1;  // last cell with cstart <= i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sh.cstart[mid] <= i) lo = mid; else hi = mid - 1;
            }
            sh.pos[i] = spos(d, S, sh.cglob[lo] + (i - sh.cstart[lo]));
        }
    }
    __syncthreads();

    // 3. home particles: cells (lay 1, rows 1..nyh, cols 1..nxh), row by row
    if (use_lds) topo_search<true>(d, sh, W, H, nxh, nyh);
    else topo_search<false>(d, sh, W, H, nxh, nyh);
}

// One thread per sorted particle; the 27 cells are walked as 9 runs of the
// sorted arrays (cells x-1, x, x+1 of a row have consecutive ids), the same
// candidate order as k_topo, candidates read through L1/L2.  The hot loop
// reads the FP32 copy of the positions and keeps only pairs within
// (r_a + r_b + search_eps), a margin larger than the FP32 rounding of the
// distance, so it never drops an overlap; the reference's exact test
// sqrt(d2) - (r_a + r_b) < 0 then runs in FP64 on the few that remain
// (rows_try).  The particle's slot words are cached in LDS from its first
// overlap on and written back once.
struct RowSlots {
    uint32_t mask, fresh, nc, err, dirty;
    bool stop, loaded;
    uint32_t ld;  // stride of the LDS slot column (threads per block)
};

// (Takes plain pointers rather than the DemDev, which a non-inlined call
// would have to copy to scratch memory.)
__device__ __noinline__ void rows_try(PosView pv, const uint32_t* __restrict__ sid, uint32_t* sw,
                                      SlotRow gslot, RowSlots& st, double4 xa, uint32_t sb, uint32_t code) {
    const double4 xb = pv_pos(pv, sb);
    const uint32_t cx = code & 3u, cy = (code >> 2) & 3u;
    const double dmx = cx == 1u ? pv.Lx : (cx == 2u ? -pv.Lx : 0.0);
    const double dmy = cy == 1u ? pv.Ly : (cy == 2u ? -pv.Ly : 0.0);
    const double ex = (xa.x - xb.x) + dmx, ey = (xa.y - xb.y) + dmy, ez = xa.z - xb.z;
    // ((e + 0.0)^2 == e^2: d2 equals dot(xab, xab) of src/contactsearch.cuh:242-248)
    const double d2 = ex * ex + ey * ey + ez * ez;
    if (!(sqrt(d2) - (xa.w + xb.w) < 0.0)) return;
    if (!st.loaded) {
        for (uint32_t m = st.mask; m; m &= m - 1) {
            const int i = __ffs(m) - 1;
            sw[i * st.ld] = gslot[i];
        }
        st.loaded = true;
    }
    // slot rule of src/contactsearch.cuh:252-271
    const uint32_t bo = sid[sb];
    int cpos = -1;
    for (uint32_t m = st.mask; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        if ((sw[i * st.ld] & SLOT_IDX_MASK) == bo) { cpos = i; break; }
    }
    if (cpos < 0) {
        const uint32_t empty = ~st.mask;
        if (empty == 0u) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; return; }
        cpos = 31 - __clz(empty);
        st.fresh |= 1u << cpos;
    }
    const uint32_t w = bo | (code << 28);
    if (sw[cpos * st.ld] != w || !((st.mask >> cpos) & 1u)) {  // unchanged kept slots are not written back
        sw[cpos * st.ld] = w;
        st.dirty |= 1u << cpos;
    }
    st.mask |= 1u << cpos;
    if (++st.nc > (uint32_t)NC) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; }
}

// Candidates [lo, hi) of one run with periodic image (dmx, dmy), in batches
// of ROW_BATCH independent loads; overlaps go to rows_try in candidate order.
constexpr int ROW_BATCH = 8;

// With BUILD, the pass also records every candidate within the list reach
// (r_a + r_b + skin, same prefilter margin) in the particle's neighbour list.
template <bool BUILD>
__device__ __forceinline__ void rows_run(const DemDev& d, const Sys& S, uint32_t lo, uint32_t hi, float dmx, float dmy,
                                         uint32_t code, uint32_t s, float4 xaf, float ra_eps, float ra_skin,
                                         double4 xa, uint32_t* sw, SlotRow gslot, RowSlots& st,
                                         uint32_t* list, uint32_t& lcnt) {
    for (uint32_t sb0 = lo; sb0 < hi; sb0 += ROW_BATCH) {
        float4 c[ROW_BATCH];
#pragma unroll
        for (int k = 0; k < ROW_BATCH; ++k)
            c[k] = sb0 + k < hi ? d.xf[sb0 + k] : make_float4(1e30f, 1e30f, 1e30f, 0.0f);
        uint32_t hits = 0, near = 0;
#pragma unroll
        for (int k = 0; k < ROW_BATCH; ++k) {
            const float ex = (xaf.x - c[k].x) + dmx, ey = (xaf.y - c[k].y) + dmy, ez = xaf.z - c[k].z;
            const float d2 = ex * ex + ey * ey + ez * ez;
            const float rs = ra_eps + c[k].w;
            hits |= d2 <= rs * rs ? 1u << k : 0u;
            if constexpr (BUILD) {
                const float rl = ra_skin + c[k].w;
                near |= d2 <= rl * rl ? 1u << k : 0u;
            }
        }
        if constexpr (BUILD) {
            for (; near; near &= near - 1) {
                const int k = __ffs(near) - 1;
                const uint32_t sb = sb0 + (uint32_t)k;
                if (sb == s) continue;
                if (lcnt < (uint32_t)NLIST) {
                    list[(size_t)lcnt * d.np] = sb;  // the partner's list index
                    d.nhint[(size_t)lcnt * d.np + s] = 0xffu;  // no slot known yet
                    const float ex = (xaf.x - c[k].x) + dmx, ey = (xaf.y - c[k].y) + dmy, ez = xaf.z - c[k].z;
                    d.nlistg[(size_t)lcnt * d.np + s] = sqrtf(ex * ex + ey * ey + ez * ez) - (xaf.w + c[k].w);
                }
                ++lcnt;
            }
        }
        if (__builtin_expect(hits != 0u, 0)) {
            for (; hits; hits &= hits - 1) {
                const uint32_t sb = sb0 + (uint32_t)(__ffs(hits) - 1);
                if (sb == s) continue;
                rows_try(pos_view(d, S), d.sid, sw, gslot, st, xa, sb, code);
                if (st.stop) return;
            }
        }
    }
}

// Distance along one axis from the coordinate v to the box of cell t (of n)
// seen through the periodic shift dm (the image of a sphere at c is c - dm),
// widened by mg.  Spheres of the edge cells may lie beyond the grid (cell
// ids clamp), so those cells are open on their outer side.
__device__ __forceinline__ float cell_gap(float v, int t, int n, float o, float cl, float dm, float mg) {
    const float lo = t == 0 ? -INFINITY : o + (float)t * cl - dm - mg;
    const float hi = t == n - 1 ? INFINITY : o + (float)(t + 1) * cl - dm + mg;
    return fmaxf(0.0f, fmaxf(lo - v, v - hi));
}

// Second pass of a list build: the spheres two cells away (the outer shell of
// the 5 x 5 x 5 cells around the particle's cell) within the list reach.  A
// pair two cells apart is more than a cell apart; when a cell is narrower
// than the largest pair plus 0.9 skin such a pair can meet before the next
// rebuild, so the host asks for this pass (list_shell) instead of shrinking
// the rebuild distance to fit the 27 cells.  It does so only when every
// periodic axis has at least 5 cells: no cell is then reached twice.
// Cells whose widened box lies farther than the list reach with the largest
// partner (reach2 = (r_a + eps + skin + r_max)^2, rounded up) hold no list
// entry and are skipped.
struct CellBox {
    float o[3], cl[3], mg, reach2;
};

__device__ __noinline__ void list_shell(const uint32_t* __restrict__ off, const float4* __restrict__ xf,
                                        uint8_t* hint,
                                        uint32_t* list, float* gap, uint32_t np, uint32_t& lcnt, int gx, int gy, int gz, int nx,
                                        int ny, int nz, int periodic, float Lx, float Ly, uint32_t cbase,
                                        float4 xaf, float ra_skin, CellBox cb3) {
    for (int dz = -2; dz <= 2; ++dz) {
        const int tz = gz + dz;
        if (tz < 0 || tz >= nz) continue;
        const float gzz = cell_gap(xaf.z, tz, nz, cb3.o[2], cb3.cl[2], 0.0f, cb3.mg);
        for (int dy = -2; dy <= 2; ++dy) {
            int ty = gy + dy;
            float dmy = 0.0f;
            if (ty < 0 || ty >= ny) {
                if (periodic != 1) continue;
                dmy = ty < 0 ? Ly : -Ly;
                ty = ty < 0 ? ty + ny : ty - ny;
            }
            const float gyy = cell_gap(xaf.y, ty, ny, cb3.o[1], cb3.cl[1], dmy, cb3.mg);
            if (gyy * gyy + gzz * gzz > cb3.reach2) continue;
            for (int dx = -2; dx <= 2; ++dx) {
                if (dx > -2 && dx < 2 && dy > -2 && dy < 2 && dz > -2 && dz < 2) continue;  // the 27 cells
                int tx = gx + dx;
                float dmx = 0.0f;
                if (tx < 0 || tx >= nx) {
                    if (periodic == 0) continue;
                    dmx = tx < 0 ? Lx : -Lx;
                    tx = tx < 0 ? tx + nx : tx - nx;
                }
                const float gxx = cell_gap(xaf.x, tx, nx, cb3.o[0], cb3.cl[0], dmx, cb3.mg);
                if (gxx * gxx + gyy * gyy + gzz * gzz > cb3.reach2) continue;
                const uint32_t c = cbase + (uint32_t)tx + (uint32_t)ty * (uint32_t)nx + (uint32_t)tz * (uint32_t)(nx * ny);
                const uint32_t hi = off[c + 1];
                for (uint32_t sb = off[c]; sb < hi; ++sb) {
                    const float4 cb = xf[sb];
                    const float ex = (xaf.x - cb.x) + dmx, ey = (xaf.y - cb.y) + dmy, ez = xaf.z - cb.z;
                    const float rl = ra_skin + cb.w;
                    const float d2 = ex * ex + ey * ey + ez * ez;
                    if (d2 <= rl * rl) {
                        if (lcnt < (uint32_t)NLIST) {
                            list[(size_t)lcnt * np] = sb;
                            hint[(size_t)lcnt * np] = 0xffu;
                            gap[(size_t)lcnt * np] = sqrtf(d2) - (xaf.w + cb.w);
                        }
                        ++lcnt;
                    }
                }
            }
        }
    }
}

// BUILD (search mode 5): runs only on steps whose neighbour lists must be
// rebuilt, and then also rebuilds the lists.
template <bool BUILD, bool B>
__global__ __launch_bounds__(BLOCK) void k_topo_rows(DemDev d) {
    extern __shared__ uint32_t sw_all[];  // NC x blockDim: slot i of thread t at [i * blockDim + t]
    if (BUILD && d.nstate[0] == 0u) return;
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= d.np) return;
    const Sys& S = sys_at<B>(d, s);
    uint32_t* sw = sw_all + threadIdx.x;
    const int nx = (int)S.num[0], ny = (int)S.num[1], nz = (int)S.num[2];
    const uint32_t key = d.skey[s] - S.cbase;
    const int gx = (int)(key % S.num[0]), gy = (int)((key / S.num[0]) % S.num[1]), gz = (int)(key / (S.num[0] * S.num[1]));
    const double4 xa = spos(d, S, s);
    const float4 xaf = d.xf[s];
    const float ra_eps = xaf.w + d.search_eps;
    const float ra_skin = ra_eps + d.list_skin;
    const SlotRow gslot = slot_row(d, s);
    uint32_t* list = BUILD ? d.nlist + s : nullptr;  // entry k at list[k * np]
    uint32_t lcnt = 0;
    // rows of cells farther than any hit (or list entry) can reach are skipped
    CellBox cb3;
    for (int a = 0; a < 3; ++a) {
        cb3.o[a] = (float)S.origo[a];
        cb3.cl[a] = (float)S.cell_len[a];
    }
    cb3.mg = d.cell_margin;
    {
        const float reach = (BUILD ? ra_skin : ra_eps) + d.rmax_f;
        cb3.reach2 = reach * reach * 1.001f;
    }
    RowSlots st;
    st.mask = d.ro.cmask[d.perm[s]];
    st.fresh = 0; st.nc = 0; st.err = 0; st.dirty = 0; st.stop = false; st.loaded = false;
    st.ld = blockDim.x;
    for (int dz = -1; dz <= 1 && !st.stop; ++dz) {
        const int tz = gz + dz;
        if (tz < 0 || tz == nz) continue;
        for (int dy = -1; dy <= 1 && !st.stop; ++dy) {
            int ty = gy + dy;
            uint32_t cy = 0;
            if (ty < 0 || ty == ny) {
                if (S.periodic != 1) continue;
                cy = ty < 0 ? 1u : 2u;
                ty = ty < 0 ? ny - 1 : 0;
            }
            {
                const float dmy = cy == 1u ? S.Lf[1] : (cy == 2u ? -S.Lf[1] : 0.0f);
                const float gyy = cell_gap(xaf.y, ty, ny, cb3.o[1], cb3.cl[1], dmy, cb3.mg);
                const float gzz = cell_gap(xaf.z, tz, nz, cb3.o[2], cb3.cl[2], 0.0f, cb3.mg);
                if (gyy * gyy + gzz * gzz > cb3.reach2) continue;
            }
            const uint32_t row = S.cbase + (uint32_t)ty * S.num[0] + (S.num[0] * S.num[1]) * (uint32_t)tz;
            if (gx >= 1 && gx + 1 < nx && cy == 0) {
                // no periodic image anywhere in the run: one contiguous range
                rows_run<BUILD>(d, S, d.off[row + gx - 1], d.off[row + gx + 2], 0.0f, 0.0f, 0u, s, xaf, ra_eps,
                                ra_skin, xa, sw, gslot, st, list, lcnt);
                continue;
            }
            for (int dx = -1; dx <= 1 && !st.stop; ++dx) {
                int tx = gx + dx;
                uint32_t cx = 0;
                if (tx < 0 || tx == nx) {
                    if (S.periodic == 0) continue;
                    cx = tx < 0 ? 1u : 2u;
                    tx = tx < 0 ? nx - 1 : 0;
                }
                const float dmx = cx == 1u ? S.Lf[0] : (cx == 2u ? -S.Lf[0] : 0.0f);
                const float dmy = cy == 1u ? S.Lf[1] : (cy == 2u ? -S.Lf[1] : 0.0f);
                rows_run<BUILD>(d, S, d.off[row + tx], d.off[row + tx + 1], dmx, dmy, cx | (cy << 2), s, xaf,
                                ra_eps, ra_skin, xa, sw, gslot, st, list, lcnt);
            }
        }
    }
    for (uint32_t m = st.dirty; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        gslot[i] = sw[i * st.ld];
    }
    d.tmask[s] = st.mask;
    d.tfresh[s] = st.fresh;
    if (st.err) atomicOr(d.flags, st.err);
    if constexpr (BUILD) {  // list index = this sorted position
        if (d.list_shell && !st.stop)
            list_shell(d.off, d.xf, d.nhint + s, list, d.nlistg + s, d.np, lcnt, gx, gy, gz, nx, ny, nz, S.periodic, S.Lf[0],
                       S.Lf[1], S.cbase, xaf, ra_skin, cb3);
        d.ncnt[s] = lcnt < (uint32_t)NLIST ? lcnt : (uint32_t)NLIST;
        d.xref[s] = make_float4(xaf.x, xaf.y, xaf.z, 0.0f);
        if (lcnt > (uint32_t)NLIST || st.stop) atomicOr(&d.nstate[2], 1u);  // list unusable: rebuild next step
    }
}

// ---------------------------------------------------- half-shell search
// Every unordered pair is tested once, by its lower sorted position, with the
// same FP32 prefilter and FP64 exact test as k_topo_rows (the exact test is
// symmetric: swapping a and b negates every component of the distance
// exactly).  An overlap is recorded in both particles' hit lists under its
// key in each one's reference visiting order -- cell offset index
// (dz+1)*9 + (dy+1)*3 + (dx+1), then sorted position (src/contactsearch.cuh:
// 340-353) -- and k_topo_apply replays each list in key order through the
// slot rule.
constexpr uint32_t NO_SPLIT = 0xffffffffu;

__device__ __noinline__ uint32_t half_hit(PosView pv, uint32_t* hcnt, uint64_t* hits, double4 xa, uint32_t s,
                                          uint32_t sb, uint32_t vi, uint32_t code) {
    const double4 xb = pv_pos(pv, sb);
    const uint32_t cx = code & 3u, cy = (code >> 2) & 3u;
    const double dmx = cx == 1u ? pv.Lx : (cx == 2u ? -pv.Lx : 0.0);
    const double dmy = cy == 1u ? pv.Ly : (cy == 2u ? -pv.Ly : 0.0);
    const double ex = (xa.x - xb.x) + dmx, ey = (xa.y - xb.y) + dmy, ez = xa.z - xb.z;
    const double d2 = ex * ex + ey * ey + ez * ez;
    if (!(sqrt(d2) - (xa.w + xb.w) < 0.0)) return 0u;
    const uint32_t rx = cx == 1u ? 2u : (cx == 2u ? 1u : 0u), ry = cy == 1u ? 2u : (cy == 2u ? 1u : 0u);
    const uint32_t rcode = rx | (ry << 2);
    uint32_t err = 0;
    const uint32_t k = atomicAdd(&hcnt[s], 1u);
    if (k < (uint32_t)NC) hits[(size_t)s * NC + k] = ((uint64_t)vi << 40) | ((uint64_t)code << 32) | sb;
    else err |= FLAG_CONTACT_OVERFLOW;
    const uint32_t k2 = atomicAdd(&hcnt[sb], 1u);
    if (k2 < (uint32_t)NC) hits[(size_t)sb * NC + k2] = ((uint64_t)(26u - vi) << 40) | ((uint64_t)rcode << 32) | s;
    else err |= FLAG_CONTACT_OVERFLOW;
    return err;
}

// Candidates (s, hi) of one run; the run's cells start their offset index at
// vi0 and advance at sorted positions m1 and m2.
__device__ __forceinline__ uint32_t half_run(const DemDev& d, const Sys& S, uint32_t lo, uint32_t hi, uint32_t m1, uint32_t m2,
                                             uint32_t vi0, float dmx, float dmy, uint32_t code, uint32_t s,
                                             float4 xaf, float ra_eps, double4 xa) {
    uint32_t err = 0;
    if (lo <= s) lo = s + 1;
    for (uint32_t sb0 = lo; sb0 < hi; sb0 += ROW_BATCH) {
        float4 c[ROW_BATCH];
#pragma unroll
        for (int k = 0; k < ROW_BATCH; ++k)
            c[k] = sb0 + k < hi ? d.xf[sb0 + k] : make_float4(1e30f, 1e30f, 1e30f, 0.0f);
        uint32_t hits = 0;
#pragma unroll
        for (int k = 0; k < ROW_BATCH; ++k) {
            const float ex = (xaf.x - c[k].x) + dmx, ey = (xaf.y - c[k].y) + dmy, ez = xaf.z - c[k].z;
            const float d2 = ex * ex + ey * ey + ez * ez;
            const float rs = ra_eps + c[k].w;
            hits |= d2 <= rs * rs ? 1u << k : 0u;
        }
        if (__builtin_expect(hits != 0u, 0)) {
            for (; hits; hits &= hits - 1) {
                const uint32_t sb = sb0 + (uint32_t)(__ffs(hits) - 1);
                const uint32_t vi = vi0 + (sb >= m1 ? 1u : 0u) + (sb >= m2 ? 1u : 0u);
                err |= half_hit(pos_view(d, S), d.hcnt, d.hits, xa, s, sb, vi, code);
            }
        }
    }
    return err;
}

template <bool B>
__global__ __launch_bounds__(BLOCK) void k_topo_half(DemDev d) {
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= d.np) return;
    const Sys& S = sys_at<B>(d, s);
    const int nx = (int)S.num[0], ny = (int)S.num[1], nz = (int)S.num[2];
    const uint32_t key = d.skey[s] - S.cbase;
    const int gx = (int)(key % S.num[0]), gy = (int)((key / S.num[0]) % S.num[1]), gz = (int)(key / (S.num[0] * S.num[1]));
    const double4 xa = spos(d, S, s);
    const float4 xaf = d.xf[s];
    const float ra_eps = xaf.w + d.search_eps;
    uint32_t err = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int tz = gz + dz;
        if (tz < 0 || tz == nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int ty = gy + dy;
            uint32_t cy = 0;
            if (ty < 0 || ty == ny) {
                if (S.periodic != 1) continue;
                cy = ty < 0 ? 1u : 2u;
                ty = ty < 0 ? ny - 1 : 0;
            }
            const uint32_t row = S.cbase + (uint32_t)ty * S.num[0] + (S.num[0] * S.num[1]) * (uint32_t)tz;
            const uint32_t vrow = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3);
            if (gx >= 1 && gx + 1 < nx && cy == 0) {
                const uint32_t hi = d.off[row + gx + 2];
                if (hi <= s + 1) continue;  // the whole run lies at or below s
                err |= half_run(d, S, d.off[row + gx - 1], hi, d.off[row + gx], d.off[row + gx + 1], vrow, 0.0f, 0.0f,
                                0u, s, xaf, ra_eps, xa);
                continue;
            }
            for (int dx = -1; dx <= 1; ++dx) {
                int tx = gx + dx;
                uint32_t cx = 0;
                if (tx < 0 || tx == nx) {
                    if (S.periodic == 0) continue;
                    cx = tx < 0 ? 1u : 2u;
                    tx = tx < 0 ? nx - 1 : 0;
                }
                const float dmx = cx == 1u ? S.Lf[0] : (cx == 2u ? -S.Lf[0] : 0.0f);
                const float dmy = cy == 1u ? S.Lf[1] : (cy == 2u ? -S.Lf[1] : 0.0f);
                err |= half_run(d, S, d.off[row + tx], d.off[row + tx + 1], NO_SPLIT, NO_SPLIT, vrow + (uint32_t)(dx + 1),
                                dmx, dmy, cx | (cy << 2), s, xaf, ra_eps, xa);
            }
        }
    }
    if (err) atomicOr(d.flags, err);
}

// Replays each particle's hits in visiting-key order through the slot rule
// (src/contactsearch.cuh:252-271); the keys (at most NC) are insertion-sorted
// in place, the slot words cached in LDS.
constexpr int APPLY_BLOCK = 128;

// Copies the occupied slot words of a row into the particle's LDS column
// (stride APPLY_BLOCK), eight independent loads at a time.
__device__ __forceinline__ void stage_slots(SlotRow gslot, uint32_t mask, uint32_t* sw) {
    constexpr int G = 8;
    for (uint32_t m = mask; m;) {
        int si[G];
        uint32_t w[G];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            si[k] = m ? __ffs(m) - 1 : 0;
            if (m) { m &= m - 1; ++cnt; }
        }
#pragma unroll
        for (int k = 0; k < G; ++k) w[k] = k < cnt ? gslot[si[k]] : 0u;
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (k < cnt) sw[si[k] * APPLY_BLOCK] = w[k];
    }
}

// Applies the n (<= NC) hit keys k of sorted particle s through the slot
// rule as if in key order.  A partner occurs at most once among the hits, so
// a hit whose partner already holds a slot lands in that slot whatever the
// order; only the new contacts take empty slots in key order, and only they
// are sorted.  sw is the particle's LDS column of slot words (stride
// APPLY_BLOCK).
__device__ __forceinline__ void apply_hits(const DemDev& d, uint32_t s, uint64_t* k, uint32_t n, uint32_t mask0,
                                           uint32_t* sw, uint32_t err) {
    const SlotRow gslot = slot_row(d, s);
    stage_slots(gslot, mask0, sw);
    uint32_t mask = mask0, fresh = 0, dirty = 0, nnew = 0;
    for (uint32_t h = 0; h < n; ++h) {
        const uint64_t v = k[h];
        const uint32_t sb = (uint32_t)(v & 0xffffffffu), code = (uint32_t)(v >> 32) & 15u;
        const uint32_t bo = d.sid[sb];
        int cpos = -1;
        for (uint32_t m = mask0; m; m &= m - 1) {
            const int i = __ffs(m) - 1;
            if ((sw[i * APPLY_BLOCK] & SLOT_IDX_MASK) == bo) { cpos = i; break; }
        }
        if (cpos >= 0) {  // kept; written back only if its image changed
            const uint32_t w = bo | (code << 28);
            if (sw[cpos * APPLY_BLOCK] != w) {
                sw[cpos * APPLY_BLOCK] = w;
                dirty |= 1u << cpos;
            }
        } else {
            k[nnew++] = v;  // compacted in place (nnew <= h)
        }
    }
    for (uint32_t i = 1; i < nnew; ++i) {
        const uint64_t v = k[i];
        uint32_t j = i;
        while (j > 0 && k[j - 1] > v) { k[j] = k[j - 1]; --j; }
        k[j] = v;
    }
    for (uint32_t h = 0; h < nnew; ++h) {
        const uint64_t v = k[h];
        const uint32_t sb = (uint32_t)(v & 0xffffffffu), code = (uint32_t)(v >> 32) & 15u;
        const uint32_t empty = ~mask;
        if (empty == 0u) { err |= FLAG_CONTACT_OVERFLOW; break; }
        const int cpos = 31 - __clz(empty);
        fresh |= 1u << cpos;
        sw[cpos * APPLY_BLOCK] = d.sid[sb] | (code << 28);
        dirty |= 1u << cpos;
        mask |= 1u << cpos;
    }
    for (uint32_t m = dirty; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        gslot[i] = sw[i * APPLY_BLOCK];
    }
    d.tmask[s] = mask;
    d.tfresh[s] = fresh;
    if (err) atomicOr(d.flags, err);
}

__global__ __launch_bounds__(APPLY_BLOCK) void k_topo_apply(DemDev d) {
    __shared__ uint32_t sw_all[NC * APPLY_BLOCK];
    const uint32_t s = xcd_block() * APPLY_BLOCK + threadIdx.x;
    if (s >= d.np) return;
    const uint32_t cnt = d.hcnt[s];
    const uint32_t mask0 = d.ro.cmask[d.perm[s]];
    if (cnt != 0u) d.hcnt[s] = 0u;  // zero for the next search
    if (cnt == 0u) {
        d.tmask[s] = mask0;
        d.tfresh[s] = 0u;
        return;
    }
    const uint32_t n = cnt < (uint32_t)NC ? cnt : (uint32_t)NC;
    apply_hits(d, s, d.hits + (size_t)s * NC, n, mask0, sw_all + threadIdx.x, 0u);
}

// ------------------------------------------------- neighbour-list search
// Search mode 5.  k_topo_rows<true> rebuilds, per particle, the list of every
// sphere within r_a + r_b + skin (original indices) and keeps the position of
// the build; the force kernel requests a rebuild once any sphere has moved
// more than list_lim since.  list_lim is at most 0.45 skin, so no pair
// outside the reach can overlap before the rebuild (both spheres would have
// had to close a gap of more than skin).  A pair more than a cell apart at the
// build cannot overlap either: the build also scans the cells two away
// (list_shell) when a cell leaves less room than 0.9 skin beyond the largest
// pair, and where it cannot (a periodic axis of fewer than 5 cells) list_lim
// is at most half that room.  Testing the list pairs then finds exactly the
// overlaps the full search finds.
// The overlaps are keyed by the reference's visiting order -- the offset of
// the partner's cell (periodic wrap included, as findDistMod resolves it,
// src/contactsearch.cuh:12-74) and its sorted position -- and replayed
// through the slot rule in that order.
__device__ __forceinline__ bool cell_offset(int g, int b, int n, bool wrap, int& dxo, uint32_t& code) {
    const int raw = b - g;
    code = 0u;
    if (raw >= -1 && raw <= 1) { dxo = raw; return true; }
    if (wrap && g == 0 && b == n - 1) { dxo = -1; code = 1u; return true; }   // x - 1 wraps to n - 1: +L
    if (wrap && g == n - 1 && b == 0) { dxo = 1; code = 2u; return true; }    // x + 1 wraps to 0: -L
    return false;
}

constexpr int LIST_BATCH = 4;


// Cell coordinates of a system-local cell id (x fastest, then y, then z).
__device__ __forceinline__ void cell_xyz(const Sys& S, uint32_t key, int& gx, int& gy, int& gz) {
    const uint32_t z = fdiv(key, S.div_xy), r = key - z * S.div_xy.d, y = fdiv(r, S.div_x);
    gx = (int)(r - y * S.div_x.d);
    gy = (int)y;
    gz = (int)z;
}

// Full FP64 treatment of one list partner that passed the prefilter: its cell
// offset and image as the reference's search would reach it and the exact
// overlap test; returns the hit key, or 0 for no overlap (a real key is
// never 0: its sorted position differs from the particle's own).
__device__ __forceinline__ uint64_t list_pair(const Sys& S, uint32_t sb, uint32_t kb, double4 x, double4 xa,
                                              int gx, int gy, int gz) {
    int bx, by, bz, dx, dy, dz;
    uint32_t cx, cy, cz;
    cell_xyz(S, kb, bx, by, bz);
    if (!cell_offset(gz, bz, (int)S.num[2], false, dz, cz)) return 0ull;
    if (!cell_offset(gy, by, (int)S.num[1], S.periodic == 1, dy, cy)) return 0ull;
    if (!cell_offset(gx, bx, (int)S.num[0], S.periodic != 0, dx, cx)) return 0ull;
    const uint32_t code = cx | (cy << 2);
    (void)wrap_xy(S, x.x, x.y);
    const double dmx = cx == 1u ? S.L[0] : (cx == 2u ? -S.L[0] : 0.0);
    const double dmy = cy == 1u ? S.L[1] : (cy == 2u ? -S.L[1] : 0.0);
    const double ex = (xa.x - x.x) + dmx, ey = (xa.y - x.y) + dmy, ez = xa.z - x.z;
    const double d2 = ex * ex + ey * ey + ez * ez;
    if (!(sqrt(d2) - (xa.w + x.w) < 0.0)) return 0ull;  // exact test of src/contactsearch.cuh:242-248
    const uint32_t vi = (uint32_t)((dz + 1) * 9 + (dy + 1) * 3 + (dx + 1));
    return ((uint64_t)vi << 40) | ((uint64_t)code << 32) | sb;
}

// An overlap whose partner already holds a slot is applied on the spot (its
// slot does not depend on the order, a partner occurs once); the others are
// collected and given empty slots in key order at the end.  Each list entry
// keeps the slot its partner was last found in, with the partner's original
// index (both written when the search finds the partner in a slot; a build
// marks them unknown): one slot-word compare confirms the slot, and only a
// stale or unknown hint scans the occupied slots.
// Held contacts are skipped outright where the search could only rewrite
// their slot word unchanged.  That is so when no sphere crossed a periodic
// face in the last step: the image through which two overlapping spheres
// meet changes only when one of them wraps, and the word was written (or
// confirmed) by the last step's search, while the pair overlapped.  It is
// also so, in any step, when the word carries no periodic image and the
// particle's cell is off the periodic faces: any overlap the search finds
// with that partner has no image either (and one it does not find leaves
// the word as it is).  Skipped entries are not counted; a contact overflow is
// still caught when a new contact finds no empty slot.
template <bool B>
__global__ __launch_bounds__(APPLY_BLOCK, 4) void k_list_search(DemDev d) {  // four waves per SIMD: <= 128 VGPRs
    __shared__ uint32_t sw_all[NC * APPLY_BLOCK];
    if (d.nstate[0] != 0u) return;  // k_topo_rows<true> rebuilds the lists this step
    const uint32_t s = xcd_block() * APPLY_BLOCK + threadIdx.x;
    if (s >= d.np) return;
    const Sys& S = sys_at<B>(d, s);
    uint32_t* sw = sw_all + threadIdx.x;
    const uint32_t p = d.perm[s];
    const float4 xaf = d.xf[s];
    const float ra_eps = xaf.w + d.search_eps;
    const uint32_t li = d.slidx[s];
    const uint32_t* list = d.nlist + li;  // entry k at list[k * np]
    uint8_t* hint = d.nhint + li;
    uint32_t* lorig = d.nlo + li;
    const uint32_t nl = d.ncnt[li];
    const uint32_t mask0 = d.ro.cmask[p];
    const bool px = S.periodic != 0, py = S.periodic == 1;
    uint64_t* hk = d.hits + (size_t)s * NC;
    uint32_t n = 0, nnew = 0, err = 0, dirty = 0;
    bool have_a = false;
    double4 xa = make_double4(0.0, 0.0, 0.0, 0.0);
    int gx = 0, gy = 0, gz = 0;
    bool have_cell = false;
    bool skip_held = false;
    const bool no_wraps = d.nstate[4] == 0u;
    const uint32_t held_bits = no_wraps ? SLOT_IDX_MASK : 0xffffffffu;  // with wraps: the word must carry no image
    const SlotRow gslot = slot_row(d, s);
    if (mask0) {
        skip_held = no_wraps;
        if (!no_wraps) {
            cell_xyz(S, d.skey[s] - S.cbase, gx, gy, gz);
            have_cell = true;
            skip_held = (!px || (gx > 0 && gx < (int)S.num[0] - 1)) && (!py || (gy > 0 && gy < (int)S.num[1] - 1));
        }
        stage_slots(gslot, mask0, sw);
    }
    // A pair whose gap at the build exceeds the two spheres' displacements
    // since then cannot overlap now: skip it without touching the partner's
    // record (margins of three FP32 prefilter widths for the rounding of the
    // gap and of the displacements)
    const float da = d.ldisp[li] + 3.0f * d.search_eps;
    // No sphere has moved more than list_lim since the build (a step after
    // one that did rebuilds the lists instead): a gap beyond this cannot
    // close, whatever the partner's displacement, which is then not read.
    const float dfar = da + d.list_lim + 3.0f * d.search_eps;
    const float* gap = d.nlistg + li;
    // the next batch's entries are loaded while the current one is processed
    uint32_t qlb[LIST_BATCH], qhb[LIST_BATCH], qlw[LIST_BATCH];
    float qg[LIST_BATCH];
    auto fetch = [&](uint32_t i0) {
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) {
            const bool in = i0 + k < nl;
            qlb[k] = in ? list[(size_t)(i0 + k) * d.np] : li;
            qg[k] = in ? gap[(size_t)(i0 + k) * d.np] : 1e30f;
            qhb[k] = in && mask0 ? hint[(size_t)(i0 + k) * d.np] : 0xffu;
            qlw[k] = in && mask0 && skip_held ? lorig[(size_t)(i0 + k) * d.np] : 0xffffffffu;
        }
    };
    fetch(0);
    for (uint32_t i0 = 0; i0 < nl; i0 += LIST_BATCH) {
        uint32_t lb[LIST_BATCH], sb[LIST_BATCH], hb[LIST_BATCH], lw[LIST_BATCH];
        float g[LIST_BATCH];
        float4 c[LIST_BATCH];
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) {
            lb[k] = qlb[k];
            g[k] = qg[k];
            hb[k] = qhb[k];
            lw[k] = qlw[k];
        }
        if (i0 + LIST_BATCH < nl) fetch(i0 + LIST_BATCH);
        uint32_t held = 0, near = 0;
        if (mask0) {
#pragma unroll
            for (int k = 0; k < LIST_BATCH; ++k)
                held |= hb[k] < 32u && ((mask0 >> hb[k]) & 1u) && (sw[(hb[k] & 31u) * APPLY_BLOCK] & held_bits) == lw[k]
                            ? 1u << k : 0u;
        }
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k)
            near |= !((held >> k) & 1u) && g[k] <= dfar && g[k] <= da + d.ldisp[lb[k]] ? 1u << k : 0u;
        if (near == 0u) continue;
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) sb[k] = (near >> k) & 1u ? d.linv[lb[k]] : s;
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) c[k] = d.xf[sb[k]];
        // FP32 prefilter at the minimum image (an overlap is always closer
        // than half a period), the same margin as the cell search
        uint32_t pass = 0;
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) {
            float ex = xaf.x - c[k].x, ey = xaf.y - c[k].y;
            const float ez = xaf.z - c[k].z;
            if (px) ex -= S.Lf[0] * rintf(ex / S.Lf[0]);
            if (py) ey -= S.Lf[1] * rintf(ey / S.Lf[1]);
            const float d2 = ex * ex + ey * ey + ez * ez;
            const float rs = ra_eps + c[k].w;
            pass |= ((near >> k) & 1u) && d2 <= rs * rs ? 1u << k : 0u;
        }
        if (pass == 0u) continue;
        if (!have_a) {
            xa = spos(d, S, s);
            if (!have_cell) cell_xyz(S, d.skey[s] - S.cbase, gx, gy, gz);
            have_a = true;
        }
        // the batch's keys, FP64 positions and original indices, independent loads
        uint32_t kb[LIST_BATCH], bo[LIST_BATCH];
        double4 xb[LIST_BATCH];
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) {
            kb[k] = d.skey[sb[k]];
            xb[k] = d.ro.x[d.perm[sb[k]]];
            bo[k] = d.sid[sb[k]];
        }
#pragma unroll
        for (int k = 0; k < LIST_BATCH; ++k) {
            if (!((pass >> k) & 1u)) continue;
            const uint64_t v = list_pair(S, sb[k], kb[k] - S.cbase, xb[k], xa, gx, gy, gz);
            if (v == 0ull) continue;
            ++n;
            int cpos = -1;
            if (hb[k] < 32u && ((mask0 >> hb[k]) & 1u) && (sw[(hb[k] & 31u) * APPLY_BLOCK] & SLOT_IDX_MASK) == bo[k]) {
                cpos = (int)hb[k];
            } else {
                for (uint32_t m = mask0; m; m &= m - 1) {
                    const int i = __ffs(m) - 1;
                    if ((sw[i * APPLY_BLOCK] & SLOT_IDX_MASK) == bo[k]) { cpos = i; break; }
                }
                if (cpos >= 0) {  // the entry's hint and partner, written together
                    hint[(size_t)(i0 + k) * d.np] = (uint8_t)cpos;
                    lorig[(size_t)(i0 + k) * d.np] = bo[k];
                }
            }
            if (cpos >= 0) {  // kept; written back only if its image changed
                const uint32_t w = bo[k] | ((uint32_t)(v >> 32) & 15u) << 28;
                if (sw[cpos * APPLY_BLOCK] != w) {
                    sw[cpos * APPLY_BLOCK] = w;
                    dirty |= 1u << cpos;
                }
            } else if (nnew < (uint32_t)NC) {
                hk[nnew++] = v;
            }
        }
    }
    if (n > (uint32_t)NC) err |= FLAG_CONTACT_OVERFLOW;
    if (!have_a) {
        d.tmask[s] = mask0;
        d.tfresh[s] = 0u;
        return;
    }
    // new contacts: empty slots in key order
    for (uint32_t i = 1; i < nnew; ++i) {
        const uint64_t v = hk[i];
        uint32_t j = i;
        while (j > 0 && hk[j - 1] > v) { hk[j] = hk[j - 1]; --j; }
        hk[j] = v;
    }
    uint32_t mask = mask0, fresh = 0;
    for (uint32_t h = 0; h < nnew; ++h) {
        const uint64_t v = hk[h];
        const uint32_t empty = ~mask;
        if (empty == 0u) { err |= FLAG_CONTACT_OVERFLOW; break; }
        const int cpos = 31 - __clz(empty);
        fresh |= 1u << cpos;
        sw[cpos * APPLY_BLOCK] = d.sid[(uint32_t)(v & 0xffffffffu)] | ((uint32_t)(v >> 32) & 15u) << 28;
        dirty |= 1u << cpos;
        mask |= 1u << cpos;
    }
    for (uint32_t m = dirty; m; m &= m - 1) {
        const int i = __ffs(m) - 1;
        gslot[i] = sw[i * APPLY_BLOCK];
    }
    d.tmask[s] = mask;
    d.tfresh[s] = fresh;
    if (err) atomicOr(d.flags, err);
}

// One wavefront per cell.  The candidates of the cell's 27-cell neighbourhood,
// in the reference's visiting order (z, y, x cells, sorted order within a
// cell; src/contactsearch.cuh:340-353), are up to 27 runs of the sorted
// arrays; candidate t of the concatenation sits in lane t % 64 of chunk
// t / 64, in registers for the first SEARCH_REG_CHUNKS chunks.  Each member of
// the cell tests all candidates at once, one lane per candidate, and applies
// the overlaps (a ballot) in lane order, i.e. in the serial loop's order.
// While a member is processed its 32 slot words sit in lanes 0..31, so the
// slot rule (src/contactsearch.cuh:252-271) is one ballot per overlap.
constexpr int SEARCH_REG_CHUNKS = 2;

struct Cand {
    double x, y, z, r, dmx, dmy;
    uint32_t sb, bid, code;
    bool ok;
};

// Candidate t of the run list held in lanes 0..26 (start, exclusive prefix,
// length, image code); `runs` has a bit per non-empty run.
__device__ __forceinline__ Cand load_cand(const DemDev& d, uint32_t t, uint32_t T, uint64_t runs, uint32_t r_lo,
                                          uint32_t r_pre, uint32_t r_len, uint32_t r_code) {
    Cand c;
    uint32_t sb = 0, code = 0;
    for (uint64_t m = runs; m; m &= m - 1) {
        const int j = __ffsll((unsigned long long)m) - 1;
        const uint32_t pre = __builtin_amdgcn_readlane(r_pre, j);
        const uint32_t len = __builtin_amdgcn_readlane(r_len, j);
        if (t - pre < len) {
            sb = __builtin_amdgcn_readlane(r_lo, j) + (t - pre);
            code = __builtin_amdgcn_readlane(r_code, j);
        }
    }
    c.ok = t < T;
    c.sb = sb;
    c.code = code;
    if (c.ok) {
        const double4 xb = spos(d, d.sys, sb);
        c.x = xb.x; c.y = xb.y; c.z = xb.z; c.r = xb.w;
        c.bid = d.sid[sb];
    } else {
        c.x = c.y = c.z = c.r = 0.0;
        c.bid = 0;
    }
    decode_image(d.sys, code, c.dmx, c.dmy);
    return c;
}

struct MemberState {
    uint32_t mask, fresh, nc, err;
    bool stop, loaded;
};

// Overlaps of the member at sorted position m with one chunk of candidates.
__device__ __forceinline__ void member_chunk(const Cand& cb, uint32_t lane, uint32_t m, double4 xa,
                                             SlotRow slot, uint32_t& w, bool& dirty, MemberState& st) {
    const double ex = (xa.x - cb.x) + cb.dmx, ey = (xa.y - cb.y) + cb.dmy, ez = xa.z - cb.z;
    const double d2 = ex * ex + ey * ey + ez * ez;
    const double rsum = xa.w + cb.r;
    bool hit = false;
    if (__builtin_expect(cb.ok && d2 <= rsum * rsum * 1.0000001 && cb.sb != m, 0))
        hit = sqrt(d2) - rsum < 0.0;  // exact test of src/contactsearch.cuh:242-248 (double sqrt)
    uint64_t hits = __ballot(hit);
    while (hits) {
        const int l = __ffsll((unsigned long long)hits) - 1;
        hits &= hits - 1;
        const uint32_t bo = __builtin_amdgcn_readlane(cb.bid, l);
        const uint32_t code = __builtin_amdgcn_readlane(cb.code, l);
        if (!st.loaded) {
            w = lane < (uint32_t)NC ? slot[(int)lane] : 0u;
            st.loaded = true;
        }
        const uint64_t match = __ballot(lane < (uint32_t)NC && ((st.mask >> lane) & 1u) && (w & SLOT_IDX_MASK) == bo);
        int cpos;
        if (match) {
            cpos = __ffsll((unsigned long long)match) - 1;
        } else {
            const uint32_t empty = ~st.mask;
            if (empty == 0u) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; return; }
            cpos = 31 - __clz(empty);
            st.fresh |= 1u << cpos;
        }
        if (lane == (uint32_t)cpos) { w = bo | (code << 28); dirty = true; }
        st.mask |= 1u << cpos;
        if (++st.nc > (uint32_t)NC) { st.err |= FLAG_CONTACT_OVERFLOW; st.stop = true; return; }
    }
}

__global__ __launch_bounds__(BLOCK) void k_topo_wave(DemDev d) {
    const Sys& S = d.sys;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t c = xcd_block() * (BLOCK / 64) + (threadIdx.x >> 6);
    if (c >= d.K) return;
    const uint32_t m0 = d.off[c], m1 = d.off[c + 1];
    if (m0 == m1) return;
    const int nx = (int)S.num[0], ny = (int)S.num[1], nz = (int)S.num[2];
    const int gx = (int)(c % S.num[0]), gy = (int)((c / S.num[0]) % S.num[1]), gz = (int)(c / (S.num[0] * S.num[1]));

    // run j = lane (0..26): cell (gx + j%3 - 1, gy + (j/3)%3 - 1, gz + j/9 - 1);
    // a row without periodic image is one run, held by its dx = -1 lane
    uint32_t r_lo = 0, r_len = 0, r_code = 0;
    if (lane < 27u) {
        const int dz = (int)(lane / 9u) - 1, dy = (int)((lane / 3u) % 3u) - 1, dx = (int)(lane % 3u) - 1;
        const int tz = gz + dz;
        int ty = gy + dy, tx = gx + dx;
        bool ok = tz >= 0 && tz < nz;
        uint32_t cy = 0, cx = 0;
        if (ok && (ty < 0 || ty == ny)) {
            if (S.periodic == 1) { cy = ty < 0 ? 1u : 2u; ty = ty < 0 ? ny - 1 : 0; }
            else ok = false;
        }
        if (ok) {
            const uint32_t row = S.cbase + (uint32_t)ty * S.num[0] + (S.num[0] * S.num[1]) * (uint32_t)tz;
            if (gx >= 1 && gx + 1 < nx && cy == 0) {
                if (dx == -1) { r_lo = d.off[row + gx - 1]; r_len = d.off[row + gx + 2] - r_lo; }
            } else {
                if (tx < 0 || tx == nx) {
                    if (S.periodic != 0) { cx = tx < 0 ? 1u : 2u; tx = tx < 0 ? nx - 1 : 0; }
                    else ok = false;
                }
                if (ok) {
                    r_lo = d.off[row + tx];
                    r_len = d.off[row + tx + 1] - r_lo;
                    r_code = cx | (cy << 2);
                }
            }
        }
    }
    uint32_t incl = r_len;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if (lane >= (uint32_t)o) incl += v;
    }
    const uint32_t r_pre = incl - r_len;
    const uint32_t T = __shfl(incl, 26, 64);
    const uint64_t runs = __ballot(r_len != 0u);

    Cand cand[SEARCH_REG_CHUNKS];
#pragma unroll
    for (int j = 0; j < SEARCH_REG_CHUNKS; ++j)
        cand[j] = load_cand(d, 64u * j + lane, T, runs, r_lo, r_pre, r_len, r_code);

    for (uint32_t m = m0; m < m1; ++m) {
        const double4 xa = spos(d, S, m);
        const SlotRow slot = slot_row(d, m);
        MemberState st;
        st.mask = d.ro.cmask[d.perm[m]];
        st.fresh = 0; st.nc = 0; st.err = 0; st.stop = false; st.loaded = false;
        uint32_t w = 0;
        bool dirty = false;
#pragma unroll
        for (int j = 0; j < SEARCH_REG_CHUNKS; ++j)
            if (64u * j < T && !st.stop) member_chunk(cand[j], lane, m, xa, slot, w, dirty, st);
        for (uint32_t base = 64u * SEARCH_REG_CHUNKS; base < T && !st.stop; base += 64u) {
            const Cand cb = load_cand(d, base + lane, T, runs, r_lo, r_pre, r_len, r_code);
            member_chunk(cb, lane, m, xa, slot, w, dirty, st);
        }
        if (dirty) slot[(int)lane] = w;
        if (lane == 0) {
            d.tmask[m] = st.mask;
            d.tfresh[m] = st.fresh;
            if (st.err) atomicOr(d.flags, st.err);
        }
    }
}

// ------------------------------------------------------------ force kernel
// One thread per sorted particle: contact forces from the occupied slots in
// slot order, wall contacts, bond forces, leapfrog.
constexpr int FORCE_BATCH = 4;

// VW: the partners' velocity records are gathered with their positions, ahead
// of the slots' processing (more registers, fewer waves: faster with many
// contacts per sphere, slower with few; the host picks by timing).  M3:
// contact model 3's pair force (FP64, one system) instead of model 2's.  The
// one-system FP64 kernel without the gathers ahead is held to three waves per
// SIMD (168 VGPRs, no spills): with two it loses 4-8 % on both a contact-free
// packing and the dense crystal.
template <class T, bool B, bool VW, bool M3 = false>
__global__ __launch_bounds__(BLOCK, (sizeof(T) == 8 && !B && !VW && !M3) ? 3 : 1) void k_force(DemDev d) {
    static_assert(!M3 || (sizeof(T) == 8 && !B), "contact model 3: FP64, one system");
    const uint32_t s = xcd_block() * blockDim.x + threadIdx.x;
    if (s >= d.np) return;
    const Sys& S = sys_at<B>(d, s);
    const uint32_t o = d.sid[s];
    const uint32_t prev = d.perm[s];  // where this particle's records are stored
    double4 xa = d.ro.x[prev];
    (void)wrap_xy(S, xa.x, xa.y);
    const double ra = xa.w;

    // contact slots after the search kernel (contact model 1 keeps none)
    const bool viscous = S.contactmodel == 1;
    uint32_t mask = viscous ? 0u : d.tmask[s];
    const uint32_t fresh = viscous ? 0u : d.tfresh[s];  // slots taken from empty this step: delta_t starts at 0
    const SlotRow slot{d.ro.cslot + prev, d.np};

    // ---- forces from the occupied slots in slot order (src/contactsearch.cuh:479-543)
    const double4 va4 = d.ro.vel[prev], wa4 = d.ro.angvel[prev];
    const V3<double> va = xyz(va4), wa = xyz(wa4);
    Acc acc;
    acc.F = mk(0.0, 0.0, 0.0);
    acc.T = mk(0.0, 0.0, 0.0);
    acc.es_dot = 0.0; acc.ev_dot = 0.0; acc.p = 0.0;
    // Slots in groups of FORCE_BATCH: the group's dependent gathers (slot
    // word -> where the partner's record is stored, pos_of still holding the
    // storage positions -> the record) are issued together, then the slots
    // are processed one by one in slot order.
    for (uint32_t m = mask; m;) {
        int si[FORCE_BATCH];
        uint32_t sw[FORCE_BATCH], pb[FORCE_BATCH];
        int cnt = 0;
#pragma unroll
        for (int k = 0; k < FORCE_BATCH; ++k) {
            si[k] = m ? __ffs(m) - 1 : -1;
            if (m) { m &= m - 1; ++cnt; }
        }
#pragma unroll
        for (int k = 0; k < FORCE_BATCH; ++k) sw[k] = k < cnt ? slot[si[k]] : 0u;
#pragma unroll
        for (int k = 0; k < FORCE_BATCH; ++k) pb[k] = k < cnt ? d.pos_of[sw[k] & SLOT_IDX_MASK] : 0u;
        using PV = PartnerVel<T>;
        double4 xbs[FORCE_BATCH];
        typename PV::type vbs[VW ? FORCE_BATCH : 1], wbs[VW ? FORCE_BATCH : 1];
#pragma unroll
        for (int k = 0; k < FORCE_BATCH; ++k) xbs[k] = k < cnt ? d.ro.x[pb[k]] : make_double4(0.0, 0.0, 0.0, 0.0);
        if constexpr (VW) {
#pragma unroll
            for (int k = 0; k < FORCE_BATCH; ++k) {
                vbs[k] = k < cnt ? PV::vel(d, pb[k]) : PV::zero();
                wbs[k] = k < cnt ? PV::angvel(d, pb[k]) : PV::zero();
            }
        }
#pragma unroll
        for (int k = 0; k < FORCE_BATCH; ++k) {
            if (k >= cnt) break;
            const int i = si[k];
            double4 xb = xbs[k];
            (void)wrap_xy(S, xb.x, xb.y);
            double dmx, dmy;
            decode_image(S, sw[k] >> 28, dmx, dmy);
            const V3<double> xab = mk((xa.x - xb.x) + dmx, (xa.y - xb.y) + dmy, (xa.z - xb.z) + 0.0);
            const double len = vlen(xab);
            const double delta_n = len - (ra + xb.w);
            if (delta_n < 0.0) {
                V3<T> dt0 = (fresh >> i) & 1u ? mk<T>(0, 0, 0) : SlotStore<T>::load(d, i, prev);
                const typename PV::type vb = VW ? vbs[VW ? k : 0] : PV::vel(d, pb[k]);
                const typename PV::type wb = VW ? wbs[VW ? k : 0] : PV::angvel(d, pb[k]);
                if constexpr (M3)
                    contact_hertz(S, acc, va, wa, PV::v3(vb), PV::v3(wb), ra, xb.w, xab, len, delta_n, dt0);
                else
                    contact_linear<T>(S, acc, cvt<T>(va), cvt<T>(wa), PV::v3(vb), PV::v3(wb), (T)ra,
                                      (T)xb.w, cvt<T>(xab), (T)len, (T)delta_n, dt0);
                SlotStore<T>::store(d, i, s, dt0);
                d.wo.cslot[(size_t)i * d.np + s] = sw[k];  // the contact lasts: its slot word moves with it
            } else {
                mask &= ~(1u << i);  // contact ended: the slot is empty again
            }
        }
    }

    // ---- contact model 1: contacts found and processed at once over the 27
    // cells in z, y, x order, sorted order within a cell, with the capillary
    // bond for gaps below db (src/contactsearch.cuh:76-173,548-576)
    if (viscous) {
        const int nx = (int)S.num[0], ny = (int)S.num[1], nz = (int)S.num[2];
        const uint32_t key = d.skey[s] - S.cbase;
        const int gx = (int)(key % S.num[0]), gy = (int)((key / S.num[0]) % S.num[1]);
        const int gz = (int)(key / (S.num[0] * S.num[1]));
        for (int dz = -1; dz <= 1; ++dz) {
            const int tz = gz + dz;
            if (tz < 0 || tz == nz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int ty = gy + dy;
                double dmy = 0.0;
                if (ty < 0 || ty == ny) {
                    if (S.periodic != 1) continue;
                    dmy = ty < 0 ? S.L[1] : -S.L[1];
                    ty = ty < 0 ? ny - 1 : 0;
                }
                for (int dx = -1; dx <= 1; ++dx) {
                    int tx = gx + dx;
                    double dmx = 0.0;
                    if (tx < 0 || tx == nx) {
                        if (S.periodic == 0) continue;
                        dmx = tx < 0 ? S.L[0] : -S.L[0];
                        tx = tx < 0 ? nx - 1 : 0;
                    }
                    const uint32_t c = S.cbase + (uint32_t)tx + (uint32_t)ty * S.num[0] + (S.num[0] * S.num[1]) * (uint32_t)tz;
                    const uint32_t hi = d.off[c + 1];
                    for (uint32_t sb = d.off[c]; sb < hi; ++sb) {
                        if (sb == s) continue;
                        const uint32_t pb = d.perm[sb];
                        double4 xb = d.ro.x[pb];
                        (void)wrap_xy(S, xb.x, xb.y);
                        const V3<double> xab = mk((xa.x - xb.x) + dmx, (xa.y - xb.y) + dmy, (xa.z - xb.z) + 0.0);
                        const double len = vlen(xab);
                        const double delta = len - (ra + xb.w);
                        if (delta < 0.0)
                            contact_viscous(S, acc, va, wa, xyz(d.ro.vel[pb]), xyz(d.ro.angvel[pb]), ra, xb.w, xab, len,
                                            delta);
                        else if (delta < S.db)
                            capillary(S, acc, ra, xb.w, delta, xab, len);
                    }
                }
            }
        }
    }

    // ---- walls (src/contactsearch.cuh:587-663)
    double wf0 = 0.0, wf1 = 0.0, wf2 = 0.0, wf3 = 0.0, wf4 = 0.0;
    {
        const double4 w0 = d.wnx[S.wbase + 0];
        const double dw = w0.w - (xa.z + ra);
        if (dw < 0.0) wf0 = contact_wall(S, acc, va, wa, ra, mk(w0.x, w0.y, w0.z), dw, d.wmvfd[S.wbase + 0].y);
    }
    {
        const double dw = xa.z - ra - S.origo[2];
        if (dw < 0.0) (void)contact_wall(S, acc, va, wa, ra, mk(0.0, 0.0, 1.0), dw, 0.0);
    }
    if (S.periodic == 0) {
        const double4 w1 = d.wnx[S.wbase + 1], w2 = d.wnx[S.wbase + 2];
        double dw = w1.w - (xa.x + ra);
        if (dw < 0.0) wf1 = contact_wall(S, acc, va, wa, ra, mk(w1.x, w1.y, w1.z), dw, d.wmvfd[S.wbase + 1].y);
        dw = xa.x - ra - w2.w;
        if (dw < 0.0) wf2 = contact_wall(S, acc, va, wa, ra, mk(w2.x, w2.y, w2.z), dw, d.wmvfd[S.wbase + 2].y);
    }
    if (S.periodic == 0 || S.periodic == 2) {
        const double4 w3 = d.wnx[S.wbase + 3], w4 = d.wnx[S.wbase + 4];
        double dw = w3.w - (xa.y + ra);
        if (dw < 0.0) wf3 = contact_wall(S, acc, va, wa, ra, mk(w3.x, w3.y, w3.z), dw, d.wmvfd[S.wbase + 3].y);
        dw = xa.y - ra - w4.w;
        if (dw < 0.0) wf4 = contact_wall(S, acc, va, wa, ra, mk(w4.x, w4.y, w4.z), dw, d.wmvfd[S.wbase + 4].y);
    }

    // ---- per-particle outputs (src/contactsearch.cuh:669-687)
    const double dt = S.dt;
    const double4 q0 = d.ro.pr0[prev], q1 = d.ro.pr1[prev], q2 = d.ro.pr2[prev];
    double4 ea = make_double4(q2.y, q2.z, q2.w, wa4.w);  // xysum.x, xysum.y, es, ev
    ea.z += acc.es_dot * dt;
    ea.w += acc.ev_dot * dt;
    d.pres[s] = acc.p;
    if (S.nw > 0) d.wall_pp[o] = wf0;
    if (S.nw > 1) d.wall_pp[(size_t)d.np + o] = wf1;
    if (S.nw > 2) d.wall_pp[2 * (size_t)d.np + o] = wf2;
    if (S.nw > 3) d.wall_pp[3 * (size_t)d.np + o] = wf3;
    if (S.nw > 4) d.wall_pp[4 * (size_t)d.np + o] = wf4;

    // ---- bond forces in ascending bond index (src/cohesion.cuh:220-226)
    V3<double> force = acc.F, torque = acc.T;
    if (d.nb0) {
        const uint32_t e0 = d.bcsr_off[o], e1 = d.bcsr_off[o + 1];
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t ent = d.bcsr_ent[e];
            const uint32_t k = ent >> 1;
            if (!(d.bonds[k].x < d.np)) continue;
            const V3<double> f = xyz(d.bond_f[k]), t = xyz(d.bond_t[k]);
            if ((ent & 1u) == 0u) { force = force + f; torque = torque - t; }
            else { force = force - f; torque = torque + t; }
        }
    }
    d.force[s] = make_double4(force.x, force.y, force.z, acc.es_dot);
    d.torque[s] = make_double4(torque.x, torque.y, torque.z, acc.ev_dot);

    // ---- leapfrog (src/integration.cuh:20-185)
    const double radius = ra;
    const double m = 4.0 / 3.0 * PI_REF * radius * radius * radius * S.rho;
    double ax = force.x / m, ay = force.y / m, az = force.z / m;
    double bx = torque.x * 1.0 / (2.0 / 5.0 * m * radius * radius);
    double by = torque.y * 1.0 / (2.0 / 5.0 * m * radius * radius);
    double bz = torque.z * 1.0 / (2.0 / 5.0 * m * radius * radius);
    ax += S.g[0]; ay += S.g[1]; az += S.g[2];
    if (va4.w > 0.0) {
        ax = 0.0; ay = 0.0; az -= S.g[2];
        if (va4.w >= SPH_FIXVEL_HOLD_Z) az = 0.0;  // extension: the z velocity held as well
        bx = 0.0; by = 0.0; bz = 0.0;
    }
    double px = xa.x, py = xa.y, pz = xa.z;
    (void)wrap_xy(S, px, py);
    double4 v0 = make_double4(q0.x, q0.y, q0.z, 0.0), w0 = make_double4(q0.w, q1.x, q1.y, 0.0);
    double4 ap = make_double4(q1.z, q1.w, q2.x, 0.0);
    v0.x += ax * dt; v0.y += ay * dt; v0.z += az * dt;
    w0.x += bx * dt; w0.y += by * dt; w0.z += bz * dt;
    px += v0.x * dt; py += v0.y * dt; pz += v0.z * dt;
    ap.x += w0.x * dt; ap.y += w0.y * dt; ap.z += w0.z * dt;
    const double vx = v0.x + 0.5 * ax * dt, vy = v0.y + 0.5 * ay * dt, vz = v0.z + 0.5 * az * dt;
    const double wx = w0.x + 0.5 * bx * dt, wy = w0.y + 0.5 * by * dt, wz = w0.z + 0.5 * bz * dt;
    double sx = 0.0, sy = 0.0;
    sx += vx * dt;
    sy += vy * dt;
    ea.x += sx;
    ea.y += sy;
    d.wo.pr0[s] = make_double4(v0.x, v0.y, v0.z, w0.x);
    d.wo.pr1[s] = make_double4(w0.y, w0.z, ap.x, ap.y);
    d.wo.pr2[s] = make_double4(ap.z, ea.x, ea.y, ea.z);
    d.wo.angvel[s] = make_double4(wx, wy, wz, ea.w);
    d.wo.vel[s] = make_double4(vx, vy, vz, va4.w);
    if constexpr (sizeof(T) == 4) {
        d.wo.vf[s] = make_float4((float)vx, (float)vy, (float)vz, 0.0f);
        d.wo.wf[s] = make_float4((float)wx, (float)wy, (float)wz, 0.0f);
    }
    d.wo.x[s] = make_double4(px, py, pz, radius);
    d.wo.id[s] = o;
    d.wo.cmask[s] = mask;

    // cell id of the new position for the next step's sort
    uint32_t cerr = 0;
    const uint32_t nkey = cell_key(S, px, py, pz, cerr);
    d.key[s] = nkey;
    if (cerr) atomicOr(d.flags_next, cerr);
    count_key(d, s, nkey);
    if (d.search_mode == 5) {
        // the next step wraps the new position: the sphere crossed a periodic
        // face (the image through which its contacts meet may change)
        double wx_ = px, wy_ = py;
        if (wrap_xy(S, wx_, wy_)) atomicOr(&d.nstate[5], 1u);
    }

    // neighbour lists (search mode 5): the list index travels with the
    // particle; request a rebuild once it moved list_lim since the build
    if (d.search_mode == 5) {
        const uint32_t li = d.nstate[0] ? s : d.slidx[s];
        d.wo.lidx[s] = li;
        const float4 xr = d.xref[li];
        float ddx = (float)px - xr.x, ddy = (float)py - xr.y;
        const float ddz = (float)pz - xr.z;
        if (S.periodic == 1 || S.periodic == 2) ddx -= S.Lf[0] * rintf(ddx / S.Lf[0]);
        if (S.periodic == 1) ddy -= S.Lf[1] * rintf(ddy / S.Lf[1]);
        const float dd2 = ddx * ddx + ddy * ddy + ddz * ddz;
        d.ldisp[li] = sqrtf(dd2);
        if (dd2 > d.list_lim * d.list_lim) atomicOr(&d.nstate[1], 1u);
    }
}

// FP32 copies of the input set's velocity records (mixed mode), by storage
// position.
__global__ __launch_bounds__(BLOCK) void k_f32_copies(DemDev d) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.np) return;
    const double4 v = d.ro.vel[q], w = d.ro.angvel[q];
    d.ro.vf[q] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.0f);
    d.ro.wf[q] = make_float4((float)w.x, (float)w.y, (float)w.z, 0.0f);
}

// After the force kernel: pos_of[id] = s for the particles whose sorted
// position changed this step (most keep theirs from one step to the next,
// which spares the scattered write).  Until this kernel pos_of holds the
// storage positions -- the previous step's sorted positions, where the
// input set's records are -- which is what the force kernel and the next
// step's bond kernel gather partner records through.
__global__ __launch_bounds__(BLOCK) void k_pos_of(DemDev d) {
    const uint32_t q = xcd_block() * blockDim.x + threadIdx.x;
    if (q >= d.np) return;
    if (d.perm[q] != q) d.pos_of[d.sid[q]] = q;
}

// ------------------------------------------------------------------- walls
// src/integration.cuh:191-221: blocks of 256 original indices, tree halving.
// In an ensemble each system's blocks start at its first index, so every
// system is summed exactly as on its own.
template <bool B>
__global__ __launch_bounds__(BLOCK) void k_wallsum(DemDev d) {
    __shared__ double cache[BLOCK];
    const uint32_t w = blockIdx.y;
    const Sys& S = B ? d.systems[d.sys_of_block[blockIdx.x]] : d.sys;
    const uint32_t i = S.pbase + (blockIdx.x - S.bbase) * BLOCK + threadIdx.x;
    double temp = 0.0;
    if (i < S.pbase + S.np) temp += d.wall_pp[(size_t)w * d.np + i];
    cache[threadIdx.x] = temp;
    __syncthreads();
    for (uint32_t h = BLOCK / 2; h != 0; h >>= 1) {
        if (threadIdx.x < h) cache[threadIdx.x] += cache[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.wall_partial[(size_t)w * d.nwall_blocks + blockIdx.x] = cache[0];
}

// src/integration.cuh:224-295, one thread per wall; then time.current += dt.
// The wall force is the serial sum of the block partials in block order; the
// block stages the partials through LDS in chunks so the summing thread reads
// LDS instead of waiting on HBM for every term.  One block per system.
constexpr uint32_t WALL_CHUNK = 4096;
template <bool B>
__global__ __launch_bounds__(BLOCK) void k_walls(DemDev d) {
    __shared__ double part[WALL_CHUNK];
    __shared__ double fz[6];
    const uint32_t w = threadIdx.x;
    const Sys& S = B ? d.systems[blockIdx.x] : d.sys;
    const uint32_t k = B ? blockIdx.x : 0u;
    const double t_current = d.t_cur[k];
    for (uint32_t wall = 0; wall < S.nw; ++wall) {
        if (d.wmode[S.wbase + wall] == 0) continue;
        const double* src = d.wall_partial + (size_t)wall * d.nwall_blocks + S.bbase;
        double acc = 0.0;
        for (uint32_t c0 = 0; c0 < S.nblocks; c0 += WALL_CHUNK) {
            const uint32_t cn = S.nblocks - c0 < WALL_CHUNK ? S.nblocks - c0 : WALL_CHUNK;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < cn; i += BLOCK) part[i] = src[c0 + i];
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t i = 0;
                for (; i + 8 <= cn; i += 8) {
                    const double p0 = part[i], p1 = part[i + 1], p2 = part[i + 2], p3 = part[i + 3];
                    const double p4 = part[i + 4], p5 = part[i + 5], p6 = part[i + 6], p7 = part[i + 7];
                    acc += p0; acc += p1; acc += p2; acc += p3;
                    acc += p4; acc += p5; acc += p6; acc += p7;
                }
                for (; i < cn; ++i) acc += part[i];
            }
        }
        if (threadIdx.x == 0) fz[wall] = acc;
    }
    __syncthreads();
    if (w < S.nw && d.wmode[S.wbase + w] != 0) {
        const int wmode = d.wmode[S.wbase + w];
        double4 w_nx = d.wnx[S.wbase + w], w_mvfd = d.wmvfd[S.wbase + w];
        double vel0 = d.wvel0[S.wbase + w];
        w_mvfd.z = fz[w];
        const double dt = S.dt;
        const double sigma_0 = w_mvfd.w + S.devs_A * sin(2.0 * PI_WALL * S.devs_f * t_current);
        const double N = -sigma_0 * S.L[0] * S.L[1];
        double acc = (w_mvfd.z + N) / w_mvfd.x;
        if (wmode == 2) acc = 0.0;
        vel0 += acc * dt;
        w_nx.w += vel0 * dt;
        w_mvfd.y = vel0 + 0.5 * acc * dt;
        d.wnx[S.wbase + w] = w_nx;
        d.wmvfd[S.wbase + w] = w_mvfd;
        d.wvel0[S.wbase + w] = vel0;
    }
    __syncthreads();
    if (threadIdx.x == 0) d.t_cur[k] = t_current + S.dt;
}

// Plain streaming copy used to calibrate bandwidth counters.
__global__ __launch_bounds__(BLOCK) void k_copy_probe(const double4* __restrict__ src, double4* __restrict__ dst,
                                                      uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Counter calibration: every thread runs `iters` rounds of 8 independent
// fused multiply-adds, i.e. exactly 16 * iters floating-point operations,
// and stores one value so that nothing is removed.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_flops_probe(T* __restrict__ out, uint32_t iters, T a, T b) {
    T acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = (T)(threadIdx.x + k);
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] = fma(acc[k], a, b);
    }
    T sum = acc[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) sum += acc[k];  // (7 adds, outside the counted work's 16 * iters)
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

inline uint32_t nblk(uint64_t n, uint32_t b = BLOCK) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n) k_fill_u32<<<nblk(n), BLOCK, 0, s>>>(p, n, v);
}
void launch_bonds(const DemDev& d, hipStream_t s) {
    if (d.nb0) k_bonds<<<nblk(d.nb0), BLOCK, 0, s>>>(d);
}
void launch_cellid(const DemDev& d, hipStream_t s) {
    if (d.systems) k_cellid<true><<<nblk(d.np), BLOCK, 0, s>>>(d);
    else k_cellid<false><<<nblk(d.np), BLOCK, 0, s>>>(d);
}
void launch_scan(const DemDev& d, hipStream_t s) {
    k_scan_reduce<<<d.scan_blocks, BLOCK, 0, s>>>(d);
    k_scan_down<<<d.scan_blocks, BLOCK, 0, s>>>(d);
}
void launch_scatter(const DemDev& d, hipStream_t s) { k_scatter<<<nblk(d.np), BLOCK, 0, s>>>(d); }
void launch_cellsort(const DemDev& d, hipStream_t s) { k_cellsort<<<nblk(d.K), BLOCK, 0, s>>>(d); }
void launch_reorder(const DemDev& d, hipStream_t s) {
    if (d.systems) k_reorder<true><<<nblk(d.np), BLOCK, 0, s>>>(d);
    else k_reorder<false><<<nblk(d.np), BLOCK, 0, s>>>(d);
}
template <bool B>
void topology(const DemDev& d, hipStream_t s) {
    if (!B && d.sys.contactmodel == 1) return;  // contact model 1 keeps no slots (src/device.cu:809)
    if (d.search_mode == 5) {  // one of the two returns at once (nstate[0])
        k_topo_rows<true, B><<<nblk(d.np, d.tpb_search), d.tpb_search, NC * 4 * d.tpb_search, s>>>(d);
        k_list_search<B><<<nblk(d.np, APPLY_BLOCK), APPLY_BLOCK, 0, s>>>(d);
    } else if (d.search_mode == 4) {  // hcnt starts at zero: k_topo_apply clears it
        k_topo_half<B><<<nblk(d.np, d.tpb_search), d.tpb_search, 0, s>>>(d);
        k_topo_apply<<<nblk(d.np, APPLY_BLOCK), APPLY_BLOCK, 0, s>>>(d);
    } else if (d.search_mode == 2) k_topo_rows<false, B><<<nblk(d.np, d.tpb_search), d.tpb_search, NC * 4 * d.tpb_search, s>>>(d);
    else if (B) return;  // (modes 1 and 3 walk one grid; the host refuses them for ensembles)
    else if (d.search_mode == 3) k_topo_wave<<<(d.K + BLOCK / 64 - 1) / (BLOCK / 64), BLOCK, 0, s>>>(d);
    else k_topo<<<d.tiles_x * d.tiles_y * d.sys.num[2], BLOCK, 0, s>>>(d);
}
void launch_topology(const DemDev& d, hipStream_t s) {
    if (d.systems) topology<true>(d, s);
    else topology<false>(d, s);
}
template <bool B, bool VW>
void forces(const DemDev& d, bool mixed, hipStream_t s) {
    if (mixed) k_force<float, B, VW><<<nblk(d.np, d.tpb_force), d.tpb_force, 0, s>>>(d);
    else if (!B && d.sys.contactmodel == 3) k_force<double, false, VW, true><<<nblk(d.np, d.tpb_force), d.tpb_force, 0, s>>>(d);
    else k_force<double, B, VW><<<nblk(d.np, d.tpb_force), d.tpb_force, 0, s>>>(d);
}
void launch_forces(const DemDev& d, int precision, hipStream_t s) {
    const bool mixed = precision == SPH_PREC_MIXED;
    if (d.systems) {
        if (d.force_vw) forces<true, true>(d, mixed, s);
        else forces<true, false>(d, mixed, s);
    } else {
        if (d.force_vw) forces<false, true>(d, mixed, s);
        else forces<false, false>(d, mixed, s);
    }
    k_pos_of<<<nblk(d.np), BLOCK, 0, s>>>(d);
}
void launch_contact(const DemDev& d, int precision, hipStream_t s) {
    launch_topology(d, s);
    launch_forces(d, precision, s);
}
void launch_copy_probe(const double4* src, double4* dst, uint64_t n, hipStream_t s) {
    k_copy_probe<<<2048, BLOCK, 0, s>>>(src, dst, n);
}
void launch_flops_probe(void* out, uint32_t blocks, uint32_t iters, bool fp64, hipStream_t s) {
    if (fp64) k_flops_probe<double><<<blocks, BLOCK, 0, s>>>(static_cast<double*>(out), iters, 0.999999, 1e-7);
    else k_flops_probe<float><<<blocks, BLOCK, 0, s>>>(static_cast<float*>(out), iters, 0.999999f, 1e-7f);
}
void launch_f32_copies(const DemDev& d, hipStream_t s) { k_f32_copies<<<nblk(d.np), BLOCK, 0, s>>>(d); }
void launch_walls(const DemDev& d, hipStream_t s) {
    if (d.systems) {
        if (d.nw) k_wallsum<true><<<dim3(d.nwall_blocks, d.nw), BLOCK, 0, s>>>(d);
        k_walls<true><<<d.nsys, BLOCK, 0, s>>>(d);
    } else {
        if (d.nw) k_wallsum<false><<<dim3(d.nwall_blocks, d.nw), BLOCK, 0, s>>>(d);
        k_walls<false><<<1, BLOCK, 0, s>>>(d);
    }
}

}  // namespace sph
