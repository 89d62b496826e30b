// adjoint.hip -- hand-written discrete adjoint of the DEM time step.
//
// No counterpart in the reference (it has no derivatives).  The objective is
// a time mean of weighted particle forces,
//
//     J = 1/N  sum_{n < N}  sum_i  w_i . F_i(n),
//
// F_i(n) the total force of step n (the shear stress of a sheared layer is
// w_i = (-1/(L0 L1), 0, 0) on the driven particles), and the gradient is
// taken with respect to theta = (k_n, k_t, gamma_n, mu) with mu = mu_s = mu_d.
//
// Forward (sph_adjoint_forward): ordinary steps, a full-state checkpoint at
// the start of every segment of `seg` steps, J accumulated on the device.
// Backward (sph_adjoint_backward): for each segment from the last, restore
// its checkpoint, recompute its steps while taping each step's sorted state,
// contact slots and incoming springs (taken after the contact search, before
// the force kernel), then run the adjoint steps in reverse over the tape.
//
// The adjoint state lambda is kept in original particle order: positions,
// half-step and full-step velocities and angular velocities, the tangential
// spring of every contact slot, and per wall its position, velocity and
// half-step velocity.  One adjoint step, per particle of the taped order:
//   1. leapfrog backward (src/integration.cuh:20-185): lambda of the next
//      step's x, vel0, vel, angvel0, angvel -> adjoint of force and torque;
//   2. the objective's w_i / N is added to the force adjoint;
//   3. for every contact of the step the pair force of src/contactmodels.cuh:
//      283-493 is recomputed and its vector-Jacobian product taken (derived by
//      hand below); the partner's share goes to its lambda with atomics;
//   4. the same for the wall contacts (src/contactmodels.cuh:10-109) and the
//      wall leapfrog (src/integration.cuh:224-295).
// The contact set, slot positions and the branches (normal clamp, Coulomb
// slip) are those of the taped forward step; the derivative is that of the
// step with its branches fixed, which is what a small central difference
// sees when no branch flips.  Bonds are not differentiated (a handle with
// bonds is rejected).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine_internal.h"
#include "vec3.h"

namespace sph {

namespace {

constexpr double PI_REF = 3.14159265358979;  // src/constants.h:14
constexpr int NWALL = 6;

// ---------------------------------------------------------------- adjoint
struct Lam {
    double4 *x, *v0, *v, *w0, *w;  // original order; .w unused
    double* dt;                    // np * NC * 3
    double* wall;                  // NWALL x (pos, vel, vel0)
};

// One taped step: its sorted input state, slots and springs after the contact
// search, before the force kernel.
struct TapeView {
    const double4 *x, *vel, *angvel;
    const uint32_t *id, *pos_of, *tmask, *tfresh, *cslot;
    const double* cdt;
    const float* cdt32;
    const double4 *wnx, *wmvfd;  // NWALL each, wall state at the step's start
};

struct AdjDev {
    TapeView t;
    Lam next, cur;
    double4 *gfa, *gfb;   // cotangents of each particle's force and torque, sorted order of the step:
                          // (gF.x, gF.y, gF.z, gT.x), (gT.y, gT.z, -, -)
    const double4* wobj;  // objective weights, original order
    double inv_n;         // 1 / number of steps of the objective
    const double* lfw;    // adjoint of each wall's force sum for this step
    double* theta;        // d J / d (k_n, k_t, gamma_n, mu)
    const int* wmode;
    double wmass[NWALL];
};

template <class T> __device__ __forceinline__ V3<double> up(V3<T> v) {
    return mk<double>((double)v.x, (double)v.y, (double)v.z);
}

// Vector-Jacobian product of contact_linear (the pair force of
// src/contactmodels.cuh:283-493 as the force kernel evaluates it).  Inputs:
// relative position x (length len, overlap delta), velocities, angular
// velocities, radii, incoming spring dt0u; cotangents gF, gT, gD of the
// force, torque and outgoing spring.  Accumulates the cotangents of the
// inputs and of (k_n, k_t, gamma_n, mu_s, mu_d).
template <class T> struct PairGrad {
    V3<T> gx, gva, gvb, gwa, gwb, gdt0u;
    T gkn, gkt, ggn, gmus, gmud;
};

template <class T>
__device__ __forceinline__ void contact_linear_vjp(const DemDev& d, V3<T> va, V3<T> wa, V3<T> vb, V3<T> wb, T ra,
                                                   T rb, V3<T> x, T len, T delta, V3<T> dt0u, V3<T> gF, V3<T> gT,
                                                   V3<T> gD, PairGrad<T>& g) {
    const T k_n = (T)d.sys.k_n, k_t = (T)d.sys.k_t, gamma_n = (T)d.sys.gamma_n, gamma_t = (T)d.sys.gamma_t;
    const T mu_s = (T)d.sys.mu_s, mu_d = (T)d.sys.mu_d, kappa = (T)d.sys.kappa, dt = (T)d.sys.dt;
    const T zero = (T)0;
    // ---- forward recompute
    const V3<T> n = x / len;
    const V3<T> vl = va - vb;
    const T ha = ra + delta / (T)2.0, hb = rb + delta / (T)2.0;
    const V3<T> ca = cross(n, wa), cb = cross(n, wb);
    const V3<T> vel = vl + ha * ca + hb * cb;
    const T vel_n = -dot(vl, n);
    const T nv = dot(n, vel);
    const V3<T> vel_t = vel - n * nv;
    const T vtl = vlen(vel_t);
    const T nd = dot(n, dt0u);
    const V3<T> dt0 = dt0u - n * nd;
    const T dt0l = vlen(dt0);
    const T sn = -k_n * delta - gamma_n * vel_n;
    V3<T> fn = sn * n;
    const bool clamped = dot(fn, n) < zero;
    if (clamped) fn = mk<T>(0, 0, 0);
    const T ck = -kappa * (T)sqrt(ra * rb);
    const V3<T> fc = ck * n;
    const bool branch = dt0l > zero || vtl > zero;
    V3<T> ft = mk<T>(0, 0, 0), dtp = mk<T>(0, 0, 0), fpre = mk<T>(0, 0, 0), that = mk<T>(0, 0, 0);
    V3<T> wv = mk<T>(0, 0, 0);
    T fpl = zero, wl = zero, lim = zero;
    bool slip = false;
    V3<T> dtout = dt0;
    if (branch) {
        dtp = dt0 + vel_t * dt;
        fpre = -k_t * dtp - gamma_t * vel_t;
        fpl = vlen(fpre);
        wv = fn - fc;
        wl = vlen(wv);
        lim = mu_s * wl;
        slip = fpl > lim;
        if (slip) {
            that = fpre / fpl;
            ft = lim * that;
            dtout = (T)-1.0 / k_t * (mu_d * wl * that + gamma_t * vel_t);
        } else {
            ft = fpre;
            dtout = dtp;
        }
    }
    // ---- backward
    V3<T> g_fn = gF, g_ft = gF, g_fc = gF;
    V3<T> g_n = mk<T>(0, 0, 0), g_vt = mk<T>(0, 0, 0), g_dt0 = mk<T>(0, 0, 0);
    T g_ha = zero, g_hb = zero, g_delta = zero, g_veln = zero;
    // torque T = cross(u, ft), u = -(ra + delta * 0.5) n
    const T hu = ra + delta * (T)0.5;
    const V3<T> u = -hu * n;
    g_ft = g_ft + cross(gT, u);
    const V3<T> g_u = cross(ft, gT);
    g_delta += -(T)0.5 * dot(g_u, n);  // hu = ra + delta/2
    g_n = g_n - hu * g_u;
    if (branch) {
        V3<T> g_dtp = mk<T>(0, 0, 0), g_fpre;
        T g_wl = zero;
        if (slip) {
            const V3<T> q = mu_d * wl * that + gamma_t * vel_t;
            const V3<T> g_q = ((T)-1.0 / k_t) * gD;
            g.gkt += dot(gD, q) / (k_t * k_t);
            g.gmud += wl * dot(that, g_q);
            g_wl += mu_d * dot(that, g_q);
            V3<T> g_that = (mu_d * wl) * g_q;
            g_vt = g_vt + gamma_t * g_q;
            const T g_lim = dot(that, g_ft);
            g_that = g_that + lim * g_ft;
            g.gmus += wl * g_lim;
            g_wl += mu_s * g_lim;
            g_fpre = (g_that - that * dot(that, g_that)) / fpl;
        } else {
            g_dtp = gD;
            g_fpre = g_ft;
        }
        g_dtp = g_dtp - k_t * g_fpre;
        g.gkt += -dot(dtp, g_fpre);
        g_vt = g_vt - gamma_t * g_fpre;
        g_dt0 = g_dt0 + g_dtp;
        g_vt = g_vt + dt * g_dtp;
        if (wl > zero) {
            const V3<T> g_w = (g_wl / wl) * wv;
            g_fn = g_fn + g_w;
            g_fc = g_fc - g_w;
        }
    } else {
        g_dt0 = g_dt0 + gD;
    }
    g_n = g_n + ck * g_fc;
    if (!clamped) {
        const T g_sn = dot(n, g_fn);
        g_n = g_n + sn * g_fn;
        g.gkn += -delta * g_sn;
        g_delta += -k_n * g_sn;
        g.ggn += -vel_n * g_sn;
        g_veln += -gamma_n * g_sn;
    }
    // dt0 = dt0u - n (n . dt0u)
    const T ngd = dot(n, g_dt0);
    g.gdt0u = g_dt0 - n * ngd;
    g_n = g_n - (nd * g_dt0 + ngd * dt0u);
    // vel_t = vel - n (n . vel)
    const T ngv = dot(n, g_vt);
    const V3<T> g_vel = g_vt - n * ngv;
    g_n = g_n - (nv * g_vt + ngv * vel);
    // vel_n = -(vl . n)
    V3<T> g_vl = -g_veln * n;
    g_n = g_n - g_veln * vl;
    // vel = vl + ha (n x wa) + hb (n x wb)
    g_vl = g_vl + g_vel;
    g_ha += dot(g_vel, ca);
    g_hb += dot(g_vel, cb);
    const V3<T> g_ca = ha * g_vel, g_cb = hb * g_vel;
    g_n = g_n + cross(wa, g_ca) + cross(wb, g_cb);
    g.gwa = g.gwa + cross(g_ca, n);
    g.gwb = g.gwb + cross(g_cb, n);
    g_delta += g_ha / (T)2.0 + g_hb / (T)2.0;
    g.gva = g.gva + g_vl;
    g.gvb = g.gvb - g_vl;
    // n = x / len, delta = len - (ra + rb)
    g.gx = g.gx + (g_n - n * dot(n, g_n)) / len + g_delta * n;
}

// Vector-Jacobian product of contact_wall (src/contactmodels.cuh:10-109):
// wall normal nw, overlap delta, wall velocity wvel added to every velocity
// component; cotangents gF, gT of the particle's force and torque and gfz of
// the returned normal force (the wall's force sum).
struct WallGrad {
    V3<double> gv, gw;
    double gdelta, gwvel, gkn;
};

__device__ __forceinline__ void contact_wall_vjp(const DemDev& d, V3<double> linvel, V3<double> angvel,
                                                 double ra, V3<double> n, double delta, double wvel,
                                                 V3<double> gF, V3<double> gT, double gfz, WallGrad& g) {
    // forward recompute
    const V3<double> c = cross(n, angvel);
    const V3<double> v0 = linvel + ra * c;
    const V3<double> vel = mk(v0.x + wvel, v0.y + wvel, v0.z + wvel);
    const double vel_n = dot(vel, n);
    const V3<double> vel_t = vel - n * dot(vel, n);
    const double vtl = vlen(vel_t);
    const double sn = -d.sys.k_n * delta - d.sys.gamma_wn * vel_n;
    V3<double> fn = sn * n;
    const bool clamped = dot(fn, n) < 0.0;
    if (clamped) fn = mk(0.0, 0.0, 0.0);
    const double fnl = vlen(fn);
    V3<double> ft = mk(0.0, 0.0, 0.0);
    bool visc_branch = false;
    double lim = 0.0, mu = 0.0;
    if (vtl > 0.0) {
        const double visc = d.sys.gamma_wt * vtl;
        mu = vtl > (double)0.001f ? d.sys.mu_wd : d.sys.mu_ws;
        lim = mu * fnl;
        visc_branch = visc < lim;
        ft = (-1.0 * (visc_branch ? visc : lim)) * vel_t / vtl;
    }
    // backward
    V3<double> g_fn = gF + gfz * n;
    V3<double> g_ft = gF + (-ra) * cross(gT, n);
    V3<double> g_vt = mk(0.0, 0.0, 0.0);
    double g_fnl = 0.0;
    if (vtl > 0.0) {
        if (visc_branch) {
            g_vt = g_vt - d.sys.gamma_wt * g_ft;  // ft = -gamma_wt vel_t
        } else {
            const V3<double> e = vel_t / vtl;
            const double g_lim = -dot(e, g_ft);
            const V3<double> g_e = -lim * g_ft;
            g_fnl += mu * g_lim;
            g_vt = g_vt + (g_e - e * dot(e, g_e)) / vtl;
        }
    }
    double g_veln = 0.0;
    g.gdelta = 0.0;
    g.gkn = 0.0;
    if (!clamped) {
        if (fnl > 0.0) g_fn = g_fn + (g_fnl / fnl) * fn;
        const double g_sn = dot(n, g_fn);
        g.gkn = -delta * g_sn;
        g.gdelta = -d.sys.k_n * g_sn;
        g_veln = -d.sys.gamma_wn * g_sn;
    }
    const V3<double> g_vel = (g_vt - n * dot(n, g_vt)) + g_veln * n;
    g.gwvel = g_vel.x + g_vel.y + g_vel.z;
    g.gv = g_vel;
    const V3<double> g_c = ra * g_vel;
    g.gw = cross(g_c, n);
}

// Wall leapfrog backward (src/integration.cuh:224-295), one thread per wall:
// next -> adjoint of the wall force sum, and the wall's own part of cur.
__global__ void k_adj_walls(DemDev d, AdjDev a, double* lfw) {
    const int w = threadIdx.x;
    if (w >= NWALL) return;
    const double* nx = a.next.wall + 3 * w;  // pos, vel, vel0
    double* cu = a.cur.wall + 3 * w;
    const double dt = d.sys.dt;
    const int mode = w < (int)d.nw ? a.wmode[w] : 0;
    if (mode == 0) {  // the wall does not move
        cu[0] = nx[0];
        cu[1] = nx[1];
        cu[2] = 0.0;
        lfw[w] = 0.0;
        return;
    }
    const double lv0t = nx[2] + nx[0] * dt + nx[1];
    const double lacc = lv0t * dt + nx[1] * (0.5 * dt);
    lfw[w] = mode == 1 ? lacc / a.wmass[w] : 0.0;
    cu[0] = nx[0];
    cu[1] = 0.0;
    cu[2] = lv0t;
}

__global__ __launch_bounds__(256) void k_fill_f64(double* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Block-wide sum of v into *dst (one atomic per block).
__device__ __forceinline__ void block_add(double v, double* dst, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += sh[k];
        if (s != 0.0) unsafeAtomicAdd(dst, s);
    }
}

// Adjoint of the leapfrog and of the objective for one taped step, one thread
// per particle: the cotangents of its half-step velocities (to cur) and of
// its force and torque (gfa, gfb, by sorted position, for k_adj_step).
__global__ __launch_bounds__(256) void k_adj_lf(DemDev d, AdjDev a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= d.np) return;
    const TapeView& t = a.t;
    const uint32_t o = t.id[s];
    const double ra = t.x[s].w;
    const bool fix = t.vel[s].w > 0.0;
    const bool hold_z = t.vel[s].w >= SPH_FIXVEL_HOLD_Z;  // extension: no z acceleration either
    const double dt = d.sys.dt;
    const double m = 4.0 / 3.0 * PI_REF * ra * ra * ra * d.sys.rho;
    const double inertia = 2.0 / 5.0 * m * ra * ra;
    const V3<double> Lx = xyz(a.next.x[o]), Lv0 = xyz(a.next.v0[o]), Lv = xyz(a.next.v[o]);
    const V3<double> Lw0 = xyz(a.next.w0[o]), Lw = xyz(a.next.w[o]);
    const V3<double> Lv0t = Lv0 + Lx * dt + Lv;
    const V3<double> La = Lv0t * dt + Lv * (0.5 * dt);
    const V3<double> Lw0t = Lw0 + Lw;
    const V3<double> Lal = Lw0t * dt + Lw * (0.5 * dt);
    V3<double> LF = fix ? mk(0.0, 0.0, hold_z ? 0.0 : La.z / m) : La / m;
    const V3<double> LT = fix ? mk(0.0, 0.0, 0.0) : Lal / inertia;
    const double4 wo = a.wobj[o];
    LF = LF + mk(wo.x, wo.y, wo.z) * a.inv_n;
    a.cur.v0[o] = make_double4(Lv0t.x, Lv0t.y, Lv0t.z, 0.0);
    a.cur.w0[o] = make_double4(Lw0t.x, Lw0t.y, Lw0t.z, 0.0);
    a.gfa[s] = make_double4(LF.x, LF.y, LF.z, LT.x);
    a.gfb[s] = make_double4(LT.y, LT.z, 0.0, 0.0);
}

// One adjoint step, one thread per particle of the taped sorted order.  A
// particle gathers every cotangent of its own state: those of its own pair
// forces and wall forces, and those of each partner's pair force with it,
// recomputed from the partner's side (the partner's slot holding this
// particle, its spring and its force cotangent) -- so every write of cur is
// a plain store and no atomics are needed.
template <class T>
__global__ __launch_bounds__(256) void k_adj_step(DemDev d, AdjDev a) {
    __shared__ double red[4];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = s < d.np;
    double th[4] = {0.0, 0.0, 0.0, 0.0};     // k_n, k_t, gamma_n, mu
    double gwp[NWALL] = {0, 0, 0, 0, 0, 0};   // wall position
    double gwv[NWALL] = {0, 0, 0, 0, 0, 0};   // wall velocity
    if (act) {
        const TapeView& t = a.t;
        const uint32_t o = t.id[s];
        const double4 xa4 = t.x[s], va4 = t.vel[s], wa4 = t.angvel[s];
        const double ra = xa4.w;
        const double4 fa = a.gfa[s], fb = a.gfb[s];
        const V3<double> LF = mk(fa.x, fa.y, fa.z), LT = mk(fa.w, fb.x, fb.y);
        V3<double> gx = xyz(a.next.x[o]), gv = mk(0.0, 0.0, 0.0), gw = mk(0.0, 0.0, 0.0);
        // 3. contacts in slot order
        const uint32_t mask = t.tmask[s], fresh = t.tfresh[s];
        const V3<double> va = xyz(va4), wa = xyz(wa4);
        // (only the slots occupied after the search: the adjoint of the step
        // before reads a spring adjoint only for a slot that was occupied at
        // the end of that step, i.e. at this step's start)
        for (uint32_t mm = mask; mm; mm &= mm - 1) {
            const int i = __ffs(mm) - 1;
            const size_t mp = (size_t)o * NC + i;
            double* cdt_out = a.cur.dt + 3 * mp;
            const uint32_t w = t.cslot[mp];
            const uint32_t bo = w & SLOT_IDX_MASK;
            const uint32_t sb = t.pos_of[bo];
            const double4 xb = t.x[sb];
            const uint32_t code = w >> 28, cx = code & 3u, cy = (code >> 2) & 3u;
            const double dmx = cx == 1u ? d.sys.L[0] : (cx == 2u ? -d.sys.L[0] : 0.0);
            const double dmy = cy == 1u ? d.sys.L[1] : (cy == 2u ? -d.sys.L[1] : 0.0);
            const V3<double> xab = mk((xa4.x - xb.x) + dmx, (xa4.y - xb.y) + dmy, (xa4.z - xb.z) + 0.0);
            const double len = vlen(xab);
            const double delta_n = len - (ra + xb.w);
            if (!(delta_n < 0.0)) {  // the force kernel ended this contact
                cdt_out[0] = cdt_out[1] = cdt_out[2] = 0.0;
                continue;
            }
            V3<T> dt0u = mk<T>(0, 0, 0);
            if (!((fresh >> i) & 1u)) {
                if (t.cdt32) dt0u = mk<T>((T)t.cdt32[3 * mp], (T)t.cdt32[3 * mp + 1], (T)t.cdt32[3 * mp + 2]);
                else dt0u = mk<T>((T)t.cdt[3 * mp], (T)t.cdt[3 * mp + 1], (T)t.cdt[3 * mp + 2]);
            }
            const double* gdp = a.next.dt + 3 * mp;
            PairGrad<T> g;
            g.gx = g.gva = g.gvb = g.gwa = g.gwb = g.gdt0u = mk<T>(0, 0, 0);
            g.gkn = g.gkt = g.ggn = g.gmus = g.gmud = (T)0;
            const V3<double> vb = xyz(t.vel[sb]), wb = xyz(t.angvel[sb]);
            contact_linear_vjp<T>(d, cvt<T>(va), cvt<T>(wa), cvt<T>(vb), cvt<T>(wb), (T)ra, (T)xb.w, cvt<T>(xab),
                                  (T)len, (T)delta_n, dt0u, cvt<T>(LF), cvt<T>(LT),
                                  mk<T>((T)gdp[0], (T)gdp[1], (T)gdp[2]), g);
            gx = gx + up(g.gx);
            gv = gv + up(g.gva);
            gw = gw + up(g.gwa);
            // the partner's pair force with this particle, from its side
            const uint32_t mb = t.tmask[sb];
            const uint32_t* bslot = t.cslot + (size_t)bo * NC;
            int jb = -1;
            for (uint32_t q = mb; q; q &= q - 1) {
                const int k = __ffs(q) - 1;
                if ((bslot[k] & SLOT_IDX_MASK) == o) { jb = k; break; }
            }
            if (jb >= 0) {
                const size_t mq = (size_t)bo * NC + jb;
                const uint32_t cb = bslot[jb] >> 28, bx = cb & 3u, by = (cb >> 2) & 3u;
                const double emx = bx == 1u ? d.sys.L[0] : (bx == 2u ? -d.sys.L[0] : 0.0);
                const double emy = by == 1u ? d.sys.L[1] : (by == 2u ? -d.sys.L[1] : 0.0);
                const V3<double> xba = mk((xb.x - xa4.x) + emx, (xb.y - xa4.y) + emy, (xb.z - xa4.z) + 0.0);
                const double lenb = vlen(xba);
                const double delb = lenb - (xb.w + ra);
                V3<T> dtb = mk<T>(0, 0, 0);
                if (!((t.tfresh[sb] >> jb) & 1u)) {
                    if (t.cdt32) dtb = mk<T>((T)t.cdt32[3 * mq], (T)t.cdt32[3 * mq + 1], (T)t.cdt32[3 * mq + 2]);
                    else dtb = mk<T>((T)t.cdt[3 * mq], (T)t.cdt[3 * mq + 1], (T)t.cdt[3 * mq + 2]);
                }
                const double4 pa = a.gfa[sb], pb = a.gfb[sb];
                const double* gdq = a.next.dt + 3 * mq;
                PairGrad<T> h;
                h.gx = h.gva = h.gvb = h.gwa = h.gwb = h.gdt0u = mk<T>(0, 0, 0);
                h.gkn = h.gkt = h.ggn = h.gmus = h.gmud = (T)0;
                contact_linear_vjp<T>(d, cvt<T>(vb), cvt<T>(wb), cvt<T>(va), cvt<T>(wa), (T)xb.w, (T)ra, cvt<T>(xba),
                                      (T)lenb, (T)delb, dtb, mk<T>((T)pa.x, (T)pa.y, (T)pa.z),
                                      mk<T>((T)pa.w, (T)pb.x, (T)pb.y), mk<T>((T)gdq[0], (T)gdq[1], (T)gdq[2]), h);
                gx = gx - up(h.gx);
                gv = gv + up(h.gvb);
                gw = gw + up(h.gwb);
            }
            const bool f = (fresh >> i) & 1u;
            cdt_out[0] = f ? 0.0 : (double)g.gdt0u.x;
            cdt_out[1] = f ? 0.0 : (double)g.gdt0u.y;
            cdt_out[2] = f ? 0.0 : (double)g.gdt0u.z;
            th[0] += (double)g.gkn;
            th[1] += (double)g.gkt;
            th[2] += (double)g.ggn;
            th[3] += (double)g.gmus + (double)g.gmud;
        }
        // 4. walls, as the force kernel evaluates them (src/contactsearch.cuh:587-663)
        WallGrad g;
        {
            const double4 w0 = t.wnx[0];
            const double dw = w0.w - (xa4.z + ra);
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(w0.x, w0.y, w0.z), dw, t.wmvfd[0].y, LF, LT,
                                 d.nw > 0 ? a.lfw[0] : 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.z -= g.gdelta; gwp[0] += g.gdelta; gwv[0] += g.gwvel;
            }
        }
        {
            const double dw = xa4.z - ra - d.sys.origo[2];
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(0.0, 0.0, 1.0), dw, 0.0, LF, LT, 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.z += g.gdelta;
            }
        }
        if (d.sys.periodic == 0) {
            const double4 w1 = t.wnx[1], w2 = t.wnx[2];
            double dw = w1.w - (xa4.x + ra);
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(w1.x, w1.y, w1.z), dw, t.wmvfd[1].y, LF, LT,
                                 d.nw > 1 ? a.lfw[1] : 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.x -= g.gdelta; gwp[1] += g.gdelta; gwv[1] += g.gwvel;
            }
            dw = xa4.x - ra - w2.w;
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(w2.x, w2.y, w2.z), dw, t.wmvfd[2].y, LF, LT,
                                 d.nw > 2 ? a.lfw[2] : 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.x += g.gdelta; gwp[2] -= g.gdelta; gwv[2] += g.gwvel;
            }
        }
        if (d.sys.periodic == 0 || d.sys.periodic == 2) {
            const double4 w3 = t.wnx[3], w4 = t.wnx[4];
            double dw = w3.w - (xa4.y + ra);
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(w3.x, w3.y, w3.z), dw, t.wmvfd[3].y, LF, LT,
                                 d.nw > 3 ? a.lfw[3] : 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.y -= g.gdelta; gwp[3] += g.gdelta; gwv[3] += g.gwvel;
            }
            dw = xa4.y - ra - w4.w;
            if (dw < 0.0) {
                contact_wall_vjp(d, va, wa, ra, mk(w4.x, w4.y, w4.z), dw, t.wmvfd[4].y, LF, LT,
                                 d.nw > 4 ? a.lfw[4] : 0.0, g);
                gv = gv + g.gv; gw = gw + g.gw; th[0] += g.gkn;
                gx.y += g.gdelta; gwp[4] -= g.gdelta; gwv[4] += g.gwvel;
            }
        }
        a.cur.x[o] = make_double4(gx.x, gx.y, gx.z, 0.0);
        a.cur.v[o] = make_double4(gv.x, gv.y, gv.z, 0.0);
        a.cur.w[o] = make_double4(gw.x, gw.y, gw.z, 0.0);
    }
    for (int k = 0; k < 4; ++k) block_add(th[k], a.theta + k, red);
    // (walls past nw stay where the reference's defaults put them: their
    // adjoints feed nothing)
    for (int w = 0; w < 5 && w < (int)d.nw; ++w) {
        block_add(gwp[w], a.cur.wall + 3 * w + 0, red);
        block_add(gwv[w], a.cur.wall + 3 * w + 1, red);
    }
}

// J += sum over particles of w_o . F (sorted order of the step just done).
__global__ __launch_bounds__(256) void k_adj_objective(DemDev d, const double4* wobj, double* J) {
    __shared__ double red[4];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (s < d.np) {
        const double4 w = wobj[d.sid[s]], f = d.force[s];
        v = w.x * f.x + w.y * f.y + w.z * f.z;
    }
    block_add(v, J, red);
}

// One taped step: the sorted input state as the force kernel sees it
// (records at perm[q], positions wrapped as the step wraps them), the
// inverse permutation, slot masks, slot words, incoming springs and walls.
struct TapeOut {
    double4 *x, *vel, *angvel, *wnx, *wmvfd;
    uint32_t *id, *pos_of, *tmask, *tfresh, *cslot;
    double* cdt;
    float* cdt32;
};

__global__ __launch_bounds__(256) void k_adj_tape(DemDev d, TapeOut t) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < d.np) {
        const uint32_t p = d.perm[q];
        double4 x = d.ro.x[p];
        if (d.sys.periodic == 1 || d.sys.periodic == 2) {
            if (x.x < d.sys.origo[0]) x.x += d.sys.L[0];
            if (x.x > d.sys.L[0]) x.x -= d.sys.L[0];
        }
        if (d.sys.periodic == 1) {
            if (x.y < d.sys.origo[1]) x.y += d.sys.L[1];
            if (x.y > d.sys.L[1]) x.y -= d.sys.L[1];
        }
        t.x[q] = x;
        t.vel[q] = d.ro.vel[p];
        t.angvel[q] = d.ro.angvel[p];
        t.id[q] = d.sid[q];
        t.pos_of[d.sid[q]] = (uint32_t)q;  // this step's sorted position of each particle
        const uint32_t mask = d.tmask[q];
        t.tmask[q] = mask;
        t.tfresh[q] = d.tfresh[q];
        // the occupied slots only: the adjoint step reads no other (its own
        // and its partners' occupied slots of this step)
        // (the input set keeps slot k's word at k * np + p and its spring at
        // 3 * (k * np + p), p the storage position; the tape keys them by
        // original index and slot)
        const size_t o = (size_t)d.sid[q] * NC;
        for (uint32_t m = mask; m; m &= m - 1) {
            const int k = __ffs(m) - 1;
            const size_t i = o + (size_t)k, src = 3 * ((size_t)k * d.np + p);
            t.cslot[i] = d.ro.cslot[(size_t)k * d.np + p];
            if (t.cdt32) {
                t.cdt32[3 * i] = d.ro.cdt32[src];
                t.cdt32[3 * i + 1] = d.ro.cdt32[src + 1];
                t.cdt32[3 * i + 2] = d.ro.cdt32[src + 2];
            } else {
                t.cdt[3 * i] = d.ro.cdt[src];
                t.cdt[3 * i + 1] = d.ro.cdt[src + 1];
                t.cdt[3 * i + 2] = d.ro.cdt[src + 2];
            }
        }
    }
    if (q < 6) {
        t.wnx[q] = d.wnx[q];
        t.wmvfd[q] = d.wmvfd[q];
    }
}

inline uint32_t nb(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

struct AdjointState {
    std::vector<void*> allocs;
    uint64_t nsteps = 0;
    uint32_t seg = 0;
    int precision = SPH_PREC_FP64;
    std::vector<Checkpoint> cps;  // one per segment
    Checkpoint end;
    // tape, seg steps
    double4 *tx = nullptr, *tv = nullptr, *tw = nullptr, *twn = nullptr, *twm = nullptr;
    uint32_t *tid = nullptr, *tpos = nullptr, *tmask = nullptr, *tfresh = nullptr, *tslot = nullptr;
    double* tcdt = nullptr;
    float* tcdt32 = nullptr;
    Lam lam[2] = {};
    double4 *gfa = nullptr, *gfb = nullptr;  // force and torque cotangents of the step being processed
    double4* wobj = nullptr;
    double* scal = nullptr;  // J, theta[4], lfw[6]
    bool have_forward = false;
    ~AdjointState() {
        for (void* p : allocs) (void)hipFree(p);
    }
};

namespace {

void free_adj(AdjointState* a) { delete a; }

template <class T> int adj_alloc(sph_dem* h, AdjointState* s, T** p, size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1) != hipSuccess) {
        h->err = "adjoint: hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed";
        return SPH_E_HIP;
    }
    s->allocs.push_back(q);
    *p = (T*)q;
    return SPH_OK;
}

int read_flags(sph_dem* h) {
    uint32_t f = 0;
    if (hipMemcpyAsync(&f, h->d.flags, 4, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
        h->err = "adjoint: flag read failed";
        return SPH_E_HIP;
    }
    if (!f) return SPH_OK;
    if (f & FLAG_NAN) { h->err = "non-finite particle position"; return SPH_E_NAN; }
    if (f & FLAG_OUT_OF_GRID) { h->err = "a particle left the grid"; return SPH_E_OUT_OF_GRID; }
    h->err = "a particle has more than " + std::to_string(NC) + " contacts";
    return SPH_E_CONTACT_OVERFLOW;
}

// one sequential step (wall update at its end), with hooks
void run_one(sph_dem* h, const StepHooks* hooks) {
    enqueue_step(h, false, h->own_cur, false, true, hooks);
    h->own_cur ^= 1;
    h->hs.t_cur += h->hs.dt;
    ++h->steps_done;
}

}  // namespace
}  // namespace sph

using sph::AdjointState;

extern "C" {

int sph_adjoint_forward(sph_t* h, uint64_t nsteps, uint32_t seg_len, const double* weights, double* J_out) {
    if (!h || !weights || !J_out || nsteps == 0 || seg_len == 0) return SPH_E_INVALID;
    if (h->hs.nb0) {
        h->err = "sph_adjoint_forward: bonds are not differentiated";
        return SPH_E_UNSUPPORTED;
    }
    if (h->hs.contactmodel != 2) {
        h->err = "sph_adjoint_forward: only contact model 2 is differentiated";
        return SPH_E_UNSUPPORTED;
    }
    if (h->hs.mu_s != h->hs.mu_d) {
        // theta holds one friction coefficient mu = mu_s = mu_d
        h->err = "sph_adjoint_forward: the gradient is taken for mu = mu_s = mu_d; set them equal";
        return SPH_E_INVALID;
    }
    const size_t n = h->hs.np;
    h->adj = std::unique_ptr<AdjointState, void (*)(AdjointState*)>(new AdjointState(), sph::free_adj);
    AdjointState* a = h->adj.get();
    a->nsteps = nsteps;
    a->seg = seg_len;
    a->precision = h->precision;
    int rc;
    if ((rc = sph::adj_alloc(h, a, &a->wobj, n)) || (rc = sph::adj_alloc(h, a, &a->scal, 16))) return rc;
    std::vector<double> w4(4 * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) w4[4 * i + c] = weights[3 * i + c];
    if (hipMemcpy(a->wobj, w4.data(), n * 32, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(a->scal, 0, 16 * 8) != hipSuccess) {
        h->err = "sph_adjoint_forward: upload failed";
        return SPH_E_HIP;
    }
    // pick the contact-search kernel on a throw-away copy of the state
    if ((rc = sph::autotune(h))) return rc;
    const uint64_t nseg = (nsteps + seg_len - 1) / seg_len;
    a->cps.resize(nseg);
    sph::StepHooks hooks;
    double* J = a->scal;
    const double4* wobj = a->wobj;
    hooks.after_force = [wobj, J](const sph::DemDev& d, hipStream_t s) {
        sph::k_adj_objective<<<sph::nb(d.np), 256, 0, s>>>(d, wobj, J);
    };
    for (uint64_t k = 0; k < nseg; ++k) {
        if ((rc = sph::save_state(h, a->cps[k]))) return rc;
        const uint64_t m = std::min<uint64_t>(seg_len, nsteps - k * seg_len);
        for (uint64_t i = 0; i < m; ++i) sph::run_one(h, &hooks);
        if ((rc = sph::read_flags(h))) return rc;
    }
    if ((rc = sph::save_state(h, a->end))) return rc;
    double Jsum = 0.0;
    if (hipMemcpy(&Jsum, J, 8, hipMemcpyDeviceToHost) != hipSuccess) {
        h->err = "sph_adjoint_forward: read-back failed";
        return SPH_E_HIP;
    }
    *J_out = Jsum / (double)nsteps;
    a->have_forward = true;
    return SPH_OK;
}

int sph_adjoint_backward(sph_t* h, double grad_out[4]) {
    if (!h || !grad_out) return SPH_E_INVALID;
    AdjointState* a = h->adj.get();
    if (!a || !a->have_forward) {
        h->err = "sph_adjoint_backward: no sph_adjoint_forward on this handle";
        return SPH_E_STATE;
    }
    if (a->precision != h->precision) {
        h->err = "sph_adjoint_backward: precision changed since the forward run";
        return SPH_E_STATE;
    }
    const size_t n = h->hs.np, nc = sph::NC, S = a->seg;
    int rc;
    // tape and adjoint buffers on first use
    if (!a->tx) {
        if ((rc = sph::adj_alloc(h, a, &a->tx, S * n)) || (rc = sph::adj_alloc(h, a, &a->tv, S * n)) ||
            (rc = sph::adj_alloc(h, a, &a->tw, S * n)) || (rc = sph::adj_alloc(h, a, &a->twn, S * 6)) ||
            (rc = sph::adj_alloc(h, a, &a->twm, S * 6)) || (rc = sph::adj_alloc(h, a, &a->tid, S * n)) ||
            (rc = sph::adj_alloc(h, a, &a->tpos, S * n)) || (rc = sph::adj_alloc(h, a, &a->tmask, S * n)) ||
            (rc = sph::adj_alloc(h, a, &a->tfresh, S * n)) || (rc = sph::adj_alloc(h, a, &a->tslot, S * n * nc)) ||
            (rc = sph::adj_alloc(h, a, &a->gfa, n)) || (rc = sph::adj_alloc(h, a, &a->gfb, n)))
            return rc;
        if (h->precision == SPH_PREC_MIXED) rc = sph::adj_alloc(h, a, &a->tcdt32, S * n * nc * 3);
        else rc = sph::adj_alloc(h, a, &a->tcdt, S * n * nc * 3);
        if (rc) return rc;
        for (auto& l : a->lam) {
            if ((rc = sph::adj_alloc(h, a, &l.x, n)) || (rc = sph::adj_alloc(h, a, &l.v0, n)) ||
                (rc = sph::adj_alloc(h, a, &l.v, n)) || (rc = sph::adj_alloc(h, a, &l.w0, n)) ||
                (rc = sph::adj_alloc(h, a, &l.w, n)) || (rc = sph::adj_alloc(h, a, &l.dt, n * nc * 3)) ||
                (rc = sph::adj_alloc(h, a, &l.wall, 3 * 6)))
                return rc;
        }
    }
    hipStream_t st = h->stream;
    auto zero = [&](double* p, size_t count) { sph::k_fill_f64<<<sph::nb(count), 256, 0, st>>>(p, count); };
    auto zero_lam = [&](sph::Lam& l) {
        zero((double*)l.x, 4 * n); zero((double*)l.v0, 4 * n); zero((double*)l.v, 4 * n);
        zero((double*)l.w0, 4 * n); zero((double*)l.w, 4 * n); zero(l.dt, n * nc * 3); zero(l.wall, 18);
    };
    zero_lam(a->lam[0]);
    zero_lam(a->lam[1]);    // (each adjoint step clears the x, v, w of lambda it consumed)
    zero(a->scal + 1, 10);  // theta, lfw
    int nxt = 0;            // lam[nxt] holds lambda of the step after the one processed
    std::vector<int> wmode(6, 0);
    for (uint32_t w = 0; w < h->hs.nw && w < 6; ++w) wmode[w] = h->hs.wmode[w];
    int* d_wmode = nullptr;
    if ((rc = sph::adj_alloc(h, a, &d_wmode, 6))) return rc;
    if (hipMemcpyAsync(d_wmode, wmode.data(), 24, hipMemcpyHostToDevice, st) != hipSuccess) return SPH_E_HIP;
    const uint64_t nseg = a->cps.size();
    for (uint64_t k = nseg; k-- > 0;) {
        const uint64_t m = std::min<uint64_t>(S, a->nsteps - k * S);
        if ((rc = sph::restore_state(h, a->cps[k]))) return rc;
        // recompute the segment, taping each step before its force kernel
        for (uint64_t j = 0; j < m; ++j) {
            sph::StepHooks hooks;
            hooks.before_force = [a, j, n](const sph::DemDev& d, hipStream_t s) {
                const size_t o1 = j * n, on = j * n * sph::NC;
                sph::TapeOut t;
                t.x = a->tx + o1; t.vel = a->tv + o1; t.angvel = a->tw + o1;
                t.wnx = a->twn + 6 * j; t.wmvfd = a->twm + 6 * j;
                t.id = a->tid + o1; t.pos_of = a->tpos + o1; t.tmask = a->tmask + o1; t.tfresh = a->tfresh + o1;
                t.cslot = a->tslot + on;
                t.cdt = a->tcdt ? a->tcdt + 3 * on : nullptr;
                t.cdt32 = a->tcdt32 ? a->tcdt32 + 3 * on : nullptr;
                sph::k_adj_tape<<<sph::nb(n), 256, 0, s>>>(d, t);
            };
            sph::run_one(h, &hooks);
        }
        if ((rc = sph::read_flags(h))) return rc;
        // adjoint steps in reverse
        for (uint64_t j = m; j-- > 0;) {
            sph::AdjDev ad{};
            const size_t o1 = j * n, on = j * n * nc;
            ad.t.x = a->tx + o1; ad.t.vel = a->tv + o1; ad.t.angvel = a->tw + o1;
            ad.t.id = a->tid + o1; ad.t.pos_of = a->tpos + o1; ad.t.tmask = a->tmask + o1;
            ad.t.tfresh = a->tfresh + o1; ad.t.cslot = a->tslot + on;
            ad.t.cdt = a->tcdt ? a->tcdt + 3 * on : nullptr;
            ad.t.cdt32 = a->tcdt32 ? a->tcdt32 + 3 * on : nullptr;
            ad.t.wnx = a->twn + 6 * j; ad.t.wmvfd = a->twm + 6 * j;
            ad.next = a->lam[nxt];
            ad.cur = a->lam[nxt ^ 1];
            ad.wobj = a->wobj;
            ad.inv_n = 1.0 / (double)a->nsteps;
            ad.lfw = a->scal + 5;
            ad.theta = a->scal + 1;
            ad.wmode = d_wmode;
            for (uint32_t w = 0; w < 6; ++w) ad.wmass[w] = w < h->hs.nw ? h->hs.wmvfd[4 * w] : 1.0;
            ad.gfa = a->gfa;
            ad.gfb = a->gfb;
            sph::k_adj_walls<<<1, 64, 0, st>>>(h->d, ad, a->scal + 5);
            sph::k_adj_lf<<<sph::nb(n), 256, 0, st>>>(h->d, ad);
            if (h->precision == SPH_PREC_MIXED) sph::k_adj_step<float><<<sph::nb(n), 256, 0, st>>>(h->d, ad);
            else sph::k_adj_step<double><<<sph::nb(n), 256, 0, st>>>(h->d, ad);
            nxt ^= 1;
        }
        if (hipGetLastError() != hipSuccess) {
            h->err = "sph_adjoint_backward: kernel launch failed";
            return SPH_E_HIP;
        }
    }
    double th[4];
    if (hipMemcpyAsync(th, a->scal + 1, 32, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        h->err = "sph_adjoint_backward: read-back failed";
        return SPH_E_HIP;
    }
    for (int k = 0; k < 4; ++k) grad_out[k] = th[k];
    // leave the handle where the forward run ended
    if ((rc = sph::restore_state(h, a->end))) return rc;
    if (hipStreamSynchronize(st) != hipSuccess) return SPH_E_HIP;
    return SPH_OK;
}

}  // extern "C"
