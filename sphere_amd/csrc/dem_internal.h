// dem_internal.h -- device layout and kernel entry points of the gfx950 DEM step.
//
// Per-particle state is kept in HBM in the cell-sorted order of the last
// step, not in the .bin file's order: every per-particle kernel then reads
// and writes consecutive records, whatever order the input numbered the
// particles in.  Each particle carries its original index (`id`), which is
// what the reference's semantics are stated in (the stable sort's tie-break,
// contact partners, bond endpoints, the wall-force reduction order and the
// output file order).  Two buffer sets alternate between steps: a step reads
// every record from one set (`ro`) at the particle's previous sorted position
// perm[s] and writes the updated record to the other (`wo`) at its new
// sorted position s, so the next step finds its input in consecutive order
// and neighbour reads never race with updates.  The sort itself only moves
// indices, an FP32 copy of the positions for the contact search's prefilter
// and the sorted ids (`sid`).  Records are 32-byte double4 so each access
// fills whole sectors.
//
// Contact history moves with the particle: a 32-bit occupancy mask and, per
// slot (SPH_NC of them), a slot word and a tangential spring (3 doubles in
// FP64, 3 floats in mixed mode), slot-major in each buffer set.  A slot word
// packs the partner's original index (low 28 bits) and the periodic image
// code (high 4 bits: 2 bits per periodic axis, 0 = none, 1 = +L, 2 = -L).
// Only occupied slots are read or written.
#ifndef SPHERE_DEM_INTERNAL_H_
#define SPHERE_DEM_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sphere_dem.h"

namespace sph {

constexpr int NC = SPH_NC;
constexpr uint32_t EMPTY_CELL = 0xffffffffu;
constexpr uint32_t SLOT_IDX_MASK = 0x0fffffffu;
constexpr uint32_t MAX_PARTICLES = SLOT_IDX_MASK;  // indices must fit 28 bits
constexpr int BLOCK = 256;
constexpr int SCAN_PER_THREAD = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_PER_THREAD;
constexpr int TOPO_CAP = 1024;          // staged neighbour particles per search block
constexpr int TOPO_MAX_CELLS = 384;     // staged cells per search block
constexpr int NLIST = 48;               // neighbour-list entries per particle (search mode 5)

// device error flags (atomicOr'd by kernels, read after each sph_step chunk)
enum : uint32_t {
    FLAG_OUT_OF_GRID = 1u,
    FLAG_CONTACT_OVERFLOW = 2u,
    FLAG_NAN = 4u,
};

// Unsigned division by an invariant divisor d >= 1 as a multiply-high, a
// subtract, an add and two shifts (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, figure 4.1): with
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1 and t = umulhi(m, n),
// n / d = (t + ((n - t) >> s1)) >> s2 for every 32-bit n, s1 = min(l, 1),
// s2 = max(l - 1, 0).  The contact search takes cell coordinates apart per
// candidate pair, where a 32-bit integer division (no single instruction on
// the GPU) costs tens of instructions.
struct FastDiv {
    uint32_t d, m, s1, s2;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (uint64_t(1) << l) < d) ++l;
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
    f.s1 = l < 1 ? l : 1;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
}
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

// One buffer set of per-particle state in storage order.
struct PState {
    double4 *x;       // x, y, z, r
    double4 *vel;     // vx, vy, vz, fixvel
    double4 *angvel;  // wx, wy, wz, ev
    // records only the particle reads, 11 doubles packed without padding:
    double4 *pr0;     // vel0.x, vel0.y, vel0.z, angvel0.x (half-step velocities)
    double4 *pr1;     // angvel0.y, angvel0.z, angpos.x, angpos.y
    double4 *pr2;     // angpos.z, xysum.x, xysum.y, es
    uint32_t *id;     // original index
    uint32_t *cmask;  // occupied contact slots
    uint32_t *lidx;   // search mode 5: the particle's neighbour list (its sorted position at the build)
    float4 *vf, *wf;  // mixed mode: FP32 copies of vel and angvel (xyz) for the pair force's partner gathers
    // contact slots, slot-major: the slot word of slot k of the particle
    // stored at q is cslot[k * np + q], its tangential spring (x, y, z) at
    // 3 * (k * np + q).  A step's searches update the particle's words in its
    // input set; the force kernel reads words and springs there and writes
    // those of the contacts that last at the particle's sorted position in
    // the output set, so the lanes of a wavefront, working on consecutive
    // particles, touch consecutive records.
    uint32_t* cslot;
    double* cdt;   // FP64 mode
    float* cdt32;  // mixed mode
};

// The records of one buffer set, as a step reads (ro) or writes (wo) them.
struct OwnSet {
    double4 *x, *vel, *angvel, *pr0, *pr1, *pr2;
    uint32_t *id, *cmask, *lidx;
    float4 *vf, *wf;
    uint32_t* cslot;
    double* cdt;
    float* cdt32;
};

// Box, cell grid, parameters and walls of one system.  A handle steps either
// one system (DemDev::sys) or an ensemble of independent systems stored back
// to back (DemDev::systems, system of index i = sys_of[i]): system k owns the
// original indices and the sorted positions [pbase, pbase + np) -- the cell
// ids of different systems never interleave, so the sort keeps every system
// in its own range -- and the cell ids [cbase, cbase + K).
struct Sys {
    double origo[3], L[3], cell_len[3];
    uint32_t num[3];
    FastDiv div_x, div_xy;  // division by num[0] and by num[0] * num[1] (cell coordinates of a cell id)
    int periodic;
    float Lf[2];        // L[0], L[1] in FP32 (periodic images in the prefilter)
    uint32_t cbase, K;  // first cell id, cells
    uint32_t pbase, np; // first index, particles
    uint32_t nw, wbase; // walls; wall w of the system is entry wbase + w of the wall arrays
    uint32_t bbase, nblocks;  // wall-sum blocks (256 original indices each) of the system
    double dt;
    double g[3];
    double k_n, k_t, gamma_n, gamma_t, mu_s, mu_d;
    double gamma_wn, gamma_wt, mu_ws, mu_wd, rho, kappa;
    double gamma_r, mu_r;   // rolling resistance (contact model 3)
    double devs_A, devs_f;
    uint32_t contactmodel;  // 1: visco-frictional, found and processed at once; 2: linear with slots;
                            // 3: Hertz-Mindlin with rolling resistance, with slots
    double db, V_b;         // capillary bond reach and liquid volume (contact model 1)
};

// Everything a step kernel needs, passed by value as one kernel argument.
struct DemDev {
    uint32_t np, nb0, nw, K;  // totals over the systems (nw: the most walls of any system)
    uint32_t bond_broken;  // sentinel written into bonds[k].x on breakage
    uint32_t nwall_blocks; // wall-sum blocks over all systems, partials per wall
    double lambda_bar, sigma_b, tau_b;
    Sys sys;               // the system of a one-system handle
    const Sys* systems;    // ensemble: one entry per system (nullptr for one system)
    const uint16_t* sys_of;    // ensemble: system of each index / sorted position
    const uint16_t* sys_of_block;  // ensemble: system of each wall-sum block
    uint32_t nsys;

    PState a, b;        // storage of the two buffer sets
    OwnSet ro, wo;      // this step's input (at perm[s]) and output (at s)
    uint32_t* sid;      // original index per sorted position of this step
    uint32_t* pos_of;   // original index -> storage position (the previous step's sorted position)
                        // until k_pos_of after the force kernel sets this step's

    // per-step outputs, sorted order of the step
    double4 *force;     // fx, fy, fz, es_dot
    double4 *torque;    // tx, ty, tz, ev_dot
    double* pres;       // p


    // cell list
    uint32_t *key, *rank, *perm, *count, *off, *scan_bsum;  // off: K+1 prefix
    uint32_t scan_blocks;

    // contact search tiles: TX x TY cells of one z plane per block
    uint32_t tile_x, tile_y, tiles_x, tiles_y;
    int search_mode;    // 1: LDS tiles, 2: rows per particle, 3: wave per cell, 4: half shell, 5: lists
    uint32_t tpb_search, tpb_force;  // threads per block of k_topo_rows / k_topo_half and of k_force (<= BLOCK)
    uint32_t force_vw;  // k_force gathers the partners' velocity records with their positions
    uint32_t* skey;     // cell id per sorted position
    float4* xf;         // FP32 copy of a.x for the search's distance prefilter
    float search_eps;   // prefilter margin (m), above the FP32 rounding of a distance
    float rmax_f;       // largest radius of any system (FP32, rounded up)
    float cell_margin;  // widening of a cell's box when its spheres are bounded by it (m)
    uint32_t* nlist;    // mode 5: np * NLIST partners (their list index), list index = sorted position at the build
    uint32_t* ncnt;     // mode 5: list length per list index
    float4* xref;       // mode 5: position at the last list build, per list index
    float* nlistg;      // mode 5: np * NLIST, each entry's gap at the build (FP32), transposed like nlist
    uint32_t* nlo;      // mode 5: np * NLIST, the entry's partner by original index, transposed like nlist
                        // (valid with the hint: both are written when the partner is found in a slot)
    uint8_t* nhint;     // mode 5: np * NLIST, the slot where the entry's partner was last found (a hint,
                        // checked against the slot word before use; >= 32 for none), transposed like nlist
    float* ldisp;       // mode 5: displacement since the build, per list index (FP32)
    uint32_t* linv;     // mode 5: list index -> sorted position of this step
    uint32_t* slidx;    // mode 5: list index per sorted position of this step
    uint32_t* nstate;   // mode 5: [0] rebuild this step, [1] rebuild requested, [2] build failed, [3] builds so far,
                        // [4] a sphere crossed a periodic face in the last step, [5] ... in this step
    float list_skin;    // mode 5: list reach beyond r_a + r_b (m)
    float list_lim;     // mode 5: displacement since the build that requests a rebuild (m)
    uint32_t list_shell;  // mode 5: list builds also scan the cells two away
    uint32_t* hcnt;     // half-shell search: overlaps found per particle
    uint64_t* hits;     // np * NC keys: offset index << 40 | image << 32 | sorted position
    uint32_t* tmask;    // slot occupancy after the search, sorted order
    uint32_t* tfresh;   // slots taken from empty by the search, sorted order

    // walls (6 entries per system; entries >= its nw hold the reference's defaults)
    double4 *wnx, *wmvfd;
    double* wvel0;
    int* wmode;
    double* wall_pp;       // nw * np, original order
    double* wall_partial;  // nw * nwall_blocks
    double* t_cur;         // device copy of time.current, one per system

    // bonds
    uint2* bonds;
    double4 *bdelta, *bomega;
    double4 *bond_f, *bond_t;
    uint32_t *bcsr_off, *bcsr_ent;  // per-particle bond lists, entry = k<<1 | side

    uint32_t* flags;
    uint32_t* flags_next;  // cell-id errors of the next step's positions, merged at its start
};

// kernel launchers (kernels.hip); all enqueue on `s`
void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
void launch_bonds(const DemDev& d, hipStream_t s);
// keys of the stored positions, with their per-cell counts and ranks (the
// counts must start at zero: k_scan_down clears them after use)
void launch_cellid(const DemDev& d, hipStream_t s);
void launch_scan(const DemDev& d, hipStream_t s);
void launch_scatter(const DemDev& d, hipStream_t s);
void launch_cellsort(const DemDev& d, hipStream_t s);
void launch_reorder(const DemDev& d, hipStream_t s);
void launch_contact(const DemDev& d, int precision, hipStream_t s);  // search + forces
void launch_topology(const DemDev& d, hipStream_t s);
void launch_forces(const DemDev& d, int precision, hipStream_t s);
// FP64 or FP32 fused multiply-adds of known count (16 * iters per thread, BLOCK threads per block)
void launch_flops_probe(void* out, uint32_t blocks, uint32_t iters, bool fp64, hipStream_t s);
void launch_walls(const DemDev& d, hipStream_t s);
// mixed mode: the input set's FP32 velocity copies from its FP64 records (after
// any change of the state outside a step; steps keep them)
void launch_f32_copies(const DemDev& d, hipStream_t s);
void launch_copy_probe(const double4* src, double4* dst, uint64_t n, hipStream_t s);

}  // namespace sph

#endif
