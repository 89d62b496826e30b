/*
 * dem_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial CPU restatement of the reference's DEM time step (sphere,
 * src/device.cu:747-1098 and the kernels it launches).  It is the checker
 * for the HIP implementation and the CPU baseline of the benchmark; only the
 * test suite, the smoke check and the benchmark's cpu_baseline leg may load
 * it.  Nothing in the product links, imports or calls it.
 *
 * Arrays are kept in the reference's own layout (AoS, original particle
 * order, NC contact slots per particle with the np sentinel) so that every
 * stage can be compared field by field with the GPU path.
 *
 * Parity pins (each checked by tests/test_oracle_*.py):
 *   cell id / sort / reorder ..... brute force (floor per particle,
 *                                  std::stable_sort, x[idx[i]] gather)
 *   topology ..................... O(N^2) overlap set with periodic images,
 *                                  hand-built slot cases
 *   contactLinear / wall ......... closed forms: restitution, rest-on-floor
 *                                  overlap m*g/k_n, stick-slip at mu_s*|f_n|,
 *                                  Newton's third law
 *   contactLinearViscous ......... (contact model 1) first-step pair force
 *                                  -k_n delta - gamma_n v_n, viscous shear
 *                                  gamma_t |v_t| below the Coulomb cap and
 *                                  mu_d |f_n| above it, no torque, Newton's
 *                                  third law, contact slots left untouched
 *   capillaryCohesion_exp ........ the closed form kappa R_geo exp(-d/lambda)
 *                                  for a pair at gap d < db; orc_exp against
 *                                  the C library's exp to 2 ulp
 *   bondsLinear .................. two-particle bond energy conservation and
 *                                  axial frequency sqrt(2 k_n A / m)
 *   contactLinear_wall friction .. sliding-to-rolling on the floor (5/7
 *                                  v0 at t = 2 v0 / (7 mu_wd g)), the
 *                                  viscous / static / dynamic caps and the
 *                                  1 mm/s switch, torque = arm x f_t, the
 *                                  moving wall's scalar velocity added to
 *                                  all three components
 *   integrate .................... leapfrog free-fall closed form; fixvel
 *                                  free flight bit-invariant velocities,
 *                                  z moved by the contact force alone;
 *                                  fixvel >= 2 (extension) holds vel_z too;
 *                                  periodic wrap before the cell id
 *   summation / integrateWalls ... velocity wall pos0 + v t, stress wall
 *                                  equilibrium
 *   whole step ................... zero-g no-contact bit invariance and the
 *                                  .bin round trip (tests/io_tests.py:8-34)
 *   porosity (both rules) ........ a sphere inside one slice, textbook cap
 *                                  volumes for each crossing branch, the
 *                                  exact rule's slice sum = total volume,
 *                                  Monte Carlo on a random packing
 *   findOverlaps ................. scipy's cKDTree pair query on a random
 *                                  packing (i < j, no periodic images)
 *   render ....................... ray-sphere hits and the shading of the
 *                                  chosen surface point from the textbook
 *                                  quadratic in FP64, occlusion order,
 *                                  cutoff and fixed-sphere colours, the
 *                                  integer height / width of the image plane
 */
#ifndef SPHERE_DEM_ORACLE_H_
#define SPHERE_DEM_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

/* fixvel at or above this value holds the z velocity too (an extension of
 * the reference's fixed horizontal velocity, src/integration.cuh:65-77). */
#define ORC_FIXVEL_HOLD_Z 2.0

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_t;

/* error flags returned by orc_step / orc_stage */
enum {
    ORC_F_OUT_OF_GRID = 1,
    ORC_F_CONTACT_OVERFLOW = 2,
    ORC_F_NAN = 4,
    ORC_F_INVALID = 8,
};

/* field ids for orc_get (shapes given in doubles or u32 per entry) */
enum {
    ORC_X = 0,          /* np x 4 f64: x, y, z, r                       */
    ORC_VEL = 1,        /* np x 4 f64: vx, vy, vz, fixvel               */
    ORC_ANGVEL = 2,     /* np x 3 f64                                   */
    ORC_FORCE = 3,      /* np x 3 f64                                   */
    ORC_TORQUE = 4,     /* np x 3 f64                                   */
    ORC_ANGPOS = 5,     /* np x 3 f64                                   */
    ORC_XYSUM = 6,      /* np x 2 f64                                   */
    ORC_VEL0 = 7,       /* np x 3 f64                                   */
    ORC_ANGVEL0 = 8,    /* np x 3 f64                                   */
    ORC_ES_DOT = 9,     /* np f64                                       */
    ORC_ES = 10,
    ORC_EV_DOT = 11,
    ORC_EV = 12,
    ORC_P = 13,
    ORC_CONTACTS = 14,  /* np x NC u32, empty slot = np                 */
    ORC_DISTMOD = 15,   /* np x NC x 4 f64 (dx, dy, dz, r_b)            */
    ORC_DELTA_T = 16,   /* np x NC x 3 f64                              */
    ORC_CELL_KEYS = 17, /* np u32, cell id per original particle        */
    ORC_SORT_IDX = 18,  /* np u32, original index per sorted position   */
    ORC_CELL_START = 19,/* K u32, 0xffffffff when empty                 */
    ORC_CELL_END = 20,  /* K u32, 0xffffffff when empty                 */
    ORC_BONDS = 21,     /* nb0 x 2 u32                                  */
    ORC_BONDS_DELTA = 22, /* nb0 x 4 f64: t.x, t.y, t.z, n              */
    ORC_BONDS_OMEGA = 23, /* nb0 x 4 f64                                */
    ORC_WALLS_NX = 24,  /* nw x 4 f64: nx, ny, nz, pos                  */
    ORC_WALLS_MVFD = 25,/* nw x 4 f64: mass, vel, force, devs           */
    ORC_WALLS_VEL0 = 26,/* nw f64                                       */
    ORC_WALL_PP = 27,   /* nw x np f64                                  */
    ORC_TIME = 28,      /* 4 f64: dt, current, total, file_dt           */
    ORC_X_SORTED = 29,  /* np x 4 f64                                   */
    ORC_NFIELDS = 30
};

/* stages of one step, in execution order */
enum {
    ORC_STAGE_BONDS = 0,      /* bondsLinear forces + state update         */
    ORC_STAGE_CELLID = 1,     /* wrap + calcParticleCellID                 */
    ORC_STAGE_SORT = 2,       /* stable sort (key, index)                  */
    ORC_STAGE_REORDER = 3,    /* cellStart/cellEnd + gather                */
    ORC_STAGE_TOPOLOGY = 4,
    ORC_STAGE_INTERACT = 5,
    ORC_STAGE_BOND_APPLY = 6, /* add bond forces in ascending bond index   */
    ORC_STAGE_INTEGRATE = 7,
    ORC_STAGE_SUMMATION = 8,
    ORC_STAGE_WALLS = 9,
    ORC_STAGE_TIME = 10,
    ORC_NSTAGES = 11
};

/* Load a .bin file; nc = contact slots per particle (reference: 32). */
orc_t* orc_open(const char* path, int nc, char* err, size_t errlen);
void orc_close(orc_t* s);
int orc_writebin(orc_t* s, const char* path);
/* n full steps; returns accumulated error flags (0 = clean). */
int orc_step(orc_t* s, uint64_t n);
int orc_stage(orc_t* s, int stage);
/* bytes needed for a field, or 0 for an unknown id */
size_t orc_field_bytes(const orc_t* s, int field);
int orc_get(const orc_t* s, int field, void* dst, size_t bytes);
int orc_set(orc_t* s, int field, const void* src, size_t bytes);
/* np, nb0, nw, nc, K */
void orc_dims(const orc_t* s, uint64_t* out5);
/* k_n, k_t, gamma_n, gamma_t, mu_s, mu_d (set: same order) */
void orc_get_params(const orc_t* s, double* out6);
void orc_set_params(orc_t* s, const double* in6);

/* The exponential the capillary force uses (see dem_oracle.cpp). */
double orc_exp(double x);

/* DEM::porosity: rule 0 as the reference computes it, 1 exact slab volumes.
 * Returns 0, 1 when the slices are too thin for rule 0, 2 for slices < 1. */
int orc_porosity(const orc_t* s, int slices, int rule, double* z_pos, double* porosity);
/* DEM::findOverlaps: pairs (i, j), i < j, i major; returns the count and
 * writes at most `capacity` pairs. */
uint64_t orc_find_overlaps(const orc_t* s, uint32_t* ij, double* delta_n, uint64_t capacity);
/* DEM::render: rgb receives width * height * 3 bytes; method 0 shaded
 * spheres, 1..5 colour map of pressure, |vel|, |angvel|, x displacement,
 * |angpos|.  Returns 0, or 2 for a bad method or an empty image. */
int orc_render(const orc_t* s, int method, float max_val, float lower_cutoff, uint32_t width, uint32_t height,
               uint8_t* rgb);

#ifdef __cplusplus
}
#endif
#endif
