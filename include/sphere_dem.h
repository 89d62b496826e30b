/*
 * sphere_dem.h -- C-ABI of the MI355X (gfx950) DEM time step.
 *
 * The reference's boundary for this path is the C++ DEM class plus the .bin
 * file (sphere, src/sphere.h:248-321; src/file_io.cpp:29-490).  Each entry
 * point below replaces one piece of that class; the reference member it
 * stands in for is named beside it.  Plain pointers and sizes only; every
 * call returns SPH_OK (0) or a negative SPH_E_* code instead of exiting.
 * A handle owns one HIP stream and its device buffers; it is not thread safe.
 * Several handles may coexist: no global constant memory is used.
 */
#ifndef SPHERE_DEM_H_
#define SPHERE_DEM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 1
/* contact slots per particle (src/constants.h:24) */
#define SPH_NC 32

typedef struct sph_dem sph_t;

enum sph_status {
    SPH_OK = 0,
    SPH_E_IO = -1,               /* file missing, short read, short write      */
    SPH_E_INVALID = -2,          /* checkValues failure or bad argument        */
    SPH_E_OUT_OF_GRID = -3,      /* particle more than one cell off the grid   */
    SPH_E_CONTACT_OVERFLOW = -4, /* more than SPH_NC contacts on a particle    */
    SPH_E_NAN = -5,              /* non-finite position                        */
    SPH_E_HIP = -6,              /* HIP runtime error                          */
    SPH_E_STATE = -7,            /* call not valid in the handle's state       */
    SPH_E_UNSUPPORTED = -8       /* input uses a feature outside this path     */
};

/* Particles with fixvel (the 4th component of SPH_VEL) > 0 keep their x and
 * y velocity, feel no gravity and do not spin up (src/integration.cuh:65-77;
 * the reference's inputs use fixvel = 1).  Extension with no reference
 * counterpart: fixvel >= SPH_FIXVEL_HOLD_Z also keeps the z velocity, a
 * velocity-controlled particle layer (tension and compression tests). */
#define SPH_FIXVEL_HOLD_Z 2.0

/* Field ids for sph_get / sph_set.  Shapes in entries of f64 or u32,
 * original particle order, the reference's own layout. */
enum sph_field {
    SPH_X = 0,          /* np x 4 f64: x, y, z, r                              */
    SPH_VEL = 1,        /* np x 4 f64: vx, vy, vz, fixvel                      */
    SPH_ANGVEL = 2,     /* np x 3 f64                                          */
    SPH_FORCE = 3,      /* np x 3 f64                                          */
    SPH_TORQUE = 4,     /* np x 3 f64                                          */
    SPH_ANGPOS = 5,     /* np x 3 f64                                          */
    SPH_XYSUM = 6,      /* np x 2 f64                                          */
    SPH_VEL0 = 7,       /* np x 3 f64, half-step velocity                      */
    SPH_ANGVEL0 = 8,    /* np x 3 f64                                          */
    SPH_ES_DOT = 9,     /* np f64                                              */
    SPH_ES = 10,
    SPH_EV_DOT = 11,
    SPH_EV = 12,
    SPH_P = 13,
    SPH_CONTACTS = 14,  /* np x NC u32, empty slot = np  (get only)            */
    SPH_DISTMOD = 15,   /* np x NC x 4 f64 (dx, dy, dz, r_b) (get only)        */
    SPH_DELTA_T = 16,   /* np x NC x 3 f64 (get only)                          */
    SPH_CELL_KEYS = 17, /* np u32, cell id per original particle (get only)    */
    SPH_SORT_IDX = 18,  /* np u32, original index per sorted position (get)    */
    SPH_CELL_START = 19,/* K u32, 0xffffffff for an empty cell (get only)      */
    SPH_CELL_END = 20,  /* K u32, 0xffffffff for an empty cell (get only)      */
    SPH_BONDS = 21,     /* nb0 x 2 u32                                         */
    SPH_BONDS_DELTA = 22, /* nb0 x 4 f64: t.x, t.y, t.z, n                     */
    SPH_BONDS_OMEGA = 23, /* nb0 x 4 f64                                       */
    SPH_WALLS_NX = 24,  /* nw x 4 f64: nx, ny, nz, pos                         */
    SPH_WALLS_MVFD = 25,/* nw x 4 f64: mass, vel, force, devs                  */
    SPH_WALLS_VEL0 = 26,/* nw f64                                              */
    SPH_WALL_PP = 27,   /* nw x np f64, per-particle wall force (get only)     */
    SPH_TIME = 28,      /* 4 f64: dt, current, total, file_dt                  */
    SPH_X_SORTED = 29,  /* np x 4 f64 (get only)                               */
    SPH_NFIELDS = 30
};

/* Arithmetic of the pair forces. */
enum sph_precision {
    SPH_PREC_FP64 = 0,  /* everything in FP64 (the reference's Float=double)  */
    SPH_PREC_MIXED = 1  /* pair forces in FP32; positions, velocities and the
                           per-particle force sums in FP64                    */
};

/* Sizes of an open handle. */
typedef struct sph_info {
    uint64_t np, nb0, nw, nc, ncells;
    uint32_t num[3];
    int32_t periodic;
    double dt, t_current, t_total, file_dt;
    uint64_t steps_done;
    int32_t precision;
} sph_info_t;

/* Contact parameters a caller may change between steps (the differentiable
 * set of the adjoint plus the tangential pair). */
typedef struct sph_params {
    double k_n, k_t, gamma_n, gamma_t, mu_s, mu_d;
} sph_params_t;

/* Per-kernel timing of the last sph_step call (hipEvent, milliseconds,
 * summed over the steps of the call; valid when profiling is enabled). */
typedef struct sph_timing {
    double bonds_ms, cellid_ms, sort_ms, reorder_ms, contact_ms, walls_ms;
    double total_ms;
    uint64_t steps;
    double search_ms, force_ms;  /* contact_ms = search_ms + force_ms */
    uint32_t tile_x, tile_y;     /* cells per contact-search block */
    int32_t search_mode;         /* contact-search kernel in use (sph_set_search) */
    int32_t search_auto;         /* 1 when the kernel is chosen by timing */
    int32_t force_variant;       /* k_force in use: 1 per-slot velocity gathers, 2 gathered ahead (sph_set_force_kernel) */
    uint32_t list_builds;        /* neighbour-list builds (search mode 5) since the handle was opened */
} sph_timing_t;

/* DEM::DEM + readbin + checkValues + allocateGlobalDeviceMemory +
 * transferToGlobalDeviceMemory (src/sphere.cpp:16-77, src/file_io.cpp:29-301,
 * src/sphere.cpp:104-246, src/device.cu:233-324, :394-488).
 * check_vals=1 applies checkValues; verbosity 0 is silent. */
int sph_open(const char* bin_path, int verbosity, int check_vals, sph_t** out);

/* The reference's dry run (DEM(..., dry=1), src/sphere.cpp:42-48): parse an
 * in-memory .bin image and apply checkValues without touching the GPU.  When
 * out is non-NULL the parsed state is serialised again into out (capacity
 * out_cap bytes, size written to *out_len) for byte-level round trips. */
int sph_dry_run(const void* bin, size_t bytes, int check_vals, void* out, size_t out_cap,
                size_t* out_len, char* err, size_t err_len);

/* Same as sph_open for an in-memory .bin image. */
int sph_open_mem(const void* bin, size_t bytes, int verbosity, int check_vals, sph_t** out);

/* ~DEM + freeGlobalDeviceMemory (src/sphere.cpp:80-100, src/device.cu:326-391). */
void sph_close(sph_t* h);

/* Message of the last failing call on this handle (or of the last failing
 * sph_open on this thread when h is NULL). */
const char* sph_last_error(const sph_t* h);

/* n time steps of the loop body of DEM::startTime (src/device.cu:731-1111)
 * without file output; time.current advances by dt per step. */
int sph_step(sph_t* h, uint64_t n);

/* DEM::startTime (src/device.cu:576-1299): steps while
 * time.current <= time.total, writing <outdir>/<sid>.outputNNNNN.bin every
 * file_dt and <outdir>/<sid>.status.dat. */
int sph_run(sph_t* h, const char* outdir, const char* sid);

/* DEM::writebin (src/file_io.cpp:304-490). */
int sph_writebin(sph_t* h, const char* path);

/* DEM::readbin (src/file_io.cpp:29-301) on an open handle: replaces the
 * simulation with the file's (host state, device buffers, step count),
 * keeping the handle's settings (precision, search kernel, graph capture,
 * profiling, verbosity) and applying checkValues as the handle was opened
 * with.  The file is parsed and checked before anything is replaced: on
 * SPH_E_IO, SPH_E_INVALID or SPH_E_UNSUPPORTED the handle is unchanged. */
int sph_readbin(sph_t* h, const char* path);

/* DEM::checkValues (src/sphere.cpp:104-246) on the current device state. */
int sph_check_values(sph_t* h);

/* transferFromGlobalDeviceMemory / transferToGlobalDeviceMemory for one
 * field (src/device.cu:394-572); dst/src are host pointers owned by the
 * caller and bytes must equal sph_field_bytes. */
size_t sph_field_bytes(const sph_t* h, int field);
int sph_get(sph_t* h, int field, void* dst, size_t bytes);
int sph_set(sph_t* h, int field, const void* src, size_t bytes);

int sph_info(const sph_t* h, sph_info_t* out);
int sph_get_params(const sph_t* h, sph_params_t* out);
/* Replaces the constant-memory upload of Params (src/device.cu:201-228). */
int sph_set_params(sph_t* h, const sph_params_t* in);

/* Chooses the pair-force arithmetic; a handle opens in SPH_PREC_FP64. */
int sph_set_precision(sph_t* h, int precision);

/* 1 = time each kernel with hipEvents in sph_step (the PROFILING flag of
 * src/debug.h:7); 0 = off (default). */
int sph_set_profiling(sph_t* h, int on);
int sph_get_timing(const sph_t* h, sph_timing_t* out);

/* 1 = replay the captured step as a HIP graph (default), 0 = plain launches. */
int sph_set_graph(sph_t* h, int on);

/* Cells per contact-search block (tile_x along x, tile_y along y); 0, 0
 * restores the automatic choice made at open from the cell occupancy. */
int sph_set_tile(sph_t* h, uint32_t tile_x, uint32_t tile_y);

/* Contact-search kernel: 1 = one block per tile with the neighbourhood staged
 * in LDS, 2 = one thread per sorted particle reading through L1/L2, 3 = one
 * wavefront per cell, one lane per candidate, 4 = half shell (each pair
 * tested once, hits replayed in the reference's order), 5 = neighbour lists
 * (pairs within r_a + r_b + skin, rebuilt when a sphere has moved 0.45 skin
 * or half of cell - 2 r_max, whichever is less;
 * overlaps replayed in the reference's order), 0 = automatic (the default):
 * every 4096 steps sph_step times a few steps with each of 2, 4 and 5 and
 * keeps the fastest.  All kernels give bit-identical results. */
int sph_set_search(sph_t* h, int mode);

/* Force kernel: 1 = each contact's partner velocities are gathered when the
 * contact is processed, 2 = gathered ahead with the partner positions, four
 * slots at a time (more registers: faster with many contacts per sphere),
 * 0 = automatic (the default): timed with the search kernels on the first
 * long enough sph_step and every 4096 steps.  All give the same bits. */
int sph_set_force_kernel(sph_t* h, int variant);

/* Measurement helper with no counterpart in the reference: copies `bytes`
 * between two fresh device buffers `reps` times with a plain streaming kernel
 * (k_copy_probe) and returns the achieved read+write rate in GB/s. */
int sph_bandwidth_probe(uint64_t bytes, int reps, double* gbs);

/* Measurement helper with no counterpart in the reference: `blocks` blocks
 * of 256 threads, each thread running iters x 8 independent fused
 * multiply-adds (16 x iters floating-point operations) in FP64 (fp64 != 0)
 * or FP32, `reps` times (k_flops_probe); returns the achieved TFLOP/s.  It
 * calibrates the VALU FLOP counters of a profiler against known work. */
int sph_flops_probe(int fp64, uint32_t blocks, uint32_t iters, int reps, double* tflops);

/* Adjoint of the time step, with no counterpart in the reference: gradients
 * of a time-mean force objective
 *     J = 1/nsteps sum_{steps} sum_i w_i . F_i
 * (F_i the particle's total force of the step, w_i = weights[3i..3i+2] in
 * the file's particle order; e.g. w_i = (-1/(L0 L1), 0, 0) on the driven
 * layer gives the mean shear stress) with respect to
 * theta = (k_n, k_t, gamma_n, mu), mu = mu_s = mu_d.
 * sph_adjoint_forward runs nsteps steps from the current state, keeping a
 * device checkpoint every seg_len steps, and returns J.  The handle is left
 * at the end of the run.  sph_adjoint_backward recomputes each segment from
 * its checkpoint with a per-step tape (about 1 KB per particle and step) and
 * runs the discrete adjoint backwards; grad[4] = dJ/dtheta.  Contact set and
 * force-law branches are those of the forward run.  Handles with bonds
 * return SPH_E_UNSUPPORTED. */
int sph_adjoint_forward(sph_t* h, uint64_t nsteps, uint32_t seg_len, const double* weights, double* J);
int sph_adjoint_backward(sph_t* h, double grad[4]);

/* Setup helper (host only): seeded non-overlapping random sequential packing
 * of n spheres with the given radii in box[3] (periodic: 0 none, 1 x+y, 2 x),
 * largest first; xyz[3n] receives the centres.  Stands in for the
 * reference's unseeded initRandomPos (python/sphere.py:856-908).  Returns 0
 * or -1 when a sphere finds no free spot in max_tries draws. */
int sph_pack_rsa(uint64_t n, const double* radii, const double* box, int periodic, uint64_t seed,
                 uint64_t max_tries, double* xyz);

/* ------------------------------------------------------- host analysis
 * The reference's analysis members, computed on the device from the
 * handle's current state (after sph_open or any sph_step). */

/* DEM::porosity (src/sphere.cpp:363-483, the `porosity` tool of
 * src/porosity.cpp): porosity of `slices` horizontal slices of equal
 * thickness (top - origo_z) / slices from z = 0, top being wall 0's current
 * position when the input has walls, else L_z; z_pos[k] is the slice's mid
 * height.  rule SPH_POROSITY_REFERENCE removes what the reference removes,
 * including V_sphere + cap (not V_sphere - cap) for a sphere whose centre is
 * in the slice but which crosses a slice plane, and refuses slices no
 * thicker than twice the last particle's radius (the reference's r_max);
 * SPH_POROSITY_EXACT removes the exact volume of every sphere inside the
 * slice and has no thickness limit.  SPH_E_INVALID for slices < 1 or the
 * refusal (message "Error! The number of z-slices is too high."). */
#define SPH_POROSITY_REFERENCE 0
#define SPH_POROSITY_EXACT 1
int sph_porosity(sph_t* h, int slices, int rule, double* z_pos, double* porosity);

/* No counterpart in the reference: sph_porosity for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * sph_porosity, bit for bit, but the call allocates nothing: its device work
 * space is one static array per GPU of SPH_POROSITY_NOALLOC_MAX doubles
 * (8 MiB, held from the library's load until the process ends), which must
 * fit 1793 + slices * (ceil(np / 2048) + 2) doubles -- e.g. up to 2,036
 * slices at 1,048,576 spheres; larger requests return SPH_E_INVALID.  Every
 * call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (from several threads or handles) race. */
#define SPH_POROSITY_NOALLOC_MAX 1048576
int sph_porosity_noalloc(sph_t* h, int slices, int rule, double* z_pos, double* porosity);

/* DEM::findOverlaps (src/sphere.cpp:553-604, used by the `forcechains`
 * tool): every pair of particles i < j (original indices) whose centres are
 * closer than r_i + r_j, periodic images not considered, ordered by i then
 * j; delta_n = |x_j - x_i| - (r_i + r_j) < 0.  *count receives the number of
 * pairs.  ij[2k], ij[2k+1] and delta_n[k] are written when capacity >=
 * *count; with ij or delta_n NULL (a size query) nothing else is written.
 * SPH_E_INVALID when the buffers are given but too small. */
int sph_find_overlaps(sph_t* h, uint32_t* ij, double* delta_n, uint64_t capacity, uint64_t* count);

/* DEM::render (src/raytracer.cuh:374-553, `sphere -r`): ray-traced image of
 * the current state from the reference's camera (eye at (2.3, -5, 1.3) x L,
 * looking at half the largest position), in FP32 as there.  method 0
 * shades every sphere; 1..5 colour by pressure, |velocity|, |angular
 * velocity|, x displacement, |angular position| over [0, max_val], with
 * lower_cutoff > 0 hiding spheres at or below it and fixed ones.  rgb
 * receives width * height * 3 bytes (R, G, B; the reference's pixel
 * order, as its PPM writes them; white background).  cull 1 bins the
 * spheres into 16 x 16-pixel tiles first, cull 0 tests every sphere for
 * every pixel as the reference does; both give the same bytes.
 * width * height <= 2^24. */
int sph_render(sph_t* h, int method, float max_val, float lower_cutoff, uint32_t width, uint32_t height, int cull,
               uint8_t* rgb);

/* ------------------------------------------------------------ ensembles
 * No counterpart in the reference, which steps one DEM per process: an
 * ensemble steps many independent systems (each a .bin image with its own
 * particle count, box, grid, parameters, time step and walls) together, one
 * launch of each kernel for all of them per step, the step captured in a HIP
 * graph.  Every member evolves exactly as it would in its own handle on the
 * same grid (bit for bit in FP64).  Members may not have bonds
 * (SPH_E_UNSUPPORTED); at most 65535 members. */
typedef struct sph_ens sph_ens_t;

/* Opens an ensemble from n in-memory .bin images (sph_ens_open: files).
 * checkValues is applied to every member.  cell_factor < 0 keeps each
 * image's grid; cell_factor >= 1 regrids every member to cells of
 * cell_factor x 2 r_max (r_max of the member; an axis keeps at least one
 * cell, a periodic one at least three); cell_factor = 0 lets the autotuner
 * choose the factor among 1, 1.25, 1.5 and 2 by timing a few steps of
 * trial ensembles (discarded) built from the same images. */
int sph_ens_open_mem(const void* const* images, const size_t* bytes, uint32_t n, double cell_factor,
                     int verbosity, sph_ens_t** out);
int sph_ens_open(const char* const* paths, uint32_t n, double cell_factor, int verbosity, sph_ens_t** out);
void sph_ens_close(sph_ens_t* e);
const char* sph_ens_last_error(const sph_ens_t* e);

/* nsteps steps of every member (sph_step for each). */
int sph_ens_step(sph_ens_t* e, uint64_t nsteps);

/* Fields of member k in that member's own numbering and sizes, as sph_get
 * returns them for a handle opened on the member alone; the sort fields and
 * SPH_DISTMOD are not available (SPH_E_UNSUPPORTED). */
size_t sph_ens_field_bytes(const sph_ens_t* e, uint32_t k, int field);
int sph_ens_get(sph_ens_t* e, uint32_t k, int field, void* dst, size_t bytes);
/* Sizes of member k (grid after any regridding). */
int sph_ens_member_info(const sph_ens_t* e, uint32_t k, sph_info_t* out);

typedef struct sph_ens_info {
    uint32_t members;
    uint64_t np, ncells, steps_done;
    double cell_factor;          /* in use; < 0: the images' own grids */
    double tuned_ms[4];          /* autotuner: ms per step of each trial factor (0: not tried) */
    int32_t search_mode;         /* contact-search kernel in use (2, 4 or 5) */
    uint32_t force_block, search_block;  /* threads per block of k_force and of the search kernel */
    int32_t precision;
} sph_ens_info_t;
int sph_ens_info(const sph_ens_t* e, sph_ens_info_t* out);

/* Kernel choices, as for one handle: precision; search mode 0 (automatic: on
 * the next sph_ens_step the autotuner times the row, half-shell and list
 * kernels and then 64, 128 and 256 threads per block for the search and
 * force kernels, on ordinary steps, and keeps the fastest), 2, 4 or 5 (1 and
 * 3 walk one grid and are refused); graph capture; per-kernel profiling. */
int sph_ens_set_precision(sph_ens_t* e, int precision);
int sph_ens_set_search(sph_ens_t* e, int mode);
int sph_ens_set_blocks(sph_ens_t* e, uint32_t force_block, uint32_t search_block);
int sph_ens_set_graph(sph_ens_t* e, int on);
int sph_ens_set_profiling(sph_ens_t* e, int on);
int sph_ens_get_timing(const sph_ens_t* e, sph_timing_t* out);

#ifdef __cplusplus
}
#endif
#endif
