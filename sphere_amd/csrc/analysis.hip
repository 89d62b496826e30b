// analysis.hip -- the reference's host analysis members, computed on the
// device from the handle's resident state: DEM::porosity
// (src/sphere.cpp:363-483) and DEM::findOverlaps (src/sphere.cpp:553-604).
//
// Both read the current buffer set (positions and radii in x, original
// index in id) in whatever order the last step left it.  The reference loops
// over every particle for every slice and over every pair; here the porosity
// visits only the slices a block's spheres can reach, and the overlap search
// bins the spheres into cells of 2 r_max and scans the 27 neighbouring cells,
// then orders the pairs by (i, j) as the reference's double loop emits them.
// Distances and volumes use the reference's expressions with FP contraction
// off, so every delta_n is bit-identical to the double loop's; porosity sums
// are taken in a fixed order (a tree per block, then a fixed pattern over
// the block sums), which
// differs from the reference's running subtraction only by rounding.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine_internal.h"
#include "../../include/sphere_dem.h"

namespace {

constexpr int ABLOCK = 256;
constexpr int PPT = 8;  // particles per thread of k_porosity
constexpr int SCAN_CHUNK = 4 * ABLOCK;

#define A_TRY(expr)                                                                \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return SPH_E_HIP;                                                      \
        }                                                                          \
    } while (0)

const sph::PState& cur(const sph_dem* h) { return h->own_cur ? h->d.a : h->d.b; }

// Device buffers of one call, freed on every return path.
struct Scratch {
    std::vector<void*> p;
    ~Scratch() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    hipError_t get(T** out, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T*>(q);
        return e;
    }
};

// ------------------------------------------------------------------ bounds
// Per block: min x, y, z, max x, y, z, max r.  The radius of original index
// n - 1 goes to *r_last.
__global__ __launch_bounds__(ABLOCK) void k_bounds(const double4* __restrict__ x, const uint32_t* __restrict__ id,
                                                   uint32_t n, double* __restrict__ part, double* __restrict__ r_last) {
    __shared__ double red[7][ABLOCK];
    double v[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (uint32_t i = blockIdx.x * ABLOCK + threadIdx.x; i < n; i += gridDim.x * ABLOCK) {
        const double4 p = x[i];
        v[0] = fmin(v[0], p.x); v[1] = fmin(v[1], p.y); v[2] = fmin(v[2], p.z);
        v[3] = fmax(v[3], p.x); v[4] = fmax(v[4], p.y); v[5] = fmax(v[5], p.z);
        v[6] = fmax(v[6], p.w);
        // a non-finite coordinate poisons the bounds so that the host refuses
        if (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z) || !isfinite(p.w)) v[6] = NAN;
        if (id[i] == n - 1) *r_last = p.w;
    }
    for (int c = 0; c < 7; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = ABLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + w]);
            for (int c = 3; c < 6; ++c) red[c][threadIdx.x] = fmax(red[c][threadIdx.x], red[c][threadIdx.x + w]);
            const double a = red[6][threadIdx.x], b = red[6][threadIdx.x + w];
            red[6][threadIdx.x] = (isnan(a) || isnan(b)) ? NAN : fmax(a, b);
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[7 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// bounds[7] and the last particle's radius, on the host, with the device
// work space given: part holds 7 * BOUNDS_BLOCKS doubles, rl one.
constexpr uint32_t BOUNDS_BLOCKS = 256;
constexpr size_t BOUNDS_WS = 7 * (size_t)BOUNDS_BLOCKS + 1;  // doubles

int bounds_in(sph_dem* h, const sph::PState& st, uint32_t n, double* part, double* rl, double out[7],
              double* r_last) {
    const uint32_t nblk = std::max<uint32_t>(1, std::min<uint32_t>(BOUNDS_BLOCKS, (n + ABLOCK - 1) / ABLOCK));
    const double zero = 0.0;
    A_TRY(hipMemcpyAsync(rl, &zero, 8, hipMemcpyHostToDevice, h->stream));
    k_bounds<<<nblk, ABLOCK, 0, h->stream>>>(st.x, st.id, n, part, rl);
    A_TRY(hipGetLastError());
    double hp[7 * BOUNDS_BLOCKS];
    A_TRY(hipMemcpyAsync(hp, part, 7 * (size_t)nblk * 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipMemcpyAsync(r_last, rl, 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipStreamSynchronize(h->stream));
    for (int c = 0; c < 7; ++c) out[c] = hp[c];
    for (uint32_t b = 1; b < nblk; ++b) {
        for (int c = 0; c < 3; ++c) out[c] = std::min(out[c], hp[7 * b + c]);
        for (int c = 3; c < 6; ++c) out[c] = std::max(out[c], hp[7 * b + c]);
        out[6] = (std::isnan(out[6]) || std::isnan(hp[7 * b + 6])) ? NAN : std::max(out[6], hp[7 * b + 6]);
    }
    return SPH_OK;
}

int bounds(sph_dem* h, const sph::PState& st, uint32_t n, Scratch& s, double out[7], double* r_last) {
    double* ws;
    A_TRY(s.get(&ws, BOUNDS_WS));
    return bounds_in(h, st, n, ws, ws + 7 * (size_t)BOUNDS_BLOCKS, out, r_last);
}

// ---------------------------------------------------------------- porosity
// sphericalCap, src/sphere.cpp:343-347
__device__ inline double cap_volume(double hc, double r) { return M_PI * hc * hc / 3.0 * (3.0 * r - hc); }

// Exact volume of the sphere between the planes lo and hi:
// pi (r^2 t - t^3 / 3) from t1 to t2, t measured from the centre.
__device__ inline double slab_volume(double zc, double r, double lo, double hi) {
    const double t1 = fmax(lo - zc, -r), t2 = fmin(hi - zc, r);
    if (!(t2 > t1)) return 0.0;
    return M_PI * (r * r * (t2 - t1) - (t2 * t2 * t2 - t1 * t1 * t1) / 3.0);
}

// Volume one sphere removes from slice [lo, hi]: the reference's branches
// (src/sphere.cpp:408-454, each subtraction a separate term) or the exact
// slab volume.
__device__ inline double removed(double zc, double r, double lo, double hi, int rule) {
    if (rule != SPH_POROSITY_REFERENCE) return slab_volume(zc, r, lo, hi);
    const double zl = zc - r, zh = zc + r;
    const double Vs = 4.0 / 3.0 * M_PI * r * r * r;
    if (lo <= zl && zh <= hi) return Vs;
    double acc = 0.0;
    if (lo >= zc && lo <= zh)
        acc += cap_volume(zh - lo, r);
    else if (lo < zc && lo > zl)
        acc += Vs + cap_volume(lo - zl, r);
    if (hi <= zc && hi >= zl)
        acc += cap_volume(hi - zl, r);
    else if (hi > zc && hi < zh)
        acc += Vs + cap_volume(zh - hi, r);
    return acc;
}

// One block: ABLOCK * PPT spheres, each thread PPT of them (strided, so the
// loads coalesce).  Every branch above needs lo <= z + r and hi >= z - r, so
// a sphere reaches at most slices floor((z - r) / h) - 1 .. floor((z + r) / h)
// + 1 (two of margin each side; the branch tests themselves decide).  The
// block visits the union of its spheres' ranges and writes each slice's
// block sum to part[slice * gridDim.x + block].
__global__ __launch_bounds__(ABLOCK) void k_porosity(const double4* __restrict__ x, uint32_t n, int slices,
                                                     double h_slice, int rule, double* __restrict__ part) {
    __shared__ double red[ABLOCK];
    __shared__ int range[2];
    double zc[PPT], rr[PPT];
    int lo_s = INT_MAX, hi_s = -1;
    const uint32_t base = blockIdx.x * (uint32_t)(ABLOCK * PPT) + threadIdx.x;
    for (int k = 0; k < PPT; ++k) {
        const uint32_t i = base + (uint32_t)k * ABLOCK;
        zc[k] = 0.0;
        rr[k] = -1.0;  // no sphere
        if (i < n) {
            const double4 p = x[i];
            zc[k] = p.z;
            rr[k] = p.w;
            const double a = fmax(floor((p.z - p.w) / h_slice) - 2.0, 0.0);
            const double b = fmin(floor((p.z + p.w) / h_slice) + 2.0, (double)(slices - 1));
            if (a <= b) {
                lo_s = min(lo_s, (int)a);
                hi_s = max(hi_s, (int)b);
            }
        }
    }
    if (threadIdx.x == 0) {
        range[0] = INT_MAX;
        range[1] = -1;
    }
    __syncthreads();
    if (hi_s >= 0) {
        atomicMin(&range[0], lo_s);
        atomicMax(&range[1], hi_s);
    }
    __syncthreads();
    const int s0 = range[0], s1 = range[1];
    for (int s = s0; s <= s1; ++s) {
        const double lo = s * h_slice, hi = lo + h_slice;
        double acc = 0.0;
        for (int k = 0; k < PPT; ++k)
            if (rr[k] >= 0.0) acc += removed(zc[k], rr[k], lo, hi, rule);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = ABLOCK / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = red[0];
        __syncthreads();
    }
}

// One wavefront per slice: lane l adds the block sums l, l + 64, ... in
// order, then the lanes combine in a fixed butterfly.
__global__ __launch_bounds__(64) void k_porosity_final(const double* __restrict__ part, uint32_t nblk, int slices,
                                                       double h_slice, double V_slice, double* __restrict__ z_pos,
                                                       double* __restrict__ por) {
    const int s = blockIdx.x;
    double sum = 0.0;
    for (uint32_t b = threadIdx.x; b < nblk; b += 64) sum += part[(size_t)s * nblk + b];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (threadIdx.x != 0) return;
    const double lo = s * h_slice;
    z_pos[s] = lo + 0.5 * h_slice;
    por[s] = (V_slice - sum) / V_slice;
}

__global__ __launch_bounds__(ABLOCK) void k_zero_f64(double* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)ABLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * ABLOCK) p[i] = 0.0;
}

__global__ __launch_bounds__(ABLOCK) void k_zero_u32(uint32_t* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)ABLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * ABLOCK) p[i] = 0u;
}

__global__ __launch_bounds__(ABLOCK) void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
    for (size_t i = blockIdx.x * (size_t)ABLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * ABLOCK) p[i] = v;
}

uint32_t fill_blocks(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>(4096, (n + ABLOCK - 1) / ABLOCK)); }

// --------------------------------------------------------- exclusive scan
// out[0..n) = exclusive prefix sums of in[0..n) (u32), in three launches:
// per-chunk scans with chunk totals, a scan of the totals, the add-back.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < ABLOCK; off <<= 1) {
        const uint32_t t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    const uint32_t incl = sh[threadIdx.x];
    *total = sh[ABLOCK - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(ABLOCK) void k_scan_chunks(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        size_t n, uint32_t* __restrict__ chunk_sum) {
    __shared__ uint32_t sh[ABLOCK];
    const size_t base = (size_t)blockIdx.x * SCAN_CHUNK + 4 * (size_t)threadIdx.x;
    uint32_t v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        t += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(t, sh, &total);
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) chunk_sum[blockIdx.x] = total;
}

// one block: exclusive scan of the chunk totals in place
__global__ __launch_bounds__(ABLOCK) void k_scan_sums(uint32_t* __restrict__ sums, uint32_t m) {
    __shared__ uint32_t sh[ABLOCK];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < m; b0 += ABLOCK) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < m ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, sh, &total);
        if (i < m) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(ABLOCK) void k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                     const uint32_t* __restrict__ chunk_sum) {
    const size_t base = (size_t)blockIdx.x * SCAN_CHUNK + 4 * (size_t)threadIdx.x;
    const uint32_t add = chunk_sum[blockIdx.x];
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

int exclusive_scan(sph_dem* h, const uint32_t* in, uint32_t* out, size_t n, Scratch& s) {
    const size_t chunks = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    uint32_t* sums;
    A_TRY(s.get(&sums, chunks));
    k_scan_chunks<<<(uint32_t)chunks, ABLOCK, 0, h->stream>>>(in, out, n, sums);
    k_scan_sums<<<1, ABLOCK, 0, h->stream>>>(sums, (uint32_t)chunks);
    k_scan_add<<<(uint32_t)chunks, ABLOCK, 0, h->stream>>>(out, n, sums);
    A_TRY(hipGetLastError());
    return SPH_OK;
}

// ---------------------------------------------------------------- overlaps
// Cells of edge c >= 2 r_max (1 + 1e-9) from the lower corner of the
// bounding box; a pair closer than r_i + r_j differs by less than one cell
// edge on every axis, so its cells are neighbours even after the rounding
// of (x - min) / c (relative 2^-52 on quotients below 2^20).
struct OGrid {
    double mn[3];
    double c;
    uint32_t nd[3];
};

__device__ inline uint32_t ocell_axis(double v, double mn, double c, uint32_t nd) {
    const double q = floor((v - mn) / c);
    return q >= (double)(nd - 1) ? nd - 1 : (q > 0.0 ? (uint32_t)q : 0u);
}

__device__ inline uint32_t ocell(const double4& p, const OGrid& g, uint32_t* ix, uint32_t* iy, uint32_t* iz) {
    *ix = ocell_axis(p.x, g.mn[0], g.c, g.nd[0]);
    *iy = ocell_axis(p.y, g.mn[1], g.c, g.nd[1]);
    *iz = ocell_axis(p.z, g.mn[2], g.c, g.nd[2]);
    return (*iz * g.nd[1] + *iy) * g.nd[0] + *ix;
}

__global__ __launch_bounds__(ABLOCK) void k_ocount_cells(const double4* __restrict__ x, uint32_t n, OGrid g,
                                                         uint32_t* __restrict__ cell_of, uint32_t* __restrict__ ccount) {
    const uint32_t i = blockIdx.x * ABLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t ix, iy, iz;
    const uint32_t c = ocell(x[i], g, &ix, &iy, &iz);
    cell_of[i] = c;
    atomicAdd(&ccount[c], 1u);
}

// Bins the spheres by cell (any order inside a cell): sorted position and
// original index.
__global__ __launch_bounds__(ABLOCK) void k_oscatter(const double4* __restrict__ x, const uint32_t* __restrict__ id,
                                                     uint32_t n, const uint32_t* __restrict__ cell_of,
                                                     uint32_t* __restrict__ cursor, double4* __restrict__ sx,
                                                     uint32_t* __restrict__ sid) {
    const uint32_t i = blockIdx.x * ABLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t q = atomicAdd(&cursor[cell_of[i]], 1u);
    sx[q] = x[i];
    sid[q] = id[i];
}

// For the sphere at binned position p (original index i): the partners
// j > i closer than r_i + r_j.  EMIT = false counts them into cnt[i],
// EMIT = true writes (i, j) and delta_n from off[i] on, in scan order.
template <bool EMIT>
__global__ __launch_bounds__(ABLOCK) void k_opairs(const double4* __restrict__ sx, const uint32_t* __restrict__ sid,
                                                   uint32_t n, OGrid g, const uint32_t* __restrict__ cstart,
                                                   uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                   uint2* __restrict__ pairs, double* __restrict__ delta_n) {
    const uint32_t p = blockIdx.x * ABLOCK + threadIdx.x;
    if (p >= n) return;
    const double4 a = sx[p];
    const uint32_t ia = sid[p];
    uint32_t ix, iy, iz;
    (void)ocell(a, g, &ix, &iy, &iz);
    uint32_t found = 0, w = EMIT ? off[ia] : 0u;
    const uint32_t z0 = iz ? iz - 1 : 0, z1 = min(iz + 1, g.nd[2] - 1);
    const uint32_t y0 = iy ? iy - 1 : 0, y1 = min(iy + 1, g.nd[1] - 1);
    const uint32_t x0 = ix ? ix - 1 : 0, x1 = min(ix + 1, g.nd[0] - 1);
    for (uint32_t cz = z0; cz <= z1; ++cz)
        for (uint32_t cy = y0; cy <= y1; ++cy) {
            const uint32_t row = (cz * g.nd[1] + cy) * g.nd[0];
            const uint32_t q0 = cstart[row + x0], q1 = cstart[row + x1 + 1];
            for (uint32_t q = q0; q < q1; ++q) {
                const uint32_t ib = sid[q];
                if (ib <= ia) continue;
                const double4 b = sx[q];
                // x_ij = x_j - x_i, src/sphere.cpp:577-585
                const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
                const double delta = sqrt(dx * dx + dy * dy + dz * dz) - (a.w + b.w);
                if (delta < 0.0) {
                    if (EMIT) {
                        pairs[w] = make_uint2(ia, ib);
                        delta_n[w] = delta;
                        ++w;
                    }
                    ++found;
                }
            }
        }
    if (!EMIT) cnt[ia] = found;
}

// Orders each sphere's pairs by partner index (insertion sort of the
// segment off[i] .. off[i + 1]; segments hold a sphere's overlaps).
__global__ __launch_bounds__(ABLOCK) void k_osort(const uint32_t* __restrict__ off, uint32_t n, uint2* __restrict__ pairs,
                                                  double* __restrict__ delta_n) {
    const uint32_t i = blockIdx.x * ABLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = off[i], e = off[i + 1];
    for (uint32_t k = b + 1; k < e; ++k) {
        const uint2 pk = pairs[k];
        const double dk = delta_n[k];
        uint32_t m = k;
        while (m > b && pairs[m - 1].y > pk.y) {
            pairs[m] = pairs[m - 1];
            delta_n[m] = delta_n[m - 1];
            --m;
        }
        pairs[m] = pk;
        delta_n[m] = dk;
    }
}

// --------------------------------------------------------------- raytracer
// DEM::render (src/raytracer.cuh:14-553) in FP32 as there.  The reference
// tests every sphere for every pixel; here the spheres are first binned
// into 16 x 16-pixel tiles by a conservative projection of a sphere grown by
// more than the FP32 error of the hit test, and each tile's pixels test only
// the spheres binned to it.  The nearest hit is the least (|t|, original
// index), which is the sphere the reference's strict "<" over ascending
// indices keeps; the hit test, roots, normal and shading use the
// reference's expressions, so the image is the same byte for byte.
constexpr int RT_TILE = 16;

struct RF3 {
    float x, y, z;
};
__host__ __device__ inline RF3 rf3(float x, float y, float z) { return {x, y, z}; }
__host__ __device__ inline RF3 operator+(RF3 a, RF3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline RF3 operator-(RF3 a, RF3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline RF3 operator-(RF3 a) { return {-a.x, -a.y, -a.z}; }
__host__ __device__ inline RF3 operator*(float s, RF3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline RF3 operator*(RF3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline RF3 operator*(RF3 a, RF3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline RF3 operator/(RF3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline float rdot(RF3 a, RF3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline float rlength(RF3 a) { return sqrtf(rdot(a, a)); }
// helper_math's host normalize: v * (1 / sqrt(v . v))
__host__ __device__ inline RF3 rnormalize(RF3 a) { return a * (1.0f / sqrtf(rdot(a, a))); }
__host__ __device__ inline RF3 rcross(RF3 a, RF3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// cameraInit (src/raytracer.cuh:281-333) with render's eye, look-at point
// and screen width (:409-418)
struct RtCam {
    RF3 eye, u, v, w, light;
    float pl, pr, pb, pt, dfoc;
    uint32_t width, height;
};

// colorbar.h:7-22
__device__ inline float cb_red(float r) { return fminf(1.0f, 0.209f * r * r * r - 2.49f * r * r + 3.0f * r + 0.0109f); }
__device__ inline float cb_green(float r) { return fminf(1.0f, -2.44f * r * r + 2.15f * r + 0.369f); }
__device__ inline float cb_blue(float r) { return fminf(1.0f, -2.21f * r * r + 1.61f * r + 0.573f); }

// rayInitPerspective (src/raytracer.cuh:27-51)
__device__ inline RF3 ray_direction(const RtCam& c, uint32_t mempos) {
    const uint32_t i = mempos % c.width;
    const uint32_t j = (uint32_t)(int)floorf((float)mempos / c.width) % c.width;
    const float p_u = c.pl + (c.pr - c.pl) * (i + 0.5f) / c.width;
    const float p_v = c.pb + (c.pt - c.pb) * (j + 0.5f) / c.height;
    return -c.dfoc * c.w + p_u * c.u + p_v * c.v;
}

// May a sphere of this colour value be drawn (rayIntersectSpheresColormap's
// cutoff, src/raytracer.cuh:193-203)?
__device__ inline bool rt_drawn(int method, float lower_cutoff, double lin, double fixvel) {
    if (method == 0 || !(lower_cutoff > 0.0)) return true;
    return (float)lin > lower_cutoff && (float)fixvel == 0.f;
}

// Pixel-independent parts of the hit test: e - c and (e - c).(e - c) - R^2.
__device__ inline float4 rt_record(const RtCam& c, const double4& x) {
    const RF3 ctr = rf3((float)x.x, (float)x.y, (float)x.z);
    const float R = (float)x.w;
    const RF3 ec = c.eye - ctr;
    return make_float4(ec.x, ec.y, ec.z, rdot(ec, ec) - R * R);
}

// Tiles (inclusive rectangle) the sphere can be hit from, or false for none.
// Delta / 4 = (d.ec)^2 - (d.d)(ec.ec - R^2) = (d.d) R^2 - |d x ec|^2, and
// its FP32 evaluation errs by at most about 14 eps (d.d)(ec.ec) (eps =
// 2^-24; d.ec does not cancel: the ray points at the sphere), so the test
// can accept rays up to sqrt(R^2 + 8.4e-7 |q|^2) from the centre (q = c -
// eye).  The sphere is grown to R' = 1.001 sqrt(R^2 + 2e-6 |q|^2) + 1e-6
// |q| (the last term covers the rounding of e - c).  Its box [q -+ R']
// projects through the eye (p_u = -D q_u / q_w, the line of pixel p) inside
// the corner ratios when it does not straddle the eye plane; otherwise
// every tile.  One pixel of margin on each side.
__device__ inline bool rt_tiles(const RtCam& c, const double4& x, uint32_t Hr, uint32_t tiles_x, uint32_t tiles_y,
                                uint32_t* t0x, uint32_t* t1x, uint32_t* t0y, uint32_t* t1y) {
    const double qx = (double)(float)x.x - c.eye.x, qy = (double)(float)x.y - c.eye.y,
                 qz = (double)(float)x.z - c.eye.z;
    const double qn = sqrt(qx * qx + qy * qy + qz * qz);
    const double R = (double)(float)x.w;
    const double Rg = 1.001 * sqrt(R * R + 2e-6 * qn * qn) + 1e-6 * qn + 1e-30;
    const double qu = qx * c.u.x + qy * c.u.y + qz * c.u.z;
    const double qv = qx * c.v.x + qy * c.v.y + qz * c.v.z;
    const double qw = qx * c.w.x + qy * c.w.y + qz * c.w.z;
    *t0x = 0; *t1x = tiles_x - 1; *t0y = 0; *t1y = tiles_y - 1;
    if (!isfinite(qn) || !isfinite(Rg) || fabs(qw) <= Rg) return true;
    const double D = c.dfoc;
    double ulo = INFINITY, uhi = -INFINITY, vlo = INFINITY, vhi = -INFINITY;
    for (int a = -1; a <= 1; a += 2)
        for (int b = -1; b <= 1; b += 2) {
            const double den = qw + b * Rg;
            const double pu = -D * (qu + a * Rg) / den, pv = -D * (qv + a * Rg) / den;
            ulo = fmin(ulo, pu); uhi = fmax(uhi, pu);
            vlo = fmin(vlo, pv); vhi = fmax(vhi, pv);
        }
    // pixel i sits at p_u = pl + (pr - pl) (i + 0.5) / W
    const double su = ((double)c.pr - c.pl) / c.width;
    const double ilo = floor((ulo - c.pl) / su - 0.5) - 1.0, ihi = ceil((uhi - c.pl) / su - 0.5) + 1.0;
    if (ihi < 0.0 || ilo > (double)(c.width - 1)) return false;
    *t0x = (uint32_t)fmax(ilo, 0.0) / RT_TILE;
    *t1x = (uint32_t)fmin(ihi, (double)(c.width - 1)) / RT_TILE;
    const double sv = ((double)c.pt - c.pb) / c.height;
    if (sv == 0.0)  // zero-height image plane: every row is the line p_v = pb
        return !(vhi < (double)c.pb - su || vlo > (double)c.pb + su);
    const double jlo = floor((vlo - c.pb) / sv - 0.5) - 1.0, jhi = ceil((vhi - c.pb) / sv - 0.5) + 1.0;
    if (jhi < 0.0 || jlo > (double)(Hr - 1)) return false;
    *t0y = (uint32_t)fmax(jlo, 0.0) / RT_TILE;
    *t1y = (uint32_t)fmin(jhi, (double)(Hr - 1)) / RT_TILE;
    return true;
}

// Binning with one LDS histogram of the T tiles per block when it fits
// (RT_LDS_TILES), so that a block's spheres, which lie close together in the
// cell-sorted state, cost one global atomic per tile they touch rather than
// one per sphere and tile.
constexpr uint32_t RT_LDS_TILES = 16384;

// Pass 1: tiles per drawn sphere into tcount, their total into *total.
template <bool LDS>
__global__ __launch_bounds__(ABLOCK) void k_rt_count(RtCam cam, const double4* __restrict__ x,
                                                     const double4* __restrict__ vel, const uint32_t* __restrict__ id,
                                                     uint32_t n, int method, float lower_cutoff,
                                                     const double* __restrict__ lin, uint32_t Hr, uint32_t tiles_x,
                                                     uint32_t tiles_y, uint32_t* __restrict__ tcount,
                                                     unsigned long long* __restrict__ total) {
    extern __shared__ uint32_t hist[];
    __shared__ unsigned long long block_total;
    const uint32_t T = tiles_x * tiles_y;
    if (LDS) {
        for (uint32_t t = threadIdx.x; t < T; t += ABLOCK) hist[t] = 0u;
        if (threadIdx.x == 0) block_total = 0ull;
        __syncthreads();
    }
    const uint32_t s = blockIdx.x * ABLOCK + threadIdx.x;
    uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    const bool ok = s < n && rt_drawn(method, lower_cutoff, lin[id[s]], vel[s].w) &&
                    rt_tiles(cam, x[s], Hr, tiles_x, tiles_y, &x0, &x1, &y0, &y1);
    if (ok) {
        for (uint32_t ty = y0; ty <= y1; ++ty)
            for (uint32_t tx = x0; tx <= x1; ++tx) atomicAdd(LDS ? &hist[ty * tiles_x + tx] : &tcount[ty * tiles_x + tx], 1u);
        const unsigned long long area = (unsigned long long)(x1 - x0 + 1) * (y1 - y0 + 1);
        atomicAdd(LDS ? &block_total : total, area);
    }
    if (LDS) {
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < T; t += ABLOCK)
            if (hist[t]) atomicAdd(&tcount[t], hist[t]);
        if (threadIdx.x == 0 && block_total) atomicAdd(total, block_total);
    }
}

// Pass 2: the records of each tile (any order inside a tile).  With LDS the
// block reserves one range per tile it touches and hands out slots from it.
template <bool LDS>
__global__ __launch_bounds__(ABLOCK) void k_rt_fill(RtCam cam, const double4* __restrict__ x,
                                                    const double4* __restrict__ vel, const uint32_t* __restrict__ id,
                                                    uint32_t n, int method, float lower_cutoff,
                                                    const double* __restrict__ lin, uint32_t Hr, uint32_t tiles_x,
                                                    uint32_t tiles_y, uint32_t* __restrict__ cursor,
                                                    float4* __restrict__ rec, uint2* __restrict__ rec_id) {
    extern __shared__ uint32_t hist[];
    const uint32_t T = tiles_x * tiles_y;
    if (LDS) {
        for (uint32_t t = threadIdx.x; t < T; t += ABLOCK) hist[t] = 0u;
        __syncthreads();
    }
    const uint32_t s = blockIdx.x * ABLOCK + threadIdx.x;
    uint32_t x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    const bool ok = s < n && rt_drawn(method, lower_cutoff, lin[id[s]], vel[s].w) &&
                    rt_tiles(cam, x[s], Hr, tiles_x, tiles_y, &x0, &x1, &y0, &y1);
    if (LDS) {
        if (ok)
            for (uint32_t ty = y0; ty <= y1; ++ty)
                for (uint32_t tx = x0; tx <= x1; ++tx) atomicAdd(&hist[ty * tiles_x + tx], 1u);
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < T; t += ABLOCK)
            if (hist[t]) hist[t] = atomicAdd(&cursor[t], hist[t]);  // the block's first slot in tile t
        __syncthreads();
    }
    if (!ok) return;
    const uint32_t orig = id[s];
    const float4 r = rt_record(cam, x[s]);
    for (uint32_t ty = y0; ty <= y1; ++ty)
        for (uint32_t tx = x0; tx <= x1; ++tx) {
            const uint32_t t = ty * tiles_x + tx;
            const uint32_t q = atomicAdd(LDS ? &hist[t] : &cursor[t], 1u);
            rec[q] = r;
            rec_id[q] = make_uint2(orig, s);
        }
}

// Without tiles: one list of every drawn sphere, shared by all tiles.
__global__ __launch_bounds__(ABLOCK) void k_rt_all(RtCam cam, const double4* __restrict__ x,
                                                   const double4* __restrict__ vel, const uint32_t* __restrict__ id,
                                                   uint32_t n, int method, float lower_cutoff,
                                                   const double* __restrict__ lin, uint32_t* __restrict__ cursor,
                                                   float4* __restrict__ rec, uint2* __restrict__ rec_id) {
    const uint32_t s = blockIdx.x * ABLOCK + threadIdx.x;
    if (s >= n) return;
    const uint32_t orig = id[s];
    if (!rt_drawn(method, lower_cutoff, lin[orig], vel[s].w)) return;
    const uint32_t q = atomicAdd(cursor, 1u);
    rec[q] = rt_record(cam, x[s]);
    rec_id[q] = make_uint2(orig, s);
}

// One block per 16 x 16-pixel tile of rows [0, Hr): every pixel tests the
// tile's records (staged through LDS 256 at a time) and shades its nearest
// hit (rayIntersectSpheres, src/raytracer.cuh:55-137, and
// rayIntersectSpheresColormap, :140-278).
__global__ __launch_bounds__(RT_TILE* RT_TILE) void k_rt_trace(
    RtCam cam, const float4* __restrict__ rec, const uint2* __restrict__ rec_id, const uint32_t* __restrict__ tstart,
    uint32_t shared_n, uint32_t tiles_x, uint32_t Hr, int method, float max_val, const double4* __restrict__ x,
    const double4* __restrict__ vel, const double* __restrict__ lin, uint8_t* __restrict__ rgb) {
    __shared__ float4 s_rec[RT_TILE * RT_TILE];
    __shared__ uint32_t s_id[RT_TILE * RT_TILE];
    const uint32_t tile = blockIdx.x;
    const uint32_t i = (tile % tiles_x) * RT_TILE + (threadIdx.x % RT_TILE);
    const uint32_t jr = (tile / tiles_x) * RT_TILE + (threadIdx.x / RT_TILE);
    const bool active = i < cam.width && jr < Hr;
    const uint32_t mempos = jr * cam.width + i;
    const RF3 d = active ? ray_direction(cam, mempos) : rf3(0.f, 0.f, 0.f);
    const float dd = rdot(d, d);
    float tdist = 1e10f, t_best = 0.0f;
    uint32_t best_id = 0xffffffffu, best_q = 0xffffffffu;
    const uint32_t lo = shared_n ? 0u : tstart[tile], hi = shared_n ? shared_n : tstart[tile + 1];
    for (uint32_t base = lo; base < hi; base += RT_TILE * RT_TILE) {
        const uint32_t q = base + threadIdx.x;
        if (q < hi) {
            s_rec[threadIdx.x] = rec[q];
            s_id[threadIdx.x] = rec_id[q].x;
        }
        __syncthreads();
        const uint32_t m = min(hi - base, (uint32_t)(RT_TILE * RT_TILE));
        if (active) {
            for (uint32_t k = 0; k < m; ++k) {
                const float4 r = s_rec[k];
                const RF3 ec = rf3(r.x, r.y, r.z);
                const float Delta = (2.0f * rdot(d, ec)) * (2.0f * rdot(d, ec)) - 4.0f * dd * r.w;
                if (!(Delta > 0.0f)) continue;
                const float t_minus = ((rdot(-d, ec) - sqrtf(rdot(d, ec) * rdot(d, ec) - dd * r.w)) / dd);
                const float at = fabsf(t_minus);
                const uint32_t oid = s_id[k];
                if (at < tdist || (at == tdist && best_id != 0xffffffffu && oid < best_id)) {
                    tdist = at;
                    t_best = t_minus;
                    best_id = oid;
                    best_q = base + k;
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    uint8_t* px = rgb + 3 * (size_t)mempos;
    if (best_id == 0xffffffffu) {  // imageInit's background (:14-24)
        px[0] = px[1] = px[2] = 255;
        return;
    }
    const uint32_t s = rec_id[best_q].y;
    const double4 xs = x[s];
    const RF3 c = rf3((float)xs.x, (float)xs.y, (float)xs.z);
    const RF3 p = cam.eye + t_best * d;
    const RF3 n = rnormalize(2.0f * (p - c));
    const float dotprod = fmaxf(0.0f, rdot(n, cam.light));
    const float I_d = 40.0f, k_d = 5.0f, k_a = 10.0f, I_a = 5.0f;
    if (method == 0) {
        px[0] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.48f);
        px[1] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.41f);
        px[2] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * 0.27f);
        return;
    }
    const float fixvel = (float)vel[s].w;
    float ratio = (float)(lin[best_id] / max_val);
    if (ratio < 0.01f) ratio = 0.01f;
    if (ratio > 0.99f) ratio = 0.99f;
    float redv = cb_red(ratio), greenv = cb_green(ratio), bluev = cb_blue(ratio);
    if (fixvel > 0.f) redv = greenv = bluev = 0.5f;
    px[0] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * redv);
    px[1] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * greenv);
    px[2] = (uint8_t)((k_d * I_d * dotprod + k_a * I_a) * bluev);
}

// Rows r >= width repeat row r mod width (the reference's row index is
// (mempos / width) mod width).
__global__ __launch_bounds__(ABLOCK) void k_rt_repeat_rows(uint8_t* __restrict__ rgb, uint32_t width, uint32_t height) {
    const size_t row_bytes = 3 * (size_t)width;
    const size_t total = row_bytes * height;
    for (size_t b = row_bytes * width + blockIdx.x * (size_t)ABLOCK + threadIdx.x; b < total;
         b += (size_t)gridDim.x * ABLOCK) {
        const size_t r = b / row_bytes, off = b % row_bytes;
        rgb[b] = rgb[(r % width) * row_bytes + off];
    }
}

int check_handle(sph_dem* h, const char* fn) {
    if (!h) return SPH_E_INVALID;
    if (h->d.nsys > 1 || !h->members.empty()) {
        h->err = std::string(fn) + ": not available on an ensemble";
        return SPH_E_STATE;
    }
    return SPH_OK;
}

}  // namespace

namespace {

// DEM::porosity with its device work space: `ws` (ws_cap doubles) or, when
// ws is null, a buffer allocated for the call.  Layout: the bounds partials
// (BOUNDS_WS), the per-slice block sums (slices x particle blocks), z_pos and
// porosity (slices each).
int porosity_run(sph_dem* h, int slices, int rule, double* z_pos, double* porosity, double* ws, size_t ws_cap,
                 const char* fn) {
    int rc;
    if ((rc = check_handle(h, fn))) return rc;
    if (slices < 1 || !z_pos || !porosity || (rule != SPH_POROSITY_REFERENCE && rule != SPH_POROSITY_EXACT)) {
        h->err = std::string(fn) + ": need slices >= 1, output arrays and rule 0 or 1";
        return SPH_E_INVALID;
    }
    const sph::HostState& hs = h->hs;
    const sph::PState& st = cur(h);
    const uint32_t n = h->d.np;
    const uint32_t nblk = (n + ABLOCK * PPT - 1) / (ABLOCK * PPT);
    const size_t parts = (size_t)slices * std::max<uint32_t>(nblk, 1);
    if (parts > ((size_t)1 << 28)) {
        h->err = std::string(fn) + ": slices x particle blocks above 2^28";
        return SPH_E_INVALID;
    }
    const size_t need = BOUNDS_WS + parts + 2 * (size_t)slices;
    Scratch s;
    if (!ws) {
        A_TRY(s.get(&ws, need));
    } else if (need > ws_cap) {
        h->err = std::string(fn) + ": " + std::to_string(slices) + " slices of " + std::to_string(n) +
                 " spheres need " + std::to_string(need) + " doubles of work space, above its maximum of " +
                 std::to_string(ws_cap);
        return SPH_E_INVALID;
    }
    double* bpart = ws;
    double* rl = bpart + 7 * (size_t)BOUNDS_BLOCKS;
    double* part = ws + BOUNDS_WS;
    double* dz = part + parts;
    double* dp = dz + slices;
    double bnd[7], r_last = 0.0;
    if (n && (rc = bounds_in(h, st, n, bpart, rl, bnd, &r_last))) return rc;
    double top = hs.L[2];
    if (hs.nw > 0) {  // wall 0's current position (src/sphere.cpp:367-371)
        double4 w0;
        A_TRY(hipMemcpyAsync(&w0, h->d.wnx, sizeof w0, hipMemcpyDeviceToHost, h->stream));
        A_TRY(hipStreamSynchronize(h->stream));
        top = w0.w;
    }
    const double h_slice = (top - hs.origo[2]) / (double)slices;
    if (rule == SPH_POROSITY_REFERENCE && h_slice <= r_last * 2.0) {
        h->err = "Error! The number of z-slices is too high.";
        return SPH_E_INVALID;
    }
    if (!(h_slice > 0.0) || !std::isfinite(h_slice)) {
        h->err = std::string(fn) + ": the slices have no positive thickness";
        return SPH_E_INVALID;
    }
    const double V_slice = h_slice * hs.L[0] * hs.L[1];
    k_zero_f64<<<fill_blocks(parts), ABLOCK, 0, h->stream>>>(part, parts);
    if (nblk) k_porosity<<<nblk, ABLOCK, 0, h->stream>>>(st.x, n, slices, h_slice, rule, part);
    k_porosity_final<<<(uint32_t)slices, 64, 0, h->stream>>>(part, nblk, slices, h_slice, V_slice, dz, dp);
    A_TRY(hipGetLastError());
    A_TRY(hipMemcpyAsync(z_pos, dz, (size_t)slices * 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipMemcpyAsync(porosity, dp, (size_t)slices * 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipStreamSynchronize(h->stream));
    return SPH_OK;
}

// Work space of sph_porosity_noalloc: one array per device, held from the
// library's load until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy (on the host the name stands for a host-side
// shadow, not device memory): it uses the device address that
// hipGetSymbolAddress gives for the current device, looked up once per
// device at the entry point's first call there.
constexpr size_t SCRATCH_DOUBLES = SPH_POROSITY_NOALLOC_MAX;  // 8 MiB
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[MAX_DEVICES] = {};

int scratch_of(sph_dem* h, double** out) {
    const int dev = h->device;
    if (dev < 0 || dev >= MAX_DEVICES) {
        h->err = "sph_porosity_noalloc: device index out of range";
        return SPH_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        int cur_dev = -1;
        A_TRY(hipGetDevice(&cur_dev));
        if (cur_dev != dev) A_TRY(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur_dev != dev) A_TRY(hipSetDevice(cur_dev));
        if (e != hipSuccess || !p) {
            h->err = std::string("sph_porosity_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SPH_E_HIP;
        }
        g_scratch_addr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_addr[dev];
    return SPH_OK;
}

}  // namespace

extern "C" int sph_porosity(sph_t* h, int slices, int rule, double* z_pos, double* porosity) {
    return porosity_run(h, slices, rule, z_pos, porosity, nullptr, 0, "sph_porosity");
}

extern "C" int sph_porosity_noalloc(sph_t* h, int slices, int rule, double* z_pos, double* porosity) {
    if (!h) return SPH_E_INVALID;
    double* ws = nullptr;
    int rc;
    if ((rc = scratch_of(h, &ws))) return rc;
    return porosity_run(h, slices, rule, z_pos, porosity, ws, SCRATCH_DOUBLES, "sph_porosity_noalloc");
}

extern "C" int sph_find_overlaps(sph_t* h, uint32_t* ij, double* delta_n, uint64_t capacity, uint64_t* count) {
    int rc;
    if ((rc = check_handle(h, "sph_find_overlaps"))) return rc;
    if (!count) {
        h->err = "sph_find_overlaps: count is NULL";
        return SPH_E_INVALID;
    }
    *count = 0;
    const sph::PState& st = cur(h);
    const uint32_t n = h->d.np;
    if (n < 2) return SPH_OK;
    Scratch s;
    double bnd[7], r_last;
    if ((rc = bounds(h, st, n, s, bnd, &r_last))) return rc;
    if (std::isnan(bnd[6]) || !std::isfinite(bnd[0] + bnd[1] + bnd[2] + bnd[3] + bnd[4] + bnd[5])) {
        h->err = "sph_find_overlaps: non-finite position or radius";
        return SPH_E_NAN;
    }
    if (!(bnd[6] > 0.0)) return SPH_OK;  // no radius, no overlap
    OGrid g;
    g.c = 2.0 * bnd[6] * (1.0 + 1e-9);
    for (;;) {
        uint64_t K = 1;
        bool fits = true;
        for (int a = 0; a < 3; ++a) {
            g.mn[a] = bnd[a];
            const double q = std::floor((bnd[3 + a] - bnd[a]) / g.c) + 1.0;
            if (q > (double)(1u << 20)) fits = false;
            g.nd[a] = fits ? (uint32_t)q : 1u;
            K *= g.nd[a];
        }
        if (fits && K <= (1u << 22)) break;
        g.c *= 1.5;  // coarser cells: more candidates per cell, same pairs
    }
    const size_t K = (size_t)g.nd[0] * g.nd[1] * g.nd[2];
    uint32_t *cell_of, *ccount, *cstart, *cursor, *sid, *cnt, *off;
    double4* sx;
    A_TRY(s.get(&cell_of, n));
    A_TRY(s.get(&ccount, K + 1));
    A_TRY(s.get(&cstart, K + 1));
    A_TRY(s.get(&cursor, K + 1));
    A_TRY(s.get(&sx, n));
    A_TRY(s.get(&sid, n));
    A_TRY(s.get(&cnt, (size_t)n + 1));
    A_TRY(s.get(&off, (size_t)n + 1));
    const uint32_t nb = (n + ABLOCK - 1) / ABLOCK;
    k_zero_u32<<<fill_blocks(K + 1), ABLOCK, 0, h->stream>>>(ccount, K + 1);
    k_ocount_cells<<<nb, ABLOCK, 0, h->stream>>>(st.x, n, g, cell_of, ccount);
    if ((rc = exclusive_scan(h, ccount, cstart, K + 1, s))) return rc;
    A_TRY(hipMemcpyAsync(cursor, cstart, (K + 1) * 4, hipMemcpyDeviceToDevice, h->stream));
    k_oscatter<<<nb, ABLOCK, 0, h->stream>>>(st.x, st.id, n, cell_of, cursor, sx, sid);
    k_zero_u32<<<1, ABLOCK, 0, h->stream>>>(cnt + n, 1);
    k_opairs<false><<<nb, ABLOCK, 0, h->stream>>>(sx, sid, n, g, cstart, cnt, nullptr, nullptr, nullptr);
    if ((rc = exclusive_scan(h, cnt, off, (size_t)n + 1, s))) return rc;
    A_TRY(hipGetLastError());
    uint32_t total = 0;
    A_TRY(hipMemcpyAsync(&total, off + n, 4, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipStreamSynchronize(h->stream));
    *count = total;
    if (!ij || !delta_n || total == 0) return SPH_OK;
    if (capacity < total) {
        h->err = "sph_find_overlaps: " + std::to_string(total) + " pairs, capacity " + std::to_string(capacity);
        return SPH_E_INVALID;
    }
    uint2* pairs;
    double* dn;
    A_TRY(s.get(&pairs, total));
    A_TRY(s.get(&dn, total));
    k_opairs<true><<<nb, ABLOCK, 0, h->stream>>>(sx, sid, n, g, cstart, nullptr, off, pairs, dn);
    k_osort<<<nb, ABLOCK, 0, h->stream>>>(off, n, pairs, dn);
    A_TRY(hipGetLastError());
    A_TRY(hipMemcpyAsync(ij, pairs, (size_t)total * 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipMemcpyAsync(delta_n, dn, (size_t)total * 8, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipStreamSynchronize(h->stream));
    return SPH_OK;
}

extern "C" int sph_render(sph_t* h, int method, float max_val, float lower_cutoff, uint32_t width, uint32_t height,
                          int cull, uint8_t* rgb) {
    int rc;
    if ((rc = check_handle(h, "sph_render"))) return rc;
    if (method < 0 || method > 5 || width == 0 || height == 0 || !rgb ||
        (uint64_t)width * height > (1ull << 24)) {
        h->err = "sph_render: need method 0..5, an output buffer and 1 <= width x height <= 2^24 pixels";
        return SPH_E_INVALID;
    }
    const sph::HostState& hs = h->hs;
    const sph::PState& st = cur(h);
    const uint32_t n = h->d.np;
    Scratch s;
    double bnd[7] = {0, 0, 0, 0, 0, 0, 0}, r_last;
    if (n && (rc = bounds(h, st, n, s, bnd, &r_last))) return rc;
    if (std::isnan(bnd[6])) {
        h->err = "sph_render: non-finite position or radius";
        return SPH_E_NAN;
    }
    // camera: render (src/raytracer.cuh:409-418) and cameraInit (:281-313)
    RtCam cam;
    cam.width = width;
    cam.height = height;
    const float imgw = (float)(hs.L[0] * 1.35f);
    const double mx[3] = {std::max(0.0, bnd[3]), std::max(0.0, bnd[4]), std::max(0.0, bnd[5])};  // DEM::maxPos
    const RF3 lookat = rf3((float)mx[0], (float)mx[1], (float)mx[2]) / 2.0f;
    cam.eye = rf3((float)(hs.L[0] * 2.3f), (float)(hs.L[1] * -5.0f), (float)(hs.L[2] * 1.3f));
    const float hw_ratio = (float)(height / width);
    cam.pl = -0.5f * imgw;
    cam.pr = 0.5f * imgw;
    cam.pb = -0.5f * imgw * hw_ratio;
    cam.pt = 0.5f * imgw * hw_ratio;
    const RF3 view = cam.eye - lookat;
    const RF3 up = rf3(0.0f, 0.0f, 1.0f);
    cam.w = -view / rlength(view);
    cam.u = rcross(up, cam.w) / rlength(rcross(up, cam.w));
    cam.v = rcross(cam.w, cam.u);
    cam.dfoc = rlength(view) * 0.8f;
    cam.light = rnormalize(-1.0f * cam.eye * rf3(1.0f, 0.2f, 0.6f));
    // colour values per original index (render, :445-503), in FP64
    std::vector<double> lin(std::max<uint32_t>(n, 1), 0.0);
    if (method != 0 && n) {
        const int field = method == 1 ? SPH_P : method == 2 ? SPH_VEL : method == 3 ? SPH_ANGVEL
                        : method == 4 ? SPH_XYSUM : SPH_ANGPOS;
        const size_t comps = method == 1 ? 1 : method == 2 ? 4 : method == 4 ? 2 : 3;
        std::vector<double> f(comps * n);
        if ((rc = sph_get(h, field, f.data(), f.size() * 8))) return rc;
        for (uint32_t i = 0; i < n; ++i) {
            const double* v = &f[comps * i];
            lin[i] = (method == 1 || method == 4) ? v[0] : std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
    }
    double* dlin;
    A_TRY(s.get(&dlin, lin.size()));
    A_TRY(hipMemcpyAsync(dlin, lin.data(), lin.size() * 8, hipMemcpyHostToDevice, h->stream));
    const uint32_t Hr = std::min(width, height);
    const uint32_t tiles_x = (width + RT_TILE - 1) / RT_TILE, tiles_y = (Hr + RT_TILE - 1) / RT_TILE;
    const size_t T = (size_t)tiles_x * tiles_y;
    const uint32_t nb = std::max<uint32_t>(1, (n + ABLOCK - 1) / ABLOCK);
    uint32_t *tcount, *tstart, *cursor;
    unsigned long long* dtotal;
    A_TRY(s.get(&tcount, T + 1));
    A_TRY(s.get(&tstart, T + 1));
    A_TRY(s.get(&cursor, T + 1));
    A_TRY(s.get(&dtotal, 1));
    k_zero_u32<<<fill_blocks(T + 1), ABLOCK, 0, h->stream>>>(tcount, T + 1);
    k_zero_u32<<<1, ABLOCK, 0, h->stream>>>(reinterpret_cast<uint32_t*>(dtotal), 2);
    unsigned long long total = 0;
    // records of one tiling: 24 bytes each; beyond 2^26 of them (every
    // sphere on many tiles) the single shared list is cheaper
    const unsigned long long max_entries = 1ull << 26;
    if (cull && n) {
        if (T <= RT_LDS_TILES)
            k_rt_count<true><<<nb, ABLOCK, T * 4, h->stream>>>(cam, st.x, st.vel, st.id, n, method, lower_cutoff, dlin,
                                                               Hr, tiles_x, (uint32_t)tiles_y, tcount, dtotal);
        else
            k_rt_count<false><<<nb, ABLOCK, 0, h->stream>>>(cam, st.x, st.vel, st.id, n, method, lower_cutoff, dlin, Hr,
                                                            tiles_x, (uint32_t)tiles_y, tcount, dtotal);
        A_TRY(hipGetLastError());
        A_TRY(hipMemcpyAsync(&total, dtotal, 8, hipMemcpyDeviceToHost, h->stream));
        A_TRY(hipStreamSynchronize(h->stream));
    }
    const bool tiled = cull && n && total <= max_entries;
    float4* rec;
    uint2* rec_id;
    uint32_t shared_n = 0;
    if (tiled) {
        if ((rc = exclusive_scan(h, tcount, tstart, T + 1, s))) return rc;
        A_TRY(s.get(&rec, total));
        A_TRY(s.get(&rec_id, total));
        A_TRY(hipMemcpyAsync(cursor, tstart, (T + 1) * 4, hipMemcpyDeviceToDevice, h->stream));
        if (T <= RT_LDS_TILES)
            k_rt_fill<true><<<nb, ABLOCK, T * 4, h->stream>>>(cam, st.x, st.vel, st.id, n, method, lower_cutoff, dlin,
                                                              Hr, tiles_x, (uint32_t)tiles_y, cursor, rec, rec_id);
        else
            k_rt_fill<false><<<nb, ABLOCK, 0, h->stream>>>(cam, st.x, st.vel, st.id, n, method, lower_cutoff, dlin, Hr,
                                                           tiles_x, (uint32_t)tiles_y, cursor, rec, rec_id);
    } else {
        A_TRY(s.get(&rec, std::max<uint32_t>(n, 1)));
        A_TRY(s.get(&rec_id, std::max<uint32_t>(n, 1)));
        k_zero_u32<<<1, ABLOCK, 0, h->stream>>>(cursor, 1);
        if (n)
            k_rt_all<<<nb, ABLOCK, 0, h->stream>>>(cam, st.x, st.vel, st.id, n, method, lower_cutoff, dlin, cursor, rec,
                                                   rec_id);
        A_TRY(hipGetLastError());
        A_TRY(hipMemcpyAsync(&shared_n, cursor, 4, hipMemcpyDeviceToHost, h->stream));
        A_TRY(hipStreamSynchronize(h->stream));
    }
    uint8_t* img;
    const size_t bytes = 3 * (size_t)width * height;
    A_TRY(s.get(&img, bytes));
    if (tiled || shared_n)
        k_rt_trace<<<(uint32_t)T, RT_TILE * RT_TILE, 0, h->stream>>>(cam, rec, rec_id, tstart, shared_n, tiles_x, Hr,
                                                                    method, max_val, st.x, st.vel, dlin, img);
    else  // nothing drawn: imageInit's background only
        k_fill_u8<<<fill_blocks(bytes), ABLOCK, 0, h->stream>>>(img, bytes, 255);
    if (height > width) k_rt_repeat_rows<<<fill_blocks(bytes), ABLOCK, 0, h->stream>>>(img, width, height);
    A_TRY(hipGetLastError());
    A_TRY(hipMemcpyAsync(rgb, img, bytes, hipMemcpyDeviceToHost, h->stream));
    A_TRY(hipStreamSynchronize(h->stream));
    return SPH_OK;
}
