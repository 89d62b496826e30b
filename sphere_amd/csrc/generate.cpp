// generate.cpp -- seeded non-overlapping random packing (setup helper).
//
// Random sequential addition: spheres are inserted largest first at
// uniformly drawn centres inside the box (walls at 0 and L on every axis;
// periodic axes use the minimum image), rejected on overlap with any earlier
// sphere.  Neighbour checks use a uniform grid of cell >= 2 r_max.  The
// random stream is splitmix64, so the packing depends only on the seed.
// This replaces the reference's unseeded initRandomPos/initRandomGridPos
// (python/sphere.py:856-908, :1058-1115) for large inputs.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" {

// radii[n] in, xyz[3n] out (original order of radii).  periodic: 0 none,
// 1 x+y, 2 x.  Returns 0, or -1 when a sphere found no free spot within
// max_tries draws.
int sph_pack_rsa(uint64_t n, const double* radii, const double* box, int periodic, uint64_t seed,
                 uint64_t max_tries, double* xyz) {
    if (n == 0) return 0;
    double rmax = 0.0;
    for (uint64_t i = 0; i < n; ++i) rmax = std::max(rmax, radii[i]);
    const double cell = 2.0 * rmax;
    int nc[3];
    for (int a = 0; a < 3; ++a) nc[a] = std::max(1, (int)std::floor(box[a] / cell));
    const double cl[3] = {box[0] / nc[0], box[1] / nc[1], box[2] / nc[2]};
    const bool per[3] = {periodic == 1 || periodic == 2, periodic == 1, false};
    std::vector<std::vector<uint32_t>> grid((size_t)nc[0] * nc[1] * nc[2]);
    std::vector<uint64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return radii[a] > radii[b]; });
    SplitMix64 rng{seed};
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t i = order[k];
        const double r = radii[i];
        bool placed = false;
        for (uint64_t t = 0; t < max_tries && !placed; ++t) {
            double p[3];
            for (int a = 0; a < 3; ++a) {
                const double lo = per[a] ? 0.0 : r, hi = per[a] ? box[a] : box[a] - r;
                p[a] = lo + (hi - lo) * rng.uniform();
            }
            int c[3];
            for (int a = 0; a < 3; ++a) c[a] = std::min(nc[a] - 1, std::max(0, (int)std::floor(p[a] / cl[a])));
            bool ok = true;
            for (int dz = -1; dz <= 1 && ok; ++dz)
                for (int dy = -1; dy <= 1 && ok; ++dy)
                    for (int dx = -1; dx <= 1 && ok; ++dx) {
                        int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                        bool skip = false;
                        for (int a = 0; a < 3; ++a) {
                            if (q[a] < 0 || q[a] >= nc[a]) {
                                if (!per[a]) { skip = true; break; }
                                q[a] = (q[a] + nc[a]) % nc[a];
                            }
                        }
                        if (skip) continue;
                        const auto& bucket = grid[((size_t)q[2] * nc[1] + q[1]) * nc[0] + q[0]];
                        for (uint32_t j : bucket) {
                            double d2 = 0.0;
                            for (int a = 0; a < 3; ++a) {
                                double dd = p[a] - xyz[3 * (size_t)j + a];
                                if (per[a]) {
                                    if (dd > 0.5 * box[a]) dd -= box[a];
                                    if (dd < -0.5 * box[a]) dd += box[a];
                                }
                                d2 += dd * dd;
                            }
                            const double rs = r + radii[j];
                            if (d2 < rs * rs) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            for (int a = 0; a < 3; ++a) xyz[3 * i + a] = p[a];
            grid[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back((uint32_t)i);
            placed = true;
        }
        if (!placed) return -1;
    }
    return 0;
}

}  // extern "C"
