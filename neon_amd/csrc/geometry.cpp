// geometry.cpp -- seeded solid geometries for the benchmark workloads.
//
// Random numbers: SplitMix64 (G. Steele, D. Lea, C. Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014), uniform doubles from the top
// 53 bits, so every host produces the same geometry for a seed.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "neon_amd.h"

namespace nx {
int set_error(int code, const std::string& msg);
}

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" int nx_geometry_spheres(const int32_t dim[3], uint32_t n_spheres, double rmin,
                                   double rmax, uint64_t seed, int32_t periodic, uint8_t* solid)
{
    if (!dim || !solid) return nx::set_error(NX_ERR_INVALID, "null pointer");
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1 || rmin < 0 || rmax < rmin)
        return nx::set_error(NX_ERR_INVALID, "nx_geometry_spheres: bad dimensions or radii");
    // draw order per sphere: cx, cy, cz, r
    std::vector<double> c(4 * (size_t)n_spheres);
    SplitMix64 rng{seed};
    for (uint32_t k = 0; k < n_spheres; k++) {
        c[4 * k + 0] = rng.uniform() * dim[0];
        c[4 * k + 1] = rng.uniform() * dim[1];
        c[4 * k + 2] = rng.uniform() * dim[2];
        c[4 * k + 3] = rmin + (rmax - rmin) * rng.uniform();
    }
    const size_t X = (size_t)dim[0], XY = (size_t)dim[0] * dim[1];
    std::fill(solid, solid + XY * dim[2], (uint8_t)0);
    // voxel (x,y,z) is solid when |(x,y,z) - c|^2 < r^2 for some sphere; a
    // periodic axis measures the distance across the wrap
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int z = (int)t; z < dim[2]; z += (int)nt)
                for (uint32_t k = 0; k < n_spheres; k++) {
                    const double cx = c[4 * k], cy = c[4 * k + 1], cz = c[4 * k + 2],
                                 r = c[4 * k + 3];
                    double dz = z - cz;
                    if (periodic & NX_PERIODIC_Z) dz -= dim[2] * std::nearbyint(dz / dim[2]);
                    const double r2 = r * r - dz * dz;
                    if (r2 <= 0) continue;
                    const double ry = std::sqrt(r2);
                    for (int y = (int)std::floor(cy - ry); y <= (int)std::ceil(cy + ry); y++) {
                        int yy = y;
                        if (periodic & NX_PERIODIC_Y) yy = ((y % dim[1]) + dim[1]) % dim[1];
                        if (yy < 0 || yy >= dim[1]) continue;
                        const double dy = y - cy;
                        const double rx2 = r2 - dy * dy;
                        if (rx2 <= 0) continue;
                        const double rx = std::sqrt(rx2);
                        for (int x = (int)std::floor(cx - rx); x <= (int)std::ceil(cx + rx); x++) {
                            const double dx = x - cx;
                            if (dx * dx + dy * dy + dz * dz >= r * r) continue;
                            int xx = x;
                            if (periodic & NX_PERIODIC_X) xx = ((x % dim[0]) + dim[0]) % dim[0];
                            if (xx < 0 || xx >= dim[0]) continue;
                            solid[(size_t)xx + X * yy + XY * z] = 1;
                        }
                    }
                }
        });
    for (auto& x : th) x.join();
    return NX_OK;
}

// Lid-driven cavity classes (RunCavityTwoPop.cu:208-224: outer layer
// bounce-back, y = Y-1 moving wall) plus n_boxes axis-aligned bounce-back
// boxes inside the fluid: per box, edges e_a uniform in {emin..emax} and a
// lower corner uniform in {1 .. dim_a-1-e_a} (draw order ex, ey, ez, x0, y0,
// z0), clipped to the interior.
extern "C" int nx_geometry_cavity(const int32_t dim[3], uint32_t n_boxes, int32_t emin,
                                  int32_t emax, uint64_t seed, uint8_t* cls)
{
    if (!dim || !cls) return nx::set_error(NX_ERR_INVALID, "null pointer");
    if (dim[0] < 3 || dim[1] < 3 || dim[2] < 3 || emin < 1 || emax < emin)
        return nx::set_error(NX_ERR_INVALID, "nx_geometry_cavity: bad dimensions or edges");
    const int X = dim[0], Y = dim[1], Z = dim[2];
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                uint8_t c = NX_CELL_BULK;
                if (x == 0 || x == X - 1 || y == 0 || y == Y - 1 || z == 0 || z == Z - 1)
                    c = y == Y - 1 ? NX_CELL_MOVINGWALL : NX_CELL_BOUNCEBACK;
                cls[(size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)] = c;
            }
    SplitMix64 rng{seed};
    auto pick = [&](int lo, int hi) {  // uniform integer in [lo, hi]
        return lo + (int)(rng.uniform() * (double)(hi - lo + 1));
    };
    for (uint32_t k = 0; k < n_boxes; k++) {
        int e[3], o[3];
        for (int a = 0; a < 3; a++) e[a] = pick(emin, emax);
        for (int a = 0; a < 3; a++) o[a] = pick(1, std::max(1, dim[a] - 1 - e[a]));
        for (int z = o[2]; z < std::min(o[2] + e[2], Z - 1); z++)
            for (int y = o[1]; y < std::min(o[1] + e[1], Y - 1); y++)
                for (int x = o[0]; x < std::min(o[0] + e[0], X - 1); x++)
                    cls[(size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)] = NX_CELL_BOUNCEBACK;
    }
    return NX_OK;
}
