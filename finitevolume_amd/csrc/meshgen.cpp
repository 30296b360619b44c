// Host-side mesh generation helper: Bridson's Poisson-disk sampling
// (R. Bridson, "Fast Poisson disk sampling in arbitrary dimensions",
// SIGGRAPH sketches 2007) in a rectangle, around an excluded NACA 4-digit
// symmetric section.  Deterministic for a seed (splitmix64 stream).  Used to
// build the Voronoi meshes of the batched airfoil cases; not part of the
// residual path.
#include <cmath>
#include <cstdint>
#include <vector>

#include "fvk.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// half thickness of a NACA 4-digit symmetric section, chord 1, leading edge
// at the origin, closed trailing edge (coefficient -0.1036)
double naca_half_thickness(double x, double t) {
    if (x <= 0.0 || x >= 1.0) return 0.0;
    return 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
}

}  // namespace

extern "C" int fvk_poisson_disk(double x0, double y0, double x1, double y1, double radius, uint64_t seed,
                                const double* seed_xy, int64_t n_seed, double thickness, double margin,
                                double* out_xy, int64_t cap, int64_t* n_out) {
    if (!(radius > 0.0) || !(x1 > x0) || !(y1 > y0) || !out_xy || !n_out || cap <= 0) return FVK_ERR_INVALID;
    const double cell = radius / std::sqrt(2.0);
    const int64_t gw = (int64_t)std::ceil((x1 - x0) / cell), gh = (int64_t)std::ceil((y1 - y0) / cell);
    if (gw <= 0 || gh <= 0 || gw * gh > (int64_t)1 << 31) return FVK_ERR_INVALID;
    std::vector<int64_t> grid((size_t)(gw * gh), -1);
    std::vector<int64_t> active;
    int64_t n = 0;
    const double r2 = radius * radius;
    auto gidx = [&](double x, double y, int64_t& gx, int64_t& gy) {
        gx = std::min(gw - 1, std::max<int64_t>(0, (int64_t)((x - x0) / cell)));
        gy = std::min(gh - 1, std::max<int64_t>(0, (int64_t)((y - y0) / cell)));
    };
    auto excluded = [&](double x, double y) {
        if (x < x0 || x >= x1 || y < y0 || y >= y1) return true;
        if (x < -margin || x > 1.0 + margin) return false;
        const double xc = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
        return std::fabs(y) <= naca_half_thickness(xc, thickness) + margin;
    };
    auto far_enough = [&](double x, double y) {
        int64_t gx, gy;
        gidx(x, y, gx, gy);
        for (int64_t j = std::max<int64_t>(0, gy - 2); j <= std::min(gh - 1, gy + 2); ++j)
            for (int64_t i = std::max<int64_t>(0, gx - 2); i <= std::min(gw - 1, gx + 2); ++i) {
                const int64_t p = grid[(size_t)(j * gw + i)];
                if (p < 0) continue;
                const double dx = out_xy[2 * p] - x, dy = out_xy[2 * p + 1] - y;
                if (dx * dx + dy * dy < r2) return false;
            }
        return true;
    };
    auto insert = [&](double x, double y) -> bool {
        if (n >= cap) return false;
        out_xy[2 * n] = x;
        out_xy[2 * n + 1] = y;
        int64_t gx, gy;
        gidx(x, y, gx, gy);
        grid[(size_t)(gy * gw + gx)] = n;
        active.push_back(n);
        ++n;
        return true;
    };
    // caller seeds (points along the excluded body) first, unconditionally
    for (int64_t k = 0; k < n_seed; ++k)
        if (!insert(seed_xy[2 * k], seed_xy[2 * k + 1])) return FVK_ERR_NOMEM;
    Rng rng{seed};
    if (n == 0) {
        double x, y;
        do {
            x = x0 + (x1 - x0) * rng.uni();
            y = y0 + (y1 - y0) * rng.uni();
        } while (excluded(x, y));
        insert(x, y);
    }
    const int kTries = 30;
    while (!active.empty()) {
        const size_t pick = (size_t)(rng.next() % active.size());
        const int64_t p = active[pick];
        const double px = out_xy[2 * p], py = out_xy[2 * p + 1];
        bool placed = false;
        for (int t = 0; t < kTries; ++t) {
            const double ang = 2.0 * M_PI * rng.uni();
            const double rr = radius * (1.0 + rng.uni());
            const double x = px + rr * std::cos(ang), y = py + rr * std::sin(ang);
            if (excluded(x, y) || !far_enough(x, y)) continue;
            if (!insert(x, y)) return FVK_ERR_NOMEM;
            placed = true;
            break;
        }
        if (!placed) {
            active[pick] = active.back();
            active.pop_back();
        }
    }
    *n_out = n;
    return FVK_OK;
}
