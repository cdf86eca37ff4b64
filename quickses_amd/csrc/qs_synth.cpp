/* Seeded synthetic molecules: atoms uniform in a ball or a spherical shell at a given number
 * density, with a minimum-distance rejection and a categorical radius mix, and jittered
 * trajectory frames.  Portable: splitmix64 and explicit double arithmetic, so every platform
 * produces the same atoms for the same seed.
 *
 * Radius mix (cumulative thresholds): C 1.70 (0.55), O 1.52 (0.75), N 1.55 (0.90),
 * H 1.20 (0.99), S 1.80 (1.00).  The element radii are those of CudaSurf.h:64-73.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

float draw_radius(SplitMix64 &rng) {
    double u = rng.uniform();
    if (u < 0.55) return 1.70f;
    if (u < 0.75) return 1.52f;
    if (u < 0.90) return 1.55f;
    if (u < 0.99) return 1.20f;
    return 1.80f;
}

/* min-distance rejection on a dense grid of `cell`-sized cells covering [-extent, extent]^3
   (per-cell linked lists of accepted points) */
struct RejectGrid {
    double cell, lo;
    int64_t dim;
    std::vector<int> head, next;
    std::vector<double> pts;
    RejectGrid(double c, double extent) : cell(c), lo(-extent - c)
    {
        dim = (int64_t)std::ceil((2.0 * extent + 2.0 * c) / c) + 2;
        head.assign((size_t)(dim * dim * dim), -1);
    }
    int64_t idx(int64_t x, int64_t y, int64_t z) const { return (x * dim + y) * dim + z; }
    int64_t coord(double v) const { return (int64_t)std::floor((v - lo) / cell); }
    bool ok(double x, double y, double z, double dmin) const {
        int64_t cx = coord(x), cy = coord(y), cz = coord(z);
        double d2 = dmin * dmin;
        for (int a = -1; a <= 1; a++)
            for (int b = -1; b <= 1; b++)
                for (int c = -1; c <= 1; c++)
                    for (int id = head[(size_t)idx(cx + a, cy + b, cz + c)]; id >= 0; id = next[id]) {
                        double ex = pts[3 * id] - x, ey = pts[3 * id + 1] - y, ez = pts[3 * id + 2] - z;
                        if (ex * ex + ey * ey + ez * ez < d2) return false;
                    }
        return true;
    }
    void add(double x, double y, double z) {
        int id = (int)(pts.size() / 3);
        pts.push_back(x); pts.push_back(y); pts.push_back(z);
        size_t k = (size_t)idx(coord(x), coord(y), coord(z));
        next.push_back(head[k]);
        head[k] = id;
    }
};

int synth(uint64_t seed, int n, double r_in, double r_out, double dmin, float *pos, float *rad) {
    SplitMix64 rng(seed);
    RejectGrid grid(dmin > 0 ? dmin : 1.0, dmin > 0 ? r_out : 0.0);
    int have = 0;
    const double ri2 = r_in * r_in, ro2 = r_out * r_out;
    uint64_t tries = 0, max_tries = (uint64_t)n * 1000ULL + 100000ULL;
    while (have < n) {
        if (++tries > max_tries) return -1;
        double x = (2.0 * rng.uniform() - 1.0) * r_out;
        double y = (2.0 * rng.uniform() - 1.0) * r_out;
        double z = (2.0 * rng.uniform() - 1.0) * r_out;
        double d2 = x * x + y * y + z * z;
        if (d2 > ro2 || d2 < ri2) continue;
        if (dmin > 0 && !grid.ok(x, y, z, dmin)) continue;
        if (dmin > 0) grid.add(x, y, z);
        pos[3 * have + 0] = (float)x;
        pos[3 * have + 1] = (float)y;
        pos[3 * have + 2] = (float)z;
        rad[have] = draw_radius(rng);
        have++;
    }
    return 0;
}

}  // namespace

extern "C" {

/* N atoms uniform in a ball of radius (3N / (4 pi rho))^(1/3); 0 on success */
int qs_synth_ball(uint64_t seed, int n, double density, double min_dist, float *pos, float *rad) {
    if (n < 0 || density <= 0) return -1;
    if (n == 0) return 0;
    double R = std::cbrt(3.0 * n / (4.0 * M_PI * density));
    return synth(seed, n, 0.0, R, min_dist, pos, rad);
}

/* N atoms uniform in the shell r_in <= |p| <= r_out */
int qs_synth_shell(uint64_t seed, int n, double r_in, double r_out, double min_dist, float *pos, float *rad) {
    if (n <= 0 || r_out <= r_in) return -1;
    return synth(seed, n, r_in, r_out, min_dist, pos, rad);
}

/* frame = base + independent N(0, sigma) per coordinate (Box-Muller on splitmix64) */
int qs_synth_jitter(uint64_t seed, int n, double sigma, const float *base, float *out) {
    SplitMix64 rng(seed);
    for (int i = 0; i < 3 * n; i += 2) {
        double u1 = rng.uniform(), u2 = rng.uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        double m = std::sqrt(-2.0 * std::log(u1));
        double z0 = m * std::cos(2.0 * M_PI * u2), z1 = m * std::sin(2.0 * M_PI * u2);
        out[i] = (float)((double)base[i] + sigma * z0);
        if (i + 1 < 3 * n) out[i + 1] = (float)((double)base[i + 1] + sigma * z1);
    }
    return 0;
}

/* atom counts log-uniform in [lo, hi]: round(exp(U(ln lo, ln hi))) */
int qs_synth_sizes(uint64_t seed, int count, double lo, double hi, int *out) {
    SplitMix64 rng(seed);
    for (int i = 0; i < count; i++) {
        double u = rng.uniform();
        out[i] = (int)std::llround(std::exp(std::log(lo) + u * (std::log(hi) - std::log(lo))));
    }
    return 0;
}

}  // extern "C"
