// Deterministic synthetic structures for the benchmark configurations.
//
// PRNG: xoshiro256** seeded through SplitMix64, so the same seed gives the
// same coordinates on every platform (no std::*_distribution, whose output is
// implementation-defined).
#include "synth.hpp"

#include <cmath>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace ausaxs::gpu {

namespace {
uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
}  // namespace

Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
}

uint64_t Rng::next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double Rng::uniform() { return (double)(next() >> 11) * 0x1.0p-53; }

void Rng::unit_vector(double v[3]) {
    // Marsaglia (1972): uniform on the sphere from the unit disk
    for (;;) {
        const double a = 2 * uniform() - 1, b = 2 * uniform() - 1, s2 = a * a + b * b;
        if (s2 >= 1 || s2 == 0) continue;
        const double f = 2 * std::sqrt(1 - s2);
        v[0] = a * f;
        v[1] = b * f;
        v[2] = 1 - 2 * s2;
        return;
    }
}

namespace {

struct CellGrid {
    double cell, lo[3];
    int dim[3];
    std::vector<int> head, next;
    CellGrid(double c, const double half[3], int n) : cell(c) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = -half[k] - c;
            dim[k] = (int)std::ceil((2 * half[k] + 2 * c) / c) + 1;
        }
        head.assign((size_t)dim[0] * dim[1] * dim[2], -1);
        next.reserve(n);
    }
    int cell_of(const double p[3], int k) const {
        int c = (int)std::floor((p[k] - lo[k]) / cell);
        return c < 0 ? 0 : (c >= dim[k] ? dim[k] - 1 : c);
    }
    void insert(const double p[3], int id) {
        const size_t c = ((size_t)cell_of(p, 0) * dim[1] + cell_of(p, 1)) * dim[2] + cell_of(p, 2);
        next.push_back(head[c]);
        head[c] = id;
    }
    template <class F>
    bool any_within(const double p[3], double r, const std::vector<double>& xyz, F skip) const {
        const int cx = cell_of(p, 0), cy = cell_of(p, 1), cz = cell_of(p, 2);
        const double r2 = r * r;
        for (int x = std::max(0, cx - 1); x <= std::min(dim[0] - 1, cx + 1); ++x)
            for (int y = std::max(0, cy - 1); y <= std::min(dim[1] - 1, cy + 1); ++y)
                for (int z = std::max(0, cz - 1); z <= std::min(dim[2] - 1, cz + 1); ++z)
                    for (int id = head[((size_t)x * dim[1] + y) * dim[2] + z]; id >= 0; id = next[id]) {
                        if (skip(id)) continue;
                        const double dx = xyz[3 * id] - p[0], dy = xyz[3 * id + 1] - p[1], dz = xyz[3 * id + 2] - p[2];
                        if (dx * dx + dy * dy + dz * dz < r2) return true;
                    }
        return false;
    }
};

inline bool inside(const double p[3], const double a[3]) {
    return (p[0] * p[0]) / (a[0] * a[0]) + (p[1] * p[1]) / (a[1] * a[1]) + (p[2] * p[2]) / (a[2] * a[2]) <= 1.0;
}

}  // namespace

void synth_globule(uint64_t seed, int n, double ax, double ay, double az, double r_excl, float* out) {
    if (n < 0 || !(ax > 0 && ay > 0 && az > 0) || !(r_excl >= 0)) throw std::runtime_error("synth_globule: bad args");
    Rng rng(seed);
    const double a[3] = {ax, ay, az};
    const double step = 3.8;
    double r = r_excl;
    CellGrid grid(std::max(r_excl, 1.0), a, n);
    std::vector<double> xyz;
    xyz.reserve(3 * (size_t)n);
    int prev = -1;
    auto random_inside = [&](double p[3]) {
        do {
            for (int k = 0; k < 3; ++k) p[k] = (2 * rng.uniform() - 1) * a[k];
        } while (!inside(p, a));
    };
    for (int i = 0; i < n; ++i) {
        double p[3];
        bool placed = false;
        if (prev >= 0) {
            for (int t = 0; t < 64 && !placed; ++t) {
                double v[3];
                rng.unit_vector(v);
                for (int k = 0; k < 3; ++k) p[k] = xyz[3 * prev + k] + step * v[k];
                if (!inside(p, a)) continue;
                placed = !grid.any_within(p, r, xyz, [&](int id) { return id == prev; });
            }
        }
        // stuck (or first atom): start a new chain at a free random point
        for (int t = 0; !placed; ++t) {
            random_inside(p);
            placed = !grid.any_within(p, r, xyz, [](int) { return false; });
            if (!placed && t > 2000) r *= 0.98, t = 0;  // densest packings: relax the exclusion
        }
        for (int k = 0; k < 3; ++k) xyz.push_back(p[k]);
        grid.insert(p, i);
        prev = i;
        // element mix: C 62% / N 17% / O 20% / S 1% -> form-factor types C=3, N=7, O=11, S=12
        const double u = rng.uniform();
        const int32_t ff = u < 0.62 ? 3 : u < 0.79 ? 7 : u < 0.99 ? 11 : 12;
        out[4 * (size_t)i + 0] = (float)p[0];
        out[4 * (size_t)i + 1] = (float)p[1];
        out[4 * (size_t)i + 2] = (float)p[2];
        std::memcpy(&out[4 * (size_t)i + 3], &ff, 4);
    }
}

void synth_shell(uint64_t seed, int m, double ax, double ay, double az, double offset, float* out) {
    if (m < 0 || !(ax > 0 && ay > 0 && az > 0)) throw std::runtime_error("synth_shell: bad args");
    Rng rng(seed);
    const int32_t water = 1;
    for (int i = 0; i < m; ++i) {
        double u[3];
        rng.unit_vector(u);
        const double p[3] = {ax * u[0], ay * u[1], az * u[2]};
        double nv[3] = {p[0] / (ax * ax), p[1] / (ay * ay), p[2] / (az * az)};
        const double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        for (int k = 0; k < 3; ++k) out[4 * (size_t)i + k] = (float)(p[k] + offset * nv[k] / nn);
        std::memcpy(&out[4 * (size_t)i + 3], &water, 4);
    }
}

}  // namespace ausaxs::gpu
