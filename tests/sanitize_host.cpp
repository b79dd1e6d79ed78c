// Host-only modules of the library under the compiler's sanitizers.
//
// The CPU-side parts of the path — the excluded-volume grid builder
// (grid.cpp, threaded marking on slabs), radial hydration (hydrate.cpp), the
// synthetic structures (synth.cpp) and the threaded range split of the layout
// passes (host_parallel.hpp) — hold no GPU code,
// so they build with g++ under AddressSanitizer + UndefinedBehaviorSanitizer,
// or ThreadSanitizer, and run here. The program drives every entry point on
// small inputs (several bodies, noise, the surface variants, empty and
// degenerate sets, the error paths) and checks a few facts of each result; a
// sanitizer report or a failed check ends it with a non-zero status.
#include "grid.hpp"
#include "host_parallel.hpp"
#include "hydrate.hpp"
#include "synth.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

using namespace ausaxs::gpu;

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        ++failures;
    }
}

struct Molecule {
    std::vector<double> xyz;
    std::vector<int32_t> ff;
};

Molecule globule(uint64_t seed, int n, double r) {
    std::vector<float> p(4 * (size_t)n);
    synth_globule(seed, n, r, r * 0.9, r * 0.8, 2.0, p.data());
    Molecule m;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) m.xyz.push_back(p[4 * i + k]);
        int32_t t;
        std::memcpy(&t, &p[4 * i + 3], 4);
        m.ff.push_back(t);
    }
    return m;
}

template <class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}

}  // namespace

int main() {
    // synthetic structures: deterministic, inside their ellipsoids
    const Molecule a = globule(3, 1500, 16.0);
    const Molecule b = globule(3, 1500, 16.0);
    expect(a.xyz == b.xyz && a.ff == b.ff, "synth_globule is deterministic");
    std::vector<float> shell(4 * 400);
    synth_shell(5, 400, 16.0, 14.4, 12.8, 3.0, shell.data());
    std::vector<double> waters;
    for (int i = 0; i < 400; ++i)
        for (int k = 0; k < 3; ++k) waters.push_back(shell[4 * i + k]);
    expect(std::isfinite(waters[0]) && std::isfinite(waters.back()), "synth_shell is finite");
    Rng rng(11);
    double u[3];
    rng.unit_vector(u);
    expect(std::fabs(u[0] * u[0] + u[1] * u[1] + u[2] * u[2] - 1) < 1e-12, "unit vector");

    // excluded-volume grid: plain, with a surface, thick surfaces, waters, tiny sets
    const int n = (int)a.ff.size();
    GridExv g = make_grid_exv(a.xyz.data(), a.ff.data(), n, waters.data(), 400);
    expect(!g.xyz.empty() && g.surface.empty(), "grid dummies");
    expect(g.volume == (double)(g.xyz.size() / 3), "grid volume is the dummy count");
    for (double thick : {0.0, 1.0, 3.0}) {
        GridExv s = make_grid_exv(a.xyz.data(), a.ff.data(), n, nullptr, 0, true, thick);
        expect(s.xyz.size() + s.surface.size() >= g.xyz.size(), "surface split keeps the dummies");
        expect(!s.surface.empty(), "a globule has a surface");
    }
    GridExv one = make_grid_exv(a.xyz.data(), a.ff.data(), 1, nullptr, 0, true, 2.0);
    expect(!one.xyz.empty() || !one.surface.empty(), "one atom has dummies");
    expect(throws([&] { make_grid_exv(a.xyz.data(), a.ff.data(), n, nullptr, 0, true, -1.0); }),
           "negative surface thickness is an error");
    GridGeometry geo = grid_geometry(a.xyz.data(), n, waters.data(), 400);
    expect(geo.bins[0] > 0 && geo.bins[1] > 0 && geo.bins[2] > 0, "grid geometry");
    expect(sphere_directions(8).size() == 26, "26 directions for 8 divisions");

    // hydration: one body, three bodies with noise, radius overrides, explicit grid
    Hydration h1 = hydrate_radial(a.xyz.data(), a.ff.data(), n, {}, nullptr);
    expect(!h1.waters.empty() && h1.body_count.size() == 1, "hydration of one body");
    std::vector<double> noise(3 * 26 * (size_t)n);
    Rng nr(17);
    for (double& v : noise) v = 0.75 * (2 * nr.uniform() - 1);
    Hydration h3 = hydrate_radial(a.xyz.data(), a.ff.data(), n, {400, 700, 400}, noise.data());
    expect(h3.body_count.size() == 3, "hydration per body");
    size_t total = 0;
    for (int c : h3.body_count) total += c;
    expect(total * 3 == h3.waters.size(), "water counts add up");
    HydrateSettings hs;
    hs.atomic_radius = 3;
    hs.hydration_radius = 3;
    hs.shell_correction = 0;
    hs.explicit_grid = true;
    for (int k = 0; k < 3; ++k) hs.grid_min[k] = -10, hs.grid_max[k] = 10;
    const double c0[3] = {0, 0, 0};
    const int32_t carbon = 3;
    Hydration hc = hydrate_radial(c0, &carbon, 1, {}, nullptr, hs);
    expect(hc.waters.size() == 6 * 3, "six waters around one carbon atom");
    expect(throws([&] { hydrate_radial(a.xyz.data(), a.ff.data(), n, {400, 400}, nullptr); }),
           "body sizes that do not add up are an error");
    const double far[3] = {50, 0, 0};
    expect(throws([&] { hydrate_radial(far, &carbon, 1, {}, nullptr, hs); }), "an atom outside the grid is an error");

    // the set-up passes' range split: disjoint chunks, every index once
    std::vector<int> hits(1 << 20, 0);
    parallel_for((int64_t)hits.size(), 1 << 12, [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k) hits[k] += 1;
    });
    bool once = true;
    for (int v : hits) once = once && v == 1;
    expect(once, "parallel_for visits every index once");

    if (failures) return 1;
    std::printf("sanitized host modules: all checks passed\n");
    return 0;
}
