// Radial hydration shell (see hydrate.hpp).
#include "hydrate.hpp"

#include "grid.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>

namespace ausaxs::gpu {

namespace {
// grid::detail::State bits (GridObj.h)
constexpr uint8_t kACenter = 1, kAArea = 2, kWCenter = 4, kWArea = 8, kVolume = 16;
constexpr double kRvol = 2.15;      // settings::grid::min_exv_radius
constexpr double kMinScore = 0.25;  // settings::grid::detail::min_score
}  // namespace

Hydration hydrate_radial(const double* atoms, const int32_t* ff, int n, const std::vector<int>& bodies,
                         const double* noise, const HydrateSettings& hs) {
    Hydration out;
    if (n < 0) throw std::runtime_error("ausaxs_hip: negative atom count");
    std::vector<int> sizes = bodies.empty() ? std::vector<int>{n} : bodies;
    int64_t total = 0;
    for (int s : sizes) {
        if (s < 0) throw std::runtime_error("ausaxs_hip: negative body size");
        total += s;
    }
    if (total != n) throw std::runtime_error("ausaxs_hip: body sizes do not add up to the atom count");
    out.body_count.assign(sizes.size(), 0);
    if (n == 0) return out;
    if (!(hs.hydration_radius > 0) || !std::isfinite(hs.shell_correction) || !(hs.atomic_radius < 64) ||
        !(hs.hydration_radius < 64))
        throw std::runtime_error("ausaxs_hip: hydration radii must be positive and finite");
    // Grid::get_atomic_radius (throws on types without a radius unless overridden)
    std::vector<double> ra(n);
    for (int i = 0; i < n; ++i) ra[i] = hs.atomic_radius < 0 ? vdw_radius(ff[i]) : hs.atomic_radius;

    // the molecule's grid from its atoms (GridBasedHydration::initialize creates
    // it after clear_hydration), or the caller's limits (Grid(Limit3D): span /
    // width bins per axis); Grid::add marking the atom centres
    double origin[3];
    int B[3];
    if (hs.explicit_grid) {
        for (int k = 0; k < 3; ++k) {
            const double span = hs.grid_max[k] - hs.grid_min[k];
            if (!(span >= 1) || !(span <= 4096)) throw std::runtime_error("ausaxs_hip: grid limits must span 1-4096 A");
            origin[k] = hs.grid_min[k];
            B[k] = (int)(span / 1.0);  // Axis3D(limits, width): bins = span / width
        }
    } else {
        const GridGeometry geo = grid_geometry(atoms, n, nullptr, 0);
        for (int k = 0; k < 3; ++k) origin[k] = geo.origin[k], B[k] = geo.bins[k];
    }
    std::vector<uint8_t> g((size_t)B[0] * B[1] * B[2], 0);
    auto at = [&](int i, int j, int k) -> uint8_t& { return g[((size_t)i * B[1] + j) * B[2] + k]; };
    auto to_bin = [&](double v, int k) { return (int)std::round((v - origin[k]) / 1.0); };
    auto to_coord = [&](int i, int k) { return origin[k] + i * 1.0; };
    auto oob = [&](int i, int j, int k) { return i < 0 || i >= B[0] || j < 0 || j >= B[1] || k < 0 || k >= B[2]; };
    std::vector<std::array<int, 3>> loc(n);
    for (int a = 0; a < n; ++a) {
        for (int k = 0; k < 3; ++k) loc[a][k] = to_bin(atoms[3 * a + k], k);
        if (oob(loc[a][0], loc[a][1], loc[a][2])) throw std::runtime_error("ausaxs_hip: an atom lies outside the grid");
        at(loc[a][0], loc[a][1], loc[a][2]) |= kACenter;
    }
    // Grid::expand_volume: SphericalExpander<true, false> per atom, in order
    // (SphericalExpander.h:29-90)
    for (int a = 0; a < n; ++a) {
        const double rvdw = ra[a], rvdw2 = std::pow(rvdw, 2), rvol2 = std::pow(kRvol, 2);
        const int br = (int)std::ceil(std::max(rvdw, kRvol) / 1.0);
        int lo[3], hi[3];
        for (int k = 0; k < 3; ++k) lo[k] = std::max(loc[a][k] - br, 0), hi[k] = std::min(loc[a][k] + br + 1, B[k]);
        const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
        for (int i = lo[0]; i < hi[0]; ++i) {
            const double x2 = std::pow(to_coord(i, 0) - ax, 2);
            for (int j = lo[1]; j < hi[1]; ++j) {
                const double x2y2 = x2 + std::pow(to_coord(j, 1) - ay, 2);
                for (int k = lo[2]; k < hi[2]; ++k) {
                    const double dist = x2y2 + std::pow(to_coord(k, 2) - az, 2);
                    uint8_t& s = at(i, j, k);
                    const bool fill = dist <= rvdw2;
                    if (fill) {
                        if (!(s == 0 || (s & kVolume))) continue;  // is_empty_or_volume
                        s |= kAArea;
                    }
                    if (!fill && dist <= rvol2) {
                        if (s != 0) continue;
                        s |= kVolume;
                    }
                }
            }
        }
    }

    // RadialHydration::prepare_rotations(8): directions and the probe offsets
    // at 1, 3 and 5 shell radii
    const double rh = hs.hydration_radius + hs.shell_correction;
    const std::vector<std::array<double, 3>> dirs = sphere_directions(8);
    const int nd = (int)dirs.size();
    std::vector<std::array<int, 3>> p1(nd), p3(nd), p5(nd);
    for (int d = 0; d < nd; ++d)
        for (int c = 0; c < 3; ++c) {
            p1[d][c] = (int)(std::round(rh * dirs[d][c]) / 1.0);
            p3[d][c] = (int)(std::round(3 * rh * dirs[d][c]) / 1.0);
            p5[d][c] = (int)(std::round(5 * rh * dirs[d][c]) / 1.0);
        }
    // GridObj::is_only_empty_or_volume
    auto open = [](uint8_t s) { return s == 0 || s == kVolume; };
    // RadialHydration::collision_check (RadialHydration.cpp:141-206)
    auto collision_check = [&](int x, int y, int z) {
        int score = 0;
        unsigned inside_1rh = 0;
        for (int d = 0; d < nd; ++d) {
            {
                const int xr = std::clamp(x + p1[d][0], 0, B[0] - 1), yr = std::clamp(y + p1[d][1], 0, B[1] - 1),
                          zr = std::clamp(z + p1[d][2], 0, B[2] - 1);
                if (!open(at(xr, yr, zr))) {
                    if (2 < ++inside_1rh) return false;
                    continue;
                }
            }
            {
                const int xr = x + p3[d][0], yr = y + p3[d][1], zr = z + p3[d][2];
                if (oob(xr, yr, zr)) {
                    score += 2;
                    continue;
                }
                if (!open(at(xr, yr, zr))) {
                    score -= 2;
                    continue;
                }
                score += 2;
            }
            {
                const int xr = x + p5[d][0], yr = y + p5[d][1], zr = z + p5[d][2];
                if (oob(xr, yr, zr)) {
                    score += 1;
                    continue;
                }
                if (!open(at(xr, yr, zr))) {
                    score -= 1;
                    continue;
                }
            }
        }
        const double max_points = 3.0 * nd;
        return !(score <= kMinScore * max_points);
    };

    // generate_explicit_hydration per body (RadialHydration::global() is false)
    const double rw2 = std::pow(hs.hydration_radius, 2);
    const int brw = (int)std::ceil(hs.hydration_radius / 1.0);
    int a0 = 0;
    for (size_t b = 0; b < sizes.size(); ++b) {
        for (int a = a0; a < a0 + sizes[b]; ++a) {
            const double reff = ra[a] + rh;
            for (int d = 0; d < nd; ++d) {
                double w[3];
                for (int c = 0; c < 3; ++c) {
                    const double nz = noise ? noise[((size_t)a * nd + d) * 3 + c] : 0.0;
                    w[c] = atoms[3 * a + c] + dirs[d][c] * reff + nz;
                }
                int wb[3];
                for (int c = 0; c < 3; ++c) wb[c] = to_bin(w[c], c);
                if (oob(wb[0], wb[1], wb[2])) continue;  // Grid::is_valid_bin
                if (!open(at(wb[0], wb[1], wb[2])) || !collision_check(wb[0], wb[1], wb[2])) continue;
                // Grid::add(water, expand): the centre, then W_AREA on empty
                // voxels within the hydration radius (SphericalExpander.h:92-140)
                at(wb[0], wb[1], wb[2]) |= kWCenter;
                int lo[3], hi[3];
                for (int k = 0; k < 3; ++k) lo[k] = std::max(wb[k] - brw, 0), hi[k] = std::min(wb[k] + brw + 1, B[k]);
                for (int i = lo[0]; i < hi[0]; ++i) {
                    const double x2 = std::pow(to_coord(i, 0) - w[0], 2);
                    for (int j = lo[1]; j < hi[1]; ++j) {
                        const double x2y2 = x2 + std::pow(to_coord(j, 1) - w[1], 2);
                        for (int k = lo[2]; k < hi[2]; ++k) {
                            const double dist = x2y2 + std::pow(to_coord(k, 2) - w[2], 2);
                            if (dist <= rw2 && at(i, j, k) == 0) at(i, j, k) |= kWArea;
                        }
                    }
                }
                out.waters.insert(out.waters.end(), w, w + 3);
                ++out.body_count[b];
            }
        }
        a0 += sizes[b];
    }
    return out;
}

}  // namespace ausaxs::gpu
