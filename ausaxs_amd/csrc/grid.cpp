// Excluded-volume dummy atoms on the integer lattice (see grid.hpp).
#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

namespace ausaxs::gpu {

double vdw_radius(int t) {
    // constants::radius::vdw (vdwTable.h) by form-factor type
    switch (t) {
        case 2: return 1.1;                          // H
        case 3: case 4: case 5: case 6: return 1.96;  // C, CH, CH2, CH3
        case 7: case 8: case 9: case 10: return 1.79; // N, NH, NH2, NH3
        case 1: case 11: return 1.71;                 // OH (water), O
        case 12: case 13: return 2.06;                // S, SH
        case 14: return 1.88;                         // OTHER -> Ar
        default: throw std::runtime_error("ausaxs_hip: no van der Waals radius for form factor type " + std::to_string(t));
    }
}

namespace {
constexpr double kCell = 1.0;     // settings::grid::cell_width
constexpr double kRvol = 2.15;    // settings::grid::min_exv_radius
constexpr uint8_t kACenter = 1, kWCenter = 2, kExv = 4;
}  // namespace

GridExv make_grid_exv(const double* atoms, const int32_t* ff, int n, const double* waters, int m) {
    GridExv out;
    if (n <= 0) return out;
    // grid minimum as in Grid(bodies) (Grid.cpp:34-104): bounding box grown from
    // the origin, expanded by max(0.5*diff*scaling, margin) and a cell, floored.
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], atoms[3 * i + k]), hi[k] = std::max(hi[k], atoms[3 * i + k]);
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], waters[3 * i + k]), hi[k] = std::max(hi[k], waters[3 * i + k]);
    double scaling = 0.25;  // settings::grid::scaling; small systems switch to 1 (Grid.cpp:87-90)
    if (hi[0] - lo[0] < 50 || hi[1] - lo[1] < 50 || hi[2] - lo[2] < 50) scaling = 1;
    const double margin = 2.14 + 1.71 + (-0.35) + 3 * kCell;  // vdw P + vdw O + shell_correction + 3 cells
    for (int k = 0; k < 3; ++k) {
        const double expand = std::max(0.5 * (hi[k] - lo[k]) * scaling, margin);
        out.origin[k] = (int)std::floor(lo[k] - expand - kCell);
    }

    // dense occupancy box around the atoms (one voxel = one lattice point)
    int bmin[3], bmax[3];
    std::vector<int> cen(3 * (size_t)n);
    std::vector<double> r2(n);
    int br_max = 0;
    for (int k = 0; k < 3; ++k) bmin[k] = std::numeric_limits<int>::max(), bmax[k] = std::numeric_limits<int>::min();
    for (int i = 0; i < n; ++i) {
        const double rv = vdw_radius(ff[i]);
        r2[i] = std::max(std::pow(rv, 2), std::pow(kRvol, 2));
        br_max = std::max(br_max, (int)std::ceil(std::max(rv, kRvol) / kCell));
        for (int k = 0; k < 3; ++k) {
            // Grid::to_bins: round((x - min)/width), half away from zero
            const int b = (int)std::round((atoms[3 * i + k] - out.origin[k]) / kCell);
            cen[3 * i + k] = b;
            bmin[k] = std::min(bmin[k], b);
            bmax[k] = std::max(bmax[k], b);
        }
    }
    int dim[3];
    for (int k = 0; k < 3; ++k) {
        bmin[k] -= br_max + 1;
        bmax[k] += br_max + 1;
        dim[k] = bmax[k] - bmin[k] + 1;
    }
    std::vector<uint8_t> g((size_t)dim[0] * dim[1] * dim[2], 0);
    auto at = [&](int i, int j, int k) -> uint8_t& {
        return g[((size_t)(i - bmin[0]) * dim[1] + (j - bmin[1])) * dim[2] + (k - bmin[2])];
    };
    auto inside = [&](int i, int j, int k) {
        return i >= bmin[0] && i <= bmax[0] && j >= bmin[1] && j <= bmax[1] && k >= bmin[2] && k <= bmax[2];
    };
    for (int i = 0; i < n; ++i) at(cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]) |= kACenter;
    for (int w = 0; w < m; ++w) {
        int b[3];
        for (int k = 0; k < 3; ++k) b[k] = (int)std::round((waters[3 * w + k] - out.origin[k]) / kCell);
        if (inside(b[0], b[1], b[2])) at(b[0], b[1], b[2]) |= kWCenter;
    }
    // atomic expansion (SphericalExpander.h:29-90): a voxel holding only a water
    // centre is never claimed by an atom
    for (int a = 0; a < n; ++a) {
        const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
        const int br = (int)std::ceil(std::sqrt(r2[a]) / kCell - 1e-12);
        const int bx = cen[3 * a], by = cen[3 * a + 1], bz = cen[3 * a + 2];
        for (int i = bx - br; i <= bx + br; ++i) {
            const double x2 = std::pow(out.origin[0] + i * kCell - ax, 2);
            for (int j = by - br; j <= by + br; ++j) {
                const double x2y2 = x2 + std::pow(out.origin[1] + j * kCell - ay, 2);
                for (int k = bz - br; k <= bz + br; ++k) {
                    const double dist = x2y2 + std::pow(out.origin[2] + k * kCell - az, 2);
                    if (dist > r2[a]) continue;
                    uint8_t& s = at(i, j, k);
                    if ((s & kWCenter) && !(s & kACenter)) continue;
                    s |= kExv;
                }
            }
        }
    }
    // RawGridExv::create: lexicographic voxel order
    for (int i = bmin[0]; i <= bmax[0]; ++i)
        for (int j = bmin[1]; j <= bmax[1]; ++j)
            for (int k = bmin[2]; k <= bmax[2]; ++k)
                if (at(i, j, k) & (kACenter | kExv)) {
                    out.xyz.push_back(out.origin[0] + i * kCell);
                    out.xyz.push_back(out.origin[1] + j * kCell);
                    out.xyz.push_back(out.origin[2] + k * kCell);
                }
    out.volume = (double)(out.xyz.size() / 3) * kCell * kCell * kCell;
    return out;
}

}  // namespace ausaxs::gpu
