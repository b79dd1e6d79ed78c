// Excluded-volume dummy atoms on the integer lattice (see grid.hpp).
#include "grid.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>

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

namespace {
// RadialLineGenerator::generate (RadialLineGenerator.cpp:15-63) for cell width
// 1 and the collision radii of RawGridWithSurfaceExv: the bin offsets of the
// points at radius r_k along each direction of an octant-reflected sphere
struct RadialLines {
    std::vector<std::array<int, 3>> b1, b2, b3;
};
RadialLines radial_lines() {
    const double r[3] = {std::sqrt(1.0) + 1e-3, 2.0, 3.0};
    RadialLines out;
    for (const auto& p : sphere_directions(8)) {
        std::array<int, 3> v[3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) v[k][c] = (int)(std::round(r[k] * p[c]) * 1.0);
        out.b1.push_back(v[0]);
        out.b2.push_back(v[1]);
        out.b3.push_back(v[2]);
    }
    return out;
}
}  // namespace

std::vector<std::array<double, 3>> sphere_directions(int divisions) {
    const double pi = 3.14159265358979323846;
    const double ang = 2 * pi / divisions;
    std::vector<std::array<double, 3>> sphere;
    for (double theta = 0; theta <= pi * 0.5; theta += ang)
        for (double phi = 0; phi <= pi * 0.5; phi += ang) {
            const double x = std::cos(phi) * std::sin(theta), y = std::sin(phi) * std::sin(theta),
                         z = std::cos(theta);
            for (int sz : {1, -1})
                for (int sy : {1, -1})
                    for (int sx : {1, -1}) sphere.push_back({sx * x, sy * y, sz * z});
        }
    // the reflections in the reference's order: (x,y,z), (-x,y,z), (x,-y,z), (-x,-y,z), then -z
    std::vector<std::array<double, 3>> rots;
    for (auto p : sphere) {
        for (double& c : p)
            if (std::abs(c) < 1e-5) c = 0;
        bool present = false;
        for (const auto& q : rots) {
            const double d = std::sqrt(std::pow(q[0] - p[0], 2) + std::pow(q[1] - p[1], 2) + std::pow(q[2] - p[2], 2));
            if (d < 1e-5) {
                present = true;
                break;
            }
        }
        if (!present) rots.push_back(p);
    }
    return rots;
}

GridGeometry grid_geometry(const double* atoms, int n, const double* waters, int m) {
    // Grid(bodies) (Grid.cpp:34-104): bounding box grown from the origin,
    // expanded by max(0.5*diff*scaling, margin) and a cell, floored / ceiled
    GridGeometry geo;
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
        geo.origin[k] = (int)std::floor(lo[k] - expand - kCell);
        geo.bins[k] = (int)std::ceil(hi[k] + expand + kCell) - geo.origin[k];  // Axis3D(nmin, nmax, width).bins
    }
    return geo;
}

GridExv make_grid_exv(const double* atoms, const int32_t* ff, int n, const double* waters, int m,
                      bool detect_surface, double surface_thickness) {
    if (!(surface_thickness >= 0) || surface_thickness > 64)
        throw std::runtime_error("ausaxs_hip: surface_thickness must lie in [0, 64] A");
    GridExv out;
    if (n <= 0) return out;
    const GridGeometry geo = grid_geometry(atoms, n, waters, m);
    int bins[3];
    for (int k = 0; k < 3; ++k) out.origin[k] = geo.origin[k], bins[k] = geo.bins[k];

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
    // The marking below has an order-free result, so it runs on slabs of x in
    // parallel: each thread marks and later collects the voxels of its own
    // slab (no two threads write one voxel) and the slabs are concatenated in
    // x order, which keeps the lexicographic voxel order.
    const int n_thr = (int)std::clamp<unsigned>(std::thread::hardware_concurrency(), 1u,
                                                (unsigned)std::clamp(dim[0] / 8, 1, 16));
    auto slab = [&](int t, int& x0, int& x1) {
        x0 = bmin[0] + (int)((int64_t)dim[0] * t / n_thr);
        x1 = bmin[0] + (int)((int64_t)dim[0] * (t + 1) / n_thr);  // exclusive
    };
    auto parallel = [&](auto&& body) {
        std::vector<std::thread> pool;
        for (int t = 1; t < n_thr; ++t) pool.emplace_back(body, t);
        body(0);
        for (auto& th : pool) th.join();
    };
    // atomic expansion (SphericalExpander.h:29-90): a voxel holding only a water
    // centre is never claimed by an atom
    parallel([&](int t) {
        int x0, x1;
        slab(t, x0, x1);
        for (int a = 0; a < n; ++a) {
            const int br = (int)std::ceil(std::sqrt(r2[a]) / kCell - 1e-12);
            const int bx = cen[3 * a], by = cen[3 * a + 1], bz = cen[3 * a + 2];
            const int i0 = std::max(bx - br, x0), i1 = std::min(bx + br, x1 - 1);
            if (i0 > i1) continue;
            const double ax = atoms[3 * a], ay = atoms[3 * a + 1], az = atoms[3 * a + 2];
            for (int i = i0; i <= i1; ++i) {
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
    });
    // GridObj::is_empty_or_water: nothing, or a water centre (water areas
    // only ever fill empty voxels); voxels outside the box hold no atom
    auto empty_or_water = [&](int i, int j, int k) {
        if (!inside(i, j, k)) return true;
        const uint8_t s = at(i, j, k);
        return s == 0 || (s & kWCenter);
    };
    auto out_of_bounds = [&](int i, int j, int k) {
        return i < 0 || i >= bins[0] || j < 0 || j >= bins[1] || k < 0 || k >= bins[2];
    };
    RadialLines lines;
    if (detect_surface) lines = radial_lines();
    // collision_check (RawGridWithSurfaceExv.cpp:7-61): true for interior dummies
    auto interior = [&](int i, int j, int k) {
        int score = 0;
        for (size_t d = 0; d < lines.b1.size(); ++d) {
            const auto& a = lines.b1[d];
            if (out_of_bounds(i + a[0], j + a[1], k + a[2])) {
                score += 7;
                continue;
            }
            if (!empty_or_water(i + a[0], j + a[1], k + a[2])) continue;
            const auto& b = lines.b2[d];
            if (out_of_bounds(i + b[0], j + b[1], k + b[2])) {
                score += 7;
                continue;
            }
            if (!empty_or_water(i + b[0], j + b[1], k + b[2])) {
                score += 3;
                continue;
            }
            const auto& c = lines.b3[d];
            if (out_of_bounds(i + c[0], j + c[1], k + c[2])) {
                score += 7;
                continue;
            }
            if (!empty_or_water(i + c[0], j + c[1], k + c[2])) {
                score += 5;
                continue;
            }
            score += 7;
        }
        return score < 42;
    };
    // RawGridExv::create / RawGridWithSurfaceExv::create: lexicographic voxel
    // order. Pass 1 classifies each slab's dummies and counts them, pass 2
    // writes them at their slab's offset into the output.
    std::vector<uint8_t> cls(g.size(), 0);  // 1 interior, 2 surface
    std::vector<int64_t> cnt_in(n_thr + 1, 0), cnt_sf(n_thr + 1, 0);
    auto idx = [&](int i, int j, int k) { return ((size_t)(i - bmin[0]) * dim[1] + (j - bmin[1])) * dim[2] + (k - bmin[2]); };
    parallel([&](int t) {
        int x0, x1;
        slab(t, x0, x1);
        int64_t ci = 0, cs = 0;
        for (int i = x0; i < x1; ++i)
            for (int j = bmin[1]; j <= bmax[1]; ++j)
                for (int k = bmin[2]; k <= bmax[2]; ++k)
                    if (at(i, j, k) & (kACenter | kExv)) {
                        const bool in = !detect_surface || interior(i, j, k);
                        cls[idx(i, j, k)] = in ? 1 : 2;
                        (in ? ci : cs) += 1;
                    }
        cnt_in[t + 1] = ci;
        cnt_sf[t + 1] = cs;
    });
    // A thicker surface (RawGridWithSurfaceExv.cpp:137-190): the reference
    // marks every dummy within expand cells of each first-pass surface dummy;
    // the same set is the dummies with a first-pass surface dummy within
    // expand cells of them (the relation is symmetric), gathered per slab.
    const int thick = (int)std::round(surface_thickness / kCell);
    if (detect_surface && thick != 1) {
        const int e = (int)(thick * 0.5), e2 = e * e;
        std::vector<uint8_t> grown(cls);
        parallel([&](int t) {
            int x0, x1;
            slab(t, x0, x1);
            int64_t ci = 0, cs = 0;
            for (int i = x0; i < x1; ++i)
                for (int j = bmin[1]; j <= bmax[1]; ++j)
                    for (int k = bmin[2]; k <= bmax[2]; ++k) {
                        uint8_t& c = grown[idx(i, j, k)];
                        if (c == 1) {
                            for (int l = -e; l <= e && c == 1; ++l)
                                for (int mm = -e; mm <= e && c == 1; ++mm)
                                    for (int nn = -e; nn <= e; ++nn) {
                                        const int a = i + l, b = j + mm, d = k + nn;
                                        if (l * l + mm * mm + nn * nn > e2 || a < bmin[0] || a > bmax[0] ||
                                            b < bmin[1] || b > bmax[1] || d < bmin[2] || d > bmax[2])
                                            continue;
                                        if (cls[idx(a, b, d)] == 2) {
                                            c = 2;
                                            break;
                                        }
                                    }
                        }
                        if (c) (c == 1 ? ci : cs) += 1;
                    }
            cnt_in[t + 1] = ci;
            cnt_sf[t + 1] = cs;
        });
        cls.swap(grown);
    }
    for (int t = 0; t < n_thr; ++t) cnt_in[t + 1] += cnt_in[t], cnt_sf[t + 1] += cnt_sf[t];
    out.xyz.resize(3 * (size_t)cnt_in[n_thr]);
    out.surface.resize(3 * (size_t)cnt_sf[n_thr]);
    parallel([&](int t) {
        int x0, x1;
        slab(t, x0, x1);
        double* pi = out.xyz.data() + 3 * cnt_in[t];
        double* ps = out.surface.data() + 3 * cnt_sf[t];
        for (int i = x0; i < x1; ++i)
            for (int j = bmin[1]; j <= bmax[1]; ++j)
                for (int k = bmin[2]; k <= bmax[2]; ++k) {
                    const uint8_t c = cls[idx(i, j, k)];
                    if (!c) continue;
                    double*& p = c == 1 ? pi : ps;
                    p[0] = out.origin[0] + i * kCell;
                    p[1] = out.origin[1] + j * kCell;
                    p[2] = out.origin[2] + k * kCell;
                    p += 3;
                }
    });
    out.volume = (double)((out.xyz.size() + out.surface.size()) / 3) * kCell * kCell * kCell;
    return out;
}

}  // namespace ausaxs::gpu
