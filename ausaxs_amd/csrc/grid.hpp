// Excluded-volume dummy atoms on the integer lattice (grid exv).
//
// Restates grid::Grid(bodies) (Grid.cpp:34-110), the atomic expansion of
// SphericalExpander<true,false> (SphericalExpander.h:29-90) and
// RawGridExv::create (RawGridExv.cpp:10-59) with the default settings
// (cell_width = exv.width = 1 A, min_exv_radius = 2.15 A, AtomicOnly).
//
// With those settings the marking procedure has an order-free result, which
// is what this builder computes directly:
//   dummies = {voxel of every atom centre (to_bins: round half away)}
//           u {lattice points p with |p - a|^2 <= max(r_vdw(a), 2.15)^2 for
//              some atom a, evaluated as ((x-ax)^2 + (y-ay)^2) + (z-az)^2 in
//              FP64, unless p is the voxel of a water centre and of no atom
//              centre}
// listed in lexicographic (x, y, z) voxel order; the grid volume is their count.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace ausaxs::gpu {

struct GridExv {
    std::vector<double> xyz;      // n x 3, integer-valued lattice points (interior ones if split)
    std::vector<double> surface;  // surface lattice points (detect_surface only)
    double volume = 0;            // Grid::get_volume(): count * cell_width^3 (both lists)
    int origin[3]{};              // grid minimum (integer)
};

// atoms: n x 3 FP64 coordinates + form-factor types; waters: m x 3 FP64.
// detect_surface: RawGridWithSurfaceExv::create (RawGridWithSurfaceExv.cpp:7-155)
// with surface_thickness 1: a dummy is on the surface when the radial lines
// from it (RadialLineGenerator, radii 1.001, 2, 3 A, 8 divisions) score at
// least 42 (7 per line that leaves the molecule, 3 or 5 for one blocked at 2
// or 3 A, 0 at 1 A), every other dummy is interior.
// surface_thickness (settings::grid::exv::surface_thickness, A; cell width 1):
// other than one cell, every dummy within int(round(thickness) * 0.5) cells
// (squared integer distance) of a surface dummy is a surface dummy too
// (RawGridWithSurfaceExv.cpp:137-190, 196-204).
GridExv make_grid_exv(const double* atoms, const int32_t* ff, int n, const double* waters, int m,
                      bool detect_surface = false, double surface_thickness = 1.0);

// constants::radius::vdw per form-factor type (FormFactorType.h:174-193)
double vdw_radius(int ff_type);

// The grid box of Grid(bodies) (Grid.cpp:34-110) at cell width 1: integer
// minimum and bin counts, from FP64 atoms and waters.
struct GridGeometry {
    int origin[3]{};
    int bins[3]{};
};
GridGeometry grid_geometry(const double* atoms, int n, const double* waters, int m);

// Unit directions of an octant of a sphere cut into `divisions` angles and
// reflected into all eight, duplicates removed, in the reference's order
// (RadialLineGenerator::generate, RadialLineGenerator.cpp:15-63, and
// RadialHydration::prepare_rotations, RadialHydration.cpp:82-139): 26 for 8.
std::vector<std::array<double, 3>> sphere_directions(int divisions);

}  // namespace ausaxs::gpu
