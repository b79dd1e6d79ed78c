// Radial hydration shell (RadialHydration, RadialHydration.cpp:40-206, on the
// grid of GridBasedHydration::hydrate, GridBasedHydration.cpp:36-79).
//
// Host code: the placement is sequential (every accepted water marks the grid
// and so changes the acceptance of the candidates after it), O(atoms x 26 x
// 78) grid lookups.
#pragma once

#include <cstdint>
#include <vector>

namespace ausaxs::gpu {

struct Hydration {
    std::vector<double> waters;   // m x 3, body after body, in placement order
    std::vector<int> body_count;  // waters per body
};

// The grid's radii, the shell correction and the grid's extent, the knobs
// the reference's own hydration tests turn (GridDebug's get_atomic_radius /
// get_hydration_radius overrides, settings::hydrate::shell_correction and the
// Grid(Limit3D) constructor, tests/feature/grid/grid.cpp:707-748). The
// defaults are the reference's: per-element van der Waals radii, the oxygen
// radius 1.71 A, -0.35 A, and the grid of Grid(bodies) from the atoms.
struct HydrateSettings {
    double atomic_radius = -1;         // < 0: vdw_radius of each atom's type
    double hydration_radius = 1.71;    // Grid::get_hydration_radius
    double shell_correction = -0.35;   // settings::hydrate::shell_correction
    bool explicit_grid = false;        // Grid(Limit3D): cells of 1 A from grid_min to grid_max
    double grid_min[3] = {0, 0, 0};
    double grid_max[3] = {0, 0, 0};
};

// atoms: n x 3 FP64 with form-factor types; bodies: sizes summing to n (empty:
// one body). noise: n x 26 x 3 displacements added to the candidate
// positions (the reference draws N(0, 0.75) per component; the reference's
// tests set it to zero), or nullptr for none.
Hydration hydrate_radial(const double* atoms, const int32_t* ff, int n, const std::vector<int>& bodies,
                         const double* noise, const HydrateSettings& hs = HydrateSettings{});

}  // namespace ausaxs::gpu
