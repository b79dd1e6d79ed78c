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

// atoms: n x 3 FP64 with form-factor types; bodies: sizes summing to n (empty:
// one body). noise: n x 26 x 3 displacements added to the candidate
// positions (the reference draws N(0, 0.75) per component; the reference's
// tests set it to zero), or nullptr for none.
Hydration hydrate_radial(const double* atoms, const int32_t* ff, int n, const std::vector<int>& bodies,
                         const double* noise);

}  // namespace ausaxs::gpu
