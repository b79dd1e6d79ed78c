// Deterministic synthetic structures (see synth.cpp).
#pragma once

#include <cstdint>
#include <cstring>

namespace ausaxs::gpu {

struct Rng {  // xoshiro256** (Blackman & Vigna 2018), SplitMix64-seeded
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();  // [0, 1)
    void unit_vector(double v[3]);
};

void synth_globule(uint64_t seed, int n, double ax, double ay, double az, double r_excl, float* out);
void synth_shell(uint64_t seed, int m, double ax, double ay, double az, double offset, float* out);

}  // namespace ausaxs::gpu
