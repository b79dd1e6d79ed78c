// Device helpers shared by the pair and lattice kernels: the reference's
// SIMD-body binning rule (correctly rounded, and the pair kernel's fast path
// with its tie fallback), and the fixed-point offset of a pair's distance
// from its bin centre.
//
// Offsets are integers in units of 2^-E A (E = delta_exp(inv_w), see
// device_api.hpp). The pair kernel works on coordinates scaled by
// 2^(E-1) (exact: a power of two), so that for a squared distance D in
// scaled units, r = rsq(D) and the root s = D r, the Newton correction of s,
// (D - s^2) / (2 s), is e * r offset units with e = fma(-s, s, D): the
// offset of the real root costs two FMAs beyond the root itself. The
// reference sums correctly rounded float distances (WeightedEntry.h:64-81
// fed by CompactCoordinatesXYZFF.h:514); the real root's offset differs from
// theirs by at most half an ulp per pair, with no bias. An uncorrected root
// would not do: v_sqrt_f32's root falls one ulp short in a fraction of cases
// that changes with the mantissa of D, which shifted the mean distance of
// the bins of a 250k-atom structure by -0.14 ulp on average and by up to
// -0.4 ulp in single bins (measured against the oracle's sums), enough to
// move its I(q) by 1.5e-4 near the minima of the wide-angle tail.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_api.hpp"

namespace ausaxs::gpu::binning {

// correctly rounded square root (IEEE): sqrtf is v_sqrt_f32 plus an
// FMA-residual correction, while __fsqrt_rn lowers to the uncorrected
// instruction (launch_bin_check verifies this exhaustively on device)
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

// The fast path bins through t = fma(s, inv_ws, kMagic): the product is exact
// inside the FMA and the sum lands in [2^23, 2^24), where floats are the
// integers, so t = kMagic + rint(s * inv_ws) (ties to even) and the bin is the
// bit pattern of t minus kMagicBits (no conversion). dl = fma(s, inv_ws,
// -(t - kMagic)) is the offset from the bin centre in bins, rounded once.
constexpr float kMagic = 12582912.0f;  // 1.5 * 2^23
constexpr uint32_t kMagicBits = 0x4B400000u;
// Tie window per bin unit: 1.75 * 2^-22 (see bin_fast)
constexpr float kTieWin = 4.172325134277344e-7f;

// Constants of one launch: binning and offsets of a histogram with
// inv_w = 1 / bin width, in the pair kernel's scaled coordinates.
struct Consts {
    float scale;     // 2^(E-1): coordinate scale of the pair loop
    float inv_ws;    // inv_w / scale (exact): bins per scaled unit
    float win;       // kTieWin * inv_ws: tie window per scaled unit of root
    float q;         // offset units per bin: 2^E / inv_w
    float m2;        // 1.5 * 2^23 + kDeltaBias: offset code of a zero offset
};

__host__ __device__ inline Consts make_consts(float inv_w) {
    const int E = delta_exp(inv_w);
    Consts c;
    c.scale = ldexpf(1.0f, E - 1);
    c.inv_ws = ldexpf(inv_w, 1 - E);
    c.win = kTieWin * c.inv_ws;
    c.q = (float)(ldexp(1.0, E) / (double)inv_w);
    c.m2 = 12582912.0f + (float)kDeltaBias;
    return c;
}

__device__ __forceinline__ unsigned bin_of(float t) { return __float_as_uint(t) - kMagicBits; }

// rsq(D) clamped to [0, 1] (the clamp is an output modifier of v_rsq_f32):
// D = 0 gives 0 * 1 instead of 0 * inf in the root and the correction, and
// below D = 1 (distances under 2^(1-E) A, two offset units, all in bin 0)
// the root is D itself and the offset code stays within 0.35 units of the
// real root's.
__device__ __forceinline__ float rsq_clamped(float D) {
    return __builtin_amdgcn_fmed3f(__builtin_amdgcn_rsqf(D), 0.f, 1.f);
}

// Fast binning of one scaled D with the root s = D * rsq(D) (v_rsq_f32 is
// within 1 ulp, so s is within 1.5 * 2^-23 relative of the real root and
// 2^-22 of the correctly rounded one; one transcendental instead of the
// sqrt and rsq both). The reference's x is the correctly rounded root times
// inv_w, rounded; the exact product s * inv_ws lies within 1.25 * 2^-22 * x
// of it, so both round to the same bin unless x is that close to a .5 tie.
// Returns true inside a window of 1.75 * 2^-22 * x around a tie, tested as
// |dl| + 1.75 * 2^-22 * x > 0.5 with one rounding (the margin covers it and
// the rounding of dl); there the caller redoes the bin with bin_tie. The test
// is an FMA with the absolute-value modifier and one compare against an
// inline constant: with every operand in a vector register the FMA issues at
// the full rate of the FP32 units, which an operand in a scalar register
// halves on gfx950.
__device__ __forceinline__ bool bin_fast(float D, const Consts& c, float& s, float& t, float& dl) {
    s = D * rsq_clamped(D);
    t = __builtin_fmaf(s, c.inv_ws, kMagic);
    dl = __builtin_fmaf(s, c.inv_ws, kMagic - t);
    return __builtin_fmaf(s, c.win, fabsf(dl)) > 0.5f;
}

// Offset code of a pair from its root s and offset dl (bins): the bit
// pattern of the float m2 + rne(dl * q) plus the Newton term e * r, rounded
// again, i.e. kLowBias + k with k = offset / 2^-E + kDeltaBias in
// [0, 2^20 + 2^14] (the correction is at most x / 4 units for a pair x bins
// out, and the host keeps x below 2^15). The two roundings and the float q
// keep k within 1 + 2^-24 q units of the real root's offset, with no bias.
// rsq_clamped(D) is the bin_fast one (the compiler keeps one v_rsq_f32).
__device__ __forceinline__ uint32_t offset_code(float D, float s, float dl, const Consts& c) {
    const float r = rsq_clamped(D);
    const float e = __builtin_fmaf(-s, s, D);
    return __float_as_uint(__builtin_fmaf(e, r, __builtin_fmaf(dl, c.q, c.m2)));
}

// Correctly rounded sqrt(d2) from a raw v_sqrt_f32 result s (within 1 ulp),
// for the tie fallback, where d2 >= (bin_width/2)^2 is normal. With s- and s+
// the neighbours of s, s- s and s s+ are exact multiples of ulp(s)^2, as is
// d2, so d2 > s- s implies d2 >= s- s + ulp^2 > ((s- + s)/2)^2 and d2 <= s- s
// implies sqrt(d2) < (s- + s)/2 (likewise for s+). The FMA residuals'
// signs are exact, so they pick the correctly rounded neighbour.
__device__ __forceinline__ float sqrt_fix(float d2, float s) {
    const float s_dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float s_up = __uint_as_float(__float_as_uint(s) + 1u);
    const float r_dn = __builtin_fmaf(-s_dn, s, d2);
    const float r_up = __builtin_fmaf(-s_up, s, d2);
    s = r_dn <= 0.f ? s_dn : s;
    return r_up > 0.f ? s_up : s;
}

// Inside the tie window: the bin of the reference's rule with the correctly
// rounded root (x = root * inv_w rounded, bin = rint(x); scaling the root and
// dividing inv_w by the same power of two leaves the product unchanged). The
// offset keeps the fast path's formula with its root s (offset_code then
// adds the Newton correction), so a pair's offset is a function of D alone
// whichever path bins it; it may pass +-0.5 bins by a few ulp.
__device__ __forceinline__ void bin_tie(float D, float s, const Consts& c, float& t, float& dl) {
    const float b = __builtin_rintf(sqrt_fix(D, __builtin_amdgcn_sqrtf(D)) * c.inv_ws);
    t = b + kMagic;
    dl = __builtin_fmaf(s, c.inv_ws, -b);
}

// One scaled D binned as the pair kernel bins it (fast path, tie fallback):
// t holds the reference's bin, the result is the offset code.
__device__ __forceinline__ uint32_t bin_pair(float D, const Consts& c, float& t) {
    float s, dl;
    if (bin_fast(D, c, s, t, dl)) bin_tie(D, s, c, t, dl);
    return offset_code(D, s, dl, c);
}

// The reference's SIMD rule with the correctly rounded root on unscaled
// d2 (tail strips): bin, and the offset code of the root (no correction).
__device__ __forceinline__ uint32_t simd_rule_exact(float d2, float inv_w, const Consts& c, float& b) {
    const float r = sqrt_rn(d2);
    b = __builtin_rintf(r * inv_w);
    return __float_as_uint(__builtin_fmaf(__builtin_fmaf(r, inv_w, -b), c.q, c.m2));
}

}  // namespace ausaxs::gpu::binning
