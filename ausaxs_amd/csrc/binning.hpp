// Device helpers shared by the pair and lattice kernels: the reference's
// SIMD-body binning rule (correctly rounded, and the pair kernel's fast path
// with its tie fallback), and the fixed-point quantization of a pair's
// offset from its bin centre.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace ausaxs::gpu::binning {

// correctly rounded square root (IEEE): sqrtf is v_sqrt_f32 plus an
// FMA-residual correction, while __fsqrt_rn lowers to the uncorrected
// instruction (launch_bin_check verifies this exhaustively on device)
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

// d = sqrt_rn(d2), x = d * inv_w, bin b = rint(x) (round half to even), dl = x - b
__device__ __forceinline__ void simd_rule_exact(float d2, float inv_w, float& x, float& b, float& dl) {
    const float r = sqrt_rn(d2);
    x = r * inv_w;
    b = __builtin_rintf(x);
    dl = x - b;
}

// The pair kernel's fast path bins through t = fma(s, inv_w, kMagic): the
// product s * inv_w is exact inside the FMA and the sum lands in [2^23, 2^24),
// where floats are the integers, so t = kMagic + rint(s * inv_w) (ties to even)
// and the bin is the bit pattern of t minus kMagicBits (no conversion).
// dl = fma(s, inv_w, -(t - kMagic)) is the offset from the bin centre,
// rounded once.
constexpr float kMagic = 12582912.0f;  // 1.5 * 2^23
constexpr uint32_t kMagicBits = 0x4B400000u;
// Tie window slope per bin unit: 1.25 * 2^-22 (see bin_fast)
constexpr float kTieSlope = -2.98023223876953125e-7f;

__device__ __forceinline__ unsigned bin_of(float t) { return __float_as_uint(t) - kMagicBits; }

// Fast binning of one d2 with the raw v_sqrt_f32 root s (within 1 ulp, 2^-23
// relative, of the correctly rounded root). The reference's x is the
// correctly rounded root times inv_w, rounded; the exact product s * inv_w
// lies within 1.5 * 2^-23 * x of it, so both round to the same bin unless x is
// that close to a .5 tie. Returns true inside a window of 1.25 * 2^-22 * x
// around a tie (the margin covers the rounding of dl and of the threshold);
// there the caller redoes the bin with bin_tie. slope_w = kTieSlope * inv_w.
__device__ __forceinline__ bool bin_fast(float d2, float inv_w, float slope_w, float& t, float& dl) {
    const float s = __builtin_amdgcn_sqrtf(d2);
    t = __builtin_fmaf(s, inv_w, kMagic);
    dl = __builtin_fmaf(s, inv_w, kMagic - t);
    return fabsf(dl) > __builtin_fmaf(s, slope_w, 0.5f);
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
// rounded root (x = r * inv_w rounded, bin = rint(x)). The offset keeps the
// fast path's formula, dl = s * inv_w - bin with the raw root s (clamped to
// [-0.5, 0.5] when the bin differs from the raw root's), so dl is a function
// of d2 alone, whichever path bins the pair.
__device__ __forceinline__ void bin_tie(float d2, float inv_w, float& t, float& dl) {
    const float s = __builtin_amdgcn_sqrtf(d2);
    const float b = __builtin_rintf(sqrt_fix(d2, s) * inv_w);
    t = b + kMagic;
    dl = __builtin_amdgcn_fmed3f(__builtin_fmaf(s, inv_w, -b), -0.5f, 0.5f);
}

// One d2 binned as the pair kernel bins it (fast path, tie fallback): the bin
// is the reference's, the offset a function of d2 alone.
__device__ __forceinline__ void bin_pair(float d2, float inv_w, float slope_w, float& t, float& dl) {
    if (bin_fast(d2, inv_w, slope_w, t, dl)) bin_tie(d2, inv_w, t, dl);
}

// kLowBias + k with k = rne(dl * 2^20) + 2^19: dl * 2^20 + 1.5625 * 2^23 lies in
// [1.5 * 2^23, 1.625 * 2^23], where floats are the integers, so the FMA rounds
// dl * 2^20 to an integer once and the bit pattern is 0x4B400000 + k
__device__ __forceinline__ uint32_t quantize(float dl) {
    return __float_as_uint(__builtin_fmaf(dl, 1048576.0f, 13107200.0f));
}

}  // namespace ausaxs::gpu::binning
