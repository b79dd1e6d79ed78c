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

__device__ __forceinline__ bool near_tie(float x, float dl) {
    // The raw sqrt is within 1 ulp (<= 2^-23 relative) of the correctly rounded
    // one, and each of the two products rounds by <= 2^-24, so x is within
    // 2^-22 * x of the correctly rounded x. Outside a window of 1.25 * 2^-22 * x
    // around a .5 tie both round to the same bin (the margin covers the
    // rounding of the threshold itself).
    return fabsf(dl) > __builtin_fmaf(x, -2.98023223876953125e-7f, 0.5f);
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

// fast-path result near a tie: redo the bin with the correctly rounded sqrt
__device__ __forceinline__ void tie_fix(float d2, float inv_w, float& x, float& b, float& dl) {
    const float r = sqrt_fix(d2, __builtin_amdgcn_sqrtf(d2));
    x = r * inv_w;
    b = __builtin_rintf(x);
    dl = x - b;
}

// The pair kernel's binning of one d2: the raw v_sqrt_f32 result, redone
// with the correctly rounded root within the tie window. The bin is the
// correctly rounded one; dl is the raw root's (within 1 ulp of x), so every
// path that bins pairs this way produces the same fixed-point offsets.
__device__ __forceinline__ void bin_fast(float d2, float inv_w, float& x, float& b, float& dl) {
    x = __builtin_amdgcn_sqrtf(d2) * inv_w;
    b = __builtin_rintf(x);
    dl = x - b;
    if (near_tie(x, dl)) tie_fix(d2, inv_w, x, b, dl);
}

// kLowBias + k with k = rne(dl * 2^20) + 2^19: dl * 2^20 + 1.5625 * 2^23 lies in
// [1.5 * 2^23, 1.625 * 2^23], where floats are the integers, so the FMA rounds
// dl * 2^20 to an integer once and the bit pattern is 0x4B400000 + k
__device__ __forceinline__ uint32_t quantize(float dl) {
    return __float_as_uint(__builtin_fmaf(dl, 1048576.0f, 13107200.0f));
}

}  // namespace ausaxs::gpu::binning
