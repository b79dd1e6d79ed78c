// Issue-rate probes: the device's own ceilings for the instruction classes
// the pair kernel is built from, measured on the device that runs it.
//
//   * FP32 adds whose operands are all vector registers (the pair kernel's
//     distance and binning arithmetic);
//   * the same adds with one scalar-register operand (gfx950 issues these at
//     half the rate);
//   * v_rsq_f32 (the transcendental of the fast root);
//   * ds_add_u64 at hashed cells of a workgroup-private LDS histogram, three
//     512-lane workgroups per CU (the pair kernel's accumulation).
//
// Each kernel keeps 8 independent chains per lane so that issue, not latency,
// bounds it; the rates are wave instructions per CU per nanosecond.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "device_api.hpp"

namespace ausaxs::gpu {
namespace {

constexpr int kProbeIter = 4096;  // loop trips per lane
constexpr int kChains = 8;

__device__ __forceinline__ float opaque_v(float x) {
    asm volatile("" : "+v"(x));
    return x;
}

// SGPR = true: the addend stays the kernel argument, a scalar register
template <bool SGPR>
__global__ __launch_bounds__(256) void probe_add_kernel(float* __restrict__ out, float b) {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (float)(threadIdx.x + k);
    const float bv = SGPR ? b : opaque_v(b);
    for (int i = 0; i < kProbeIter; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = a[k] + bv;
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void probe_rsq_kernel(float* __restrict__ out, float b) {
    float a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = (float)(threadIdx.x + k) + b;
    for (int i = 0; i < kProbeIter; ++i) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = __builtin_amdgcn_rsqf(a[k]);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// One ds_add_u64 per lane per trip at cell umulhi(h, nb) of an LCG h: cells
// in [0, nb), spread like the bins of a pair histogram's far tail.
__global__ __launch_bounds__(512) void probe_lds_kernel(float* __restrict__ out, int nb) {
    extern __shared__ uint64_t cells[];
    for (int e = threadIdx.x; e < nb; e += blockDim.x) cells[e] = 0;
    __syncthreads();
    uint32_t h = (threadIdx.x * 2654435761u) ^ (blockIdx.x * 40503u) ^ 0x9E3779B9u;
    for (int i = 0; i < kProbeIter; ++i) {
        h = h * 1664525u + 1013904223u;
        const uint32_t c = __umulhi(h, (uint32_t)nb);
        __hip_atomic_fetch_add(cells + c, (uint64_t(1) << 42) + (h >> 12), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    uint64_t s = 0;
    for (int e = threadIdx.x; e < nb; e += blockDim.x) s += cells[e];
    out[blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
}

}  // namespace

size_t probe_out_floats(int cus) { return (size_t)cus * 8 * 512; }

hipError_t launch_issue_probe(int which, int cus, int nbins, float* out, hipStream_t s, double* wave_instr) {
    switch (which) {
        case 0:
        case 1:
        case 2: {
            const int blocks = cus * 8;  // 8 waves per SIMD
            if (which == 0) hipLaunchKernelGGL(probe_add_kernel<false>, dim3(blocks), dim3(256), 0, s, out, 0.999f);
            else if (which == 1) hipLaunchKernelGGL(probe_add_kernel<true>, dim3(blocks), dim3(256), 0, s, out, 0.999f);
            else hipLaunchKernelGGL(probe_rsq_kernel, dim3(blocks), dim3(256), 0, s, out, 1.0f);
            *wave_instr = (double)blocks * 4 * kProbeIter * kChains;
            break;
        }
        case 3: {
            if (nbins < 1 || (size_t)nbins * sizeof(uint64_t) > 64 * 1024) return hipErrorInvalidValue;
            const int blocks = cus * 3;
            hipLaunchKernelGGL(probe_lds_kernel, dim3(blocks), dim3(512), (size_t)nbins * sizeof(uint64_t), s, out,
                               nbins);
            *wave_instr = (double)blocks * 8 * kProbeIter;
            break;
        }
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
