// Debye contraction on MFMA for gfx950 (CDNA4).
//
//   out[r][k] = sum_b H[r][b] * sinc(q_k * d_b)
//
// restates CompositeDistanceHistogramFFAvgBase::cache_refresh_sinqd
// (CompositeDistanceHistogramFFAvgBase.cpp:427-467, std::inner_product over a
// VectorDebyeTable, VectorDebyeTable.cpp:24-38) as a GEMM with
// M = q points, N = histogram rows, K = distance bins.
//
// Operands are split into bf16 terms, x = x0 + x1 (+ x2), each term holding the
// next 8 significant bits; products of terms with t_a + t_b < terms are summed
// by v_mfma_f32_16x16x32_bf16 in FP32 (bf16 x bf16 products are exact in
// FP32), and every 32-wide K step is added to an FP64 accumulator, so FP32
// rounding never spans more than one K step. terms = 3 carries ~24 bits per
// operand (the FP32-level split); terms = 2 is the hi+lo split.
#include "device_api.hpp"

#include <algorithm>
#include <mutex>

namespace ausaxs::gpu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMaxTerms = 3;

__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ double bf16_val(uint16_t h) { return (double)__uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split_store(double v, int terms, uint16_t* dst, size_t plane) {
    double r = v;
    for (int t = 0; t < terms; ++t) {
        const uint16_t h = bf16_rne((float)r);
        dst[t * plane] = h;
        r -= bf16_val(h);
    }
}

// sinc table (VectorDebyeTable.cpp:24-38) split into bf16 planes [terms][nq_pad][K_pad]
__device__ __forceinline__ void sinc_split(const double* __restrict__ q, const double* __restrict__ d, int nq, int K,
                                           int nq_pad, int K_pad, int terms, uint16_t* __restrict__ A, size_t e) {
    const size_t plane = (size_t)nq_pad * K_pad;
    if (e >= plane) return;
    const int k = (int)(e / K_pad), b = (int)(e % K_pad);
    double s = 0.0;
    if (k < nq && b < K) {
        const double qd = q[k] * d[b];
        if (qd < 1e-3) {
            const double qd2 = qd * qd;
            s = 1 - qd2 * (1. / 6) + qd2 * qd2 * (1. / 120);
        } else {
            s = sin(qd) / qd;
        }
    }
    split_store(s, terms, A + e, plane);
}

// histogram rows split into bf16 planes [terms][rows_pad][K_pad]
__device__ __forceinline__ void h_split(const double* __restrict__ H, int rows, int K, int rows_pad, int K_pad,
                                        int terms, uint16_t* __restrict__ B, size_t e) {
    const size_t plane = (size_t)rows_pad * K_pad;
    if (e >= plane) return;
    const int r = (int)(e / K_pad), b = (int)(e % K_pad);
    const double v = (r < rows && b < K) ? H[(size_t)r * K + b] : 0.0;
    split_store(v, terms, B + e, plane);
}

// Both operand splits in one launch: blocks [0, a_blocks) take the sinc
// table, the rest the histogram rows (one launch less per contraction: on a
// device shared with other streams' long kernels, every launch waits for free
// compute units).
__global__ __launch_bounds__(256) void debye_split_kernel(const double* __restrict__ q, const double* __restrict__ d,
                                                          const double* __restrict__ H, int nq, int rows, int K,
                                                          int nq_pad, int rows_pad, int K_pad, int terms,
                                                          unsigned a_blocks, uint16_t* __restrict__ A,
                                                          uint16_t* __restrict__ B) {
    if (blockIdx.x < a_blocks)
        sinc_split(q, d, nq, K, nq_pad, K_pad, terms, A, (size_t)blockIdx.x * blockDim.x + threadIdx.x);
    else
        h_split(H, rows, K, rows_pad, K_pad, terms, B, (size_t)(blockIdx.x - a_blocks) * blockDim.x + threadIdx.x);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, v);
}

// One workgroup per 16 (q) x 16 (rows) output tile, one wave per K chunk
// (wave w of `chunks` takes K steps [w k_steps, (w + 1) k_steps)): the waves'
// FP64 tiles meet in LDS and wave 0 sums them in chunk order, so enough waves
// run to fill the chip, the result is deterministic, and the contraction is
// one launch. Lane l feeds row/column l%16 and K sub-block l/16 (8
// consecutive bf16); the accumulator holds D[4*(l/16)+i][l%16]. The term count
// is a template parameter: with it known at compile time the operand registers
// are indexed statically, and a step's loads are all in flight at once (the
// next step's are issued before this step's MFMAs).
constexpr int kMaxChunks = 16;
template <int TERMS>
__global__ __launch_bounds__(64 * kMaxChunks) void debye_mfma_kernel(const uint16_t* __restrict__ A,
                                                                     const uint16_t* __restrict__ B, int nq, int rows,
                                                                     int nq_pad, int rows_pad, int K_pad, int k_steps,
                                                                     double* __restrict__ out) {
    __shared__ double tiles[kMaxChunks][64][4];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q0 = blockIdx.x * 16, r0 = blockIdx.y * 16;
    const int m = l & 15, kg = l >> 4;
    const size_t pa = (size_t)nq_pad * K_pad, pb = (size_t)rows_pad * K_pad;
    const uint16_t* arow = A + (size_t)(q0 + m) * K_pad + kg * 8;
    const uint16_t* brow = B + (size_t)(r0 + m) * K_pad + kg * 8;
    double acc64[4] = {0, 0, 0, 0};
    const int kb = w * k_steps * 32, ke = min(K_pad, kb + k_steps * 32);
    if (kb < ke) {
        bf16x8 a[TERMS], b[TERMS];
#pragma unroll
        for (int t = 0; t < TERMS; ++t) {
            a[t] = load8(arow + t * pa + kb);
            b[t] = load8(brow + t * pb + kb);
        }
        for (int k0 = kb; k0 < ke; k0 += 32) {
            // the next step's operands (the last step reloads its own: in bounds, unused)
            const int kn = k0 + 32 < ke ? k0 + 32 : k0;
            bf16x8 na[TERMS], nb[TERMS];
#pragma unroll
            for (int t = 0; t < TERMS; ++t) {
                na[t] = load8(arow + t * pa + kn);
                nb[t] = load8(brow + t * pb + kn);
            }
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            // smallest products first, largest last
            if constexpr (TERMS >= 3) {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
            }
            if constexpr (TERMS >= 2) {
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) acc64[i] += (double)acc[i];
#pragma unroll
            for (int t = 0; t < TERMS; ++t) a[t] = na[t], b[t] = nb[t];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) tiles[w][l][i] = acc64[i];
    __syncthreads();
    if (w != 0) return;
    const int chunks = blockDim.x >> 6;
    const int r = r0 + m;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = 0;
        for (int c = 0; c < chunks; ++c) s += tiles[c][l][i];
        const int k = q0 + 4 * kg + i;
        if (r < rows && k < nq) out[(size_t)r * nq + k] = s;
    }
}

__global__ __launch_bounds__(256) void debye_f64_kernel(const double* __restrict__ H, const double* __restrict__ d,
                                                        const double* __restrict__ q, int rows, int K, int nq,
                                                        double* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * nq) return;
    const int r = e / nq, k = e % nq;
    double s = 0.0;
    for (int b = 0; b < K; ++b) {
        const double qd = q[k] * d[b];
        double v;
        if (qd < 1e-3) {
            const double qd2 = qd * qd;
            v = 1 - qd2 * (1. / 6) + qd2 * qd2 * (1. / 120);
        } else {
            v = sin(qd) / qd;
        }
        s += H[(size_t)r * K + b] * v;
    }
    out[e] = s;
}

inline int pad_to(int x, int m) { return (x + m - 1) / m * m; }

}  // namespace

// K chunks per output tile (waves of its workgroup): enough waves for ~8
// per CU on 256 CUs, at most kMaxChunks
int k_chunks(int rows, int K, int nq) {
    const int tiles = (pad_to(nq, 16) / 16) * (pad_to(rows, 16) / 16);
    const int steps = pad_to(K, 32) / 32;
    return std::max(1, std::min({steps, kMaxChunks, (2048 + tiles - 1) / std::max(tiles, 1)}));
}

// Work space of the allocation-free contraction: the split operands of one
// call. Every location a call reads it writes first in the same call
// (debye_split fills both padded planes), so nothing needs clearing between
// calls; calls that overlap on one GPU race on it.
static __device__ __align__(256) unsigned char scratch[kDebyeNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // per device: the device address of `scratch`
}  // namespace

// In host code `scratch` names a host-side stand-in; its address on the
// current device comes from hipGetSymbolAddress, once per device, and only
// that address is passed to kernels.
hipError_t debye_noalloc_scratch(void** ptr) {
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev); e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)); e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

size_t debye_scratch_bytes(int rows, int K, int nq) {
    const size_t K_pad = pad_to(K, 32), nq_pad = pad_to(nq, 16), rows_pad = pad_to(rows, 16);
    const size_t split = (size_t)kMaxTerms * (nq_pad + rows_pad) * K_pad * sizeof(uint16_t);
    return (split + 255) / 256 * 256;
}

hipError_t launch_debye(const DebyeLaunch& p, hipStream_t s) {
    if (p.rows == 0 || p.nq == 0) return hipSuccess;
    if (p.terms < 1 || p.terms > kMaxTerms) return hipErrorInvalidValue;
    const int K_pad = pad_to(p.K, 32), nq_pad = pad_to(p.nq, 16), rows_pad = pad_to(p.rows, 16);
    uint16_t* A = static_cast<uint16_t*>(p.scratch);
    uint16_t* B = A + (size_t)kMaxTerms * nq_pad * K_pad;
    const size_t na = (size_t)nq_pad * K_pad, nb = (size_t)rows_pad * K_pad;
    const int chunks = k_chunks(p.rows, p.K, p.nq);
    const int steps = K_pad / 32;
    const int per = (steps + chunks - 1) / chunks;
    const int used = (steps + per - 1) / per;  // chunks that hold at least one K step
    const unsigned a_blocks = (unsigned)((na + 255) / 256), b_blocks = (unsigned)((nb + 255) / 256);
    hipLaunchKernelGGL(debye_split_kernel, dim3(a_blocks + b_blocks), dim3(256), 0, s, p.q, p.d, p.H, p.nq, p.rows,
                       p.K, nq_pad, rows_pad, K_pad, p.terms, a_blocks, A, B);
    // one workgroup per output tile, one wave per K chunk that holds a step
    const dim3 grid(nq_pad / 16, rows_pad / 16);
    const dim3 block(64 * used);
    if (p.terms == 3)
        hipLaunchKernelGGL(debye_mfma_kernel<3>, grid, block, 0, s, A, B, p.nq, p.rows, nq_pad, rows_pad, K_pad, per,
                           p.out);
    else if (p.terms == 2)
        hipLaunchKernelGGL(debye_mfma_kernel<2>, grid, block, 0, s, A, B, p.nq, p.rows, nq_pad, rows_pad, K_pad, per,
                           p.out);
    else
        hipLaunchKernelGGL(debye_mfma_kernel<1>, grid, block, 0, s, A, B, p.nq, p.rows, nq_pad, rows_pad, K_pad, per,
                           p.out);
    return hipGetLastError();
}

hipError_t launch_debye_f64(const DebyeLaunch& p, hipStream_t s) {
    const int n = p.rows * p.nq;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(debye_f64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p.H, p.d, p.q, p.rows, p.K, p.nq,
                       p.out);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
