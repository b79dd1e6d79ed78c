// Pair-distance histogram kernels for gfx950 (CDNA4).
//
// Restates the reference's fused evaluate+increment (CompactCoordinatesXYZFF.h
// evaluate_{16,8,4,1}, HistogramManagerMTFFAvg.cpp:214-280) as:
//   * one workgroup per rectangle of the pair space (PairWork);
//   * row points in VGPRs (2 per lane); column points staged through LDS in
//     tiles of 256 and read back as wave-wide broadcasts, so each lane
//     evaluates 2 pairs per column point (LDS reads and LDS atomics retire in
//     order, so the broadcast read of the next column never waits on them);
//   * a workgroup-private histogram of packed uint64 cells in LDS updated
//     with ds_add_u64 (one atomic per pair);
//   * a flush of the LDS histogram to a partial slot, summed by merge_kernel.
//
// Bit-exact binning (SIMD-body rule of the reference):
//   d2 = (dx*dx + dy*dy) + dz*dz   separately rounded (file built with
//                                  -ffp-contract=off)
//   d  = sqrt_rn(d2), x = d*inv_w, bin = rint(x)  (round half to even)
// v_sqrt_f32 is accurate to 1 ulp but not correctly rounded; it is used when
// x is at least 4 ulp away from a .5 tie, otherwise the correctly rounded
// square root (sqrtf: v_sqrt_f32 plus an FMA-residual correction; note that
// __fsqrt_rn lowers to the uncorrected instruction) decides.
// launch_bin_check verifies both exhaustively on device.
#include "device_api.hpp"

namespace ausaxs::gpu {
namespace {

constexpr int kThreads = 256;
constexpr int kRowsPerLane = 2;
constexpr int kTileI = kThreads * kRowsPerLane;  // row points per work item
constexpr int kColTile = kThreads;               // column points per LDS tile

// correctly rounded square root (IEEE); see the note above
__device__ __forceinline__ float sqrt_rn(float x) { return sqrtf(x); }

__device__ __forceinline__ void simd_rule_fast(float dx, float dy, float dz, float inv_w, float& x, float& b,
                                               float& dl, float& d2) {
    const float sx = dx * dx;
    const float sy = dy * dy;
    const float sz = dz * dz;
    d2 = (sx + sy) + sz;
    const float r = __builtin_amdgcn_sqrtf(d2);
    x = r * inv_w;
    b = __builtin_rintf(x);
    dl = x - b;
}

__device__ __forceinline__ bool near_tie(float x, float dl) {
    // 2^-21 * x >= 4 ulp(x): the raw sqrt is within 1 ulp, the product adds < 3 ulp.
    return fabsf(dl) > __builtin_fmaf(x, -4.76837158203125e-7f, 0.5f);
}

__device__ __forceinline__ void simd_rule_exact(float d2, float inv_w, float& x, float& b, float& dl) {
    const float r = sqrt_rn(d2);
    x = r * inv_w;
    b = __builtin_rintf(x);
    dl = x - b;
}

// q + 2^19 with q = round(dl * 2^20); dl in [-0.5, 0.5] -> [0, 2^20]
__device__ __forceinline__ uint32_t quantize(float dl) {
    return (uint32_t)__builtin_fmaf(dl, 1048576.0f, 524288.5f);
}

__device__ __forceinline__ uint64_t cell_add(float dl) {
    return (uint64_t(1) << kCountShift) | (uint64_t)quantize(dl);
}

__device__ __forceinline__ void lds_add(uint64_t* cell, uint64_t v) {
    __hip_atomic_fetch_add(cell, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool FULL, bool DIAG, bool ORIENT>
__device__ __forceinline__ void run_item(const float4* __restrict__ pts, const PairWork& w, uint64_t* hist,
                                         float4* cols, float inv_w, int nbins) {
    const int t = threadIdx.x;
    float xi[kRowsPerLane], yi[kRowsPerLane], zi[kRowsPerLane];
    int tagi[kRowsPerLane], idx[kRowsPerLane];
    bool valid[kRowsPerLane];
#pragma unroll
    for (int k = 0; k < kRowsPerLane; ++k) {
        // lanes of one wave take points 8 apart: spatially spread rows give
        // spread distances to a common column point, so fewer lanes collide
        // on one LDS cell.
        const int li = k * kThreads + t;
        const int i = w.i_begin + ((li & 63) * 8 + (li >> 6));
        valid[k] = FULL || i < w.i_end;
        const float4 p = valid[k] ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        xi[k] = p.x;
        yi[k] = p.y;
        zi[k] = p.z;
        tagi[k] = __float_as_int(p.w);
        idx[k] = i;
    }
    const int stride = nbins + 2;  // slot nbins: overflow, slot nbins+1: masked lanes
    const int off1 = ORIENT ? stride : 0;
    const unsigned ovf = (unsigned)nbins;

    for (int jc = w.j_begin; jc < w.j_end; jc += kColTile) {
        const int nc = min(kColTile, w.j_end - jc);
        __syncthreads();  // previous tile consumed
        if (t < nc) cols[t] = pts[jc + t];
        __syncthreads();
        // the read of column c+1 is issued before the atomics of column c, so
        // waiting for it never waits for those atomics (LDS ops retire in order)
        float4 pn = cols[0];
        // completes the first read before the loop, so the loop header needs
        // no wait that would also drain the atomics in flight
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
            const float4 pj = pn;
            pn = cols[c + 1];  // cols has kColTile+1 slots; the last read is unused
            const int j = jc + c;
            const int tagj = __float_as_int(pj.w);
            float x[kRowsPerLane], b[kRowsPerLane], dl[kRowsPerLane], d2[kRowsPerLane];
            bool tie = false;
#pragma unroll
            for (int k = 0; k < kRowsPerLane; ++k) {
                simd_rule_fast(xi[k] - pj.x, yi[k] - pj.y, zi[k] - pj.z, inv_w, x[k], b[k], dl[k], d2[k]);
                tie |= near_tie(x[k], dl[k]);
            }
            if (tie) {
#pragma unroll
                for (int k = 0; k < kRowsPerLane; ++k)
                    if (near_tie(x[k], dl[k])) simd_rule_exact(d2[k], inv_w, x[k], b[k], dl[k]);
            }
#pragma unroll
            for (int k = 0; k < kRowsPerLane; ++k) {
                unsigned bin = min((unsigned)(int)b[k], ovf);
                if (!FULL || DIAG) {
                    bool keep = valid[k];
                    if (DIAG) keep = keep && (j > idx[k]);
                    bin = keep ? bin : ovf + 1u;
                }
                int off = 0;
                if (ORIENT) off = tagi[k] < tagj ? 0 : off1;
                lds_add(&hist[off + (int)bin], cell_add(dl[k]));
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void pair_hist_kernel(const float4* __restrict__ pts,
                                                             const PairWork* __restrict__ items,
                                                             uint64_t* __restrict__ partial, float inv_w,
                                                             int nbins) {
    extern __shared__ uint64_t hist[];
    __shared__ float4 cols[kColTile + 1];
    const PairWork w = items[blockIdx.x];
    const int stride = nbins + 2;
    const bool orient = (w.flags & kOrient) != 0;
    const int cells = (orient ? 2 : 1) * stride;
    for (int e = threadIdx.x; e < cells; e += kThreads) hist[e] = 0;
    __syncthreads();

    const bool full = (w.i_end - w.i_begin) >= kTileI;
    const bool diag = (w.flags & kDiag) != 0;
    if (diag) {
        if (orient) run_item<false, true, true>(pts, w, hist, cols, inv_w, nbins);
        else run_item<false, true, false>(pts, w, hist, cols, inv_w, nbins);
    } else if (full) {
        if (orient) run_item<true, false, true>(pts, w, hist, cols, inv_w, nbins);
        else run_item<true, false, false>(pts, w, hist, cols, inv_w, nbins);
    } else {
        if (orient) run_item<false, false, true>(pts, w, hist, cols, inv_w, nbins);
        else run_item<false, false, false>(pts, w, hist, cols, inv_w, nbins);
    }
    __syncthreads();

    // partial slot layout: [2][nbins+1] (the masked-lane slot is dropped)
    uint64_t* out = partial + (size_t)w.slot * 2 * (nbins + 1);
    const int out_stride = nbins + 1;
    for (int r = 0; r < (orient ? 2 : 1); ++r)
        for (int e = threadIdx.x; e < out_stride; e += kThreads) out[r * out_stride + e] = hist[r * stride + e];
}

__global__ __launch_bounds__(256) void merge_kernel(const uint64_t* __restrict__ partial,
                                                    const int32_t* __restrict__ row_ptr,
                                                    const int32_t* __restrict__ row_list, int nbins,
                                                    uint64_t* __restrict__ cnt, int64_t* __restrict__ dsum) {
    const int row = blockIdx.y;
    const int stride = nbins + 1;
    const int bin = blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= stride) return;
    uint64_t c = 0;
    int64_t s = 0;
    const int e0 = row_ptr[row], e1 = row_ptr[row + 1];
    for (int e = e0; e < e1; ++e) {
        const int ent = row_list[e];
        const uint64_t v = partial[(size_t)ent * stride + bin];
        const uint64_t n = v >> kCountShift;
        c += n;
        s += (int64_t)(v & kLowMask) - (int64_t)n * kDeltaOffset;
    }
    cnt[(size_t)row * stride + bin] = c;
    dsum[(size_t)row * stride + bin] = s;
}

// Strip kernel: rows x tail columns with the reference's per-pair rule.
__global__ __launch_bounds__(256) void strip_kernel(StripLaunch p) {
    const StripWork& w = p.w;
    const int nt = w.t_end - w.t_begin;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long long)w.n_rows * nt) return;
    const int i = (int)(tid / nt);
    const int t = w.t_begin + (int)(tid % nt);
    if (w.self && i >= t) return;
    const float4 a = p.rows[i];
    const float4 c = p.cols[t];
    const float dx = a.x - c.x, dy = a.y - c.y, dz = a.z - c.z;
    bool scalar_rule = true;
    if (w.self) scalar_rule = t >= w.n_cols_total - ((w.n_cols_total - i - 1) % 4);
    float x, b, dl;
    if (scalar_rule) {
        // evaluate_scalar with FMA contraction and round() lowered to
        // trunc(x + 0.49999997f) (CompactCoordinatesXYZFF.h:313-329)
        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
        x = sqrt_rn(d2) * p.inv_w;
        b = __builtin_truncf(x + 0.49999997f);
        dl = x - b;
    } else {
        const float d2 = (dx * dx + dy * dy) + dz * dz;
        simd_rule_exact(d2, p.inv_w, x, b, dl);
    }
    const int row = p.rowmap[__float_as_int(a.w) * w.n_col_classes + __float_as_int(c.w)];
    if (row < 0) return;  // the host guarantees a row for every present class pair
    const unsigned bin = min((unsigned)(int)b, (unsigned)p.nbins);
    const uint32_t low = quantize(dl);
    const size_t cell = (size_t)row * p.stride + bin;
    atomicAdd((unsigned long long*)&p.cnt[cell], 1ull);
    atomicAdd((unsigned long long*)&p.dsum[cell], (unsigned long long)((int64_t)low - kDeltaOffset));
}

__global__ __launch_bounds__(256) void bin_check_kernel(uint32_t lo, uint32_t hi, float inv_w,
                                                        unsigned long long* mism) {
    unsigned long long bad_fast = 0, bad_cr = 0;
    const uint64_t n = (uint64_t)hi - lo + 1;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = lo + (uint32_t)e;
        const float d2 = __uint_as_float(u);
        // correctly rounded reference, verified in FP64: r is the IEEE sqrt iff
        // (r - ulp/2)^2 <= d2 <= (r + ulp/2)^2 (exact in double: 50-bit products)
        const float r = sqrt_rn(d2);
        const double rd = (double)r;
        const double up = ((double)__uint_as_float(__float_as_uint(r) + 1) - rd) * 0.5;
        const double dn = r > 0.f ? (rd - (double)__uint_as_float(__float_as_uint(r) - 1)) * 0.5 : 0.0;
        const double lo2 = (rd - dn) * (rd - dn), hi2 = (rd + up) * (rd + up);
        const double dd = (double)d2;
        const bool even = (__float_as_uint(r) & 1u) == 0;
        const bool ok = even ? (lo2 <= dd && dd <= hi2) : (lo2 < dd && dd < hi2);
        if (!ok && d2 > 0.f) ++bad_cr;
        const float xr = r * inv_w;
        const float br = __builtin_rintf(xr);
        float x = __builtin_amdgcn_sqrtf(d2) * inv_w;
        float b = __builtin_rintf(x);
        float dl = x - b;
        if (near_tie(x, dl)) {
            float d2c = d2;
            simd_rule_exact(d2c, inv_w, x, b, dl);
        }
        if (b != br) ++bad_fast;
    }
    if (bad_fast) atomicAdd(&mism[0], bad_fast);
    if (bad_cr) atomicAdd(&mism[1], bad_cr);
}

// plain streaming copy of a known number of bytes (calibration of HBM traffic counters)
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

}  // namespace

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / sizeof(uint4);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
    return hipGetLastError();
}

hipError_t launch_pair_hist(const PairLaunch& p, hipStream_t s) {
    if (p.n_items == 0) return hipSuccess;
    const size_t lds = (size_t)2 * (p.nbins + 2) * sizeof(uint64_t);
    hipLaunchKernelGGL(pair_hist_kernel, dim3(p.n_items), dim3(kThreads), lds, s, p.pts, p.items, p.partial, p.inv_w,
                       p.nbins);
    return hipGetLastError();
}

hipError_t launch_merge(const uint64_t* partial, const int32_t* row_ptr, const int32_t* row_list, int n_rows,
                        int nbins, uint64_t* cnt, int64_t* dsum, hipStream_t s) {
    if (n_rows == 0) return hipSuccess;
    dim3 grid((nbins + 1 + 255) / 256, n_rows);
    hipLaunchKernelGGL(merge_kernel, grid, dim3(256), 0, s, partial, row_ptr, row_list, nbins, cnt, dsum);
    return hipGetLastError();
}

hipError_t launch_strip(const StripLaunch& p, hipStream_t s) {
    const long long n = (long long)p.w.n_rows * (p.w.t_end - p.w.t_begin);
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(strip_kernel, dim3(blocks), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_bin_check(uint32_t lo_bits, uint32_t hi_bits, float inv_w, unsigned long long* mism,
                            hipStream_t s) {
    hipLaunchKernelGGL(bin_check_kernel, dim3(4096), dim3(256), 0, s, lo_bits, hi_bits, inv_w, mism);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
