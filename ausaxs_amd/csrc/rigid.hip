// Kernels of the batched rigid-body scan (PartialHistogramManagerMT on many
// conformations at once): body transforms, per-conformation finalisation of
// the partial histograms and the Debye sum on each conformation's own
// weighted distance axis.
#include "rigid_kernels.hpp"

namespace ausaxs::gpu {
namespace {

// x' = float(((R00 x + R01 y) + R02 z) + t0), likewise y, z, in FP64 with no
// contraction (this file is built with -ffp-contract=off): the float points
// a CompactCoordinates rebuild of the moved FP64 atoms holds.
__global__ __launch_bounds__(256) void transform_kernel(TransformLaunch p) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (k >= p.n_slots) return;
    const int b = p.slot_body[k];
    const bool moving = p.body_moving[b] != 0;
    if (!moving && c > 0) return;  // static bodies: one shared copy
    const double* t = p.T + ((size_t)c * p.n_bodies + b) * 12;
    const int src = p.slot_src[k];
    const double x = p.xyz[3 * (size_t)src], y = p.xyz[3 * (size_t)src + 1], z = p.xyz[3 * (size_t)src + 2];
    const float vx = (float)(((t[0] * x + t[1] * y) + t[2] * z) + t[9]);
    const float vy = (float)(((t[3] * x + t[4] * y) + t[5] * z) + t[10]);
    const float vz = (float)(((t[6] * x + t[7] * y) + t[8] * z) + t[11]);
    const size_t off = moving ? (size_t)c * p.rep_stride : 0;
    p.pts[p.slot_dst[k] + off] = make_float4(vx, vy, vz, __int_as_float(p.slot_tag[k]));
    p.orig[p.slot_dsto[k] + off] = make_float4(vx, vy, vz, __int_as_float(p.slot_cls[k]));
    if (p.pts_w) {
        p.pts_w[p.slot_dst[k] + off] = make_float4(vx, vy, vz, p.slot_w[k]);
        p.orig_w[p.slot_dsto[k] + off] = make_float4(vx, vy, vz, p.slot_w[k]);
    }
}

// One workgroup per conformation: aa/aw/ww/p_tot entries per bin from the
// conformation's rows, the shared rows and the cached self terms (and, with
// per-conformation waters, the water set's own self term); the
// weighted axis, H = aa + cw aw + cw^2 ww, max_bin and the overflow flag.
__global__ __launch_bounds__(256) void finalize_kernel(FinalizeLaunch p) {
    const int c = blockIdx.x;
    const int stride = p.nbins + 1;
    __shared__ int s_max[256];
    __shared__ int s_ovf[256];
    int local_max = 10, ovf = 0;
    for (int b = threadIdx.x; b <= p.nbins; b += blockDim.x) {
        if (b == p.nbins) {
            for (int r = 0; r < p.R; ++r) {
                ovf |= p.cnt[((size_t)c * p.R + r) * stride + b] != 0;
                ovf |= p.cnt[((size_t)p.n_conf * p.R + r) * stride + b] != 0;
            }
            continue;
        }
        double aa_v = 0, aa_c = 0, aa_d = 0, aw_v = 0, aw_c = 0, aw_d = 0, ww_v = 0, ww_c = 0, ww_d = 0;
        if (b < p.base_nb) {
            aa_v = p.base[b];
            aa_c = p.base[p.base_nb + b];
            aa_d = p.base[2 * p.base_nb + b];
            ww_v = p.base[3 * p.base_nb + b];
            ww_c = p.base[4 * p.base_nb + b];
            ww_d = p.base[5 * p.base_nb + b];
        }
        for (int rep = 0; rep < 2; ++rep) {
            const size_t row0 = rep == 0 ? (size_t)c * p.R : (size_t)p.n_conf * p.R;
            for (int r = 0; r < p.R; ++r) {
                const size_t cell = (row0 + r) * stride + b;
                const uint64_t n = p.cnt[cell];
                if (!n) continue;
                // add<2> (SimpleCalculator.h:195-207): value 2 n w_i w_j, count 2 n, 2 sum d
                const double nv = 2.0 * (double)n;
                const double d = 2.0 * distance_sum(n, b, p.dsum[cell], (float)p.inv_w);
                const double v = p.wsum ? 2.0 * weight_value(p.wsum[cell], p.wstrip[cell], p.wscale) : nv * p.row_w[r];
                const int kind = p.row_kind[r];
                if (kind == 0) {
                    aa_v += v;
                    aa_c += nv;
                    aa_d += d;
                } else if (kind == 1) {
                    aw_v += v;
                    aw_c += nv;
                    aw_d += d;
                } else {
                    ww_v += v;
                    ww_c += nv;
                    ww_d += d;
                }
            }
        }
        if (b == 0 && p.ww_self) {
            ww_v += p.ww_self[2 * c];
            ww_c += p.ww_self[2 * c + 1];
        }
        const double tv = (aa_v + aw_v) + ww_v, tc = (aa_c + aw_c) + ww_c, td = (aa_d + aw_d) + ww_d;
        double dbar = b * p.width;
        if (p.weighted) dbar = ((double)(td == 0) * dbar + td) / (double)((tc == 0) + tc);
        if (b == 0) dbar = 0;
        const size_t o = (size_t)c * p.nbins + b;
        p.dbar[o] = dbar;
        p.H[o] = aa_v + p.cw * aw_v + p.cw * p.cw * ww_v;
        if (b >= 10 && tv != 0) local_max = max(local_max, b + 1);
    }
    s_max[threadIdx.x] = local_max;
    s_ovf[threadIdx.x] = ovf;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_max[threadIdx.x] = max(s_max[threadIdx.x], s_max[threadIdx.x + w]);
            s_ovf[threadIdx.x] |= s_ovf[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        p.max_bin[c] = s_max[0];
        p.flags[c] = s_ovf[0];
    }
}

// I_c(q) = exp(-q^2) sum_{b < max_bin_c} H_c[b] sinc(q dbar_c[b]) in FP64
// (DistanceHistogram.cpp:54-87; sinc as VectorDebyeTable.cpp:24-38).
__global__ __launch_bounds__(64) void debye_batch_kernel(DebyeBatchLaunch p) {
    const int c = blockIdx.x;
    const int k = blockIdx.y * blockDim.x + threadIdx.x;
    if (k >= p.nq) return;
    const double q = p.q[k];
    const int mb = p.max_bin[c];
    const double* H = p.H + (size_t)c * p.nbins;
    const double* d = p.dbar + (size_t)c * p.nbins;
    double acc = 0;
    for (int b = 0; b < mb; ++b) {
        const double x = q * d[b];
        double s;
        if (x < 1e-3) {
            const double x2 = x * x;
            s = 1 - x2 * (1.0 / 6) + x2 * x2 * (1.0 / 120);
        } else {
            s = sin(x) / x;
        }
        acc += H[b] * s;
    }
    p.I[(size_t)c * p.nq + k] = acc * exp(-q * q);
}

}  // namespace

hipError_t launch_transform(const TransformLaunch& p, hipStream_t s) {
    if (p.n_slots == 0 || p.n_conf == 0) return hipSuccess;
    hipLaunchKernelGGL(transform_kernel, dim3((p.n_slots + 255) / 256, p.n_conf), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_finalize(const FinalizeLaunch& p, hipStream_t s) {
    if (p.n_conf == 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_kernel, dim3(p.n_conf), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_debye_batch(const DebyeBatchLaunch& p, hipStream_t s) {
    if (p.n_conf == 0 || p.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(debye_batch_kernel, dim3(p.n_conf, (p.nq + 63) / 64), dim3(64), 0, s, p);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
