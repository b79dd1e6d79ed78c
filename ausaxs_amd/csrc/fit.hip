// Scale/offset least squares in FP64 (LinearLeastSquares::fit_params_only and
// Fitter::chi2) for many model curves at once, and the exact O(N^2) Debye
// sum (hist::exact_debye_transform) used to validate binned intensities.
#include "device_api.hpp"

namespace ausaxs::gpu {
namespace {

constexpr int kFitThreads = 256;

// fixed-order tree reduction of 5 FP64 sums over one workgroup
__device__ void block_sum(double (&v)[5], double (*sh)[kFitThreads]) {
    for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kFitThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 5; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    for (int k = 0; k < 5; ++k) v[k] = sh[k][0];
    __syncthreads();
}

// One workgroup per model curve: y ~ a x + b with weights 1/err^2
// (LinearLeastSquares.cpp:26-44), then chi2 = sum ((y - (a x + b)) / err)^2
// (Fitter.cpp:10-13 with LinearLeastSquares::get_residuals). out: a, b,
// var(a) = S/delta, var(b) = Sxx/delta, chi2.
__global__ __launch_bounds__(kFitThreads) void lsq_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                          const double* __restrict__ err, int n,
                                                          double* __restrict__ out) {
    __shared__ double sh[5][kFitThreads];
    const double* xm = x + (size_t)blockIdx.x * n;
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += kFitThreads) {
        const double is = err ? 1.0 / err[i] : 1.0;
        const double w = is * is;
        v[0] += w;
        v[1] += xm[i] * w;
        v[2] += y[i] * w;
        v[3] += (xm[i] * xm[i]) * w;
        v[4] += (xm[i] * y[i]) * w;
    }
    block_sum(v, sh);
    const double S = v[0], Sx = v[1], Sy = v[2], Sxx = v[3], Sxy = v[4];
    const double delta = S * Sxx - Sx * Sx;
    const double a = (S * Sxy - Sx * Sy) / delta;
    const double b = (Sxx * Sy - Sx * Sxy) / delta;
    double c[5] = {0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += kFitThreads) {
        const double is = err ? 1.0 / err[i] : 1.0;
        const double r = (y[i] - (a * xm[i] + b)) * is;
        c[0] += r * r;
    }
    block_sum(c, sh);
    if (threadIdx.x == 0) {
        double* o = out + (size_t)blockIdx.x * 5;
        o[0] = a;
        o[1] = b;
        o[2] = S / delta;
        o[3] = Sxx / delta;
        o[4] = c[0];
    }
}

// Exact Debye: workgroup (block of rows, q). Row i sums over j > i the
// contribution of 2 w_i w_j sinc(q d_ij) with d_ij from the reference's loop
// geometry (evaluate_8 / evaluate_4 SIMD rule, scalar FMA rule on the last
// (n - i - 1) mod 4 columns), plus w_i^2; the workgroup's rows are reduced in
// a fixed order into partial[q][block].
__global__ __launch_bounds__(kFitThreads) void exact_debye_kernel(const float4* __restrict__ p, int n,
                                                                  const double* __restrict__ q,
                                                                  double* __restrict__ partial) {
    __shared__ double sh[5][kFitThreads];
    const int i = blockIdx.x * kFitThreads + threadIdx.x;
    const double qk = q[blockIdx.y];
    double v[5] = {0, 0, 0, 0, 0};
    if (i < n) {
        const float4 a = p[i];
        const int tail = n - ((n - i - 1) % 4);
        double sum = 0;
        for (int j = i + 1; j < n; ++j) {
            const float4 c = p[j];
            const float dx = a.x - c.x, dy = a.y - c.y, dz = a.z - c.z;
            float d2;
            if (j < tail) d2 = (dx * dx + dy * dy) + dz * dz;
            else d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
            const float d = sqrtf(d2);  // correctly rounded
            const float w = 2 * (a.w * c.w);
            const double qr = qk * d;
            sum += qr < 1e-9 ? (double)w : w * sin(qr) / qr;
        }
        sum += (double)a.w * (double)a.w;
        v[0] = sum;
    }
    block_sum(v, sh);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

__global__ __launch_bounds__(256) void exact_debye_finish(const double* __restrict__ partial, int blocks,
                                                          const double* __restrict__ q, int nq,
                                                          double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nq) return;
    double s = 0;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)k * blocks + b];
    out[k] = s * exp(-q[k] * q[k]);
}

}  // namespace

hipError_t launch_lsq(const double* x, const double* y, const double* err, int n, int n_models, double* out,
                      hipStream_t s) {
    if (n_models <= 0) return hipSuccess;
    hipLaunchKernelGGL(lsq_kernel, dim3(n_models), dim3(kFitThreads), 0, s, x, y, err, n, out);
    return hipGetLastError();
}

int exact_debye_blocks(int n) { return (n + kFitThreads - 1) / kFitThreads; }

hipError_t launch_exact_debye(const float4* pts, int n, const double* q, int nq, double* partial, double* out,
                              hipStream_t s) {
    if (nq <= 0) return hipSuccess;
    const int blocks = exact_debye_blocks(n);
    if (blocks > 0) {
        hipLaunchKernelGGL(exact_debye_kernel, dim3(blocks, nq), dim3(kFitThreads), 0, s, pts, n, q, partial);
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(exact_debye_finish, dim3((nq + 255) / 256), dim3(256), 0, s, partial, blocks, q, nq, out);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
