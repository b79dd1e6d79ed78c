// Pair histograms of point sets on the integer lattice (gfx950).
//
// The grid's excluded-volume dummies sit on integer coordinates, so the
// distance of a dummy pair depends only on its displacement D = p_j - p_i,
// and every pair with the same D falls into the same bin with the same
// fixed-point offset: the reference's SIMD-body rule is exact integer
// arithmetic up to the square root (dx, dx*dx and the sums stay below 2^24).
// The histogram of a set's body pairs is then sum_D A(D) * [bin of |D|^2],
// with A the autocorrelation of the set's occupancy grid (the
// cross-correlation of two grids for a cross job). A is computed as
// IFFT(|FFT(occ)|^2) in FP64 on a grid padded to at least 2E - 1 per axis
// (E = extent), so that no displacement wraps. Its entries are pair counts;
// the FP64 transforms leave them within a small multiple of 2^-52 * n * log L
// of those integers, far below 1/2 for the sizes allowed, and the s-histogram
// kernel checks it (|A - rint(A)| <= 0.25, else a flag the host turns into an
// error). Counts and delta sums come out identical to evaluating every pair.
//
// FFT: radix-2 decimation in time on lines of length L (a power of two, at
// most 1024) held in LDS, G lines per workgroup (G * L <= 2048 complex values,
// 32 KB, so that several workgroups share a CU); loads run along the lines'
// common contiguous direction.
#include "binning.hpp"
#include "device_api.hpp"

#include <algorithm>

namespace ausaxs::gpu {
namespace {

using binning::bin_fast;
using binning::quantize;

__global__ __launch_bounds__(256) void lattice_zero_kernel(double2* __restrict__ buf, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        buf[e] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void lattice_zero_u64_kernel(unsigned long long* __restrict__ buf, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        buf[e] = 0ull;
}

// occupancy: +1 at every point (duplicates add up); the host has checked
// that every point is an integer inside the grid
__global__ __launch_bounds__(256) void lattice_fill_kernel(const float4* __restrict__ pts, int n, LatticeGeom g,
                                                           double2* __restrict__ buf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pts[i];
    const int ix = (int)p.x - g.origin[0], iy = (int)p.y - g.origin[1], iz = (int)p.z - g.origin[2];
    const size_t idx = ((size_t)ix * g.L[1] + iy) * g.L[2] + iz;
    atomicAdd(&buf[idx].x, 1.0);
}

// One FFT pass along an axis: lines of L elements at element stride S; line l
// starts at (l / S) * S * L + (l % S). tw[k] = exp(-2 pi i k / L), k < L / 2.
__global__ __launch_bounds__(256) void fft_lines_kernel(double2* __restrict__ buf, int L, int logL, long long S,
                                                        long long nlines, int G,
                                                        const double2* __restrict__ tw, int inverse) {
    extern __shared__ double2 sh[];  // [G][L]
    const long long l0 = (long long)blockIdx.x * G;
    const int total = G * L;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int g = S == 1 ? e / L : e % G;
        const int k = S == 1 ? e % L : e / G;
        const long long l = l0 + g;
        if (l >= nlines) continue;
        const long long base = (l / S) * S * L + (l % S);
        const int r = (int)(__brev((unsigned)k) >> (32 - logL));
        sh[g * L + r] = buf[base + (long long)k * S];
    }
    __syncthreads();
    const int half = L / 2;
    for (int h = 1; h < L; h <<= 1) {
        const int tstride = L / (2 * h);
        for (int p = threadIdx.x; p < G * half; p += blockDim.x) {
            const int g = p / half, q = p % half;
            const int j = q & (h - 1);
            const int i0 = g * L + (q / h) * 2 * h + j, i1 = i0 + h;
            double2 w = tw[j * tstride];
            if (inverse) w.y = -w.y;
            const double2 a = sh[i0], b = sh[i1];
            const double2 bw = make_double2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
            sh[i0] = make_double2(a.x + bw.x, a.y + bw.y);
            sh[i1] = make_double2(a.x - bw.x, a.y - bw.y);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int g = S == 1 ? e / L : e % G;
        const int k = S == 1 ? e % L : e / G;
        const long long l = l0 + g;
        if (l >= nlines) continue;
        const long long base = (l / S) * S * L + (l % S);
        buf[base + (long long)k * S] = sh[g * L + k];
    }
}

// self: a = |a|^2; cross: a = conj(a) * b (the transform of sum_v a(v) b(v + D))
__global__ __launch_bounds__(256) void lattice_spectrum_kernel(double2* __restrict__ a, const double2* __restrict__ b,
                                                               size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const double2 x = a[e];
        if (b) {
            const double2 y = b[e];
            a[e] = make_double2(x.x * y.x + x.y * y.y, x.x * y.y - x.y * y.x);
        } else {
            a[e] = make_double2(x.x * x.x + x.y * x.y, 0.0);
        }
    }
}

// Correlation entries -> pair counts per squared displacement s. Self jobs
// keep one of D and -D (the pair once) and, at D = 0, the pairs of coincident
// points, (A(0) - n) / 2.
__global__ __launch_bounds__(256) void lattice_shist_kernel(const double2* __restrict__ c, LatticeGeom g, double scale,
                                                            int self, long long n_self,
                                                            unsigned long long* __restrict__ scount, int smax,
                                                            int* __restrict__ flag) {
    const size_t total = (size_t)g.L[0] * g.L[1] * g.L[2];
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const double v = c[idx].x * scale;
        const double r = rint(v);
        if (fabs(v - r) > 0.25 || r < 0.0) {
            atomicOr(flag, 1);
            continue;
        }
        long long cnt = (long long)r;
        if (cnt == 0) continue;
        const int k = (int)(idx % g.L[2]);
        const size_t t = idx / g.L[2];
        const int j = (int)(t % g.L[1]);
        const int i = (int)(t / g.L[1]);
        const int dx = i <= g.L[0] / 2 ? i : i - g.L[0];
        const int dy = j <= g.L[1] / 2 ? j : j - g.L[1];
        const int dz = k <= g.L[2] / 2 ? k : k - g.L[2];
        if (self) {
            if (dx == 0 && dy == 0 && dz == 0) {
                cnt = (cnt - n_self) / 2;
                if (cnt <= 0) continue;
            } else if (!(dx > 0 || (dx == 0 && (dy > 0 || (dy == 0 && dz > 0))))) {
                continue;
            }
        }
        const int s = dx * dx + dy * dy + dz * dz;
        if (s > smax) {
            atomicOr(flag, 2);
            continue;
        }
        atomicAdd(&scount[s], (unsigned long long)cnt);
    }
}

// Pair counts per squared displacement -> one histogram row: bin and offset
// of d2 = s (exact in float below 2^24) as the pair kernel bins a body pair.
__global__ __launch_bounds__(256) void lattice_bins_kernel(const unsigned long long* __restrict__ scount, int smax,
                                                           float inv_w, int nbins, uint64_t* __restrict__ cnt,
                                                           int64_t* __restrict__ dsum) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= smax; s += gridDim.x * blockDim.x) {
        const unsigned long long c = scount[s];
        if (!c) continue;
        float x, b, dl;
        bin_fast((float)s, inv_w, x, b, dl);
        const unsigned bin = min((unsigned)(int)b, (unsigned)nbins);
        const int64_t k = (int64_t)(quantize(dl) - (uint32_t)kLowBias) - kDeltaOffset;
        atomicAdd((unsigned long long*)&cnt[bin], c);
        atomicAdd((unsigned long long*)&dsum[bin], (unsigned long long)((int64_t)c * k));
    }
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 65536); }

}  // namespace

hipError_t launch_lattice_zero(double2* buf, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(lattice_zero_kernel, dim3(grid_for(n)), dim3(256), 0, s, buf, n);
    return hipGetLastError();
}

hipError_t launch_lattice_zero_u64(unsigned long long* buf, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(lattice_zero_u64_kernel, dim3(grid_for(n)), dim3(256), 0, s, buf, n);
    return hipGetLastError();
}

hipError_t launch_lattice_fill(const float4* pts, int n, const LatticeGeom& g, double2* buf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(lattice_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pts, n, g, buf);
    return hipGetLastError();
}

hipError_t launch_fft3(double2* buf, const LatticeGeom& g, const double2* const tw[3], bool inverse, hipStream_t s) {
    const long long total = (long long)g.L[0] * g.L[1] * g.L[2];
    long long S = 1;
    for (int a = 2; a >= 0; --a) {
        const int L = g.L[a];
        if (L > 1) {
            if (L > kLatticeMaxL || (L & (L - 1))) return hipErrorInvalidValue;
            int logL = 0;
            while ((1 << logL) < L) ++logL;
            const int G = std::max(1, std::min(16, 2048 / L));
            const long long nlines = total / L;
            const unsigned blocks = (unsigned)((nlines + G - 1) / G);
            hipLaunchKernelGGL(fft_lines_kernel, dim3(blocks), dim3(256), (size_t)G * L * sizeof(double2), s, buf, L,
                               logL, S, nlines, G, tw[a], inverse ? 1 : 0);
            if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
        }
        S *= L;
    }
    return hipSuccess;
}

hipError_t launch_lattice_spectrum(double2* a, const double2* b, size_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(lattice_spectrum_kernel, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
    return hipGetLastError();
}

hipError_t launch_lattice_shist(const double2* c, const LatticeGeom& g, bool self, long long n_self,
                                unsigned long long* scount, int smax, int* flag, hipStream_t s) {
    const size_t total = (size_t)g.L[0] * g.L[1] * g.L[2];
    const double scale = 1.0 / (double)total;
    hipLaunchKernelGGL(lattice_shist_kernel, dim3(grid_for(total)), dim3(256), 0, s, c, g, scale, self ? 1 : 0,
                       n_self, scount, smax, flag);
    return hipGetLastError();
}

hipError_t launch_lattice_bins(const unsigned long long* scount, int smax, float inv_w, int nbins, uint64_t* cnt,
                               int64_t* dsum, hipStream_t s) {
    hipLaunchKernelGGL(lattice_bins_kernel, dim3(grid_for((size_t)smax + 1)), dim3(256), 0, s, scount, smax, inv_w,
                       nbins, cnt, dsum);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
