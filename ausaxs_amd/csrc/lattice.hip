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
// of those integers, far below 1/2 for the sizes allowed, and the count pass
// checks it (|A - rint(A)| <= 0.25, else a flag the host turns into an
// error). Counts and delta sums come out identical to evaluating every pair.
//
// FFT: decimation in time (radix-4 stages, one radix-2 stage when log2 L is
// odd) on lines of length L (a power of two, at most 1024) held in LDS, G
// lines per workgroup (G * L <= 2048 complex values, 32 KB, so that several
// workgroups share a CU); loads run along the lines' common contiguous
// direction. Lines known to be zero (forward: outside the
// occupied box) or not needed (inverse: leading only to displacements beyond
// the extents) are skipped; the forward transform's last pass stores the
// power (or cross) spectrum and the inverse transform's last pass turns its
// entries into pair counts, so neither needs a pass of its own.
#include "binning.hpp"
#include "device_api.hpp"

#include <algorithm>

namespace ausaxs::gpu {
namespace {

using binning::bin_of;
using binning::bin_pair;
using binning::Consts;
using binning::make_consts;

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

// One FFT pass over the lines of one axis and what its store does.
struct FftPass {
    double2* buf;
    const double2* tw;     // exp(-2 pi i k / L), k < L / 2
    const double2* other;  // store 2: the other transform
    int L, logL, G;        // line length, log2 L, lines per workgroup
    long long S, nlines;   // element stride; line l starts at (l / S) * S * L + (l % S)
    int inverse;
    // a line is computed iff keep(p, pe, pL) and keep(q, qe, qL), with
    // outer = l / S, p = outer / qdim, q = outer % qdim and
    // keep(v, e, Lv) = v < e or (Lv > 0 and v > Lv - e)
    int qdim, pe, pL, qe, qL;
    int store;  // 0: the transform; 1: |F|^2; 2: conj(F) * other; 3: pair counts
    // store 3 (the z pass): line l = x * Ly + y, element k = z
    int Lx, Ly, Ez;
    int self;
    long long n_self;
    double scale;
    unsigned long long* scount;
    int smax;
    int* flag;
};

__device__ __forceinline__ bool keep_coord(int v, int e, int Lv) { return v < e || (Lv > 0 && v > Lv - e); }

__global__ __launch_bounds__(256) void fft_lines_kernel(FftPass p) {
    extern __shared__ double2 sh[];  // [G][L]
    const int L = p.L, G = p.G;
    const long long S = p.S;
    const long long l0 = (long long)blockIdx.x * G;
    auto kept = [&](long long l) {
        if (l >= p.nlines) return false;
        const long long outer = l / S;
        const int pc = (int)(outer / p.qdim), qc = (int)(outer % p.qdim);
        return keep_coord(pc, p.pe, p.pL) && keep_coord(qc, p.qe, p.qL);
    };
    bool any = false;
    for (int g = 0; g < G; ++g) any = any || kept(l0 + g);
    if (!any) return;  // the same for every thread of the workgroup, before any barrier
    const int total = G * L;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int g = S == 1 ? e / L : e % G;
        const int k = S == 1 ? e % L : e / G;
        const long long l = l0 + g;
        if (!kept(l)) continue;
        const long long base = (l / S) * S * L + (l % S);
        const int r = (int)(__brev((unsigned)k) >> (32 - p.logL));
        sh[g * L + r] = p.buf[base + (long long)k * S];
    }
    __syncthreads();
    auto cmul = [](double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    int h = 1;
    if (p.logL & 1) {
        // one radix-2 stage first when log2 L is odd
        const int half = L / 2;
        for (int t = threadIdx.x; t < G * half; t += blockDim.x) {
            const int g = t / half, q = t % half;
            const int i0 = g * L + 2 * q, i1 = i0 + 1;
            const double2 a = sh[i0], b = sh[i1];
            sh[i0] = make_double2(a.x + b.x, a.y + b.y);
            sh[i1] = make_double2(a.x - b.x, a.y - b.y);
        }
        __syncthreads();
        h = 2;
    }
    // radix-4: the radix-2 stages of half-sizes h and 2h in registers, one LDS
    // round trip and one barrier for both. Stage h pairs (i0, i1), (i2, i3)
    // with W_2h^j; stage 2h pairs (i0, i2) with W_4h^j and (i1, i3) with
    // W_4h^(j+h) = W_4h^j * (-i) (conjugates for the inverse).
    const int quarter = L / 4;
    for (; h < L; h *= 4) {
        for (int t = threadIdx.x; t < G * quarter; t += blockDim.x) {
            const int g = t / quarter, q = t % quarter;
            const int j = q & (h - 1);
            const int i0 = g * L + (q / h) * 4 * h + j, i1 = i0 + h, i2 = i0 + 2 * h, i3 = i0 + 3 * h;
            double2 w1 = p.tw[j * (L / (2 * h))], w2 = p.tw[j * (L / (4 * h))];
            if (p.inverse) w1.y = -w1.y, w2.y = -w2.y;
            const double2 w3 = p.inverse ? make_double2(-w2.y, w2.x) : make_double2(w2.y, -w2.x);
            const double2 a0 = sh[i0], a2 = sh[i2];
            const double2 a1 = cmul(sh[i1], w1), a3 = cmul(sh[i3], w1);
            const double2 b0 = make_double2(a0.x + a1.x, a0.y + a1.y), b1 = make_double2(a0.x - a1.x, a0.y - a1.y);
            const double2 b2 = cmul(make_double2(a2.x + a3.x, a2.y + a3.y), w2);
            const double2 b3 = cmul(make_double2(a2.x - a3.x, a2.y - a3.y), w3);
            sh[i0] = make_double2(b0.x + b2.x, b0.y + b2.y);
            sh[i2] = make_double2(b0.x - b2.x, b0.y - b2.y);
            sh[i1] = make_double2(b1.x + b3.x, b1.y + b3.y);
            sh[i3] = make_double2(b1.x - b3.x, b1.y - b3.y);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int g = S == 1 ? e / L : e % G;
        const int k = S == 1 ? e % L : e / G;
        const long long l = l0 + g;
        if (!kept(l)) continue;
        const long long idx = (l / S) * S * L + (l % S) + (long long)k * S;
        const double2 v = sh[g * L + k];
        if (p.store == 0) {
            p.buf[idx] = v;
        } else if (p.store == 1) {
            p.buf[idx] = make_double2(v.x * v.x + v.y * v.y, 0.0);
        } else if (p.store == 2) {
            // the transform of sum_v a(v) b(v + D): conj(F_a) * F_b
            const double2 o = p.other[idx];
            p.buf[idx] = make_double2(v.x * o.x + v.y * o.y, v.x * o.y - v.y * o.x);
        } else {
            if (!keep_coord(k, p.Ez, L)) continue;
            const double c = v.x * p.scale;
            const double r = rint(c);
            if (fabs(c - r) > 0.25 || r < 0.0) {
                atomicOr(p.flag, 1);
                continue;
            }
            long long cnt = (long long)r;
            if (cnt == 0) continue;
            const int xi = (int)(l / p.Ly), yi = (int)(l % p.Ly);
            const int dx = xi <= p.Lx / 2 ? xi : xi - p.Lx;
            const int dy = yi <= p.Ly / 2 ? yi : yi - p.Ly;
            const int dz = k <= L / 2 ? k : k - L;
            if (p.self) {
                // one of D and -D (the pair once); at D = 0 the pairs of
                // coincident points, (A(0) - n) / 2
                if (dx == 0 && dy == 0 && dz == 0) {
                    cnt = (cnt - p.n_self) / 2;
                    if (cnt <= 0) continue;
                } else if (!(dx > 0 || (dx == 0 && (dy > 0 || (dy == 0 && dz > 0))))) {
                    continue;
                }
            }
            const int s2 = dx * dx + dy * dy + dz * dz;
            if (s2 > p.smax) {
                atomicOr(p.flag, 2);
                continue;
            }
            atomicAdd(&p.scount[s2], (unsigned long long)cnt);
        }
    }
}

// Pair counts per squared displacement -> one histogram row: bin and offset
// of d2 = s (exact in float below 2^24) as the pair kernel bins a body pair
// (on d2 scaled by the pair loop's power of four).
__global__ __launch_bounds__(256) void lattice_bins_kernel(const unsigned long long* __restrict__ scount, int smax,
                                                           float inv_w, int nbins, uint64_t* __restrict__ cnt,
                                                           int64_t* __restrict__ dsum) {
    const Consts cs = make_consts(inv_w);
    const float sc2 = cs.scale * cs.scale;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= smax; s += gridDim.x * blockDim.x) {
        const unsigned long long c = scount[s];
        if (!c) continue;
        float t;
        const uint32_t code = bin_pair((float)s * sc2, cs, t);
        const unsigned bin = min(bin_of(t), (unsigned)nbins);
        const int64_t k = (int64_t)(code - (uint32_t)kLowBias) - kDeltaBias;
        atomicAdd((unsigned long long*)&cnt[bin], c);
        atomicAdd((unsigned long long*)&dsum[bin], (unsigned long long)((int64_t)c * k));
    }
}

unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>((n + 255) / 256, 65536); }

// one pass along axis a of a grid with element stride S on that axis
hipError_t fft_pass(FftPass p, const LatticeGeom& g, int a, long long S, hipStream_t s) {
    const long long total = (long long)g.L[0] * g.L[1] * g.L[2];
    const int L = g.L[a];
    if (L < 2 || L > kLatticeMaxL || (L & (L - 1))) return hipErrorInvalidValue;
    p.L = L;
    p.logL = 0;
    while ((1 << p.logL) < L) ++p.logL;
    p.G = std::max(1, std::min(16, 2048 / L));
    p.S = S;
    p.nlines = total / L;
    const unsigned blocks = (unsigned)((p.nlines + p.G - 1) / p.G);
    hipLaunchKernelGGL(fft_lines_kernel, dim3(blocks), dim3(256), (size_t)p.G * L * sizeof(double2), s, p);
    return hipGetLastError();
}

FftPass keep_all(double2* buf, const double2* tw) {
    FftPass p{};
    p.buf = buf;
    p.tw = tw;
    p.qdim = 1, p.pe = 1 << 30, p.pL = 0, p.qe = 1, p.qL = 0;
    return p;
}

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

hipError_t launch_fft3_forward(double2* buf, const LatticeGeom& g, const double2* const tw[3], int store,
                               const double2* other, hipStream_t s) {
    for (int a = 0; a < 3; ++a)
        if (g.L[a] < 2) return hipErrorInvalidValue;
    if (store < 0 || store > 2 || (store == 2 && !other)) return hipErrorInvalidValue;
    const long long S[3] = {(long long)g.L[1] * g.L[2], g.L[2], 1};
    // z pass: the occupancy is zero outside x < Ex, y < Ey
    FftPass pz = keep_all(buf, tw[2]);
    pz.qdim = g.L[1], pz.pe = g.E[0], pz.qe = g.E[1];
    if (hipError_t e = fft_pass(pz, g, 2, S[2], s); e != hipSuccess) return e;
    // y pass: zero outside x < Ex
    FftPass py = keep_all(buf, tw[1]);
    py.pe = g.E[0];
    if (hipError_t e = fft_pass(py, g, 1, S[1], s); e != hipSuccess) return e;
    // x pass: every line, with the requested store
    FftPass px = keep_all(buf, tw[0]);
    px.store = store;
    px.other = other;
    return fft_pass(px, g, 0, S[0], s);
}

hipError_t launch_fft3_counts(double2* buf, const LatticeGeom& g, const double2* const tw[3], bool self,
                              long long n_self, unsigned long long* scount, int smax, int* flag, hipStream_t s) {
    for (int a = 0; a < 3; ++a)
        if (g.L[a] < 2) return hipErrorInvalidValue;
    const long long S[3] = {(long long)g.L[1] * g.L[2], g.L[2], 1};
    // x pass: every line
    FftPass px = keep_all(buf, tw[0]);
    px.inverse = 1;
    if (hipError_t e = fft_pass(px, g, 0, S[0], s); e != hipSuccess) return e;
    // y pass: the lines whose x leads to |Dx| < Ex
    FftPass py = keep_all(buf, tw[1]);
    py.inverse = 1;
    py.pe = g.E[0], py.pL = g.L[0];
    if (hipError_t e = fft_pass(py, g, 1, S[1], s); e != hipSuccess) return e;
    // z pass: the lines with |Dx| < Ex and |Dy| < Ey, entries with |Dz| < Ez to counts
    FftPass pz = keep_all(buf, tw[2]);
    pz.inverse = 1;
    pz.qdim = g.L[1], pz.pe = g.E[0], pz.pL = g.L[0], pz.qe = g.E[1], pz.qL = g.L[1];
    pz.store = 3;
    pz.Lx = g.L[0], pz.Ly = g.L[1], pz.Ez = g.E[2];
    pz.self = self ? 1 : 0;
    pz.n_self = n_self;
    pz.scale = 1.0 / ((double)g.L[0] * g.L[1] * g.L[2]);
    pz.scount = scount;
    pz.smax = smax;
    pz.flag = flag;
    return fft_pass(pz, g, 2, S[2], s);
}

hipError_t launch_lattice_bins(const unsigned long long* scount, int smax, float inv_w, int nbins, uint64_t* cnt,
                               int64_t* dsum, hipStream_t s) {
    hipLaunchKernelGGL(lattice_bins_kernel, dim3(grid_for((size_t)smax + 1)), dim3(256), 0, s, scount, smax, inv_w,
                       nbins, cnt, dsum);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
