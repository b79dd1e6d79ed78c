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
// FFT: mixed-radix Stockham (radix 4, 2, 3 and 5 stages; natural order in
// and out, no bit reversal) on lines of length L = 2^a 3^b 5^c (at most
// kLatticeMaxL), the smallest such length >= 2E - 1, so the padded grid is at
// most ~1.2x the occupied box per axis instead of up to 2x for powers of two.
// G lines of a pass are held in LDS (G * L <= kFftLdsElems complex values);
// each stage reads its butterflies' inputs into registers, waits, and writes
// their outputs in place. Loads run along the lines' common contiguous
// direction. Each pass reads only what can be non-zero and writes only what a
// later pass reads:
//   forward z: lines x < Ex, y < Ey; reads z < Ez (the filled box)
//   forward y: lines x < Ex; reads y < Ey
//   x (fused): every line; reads x < Ex, transforms, forms the power (or
//              cross) spectrum, transforms back, writes the displacements
//              the correlation needs (0 <= Dx < Ex for one set, |Dx| < Ex for
//              two)
//   inverse y: lines with those Dx; writes |Dy| < Ey
//   inverse z: lines with those Dx and |Dy| < Ey; entries |Dz| < Ez become
//              pair counts per squared displacement.
// So only the occupied box needs clearing before the fill.
#include "binning.hpp"
#include "device_api.hpp"

#include <algorithm>

namespace ausaxs::gpu {
namespace {

using binning::bin_of;
using binning::bin_pair;
using binning::Consts;
using binning::make_consts;

// clears the occupied box x < Ex, y < Ey, z < Ez of a grid (the only part
// the fill and the forward z pass read)
__global__ __launch_bounds__(256) void lattice_zero_box_kernel(double2* __restrict__ buf, LatticeGeom g) {
    const size_t n = (size_t)g.E[0] * g.E[1] * g.E[2];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const size_t z = e % g.E[2], xy = e / g.E[2];
        const size_t y = xy % g.E[1], x = xy / g.E[1];
        buf[(x * g.L[1] + y) * g.L[2] + z] = make_double2(0.0, 0.0);
    }
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

constexpr int kFftThreads = 256;
constexpr int kFftLdsElems = 2048;  // complex values of LDS per workgroup (32 KB)
constexpr int kFftMaxG = 16;
constexpr int kFftMaxRadices = 12;

// One FFT pass over the lines of one axis and what it stores.
struct FftPass {
    double2* buf;
    const double2* tw;     // exp(-2 pi i m / L), m < L
    const double2* other;  // mode 1 with cross: the other set's finished transform
    int L, G, nrad;
    int Lp;                   // LDS stride of a line: L + L / 8 + 1 (lds_pad)
    int rad[kFftMaxRadices];  // radices of the stages, product L
    long long S, nlines;      // element stride; line l starts at (l / S) * S * L + (l % S)
    // a line is computed iff keep(p, pe, pL) and keep(q, qe, qL), with
    // outer = l / S, p = outer / qdim, q = outer % qdim and
    // keep(v, e, Lv) = v < e or (Lv > 0 and v > Lv - e)
    int qdim, pe, pL, qe, qL;
    int load_n;            // elements k < load_n are read, the rest are zero
    int store_e, store_L;  // elements k with keep(k, store_e, store_L) are written (mode 3: counted)
    int mode;              // 0: one transform; 1: forward, spectrum, inverse; 3: inverse, then pair counts
    int inverse;           // mode 0: the inverse transform
    int cross;             // mode 1: conj(F) * other instead of |F|^2
    // mode 3 (the z pass): line l = x * Ly + y, element k = z
    int Lx, Ly;
    int self;
    long long n_self;
    double scale;
    unsigned long long* scount;
    int smax;
    int* flag;
};

__device__ __forceinline__ bool keep_coord(int v, int e, int Lv) { return v < e || (Lv > 0 && v > Lv - e); }

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double c) { return make_double2(a.x * c, a.y * c); }
// (FMAs written out: the file is built without contraction, and the
// transforms only need their counts to round to the right integers)
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(__builtin_fma(a.x, b.x, -a.y * b.y), __builtin_fma(a.x, b.y, a.y * b.x));
}
// -i a (forward) or +i a (inverse)
__device__ __forceinline__ double2 mul_i(double2 a, bool inv) {
    return inv ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// DFT of R points, y_s = sum_t a_t w^(s t), w = exp(-+2 pi i / R)
template <int R>
__device__ __forceinline__ void bfly(double2 (&a)[R], bool inv);
template <>
__device__ __forceinline__ void bfly<2>(double2 (&a)[2], bool) {
    const double2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}
template <>
__device__ __forceinline__ void bfly<3>(double2 (&a)[3], bool inv) {
    constexpr double s3 = 0.86602540378443864676;  // sin(2 pi / 3)
    const double2 t = cadd(a[1], a[2]);
    const double2 u = mul_i(cscale(csub(a[1], a[2]), s3), inv);
    const double2 m = csub(a[0], cscale(t, 0.5));
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, u);
    a[2] = csub(m, u);
}
template <>
__device__ __forceinline__ void bfly<4>(double2 (&a)[4], bool inv) {
    const double2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    const double2 t2 = cadd(a[1], a[3]), t3 = mul_i(csub(a[1], a[3]), inv);
    a[0] = cadd(t0, t2);
    a[1] = cadd(t1, t3);
    a[2] = csub(t0, t2);
    a[3] = csub(t1, t3);
}
template <>
__device__ __forceinline__ void bfly<5>(double2 (&a)[5], bool inv) {
    constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;  // cos(2 pi / 5), cos(4 pi / 5)
    constexpr double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;   // sin(2 pi / 5), sin(4 pi / 5)
    const double2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    const double2 u1 = csub(a[1], a[4]), u2 = csub(a[2], a[3]);
    const double2 m1 = cadd(a[0], cadd(cscale(t1, c1), cscale(t2, c2)));
    const double2 m2 = cadd(a[0], cadd(cscale(t1, c2), cscale(t2, c1)));
    const double2 n1 = mul_i(cadd(cscale(u1, s1), cscale(u2, s2)), inv);
    const double2 n2 = mul_i(csub(cscale(u1, s2), cscale(u2, s1)), inv);
    a[0] = cadd(a[0], cadd(t1, t2));
    a[1] = cadd(m1, n1);
    a[2] = cadd(m2, n2);
    a[3] = csub(m2, n2);
    a[4] = csub(m1, n1);
}

// LDS position of element k of a line: one spare slot after every 8
// elements, and a line stride Lp = L + L / 8 + 1, so that the 16-byte
// accesses of a wave spread over all banks in every stage (strides of 4, 16
// or 64 elements would otherwise land on a few banks) and in the strided
// passes' loads and stores (lines side by side).
__device__ __forceinline__ int lds_pad(int k) { return k + (k >> 3); }

// a / b for 0 <= a < 2^16 and 1 <= b <= 2^12 from inv_b ~ 1 / b (within a
// few ulp): (a + 1/2) / b lies at least 1 / (2b) away from the integers, far
// more than the float error, so truncation gives the quotient. Integer
// division by a value known only at run time costs tens of instructions.
__device__ __forceinline__ int fdiv(int a, float inv_b) { return (int)(((float)a + 0.5f) * inv_b); }

// One radix-R stage over the G lines in LDS after stages of product p:
// butterfly i (k = i mod p) takes x[i + r L/R], r < R, times w_pR^(r k),
// and writes y_r to (i - k) R + k + r p. All inputs are read before the
// first write (registers, then a barrier), so the stage works in place.
template <int R>
__device__ __forceinline__ void fft_stage(double2* sh, int G, int L, int Lp, int p, const double2* __restrict__ tw,
                                          bool inv) {
    constexpr int J = (kFftLdsElems / R + kFftThreads - 1) / kFftThreads;  // butterflies per thread, at most
    const int nb = L / R, total = G * nb, step = L / (p * R);
    const float inv_nb = 1.0f / (float)nb, inv_p = 1.0f / (float)p;
    double2 y[J][R];
    int at[J];  // LDS position of the butterfly's first output
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int t = threadIdx.x + j * kFftThreads;
        if (t < total) {
            const int g = fdiv(t, inv_nb), i = t - g * nb;
            const int m = fdiv(i, inv_p), k = i - m * p;
            const double2* line = sh + g * Lp;
#pragma unroll
            for (int r = 0; r < R; ++r) y[j][r] = line[lds_pad(i + r * nb)];
            if (k) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    double2 w = tw[r * k * step];
                    if (inv) w.y = -w.y;
                    y[j][r] = cmul(y[j][r], w);
                }
            }
            bfly<R>(y[j], inv);
            at[j] = g * Lp + m * p * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int t = threadIdx.x + j * kFftThreads;
        if (t < total) {
            // line g, element (i - k) R + k + r p, padded within the line
            const int g = fdiv(t, inv_nb);
            double2* line = sh + g * Lp;
            const int base = at[j] - g * Lp;
#pragma unroll
            for (int r = 0; r < R; ++r) line[lds_pad(base + r * p)] = y[j][r];
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void fft_lines(double2* sh, const FftPass& p, bool inv) {
    int span = 1;
    for (int s = 0; s < p.nrad; ++s) {
        const int R = p.rad[s];
        if (R == 4) fft_stage<4>(sh, p.G, p.L, p.Lp, span, p.tw, inv);
        else if (R == 2) fft_stage<2>(sh, p.G, p.L, p.Lp, span, p.tw, inv);
        else if (R == 3) fft_stage<3>(sh, p.G, p.L, p.Lp, span, p.tw, inv);
        else fft_stage<5>(sh, p.G, p.L, p.Lp, span, p.tw, inv);
        span *= R;
    }
}

__global__ __launch_bounds__(kFftThreads) void fft_lines_kernel(FftPass p) {
    extern __shared__ double2 sh[];  // [G][Lp], element k of line g at g * Lp + lds_pad(k)
    // per line of the workgroup: its first element, or -1 when it is not
    // computed; (dx, dy) of the count pass. Computed once, so the element
    // loops need no 64-bit division.
    __shared__ long long s_base[kFftMaxG];
    __shared__ int s_dxy[kFftMaxG][2];
    __shared__ int s_any;
    const int L = p.L, G = p.G, Lp = p.Lp;
    const long long S = p.S;
    const long long l0 = (long long)blockIdx.x * G;
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    if (threadIdx.x < G) {
        const int g = threadIdx.x;
        const long long l = l0 + g;
        bool keep = l < p.nlines;
        long long base = -1;
        if (keep) {
            const long long outer = l / S;
            const int pc = (int)(outer / p.qdim), qc = (int)(outer % p.qdim);
            keep = keep_coord(pc, p.pe, p.pL) && keep_coord(qc, p.qe, p.qL);
            if (keep) base = outer * S * L + (l - outer * S);
        }
        s_base[g] = base;
        if (keep) s_any = 1;
        if (p.mode == 3) {
            const int xi = (int)(l / p.Ly), yi = (int)(l % p.Ly);
            s_dxy[g][0] = xi <= p.Lx / 2 ? xi : xi - p.Lx;
            s_dxy[g][1] = yi <= p.Ly / 2 ? yi : yi - p.Ly;
        }
    }
    __syncthreads();
    if (!s_any) return;  // the same for every thread of the workgroup
    const int total = G * L;
    // element e of the workgroup: line g, position k (consecutive threads
    // along the lines' contiguous direction)
    const float inv_L = 1.0f / (float)L, inv_G = 1.0f / (float)G;
    auto line_of = [&](int e) { return S == 1 ? fdiv(e, inv_L) : e - fdiv(e, inv_G) * G; };
    auto pos_of = [&](int e) { return S == 1 ? e - fdiv(e, inv_L) * L : fdiv(e, inv_G); };
    {
        // every load of the thread in flight before the first LDS write
        constexpr int J = kFftLdsElems / kFftThreads;
        double2 v[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int e = threadIdx.x + j * kFftThreads;
            const int g = e < total ? line_of(e) : 0, k = pos_of(e);
            const long long base = s_base[g];
            v[j] = e < total && k < p.load_n && base >= 0 ? p.buf[base + (long long)k * S] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int e = threadIdx.x + j * kFftThreads;
            if (e < total) sh[line_of(e) * Lp + lds_pad(pos_of(e))] = v[j];
        }
    }
    __syncthreads();
    fft_lines(sh, p, p.mode == 0 ? p.inverse != 0 : p.mode == 3);
    if (p.mode == 1) {
        // the power spectrum |F|^2, or the transform of sum_v a(v) b(v + D),
        // conj(F_a) * F_b; then back along this axis
        for (int e = threadIdx.x; e < total; e += kFftThreads) {
            const int g = line_of(e), k = pos_of(e);
            const long long base = s_base[g];
            if (base < 0) continue;
            const double2 v = sh[g * Lp + lds_pad(k)];
            if (p.cross) {
                const double2 o = p.other[base + (long long)k * S];
                sh[g * Lp + lds_pad(k)] = make_double2(v.x * o.x + v.y * o.y, v.x * o.y - v.y * o.x);
            } else {
                sh[g * Lp + lds_pad(k)] = make_double2(v.x * v.x + v.y * v.y, 0.0);
            }
        }
        __syncthreads();
        fft_lines(sh, p, true);
    }
    if (p.mode != 3) {
        for (int e = threadIdx.x; e < total; e += kFftThreads) {
            const int g = line_of(e), k = pos_of(e);
            const long long base = s_base[g];
            if (base >= 0 && keep_coord(k, p.store_e, p.store_L))
                p.buf[base + (long long)k * S] = sh[g * Lp + lds_pad(k)];
        }
        return;
    }
    // pair counts: entries +dz and -dz (0 <= dz < Ez) of a line share the
    // squared displacement, so one thread takes both and adds once
    const int nz = p.store_e;
    const float inv_nz = 1.0f / (float)nz;
    for (int e = threadIdx.x; e < G * nz; e += kFftThreads) {
        const int g = fdiv(e, inv_nz), dz = e - g * nz;
        if (s_base[g] < 0) continue;
        const double cp = sh[g * Lp + lds_pad(dz)].x * p.scale;
        const double cm = dz ? sh[g * Lp + lds_pad(L - dz)].x * p.scale : 0.0;
        const double rp = rint(cp), rm = rint(cm);
        if (fabs(cp - rp) > 0.25 || rp < 0.0 || fabs(cm - rm) > 0.25 || rm < 0.0) {
            atomicOr(p.flag, 1);
            continue;
        }
        const int dx = s_dxy[g][0], dy = s_dxy[g][1];
        long long cnt = (long long)rp + (long long)rm;
        if (p.self) {
            // one of D and -D (the pair once): lines with dx > 0, or dx = 0
            // and dy > 0, count both signs of dz; the line dx = dy = 0 counts
            // dz > 0, and at D = 0 the pairs of coincident points, (A(0) - n) / 2
            if (dx == 0 && dy == 0) {
                cnt = dz ? (long long)rp : ((long long)rp - p.n_self) / 2;
            } else if (!(dx > 0 || (dx == 0 && dy > 0))) {
                continue;
            }
        }
        if (cnt <= 0) continue;
        const int s2 = dx * dx + dy * dy + dz * dz;
        if (s2 > p.smax) {
            atomicOr(p.flag, 2);
            continue;
        }
        atomicAdd(&p.scount[s2], (unsigned long long)cnt);
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
    if (!lattice_length_ok(L)) return hipErrorInvalidValue;
    p.L = L;
    p.nrad = 0;
    int x = L;
    while (x % 4 == 0) p.rad[p.nrad++] = 4, x /= 4;
    if (x % 2 == 0) p.rad[p.nrad++] = 2, x /= 2;
    while (x % 3 == 0) p.rad[p.nrad++] = 3, x /= 3;
    while (x % 5 == 0) p.rad[p.nrad++] = 5, x /= 5;
    p.G = std::max(1, std::min(kFftMaxG, kFftLdsElems / L));
    p.Lp = L + L / 8 + 1;
    p.S = S;
    p.nlines = total / L;
    const unsigned blocks = (unsigned)((p.nlines + p.G - 1) / p.G);
    hipLaunchKernelGGL(fft_lines_kernel, dim3(blocks), dim3(kFftThreads), (size_t)p.G * p.Lp * sizeof(double2), s, p);
    return hipGetLastError();
}

// every line, every element read and written
FftPass all_lines(double2* buf, const double2* tw, int L) {
    FftPass p{};
    p.buf = buf;
    p.tw = tw;
    p.qdim = 1, p.pe = 1 << 30, p.pL = 0, p.qe = 1, p.qL = 0;
    p.load_n = L;
    p.store_e = L, p.store_L = 0;
    return p;
}

bool geom_ok(const LatticeGeom& g) {
    for (int a = 0; a < 3; ++a)
        if (!lattice_length_ok(g.L[a]) || g.E[a] < 1 || 2 * g.E[a] - 1 > g.L[a]) return false;
    return true;
}

// the forward z and y passes over the filled box
hipError_t forward_zy(double2* buf, const LatticeGeom& g, const double2* const tw[3], hipStream_t s) {
    const long long S[3] = {(long long)g.L[1] * g.L[2], g.L[2], 1};
    // z pass: lines x < Ex, y < Ey; the occupancy is zero beyond z = Ez
    FftPass pz = all_lines(buf, tw[2], g.L[2]);
    pz.qdim = g.L[1], pz.pe = g.E[0], pz.qe = g.E[1];
    pz.load_n = g.E[2];
    if (hipError_t e = fft_pass(pz, g, 2, S[2], s); e != hipSuccess) return e;
    // y pass: lines x < Ex; zero beyond y = Ey
    FftPass py = all_lines(buf, tw[1], g.L[1]);
    py.pe = g.E[0];
    py.load_n = g.E[1];
    return fft_pass(py, g, 1, S[1], s);
}

}  // namespace

bool lattice_length_ok(int L) {
    if (L < 2 || L > kLatticeMaxL) return false;
    for (int f : {2, 3, 5})
        while (L % f == 0) L /= f;
    return L == 1;
}

int lattice_length(int E) {
    int L = std::max(2, 2 * E - 1);
    while (L <= kLatticeMaxL && !lattice_length_ok(L)) ++L;
    return L <= kLatticeMaxL ? L : -1;
}

hipError_t launch_lattice_zero_box(double2* buf, const LatticeGeom& g, hipStream_t s) {
    const size_t n = (size_t)g.E[0] * g.E[1] * g.E[2];
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(lattice_zero_box_kernel, dim3(grid_for(n)), dim3(256), 0, s, buf, g);
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

hipError_t launch_fft3_forward(double2* buf, const LatticeGeom& g, const double2* const tw[3], hipStream_t s) {
    if (!geom_ok(g)) return hipErrorInvalidValue;
    if (hipError_t e = forward_zy(buf, g, tw, s); e != hipSuccess) return e;
    // x pass: every line, zero beyond x = Ex, the whole transform stored
    FftPass px = all_lines(buf, tw[0], g.L[0]);
    px.load_n = g.E[0];
    return fft_pass(px, g, 0, (long long)g.L[1] * g.L[2], s);
}

hipError_t launch_fft3_correlate(double2* buf, const LatticeGeom& g, const double2* const tw[3], const double2* other,
                                 long long n_self, unsigned long long* scount, int smax, int* flag, hipStream_t s) {
    if (!geom_ok(g)) return hipErrorInvalidValue;
    const bool self = other == nullptr;
    const long long S[3] = {(long long)g.L[1] * g.L[2], g.L[2], 1};
    if (hipError_t e = forward_zy(buf, g, tw, s); e != hipSuccess) return e;
    // the displacements of x kept from here on: 0 <= Dx < Ex for one set (the
    // other half is its mirror image), |Dx| < Ex for two
    const int xL = self ? 0 : g.L[0];
    // x pass: every line; forward, spectrum, inverse; those Dx written
    FftPass px = all_lines(buf, tw[0], g.L[0]);
    px.mode = 1;
    px.cross = self ? 0 : 1;
    px.other = other;
    px.load_n = g.E[0];
    px.store_e = g.E[0], px.store_L = xL;
    if (hipError_t e = fft_pass(px, g, 0, S[0], s); e != hipSuccess) return e;
    // inverse y pass: the lines of those Dx; |Dy| < Ey written
    FftPass py = all_lines(buf, tw[1], g.L[1]);
    py.inverse = 1;
    py.pe = g.E[0], py.pL = xL;
    py.store_e = g.E[1], py.store_L = g.L[1];
    if (hipError_t e = fft_pass(py, g, 1, S[1], s); e != hipSuccess) return e;
    // inverse z pass: the lines of those Dx and |Dy| < Ey; entries |Dz| < Ez to counts
    FftPass pz = all_lines(buf, tw[2], g.L[2]);
    pz.mode = 3;
    pz.qdim = g.L[1], pz.pe = g.E[0], pz.pL = xL, pz.qe = g.E[1], pz.qL = g.L[1];
    pz.store_e = g.E[2], pz.store_L = g.L[2];
    pz.Lx = g.L[0], pz.Ly = g.L[1];
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
