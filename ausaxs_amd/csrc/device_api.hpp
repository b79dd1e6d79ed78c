// Device-side entry points of the MI355X Debye hot path (gfx950).
//
// Host code (managers, C ABI) calls these launchers; the kernels themselves
// live in pair_hist.hip and debye.hip. All launchers are asynchronous on the
// given stream and return hipError_t.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace ausaxs::gpu {

// Fixed-point encoding of one LDS/partial histogram cell (one uint64). Each
// pair adds kCellUnit + k, with k = round(offset / 2^-E) + kDeltaBias, where
// offset = d - b / inv_w is the pair's distance from its bin centre in A and
// E = delta_exp(inv_w) (|offset| <= 0.5 bins = 2^19 units, plus the pair
// kernel's Newton correction of at most 2^13 units). The low word of the
// addend is the bit pattern of a float in [1.5 * 2^23, 2^24), where floats are
// the integers, kLowBias + k, so the kernel adds it without a float-to-int
// conversion. A cell holding n pairs is T = n * (2^42 + kLowBias) + sum k with
// sum k <= n * kDeltaMax < 2^42 + kLowBias, so n = floor(T / (2^42 +
// kLowBias)) and T stays below 2^64 while n <= kMaxPairsPerItem. The mean
// distance of a bin is known to half a unit (2^-21 bins or finer) even when
// every pair in it shares one offset (lattices, symmetric test bodies).
// Integer accumulation keeps every histogram bit-reproducible.
constexpr int64_t kDeltaBias = (int64_t(1) << 19) + (int64_t(1) << 13);
constexpr int64_t kDeltaMax = (int64_t(1) << 20) + (int64_t(1) << 14);
constexpr uint64_t kLowBias = 0x4B400000u;  // bits of 1.5 * 2^23
constexpr uint64_t kCellUnit = (uint64_t(1) << 42) + kLowBias;
constexpr int64_t kMaxPairsPerItem = 4128768;  // 512 x 8064; * kDeltaMax < kCellUnit
static_assert((uint64_t)kMaxPairsPerItem * (uint64_t)kDeltaMax < kCellUnit, "cell count decode");
static_assert((uint64_t)kMaxPairsPerItem <= ~uint64_t(0) / (kCellUnit + (uint64_t)kDeltaMax), "cell range");
// Pair distances reach at most this many bins (the bound on the Newton
// correction above); larger extents are rejected on the host.
constexpr int kMaxExtentBins = 1 << 15;
// Default columns per pair work item. Items of 512 x 2000 pairs (about 0.4 ms
// of one workgroup) ran the 250k-atom launch 5 % faster than items of
// 512 x 8000, and the 20k-atom grid launch 4 % faster, at four times the
// partial-histogram slots for the merge; 1000 and 1500 were slower again.
constexpr int kDefaultTileJ = 2000;

// Exponent of the offset unit 2^-E A: the largest E with 2^E / inv_w <= 2^20,
// so that half a bin is at most 2^19 units (E = 23 at 0.1 A, 22 at 0.25 A).
__host__ __device__ inline int delta_exp(float inv_w) {
    int ex = 0;
    (void)frexpf(inv_w, &ex);  // inv_w = m 2^ex, m in [0.5, 1)
    return ex + 19;
}
__host__ __device__ inline double delta_unit(float inv_w) { return ldexp(1.0, -delta_exp(inv_w)); }
// Sum of the distances (A) of n pairs in bin b whose offsets sum to ds units,
// with the unit (delta_unit) evaluated by the caller: loops over many bins
// take it once instead of a frexpf and an ldexp per bin.
__host__ __device__ inline double distance_sum_u(uint64_t n, int b, int64_t ds, float inv_w, double unit) {
    return (double)n * (double)b / (double)inv_w + (double)ds * unit;
}
__host__ __device__ inline double distance_sum(uint64_t n, int b, int64_t ds, float inv_w) {
    return distance_sum_u(n, b, ds, inv_w, delta_unit(inv_w));
}

// One workgroup's rectangle of the pair space.
struct PairWork {
    int32_t i_begin, i_end;  // row points (indices into the point buffer)
    int32_t j_begin, j_end;  // column points
    int32_t row0, row1;      // global histogram rows: row0 if tag_i < tag_j, else row1
    int32_t flags;           // kDiag | kOrient
    int32_t slot;            // partial slot (= item index)
};
constexpr int32_t kDiag = 1;    // same point set and overlapping ranges: count j > i only
constexpr int32_t kOrient = 2;  // two output rows chosen by original-index order
// two output rows chosen by the row point's position: rows before i_begin +
// split_rows(flags) use row0, the rest row1 (a row tile spanning two classes
// of a cross job)
constexpr int32_t kSplit = 4;
constexpr int kSplitShift = 8;
// Columns of a lattice set given as z-runs (PairLaunch::runs): [j_begin,
// j_end) are run indices, and the item's column count is run_cols(flags).
// A run is {x, y, z0, length as int bits}: the lattice points (x, y, z0 + t),
// t < length, in the set's original order.
constexpr int32_t kRuns = 8;
constexpr int kRunColsShift = 18;
__host__ __device__ inline int split_rows(int32_t flags) { return (flags >> kSplitShift) & 0x3FF; }
__host__ __device__ inline int run_cols(int32_t flags) { return (flags >> kRunColsShift) & 0x1FFF; }
// columns of an item
__host__ __device__ inline int item_cols(const PairWork& w) {
    return (w.flags & kRuns) ? run_cols(w.flags) : w.j_end - w.j_begin;
}

// Pairs involving the reference's scalar-tail columns (last <= 3 points of a
// loop), evaluated with per-pair rule selection and added with global atomics.
struct StripWork {
    int32_t n_rows;        // row points (original order)
    int32_t t_begin;       // first column index handled (original order)
    int32_t t_end;         // one past last
    int32_t n_cols_total;  // column set size (n for self loops, m for cross)
    int32_t self;          // 1: self loop (i < t), rule by (n-i-1) mod 4; 0: cross (always scalar rule)
    int32_t n_col_classes; // row = rowmap[class_i * n_col_classes + class_t]
};

struct PairLaunch {
    const float4* pts;        // point buffer (x, y, z, tag bits)
    const float4* runs;       // z-runs of lattice column sets (kRuns items; may be null otherwise)
    const PairWork* items;    // device
    int n_items;
    uint64_t* partial;        // [n_items][2][nbins+1]
    float inv_w;
    int nbins;                // LDS bins per row; slot nbins collects overflow
    int clamp;                // 0 only if every pair's bin is known to be < nbins
    int grid;                 // workgroups; 0 or >= n_items: one per item, else persistent (strided)
    int rows = 2;             // LDS histogram rows: 2 if any item has kOrient or kSplit, else 1
    int copies = 1;           // LDS histogram copies (1, 2, 4 or 8), lane l adds into copy l mod copies
    // Weight pass (XYZW sets of any weights): pts holds (x, y, z, weight) and
    // every pair adds weight_code(w_i * w_j) instead of a count cell (no
    // oriented or split items); the merge then runs with raw = 1.
    int wmode = 0;
    double wscale = 1;        // 2^s: fixed-point scale of the weight products
};

// Weight products as integers: w_i * w_j (float, as the reference's XYZW
// evaluate forms it) times wscale, rounded to nearest, is k = weight_code
// (two's complement in a uint64): the double kWeightMagic + w_i w_j wscale
// lies in [2^52, 2^53), where doubles are the integers, and its bit pattern
// minus kWeightMagicBits is k. With |w_i w_j| wscale <= 2^40 each product is
// kept to 2^-41 of the largest one, and the k of one work item (at most
// kMaxPairsPerItem < 2^22 pairs) or one strip cell sum exactly in 64 bits;
// the merge adds the items' sums in FP64 (MergeLaunch::raw) and the strips'
// integer sums are added to that on the host side.
constexpr double kWeightMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr uint64_t kWeightMagicBits = 0x4338000000000000ull;
constexpr int kWeightBits = 40;
__host__ __device__ inline uint64_t weight_code(float product, double wscale) {
    const double t = __builtin_fma((double)product, wscale, kWeightMagic);
    uint64_t u;
    __builtin_memcpy(&u, &t, 8);
    return u - kWeightMagicBits;
}
// wscale for weights of magnitude up to wmax: a power of two with
// wmax^2 * wscale <= 2^kWeightBits
inline double weight_scale(double wmax) {
    if (!(wmax > 0)) return 1.0;
    int ex = 0;
    (void)std::frexp(wmax, &ex);  // wmax < 2^ex
    return std::ldexp(1.0, kWeightBits - 2 * ex);
}
// The summed weight products of a histogram cell: items' FP64 sum (bits
// of a double, as the raw merge writes it) plus the strips' integer sum.
__host__ __device__ inline double weight_value(int64_t item_bits, int64_t strip_k, double wscale) {
    double items;
    __builtin_memcpy(&items, &item_bits, 8);
    return (items + (double)strip_k) / wscale;
}

// LDS bytes of one pair_hist_kernel workgroup (static column tile included)
constexpr size_t kStaticLds = (512 + 2) * 16;
constexpr size_t pair_lds_bytes(int nbins, int rows, int copies) {
    return (size_t)copies * rows * (nbins + 2) * sizeof(uint64_t) + kStaticLds;
}

hipError_t launch_pair_hist(const PairLaunch& p, hipStream_t s);

// Sum partial slots into final rows in two passes. row_list holds partial
// entries (item*2 + local_row) grouped by output row; the entries of one row
// are cut into chunks of at most kMergeChunk: chunk c = row_list[chunk_ptr[c]
// .. chunk_ptr[c+1]), and row r owns chunks row_chunk[r] .. row_chunk[r+1].
// tmp_cnt/tmp_dsum: [n_chunks][nbins+1]; cnt/dsum: [n_rows][nbins+1]. Every
// output cell is written (rows without entries become zero). The chunk pass
// reads entries in batches of kMergeBatch, so row_list carries kMergeBatch - 1
// entries of slack past its last chunk (any valid partial row; their values
// are read and dropped).
constexpr int kMergeChunk = 32;
constexpr int kMergeBatch = 16;
struct MergeLaunch {
    const uint64_t* partial;
    const int32_t* row_list;
    const int32_t* chunk_ptr;  // [n_chunks+1]
    const int32_t* row_chunk;  // [n_rows+1]
    int n_chunks, n_rows, nbins;
    uint64_t* tmp_cnt;
    int64_t* tmp_dsum;
    uint64_t* cnt;
    int64_t* dsum;
    int raw = 0;  // 1: partials of a weight pass (integer k sums per item), summed in FP64 into dsum as
                  //    the bits of doubles (cnt zero)
    int direct = 0;  // 1: every row has exactly one chunk (chunk r is row r): the chunk sums are the
                     //    rows, written to cnt/dsum at once, and merge_rows_kernel is not launched
};
hipError_t launch_merge(const MergeLaunch& m, hipStream_t s);

struct StripLaunch {
    const float4* rows;       // original-order row points (tag = class id bits)
    const float4* cols;       // original-order column points
    StripWork w;
    const int32_t* rowmap;    // device
    uint64_t* cnt;
    int64_t* dsum;
    float inv_w;
    int nbins;
    int stride;               // nbins+1
    // replicas (grid.y): replica r reads rows + r*rows_rep, cols + r*cols_rep
    // (points) and adds r*out_rep to the histogram row
    int n_rep = 1;
    int64_t rows_rep = 0, cols_rep = 0;
    int out_rep = 0;
    // >= 0: every pair of the strip goes to this histogram row (one class on
    // each side), which lets long strips accumulate in LDS first
    int fixed_row = -1;
    // weight pass: points hold (x, y, z, weight), every pair goes to rowmap[0]
    // and adds weight_code(w_i * w_j) to dsum, an integer sum that must start
    // at zero (cnt untouched)
    int wmode = 0;
    double wscale = 1;
};
hipError_t launch_strip(const StripLaunch& p, hipStream_t s);
// pairs per workgroup of a strip that accumulates in an LDS histogram first
constexpr int kStripBlockPairs = 4096;
// a strip with one output row, no replicas, a histogram that fits 64 KB of
// LDS and more than kStripBlockPairs pairs accumulates in LDS first (long
// strips crowd into few bins, where one global atomic per pair serialises)
inline bool strip_uses_lds(const StripLaunch& p) {
    const long long n = (long long)p.w.n_rows * (p.w.t_end - p.w.t_begin);
    return p.fixed_row >= 0 && p.n_rep == 1 && (size_t)(p.nbins + 1) * sizeof(uint64_t) <= 65536 &&
           n > kStripBlockPairs;
}
inline int64_t strip_lds_blocks(const StripLaunch& p) {
    const int64_t n = (int64_t)p.w.n_rows * (p.w.t_end - p.w.t_begin);
    return n > 0 ? (n + kStripBlockPairs - 1) / kStripBlockPairs : 0;
}
// Strips table[0..n) (device) in one launch, each with n_rep = 1: strip k
// takes blocks [first[k], first[k+1]) of 256 pairs, first[n] = blocks
// (device, int64). See strip_table_blocks for the host-side prefix.
hipError_t launch_strip_table(const StripLaunch* table, const int64_t* first, int n, int64_t blocks,
                              hipStream_t s);
// blocks of 256 pairs a strip takes in launch_strip_table
inline int64_t strip_table_blocks(const StripLaunch& p) {
    const int64_t n = (int64_t)p.w.n_rows * (p.w.t_end - p.w.t_begin);
    return n > 0 ? (n + 255) / 256 : 0;
}
// LDS strips (strip_uses_lds) table[0..n) in one launch: strip k takes
// blocks [first[k], first[k+1]) of kStripBlockPairs pairs (strip_lds_blocks),
// every strip's histogram nbins + 1 cells wide.
hipError_t launch_strip_lds_table(const StripLaunch* table, const int64_t* first, int n, int64_t blocks, int nbins,
                                  hipStream_t s);

// Exhaustive check of the binning arithmetic over float d2 in [lo, hi] (bit
// patterns). out[0]: lanes where the pair kernel's bin (raw v_sqrt + tie
// fallback, scaled coordinates) differs from the correctly rounded reference;
// out[1]: where the device's correctly rounded sqrt is not the IEEE result
// (checked in FP64); out[2]: where the fallback's residual correction is not
// it; out[3]: where the pair kernel's offset code is more than 1 + 2^-24 q
// units (q = units per bin) from the exact offset of the real root (FP64); out[4]: the sum of those code
// errors in units of 2^-16 (two's complement); out[5]: the same sum for the
// correctly rounded root's offset (the reference's distance).
hipError_t launch_bin_check(uint32_t lo_bits, uint32_t hi_bits, float inv_w, unsigned long long* out,
                            hipStream_t s);

// Copies device buffers (32-bit words), back to back, into dst: pinned host
// memory mapped for the device (its device address), so small results reach
// the host through the kernel's own stores, without a DMA transfer.
constexpr int kCopyOutMax = 6;
struct CopyOut {
    const uint32_t* src[kCopyOutMax];
    size_t words[kCopyOutMax];
    int n = 0;
    uint32_t* dst = nullptr;
};
hipError_t launch_copy_out(const CopyOut& c, hipStream_t s);

// Plain copy of `bytes` (a multiple of 16) from src to dst: the reference
// stream for calibrating HBM traffic counters.
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

// Issue-rate probes (probe.hip): which = 0 FP32 adds with vector-register
// operands, 1 the same with a scalar-register operand, 2 v_rsq_f32, 3
// ds_add_u64 at hashed cells of an nbins-cell LDS histogram (three 512-lane
// workgroups per CU, nbins * 8 <= 64 KB). out: probe_out_floats(cus) floats
// (results kept so the work is not removed). *wave_instr: the wave
// instructions of the measured class that one launch issues.
size_t probe_out_floats(int cus);
hipError_t launch_issue_probe(int which, int cus, int nbins, float* out, hipStream_t s, double* wave_instr);

// Pair histograms of integer-lattice point sets by FFT correlation
// (lattice.hip). Grid index (i, j, k) holds the point origin + (i, j, k);
// L: line lengths 2^a 3^b 5^c (x slowest, z fastest), each in
// [2, kLatticeMaxL]; E: extent of the occupied box (points at indices [0, E)
// per axis), 2E - 1 <= L so that no displacement wraps.
constexpr int kLatticeMaxL = 1024;
struct LatticeGeom {
    int origin[3];
    int L[3];
    int E[3];
};
// L is a transform length the kernels take
bool lattice_length_ok(int L);
// the shortest such length >= 2E - 1 (and >= 2), or -1 beyond kLatticeMaxL
int lattice_length(int E);
// clears the occupied box (x < Ex, y < Ey, z < Ez); nothing else needs clearing
hipError_t launch_lattice_zero_box(double2* buf, const LatticeGeom& g, hipStream_t s);
hipError_t launch_lattice_zero_u64(unsigned long long* buf, size_t n, hipStream_t s);
// occupancy += 1 at every point (integer coordinates inside the occupied box)
hipError_t launch_lattice_fill(const float4* pts, int n, const LatticeGeom& g, double2* buf, hipStream_t s);
// in-place forward 3D FFT, exp(-2 pi i k x / L), of a filled occupancy grid;
// tw[a]: exp(-2 pi i m / L[a]) for m < L[a]. The whole transform is stored.
hipError_t launch_fft3_forward(double2* buf, const LatticeGeom& g, const double2* const tw[3], hipStream_t s);
// pair counts per squared displacement, scount[s] += (s <= smax), of a
// filled occupancy grid: with itself (other = null; each pair once, n_self
// points) or with a set whose finished transform is `other` (cross
// correlation sum_v a(v) b(v + D)), for displacements |D| < E per axis.
// flag |= 1 on a non-integer count, 2 on s > smax. buf is left partly
// transformed.
hipError_t launch_fft3_correlate(double2* buf, const LatticeGeom& g, const double2* const tw[3], const double2* other,
                                 long long n_self, unsigned long long* scount, int smax, int* flag, hipStream_t s);
// scount -> counts and delta sums of one histogram row (the pair kernel's
// binning), added to cnt / dsum [nbins + 1] with atomics
hipError_t launch_lattice_bins(const unsigned long long* scount, int smax, float inv_w, int nbins, uint64_t* cnt,
                               int64_t* dsum, hipStream_t s);

// Debye contraction on MFMA: out[r][k] = sum_b H[r][b] * sinc(q_k * d_b).
struct DebyeLaunch {
    const double* H;   // [rows][K] device
    const double* d;   // [K] device
    const double* q;   // [nq] device
    int rows, K, nq;
    int terms;         // bf16 terms per operand (2: hi+lo, 3: hi+mid+lo)
    double* out;       // [rows][nq] device
    void* scratch;     // workspace, see debye_scratch_bytes
};
size_t debye_scratch_bytes(int rows, int K, int nq);
// The allocation-free contraction's work space: a static device array of
// kDebyeNoAllocBytes on every GPU the library runs on (debye.hip), whose
// address on the current device is looked up once per device.
constexpr size_t kDebyeNoAllocBytes = size_t(32) << 20;
hipError_t debye_noalloc_scratch(void** ptr);
hipError_t launch_debye(const DebyeLaunch& p, hipStream_t s);
// Plain FP64 contraction (same operands), used as the device reference path.
hipError_t launch_debye_f64(const DebyeLaunch& p, hipStream_t s);

// FP64 weighted least squares y ~ a x + b for n_models curves x[m][n]
// against one data set y[n] (err may be NULL: unit errors); out[m][5] =
// a, b, S/delta, Sxx/delta, chi2.
hipError_t launch_lsq(const double* x, const double* y, const double* err, int n, int n_models, double* out,
                      hipStream_t s);
// Exact O(N^2) Debye sum; partial: [nq][exact_debye_blocks(n)] workspace.
int exact_debye_blocks(int n);
hipError_t launch_exact_debye(const float4* pts, int n, const double* q, int nq, double* partial, double* out,
                              hipStream_t s);

}  // namespace ausaxs::gpu
