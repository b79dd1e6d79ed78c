// Pair-distance histogram kernels for gfx950 (CDNA4).
//
// Restates the reference's fused evaluate+increment (CompactCoordinatesXYZFF.h
// evaluate_{16,8,4,1}, HistogramManagerMTFFAvg.cpp:214-280) as:
//   * one workgroup per rectangle of the pair space (PairWork): 512 lanes in
//     two parts that hold the same 512 row points and split each column tile;
//   * row points in VGPRs (2 per lane); column points staged through LDS in
//     tiles of 512 and read back as wave-wide broadcasts in double-buffered
//     groups of 2, so each lane evaluates 2 pairs per column point (LDS reads
//     and LDS atomics retire in order, so the broadcast reads of the next
//     group never wait on the atomics of the current one);
//   * a workgroup-private histogram of packed uint64 cells in LDS updated
//     with ds_add_u64 (one atomic per pair);
//   * a flush of the LDS histogram to a partial slot, summed by the merge kernels.
//
// Bit-exact binning (SIMD-body rule of the reference):
//   d2 = (dx*dx + dy*dy) + dz*dz   separately rounded (file built with
//                                  -ffp-contract=off)
//   d  = sqrt_rn(d2), x = d*inv_w, bin = rint(x)  (round half to even)
// The loop runs on coordinates scaled by a power of two (binning.hpp), which
// scales d2 and the roots exactly. The root s = d2 * rsq(d2) (within 1.5 ulp
// of the real root) is used when x is more than 1.75 * 2^-22 * x away from a
// .5 tie, and the bin then comes from one FMA, t = s * inv_w + 1.5 * 2^23,
// whose bit pattern is the bin plus a constant (folded into the LDS base
// address). Inside the window the correctly rounded square root (v_sqrt_f32
// plus an FMA-residual correction; note that __fsqrt_rn lowers to the
// uncorrected instruction) decides. The offset that goes into the bin's
// distance sum is that of the real root (s plus its Newton correction, two
// FMAs with the same rsq), not of s, whose error would bias the bin centres.
// launch_bin_check verifies the bins and the offsets exhaustively on device.
#include "binning.hpp"
#include "device_api.hpp"

namespace ausaxs::gpu {
namespace {

using binning::bin_fast;
using binning::bin_of;
using binning::bin_pair;
using binning::bin_tie;
using binning::Consts;
using binning::kMagic;
using binning::kMagicBits;
using binning::make_consts;
using binning::offset_code;
using binning::simd_rule_exact;
using binning::sqrt_fix;
using binning::sqrt_rn;

// Parts: groups of 256 lanes that hold the same row points, take separate
// shares of each column tile and add into one LDS histogram. Two parts put 6
// waves on each SIMD where the histogram of a 0.1 A workload (two rows of
// ~2,500 bins, 40 KB) allows three workgroups per CU; one part left the SIMDs
// with 3 waves and 12 % of their issue slots idle, four parts measure the same
// as two.
constexpr int kParts = 2;
constexpr int kRowLanes = 256;                   // lanes holding distinct row points
constexpr int kThreads = kRowLanes * kParts;
constexpr int kRowsPerLane = 2;
constexpr int kTileI = kRowLanes * kRowsPerLane;  // row points per work item
constexpr int kColTile = 512;                    // column points per LDS tile (each part takes a share)
constexpr int kColGroup = 2;                     // column points per double-buffered group
static_assert(kColTile + kColGroup == kStaticLds / sizeof(float4), "static LDS size");

// d2 of the reference's SIMD body: separately rounded squares and sums
__device__ __forceinline__ float simd_d2(float dx, float dy, float dz) {
    const float sx = dx * dx;
    const float sy = dy * dy;
    const float sz = dz * dz;
    return (sx + sy) + sz;
}

// addend of one pair: count 2^42 plus the offset code (kLowBias + k)
__device__ __forceinline__ uint64_t cell_add(uint32_t code) { return (uint64_t(1) << 42) + (uint64_t)code; }

// count and sum of k of one cell (see device_api.hpp), without branches.
// kCellUnit = 2^22 * M (kLowBias = 301 * 2^22), so n = floor(v / kCellUnit)
// = floor(w / M) with w = v >> 22 < 2^42, exact in FP64; the FP64 quotient is
// within 2^-30 of w / M, so its truncation is n, n - 1 or n + 1, and the
// integer remainder w - c * M picks n.
constexpr uint64_t kCellM = kCellUnit >> 22;
static_assert(kCellM << 22 == kCellUnit, "cell unit is a multiple of 2^22");
__device__ __forceinline__ void cell_decode(uint64_t v, uint64_t& n, uint64_t& ksum) {
    const uint64_t w = v >> 22;
    int64_t c = (int64_t)((double)w * (1.0 / (double)kCellM));
    const int64_t r = (int64_t)w - c * (int64_t)kCellM;
    c += (int64_t)(r >= (int64_t)kCellM) - (int64_t)(r < 0);
    n = (uint64_t)c;
    ksum = v - (((uint64_t)c * kCellM) << 22);
}

__device__ __forceinline__ void lds_add(uint64_t* cell, uint64_t v) {
    __hip_atomic_fetch_add(cell, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS byte addresses are 32-bit. The pair loop addresses a cell as
// base + 8 * bits(t), with the magic constant's bits folded into base (mod
// 2^32), so the address of a pair is one shift-and-add.
using lds_u64 = __attribute__((address_space(3))) uint64_t;
__device__ __forceinline__ uint32_t lds_byte_addr(uint64_t* p) { return (uint32_t)(uintptr_t)(lds_u64*)p; }
__device__ __forceinline__ void lds_add_at(uint32_t addr, uint64_t v) {
    __hip_atomic_fetch_add((lds_u64*)(uintptr_t)addr, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Row points of one lane (kRowsPerLane of them) for the current item; roff:
// LDS offset of the row each point adds into (split items).
struct Rows {
    float x[kRowsPerLane], y[kRowsPerLane], z[kRowsPerLane];
    float tagf[kRowsPerLane];  // original index as a float (exact below 2^24); weight pass: the weight
    int idx[kRowsPerLane];
    bool valid[kRowsPerLane];
    uint32_t base[kRowsPerLane];  // LDS byte address of the row's cell 0, minus 8 * kMagicBits (mod 2^32)
};

// The launch constants that enter per-pair arithmetic, held in vector
// registers for the whole item. On gfx950 an FP32 add, multiply or FMA whose
// operands are all vector registers (inline constants and literals included)
// issues at twice the rate of the same instruction with a scalar-register
// operand, so inv_ws, win, q and m2, uniform as they are, must not stay in
// scalar registers where the compiler would keep them. The empty asm makes
// the copies opaque: they are made once per item.
__device__ __forceinline__ Consts vgpr_consts(const Consts& cs) {
    Consts cv = cs;
    asm volatile("" : "+v"(cv.inv_ws), "+v"(cv.win), "+v"(cv.q), "+v"(cv.m2));
    return cv;
}

// a uniform float held in a vector register (see vgpr_consts)
__device__ __forceinline__ float vgpr_float(float x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Bins and adds the kRowsPerLane pairs of one column whose squared distances
// (scaled) are d2: fast path, tie fallback, LDS add.
// Oriented items hold the column's original index in tag_j as a float (the
// staging converts it) and add into row 1 when the row point's index is the
// larger: sat(tag_i - tag_j) is 1 exactly then (distinct integers differ by
// at least 1), and t + sat * off1 stays an exact integer-valued float, so the
// row choice costs two VALU operations per pair and no compare or select
// (off1, the row stride in bins, is a float in a vector register).
// The weight pass (WEIGHT) bins the same way and adds the pair's weight
// product (weight_code of w_i * tag_j) instead of its count cell.
template <bool FULL, bool DIAG, bool ORIENT, bool CLAMP, bool WEIGHT>
__device__ __forceinline__ void add_pairs(const Rows& r, const float (&d2)[kRowsPerLane], float tag_j, int j,
                                          const Consts& cs, unsigned ovf, float off1, double wscale) {
    float t[kRowsPerLane], dl[kRowsPerLane], s[kRowsPerLane];
    bool near[kRowsPerLane];
    bool tie = false;
#pragma unroll
    for (int k = 0; k < kRowsPerLane; ++k) {
        near[k] = bin_fast(d2[k], cs, s[k], t[k], dl[k]);
        tie |= near[k];
    }
    if (tie) {
#pragma unroll
        for (int k = 0; k < kRowsPerLane; ++k)
            if (near[k]) bin_tie(d2[k], s[k], cs, t[k], dl[k]);
    }
#pragma unroll
    for (int k = 0; k < kRowsPerLane; ++k) {
        // u = kMagicBits + bin (+ off1 for row 1 of oriented items)
        uint32_t u;
        if (ORIENT) {
            float tt = t[k];
            if (CLAMP) tt = fminf(tt, kMagic + (float)ovf);
            const float sel = __builtin_amdgcn_fmed3f(r.tagf[k] - tag_j, 0.f, 1.f);
            u = __float_as_uint(__builtin_fmaf(sel, off1, tt));
        } else {
            u = __float_as_uint(t[k]);
            if (CLAMP) u = min(u, kMagicBits + ovf);
        }
        if (!FULL || DIAG) {
            bool keep = r.valid[k];
            if (DIAG) keep = keep && (j > r.idx[k]);
            u = keep ? u : kMagicBits + ovf + 1u;
        }
        if (WEIGHT) lds_add_at(r.base[k] + (u << 3), weight_code(r.tagf[k] * tag_j, wscale));
        else lds_add_at(r.base[k] + (u << 3), cell_add(offset_code(d2[k], s[k], dl[k], cs)));
    }
}

// All kRowsPerLane pairs of one column point.
template <bool FULL, bool DIAG, bool ORIENT, bool CLAMP, bool SPLIT, bool WEIGHT>
__device__ __forceinline__ void pair_column(const Rows& r, const float4 pj, int j, uint64_t* hist, const Consts& cs,
                                            unsigned ovf, float off1, double wscale) {
    float d2[kRowsPerLane];
#pragma unroll
    for (int k = 0; k < kRowsPerLane; ++k) d2[k] = simd_d2(r.x[k] - pj.x, r.y[k] - pj.y, r.z[k] - pj.z);
    add_pairs<FULL, DIAG, ORIENT, CLAMP, WEIGHT>(r, d2, pj.w, j, cs, ovf, off1, wscale);
}

// The row points of this lane for item w (scaled coordinates, tags, LDS row
// bases). Lanes of one wave take points 8 apart: spatially spread rows give
// spread distances to a common column point, so fewer lanes collide on one
// LDS cell.
template <bool FULL, bool SPLIT, bool WEIGHT>
__device__ __forceinline__ void load_rows(const float4* __restrict__ pts, const PairWork& w, uint64_t* hist,
                                          const Consts& cs, int nbins, Rows& r) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < kRowsPerLane; ++k) {
        const int li = k * kRowLanes + (t & (kRowLanes - 1));
        const int i = w.i_begin + ((li & 63) * 8 + (li >> 6));
        r.valid[k] = FULL || i < w.i_end;
        const float4 p = r.valid[k] ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        r.x[k] = p.x * cs.scale;
        r.y[k] = p.y * cs.scale;
        r.z[k] = p.z * cs.scale;
        r.tagf[k] = WEIGHT ? p.w : (float)__float_as_int(p.w);
        r.idx[k] = i;
        // kSplit: rows from the second class add into the second LDS row
        const int roff = SPLIT && i >= w.i_begin + split_rows(w.flags) ? nbins + 2 : 0;
        r.base[k] = lds_byte_addr(hist + roff) - kMagicBits * 8u;
    }
}

// One work item whose columns are z-runs of lattice points (kRuns). For each
// run the lanes form dx^2 + dy^2 once per row; along the run only dz changes,
// so a pair costs dz, dz^2 and one add before the binning (the reference's
// ((dx dx + dy dy) + dz dz), separately rounded, as for any other pair). The
// run's z advances by the scaled lattice step, exactly. Run descriptors pass
// through LDS 512 at a time; the parts take alternate runs.
template <bool FULL, bool CLAMP, bool SPLIT>
__device__ __forceinline__ void run_item_runs(const float4* __restrict__ pts, const float4* __restrict__ runs,
                                              const PairWork& w, uint64_t* hist, float4* stage, const Consts& cs,
                                              int nbins) {
    const int t = threadIdx.x;
    const int part = t / kRowLanes;
    Rows r;
    load_rows<FULL, SPLIT, false>(pts, w, hist, cs, nbins, r);
    const unsigned ovf = (unsigned)nbins;
    const Consts cv = vgpr_consts(cs);
    const float zstep = vgpr_float(cs.scale);  // the run's z step (scaled lattice step)
    for (int rb = w.j_begin; rb < w.j_end; rb += kColTile) {
        const int n = min(kColTile, w.j_end - rb);
        __syncthreads();  // previous descriptors consumed
        if (t < n) {
            const float4 d = runs[rb + t];
            stage[t] = make_float4(d.x * cs.scale, d.y * cs.scale, d.z * cs.scale, d.w);
        }
        __syncthreads();
        for (int u = part; u < n; u += kParts) {
            // the run is the same for every lane of the wave: scalar registers
            const float4 dv = stage[u];
            const float dxr = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(dv.x)));
            const float dyr = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(dv.y)));
            const int len = __builtin_amdgcn_readfirstlane(__float_as_int(dv.w));
            float sxy[kRowsPerLane];
#pragma unroll
            for (int k = 0; k < kRowsPerLane; ++k) {
                const float dx = r.x[k] - dxr, dy = r.y[k] - dyr;
                sxy[k] = dx * dx + dy * dy;
            }
            // (read into a scalar register first: the loop then waits for no LDS load)
            float z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(dv.z)));
            for (int c = 0; c < len; ++c) {
                float d2[kRowsPerLane];
#pragma unroll
                for (int k = 0; k < kRowsPerLane; ++k) {
                    const float dz = r.z[k] - z;
                    d2[k] = sxy[k] + dz * dz;
                }
                add_pairs<FULL, false, false, CLAMP, false>(r, d2, 0.f, 0, cv, ovf, 0.f, 1.0);
                z += zstep;
            }
        }
    }
}

// One work item. Column points pass through LDS 512 at a time, each part
// taking a share of the tile; within its share a part consumes them in
// double-buffered groups of kColGroup: the reads of one
// group are issued before the atomics of the other, so a wait for them never
// waits for atomics (LDS operations retire in order) and no register copies
// are needed between groups.
template <bool FULL, bool DIAG, bool ORIENT, bool CLAMP, bool SPLIT = false, bool WEIGHT = false>
__device__ __forceinline__ void run_item(const float4* __restrict__ pts, const PairWork& w, uint64_t* hist,
                                         float4* cols, const Consts& cs, int nbins, double wscale) {
    const int t = threadIdx.x;
    const int part = t / kRowLanes;  // every part holds the same row points
    Rows r;
    load_rows<FULL, SPLIT, WEIGHT>(pts, w, hist, cs, nbins, r);
    // row stride: slot nbins overflow, nbins+1 masked lanes
    const float off1 = ORIENT ? vgpr_float((float)(nbins + 2)) : 0.f;
    const unsigned ovf = (unsigned)nbins;
    const Consts cv = vgpr_consts(cs);

    // the next column tile is loaded into a register while the current one is
    // evaluated, so the global-memory latency of the tile loads stays hidden
    // when the point buffer does not fit the caches (large batches)
    // (an unconditional load of a clamped, always valid index: a conditional
    // one would be merged through a register copy that waits for it at once)
    float4 pre = pts[min(w.j_begin + t, w.j_end - 1)];
    for (int jc = w.j_begin; jc < w.j_end; jc += kColTile) {
        const int nc = min(kColTile, w.j_end - jc);
        __syncthreads();  // previous tile consumed
        if (t < nc)
            cols[t] = make_float4(pre.x * cs.scale, pre.y * cs.scale, pre.z * cs.scale,
                                  ORIENT ? (float)__float_as_int(pre.w) : pre.w);
        __syncthreads();
        pre = pts[min(jc + kColTile + t, w.j_end - 1)];  // past the item's end: unused
        // the tile's columns split between the parts: part p takes
        // [p m, (p + 1) m) clipped to nc, m a multiple of 2 * kColGroup (the
        // last parts may be empty)
        const int m = (nc + 2 * kColGroup * kParts - 1) / (2 * kColGroup * kParts) * (2 * kColGroup);
        const int c0 = min(part * m, nc);
        const int n = min(c0 + m, nc) - c0;
        const float4* cc = cols + c0;
        const int jh = jc + c0;
        const int n2 = n & ~(2 * kColGroup - 1);
        float4 a[kColGroup], bq[kColGroup];
#pragma unroll
        for (int u = 0; u < kColGroup; ++u) a[u] = cc[u];
        // completes the first reads before the loop, so the loop needs no
        // wait that would also drain atomics in flight
        __syncthreads();
        for (int c = 0; c < n2; c += 2 * kColGroup) {
#pragma unroll
            for (int u = 0; u < kColGroup; ++u) bq[u] = cc[c + kColGroup + u];
#pragma unroll
            for (int u = 0; u < kColGroup; ++u)
                pair_column<FULL, DIAG, ORIENT, CLAMP, SPLIT, WEIGHT>(r, a[u], jh + c + u, hist, cv, ovf, off1, wscale);
            // may read up to kColGroup slots past the tile: cols is padded
#pragma unroll
            for (int u = 0; u < kColGroup; ++u) a[u] = cc[c + 2 * kColGroup + u];
#pragma unroll
            for (int u = 0; u < kColGroup; ++u)
                pair_column<FULL, DIAG, ORIENT, CLAMP, SPLIT, WEIGHT>(r, bq[u], jh + c + kColGroup + u, hist, cv, ovf, off1,
                                                                     wscale);
        }
        for (int c = n2; c < n; ++c)
            pair_column<FULL, DIAG, ORIENT, CLAMP, SPLIT, WEIGHT>(r, cc[c], jh + c, hist, cv, ovf, off1, wscale);
    }
}

template <bool CLAMP>
__device__ __forceinline__ void run_any(const float4* __restrict__ pts, const float4* __restrict__ runs,
                                        const PairWork& w, uint64_t* hist, float4* cols, const Consts& cs,
                                        int nbins) {
    const bool orient = (w.flags & kOrient) != 0;
    if (w.flags & kRuns) {
        // a cross job against lattice columns (never oriented)
        const bool full = w.i_end - w.i_begin >= kTileI;
        if (w.flags & kSplit) {
            if (full) run_item_runs<true, CLAMP, true>(pts, runs, w, hist, cols, cs, nbins);
            else run_item_runs<false, CLAMP, true>(pts, runs, w, hist, cols, cs, nbins);
        } else {
            if (full) run_item_runs<true, CLAMP, false>(pts, runs, w, hist, cols, cs, nbins);
            else run_item_runs<false, CLAMP, false>(pts, runs, w, hist, cols, cs, nbins);
        }
    } else if (w.flags & kDiag) {
        if (orient) run_item<false, true, true, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
        else run_item<false, true, false, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
    } else if (w.flags & kSplit) {
        // a row tile spanning two classes of a cross job (never oriented)
        if (w.i_end - w.i_begin >= kTileI) run_item<true, false, false, CLAMP, true>(pts, w, hist, cols, cs, nbins, 1.0);
        else run_item<false, false, false, CLAMP, true>(pts, w, hist, cols, cs, nbins, 1.0);
    } else if (w.i_end - w.i_begin >= kTileI) {
        if (orient) run_item<true, false, true, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
        else run_item<true, false, false, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
    } else {
        if (orient) run_item<false, false, true, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
        else run_item<false, false, false, CLAMP>(pts, w, hist, cols, cs, nbins, 1.0);
    }
}

// The weight pass: single-class point sets (no oriented or split items).
template <bool CLAMP>
__device__ __forceinline__ void run_weighted(const float4* __restrict__ pts, const PairWork& w, uint64_t* hist,
                                             float4* cols, const Consts& cs, int nbins, double wscale) {
    if (w.flags & kDiag) run_item<false, true, false, CLAMP, false, true>(pts, w, hist, cols, cs, nbins, wscale);
    else if (w.i_end - w.i_begin >= kTileI)
        run_item<true, false, false, CLAMP, false, true>(pts, w, hist, cols, cs, nbins, wscale);
    else run_item<false, false, false, CLAMP, false, true>(pts, w, hist, cols, cs, nbins, wscale);
}

// One workgroup per work item (grid = items), or a persistent grid that
// strides over the item list (grid < items): both visit every item once.
// clamp = 0 is allowed only when the host has bounded every pair's bin below
// nbins (bounding-box diagonal); otherwise bins are clamped into slot nbins.
// The LDS holds `copies` histograms of `rows` rows each; lane l of a wave adds
// into copy l mod copies, so lanes whose pairs fall into the same bin collide
// on one LDS cell `copies` times less often (pairs of two bodies crowd into few
// bins). An item with one output row in a two-row launch uses the same LDS as
// 2 * copies one-row histograms. The copies are summed at the flush. The
// kernel may use up to 80 VGPRs (6 waves per SIMD, what the LDS of a 0.1 A
// workload allows anyway); with column groups of 2 it needs 64 and spills
// nothing (groups of 4: 80 VGPRs and 0.7 % slower on the 250k-atom workload;
// 1,024-thread workgroups at 8 waves per SIMD: 2 % slower).
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(6, 8))) void pair_hist_kernel(const float4* __restrict__ pts,
                                                             const float4* __restrict__ runs,
                                                             const PairWork* __restrict__ items, int n_items,
                                                             uint64_t* __restrict__ partial, Consts cs,
                                                             int nbins, int clamp, int rows, int copies,
                                                             int wmode, double wscale) {
    extern __shared__ uint64_t hist[];
    __shared__ float4 cols[kColTile + kColGroup];
    const int stride = nbins + 2;
    const int out_stride = nbins + 1;
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const PairWork w = items[it];
        const bool two = (w.flags & (kOrient | kSplit)) != 0;
        // this item's histograms: n_copies of item_rows rows each
        const int item_rows = two ? 2 : 1;
        const int n_copies = two ? copies : copies * rows;
        const int copy_stride = item_rows * stride;
        uint64_t* mine = hist + (threadIdx.x & (n_copies - 1)) * copy_stride;
        __syncthreads();  // previous item's flush done
        for (int e = threadIdx.x; e < copies * rows * stride; e += kThreads) hist[e] = 0;
        __syncthreads();
        if (wmode) {
            if (clamp) run_weighted<true>(pts, w, mine, cols, cs, nbins, wscale);
            else run_weighted<false>(pts, w, mine, cols, cs, nbins, wscale);
        } else if (clamp) {
            run_any<true>(pts, runs, w, mine, cols, cs, nbins);
        } else {
            run_any<false>(pts, runs, w, mine, cols, cs, nbins);
        }
        __syncthreads();
        // partial slot layout: [2][nbins+1] (the masked-lane slot is dropped)
        uint64_t* out = partial + (size_t)w.slot * 2 * out_stride;
        for (int rr = 0; rr < item_rows; ++rr)
            for (int e = threadIdx.x; e < out_stride; e += kThreads) {
                uint64_t v = hist[rr * stride + e];
                for (int c = 1; c < n_copies; ++c) v += hist[c * copy_stride + rr * stride + e];
                out[rr * out_stride + e] = v;
            }
    }
}

// Merge phase 1: chunk c = sum of its partial rows row_list[chunk_ptr[c]] ..
// (at most 32), one thread per bin. Each batch issues all its loads before
// decoding any of them (the decode has no branches), so a thread keeps
// kMergeBatch loads in flight instead of waiting on each in turn.
__global__ __launch_bounds__(256) void merge_chunks_kernel(const uint64_t* __restrict__ partial,
                                                           const int32_t* __restrict__ row_list,
                                                           const int32_t* __restrict__ chunk_ptr, int nbins,
                                                           uint64_t* __restrict__ tcnt, int64_t* __restrict__ tdsum,
                                                           int raw) {
    const int c = blockIdx.y;
    const int stride = nbins + 1;
    const int bin = blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= stride) return;
    const uint64_t* __restrict__ col = partial + bin;
    const int e0 = chunk_ptr[c], e1 = chunk_ptr[c + 1];
    if (raw) {
        // weight pass: the items' exact integer sums, added in FP64 in a fixed order
        double w = 0;
        for (int e = e0; e < e1; e += kMergeBatch) {
            int64_t v[kMergeBatch];
#pragma unroll
            for (int u = 0; u < kMergeBatch; ++u) v[u] = (int64_t)col[(size_t)row_list[e + u] * stride];
#pragma unroll
            for (int u = 0; u < kMergeBatch; ++u) w += e + u < e1 ? (double)v[u] : 0.0;
        }
        tcnt[(size_t)c * stride + bin] = 0;
        tdsum[(size_t)c * stride + bin] = __double_as_longlong(w);
        return;
    }
    uint64_t n_sum = 0, low_sum = 0;
    for (int e = e0; e < e1; e += kMergeBatch) {
        uint64_t v[kMergeBatch];
        // unconditional loads (row_list has kMergeBatch - 1 entries of slack);
        // entries past the chunk are dropped as zero cells, which decode to 0, 0
#pragma unroll
        for (int u = 0; u < kMergeBatch; ++u) v[u] = col[(size_t)row_list[e + u] * stride];
#pragma unroll
        for (int u = 0; u < kMergeBatch; ++u) {
            uint64_t n, k;
            cell_decode(e + u < e1 ? v[u] : 0, n, k);
            n_sum += n;
            low_sum += k;
        }
    }
    tcnt[(size_t)c * stride + bin] = n_sum;
    tdsum[(size_t)c * stride + bin] = (int64_t)low_sum - (int64_t)n_sum * kDeltaBias;
}

// Merge phase 2: output row r = sum of its chunks row_chunk[r] .. row_chunk[r+1].
__global__ __launch_bounds__(256) void merge_rows_kernel(const uint64_t* __restrict__ tcnt,
                                                         const int64_t* __restrict__ tdsum,
                                                         const int32_t* __restrict__ row_chunk, int nbins,
                                                         uint64_t* __restrict__ cnt, int64_t* __restrict__ dsum,
                                                         int raw) {
    const int row = blockIdx.y;
    const int stride = nbins + 1;
    const int bin = blockIdx.x * blockDim.x + threadIdx.x;
    if (bin >= stride) return;
    uint64_t c = 0;
    int64_t s = 0;
    const int c0 = row_chunk[row], c1 = row_chunk[row + 1];
    if (raw) {  // weight pass: FP64 chunk sums (bits of doubles)
        double w = 0;
        for (int k = c0; k < c1; ++k) w += __longlong_as_double(tdsum[(size_t)k * stride + bin]);
        cnt[(size_t)row * stride + bin] = 0;
        dsum[(size_t)row * stride + bin] = __double_as_longlong(w);
        return;
    }
#pragma unroll 4
    for (int k = c0; k < c1; ++k) {
        c += tcnt[(size_t)k * stride + bin];
        s += tdsum[(size_t)k * stride + bin];
    }
    cnt[(size_t)row * stride + bin] = c;
    dsum[(size_t)row * stride + bin] = s;
}

// One strip pair with the reference's per-pair rule: bin (clamped into the
// overflow slot nbins) and offset code (kLowBias + k) of pair (row i, column
// t). The root is correctly rounded, so its offset needs no correction.
__device__ __forceinline__ void strip_pair(const StripLaunch& p, const Consts& cs, const float4& a, const float4& c,
                                           int i, int t, unsigned& bin, uint32_t& code) {
    const StripWork& w = p.w;
    const float dx = a.x - c.x, dy = a.y - c.y, dz = a.z - c.z;
    bool scalar_rule = true;
    if (w.self) scalar_rule = t >= w.n_cols_total - ((w.n_cols_total - i - 1) % 4);
    float b;
    if (scalar_rule) {
        // evaluate_scalar with FMA contraction and round() lowered to
        // trunc(x + 0.49999997f) (CompactCoordinatesXYZFF.h:313-329)
        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
        const float r = sqrt_rn(d2);
        b = __builtin_truncf(r * p.inv_w + 0.49999997f);
        code = __float_as_uint(__builtin_fmaf(__builtin_fmaf(r, p.inv_w, -b), cs.q, cs.m2));
    } else {
        const float d2 = (dx * dx + dy * dy) + dz * dz;
        code = simd_rule_exact(d2, p.inv_w, cs, b);
    }
    bin = min((unsigned)(int)b, (unsigned)p.nbins);
}

// One strip pair per thread (pair tid of replica rep), added with global atomics.
__device__ __forceinline__ void strip_one(const StripLaunch& p, long long tid, int rep) {
    const StripWork& w = p.w;
    const int nt = w.t_end - w.t_begin;
    if (tid >= (long long)w.n_rows * nt) return;
    const int i = (int)(tid / nt);
    const int t = w.t_begin + (int)(tid % nt);
    if (w.self && i >= t) return;
    const float4 a = p.rows[p.rows_rep * rep + i];
    const float4 c = p.cols[p.cols_rep * rep + t];
    int row = p.wmode ? p.rowmap[0] : p.rowmap[__float_as_int(a.w) * w.n_col_classes + __float_as_int(c.w)];
    if (row < 0) return;  // the host guarantees a row for every present class pair
    row += rep * p.out_rep;
    unsigned bin;
    uint32_t code;
    strip_pair(p, make_consts(p.inv_w), a, c, i, t, bin, code);
    const size_t cell = (size_t)row * p.stride + bin;
    if (p.wmode) {
        atomicAdd((unsigned long long*)&p.dsum[cell], (unsigned long long)weight_code(a.w * c.w, p.wscale));
        return;
    }
    atomicAdd((unsigned long long*)&p.cnt[cell], 1ull);
    atomicAdd((unsigned long long*)&p.dsum[cell], (unsigned long long)((int64_t)(code - (uint32_t)kLowBias) - kDeltaBias));
}

// Strip kernel: rows x tail columns, one thread per pair, global atomics.
__global__ __launch_bounds__(256) void strip_kernel(StripLaunch p) {
    strip_one(p, (long long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// Many strips in one launch: strip k owns blocks [first[k], first[k+1]) of a
// one-dimensional grid, 256 pairs per block (replica 0 only). A block finds
// its strip by binary search over first; the descriptor is read from global
// memory (every thread of the block reads the same bytes).
__global__ __launch_bounds__(256) void strip_table_kernel(const StripLaunch* table, const int64_t* first, int n) {
    const int64_t blk = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const StripLaunch p = table[lo];
    strip_one(p, (blk - first[lo]) * 256 + threadIdx.x, 0);
}

// Block `block` of a strip with one output row (fixed_row) and no replicas:
// its kStripBlockPairs pairs accumulate in an LDS histogram of packed cells
// (as in the pair kernel), flushed with two global atomics per occupied bin.
// Long strips (a dummy set's rows against its last columns) crowd into few
// bins, where one global atomic per pair serialises. Every thread of the
// workgroup calls it (it synchronises the workgroup).
__device__ __forceinline__ void strip_lds_block(const StripLaunch& p, long long block, uint64_t* cells) {
    const StripWork& w = p.w;
    const int nt = w.t_end - w.t_begin;
    for (int b = threadIdx.x; b <= p.nbins; b += blockDim.x) cells[b] = 0;
    __syncthreads();
    const long long total = (long long)w.n_rows * nt;
    const long long start = block * kStripBlockPairs;
    const long long end = min(start + kStripBlockPairs, total);
    const Consts cs = make_consts(p.inv_w);
    for (long long tid = start + threadIdx.x; tid < end; tid += blockDim.x) {
        const int i = (int)(tid / nt);
        const int t = w.t_begin + (int)(tid % nt);
        if (w.self && i >= t) continue;
        unsigned bin;
        uint32_t code;
        const float4 a = p.rows[i], c = p.cols[t];
        strip_pair(p, cs, a, c, i, t, bin, code);
        lds_add(&cells[bin], p.wmode ? weight_code(a.w * c.w, p.wscale) : cell_add(code));
    }
    __syncthreads();
    const size_t row = (size_t)p.fixed_row * p.stride;
    for (int b = threadIdx.x; b <= p.nbins; b += blockDim.x) {
        const uint64_t v = cells[b];
        if (!v) continue;
        if (p.wmode) {  // weight pass: the cell's integer sum
            atomicAdd((unsigned long long*)&p.dsum[row + b], (unsigned long long)v);
            continue;
        }
        uint64_t n, ksum;
        cell_decode(v, n, ksum);
        atomicAdd((unsigned long long*)&p.cnt[row + b], (unsigned long long)n);
        atomicAdd((unsigned long long*)&p.dsum[row + b],
                  (unsigned long long)((int64_t)ksum - (int64_t)n * kDeltaBias));
    }
}

__global__ __launch_bounds__(256) void strip_lds_kernel(StripLaunch p) {
    extern __shared__ uint64_t cells[];  // [nbins + 1]
    strip_lds_block(p, blockIdx.x, cells);
}

// Many LDS strips in one launch (the layout of strip_table_kernel with
// kStripBlockPairs pairs per block); the block's strip is uniform over the
// workgroup, so all its threads reach strip_lds_block's barriers.
__global__ __launch_bounds__(256) void strip_lds_table_kernel(const StripLaunch* table, const int64_t* first, int n) {
    extern __shared__ uint64_t cells[];  // [nbins + 1]
    const int64_t blk = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const StripLaunch p = table[lo];
    strip_lds_block(p, blk - first[lo], cells);
}

__global__ __launch_bounds__(256) void bin_check_kernel(uint32_t lo, uint32_t hi, float inv_w,
                                                        unsigned long long* out) {
    unsigned long long bad_fast = 0, bad_cr = 0, bad_fix = 0, bad_off = 0;
    long long err_sum = 0, err_cr_sum = 0;  // units of 2^-16 offset units
    const Consts cs = make_consts(inv_w);
    const float sc2 = cs.scale * cs.scale;
    const double unit_inv = ldexp(1.0, delta_exp(inv_w));
    // two roundings to whole units, and the rounding of dl and of q: 1 + 2^-24 q
    const double off_tol = 1.0 + ldexp((double)cs.q, -24) + 1e-6;
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
        // the pair kernel's path on the scaled d2 (exact: a power of four)
        float t;
        const uint32_t code = bin_pair(d2 * sc2, cs, t);
        if (t - kMagic != br) ++bad_fast;
        // its offset against the real root's, computed in FP64
        const double b = (double)(t - kMagic);
        const double k_true = (sqrt(dd) - b / (double)inv_w) * unit_inv + (double)kDeltaBias;
        const double err = (double)(code - (uint32_t)kLowBias) - k_true;
        if (fabs(err) > off_tol) ++bad_off;
        err_sum += (long long)rint(err * 65536.0);
        err_cr_sum += (long long)rint((rd - sqrt(dd)) * unit_inv * 65536.0);
        // the residual correction of the tie fallback, on every normal d2
        if (d2 >= 1e-30f && sqrt_fix(d2, __builtin_amdgcn_sqrtf(d2)) != r) ++bad_fix;
    }
    if (bad_fast) atomicAdd(&out[0], bad_fast);
    if (bad_cr) atomicAdd(&out[1], bad_cr);
    if (bad_fix) atomicAdd(&out[2], bad_fix);
    if (bad_off) atomicAdd(&out[3], bad_off);
    if (err_sum) atomicAdd(&out[4], (unsigned long long)err_sum);
    if (err_cr_sum) atomicAdd(&out[5], (unsigned long long)err_cr_sum);
}

// plain streaming copy of a known number of bytes (calibration of HBM traffic counters)
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

// Gathers up to kCopyOutMax device buffers, back to back, into one buffer
// (pinned host memory: the kernel's stores cross the bus, no DMA engine)
__global__ __launch_bounds__(256) void copy_out_kernel(CopyOut c) {
    size_t base = 0;
    for (int b = 0; b < c.n; ++b) {
        const uint32_t* __restrict__ src = c.src[b];
        const size_t n = c.words[b];
        for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
            c.dst[base + e] = src[e];
        base += n;
    }
}

}  // namespace

hipError_t launch_copy_out(const CopyOut& c, hipStream_t s) {
    if (c.n < 0 || c.n > kCopyOutMax || !c.dst) return hipErrorInvalidValue;
    size_t total = 0;
    for (int b = 0; b < c.n; ++b) total += c.words[b];
    if (!total) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 256);
    hipLaunchKernelGGL(copy_out_kernel, dim3(blocks), dim3(256), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / sizeof(uint4);
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n);
    return hipGetLastError();
}

hipError_t launch_pair_hist(const PairLaunch& p, hipStream_t s) {
    if (p.n_items == 0) return hipSuccess;
    if ((p.rows != 1 && p.rows != 2) || p.copies < 1 || p.copies > 8 || (p.copies & (p.copies - 1)))
        return hipErrorInvalidValue;
    const size_t lds = pair_lds_bytes(p.nbins, p.rows, p.copies) - kStaticLds;
    const int grid = p.grid > 0 ? std::min(p.grid, p.n_items) : p.n_items;
    hipLaunchKernelGGL(pair_hist_kernel, dim3(grid), dim3(kThreads), lds, s, p.pts, p.runs, p.items, p.n_items, p.partial,
                       make_consts(p.inv_w), p.nbins, p.clamp, p.rows, p.copies, p.wmode, p.wscale);
    return hipGetLastError();
}

hipError_t launch_merge(const MergeLaunch& m, hipStream_t s) {
    if (m.n_rows == 0) return hipSuccess;
    const unsigned bx = (unsigned)((m.nbins + 1 + 255) / 256);
    if (m.direct) {
        // the one chunk of each row is the row (the sums merge_rows would form: 0 + the chunk)
        if (m.n_chunks != m.n_rows) return hipErrorInvalidValue;
        hipLaunchKernelGGL(merge_chunks_kernel, dim3(bx, m.n_chunks), dim3(256), 0, s, m.partial, m.row_list,
                           m.chunk_ptr, m.nbins, m.cnt, m.dsum, m.raw);
        return hipGetLastError();
    }
    if (m.n_chunks > 0) {
        hipLaunchKernelGGL(merge_chunks_kernel, dim3(bx, m.n_chunks), dim3(256), 0, s, m.partial, m.row_list,
                           m.chunk_ptr, m.nbins, m.tmp_cnt, m.tmp_dsum, m.raw);
        if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(merge_rows_kernel, dim3(bx, m.n_rows), dim3(256), 0, s, m.tmp_cnt, m.tmp_dsum, m.row_chunk,
                       m.nbins, m.cnt, m.dsum, m.raw);
    return hipGetLastError();
}

hipError_t launch_strip(const StripLaunch& p, hipStream_t s) {
    const long long n = (long long)p.w.n_rows * (p.w.t_end - p.w.t_begin);
    if (n <= 0) return hipSuccess;
    if (p.n_rep <= 0) return hipSuccess;
    if (strip_uses_lds(p)) {
        const unsigned blocks = (unsigned)strip_lds_blocks(p);
        hipLaunchKernelGGL(strip_lds_kernel, dim3(blocks), dim3(256), (size_t)(p.nbins + 1) * sizeof(uint64_t), s,
                           p);
        return hipGetLastError();
    }
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(strip_kernel, dim3(blocks, p.n_rep), dim3(256), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_strip_table(const StripLaunch* table, const int64_t* first, int n, int64_t blocks,
                              hipStream_t s) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(strip_table_kernel, dim3((unsigned)blocks), dim3(256), 0, s, table, first, n);
    return hipGetLastError();
}

hipError_t launch_strip_lds_table(const StripLaunch* table, const int64_t* first, int n, int64_t blocks, int nbins,
                                  hipStream_t s) {
    if (n <= 0 || blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffff || nbins < 0 || (size_t)(nbins + 1) * sizeof(uint64_t) > 65536) return hipErrorInvalidValue;
    hipLaunchKernelGGL(strip_lds_table_kernel, dim3((unsigned)blocks), dim3(256), (size_t)(nbins + 1) * sizeof(uint64_t),
                       s, table, first, n);
    return hipGetLastError();
}

hipError_t launch_bin_check(uint32_t lo_bits, uint32_t hi_bits, float inv_w, unsigned long long* out,
                            hipStream_t s) {
    hipLaunchKernelGGL(bin_check_kernel, dim3(4096), dim3(256), 0, s, lo_bits, hi_bits, inv_w, out);
    return hipGetLastError();
}

}  // namespace ausaxs::gpu
