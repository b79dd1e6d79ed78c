// Host side of the pair-histogram engine: HIP context, resident point sets,
// work-list construction and kernel launches.
#pragma once

#include "device_api.hpp"
#include "pool.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ausaxs::gpu {

void check(hipError_t e, const char* what);

// bit casts for the fourth point slot (class id / original index stored as int bits)
inline int32_t bits_of(float f) {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i;
}
inline float float_of(int32_t i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

// Process-wide HIP context on the current device. Work is issued on the
// calling thread's stream: a serving slot's own stream when the thread has
// one (set_thread_stream), else the context's default stream. All streams
// have the default priority.
struct Context {
    int device = 0;
    hipStream_t default_stream = nullptr;
    hipDeviceProp_t prop{};
    static Context& get();
    static void set_thread_stream(hipStream_t s);
    hipStream_t stream() const;
    void sync() const { check(hipStreamSynchronize(stream()), "hipStreamSynchronize"); }

  private:
    Context();
};

// Device buffer from the caching pool (pool.hpp): growing it or dropping it
// hands the old block back to the pool, not to HIP.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
    ~DevBuf() { pool_free_device(p); }
    void ensure(size_t count) {
        if (count <= n && p) return;
        pool_free_device(p);
        p = nullptr;
        n = 0;
        p = static_cast<T*>(pool_alloc_device(std::max<size_t>(count, 1) * sizeof(T)));
        n = std::max<size_t>(count, 1);
    }
    size_t bytes() const { return p ? pool_size_class(n * sizeof(T)) : 0; }
    void upload(const T* src, size_t count, hipStream_t s) {
        ensure(count);
        if (count) check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    }
};

// Pinned host buffer (from the caching pool), for copies captured in graphs.
template <class T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { pool_free_host(p); }
    void ensure(size_t count) {
        if (count <= n && p) return;
        pool_free_host(p);
        p = nullptr;
        n = 0;
        p = static_cast<T*>(pool_alloc_host(std::max<size_t>(count, 1) * sizeof(T)));
        n = std::max<size_t>(count, 1);
    }
};

// A per-thread pinned host buffer that kernels read and write directly
// through its device address (small operands and results cross the bus in
// the kernels' own loads and stores: no DMA transfer to queue behind other
// streams' copies). Grows by half again when too small.
struct MappedStage {
    HostBuf<char> buf;
    void* dev = nullptr;
    char* ensure(size_t bytes) {
        if (bytes > buf.n || !buf.p) {
            buf.ensure(std::max(bytes, buf.n + buf.n / 2));
            check(hipHostGetDevicePointer(&dev, buf.p, 0), "hipHostGetDevicePointer");
        }
        return buf.p;
    }
};

// A point set in the reference's original order with a class id per point
// (form-factor type for XYZFF sets, weight class for XYZW sets).
struct PointSet {
    std::vector<float4> orig;     // x, y, z, class id (int bits)
    int n_classes = 1;
    int size() const { return (int)orig.size(); }
};

// A pair job: self pairs of one set, or rows x columns of two sets. Global
// histogram row of a pair = rowmap[c_first * n_col_classes + c_second] where
// "first" is the row point (cross) or the lower original index (self).
struct PairJob {
    int row_set = 0;
    int col_set = -1;             // -1: self job
    std::vector<int32_t> rowmap;  // [n_row_classes][n_col_classes]
    // one class on each side and integer-lattice points (grid dummies): the
    // body pairs may be histogrammed by lattice correlation instead of the
    // pair kernel (the tail pairs keep their strips)
    bool lattice = false;
    // a cross job whose one-class column set holds integer lattice points in
    // z-runs (grid dummies in voxel order): its body columns go to the pair
    // kernel as runs (kRuns items), when the points allow it
    bool col_runs = false;
};

// Class-sorted layout of one point set inside a point buffer. Within each
// class the points follow the reference's loop geometry: body (all rules
// SIMD), then the self-loop tail candidates, then the cross-loop scalar
// columns (the last m mod 4 points).
struct SetLayout {
    int base = 0;                            // offset of the set in the buffer
    std::vector<int> seg_start, seg_len;     // per class: full segment
    std::vector<int> seg_body, seg_colbody;  // per class: self-body length, column-body length
};
// Appends the class-sorted points of ps to buf (w = original index bits);
// perm (optional) receives the original index of every appended point.
SetLayout layout_set(const PointSet& ps, std::vector<float4>& buf, std::vector<int32_t>* perm = nullptr);

// Tail pairs of one job, evaluated by the strip kernel on original-order points.
struct StripSpec {
    int set_rows, set_cols;
    StripWork w;
    int rowmap_off;  // into WorkList::rowmaps
};

// The z-runs of the first mb points of ps (original order; consecutive points
// (x, y, z), (x, y, z + 1), ...), each at most max_cols long, appended to
// runs as {x, y, z0, length bits}; false (nothing appended) unless every
// point is an integer lattice point and the runs average at least 4 points.
bool lattice_runs(const PointSet& ps, int mb, int max_cols, std::vector<float4>& runs);

// Work items, strips and pair count of a list of jobs.
struct WorkList {
    std::vector<float4> runs;  // z-runs of kRuns items (lattice_runs)
    std::vector<std::array<int, 3>> set_runs;  // {column set, first run, end run} (end -1: not runs)
    std::vector<PairWork> items;
    std::vector<StripSpec> strips;
    std::vector<int32_t> rowmaps;
    int64_t pairs = 0;          // all pairs of the jobs
    int64_t strip_pairs = 0;    // of which in strips
    int64_t lattice_pairs = 0;  // of which by lattice correlation
};
// split: row tiles of cross jobs may span two row classes (kSplit items);
// body_items false: the body pairs get no work items (lattice path)
void append_job(WorkList& wl, const PairJob& job, const std::vector<PointSet>& sets,
                const std::vector<SetLayout>& layouts, int n_rows, int tile_j, bool split, bool body_items = true);

// A job whose body pairs are histogrammed on the integer lattice: rows are
// the first n_rows points of row_set (original order), columns the first
// n_cols of col_set (self: col_set = -1, the row points themselves).
struct LatticeJob {
    int row_set = 0, col_set = -1;
    int n_rows = 0, n_cols = 0;
    int out_row = 0;
    LatticeGeom g{};
    int smax = 0;
    int tw_off[3] = {0, 0, 0};  // twiddle tables in the engine's table buffer
    int64_t pairs = 0;
};
// The lattice job of a one-class job whose body points are all integers with
// a small enough span, or false. mode: 0 never, 1 when the body pairs
// outweigh the transforms, 2 whenever possible.
bool plan_lattice(const PairJob& job, const std::vector<PointSet>& sets, int mode, LatticeJob& out);

// Row work the pair kernel spends on a list: kTileI row lanes per item times
// its columns (padding of partial row tiles included).
int64_t padded_work(const std::vector<PairWork>& items);
// Whether the list with split row tiles is worth its cost. kSplit items need a
// second LDS histogram row, which halves the workgroups a CU holds, so when
// the plain list needs one row only, split tiles must save at least a tenth
// of the padded work. Otherwise they must save 2%: on a 250k-atom structure
// (padding below 0.1%) split tiles measured 5% slower at 8000-column tiles.
bool prefer_split(int64_t work_plain, int rows_plain, int64_t work_split, int rows_split);
// largest items first: the tail of a launch is made of short items
void sort_items(std::vector<PairWork>& items);

// Merge plan: partial entries grouped by output row and cut into chunks.
struct MergePlan {
    std::vector<int32_t> row_list, chunk_ptr, row_chunk;
    int n_chunks() const { return (int)chunk_ptr.size() - 1; }
    // every row has exactly one chunk (then chunk r is row r)
    bool one_chunk_per_row() const {
        for (size_t r = 0; r + 1 < row_chunk.size(); ++r)
            if (row_chunk[r + 1] - row_chunk[r] != 1) return false;
        return row_chunk.size() > 1;
    }
};
// assigns slot = index to every item and plans the merge into n_rows rows
MergePlan plan_merge(std::vector<PairWork>& items, int n_rows);

// throws unless a work item of 512 rows x tile_j columns stays below the
// fixed-point capacity of one LDS cell
void check_tile_j(int tile_j);
// LDS bytes per workgroup of pair_hist_kernel for nbins bins and 2 rows;
// throws if the device offers less.
size_t check_pair_lds(int nbins);
// LDS histogram copies for a launch: the most (up to 8) that keep a
// workgroup within 64 KB of LDS, so at least two fit on one CU (measured on
// the 20k-atom grid workload: 40 KB gave 5.79 ms per pair launch, 64 KB 5.44 ms,
// 96 KB 9.88 ms; the 250k-atom workload's single copy is unchanged)
int pick_copies(int nbins, int rows);
// 2 if any item has two output rows (orientation), else 1
int item_rows(const std::vector<PairWork>& items);

struct EngineTiming {
    float pair_ms = 0;          // pair_hist_kernel on the engine stream (HIP events)
    float merge_ms = 0;         // the two merge kernels after it
    float total_ms = 0;         // pair + merge + lattice + strips
    int64_t merge_bytes = 0;    // the merge's HBM bytes: partial rows read, chunk sums written and read, rows written
    int64_t pairs = 0;          // pairs in the histograms
    int64_t kernel_pairs = 0;   // of which evaluated by pair_hist_kernel
    int64_t lattice_pairs = 0;  // of which counted by lattice correlation
    int n_items = 0;
    int n_lattice = 0;          // lattice jobs
};

// Weight-pass results of a weighted engine, [n_rows][nbins+1] each: the work
// items' FP64 sums (bits of doubles) and the tail strips' integer sums of
// weight_code (device_api.hpp); weight_value combines them.
struct WeightSums {
    std::vector<int64_t> items, strips;
};

// What a PairEngine is built from. Structures keep theirs, so that the work
// of several small structures can run as one engine (one pair launch): their
// sets side by side, jobs renumbered, rows offset (merge_specs).
struct EngineSpec {
    std::vector<PointSet> sets;
    std::vector<PairJob> jobs;
    int n_rows = 0;
    float inv_w = 1;
    int nbins = 10;
    int tile_j = kDefaultTileJ;
    int lattice = 1;
    std::vector<std::vector<float>> weights;
    int clamp = -1;  // -1: from the sets' bounding box; else whether bins must be clamped
};
// A read-only view of an engine's inputs (an EngineSpec, or the sets and jobs
// a PairEngine already holds).
struct SpecView {
    const std::vector<PointSet>* sets = nullptr;
    const std::vector<PairJob>* jobs = nullptr;
    int n_rows = 0;
    float inv_w = 1;
    int nbins = 10;
    int tile_j = kDefaultTileJ;
    int lattice = 1;
    bool weighted = false;
    int clamp = -1;
};
inline SpecView view_of(const EngineSpec& s) {
    return SpecView{&s.sets, &s.jobs, s.n_rows, s.inv_w, s.nbins, s.tile_j, s.lattice, !s.weights.empty(), s.clamp};
}
// The specs side by side: spec k's sets, jobs and rows follow those of
// specs 0..k-1 (row0[k] = its first row); nbins is the largest, clamp the
// OR. All must share inv_w, tile_j and lattice and have no weights.
EngineSpec merge_specs(const std::vector<SpecView>& specs, std::vector<int>& row0);

// Coordinates resident on the device with their work lists built once; run()
// recomputes all jobs into [n_rows][nbins+1] (count, fixed-point delta sum).
// With per-point weights (one vector per set, all sets single-class) a second
// pass over the same work lists sums the pairs' weight products into
// [n_rows][nbins+1] (weight_code sums, see device_api.hpp).
class PairEngine {
  public:
    // lattice: lattice-path mode for PairJob::lattice jobs (see plan_lattice)
    PairEngine(std::vector<PointSet> sets, std::vector<PairJob> jobs, int n_rows, float inv_w, int nbins,
               int tile_j = kDefaultTileJ, int lattice = 1, std::vector<std::vector<float>> weights = {});
    explicit PairEngine(EngineSpec spec);
    ~PairEngine();
    PairEngine(const PairEngine&) = delete;
    PairEngine& operator=(const PairEngine&) = delete;

    // enqueue all kernels on the context stream; results stay on device
    void enqueue(bool timed);
    // copy results to host (synchronizes); weighted engines also fill wsum
    void download(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum, WeightSums* wsum = nullptr);
    void run(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum, bool timed = false, WeightSums* wsum = nullptr);
    bool weighted() const { return weighted_; }
    double wscale() const { return wscale_; }
    EngineTiming timing();

    int n_rows() const { return n_rows_; }
    int nbins() const { return nbins_; }
    float inv_w() const { return inv_w_; }
    int64_t pairs() const { return wl_.pairs; }
    // pairs the pair kernel evaluates (without strips and lattice jobs)
    int64_t kernel_pairs() const { return wl_.pairs - wl_.lattice_pairs - wl_.strip_pairs; }
    int n_lattice() const { return (int)lat_.size(); }
    int n_items() const { return (int)wl_.items.size(); }
    const uint64_t* device_cnt() const { return d_cnt_.p; }
    const int64_t* device_dsum() const { return d_dsum_.p; }
    bool overflowed(const std::vector<uint64_t>& cnt) const;
    // workgroups of the pair kernel: 0 = one per item, else a persistent grid
    void set_grid(int g) { grid_ = g; }
    bool clamped() const { return clamp_; }
    // device bytes the engine holds (pool size classes)
    size_t device_bytes() const;
    // the engine's inputs
    SpecView view() const {
        return SpecView{&sets_, &jobs_, n_rows_, inv_w_, nbins_, tile_j_, lattice_mode_, weighted_, clamp_ ? 1 : 0};
    }

  private:
    std::vector<PointSet> sets_;
    std::vector<PairJob> jobs_;
    std::vector<SetLayout> layout_;
    WorkList wl_;
    MergePlan mp_;
    int n_rows_, nbins_;
    float inv_w_;
    int tile_j_ = kDefaultTileJ;
    bool clamp_ = true;
    int grid_ = 0;
    int rows_ = 2, copies_ = 1;
    int lattice_mode_ = 1;
    std::vector<LatticeJob> lat_;
    DevBuf<double2> d_lat_a_, d_lat_b_, d_tw_;
    DevBuf<unsigned long long> d_scount_;
    DevBuf<int> d_lat_flag_;

    DevBuf<float4> d_pts_;
    DevBuf<float4> d_runs_;  // z-runs of the lattice column sets
    std::vector<DevBuf<float4>> d_orig_;
    // weight pass: the same buffers with weights in the fourth slot, and its sums
    std::vector<std::vector<float>> weights_;
    bool weighted_ = false;
    double wscale_ = 1;
    DevBuf<float4> d_pts_w_;
    std::vector<DevBuf<float4>> d_orig_w_;
    DevBuf<uint64_t> d_wcnt_;
    DevBuf<int64_t> d_wsum_, d_wstrip_;
    DevBuf<PairWork> d_items_;
    DevBuf<int32_t> d_row_list_, d_chunk_ptr_, d_row_chunk_, d_rowmaps_;
    DevBuf<uint64_t> d_partial_, d_cnt_, d_tmp_cnt_;
    DevBuf<int64_t> d_dsum_, d_tmp_dsum_;
    // The strips of one pass, resident, launched as at most two kernels (an
    // atomic table and an LDS table) instead of one launch per strip: a
    // launch waits for free compute units, which on a device busy with other
    // streams' pair kernels costs far more than the strip's own work.
    struct StripTables {
        DevBuf<StripLaunch> atomic, lds;
        DevBuf<int64_t> atomic_first, lds_first;
        int n_atomic = 0, n_lds = 0;
        int64_t atomic_blocks = 0, lds_blocks = 0;
        size_t bytes() const { return atomic.bytes() + lds.bytes() + atomic_first.bytes() + lds_first.bytes(); }
    };
    StripTables strips_, wstrips_;
    hipEvent_t ev_[4]{};
    bool have_timing_ = false;

    void build(int tile_j);
    void build_strip_tables(const std::vector<StripLaunch>& all, StripTables& t, hipStream_t s);
    void launch_strip_tables(const StripTables& t, hipStream_t s, const char* what) const;
};

// Upper bound on the bin of any pair in the given points (bounding-box
// diagonal), used to size the LDS histograms.
int bin_bound(const std::vector<const PointSet*>& sets, float inv_w);

}  // namespace ausaxs::gpu
