// Host side of the pair-histogram engine: HIP context, resident point sets,
// work-list construction and kernel launches.
#pragma once

#include "device_api.hpp"

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

// Process-wide HIP context on the current device: one default-priority stream.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    static Context& get();
    void sync() { check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

  private:
    Context();
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void ensure(size_t count) {
        if (count <= n && p) return;
        if (p) check(hipFree(p), "hipFree");
        p = nullptr;
        n = 0;
        check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
        n = std::max<size_t>(count, 1);
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        ensure(count);
        if (count) check(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
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
};

struct EngineTiming {
    float pair_ms = 0;   // pair_hist_kernel on the engine stream (HIP events)
    float total_ms = 0;  // pair + merge + strips
    int64_t pairs = 0;   // algorithmic pair evaluations
    int n_items = 0;
};

// Coordinates resident on the device with their work lists built once; run()
// recomputes all jobs into [n_rows][nbins+1] (count, fixed-point delta sum).
class PairEngine {
  public:
    PairEngine(std::vector<PointSet> sets, std::vector<PairJob> jobs, int n_rows, float inv_w, int nbins,
               int tile_j = 8000);
    PairEngine(const PairEngine&) = delete;
    PairEngine& operator=(const PairEngine&) = delete;

    // enqueue all kernels on the context stream; results stay on device
    void enqueue(bool timed);
    // copy results to host (synchronizes)
    void download(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum);
    void run(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum, bool timed = false);
    EngineTiming timing();

    int n_rows() const { return n_rows_; }
    int nbins() const { return nbins_; }
    float inv_w() const { return inv_w_; }
    int64_t pairs() const { return pairs_; }
    int n_items() const { return (int)items_.size(); }
    const uint64_t* device_cnt() const { return d_cnt_.p; }
    const int64_t* device_dsum() const { return d_dsum_.p; }
    bool overflowed(const std::vector<uint64_t>& cnt) const;
    // workgroups of the pair kernel: 0 = one per item, else a persistent grid
    void set_grid(int g) { grid_ = g; }
    bool clamped() const { return clamp_; }

  private:
    struct Layout {
        int base = 0;                            // offset in the point buffer
        std::vector<int> seg_start, seg_len;     // per class: full segment
        std::vector<int> seg_body, seg_colbody;  // per class: self-body length, column-body length
    };
    std::vector<PointSet> sets_;
    std::vector<PairJob> jobs_;
    std::vector<Layout> layout_;
    std::vector<PairWork> items_;
    std::vector<int32_t> row_list_, chunk_ptr_, row_chunk_;
    struct Strip {
        int set_rows, set_cols;
        StripWork w;
        int rowmap_off;
    };
    std::vector<Strip> strips_;
    std::vector<int32_t> rowmaps_;
    int n_rows_, nbins_;
    float inv_w_;
    bool clamp_ = true;
    int grid_ = 0;
    int64_t pairs_ = 0;

    DevBuf<float4> d_pts_;
    std::vector<DevBuf<float4>> d_orig_;
    DevBuf<PairWork> d_items_;
    DevBuf<int32_t> d_row_list_, d_chunk_ptr_, d_row_chunk_, d_rowmaps_;
    DevBuf<uint64_t> d_partial_, d_cnt_, d_tmp_cnt_;
    DevBuf<int64_t> d_dsum_, d_tmp_dsum_;
    hipEvent_t ev_[3]{};
    bool have_timing_ = false;

    void build(int tile_j);
    void add_items(int i0, int i1, int j0, int j1, int row0, int row1, bool diag, int tile_j);
};

// Upper bound on the bin of any pair in the given points (bounding-box
// diagonal), used to size the LDS histograms.
int bin_bound(const std::vector<const PointSet*>& sets, float inv_w);

}  // namespace ausaxs::gpu
