// Batched rigid-body scan: PartialHistogramManagerMT (PartialHistogramManagerMT.cpp:42-341)
// over many conformations in one launch sequence.
//
// Bodies are XYZW point sets (SimpleCalculator conventions) given once in
// FP64; each conformation moves every body by its own rigid transform. As in
// the reference manager, the self-correlation of each body (and of the
// waters) is computed once, from the coordinates given at creation, and only
// the inter-body and body-water blocks are recomputed per conformation
// (calc_aa / calc_aw, :96-108); blocks between bodies that do not move
// across the scan are computed once per scan. All conformations of a batch
// run as one replayed hipGraph: transforms in, body transform kernel, the
// persistent pair kernel over every conformation's work items, merge, tail
// strips, per-conformation finalisation and Debye sums, I(q) out.
#pragma once

#include "hist.hpp"
#include "rigid_kernels.hpp"

#include <map>
#include <memory>
#include <tuple>
#include <vector>

namespace ausaxs::gpu {

struct RigidInfo {
    int tile_j = 0, grid = 0;  // launch shape in use (autotuned when settings.tile_j == 0)
    float pair_ms = 0;         // pair kernel of the last batch (HIP events, outside graphs: tuning)
    float total_ms = 0;        // last scan, all batches (HIP events around the graph launches)
    int64_t pairs = 0;         // pair evaluations per scan (all conformations)
    int n_items = 0;           // work items of one batch
    int batch = 0;             // conformations per graph launch
    int tuned_candidates = 0;
    int copies = 1;            // LDS histogram copies of the pair kernel
};

class RigidScan {
  public:
    // xyz: n x 3 FP64 (bodies concatenated), w: n weights, sizes: per body.
    // waters (optional): m x 3 FP64 + weights, fixed in space.
    RigidScan(const double* xyz, const float* w, const int32_t* sizes, int n_bodies, const double* wxyz,
              const float* ww, int n_waters, const HistSettings& s);
    ~RigidScan();
    RigidScan(const RigidScan&) = delete;
    RigidScan& operator=(const RigidScan&) = delete;

    // T: [n_conf][n_bodies][12] (R row-major, t); I: [n_conf][nq]
    void scan(const double* T, int n_conf, const double* q, int nq, double cw, double* I);
    // The hydration-modified step (PartialHistogramManagerMT.cpp:83-105, after
    // the re-hydration of SimpleController.cpp:78): conformation c has its own
    // waters, counts[c] rows of wxyz (FP64 n x 3, fixed in space, in
    // conformation order) with weights ww, which replace the waters given at
    // creation; aw of every body and ww are recomputed for each conformation.
    void scan_hydrated(const double* T, int n_conf, const double* wxyz, const float* ww, const int32_t* counts,
                       const double* q, int nq, double cw, double* I);
    // composite histogram of conformation c of the last scan
    std::unique_ptr<CompositeDistanceHistogramHIP> histogram(int c);
    const RigidInfo& info() const { return info_; }
    int n_bodies() const { return (int)sizes_.size(); }

  private:
    struct Plan;
    struct Hydration;  // per-conformation waters of one batch
    HistSettings s_;
    float inv_w_;
    std::vector<int32_t> sizes_, body_off_;
    std::vector<double> xyz_;           // FP64 body coordinates
    std::vector<float> w_;              // weights
    std::vector<float> class_w_;        // weight of each class
    std::vector<int32_t> point_cls_;    // class of each body point
    PointSet waters_;                   // float waters (class ids)
    std::vector<float> water_w_;        // water weights
    bool weighted_ = false;             // more than kMaxWeightClasses distinct weights: weight pass
    double wscale_ = 1;                 // its fixed-point scale
    std::vector<int> water_classes_;    // classes present in the waters
    int C_ = 1;                         // classes
    int R_ = 0, R_aa_ = 0;              // rows per conformation
    std::vector<int32_t> row_kind_;
    std::vector<double> row_w_;
    // cached self terms: [6][base_nb] (aa value, count, sum d, ww value, count, sum d)
    std::vector<WeightedEntry> base_aa_, base_ww_;
    int base_nb_ = 0;
    DevBuf<double> d_xyz_, d_base_;
    DevBuf<int32_t> d_row_kind_;
    DevBuf<double> d_row_w_;
    std::vector<double> center_;        // per body: bounding-sphere centre (3) and radius
    double wcenter_[3] = {0, 0, 0}, wradius_ = -1;
    std::map<std::tuple<int, std::vector<int>, int>, std::unique_ptr<Plan>> plans_;
    std::unique_ptr<Plan> hplan_;       // plan of the last hydration-modified batch
    DevBuf<double> d_base_aa_;          // the cached self terms without the creation waters' ww
    Plan* last_ = nullptr;
    int last_conf0_ = 0;                // first conformation of the last batch
    int last_n_conf_ = 0;
    RigidInfo info_;

    int tri(int a, int b) const;  // unordered class-pair row
    Plan& plan_for(int n_conf, const std::vector<int>& moving, int nbins, const double* T);
    std::unique_ptr<Plan> build_plan(int n_conf, const std::vector<int>& moving, int nbins, int tile_j,
                                     const Hydration* hy = nullptr);
    int bound_bins(const double* T, int n_conf, const Hydration* hy = nullptr) const;
    std::vector<int> moving_bodies(const double* T, int n_conf) const;
    void record(Plan& p, int c0, int nc);
    void run_batch(Plan& p, const double* T, const double* q, int nq, double cw, double* I);
    void enqueue(Plan& p, int nq, double cw, bool copies);
    float time_pair(Plan& p, int grid, int copies);
};

}  // namespace ausaxs::gpu
