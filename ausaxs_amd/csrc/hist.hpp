// Histogram managers and composite histograms of the MI355X Debye path.
//
// Mirrors the reference's seam (IHistogramManager::calculate_all returning an
// ICompositeDistanceHistogram, IHistogramManager.h:16-29) for the managers
//   HistogramManagerMTFFAvg       (HistogramManagerMTFFAvg.cpp:194-371)
//   HistogramManagerMTFFExplicit  (HistogramManagerMTFFExplicit.cpp:216-436)
//   HistogramManagerMT            (HistogramManagerMT.cpp:23-75)
// with the pair work on the GPU (engine.hpp) and the O(rows x bins)
// bookkeeping (self terms, p_tot, max_bin, exv rows) on the host in FP64.
#pragma once

#include "engine.hpp"

#include <array>
#include <memory>
#include <vector>

namespace ausaxs::gpu {

constexpr int kMaxFF = 15;  // settings/FormFactorSettings.h:12
constexpr int kExvBin = 0;  // form_factor/FormFactorType.h:35
constexpr int kWaterBin = 1;
// XYZW sets with at most this many distinct weights keep one histogram row
// per weight-class pair; more take the weighted engine (engine.hpp)
constexpr size_t kMaxWeightClasses = 64;

struct WeightedEntry {  // distribution/detail/WeightedEntry.h:16-50
    double value = 0;
    uint32_t count = 0;
    double bin_center = 0;
};

void entry_add(WeightedEntry& a, const WeightedEntry& b);

// A [n1][n2][width] table of entries stored by rows, holding only the rows
// that were touched (form-factor pairs present in a structure); absent rows
// read as zero. Row storage never moves once created.
class EntryRows {
  public:
    void reset(int n1, int n2, int width);
    int n1() const { return n1_; }
    int n2() const { return n2_; }
    int width() const { return width_; }
    size_t size() const { return (size_t)n1_ * n2_ * width_; }
    const WeightedEntry* row(int a, int b) const;  // nullptr if absent
    WeightedEntry* row(int a, int b);
    WeightedEntry* touch(int a, int b);  // the row, zero-initialised on first use
    void clear(int a, int b);            // make a row absent (zero)
    void truncate(int width);            // keep the first `width` bins
    void dense(WeightedEntry* out) const;  // [n1][n2][width] with zero rows
    template <class F>
    void for_each(F&& f) const {  // f(a, b, row) over present rows in (a, b) order
        for (int a = 0; a < n1_; ++a)
            for (int b = 0; b < n2_; ++b)
                if (const WeightedEntry* r = row(a, b)) f(a, b, r);
    }

  private:
    int n1_ = 0, n2_ = 0, width_ = 0;
    std::vector<int32_t> slot_;
    std::vector<std::vector<WeightedEntry>> rows_;
};
// HistogramManagerMTFFAvg.cpp:340-352: last bin >= 10 with a non-zero value, + 1 (at least 10)
unsigned max_bin_of(const std::vector<WeightedEntry>& p);
// WeightedDistribution1D::get_weighted_axis (WeightedDistribution1D.cpp:51-58) with
// d[0] = 0 (DistanceHistogram.cpp:44-47); plain bin positions if !weighted
std::vector<double> axis_of(const std::vector<WeightedEntry>& p, double bin_width, bool weighted);

struct HistSettings {
    double bin_width = 0.25;    // settings::axes::bin_width (d_axis width)
    uint32_t bin_count = 8000;  // settings::axes::bin_count
    bool weighted_bins = true;  // settings::hist::weighted_bins
    int debye_terms = 3;        // bf16 terms per operand; 0 = FP64 contraction
    int tile_j = kDefaultTileJ; // columns per work item (<= 8064)
    int lattice = 1;            // grid dummy-dummy pairs by lattice correlation: 0 off, 1 auto, 2 always
    double surface_thickness = 1.0;  // grid-surface excluded volume: surface layer thickness (A)
};
// WidthControllers.h:10-27: inverse width as float
float inv_width(const HistSettings& s);

struct FreeParams {  // CompositeDistanceHistogramFFAvgBase free parameters
    double cw = 1, cx = 1, crho = 1, dw_atomic = 0, dw_exv = 0;
};

enum class Mode : int { FFAvg = 0, FFExplicit = 1, Simple = 2, FFGrid = 3, FFGridSurface = 4 };

// form factors restated from FormFactor.h:33-39 / FormFactorTable.h:28-120 /
// ExvFormFactor.h:16-42 (default MinimumFluctuation_implicit_H volumes)
double ff_atomic(int type, double q);
double ff_exv(int type, double q);

// The reference's default q axis and its sub-axis between qmin and qmax
// (ConstantsAxes.h:14, Axis.cpp:40-59).
std::vector<double> default_q_axis(double qmin, double qmax);

class CompositeDistanceHistogramHIP {
  public:
    Mode mode = Mode::FFExplicit;
    unsigned max_bin = 0;
    double bin_width = 0.25;
    bool weighted_bins = true;
    int debye_terms = 3;
    // FF modes: aa [15][15][max_bin], aw [15][1][max_bin]; Simple: [1][1][max_bin]
    EntryRows aa, aw;
    std::vector<WeightedEntry> ww, p_tot;  // [max_bin]
    std::vector<double> d_axis;
    // FFGrid only: distance axes of the atom-dummy and dummy-dummy rows
    // (CompositeDistanceHistogramFFGrid.cpp:130-136), first max_bin entries
    std::vector<double> d_axis_ax, d_axis_xx;
    std::vector<WeightedEntry> p_tot_ax, p_xx;
    // FFGridSurface only: the dummy profiles of interior and surface dummies
    // (CompositeDistanceHistogramFFGridSurface.cpp:33-46), first max_bin entries;
    // the surface terms scale with the excluded-volume parameter
    EntryRows ax_interior, ax_surface;  // [15][1][max_bin]
    std::vector<WeightedEntry> xx_interior, xx_surface, xx_cross, wx_interior, wx_surface;
    double grid_volume = 0;
    // FFGrid: volume V of the dummies' ExvFormFactor (width^3; (width k)^3
    // for dummies stretched by k, CompositeDistanceHistogramFFGridScalableExv.cpp:9-19)
    double grid_exv_volume = 1;
    FreeParams params;

    // ICompositeDistanceHistogram(Exv)::debye_transform on an explicit q grid;
    // profiles (optional, [6][nq]) = {aa, ax, aw, xx, wx, ww}.
    std::vector<double> debye_transform(const std::vector<double>& q, std::vector<double>* profiles = nullptr);
    // sinqd rows on the q grid, computed on the GPU (cached per q grid)
    const std::vector<double>& sinqd(const std::vector<double>& q);

    double last_debye_ms = 0;

  private:
    std::vector<double> cached_q_, cached_sinqd_;
    // (kind, f1, f2): kind 0 aa, 1 aw, 2 ww; axis_of_row_: 0 d_axis, 1 d_axis_ax, 2 d_axis_xx
    std::vector<std::array<int, 3>> row_keys_;
    std::vector<int> axis_of_row_;
};

// Resident structure: coordinates uploaded and work lists built once;
// calculate() is HistogramManager*::calculate_all on the resident data.
class ResidentStructure {
  public:
    // atoms: n x 4 floats {x,y,z, ff (int bits) | weight}; waters likewise.
    // FFGrid: exv_xyz holds the n_exv dummy lattice points (grid.hpp) and
    // grid_volume the grid volume used for the FFAvg base (Z_exv).
    // FFGridSurface: exv_xyz the interior dummies, surf_xyz the surface ones
    ResidentStructure(Mode mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                      const HistSettings& s, double z_exv, const double* exv_xyz = nullptr, int n_exv = 0,
                      double grid_volume = 0, const double* surf_xyz = nullptr, int n_surf = 0);
    std::unique_ptr<CompositeDistanceHistogramHIP> calculate(bool timed = false);
    // the structure's own engine, built (and its points uploaded) on first use
    PairEngine& engine();
    bool has_engine() const { return engine_ != nullptr; }
    // the engine's inputs (before or after the engine was built)
    SpecView spec_view() const { return engine_ ? engine_->view() : view_of(spec_); }
    Mode mode() const { return mode_; }
    void set_grid_exv_volume(double v) { grid_exv_volume_ = v; }
    double z_exv() const { return z_exv_; }
    bool weighted() const { return weighted_; }
    // pairs of all jobs (the histograms' pair count)
    int64_t pairs() const { return pairs_; }
    // Several structures' pair work in one engine (one pair launch, one
    // merge): each result equals its calculate() alone (integer histograms).
    // The structures must share bin width, tile width and lattice mode and
    // have no per-point weights. *pairs / *kernel_pairs: the merged engine's.
    static std::vector<std::unique_ptr<CompositeDistanceHistogramHIP>> calculate_merged(
        const std::vector<ResidentStructure*>& members, int64_t* pairs = nullptr, int64_t* kernel_pairs = nullptr);

  private:
    Mode mode_;
    HistSettings s_;
    double z_exv_;
    int n_atoms_, n_waters_, n_exv_ = 0, n_surf_ = 0;
    double grid_volume_ = 0, grid_exv_volume_ = 1;
    std::array<uint32_t, kMaxFF> ff_count_{};     // FF modes: atoms per form-factor type
    double self_w2_atoms_ = 0, self_w2_waters_ = 0;  // Simple mode: sum of w*w (float products)
    std::vector<float> atom_class_w_, water_class_w_;
    std::vector<std::array<int, 3>> rows_;        // global row -> (kind, a, b)
    EngineSpec spec_;                             // what the engine is built from (moved into it)
    int nbins_ = 10, n_rows_ = 0;
    float inv_w_ = 1;
    int64_t pairs_ = 0;
    std::unique_ptr<PairEngine> engine_;
    std::vector<uint64_t> cnt_;
    std::vector<int64_t> dsum_;
    bool weighted_ = false;       // Simple mode with more than kMaxWeightClasses distinct weights
    WeightSums wsum_;             // weight-pass sums (weighted_)
    // the composite histogram from cnt_ / dsum_ ([rows][nbins_ + 1])
    std::unique_ptr<CompositeDistanceHistogramHIP> finalize();
};

// HistogramManagerMTFFGrid structure from FP64 atoms (n x 3, form-factor
// types 1..14) and waters (m x 3): grid dummies (grid.hpp) and float casts
// of the coordinates as CompactCoordinatesFF holds them.
std::unique_ptr<ResidentStructure> make_grid_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                       const double* waters_xyz, int m, const HistSettings& s);
// HistogramManagerMTFFAvg from FP64 atoms and waters with the reference's own
// excluded-volume charge Z_exv = V_grid * 0.334 / N (HistogramManagerMTFFAvg.cpp:
// 355-359; V_grid of the molecule's grid, Molecule.cpp:147-150)
constexpr double kWaterChargeDensity = 0.334;  // constants::charge::density::water, e/A^3
std::unique_ptr<ResidentStructure> make_ffavg_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                        const double* waters_xyz, int m, const HistSettings& s);
// HistogramManagerMTFFGridSurface (HistogramManagerMTFFGridSurface.cpp:16-274):
// the dummies split into interior and surface ones (RawGridWithSurfaceExv)
std::unique_ptr<ResidentStructure> make_grid_surface_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                               const double* waters_xyz, int m,
                                                               const HistSettings& s);
// HistogramManagerMTFFGridScalableExv (HistogramManagerMTFFGridScalableExv.cpp:16-244):
// the FFGrid histograms with the dummies stretched by a factor k about the
// origin (float coordinates times k, as CompactCoordinatesFF stores them) and
// the dummy form factor of volume k^3. The grid and its dummies are built
// once; calculate(k) recomputes the pair histograms on the GPU.
class ScalableGridStructure {
  public:
    ScalableGridStructure(const double* atoms_xyz, const int32_t* ff, int n, const double* waters_xyz, int m,
                          const HistSettings& s);
    std::unique_ptr<CompositeDistanceHistogramHIP> calculate(double k);
    int n_dummies() const { return (int)(dummies_.size() / 3); }
    int64_t pairs() const { return st_ ? st_->engine().pairs() : 0; }

  private:
    std::vector<float> atoms_, waters_;
    std::vector<double> dummies_;
    double grid_volume_ = 0;
    int n_ = 0, m_ = 0;
    HistSettings s_;
    double k_ = 0;
    std::unique_ptr<ResidentStructure> st_;
};

// The FF-manager points of FP64 atoms/waters: {float x, y, z, type bits},
// waters typed WATER.
void pack_ff_points(const double* xyz, const int32_t* ff, int n, int32_t fixed_type, std::vector<float>& out);

// single pair set with the reference's loop geometry (test entry)
void raw_pair_histogram(const float* a, int n, const float* b, int m, const HistSettings& s, uint64_t* count,
                        double* dsum);

}  // namespace ausaxs::gpu
