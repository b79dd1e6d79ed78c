// Histograms of bodies with symmetry copies (SymmetryManagerMT,
// SymmetryManagerMT.cpp:39-170).
//
// Every body holds its atoms once plus, per symmetry, `reps` copies made by
// affine maps (ISymmetry::AffineTransform, ISymmetry.h:17-25) applied in FP64
// to the float coordinates (SymmetryHelpers.cpp:16-43). Copy pairs with
// identical geometry are evaluated once and counted `scale` times (the pair
// schedule, ISymmetry.h:26-39). On the GPU every body copy is a resident point
// set and every enqueue of the reference manager one pair job of a single
// PairEngine launch; a job's multiplicity selects the block of histogram rows
// it adds into, and the blocks are combined with their integer factors on the
// host, so counts stay exact.
#pragma once

#include "hist.hpp"

#include <array>
#include <memory>
#include <vector>

namespace ausaxs::gpu {

// One copy-pair class of a schedule: repetitions a and b (0 = the original
// body) stand for `scale` geometrically identical pairs (SymmetricDuplicatePair).
struct CopyPair {
    int a = 0, b = 0, scale = 1;
};

struct SymmetrySpec {
    int reps = 0;                     // ISymmetry::repetitions()
    std::vector<double> transforms;   // reps x 12: R (row-major 3x3), t; copy k = transforms[k-1]
    std::vector<CopyPair> schedule;   // ISymmetry::internal_pair_schedule()
};

// ISymmetry::internal_pair_schedule (ISymmetry.cpp:43-66): the cyclic-chain
// default, pairs (0, k) with their multiplicities.
std::vector<CopyPair> chain_pair_schedule(int reps, bool closed);
// compute_pair_schedule (PairScheduleUtil.cpp:33-60): placements (12 doubles
// each, placement 0 = the original) grouped by their relative transform,
// rounded to 1e-4, up to inversion.
std::vector<CopyPair> placement_pair_schedule(const std::vector<double>& placements);

// CyclicSymmetry (CyclicSymmetry.cpp:85-118): copies k = 1..reps of a body
// with centre of mass cm, offset t_i, per-step translation t_r and rotation
// by `angle` about `axis`; 12 doubles per copy. closed (optional) receives
// is_closed().
std::vector<double> cyclic_transforms(const double* cm, const double* offset, const double* step, const double* axis,
                                      double angle, int reps, bool* closed);

// Histograms of bodies with symmetry copies, on resident point sets.
//
// Full mode restates SymmetryManagerMT::calculate (SymmetryManagerMT.cpp:39-170).
// Partial mode restates PartialSymmetryManagerMT (PartialSymmetryManagerMT.cpp:78-450):
// the same pairs with its own row/column roles (calc_aa, calc_aw:
// rows over the lower body, over the lower symmetry, over the waters), and the
// self term of a body kept from the coordinates and copy count of its last
// internal modification (calc_aa_self runs only then). Changing bodies,
// symmetries or waters marks the structure for a rebuild; calculate() then
// evaluates every block on the current coordinates, which gives what the
// reference's recomputed and cached blocks hold.
class SymmetryStructure {
  public:
    // atoms: sum(sizes) x 4 floats {x, y, z, w}, body after body; syms[b]: the
    // symmetries of body b; waters: n_waters x 4 floats.
    SymmetryStructure(const float* atoms, const std::vector<int>& sizes, std::vector<std::vector<SymmetrySpec>> syms,
                      const float* waters, int n_waters, const HistSettings& s, bool partial = false);
    std::unique_ptr<CompositeDistanceHistogramHIP> calculate(bool timed = false);
    PairEngine& engine();
    bool partial() const { return partial_; }

    // state changes (partial mode: IPartialHistogramManager's state manager).
    // internal: the body's own distances changed (weights or a non-rigid
    // move): its self term is recomputed too.
    void set_body(int b, const float* pts, bool internal);
    void set_symmetry(int b, int s, SymmetrySpec spec);
    void set_waters(const float* pts, int n);

  private:
    HistSettings s_;
    bool partial_;
    std::vector<std::vector<float>> bodies_, self_bodies_;  // current and self-term coordinates (x 4)
    std::vector<int> self_scale_;                           // 1 + copies when the self term was taken
    std::vector<std::vector<SymmetrySpec>> syms_;
    std::vector<float> waters_;
    bool dirty_ = true;

    std::vector<double> body_self_w2_;  // per body: scale * sum w^2 (SimpleCalculator.h:144-160)
    double water_self_w2_ = 0;
    std::vector<float> atom_class_w_, water_class_w_;
    // global row -> (kind 0 aa / 1 aw / 2 ww, class a, class b, factor)
    std::vector<std::array<int, 4>> rows_;
    std::unique_ptr<PairEngine> engine_;
    std::vector<uint64_t> cnt_;
    std::vector<int64_t> dsum_;

    void rebuild();
    static void check_spec(const SymmetrySpec& sp);
};

}  // namespace ausaxs::gpu
