// Fit of a composite histogram to measured I(q) (SmartFitter, SmartFitter.cpp:1-252).
//
// Each chi2 evaluation applies the free parameters to the histogram
// (EnabledFitParameters::apply_pars, SmartFitter.cpp:19-30), transforms it on
// the model q axis with the sinqd rows kept from the first transform (the
// GEMM runs once per histogram), splines the curve onto the data q values
// (math::CubicSpline) and solves the inner linear fit y ~ a*model + b
// (LinearLeastSquares.cpp:26-44) on the GPU, for a whole batch of parameter
// vectors at once. The 1-D minimizers of the reference (mini::Golden,
// mini::Scan) run on top of the batched evaluation: Scan's landscape is one
// batch.
#pragma once

#include "hist.hpp"

#include <memory>
#include <vector>

namespace ausaxs::gpu {

// The natural cubic spline of math::CubicSpline (CubicSpline.cpp): zero
// curvature at both end knots, evaluated on the interval of the last knot at
// or below the argument (the end intervals' cubics continue outside).
class CubicSpline {
  public:
    CubicSpline(const std::vector<double>& x, const std::vector<double>& y);
    double operator()(double z) const;

  private:
    std::vector<double> x_, y_;
    std::vector<double> a1_, a2_, a3_;  // per interval: S = y_k + a1 t + a2 t^2 + a3 t^3
};

// Free parameters in the reference's order (SmartFitter.cpp:19-30); bit k of
// an enabled mask turns parameter k on.
enum FitParamBit : unsigned {
    kFitHydration = 1,       // apply_water_scaling_factor, limits [0, 10], guess 1
    kFitExcludedVolume = 2,  // apply_excluded_volume_scaling_factor, [0.92, 1.08], guess 1
    kFitSolventDensity = 4,  // apply_solvent_density_scaling_factor, [0.5, 2], guess 1
    kFitAtomicDW = 8,        // apply_atomic_debye_waller_factor, [0, 5], guess 0
    kFitExvDW = 16,          // apply_exv_debye_waller_factor, [0, 5], guess 0
};

enum class FitAlgorithm : int { Golden = 0, Scan = 2 };  // mini::algorithm values

struct FitOutcome {
    std::vector<double> params, param_err;  // fitted free parameters (centre, half interval)
    double a = 0, b = 0, a_err = 0, b_err = 0, chi2 = 0;
    int dof = 0, fevals = 0;
    std::vector<double> model;      // spliced model curve at the data q values
    std::vector<double> I_fit;      // a*model + b
    std::vector<double> residuals;  // (y - I_fit) / err
};

class HistogramFitter {
  public:
    // h must outlive the fitter. q_model empty: the default axis on
    // [1e-4, 0.5] (DistanceHistogram::get_q_axis). err empty: unit errors.
    HistogramFitter(CompositeDistanceHistogramHIP& h, std::vector<double> q, std::vector<double> I,
                    std::vector<double> err, unsigned enabled, std::vector<double> q_model = {});
    // A grid structure with stretchable dummies (must outlive the fitter):
    // the excluded-volume parameter recomputes the histograms with the
    // dummies stretched by its value (CompositeDistanceHistogramFFGridScalableExv::
    // apply_excluded_volume_scaling_factor) instead of scaling the exv terms.
    HistogramFitter(ScalableGridStructure& s, std::vector<double> q, std::vector<double> I,
                    std::vector<double> err, unsigned enabled, std::vector<double> q_model = {});
    ~HistogramFitter();

    // enabled parameters after validate_model (SmartFitter.cpp:32-49):
    // FF-free (simple) histograms drop the excluded-volume parameters
    unsigned enabled() const { return enabled_; }
    int n_params() const;
    // chi2 of P parameter vectors (params: P x n_params); out[5*p..] = a, b,
    // var a, var b, chi2 (one lsq batch on the GPU)
    void evaluate(const double* params, int P, double* out);
    // the spliced model curve of one parameter vector
    std::vector<double> model_curve(const double* params);
    FitOutcome fit(FitAlgorithm alg, int scan_evals = 100);

  private:
    CompositeDistanceHistogramHIP* h_;
    ScalableGridStructure* scalable_ = nullptr;
    std::unique_ptr<CompositeDistanceHistogramHIP> owned_;
    double owned_k_ = 1;
    std::vector<double> q_, y_, err_, q_model_;
    unsigned enabled_;
    FreeParams base_;
    int fevals_ = 0;
    DevBuf<double> d_y_, d_err_, d_models_, d_out_;

    void init();
    void apply(const double* p);
    double chi2_one(double v);  // 1-D objective (mini::Minimizer function)
    struct Interval {
        double min, max;
    };
    Interval golden(Interval bounds);  // Golden::search (Golden.cpp:22-63)
};

}  // namespace ausaxs::gpu
