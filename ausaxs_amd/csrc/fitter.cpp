// Histogram fitter (see fitter.hpp).
#include "fitter.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace ausaxs::gpu {

// Natural cubic spline through the knots (x_k, y_k): S is a cubic on every
// interval, S, S' and S'' are continuous at the inner knots and S'' = 0 at
// both ends. With h_k = x_{k+1} - x_k, the chord slopes s_k = (y_{k+1} - y_k)
// / h_k and the knot curvatures M_k = S''(x_k), continuity of S' gives the
// tridiagonal system
//   h_{k-1} M_{k-1} + 2 (h_{k-1} + h_k) M_k + h_k M_{k+1} = 6 (s_k - s_{k-1})
// for the inner knots, M_0 = M_{n-1} = 0, solved here by forward elimination
// and back substitution (diagonally dominant, no pivoting). On interval k,
// S(x_k + t) = y_k + t (a1 + t (a2 + t a3)) with a1 = s_k - h_k (2 M_k +
// M_{k+1}) / 6, a2 = M_k / 2, a3 = (M_{k+1} - M_k) / (6 h_k).
CubicSpline::CubicSpline(const std::vector<double>& x, const std::vector<double>& y) : x_(x), y_(y) {
    const int n = (int)x.size();
    if (n != (int)y.size()) throw std::runtime_error("ausaxs_hip: spline x and y differ in size");
    if (n < 4) throw std::runtime_error("ausaxs_hip: a spline needs at least four points");
    const int m = n - 1;  // intervals
    std::vector<double> h(m), slope(m);
    for (int k = 0; k < m; ++k) {
        h[k] = x[k + 1] - x[k];
        if (!(h[k] > 0)) throw std::runtime_error("ausaxs_hip: spline knots must increase strictly");
        slope[k] = (y[k + 1] - y[k]) / h[k];
    }
    // inner unknowns M_1 .. M_{n-2}: diagonal, super-diagonal (after
    // elimination) and right-hand side
    std::vector<double> curv(n, 0.0), diag(n, 0.0), upper(n, 0.0), rhs(n, 0.0);
    for (int k = 1; k < m; ++k) {
        diag[k] = 2 * (h[k - 1] + h[k]);
        upper[k] = h[k];
        rhs[k] = 6 * (slope[k] - slope[k - 1]);
        if (k > 1) {
            // eliminate the sub-diagonal entry h_{k-1} with row k-1
            const double f = h[k - 1] / diag[k - 1];
            diag[k] -= f * upper[k - 1];
            rhs[k] -= f * rhs[k - 1];
        }
    }
    for (int k = m - 1; k >= 1; --k) curv[k] = (rhs[k] - (k + 1 < m ? upper[k] * curv[k + 1] : 0.0)) / diag[k];
    a1_.resize(m);
    a2_.resize(m);
    a3_.resize(m);
    for (int k = 0; k < m; ++k) {
        a1_[k] = slope[k] - h[k] * (2 * curv[k] + curv[k + 1]) / 6;
        a2_[k] = curv[k] / 2;
        a3_[k] = (curv[k + 1] - curv[k]) / (6 * h[k]);
    }
}

double CubicSpline::operator()(double z) const {
    // the interval whose left knot is the last one at or below z; below the
    // first knot the first interval and at or past the last knot the last
    // one, whose cubics continue outside the knots
    const int n = (int)x_.size();
    int k = (int)(std::upper_bound(x_.begin(), x_.end(), z) - x_.begin()) - 1;
    k = std::min(std::max(k, 0), n - 2);
    const double t = z - x_[k];
    return y_[k] + t * (a1_[k] + t * (a2_[k] + t * a3_[k]));
}

namespace {
struct ParamSpec {
    unsigned bit;
    double lo, hi, guess;
};
// SmartFitter::get_default_guess (SmartFitter.cpp:60-109) with the limits of
// ICompositeDistanceHistogram.h:70 and ICompositeDistanceHistogramExv.cpp:8-12
constexpr ParamSpec kSpecs[5] = {{kFitHydration, 0, 10, 1},
                                 {kFitExcludedVolume, 0.92, 1.08, 1},
                                 {kFitSolventDensity, 0.5, 2, 1},
                                 {kFitAtomicDW, 0, 5, 0},
                                 {kFitExvDW, 0, 5, 0}};
}  // namespace

HistogramFitter::HistogramFitter(CompositeDistanceHistogramHIP& h, std::vector<double> q, std::vector<double> I,
                                 std::vector<double> err, unsigned enabled, std::vector<double> q_model)
    : h_(&h), q_(std::move(q)), y_(std::move(I)), err_(std::move(err)), q_model_(std::move(q_model)),
      enabled_(enabled & 31u), base_(h.params) {
    init();
}

HistogramFitter::HistogramFitter(ScalableGridStructure& s, std::vector<double> q, std::vector<double> I,
                                 std::vector<double> err, unsigned enabled, std::vector<double> q_model)
    : scalable_(&s), owned_(s.calculate(1.0)), q_(std::move(q)), y_(std::move(I)), err_(std::move(err)),
      q_model_(std::move(q_model)), enabled_(enabled & 31u) {
    h_ = owned_.get();
    base_ = h_->params;
    init();
}

void HistogramFitter::init() {
    if (q_.size() != y_.size() || (!err_.empty() && err_.size() != q_.size()))
        throw std::runtime_error("ausaxs_hip: data q, I and errors differ in size");
    if (q_.size() < 3) throw std::runtime_error("ausaxs_hip: a fit needs at least three data points");
    if (q_model_.empty()) q_model_ = default_q_axis(1e-4, 0.5);
    if (q_model_.size() < 4) throw std::runtime_error("ausaxs_hip: the model q axis needs at least four points");
    // validate_model: a histogram without excluded-volume terms fits the
    // hydration factor only
    if (h_->mode == Mode::Simple) enabled_ &= kFitHydration;
    Context& ctx = Context::get();
    d_y_.upload(y_.data(), y_.size(), ctx.stream());
    if (!err_.empty()) d_err_.upload(err_.data(), err_.size(), ctx.stream());
    ctx.sync();
}

HistogramFitter::~HistogramFitter() { h_->params = base_; }

int HistogramFitter::n_params() const {
    int n = 0;
    for (const auto& s : kSpecs) n += (enabled_ & s.bit) != 0;
    return n;
}

void HistogramFitter::apply(const double* p) {
    FreeParams f = base_;
    int k = 0;
    if (enabled_ & kFitHydration) f.cw = p[k++];
    if (enabled_ & kFitExcludedVolume) {
        const double v = p[k++];
        if (!scalable_) {
            f.cx = v;
        } else if (v != owned_k_) {
            // the stretched dummies' histograms replace the current ones
            owned_ = scalable_->calculate(v);
            owned_k_ = v;
            h_ = owned_.get();
        }
    }
    if (enabled_ & kFitSolventDensity) f.crho = p[k++];
    if (enabled_ & kFitAtomicDW) f.dw_atomic = p[k++];
    if (enabled_ & kFitExvDW) f.dw_exv = p[k++];
    h_->params = f;
}

std::vector<double> HistogramFitter::model_curve(const double* params) {
    apply(params);
    // SmartFitter::splice (SmartFitter.cpp:246-252)
    const CubicSpline s(q_model_, h_->debye_transform(q_model_));
    std::vector<double> m(q_.size());
    for (size_t i = 0; i < q_.size(); ++i) m[i] = s(q_[i]);
    return m;
}

void HistogramFitter::evaluate(const double* params, int P, double* out) {
    if (P <= 0) return;
    const int k = n_params(), n = (int)q_.size();
    std::vector<double> models((size_t)P * n);
    for (int p = 0; p < P; ++p) {
        const std::vector<double> m = model_curve(params + (size_t)p * k);
        std::copy(m.begin(), m.end(), models.begin() + (size_t)p * n);
    }
    h_->params = base_;
    fevals_ += P;
    Context& ctx = Context::get();
    d_models_.upload(models.data(), models.size(), ctx.stream());
    d_out_.ensure((size_t)P * 5);
    check(launch_lsq(d_models_.p, d_y_.p, err_.empty() ? nullptr : d_err_.p, n, P, d_out_.p, ctx.stream()),
          "lsq_kernel");
    check(hipMemcpyAsync(out, d_out_.p, (size_t)P * 5 * sizeof(double), hipMemcpyDeviceToHost, ctx.stream()),
          "D2H");
    ctx.sync();
}

double HistogramFitter::chi2_one(double v) {
    double out[5];
    evaluate(&v, 1, out);
    return out[4];
}

HistogramFitter::Interval HistogramFitter::golden(Interval bounds) {
    // Golden.cpp:22-63: golden-section search to an interval below tol = 1e-4
    // (Minimizer.h:76)
    const double tol = 1e-4;
    const double phi = (1 + std::sqrt(5.0)) / 2, invphi = 1 / phi, invphi2 = invphi * invphi;
    double a = bounds.min, b = bounds.max;
    const double sum = a + b;
    a = std::min(a, b);
    b = sum - a;
    double diff = b - a;
    if (diff < tol) return {a, b};
    const unsigned n = (unsigned)std::ceil(std::log(tol / diff) / std::log(invphi));
    double c = a + invphi2 * diff, d = a + invphi * diff;
    double fc = chi2_one(c), fd = chi2_one(d);
    for (unsigned k = 0; k + 1 < n; ++k) {
        if (fc < fd) {
            b = d, d = c, fd = fc;
            diff = invphi * diff;
            c = a + invphi2 * diff;
            fc = chi2_one(c);
        } else {
            a = c, c = d, fc = fd;
            diff = invphi * diff;
            d = a + invphi * diff;
            fd = chi2_one(d);
        }
    }
    return fc < fd ? Interval{a, d} : Interval{c, b};
}

FitOutcome HistogramFitter::fit(FitAlgorithm alg, int scan_evals) {
    fevals_ = 0;
    FitOutcome r;
    const int k = n_params();
    if (k > 1) throw std::runtime_error("ausaxs_hip: the golden-section and scan minimizers fit one parameter");
    if (k == 1) {
        const ParamSpec* sp = nullptr;
        for (const auto& s : kSpecs)
            if (enabled_ & s.bit) sp = &s;
        Interval iv{sp->lo, sp->hi};
        // CompositeDistanceHistogramFFGridSurface::get_excluded_volume_scaling_factor_limits (:47-49)
        if (sp->bit == kFitExcludedVolume && h_->mode == Mode::FFGridSurface) iv = Interval{0, 2};
        if (alg == FitAlgorithm::Scan) {
            // Scan::landscape (Scan.cpp:26-36) as one batch, its minimum
            // (Dataset::find_minimum: first smallest), then Golden on one
            // step either side (Scan.cpp:48-59)
            if (scan_evals < 1) throw std::runtime_error("ausaxs_hip: scan needs at least one evaluation");
            std::vector<double> xs;
            const double span = iv.max - iv.min;
            for (double v = iv.min; v < iv.max; v += span / scan_evals) xs.push_back(v);
            std::vector<double> out(5 * xs.size());
            evaluate(xs.data(), (int)xs.size(), out.data());
            size_t imin = 0;
            for (size_t i = 1; i < xs.size(); ++i)
                if (out[5 * i + 4] < out[5 * imin + 4]) imin = i;
            const double width = (xs.back() - xs.front()) / (double)xs.size();
            iv = Interval{std::max(xs[imin] - width, iv.min), std::min(xs[imin] + width, iv.max)};
        } else if (alg != FitAlgorithm::Golden) {
            throw std::runtime_error("ausaxs_hip: unknown minimizer (0 = golden, 2 = scan)");
        }
        const Interval best = golden(iv);
        const double centre = best.min + (best.max - best.min) / 2;
        r.params = {centre};
        r.param_err = {(best.max - best.min) / 2};
        (void)chi2_one(centre);  // Golden::minimize_override evaluates the centre
    }
    // the inner linear fit at the optimum (SmartFitter::fit, :148-170)
    r.fevals = fevals_;
    double out[5];
    evaluate(r.params.data(), 1, out);
    r.model = model_curve(r.params.data());
    h_->params = base_;
    r.a = out[0], r.b = out[1], r.a_err = std::sqrt(out[2]), r.b_err = std::sqrt(out[3]), r.chi2 = out[4];
    r.dof = (int)q_.size() - 2 - k;
    r.I_fit.resize(q_.size());
    r.residuals.resize(q_.size());
    for (size_t i = 0; i < q_.size(); ++i) {
        r.I_fit[i] = r.a * r.model[i] + r.b;
        r.residuals[i] = (y_[i] - r.I_fit[i]) / (err_.empty() ? 1.0 : err_[i]);
    }
    return r;
}

}  // namespace ausaxs::gpu
