// Histogram managers and composite histograms (see hist.hpp).
#include "hist.hpp"

#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <set>

namespace ausaxs::gpu {

float inv_width(const HistSettings& s) {
    // ConstantWidth / VariableWidth (WidthControllers.h:10-27): 1./width in
    // double, returned as float.
    return (float)(1.0 / s.bin_width);
}

// ---------------------------------------------------------------------------
// form factors
// ---------------------------------------------------------------------------
namespace {

struct Gauss5 {
    double a[5], b[5], c;
};
constexpr double kPi = 3.14159265358979323846;
constexpr double kS2Q = 1. / (4 * 4 * kPi * kPi);  // FormFactorTable.h:15-23

// FormFactorTable.h:28-120 (b in s units; converted on evaluation)
const Gauss5 kH{{0.489918, 0.262003, 0.196767, 0.049879, 0}, {20.6593, 7.74039, 49.5519, 2.20159, 0}, 0.001305};
const Gauss5 kO{{2.960427, 2.508818, 0.637853, 0.722838, 1.142756},
                {14.182259, 5.936858, 0.112726, 34.958481, 0.390240},
                0.027014};
const Gauss5 kN{{11.893780, 3.277479, 1.858092, 0.858927, 0.912985},
                {0.000158, 10.232723, 30.344690, 0.656065, 0.217287},
                -11.804902};
const Gauss5 kS{{6.362157, 5.154568, 1.473732, 1.635073, 1.209372},
                {1.514347, 22.092528, 0.061373, 55.445176, 0.646925},
                0.154722};
const Gauss5 kC{{2.657506, 1.078079, 1.490909, -4.241070, 0.713791},
                {14.780758, 0.776775, 42.086843, -0.000294, 0.239535},
                4.297983};
const Gauss5 kAr{{7.188004, 6.638454, 0.454180, 1.929593, 1.523654},
                 {0.956221, 15.339877, 15.339862, 39.043824, 0.062409},
                 0.265954};
const Gauss5 kCH{{2.909530, 0.485267, 1.516151, 0.206905, 1.541626},
                 {13.933084, 23.221524, 41.990403, 4.974183, 0.679266},
                 0.337670};
const Gauss5 kCH2{{3.275723, 0.870037, 1.534606, 0.395078, 1.544562},
                  {13.408502, 23.785175, 41.922444, 5.019072, 0.724439},
                  0.377096};
const Gauss5 kCH3{{3.681341, 1.228691, 1.549320, 0.574033, 1.554377},
                  {13.026207, 24.131974, 41.869426, 4.984373, 0.765769},
                  0.409294};
const Gauss5 kOH{{0.456221, 3.219608, 0.812773, 2.666928, 1.380927},
                 {21.503498, 13.397134, 34.547137, 5.826620, 0.412902},
                 0.463202};
const Gauss5 kNH{{1.650531, 0.429639, 2.144736, 1.851894, 1.408921},
                 {10.603730, 6.987283, 29.939901, 10.573859, 0.611678},
                 0.510589};
const Gauss5 kNH2{{1.904157, 1.942536, 2.435585, 0.730512, 1.379728},
                  {10.803702, 10.792421, 29.610479, 6.847755, 0.709687},
                  0.603738};
const Gauss5 kNH3{{1.882162, 1.933200, 2.465843, 0.927311, 1.190889},
                  {10.975157, 10.956008, 29.208572, 6.663555, 0.843650},
                  0.597322};
const Gauss5 kSH{{0.570042, 6.337416, 1.641643, 5.398549, 1.527982},
                 {11.447986, 1.197657, 55.401032, 22.420955, 2.356552},
                 1.523944};

double gauss5(const Gauss5& g, double q) {
    double s = 0;
    for (int i = 0; i < 5; ++i) s += g.a[i] * std::exp(-(g.b[i] * kS2Q) * q * q);
    return s + g.c;
}

const Gauss5* atomic_table(int type) {
    // lookup::atomic::raw::get (FormFactor.h:93-117)
    switch (type) {
        case 1: return &kOH;  // WATER == OH -> OH_alc
        case 2: return &kH;
        case 3: return &kC;
        case 4: return &kCH;   // CH_sp3
        case 5: return &kCH2;  // CH2_sp3
        case 6: return &kCH3;  // CH3_sp3
        case 7: return &kN;
        case 8: return &kNH;
        case 9: return &kNH2;
        case 10: return &kNH3;  // NH3_plus
        case 11: return &kO;
        case 12: return &kS;
        case 13: return &kSH;
        case 14: return &kAr;  // OTHER
        default: return nullptr;
    }
}

double exv_volume(int type) {
    // constants::exv::MinimumFluctuation_implicit_H; OTHER -> volume(vdw Ar = 1.88)
    static const double v[kMaxFF] = {0,      20.911, 0,      12.352, 11.640, 34.583, 41.851, 0.027,
                                     2.181, 20.562, 20.722, 14.238, 15.413, 28.529, 0};
    if (type == 14) return 4 * kPi / 3 * 1.88 * 1.88 * 1.88;
    if (type < 1 || type >= kMaxFF) throw std::runtime_error("ausaxs_hip: no excluded-volume entry for this type");
    return v[type];
}

}  // namespace

double ff_atomic(int type, double q) {
    if (type == kExvBin) {
        const double r = 1.62;  // constants::radius::average_atomic_radius
        return std::exp(-(r * r / 2) * q * q);
    }
    const Gauss5* g = atomic_table(type);
    if (!g) throw std::runtime_error("ausaxs_hip: invalid form factor type " + std::to_string(type));
    return gauss5(*g, q);
}

double ff_exv(int type, double q) {
    const double v = exv_volume(type);
    const double exponent = std::pow(v, 2. / 3) / (4 * kPi);
    return v * 0.334 * std::exp(-exponent * q * q);
}

std::vector<double> default_q_axis(double qmin, double qmax) {
    const double amin = 1e-4, amax = 1, bins = 200, width = (amax - amin) / bins;
    auto get_bin = [&](double v) -> unsigned {
        if (v <= amin) return 0;
        if (v >= amax) return (unsigned)bins;
        return (unsigned)std::floor((v + 1e-6 - amin) / width);
    };
    const unsigned b0 = get_bin(qmin), b1 = get_bin(qmax);
    std::vector<double> q;
    for (unsigned i = b0; i < b1; ++i) q.push_back(amin + i * width);
    return q;
}

// ---------------------------------------------------------------------------
// weighted entries
// ---------------------------------------------------------------------------
namespace {
inline void add(WeightedEntry& a, const WeightedEntry& b) { entry_add(a, b); }
}  // namespace

void entry_add(WeightedEntry& a, const WeightedEntry& b) {
    a.value += b.value;
    a.count += b.count;
    a.bin_center += b.bin_center;
}

void EntryRows::reset(int n1, int n2, int width) {
    n1_ = n1, n2_ = n2, width_ = width;
    slot_.assign((size_t)n1 * n2, -1);
    rows_.clear();
}

const WeightedEntry* EntryRows::row(int a, int b) const {
    const int32_t k = slot_[(size_t)a * n2_ + b];
    return k < 0 ? nullptr : rows_[k].data();
}

WeightedEntry* EntryRows::row(int a, int b) {
    const int32_t k = slot_[(size_t)a * n2_ + b];
    return k < 0 ? nullptr : rows_[k].data();
}

WeightedEntry* EntryRows::touch(int a, int b) {
    int32_t& k = slot_[(size_t)a * n2_ + b];
    if (k < 0) {
        k = (int32_t)rows_.size();
        rows_.emplace_back((size_t)width_);
    }
    return rows_[k].data();
}

void EntryRows::clear(int a, int b) {
    int32_t& k = slot_[(size_t)a * n2_ + b];
    if (k >= 0) std::fill(rows_[k].begin(), rows_[k].end(), WeightedEntry{});
    k = -1;
}

void EntryRows::truncate(int width) {
    width_ = width;
    for (auto& r : rows_) r.resize((size_t)width);
}

void EntryRows::dense(WeightedEntry* out) const {
    for (int a = 0; a < n1_; ++a)
        for (int b = 0; b < n2_; ++b) {
            WeightedEntry* dst = out + ((size_t)a * n2_ + b) * width_;
            if (const WeightedEntry* r = row(a, b)) std::copy_n(r, width_, dst);
            else std::fill_n(dst, width_, WeightedEntry{});
        }
}

namespace {

inline void scale(WeightedEntry& e, double f) {  // WeightedEntry.h operator*
    e.value *= f;
    e.count = (uint32_t)(f * e.count);
    e.bin_center *= f;
}

}  // namespace

unsigned max_bin_of(const std::vector<WeightedEntry>& p) {
    for (int i = (int)p.size() - 1; i >= 10; --i)
        if (p[i].value != 0) return (unsigned)i + 1;
    return 10;
}

std::vector<double> axis_of(const std::vector<WeightedEntry>& p, double bin_width, bool weighted) {
    const int n = (int)p.size();
    std::vector<double> d(n);
    const double width = (n * bin_width) / n;
    for (int i = 0; i < n; ++i) {
        const double dv = i * width;
        if (weighted) {
            const double bc = p[i].bin_center;
            d[i] = ((double)(!bc) * dv + bc) / (double)((!p[i].count) + p[i].count);
        } else {
            d[i] = dv;
        }
    }
    if (n) d[0] = 0;
    return d;
}

namespace {

std::vector<float4> to_points(const float* xyzw, int n) {
    std::vector<float4> v(n);
    for (int i = 0; i < n; ++i) v[i] = make_float4(xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2], 0.f);
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// ResidentStructure
// ---------------------------------------------------------------------------
ResidentStructure::ResidentStructure(Mode mode, const float* atoms, int n_atoms, const float* waters, int n_waters,
                                     const HistSettings& s, double z_exv, const double* exv_xyz, int n_exv,
                                     double grid_volume, const double* surf_xyz, int n_surf)
    : mode_(mode), s_(s), z_exv_(z_exv), n_atoms_(n_atoms), n_waters_(n_waters), n_exv_(n_exv), n_surf_(n_surf),
      grid_volume_(grid_volume) {
    if (n_atoms < 0 || n_waters < 0 || n_exv < 0 || n_surf < 0)
        throw std::runtime_error("ausaxs_hip: negative point count");
    if (!(s.bin_width > 0) || s.bin_count < 10) throw std::runtime_error("ausaxs_hip: invalid bin settings");
    const bool grid = mode == Mode::FFGrid || mode == Mode::FFGridSurface;
    if (grid && n_exv > 0 && !exv_xyz) throw std::runtime_error("ausaxs_hip: FFGrid needs dummies");
    if (mode == Mode::FFGridSurface && n_surf > 0 && !surf_xyz)
        throw std::runtime_error("ausaxs_hip: FFGridSurface needs surface dummies");
    const float inv_w = inv_width(s);
    PointSet A, W, X, XS;
    A.orig = to_points(atoms, n_atoms);
    W.orig = to_points(waters, n_waters);
    std::vector<PairJob> jobs;
    std::vector<std::vector<float>> weights;  // weighted engine only
    int n_rows = 0;

    if (mode != Mode::Simple) {
        std::set<int> present;
        for (int i = 0; i < n_atoms; ++i) {
            int32_t f;
            std::memcpy(&f, &atoms[4 * i + 3], 4);
            // CompactCoordinatesFF::setup (CompactCoordinatesFF.h:21-32) throws on UNKNOWN
            if (f < 1 || f >= kMaxFF)
                throw std::runtime_error("ausaxs_hip: atom " + std::to_string(i) + " has form factor type " +
                                         std::to_string(f) + " outside [1, 14]");
            ++ff_count_[f];
            A.orig[i].w = float_of(f);
            present.insert(f);
        }
        A.n_classes = kMaxFF;
        W.n_classes = 1;
        PairJob aa, aw, ww;
        aa.row_set = 0;
        aa.rowmap.assign(kMaxFF * kMaxFF, -1);
        aw.row_set = 0;
        aw.col_set = 1;
        aw.rowmap.assign(kMaxFF, -1);
        ww.row_set = 1;
        ww.rowmap.assign(1, -1);
        for (int f1 : present)
            for (int f2 : present) {
                aa.rowmap[f1 * kMaxFF + f2] = n_rows++;
                rows_.push_back({0, f1, f2});
            }
        for (int f : present) {
            aw.rowmap[f] = n_rows++;
            rows_.push_back({1, f, 0});
        }
        ww.rowmap[0] = n_rows++;
        rows_.push_back({2, 0, 0});
        jobs = {aa, aw, ww};
        if (mode == Mode::FFGridSurface) {
            // interior (set 2) and surface (set 3) dummies (HistogramManagerMTFFGridSurface.cpp:30-47):
            // xx interior, surface and their cross (rows interior), ax and wx
            // against each (:52-196)
            X.orig.resize(n_exv);
            for (int i = 0; i < n_exv; ++i)
                X.orig[i] = make_float4((float)exv_xyz[3 * i], (float)exv_xyz[3 * i + 1], (float)exv_xyz[3 * i + 2], 0.f);
            XS.orig.resize(n_surf);
            for (int i = 0; i < n_surf; ++i)
                XS.orig[i] =
                    make_float4((float)surf_xyz[3 * i], (float)surf_xyz[3 * i + 1], (float)surf_xyz[3 * i + 2], 0.f);
            X.n_classes = XS.n_classes = 1;
            auto one = [&](int row_set, int col_set, int kind) {
                PairJob j;
                j.row_set = row_set;
                j.col_set = col_set;
                j.lattice = row_set >= 2 && col_set != 1;  // dummies only
                j.col_runs = row_set < 2 && col_set >= 2;   // waters against dummies
                j.rowmap = {n_rows++};
                rows_.push_back({kind, 0, 0});
                jobs.push_back(std::move(j));
            };
            one(2, -1, 3);  // xx interior
            one(3, -1, 6);  // xx surface
            one(2, 3, 7);   // xx interior x surface
            for (int set : {2, 3}) {
                PairJob ax;
                ax.row_set = 0;
                ax.col_set = set;
                ax.col_runs = true;
                ax.rowmap.assign(kMaxFF, -1);
                for (int f : present) {
                    ax.rowmap[f] = n_rows++;
                    rows_.push_back({set == 2 ? 4 : 8, f, 0});
                }
                jobs.push_back(std::move(ax));
            }
            one(1, 2, 5);  // wx interior
            one(1, 3, 9);  // wx surface
        }
        if (mode == Mode::FFGrid) {
            // dummies (CompactCoordinatesFF of EXCLUDED_VOLUME atoms, HistogramManagerMTFFGrid.cpp:49-57)
            X.orig.resize(n_exv);
            for (int i = 0; i < n_exv; ++i)
                X.orig[i] = make_float4((float)exv_xyz[3 * i], (float)exv_xyz[3 * i + 1], (float)exv_xyz[3 * i + 2], 0.f);
            X.n_classes = 1;
            // xx (self, :68-89), ax (atoms x dummies by atom type, :92-113), wx (waters x dummies, :116-137)
            PairJob xx, ax, wx;
            xx.row_set = 2;
            xx.lattice = true;
            xx.rowmap = {n_rows++};
            rows_.push_back({3, 0, 0});
            ax.row_set = 0;
            ax.col_set = 2;
            ax.rowmap.assign(kMaxFF, -1);
            for (int f : present) {
                ax.rowmap[f] = n_rows++;
                rows_.push_back({4, f, 0});
            }
            wx.row_set = 1;
            wx.col_set = 2;
            ax.col_runs = wx.col_runs = true;  // dummies in voxel order: z-runs
            wx.rowmap = {n_rows++};
            rows_.push_back({5, 0, 0});
            jobs.push_back(xx);
            jobs.push_back(ax);
            jobs.push_back(wx);
        }
    } else {
        // XYZW weights: one class per distinct weight value; w_i*w_j of a pair
        // is then the float product of its two class weights. Sets with more
        // than 64 distinct weights take the weighted engine instead: one class
        // each, and a second pass that sums every pair's weight product.
        auto classify = [](const float* p, int n, PointSet& ps, std::vector<float>& w, std::vector<float>& cw) {
            std::map<uint32_t, int> id;
            w.resize(n);
            for (int i = 0; i < n; ++i) {
                const float wi = p[4 * i + 3];
                uint32_t key;
                std::memcpy(&key, &wi, 4);
                auto it = id.find(key);
                if (it == id.end()) {
                    it = id.emplace(key, (int)cw.size()).first;
                    cw.push_back(wi);
                }
                w[i] = wi;
                ps.orig[i].w = float_of(it->second);
            }
            if (cw.empty()) cw.push_back(0.f);  // empty set: one (unused) class keeps row tables non-empty
            ps.n_classes = std::max<int>(1, (int)cw.size());
        };
        std::vector<float> atom_w, water_w;
        classify(atoms, n_atoms, A, atom_w, atom_class_w_);
        classify(waters, n_waters, W, water_w, water_class_w_);
        if (atom_class_w_.size() > kMaxWeightClasses || water_class_w_.size() > kMaxWeightClasses) {
            weighted_ = true;
            for (auto* ps : {&A, &W}) {
                ps->n_classes = 1;
                for (auto& v : ps->orig) v.w = float_of(0);
            }
            atom_class_w_.assign(1, 0.f);
            water_class_w_.assign(1, 0.f);
            weights.push_back(atom_w);
            weights.push_back(water_w);
        }
        // SimpleCalculator.h:144-160: sum of w*w (float products) in point order
        for (float x : atom_w) self_w2_atoms_ += (double)(x * x);
        for (float x : water_w) self_w2_waters_ += (double)(x * x);
        const int CA = A.n_classes, CW = W.n_classes;
        PairJob aa, ww, aw;
        aa.row_set = 0;
        aa.rowmap.assign(CA * CA, -1);
        for (int a = 0; a < CA; ++a)
            for (int b = a; b < CA; ++b) {
                aa.rowmap[a * CA + b] = aa.rowmap[b * CA + a] = n_rows++;
                rows_.push_back({0, a, b});
            }
        ww.row_set = 1;
        ww.rowmap.assign(CW * CW, -1);
        for (int a = 0; a < CW; ++a)
            for (int b = a; b < CW; ++b) {
                ww.rowmap[a * CW + b] = ww.rowmap[b * CW + a] = n_rows++;
                rows_.push_back({2, a, b});
            }
        // SimpleCalculator::enqueue_calculate_cross(atoms, waters): rows over waters
        aw.row_set = 1;
        aw.col_set = 0;
        aw.rowmap.assign(CW * CA, -1);
        for (int a = 0; a < CW; ++a)
            for (int b = 0; b < CA; ++b) {
                aw.rowmap[a * CA + b] = n_rows++;
                rows_.push_back({1, a, b});
            }
        jobs = {aa, ww, aw};
    }

    const int bound = bin_bound({&A, &W, &X, &XS}, inv_w);
    // at least the 10 bins every truncated histogram keeps (max_bin >= 10)
    const int nbins = (int)std::min<int64_t>((int64_t)s.bin_count, std::max(bound, 10));
    std::vector<PointSet> sets;
    sets.push_back(std::move(A));
    sets.push_back(std::move(W));
    if (grid) sets.push_back(std::move(X));
    if (mode == Mode::FFGridSurface) sets.push_back(std::move(XS));
    for (const PairJob& j : jobs) {
        const int64_t a = sets[j.row_set].size();
        pairs_ += j.col_set < 0 ? a * (a - 1) / 2 : a * sets[j.col_set].size();
    }
    spec_ = EngineSpec{std::move(sets), std::move(jobs), n_rows, inv_w, nbins, s.tile_j, s.lattice,
                       std::move(weights), nbins < bound ? 1 : 0};
    nbins_ = nbins;
    n_rows_ = n_rows;
    inv_w_ = inv_w;
    if (weighted_) engine();  // the weight pass checks its weights at construction
}

PairEngine& ResidentStructure::engine() {
    // the engine takes the spec's sets and jobs (spec_view reads them there)
    if (!engine_) engine_ = std::make_unique<PairEngine>(std::move(spec_));
    return *engine_;
}

std::unique_ptr<CompositeDistanceHistogramHIP> ResidentStructure::calculate(bool timed) {
    engine().run(cnt_, dsum_, timed, weighted_ ? &wsum_ : nullptr);
    return finalize();
}

std::vector<std::unique_ptr<CompositeDistanceHistogramHIP>> ResidentStructure::calculate_merged(
    const std::vector<ResidentStructure*>& members, int64_t* pairs, int64_t* kernel_pairs) {
    std::vector<SpecView> specs;
    for (ResidentStructure* m : members) {
        if (m->weighted_) throw std::runtime_error("ausaxs_hip: structures with per-point weights run alone");
        specs.push_back(m->spec_view());
    }
    std::vector<int> row0;
    PairEngine eng(merge_specs(specs, row0));
    std::vector<uint64_t> cnt;
    std::vector<int64_t> dsum;
    eng.run(cnt, dsum);
    if (pairs) *pairs = eng.pairs();
    if (kernel_pairs) *kernel_pairs = eng.kernel_pairs();
    // each member's rows at its own width; pairs at or beyond it go to its
    // overflow slot (nbins), as its own engine would put them
    const size_t ms = (size_t)eng.nbins() + 1;
    std::vector<std::unique_ptr<CompositeDistanceHistogramHIP>> out;
    for (size_t k = 0; k < members.size(); ++k) {
        ResidentStructure& m = *members[k];
        const int nb = m.nbins_;
        const size_t st = (size_t)nb + 1;
        m.cnt_.assign((size_t)m.n_rows_ * st, 0);
        m.dsum_.assign(m.cnt_.size(), 0);
        for (int r = 0; r < m.n_rows_; ++r) {
            const size_t src = (size_t)(row0[k] + r) * ms, dst = (size_t)r * st;
            for (int b = 0; b < nb; ++b) {
                m.cnt_[dst + b] = cnt[src + b];
                m.dsum_[dst + b] = dsum[src + b];
            }
            for (size_t b = nb; b < ms; ++b) m.cnt_[dst + nb] += cnt[src + b];
        }
        out.push_back(m.finalize());
    }
    return out;
}

std::unique_ptr<CompositeDistanceHistogramHIP> ResidentStructure::finalize() {
    const int nbins = nbins_;
    const int stride = nbins + 1;
    for (int r = 0; r < n_rows_; ++r)
        if (cnt_[(size_t)r * stride + nbins])
            throw std::runtime_error("ausaxs_hip: a pair distance falls beyond bin_count*bin_width (" +
                                     std::to_string(s_.bin_count) + " bins of " + std::to_string(s_.bin_width) +
                                     " A); increase bin_count or bin_width");
    const size_t nb = nbins;
    const float inv_w = inv_w_;
    const double unit = delta_unit(inv_w);

    // entry of one global row and bin, times the reference's template factor
    auto entry = [&](int row, int b, int factor) {
        const uint64_t c = cnt_[(size_t)row * stride + b];
        const int64_t s = dsum_[(size_t)row * stride + b];
        WeightedEntry e;
        e.value = (double)factor * (double)c;
        e.count = (uint32_t)(factor * c);
        e.bin_center = factor * distance_sum_u(c, b, s, inv_w, unit);
        return e;
    };

    auto res = std::make_unique<CompositeDistanceHistogramHIP>();
    res->mode = mode_;
    res->bin_width = s_.bin_width;
    res->weighted_bins = s_.weighted_bins;
    res->debye_terms = s_.debye_terms;

    if (mode_ != Mode::Simple) {
        // HistogramManagerMTFFGrid starts from HistogramManagerMTFFAvg<true, vbw>
        // (HistogramManagerMTFFGrid.cpp:47): FFAvg conventions, always weighted.
        const bool avg = mode_ != Mode::FFExplicit;
        const bool grid = mode_ == Mode::FFGrid || mode_ == Mode::FFGridSurface;
        const bool surf = mode_ == Mode::FFGridSurface;
        const double z = grid ? (n_atoms_ ? grid_volume_ * 0.334 / n_atoms_ : 0.0) : z_exv_;
        // working width: bins up to the last non-empty one of any row
        int nbw = 10;
        for (size_t r = 0; r < rows_.size(); ++r)
            for (int b = nbins - 1; b >= nbw; --b)
                if (cnt_[r * stride + b]) {
                    nbw = b + 1;
                    break;
                }
        nbw = std::min(nbw, nbins);
        const size_t nbv = nbw;
        EntryRows aa, aw, ax, axs;
        aa.reset(kMaxFF, kMaxFF, nbw);
        aw.reset(kMaxFF, 1, nbw);
        ax.reset(kMaxFF, 1, nbw);
        axs.reset(kMaxFF, 1, nbw);
        // xx / ax / wx: the (interior) dummies; xxs, xxc, axs, wxs: surface ones
        std::vector<WeightedEntry> ww(nbv), xx, wx, xxs, xxc, wxs;
        if (grid) xx.resize(nbv), wx.resize(nbv);
        if (surf) xxs.resize(nbv), xxc.resize(nbv), wxs.resize(nbv);
        for (size_t r = 0; r < rows_.size(); ++r) {
            const auto& k = rows_[r];
            WeightedEntry* dst = nullptr;
            int factor = 2;
            switch (k[0]) {
                case 0: dst = aa.touch(k[1], k[2]); break;
                case 1: dst = aw.touch(k[1], 0), factor = 1; break;
                case 2: dst = ww.data(); break;
                case 3: dst = xx.data(); break;
                case 4: dst = ax.touch(k[1], 0), factor = 1; break;
                case 5: dst = wx.data(), factor = 1; break;
                case 6: dst = xxs.data(); break;
                case 7: dst = xxc.data(); break;
                case 8: dst = axs.touch(k[1], 0), factor = 1; break;
                default: dst = wxs.data(), factor = 1; break;
            }
            for (int b = 0; b < nbw; ++b)
                if (cnt_[r * stride + b]) dst[b] = entry((int)r, b, factor);
        }
        if (avg) {
            // HistogramManagerMTFFAvg.cpp:27-34,111-118: every aa pair also lands in
            // [ff_i][exv] (ff of the lower original index) and [exv][exv]; every
            // aw pair in aw[exv]. Derived from the ordered cells.
            std::vector<std::array<int, 2>> keys;
            aa.for_each([&](int f, int g, const WeightedEntry*) {
                if (f >= 1 && g >= 1) keys.push_back({f, g});
            });
            WeightedEntry* xxrow = aa.touch(kExvBin, kExvBin);
            for (auto [f, g] : keys) {
                const WeightedEntry* src = aa.row(f, g);
                WeightedEntry* fx = aa.touch(f, kExvBin);
                for (int b = 0; b < nbw; ++b) {
                    add(fx[b], src[b]);
                    add(xxrow[b], src[b]);
                }
            }
            WeightedEntry* wexv = aw.touch(kExvBin, 0);
            for (int f = 1; f < kMaxFF; ++f)
                if (const WeightedEntry* src = aw.row(f, 0))
                    for (int b = 0; b < nbw; ++b) add(wexv[b], src[b]);
        }
        // self correlations (FFAvg.cpp:308-323, FFExplicit.cpp:331-346): one
        // (1, 1, 0) per atom of a type; integer sums, so one add per type is
        // the same
        for (int f = 0; f < kMaxFF; ++f)
            if (ff_count_[f]) add(aa.touch(f, f)[0], WeightedEntry{(double)ff_count_[f], ff_count_[f], 0});
        if (avg) add(aa.touch(kExvBin, kExvBin)[0], WeightedEntry{(double)n_atoms_, (uint32_t)n_atoms_, 0});
        add(ww[0], WeightedEntry{(double)n_waters_, (uint32_t)n_waters_, 0});

        std::vector<WeightedEntry> ptot(nbv);
        aa.for_each([&](int f1, int f2, const WeightedEntry* src) {
            if (f1 >= 1 && f2 >= 1)
                for (int b = 0; b < nbw; ++b) add(ptot[b], src[b]);
        });
        for (int f = 1; f < kMaxFF; ++f)
            if (const WeightedEntry* src = aw.row(f, 0))
                for (int b = 0; b < nbw; ++b) add(ptot[b], src[b]);
        for (int b = 0; b < nbw; ++b) add(ptot[b], ww[b]);
        const unsigned base = std::min<unsigned>(max_bin_of(ptot), (unsigned)nbw);

        if (avg) {  // FFAvg.cpp:355-364
            for (int f = 1; f < kMaxFF; ++f)
                if (WeightedEntry* r = aa.row(f, kExvBin))
                    for (unsigned b = 0; b < base; ++b) scale(r[b], z);
            if (WeightedEntry* r = aa.row(kExvBin, kExvBin))
                for (unsigned b = 0; b < base; ++b) scale(r[b], z * z);
            if (WeightedEntry* r = aw.row(kExvBin, 0))
                for (unsigned b = 0; b < base; ++b) scale(r[b], z);
        }
        unsigned max_bin = base;
        ptot.resize(base);
        if (grid) {
            // HistogramManagerMTFFGrid.cpp:159-224; HistogramManagerMTFFGridSurface.cpp:203-272
            add(xx[0], WeightedEntry{(double)n_exv_, (uint32_t)n_exv_, 0});
            if (surf) add(xxs[0], WeightedEntry{(double)n_surf_, (uint32_t)n_surf_, 0});
            unsigned mbx = 10;
            for (int i = nbw - 1; i >= 10; --i)
                if (xx[i].value != 0 || wx[i].value != 0 || (surf && (xxs[i].value != 0 || wxs[i].value != 0))) {
                    mbx = (unsigned)i + 1;
                    break;
                }
            // p_tot = FFAvg weighted counts (FFAvgBase.cpp get_weighted_counts: raw
            // products at the first q of the default axis, cw = 1), centres = FFAvg d axis
            const std::vector<double> base_axis = axis_of(ptot, s_.bin_width, true);
            const double q0 = 1e-4;
            std::vector<double> wc(base, 0.0);
            aa.for_each([&](int f1, int f2, const WeightedEntry* src) {
                if (f1 < 1 || f2 < 1) return;
                const double w = ff_atomic(f1, q0) * ff_atomic(f2, q0);
                for (unsigned b = 0; b < base; ++b) wc[b] += src[b].value * w;
            });
            for (int f = 1; f < kMaxFF; ++f)
                if (const WeightedEntry* src = aw.row(f, 0)) {
                    const double w = 2 * 1.0 * (ff_atomic(f, q0) * ff_atomic(kWaterBin, q0));
                    for (unsigned b = 0; b < base; ++b) wc[b] += src[b].value * w;
                }
            {
                const double w = 1.0 * 1.0 * (ff_atomic(kWaterBin, q0) * ff_atomic(kWaterBin, q0));
                for (unsigned b = 0; b < base; ++b) wc[b] += ww[b].value * w;
            }
            max_bin = std::max(mbx, base);
            std::vector<WeightedEntry> grid_ptot(max_bin);
            for (unsigned b = 0; b < base; ++b) grid_ptot[b] = WeightedEntry{wc[b], 0, base_axis[b]};
            // FFAvg rows beyond the FFAvg size are zero; exv rows become ax / wx / xx
            for (int f1 = 1; f1 < kMaxFF; ++f1) {
                for (int f2 = 1; f2 < kMaxFF; ++f2)
                    if (WeightedEntry* r = aa.row(f1, f2))
                        for (unsigned b = base; b < max_bin; ++b) r[b] = WeightedEntry{};
                if (WeightedEntry* r = aw.row(f1, 0))
                    for (unsigned b = base; b < max_bin; ++b) r[b] = WeightedEntry{};
            }
            for (unsigned b = base; b < max_bin; ++b) ww[b] = WeightedEntry{};
            std::vector<WeightedEntry> pax((size_t)max_bin), pxx((size_t)max_bin);
            for (unsigned b = 0; b < max_bin; ++b) {
                add(pax[b], wx[b]);
                if (surf) add(pax[b], wxs[b]);
            }
            for (int f = 0; f < kMaxFF; ++f) {
                const WeightedEntry* src = ax.row(f, 0);
                const WeightedEntry* srs = surf ? axs.row(f, 0) : nullptr;
                for (unsigned b = 0; b < max_bin; ++b) {
                    if (src) add(pax[b], src[b]);
                    if (srs) add(pax[b], srs[b]);
                }
            }
            for (unsigned b = 0; b < max_bin; ++b) {
                pxx[b] = xx[b];
                if (surf) add(pxx[b], xxs[b]), add(pxx[b], xxc[b]);
            }
            // the exv rows become the dummy histograms (FFGrid), or their
            // sums at unit surface scaling (FFGridSurface's initialize)
            for (int f = 0; f < kMaxFF; ++f) {
                const WeightedEntry* src = ax.row(f, 0);
                const WeightedEntry* srs = surf ? axs.row(f, 0) : nullptr;
                if (!src && !srs) {
                    aa.clear(f, kExvBin);
                    continue;
                }
                WeightedEntry* dst = aa.touch(f, kExvBin);
                for (unsigned b = 0; b < max_bin; ++b) {
                    dst[b] = src ? src[b] : WeightedEntry{};
                    if (srs) add(dst[b], srs[b]);
                }
            }
            {
                WeightedEntry* dst = aw.touch(kExvBin, 0);
                for (unsigned b = 0; b < max_bin; ++b) {
                    dst[b] = wx[b];
                    if (surf) add(dst[b], wxs[b]);
                }
            }
            std::copy_n(pxx.data(), max_bin, aa.touch(kExvBin, kExvBin));
            if (surf) {
                res->ax_interior.reset(kMaxFF, 1, (int)max_bin);
                res->ax_surface.reset(kMaxFF, 1, (int)max_bin);
                for (int f = 1; f < kMaxFF; ++f) {
                    if (const WeightedEntry* src = ax.row(f, 0)) std::copy_n(src, max_bin, res->ax_interior.touch(f, 0));
                    if (const WeightedEntry* src = axs.row(f, 0)) std::copy_n(src, max_bin, res->ax_surface.touch(f, 0));
                }
                res->xx_interior.assign(xx.begin(), xx.begin() + max_bin);
                res->xx_surface.assign(xxs.begin(), xxs.begin() + max_bin);
                res->xx_cross.assign(xxc.begin(), xxc.begin() + max_bin);
                res->wx_interior.assign(wx.begin(), wx.begin() + max_bin);
                res->wx_surface.assign(wxs.begin(), wxs.begin() + max_bin);
            }
            // both axes live on histograms of bin_count entries (WeightedDistribution1D
            // of the manager), and neither has its first bin forced to zero
            auto axis_full = [&](const std::vector<WeightedEntry>& p) {
                const double n = (double)s_.bin_count;
                const double width = (n * s_.bin_width) / n;
                std::vector<double> d(max_bin);
                for (unsigned b = 0; b < max_bin; ++b) {
                    const double bc = p[b].bin_center;
                    d[b] = ((double)(!bc) * (b * width) + bc) / (double)((!p[b].count) + p[b].count);
                }
                return d;
            };
            res->p_tot_ax = pax;
            res->p_xx = pxx;
            res->d_axis_ax = axis_full(pax);
            res->d_axis_xx = axis_full(res->p_xx);
            res->grid_volume = grid_volume_;
            res->grid_exv_volume = grid_exv_volume_;
            ptot = std::move(grid_ptot);
        }
        res->max_bin = max_bin;
        aa.truncate(max_bin);
        aw.truncate(max_bin);
        res->aa = std::move(aa);
        res->aw = std::move(aw);
        res->ww.assign(ww.begin(), ww.begin() + max_bin);
        res->p_tot = std::move(ptot);
    } else {
        std::vector<WeightedEntry> aa(nb), aw(nb), ww(nb);
        for (size_t r = 0; r < rows_.size(); ++r) {
            const auto& k = rows_[r];
            double wprod;
            std::vector<WeightedEntry>* dst;
            if (k[0] == 0) wprod = (double)(atom_class_w_[k[1]] * atom_class_w_[k[2]]), dst = &aa;
            else if (k[0] == 2) wprod = (double)(water_class_w_[k[1]] * water_class_w_[k[2]]), dst = &ww;
            else wprod = (double)(water_class_w_[k[1]] * atom_class_w_[k[2]]), dst = &aw;
            for (int b = 0; b < nbins; ++b) {
                WeightedEntry e = entry((int)r, b, 2);  // add<2> (TemplateHelperSimple.h; SimpleCalculator.h:195-207)
                if (weighted_) {
                    // the summed weight products of the bin's pairs
                    const size_t cell = (size_t)r * stride + b;
                    e.value = 2.0 * weight_value(wsum_.items[cell], wsum_.strips[cell], engine_->wscale());
                } else {
                    e.value *= wprod;
                }
                add((*dst)[b], e);
            }
        }
        // SimpleCalculator.h:144-160: bin 0 += (sum w^2, (unsigned) sum w^2, 0)
        const double ta = self_w2_atoms_, tw = self_w2_waters_;
        add(aa[0], WeightedEntry{ta, (uint32_t)ta, 0});
        add(ww[0], WeightedEntry{tw, (uint32_t)tw, 0});
        std::vector<WeightedEntry> ptot(nb);
        for (size_t b = 0; b < nb; ++b) {
            ptot[b] = aa[b];
            add(ptot[b], ww[b]);
            add(ptot[b], aw[b]);
        }
        const unsigned max_bin = std::min<unsigned>(max_bin_of(ptot), (unsigned)nbins);
        res->max_bin = max_bin;
        res->aa.reset(1, 1, (int)max_bin);
        res->aw.reset(1, 1, (int)max_bin);
        std::copy_n(aa.begin(), max_bin, res->aa.touch(0, 0));
        std::copy_n(aw.begin(), max_bin, res->aw.touch(0, 0));
        res->ww.assign(ww.begin(), ww.begin() + max_bin);
        res->p_tot.assign(ptot.begin(), ptot.begin() + max_bin);
    }
    res->d_axis = axis_of(res->p_tot, s_.bin_width,
                          s_.weighted_bins || mode_ == Mode::FFGrid || mode_ == Mode::FFGridSurface);
    return res;
}

// ---------------------------------------------------------------------------
// Debye transform
// ---------------------------------------------------------------------------
const std::vector<double>& CompositeDistanceHistogramHIP::sinqd(const std::vector<double>& q) {
    if (!cached_sinqd_.empty() && q == cached_q_) return cached_sinqd_;
    const int K = (int)max_bin;
    const bool grid = mode == Mode::FFGrid;
    // rows grouped by distance axis; H[axis] holds that group's rows
    std::vector<double> H[3];
    std::vector<std::array<int, 3>> keys[3];
    auto push_row = [&](int axis, int kind, int f1, int f2, const WeightedEntry* src) {
        bool nz = false;
        for (int b = 0; b < K; ++b) nz |= src[b].value != 0;
        if (!nz) return;
        keys[axis].push_back({kind, f1, f2});
        const size_t o = H[axis].size();
        H[axis].resize(o + (size_t)K);
        double* h = H[axis].data() + o;
        for (int b = 0; b < K; ++b) h[b] = src[b].value;
    };
    if (mode == Mode::Simple) {
        if (const WeightedEntry* r = aa.row(0, 0)) push_row(0, 0, 0, 0, r);
        if (const WeightedEntry* r = aw.row(0, 0)) push_row(0, 1, 0, 0, r);
        push_row(0, 2, 0, 0, ww.data());
    } else if (mode == Mode::FFGridSurface) {
        // CompositeDistanceHistogramFFGridSurface.cpp:104-266: atomic rows on
        // p_tot's axis; the interior and surface dummy profiles on the ax / xx
        // axes, kept apart so the surface scaling applies per q
        aa.for_each([&](int f1, int f2, const WeightedEntry* r) {
            if (f1 >= 1 && f2 >= 1) push_row(0, 0, f1, f2, r);
        });
        aw.for_each([&](int f, int, const WeightedEntry* r) {
            if (f >= 1) push_row(0, 1, f, 0, r);
        });
        push_row(0, 2, 0, 0, ww.data());
        ax_interior.for_each([&](int f, int, const WeightedEntry* r) { push_row(1, 10, f, 0, r); });
        ax_surface.for_each([&](int f, int, const WeightedEntry* r) { push_row(1, 11, f, 0, r); });
        push_row(1, 12, 0, 0, wx_interior.data());
        push_row(1, 13, 0, 0, wx_surface.data());
        push_row(2, 14, 0, 0, xx_interior.data());
        push_row(2, 15, 0, 0, xx_surface.data());
        push_row(2, 16, 0, 0, xx_cross.data());
    } else {
        // FFGrid (CompositeDistanceHistogramFFGrid.cpp:77-119): [ff][exv] and
        // aw[exv] on the ax axis, [exv][exv] on the xx axis, the rest on p_tot's
        aa.for_each([&](int f1, int f2, const WeightedEntry* r) {
            const int axis = !grid ? 0 : (f1 == kExvBin && f2 == kExvBin) ? 2 : (f2 == kExvBin ? 1 : 0);
            push_row(axis, 0, f1, f2, r);
        });
        aw.for_each([&](int f, int, const WeightedEntry* r) { push_row(grid && f == kExvBin ? 1 : 0, 1, f, 0, r); });
        push_row(0, 2, 0, 0, ww.data());
    }
    const std::vector<double>* axes[3] = {&d_axis, &d_axis_ax, &d_axis_xx};
    const int nq = (int)q.size();
    row_keys_.clear();
    axis_of_row_.clear();
    for (int a = 0; a < 3; ++a)
        for (auto& k : keys[a]) row_keys_.push_back(k), axis_of_row_.push_back(a);
    const int rows = (int)row_keys_.size();
    cached_sinqd_.assign((size_t)rows * nq, 0.0);
    cached_q_ = q;
    if (rows == 0 || nq == 0) return cached_sinqd_;

    Context& ctx = Context::get();
    hipStream_t st = ctx.stream();
    // Per-thread buffers and events, reused across calls (serving slots run
    // concurrently, each on its own thread and stream). The operands ([q |
    // per group: H rows, d axis]) are written into a per-thread pinned
    // buffer, which one kernel copies to the device; the groups'
    // contractions run back to back on the stream (one scratch area, used in
    // stream order) and write their results straight into the pinned buffer;
    // then one wait. No DMA transfer is involved: asynchronous copies from
    // many serving threads queue on the DMA engines (pinned hipMemcpyAsync
    // raised the median serving latency from 2.7 to 4.0 ms), and pageable
    // ones cost tens of microseconds of host time each.
    thread_local DevBuf<double> dev;
    thread_local DevBuf<char> scratch;
    thread_local MappedStage stage;
    thread_local hipEvent_t ev[2] = {nullptr, nullptr};
    for (auto& e : ev)
        if (!e) check(hipEventCreate(&e), "hipEventCreate");
    size_t n_in = (size_t)nq, scratch_bytes = 0;
    for (int a = 0; a < 3; ++a) {
        const int ra = (int)keys[a].size();
        if (ra == 0) continue;
        if ((int)axes[a]->size() < K) throw std::runtime_error("ausaxs_hip: distance axis shorter than the histogram");
        n_in += (size_t)ra * K + K;
        scratch_bytes = std::max(scratch_bytes, debye_scratch_bytes(ra, K, nq));
    }
    const size_t n_out = (size_t)rows * nq;
    dev.ensure(n_in);
    double* const pin = reinterpret_cast<double*>(stage.ensure((n_in + n_out) * sizeof(double)));
    double* const pin_dev = static_cast<double*>(stage.dev);
    scratch.ensure(scratch_bytes);
    std::memcpy(pin, q.data(), (size_t)nq * sizeof(double));
    size_t off = nq;
    size_t h_off[3] = {0, 0, 0}, d_off[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        if (keys[a].empty()) continue;
        h_off[a] = off;
        std::memcpy(pin + off, H[a].data(), H[a].size() * sizeof(double));
        off += H[a].size();
        d_off[a] = off;
        std::memcpy(pin + off, axes[a]->data(), (size_t)K * sizeof(double));
        off += K;
    }
    CopyOut up{};
    up.src[0] = reinterpret_cast<const uint32_t*>(pin_dev);
    up.words[0] = n_in * 2;
    up.n = 1;
    up.dst = reinterpret_cast<uint32_t*>(dev.p);
    check(hipEventRecord(ev[0], st), "hipEventRecord");
    check(launch_copy_out(up, st), "copy_out_kernel");
    double* out = pin_dev + n_in;
    size_t row0 = 0;
    for (int a = 0; a < 3; ++a) {
        const int ra = (int)keys[a].size();
        if (ra == 0) continue;
        DebyeLaunch dl{dev.p + h_off[a], dev.p + d_off[a], dev.p, ra, K, nq, debye_terms, out + row0 * nq, scratch.p};
        check(debye_terms > 0 ? launch_debye(dl, st) : launch_debye_f64(dl, st), "debye");
        row0 += ra;
    }
    check(hipEventRecord(ev[1], st), "hipEventRecord");
    ctx.sync();
    std::memcpy(cached_sinqd_.data(), pin + n_in, n_out * sizeof(double));
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
    last_debye_ms = ms;
    return cached_sinqd_;
}

std::vector<double> CompositeDistanceHistogramHIP::debye_transform(const std::vector<double>& q,
                                                                   std::vector<double>* profiles) {
    const std::vector<double>& S = sinqd(q);
    const int nq = (int)q.size();
    std::vector<double> prof((size_t)6 * nq, 0.0);
    double* aa_ = prof.data();
    double* ax_ = aa_ + nq;
    double* aw_ = aa_ + 2 * nq;
    double* xx_ = aa_ + 3 * nq;
    double* wx_ = aa_ + 4 * nq;
    double* ww_ = aa_ + 5 * nq;
    const FreeParams& p = params;
    if (mode == Mode::Simple) {
        // DistanceHistogram::debye_transform (DistanceHistogram.cpp:54-87) with
        // p = aa + k*aw + k^2*ww (CompositeDistanceHistogram.cpp:55-57)
        for (size_t r = 0; r < row_keys_.size(); ++r) {
            const int kind = row_keys_[r][0];
            double* dst = kind == 0 ? aa_ : kind == 1 ? aw_ : ww_;
            for (int k = 0; k < nq; ++k) dst[k] += S[r * nq + k] * std::exp(-q[k] * q[k]);
        }
        std::vector<double> I(nq);
        for (int k = 0; k < nq; ++k) I[k] = aa_[k] + p.cw * aw_[k] + p.cw * p.cw * ww_[k];
        if (profiles) *profiles = prof;
        return I;
    }
    const double rm2 = 1.62 * 1.62 / 4;  // FFExplicitBase.cpp:54-57
    const double gx_q0 = grid_exv_volume * 0.334, gx_exp = std::pow(grid_exv_volume, 2. / 3) / (4 * kPi);
    const bool surface = mode == Mode::FFGridSurface;
    // the form factors depend on q only: evaluated once per q grid and thread
    thread_local std::vector<double> ff_q, cached_ff, cached_fe;
    if (ff_q != q) {
        cached_ff.assign((size_t)nq * kMaxFF, 0.0);
        cached_fe.assign((size_t)nq * kMaxFF, 0.0);
        for (int k = 0; k < nq; ++k)
            for (int f = 0; f < kMaxFF; ++f) {
                cached_ff[(size_t)k * kMaxFF + f] = ff_atomic(f, q[k]);
                if (f >= 1) cached_fe[(size_t)k * kMaxFF + f] = ff_exv(f, q[k]);
            }
        ff_q = q;
    }
    for (int k = 0; k < nq; ++k) {
        const double qk = q[k];
        const double cx = mode != Mode::FFExplicit ? p.cx : std::pow(p.cx, 3) * std::exp(-rm2 * (p.cx * p.cx - 1) * qk * qk);
        const double* fa = cached_ff.data() + (size_t)k * kMaxFF;
        const double* fe = cached_fe.data() + (size_t)k * kMaxFF;
        const double fw = fa[kWaterBin];
        // exv form factor of the FFAvg table, or of the grid dummies
        // (ExvFormFactor(V), ExvFormFactor.h:23-34; V = 1 A^3 unless stretched)
        const double fx =
            mode == Mode::FFGrid || surface ? gx_q0 * std::exp(-gx_exp * qk * qk) : fa[kExvBin];
        const double cr = cx * p.crho, cr2 = cr * cr;  // pow(cx * crho, 2): the product rounded once
        if (surface) {
            // CompositeDistanceHistogramFFGridSurface::cache_refresh_intensity_profiles
            // (:137-266): the surface dummies scale by exv_factor(q, cx) =
            // cx^3 exp(-rm2 (cx^2 - 1) q^2), xx = xx_i + c^2 xx_s + c xx_c
            const double c = std::pow(p.cx, 3) * std::exp(-rm2 * (std::pow(p.cx, 2) - 1) * qk * qk);
            const double c2 = std::pow(c, 2);
            for (size_t r = 0; r < row_keys_.size(); ++r) {
                const int kind = row_keys_[r][0], f1 = row_keys_[r][1], f2 = row_keys_[r][2];
                const double sv = S[r * nq + k];
                switch (kind) {
                    case 0: aa_[k] += sv * fa[f1] * fa[f2]; break;
                    case 1: aw_[k] += 2 * p.cw * sv * (fa[f1] * fw); break;
                    case 2: ww_[k] += p.cw * p.cw * sv * (fw * fw); break;
                    case 10: ax_[k] += 2 * p.crho * sv * (fa[f1] * fx); break;
                    case 11: ax_[k] += 2 * p.crho * (c * sv) * (fa[f1] * fx); break;
                    case 12: wx_[k] += 2 * p.crho * sv * p.cw * (fw * fx); break;
                    case 13: wx_[k] += 2 * p.crho * (c * sv) * p.cw * (fw * fx); break;
                    case 14: xx_[k] += p.crho * p.crho * sv * (fx * fx); break;
                    case 15: xx_[k] += p.crho * p.crho * (c2 * sv) * (fx * fx); break;
                    default: xx_[k] += p.crho * p.crho * (c * sv) * (fx * fx); break;
                }
                (void)f2;
            }
            continue;
        }
        for (size_t r = 0; r < row_keys_.size(); ++r) {
            const int kind = row_keys_[r][0], f1 = row_keys_[r][1], f2 = row_keys_[r][2];
            const double s = S[r * nq + k];
            if (mode != Mode::FFExplicit) {
                // FFAvgBase.cpp:470-558 with the raw_atomic_table products; FFGrid
                // replaces the exv form factor by ExvFormFactor(V = 1 A^3)
                // (CompositeDistanceHistogramFFGrid.cpp:36-75)
                if (kind == 0) {
                    if (f1 >= 1 && f2 >= 1) aa_[k] += s * fa[f1] * fa[f2];
                    else if (f1 >= 1 && f2 == kExvBin) ax_[k] += 2 * p.crho * cx * s * (fa[f1] * fx);
                    else if (f1 == kExvBin && f2 == kExvBin) xx_[k] += cr2 * s * (fx * fx);
                } else if (kind == 1) {
                    if (f1 >= 1) aw_[k] += 2 * p.cw * s * (fa[f1] * fw);
                    else wx_[k] += 2 * p.crho * cx * p.cw * s * (fx * fw);
                } else {
                    ww_[k] += p.cw * p.cw * s * fw * fw;
                }
            } else {
                // FFExplicitBase.cpp:105-212: raw_atomic, raw_cross (atomic x exv), raw_exv tables
                if (kind == 0) {
                    if (f1 < 1 || f2 < 1) continue;
                    aa_[k] += s * fa[f1] * fa[f2];
                    ax_[k] += p.crho * cx * s * (fa[f1] * fe[f2] + fa[f2] * fe[f1]);
                    xx_[k] += cr2 * s * fe[f1] * fe[f2];
                } else if (kind == 1) {
                    if (f1 < 1) continue;
                    aw_[k] += 2 * p.cw * s * fa[f1] * fw;
                    wx_[k] += p.crho * cx * p.cw * s * (fw * fe[f1] + fa[f1] * fe[kWaterBin]);
                } else {
                    ww_[k] += p.cw * p.cw * s * fw * fw;
                }
            }
        }
        if (p.dw_atomic != 0 || p.dw_exv != 0) {  // FFAvgBase.cpp:310-374
            const double ba = std::exp(-qk * qk * p.dw_atomic * p.dw_atomic * 0.5);
            const double bx = std::exp(-qk * qk * p.dw_exv * p.dw_exv * 0.5);
            aa_[k] *= ba * ba;
            xx_[k] *= bx * bx;
            ax_[k] *= ba * bx;
            aw_[k] *= ba;
            wx_[k] *= bx;
        }
    }
    std::vector<double> I(nq);
    for (int k = 0; k < nq; ++k) I[k] = aa_[k] - ax_[k] + aw_[k] + xx_[k] - wx_[k] + ww_[k];
    if (profiles) *profiles = prof;
    return I;
}

// ---------------------------------------------------------------------------
// raw pair set
// ---------------------------------------------------------------------------
void pack_ff_points(const double* xyz, const int32_t* ff, int n, int32_t fixed_type, std::vector<float>& out) {
    out.resize(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) out[4 * (size_t)i + k] = (float)xyz[3 * (size_t)i + k];
        const int32_t t = ff ? ff[i] : fixed_type;
        std::memcpy(&out[4 * (size_t)i + 3], &t, 4);
    }
}

std::unique_ptr<ResidentStructure> make_grid_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                       const double* waters_xyz, int m, const HistSettings& s) {
    for (int i = 0; i < n; ++i)
        if (ff[i] < 1 || ff[i] >= kMaxFF)
            throw std::runtime_error("ausaxs_hip: atom form factor type outside [1, 14]");
    GridExv g = make_grid_exv(atoms_xyz, ff, n, waters_xyz, m);
    std::vector<float> a, w;
    pack_ff_points(atoms_xyz, ff, n, 0, a);
    pack_ff_points(waters_xyz, nullptr, m, kWaterBin, w);
    return std::make_unique<ResidentStructure>(Mode::FFGrid, a.data(), n, w.data(), m, s, 0.0, g.xyz.data(),
                                               (int)(g.xyz.size() / 3), g.volume);
}

std::unique_ptr<ResidentStructure> make_ffavg_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                        const double* waters_xyz, int m, const HistSettings& s) {
    for (int i = 0; i < n; ++i)
        if (ff[i] < 1 || ff[i] >= kMaxFF)
            throw std::runtime_error("ausaxs_hip: atom form factor type outside [1, 14]");
    // HistogramManagerMTFFAvg.cpp:355-359: Z_exv = V_grid * 0.334 e/A^3 / N,
    // V_grid = Molecule::get_volume_grid() (Molecule.cpp:147-150): the
    // volume of the molecule's grid (Grid(bodies), waters included) after
    // expand_volume, which the excluded-volume grid builder counts
    const double v_grid = n ? make_grid_exv(atoms_xyz, ff, n, waters_xyz, m).volume : 0.0;
    const double z_exv = n ? v_grid * kWaterChargeDensity / n : 0.0;
    std::vector<float> a, w;
    pack_ff_points(atoms_xyz, ff, n, 0, a);
    pack_ff_points(waters_xyz, nullptr, m, kWaterBin, w);
    return std::make_unique<ResidentStructure>(Mode::FFAvg, a.data(), n, w.data(), m, s, z_exv);
}

std::unique_ptr<ResidentStructure> make_grid_surface_structure(const double* atoms_xyz, const int32_t* ff, int n,
                                                               const double* waters_xyz, int m,
                                                               const HistSettings& s) {
    for (int i = 0; i < n; ++i)
        if (ff[i] < 1 || ff[i] >= kMaxFF)
            throw std::runtime_error("ausaxs_hip: atom form factor type outside [1, 14]");
    GridExv g = make_grid_exv(atoms_xyz, ff, n, waters_xyz, m, true, s.surface_thickness);
    std::vector<float> a, w;
    pack_ff_points(atoms_xyz, ff, n, 0, a);
    pack_ff_points(waters_xyz, nullptr, m, kWaterBin, w);
    return std::make_unique<ResidentStructure>(Mode::FFGridSurface, a.data(), n, w.data(), m, s, 0.0, g.xyz.data(),
                                               (int)(g.xyz.size() / 3), g.volume, g.surface.data(),
                                               (int)(g.surface.size() / 3));
}

ScalableGridStructure::ScalableGridStructure(const double* atoms_xyz, const int32_t* ff, int n,
                                             const double* waters_xyz, int m, const HistSettings& s)
    : n_(n), m_(m), s_(s) {
    for (int i = 0; i < n; ++i)
        if (ff[i] < 1 || ff[i] >= kMaxFF)
            throw std::runtime_error("ausaxs_hip: atom form factor type outside [1, 14]");
    GridExv g = make_grid_exv(atoms_xyz, ff, n, waters_xyz, m);
    dummies_ = std::move(g.xyz);
    grid_volume_ = g.volume;
    pack_ff_points(atoms_xyz, ff, n, 0, atoms_);
    pack_ff_points(waters_xyz, nullptr, m, kWaterBin, waters_);
}

std::unique_ptr<CompositeDistanceHistogramHIP> ScalableGridStructure::calculate(double k) {
    if (!(k > 0) || !std::isfinite(k)) throw std::runtime_error("ausaxs_hip: the excluded-volume scale must be positive");
    if (!st_ || k != k_) {
        // eval_scaled_exv: coord.value.pos *= scale on the float dummies
        std::vector<double> x(dummies_.size());
        for (size_t i = 0; i < x.size(); ++i) x[i] = (double)(float)dummies_[i] * k;
        st_ = std::make_unique<ResidentStructure>(Mode::FFGrid, atoms_.data(), n_, waters_.data(), m_, s_, 0.0,
                                                  x.data(), (int)(x.size() / 3), grid_volume_);
        // V = (exv.width * k)^3 with the default 1 A width (CompositeDistanceHistogramFFGridScalableExv.cpp:17)
        st_->set_grid_exv_volume(std::pow(1.0 * k, 3));
        k_ = k;
    }
    return st_->calculate();
}

void raw_pair_histogram(const float* a, int n, const float* b, int m, const HistSettings& s, uint64_t* count,
                        double* dsum) {
    const float inv_w = inv_width(s);
    PointSet A, B;
    A.orig = to_points(a, n);
    std::vector<PointSet> sets;
    PairJob job;
    job.row_set = 0;
    job.rowmap = {0};
    if (b) {
        B.orig = to_points(b, m);
        job.col_set = 1;
    }
    const int nbins = (int)s.bin_count;
    sets.push_back(std::move(A));
    if (b) sets.push_back(std::move(B));
    PairEngine eng(std::move(sets), {job}, 1, inv_w, nbins, s.tile_j);
    std::vector<uint64_t> c;
    std::vector<int64_t> ds;
    eng.run(c, ds);
    if (eng.overflowed(c)) throw std::runtime_error("ausaxs_hip: pair distance beyond bin_count*bin_width");
    const double unit = delta_unit(inv_w);
    for (int k = 0; k < nbins; ++k) {
        count[k] = c[k];
        dsum[k] = distance_sum_u(c[k], (int)k, ds[k], inv_w, unit);
    }
}

}  // namespace ausaxs::gpu
