// Batched rigid-body scan (see rigid.hpp).
#include "rigid.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace ausaxs::gpu {

namespace {
constexpr int kMaxBatch = 1024;  // conformations per graph launch

// bounding sphere (centre, radius) of n points given as x,y,z with stride
void sphere_of(const double* p, int n, int stride, double c[3], double& r) {
    c[0] = c[1] = c[2] = 0;
    r = n ? 0 : -1;
    if (!n) return;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) c[k] += p[(size_t)i * stride + k];
    for (int k = 0; k < 3; ++k) c[k] /= n;
    for (int i = 0; i < n; ++i) {
        double d2 = 0;
        for (int k = 0; k < 3; ++k) d2 += (p[(size_t)i * stride + k] - c[k]) * (p[(size_t)i * stride + k] - c[k]);
        r = std::max(r, std::sqrt(d2));
    }
}

// a set's layout moved to another place of the point buffer
void shift_layout(SetLayout& L, int64_t off) {
    L.base = (int)(L.base + off);
    for (int& v : L.seg_start) v = (int)(v + off);
}
}  // namespace

struct RigidScan::Hydration {
    std::vector<PointSet> sets;         // per conformation: float waters, weight class ids
    std::vector<std::vector<float>> w;  // per conformation: weights, original order
    std::vector<double> sphere;         // per conformation: bounding-sphere centre (3) and radius
};

struct RigidScan::Plan {
    int n_conf = 0, nbins = 0, tile_j = 0, grid = 0;
    int rows = 1, copies = 1;  // LDS histogram rows and copies of the pair kernel
    std::vector<int> moving;
    int64_t S = 0, M = 0;  // shared points, points per replica
    int n_slots = 0;
    std::vector<PairWork> items;
    MergePlan mp;
    std::vector<StripLaunch> strips;  // device pointers filled at build
    std::vector<StripLaunch> wstrips;  // weighted scans: the same strips of the weight pass
    int64_t pairs = 0;
    int nq_cap = 0;
    double graph_cw = 0;
    int graph_nq = -1;

    DevBuf<float4> pts, orig;
    DevBuf<float4> pts_w, orig_w;  // weighted scans: weights in the fourth slot
    DevBuf<float> slot_w;
    DevBuf<uint64_t> wcnt;
    DevBuf<int64_t> wsum, wstrip;
    DevBuf<PairWork> d_items;
    DevBuf<int32_t> row_list, chunk_ptr, row_chunk, rowmaps;
    DevBuf<uint64_t> partial, cnt, tmp_cnt;
    DevBuf<int64_t> dsum, tmp_dsum;
    DevBuf<int32_t> slot_src, slot_body, slot_dst, slot_dsto, slot_tag, slot_cls, body_moving;
    DevBuf<double> T, q, H, dbar, I;
    DevBuf<int> max_bin, flags;
    HostBuf<double> hT, hq, hI;
    HostBuf<int> hflags;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[4]{};

    // rows per conformation: [aa class pairs | aw water class x class] or, in a
    // hydration-modified batch, [aa | aw, every water class x class | ww class pairs]
    int R = 0;
    bool hydrated = false;
    std::vector<int32_t> row_kind;  // hydration-modified batches (else the scan's own)
    std::vector<double> row_w;
    DevBuf<int32_t> d_row_kind;
    DevBuf<double> d_row_w;
    std::vector<double> ww_self;    // [n_conf][2]: each water set's self term
    DevBuf<double> d_ww_self;
    // tail strips of the per-conformation water jobs, one launch for all
    int n_table = 0;
    int64_t table_blocks = 0;
    DevBuf<StripLaunch> d_table, d_wtable;
    DevBuf<int64_t> d_first;

    Plan() {
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    }
    ~Plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    void drop_graph() {
        if (exec) check(hipGraphExecDestroy(exec), "hipGraphExecDestroy");
        if (graph) check(hipGraphDestroy(graph), "hipGraphDestroy");
        exec = nullptr;
        graph = nullptr;
        graph_nq = -1;
    }
};

RigidScan::~RigidScan() = default;

int RigidScan::tri(int a, int b) const {
    if (a > b) std::swap(a, b);
    // rows of (a, b), a <= b, in row-major upper-triangle order
    return a * C_ - a * (a - 1) / 2 + (b - a);
}

RigidScan::RigidScan(const double* xyz, const float* w, const int32_t* sizes, int n_bodies, const double* wxyz,
                     const float* ww, int n_waters, const HistSettings& s)
    : s_(s), inv_w_(inv_width(s)) {
    if (n_bodies < 1) throw std::runtime_error("ausaxs_hip: a rigid-body scan needs at least one body");
    if (n_waters < 0) throw std::runtime_error("ausaxs_hip: negative water count");
    if (!(s.bin_width > 0) || s.bin_count < 10) throw std::runtime_error("ausaxs_hip: invalid bin settings");
    if (s.tile_j != 0) check_tile_j(s.tile_j);
    body_off_.assign(1, 0);
    for (int b = 0; b < n_bodies; ++b) {
        if (sizes[b] < 0) throw std::runtime_error("ausaxs_hip: negative body size");
        sizes_.push_back(sizes[b]);
        body_off_.push_back(body_off_.back() + sizes[b]);
    }
    const int n = body_off_.back();
    xyz_.assign(xyz, xyz + 3 * (size_t)n);
    w_.assign(w, w + n);
    for (double v : xyz_)
        if (!std::isfinite(v)) throw std::runtime_error("ausaxs_hip: body coordinates must be finite");

    // weight classes over bodies and waters
    std::map<uint32_t, int> id;
    auto class_of = [&](float wi) {
        uint32_t key;
        std::memcpy(&key, &wi, 4);
        auto it = id.find(key);
        if (it == id.end()) {
            it = id.emplace(key, (int)class_w_.size()).first;
            class_w_.push_back(wi);
        }
        return it->second;
    };
    point_cls_.resize(n);
    for (int i = 0; i < n; ++i) point_cls_[i] = class_of(w_[i]);
    waters_.orig.resize(n_waters);
    std::vector<float> water_w(n_waters);
    std::vector<bool> in_water;
    for (int i = 0; i < n_waters; ++i) {
        water_w[i] = ww[i];
        const int c = class_of(ww[i]);
        waters_.orig[i] = make_float4((float)wxyz[3 * i], (float)wxyz[3 * i + 1], (float)wxyz[3 * i + 2],
                                      float_of(c));
        if ((int)in_water.size() <= c) in_water.resize(c + 1, false);
        in_water[c] = true;
    }
    if (class_w_.size() > kMaxWeightClasses) {
        // more distinct weights than class-pair rows can hold: one class, and
        // a weight pass that sums every pair's weight product
        weighted_ = true;
        class_w_.assign(1, 0.f);
        std::fill(point_cls_.begin(), point_cls_.end(), 0);
        for (auto& v : waters_.orig) v.w = float_of(0);
        in_water.assign(n_waters ? 1 : 0, true);
    }
    C_ = std::max<int>(1, (int)class_w_.size());
    if (class_w_.empty()) class_w_.push_back(0.f);
    waters_.n_classes = C_;
    water_w_ = water_w;
    if (weighted_) {
        double wmax = 0;
        for (float v : w_) wmax = std::max(wmax, (double)std::fabs(v));
        for (float v : water_w_) wmax = std::max(wmax, (double)std::fabs(v));
        for (float v : w_)
            if (!std::isfinite(v)) throw std::runtime_error("ausaxs_hip: weights must be finite");
        wscale_ = weight_scale(wmax);
    }
    for (int c = 0; c < (int)in_water.size(); ++c)
        if (in_water[c]) water_classes_.push_back(c);

    // rows of one conformation: aa class pairs, then aw (water class x class)
    R_aa_ = C_ * (C_ + 1) / 2;
    R_ = R_aa_ + (int)water_classes_.size() * C_;
    row_kind_.assign(R_, 0);
    row_w_.assign(R_, 0.0);
    for (int a = 0; a < C_; ++a)
        for (int b = a; b < C_; ++b) row_w_[tri(a, b)] = (double)(class_w_[a] * class_w_[b]);
    for (size_t k = 0; k < water_classes_.size(); ++k)
        for (int b = 0; b < C_; ++b) {
            const int r = R_aa_ + (int)k * C_ + b;
            row_kind_[r] = 1;
            row_w_[r] = (double)(class_w_[water_classes_[k]] * class_w_[b]);
        }

    // bounding spheres for the bin bound of each scan
    center_.resize(4 * (size_t)n_bodies);
    for (int b = 0; b < n_bodies; ++b)
        sphere_of(&xyz_[3 * (size_t)body_off_[b]], sizes_[b], 3, &center_[4 * b], center_[4 * b + 3]);
    if (n_waters) {
        std::vector<double> wd(3 * (size_t)n_waters);
        for (int i = 0; i < n_waters; ++i)
            for (int k = 0; k < 3; ++k) wd[3 * i + k] = (&waters_.orig[i].x)[k];
        sphere_of(wd.data(), n_waters, 3, wcenter_, wradius_);
    }

    // cached self terms (calc_self_correlation, calc_ww; PartialHistogramManagerMT.cpp:259-294)
    std::vector<PointSet> sets(n_bodies + 1);
    std::vector<PairJob> jobs;
    std::vector<int32_t> aa_map((size_t)C_ * C_), ww_map((size_t)C_ * C_);
    for (int a = 0; a < C_; ++a)
        for (int b = 0; b < C_; ++b) {
            aa_map[a * C_ + b] = tri(a, b);
            ww_map[a * C_ + b] = R_aa_ + tri(a, b);
        }
    for (int b = 0; b < n_bodies; ++b) {
        PointSet& ps = sets[b];
        ps.n_classes = C_;
        ps.orig.resize(sizes_[b]);
        for (int i = 0; i < sizes_[b]; ++i) {
            const size_t g = (size_t)body_off_[b] + i;
            ps.orig[i] = make_float4((float)xyz_[3 * g], (float)xyz_[3 * g + 1], (float)xyz_[3 * g + 2],
                                     float_of(point_cls_[g]));
        }
        jobs.push_back(PairJob{b, -1, aa_map});
    }
    sets[n_bodies] = waters_;
    jobs.push_back(PairJob{n_bodies, -1, ww_map});
    std::vector<const PointSet*> ptrs;
    for (auto& ps : sets) ptrs.push_back(&ps);
    base_nb_ = std::max(10, std::min<int>((int)s_.bin_count, bin_bound(ptrs, inv_w_)));
    std::vector<uint64_t> cnt;
    std::vector<int64_t> dsum;
    WeightSums wsum;
    double wscale = 1;
    {
        std::vector<std::vector<float>> weights;
        if (weighted_) {
            for (int b = 0; b < n_bodies; ++b) weights.emplace_back(w_.begin() + body_off_[b], w_.begin() + body_off_[b + 1]);
            weights.push_back(water_w_);
        }
        PairEngine eng(sets, jobs, 2 * R_aa_, inv_w_, base_nb_, s_.tile_j ? s_.tile_j : kDefaultTileJ, 1, std::move(weights));
        eng.run(cnt, dsum, false, weighted_ ? &wsum : nullptr);
        if (eng.overflowed(cnt))
            throw std::runtime_error("ausaxs_hip: a pair distance falls beyond bin_count*bin_width");
        wscale = eng.wscale();
    }
    const int stride = base_nb_ + 1;
    const double unit = delta_unit(inv_w_);
    base_aa_.assign(base_nb_, WeightedEntry{});
    base_ww_.assign(base_nb_, WeightedEntry{});
    for (int r = 0; r < 2 * R_aa_; ++r) {
        const int rr = r % R_aa_;
        std::vector<WeightedEntry>& dst = r < R_aa_ ? base_aa_ : base_ww_;
        for (int b = 0; b < base_nb_; ++b) {
            const uint64_t c = cnt[(size_t)r * stride + b];
            if (!c) continue;
            const int64_t ds = dsum[(size_t)r * stride + b];
            WeightedEntry e;  // add<2> (SimpleCalculator.h:195-207)
            const size_t cell = (size_t)r * stride + b;
            e.value = weighted_ ? 2.0 * weight_value(wsum.items[cell], wsum.strips[cell], wscale)
                                : 2.0 * (double)c * row_w_[rr];
            e.count = (uint32_t)(2 * c);
            e.bin_center = 2.0 * distance_sum_u(c, b, ds, inv_w_, unit);
            entry_add(dst[b], e);
        }
    }
    // SimpleCalculator.h:144-160: each self job adds (sum w^2, (unsigned) sum w^2, 0) at bin 0
    auto self_term = [](const float* v, int m) {
        double t = 0;
        for (int i = 0; i < m; ++i) t += (double)(v[i] * v[i]);
        return WeightedEntry{t, (uint32_t)t, 0};
    };
    for (int b = 0; b < n_bodies; ++b) entry_add(base_aa_[0], self_term(&w_[body_off_[b]], sizes_[b]));
    entry_add(base_ww_[0], self_term(water_w.data(), n_waters));

    Context& ctx = Context::get();
    std::vector<double> base(6 * (size_t)base_nb_);
    for (int b = 0; b < base_nb_; ++b) {
        base[b] = base_aa_[b].value;
        base[base_nb_ + b] = base_aa_[b].count;
        base[2 * base_nb_ + b] = base_aa_[b].bin_center;
        base[3 * base_nb_ + b] = base_ww_[b].value;
        base[4 * base_nb_ + b] = base_ww_[b].count;
        base[5 * base_nb_ + b] = base_ww_[b].bin_center;
    }
    d_base_.upload(base.data(), base.size(), ctx.stream());
    d_xyz_.upload(xyz_.data(), xyz_.size(), ctx.stream());
    d_row_kind_.upload(row_kind_.data(), row_kind_.size(), ctx.stream());
    d_row_w_.upload(row_w_.data(), row_w_.size(), ctx.stream());
    ctx.sync();
}

int RigidScan::bound_bins(const double* T, int n_conf, const Hydration* hy) const {
    // bounding spheres moved by each conformation; ||R x|| <= ||R||_F ||x||
    const int B = n_bodies();
    double dmax = hy ? 0.0 : 2 * std::max(wradius_, 0.0), cmax = 0;
    std::vector<double> c(3 * (size_t)B), r(B);
    for (int k = 0; k < n_conf; ++k) {
        // the waters of this conformation
        const double* wc = hy ? &hy->sphere[4 * (size_t)k] : wcenter_;
        const double wr = hy ? hy->sphere[4 * (size_t)k + 3] : wradius_;
        dmax = std::max(dmax, 2 * wr);
        for (int e = 0; e < 3 && wr >= 0; ++e) cmax = std::max(cmax, std::fabs(wc[e]) + wr);
        for (int b = 0; b < B; ++b) {
            const double* t = T + ((size_t)k * B + b) * 12;
            double f = 0;
            for (int e = 0; e < 9; ++e) f += t[e] * t[e];
            const double* cb = &center_[4 * b];
            for (int e = 0; e < 3; ++e) {
                c[3 * b + e] = t[3 * e] * cb[0] + t[3 * e + 1] * cb[1] + t[3 * e + 2] * cb[2] + t[9 + e];
                cmax = std::max(cmax, std::fabs(c[3 * b + e]));
            }
            r[b] = sizes_[b] ? cb[3] * std::sqrt(f) : -1;
            if (!std::isfinite(r[b]) || !std::isfinite(c[3 * b]) || !std::isfinite(c[3 * b + 1]) ||
                !std::isfinite(c[3 * b + 2]))
                throw std::runtime_error("ausaxs_hip: transforms must be finite");
        }
        for (int i = 0; i < B; ++i) {
            if (r[i] < 0) continue;
            dmax = std::max(dmax, 2 * r[i]);
            for (int j = 0; j < i; ++j) {
                if (r[j] < 0) continue;
                const double dx = c[3 * i] - c[3 * j], dy = c[3 * i + 1] - c[3 * j + 1], dz = c[3 * i + 2] - c[3 * j + 2];
                dmax = std::max(dmax, std::sqrt(dx * dx + dy * dy + dz * dz) + r[i] + r[j]);
            }
            if (wr >= 0) {
                const double dx = c[3 * i] - wc[0], dy = c[3 * i + 1] - wc[1], dz = c[3 * i + 2] - wc[2];
                dmax = std::max(dmax, std::sqrt(dx * dx + dy * dy + dz * dz) + r[i] + wr);
            }
        }
    }
    // float rounding of the moved coordinates: a few ulp of their magnitude
    dmax += 1e-5 * (cmax + dmax) + 1e-3;
    const double x = dmax * (double)inv_w_ * (1 + 1e-5) + 2;
    if (!(x < 1e9) || x > kMaxExtentBins)
        throw std::runtime_error("ausaxs_hip: conformations span more than " + std::to_string(kMaxExtentBins) +
                                 " bins");
    int nb = (int)std::ceil(x);
    nb = (nb + 15) / 16 * 16;  // a little slack: nearby scans reuse one plan
    return std::max(10, std::min<int>(nb, (int)s_.bin_count));
}

std::unique_ptr<RigidScan::Plan> RigidScan::build_plan(int n_conf, const std::vector<int>& moving, int nbins,
                                                       int tile_j, const Hydration* hy) {
    check_pair_lds(nbins);
    check_tile_j(tile_j);
    const int B = n_bodies();
    auto P = std::make_unique<Plan>();
    Plan& p = *P;
    p.n_conf = n_conf;
    p.nbins = nbins;
    p.tile_j = tile_j;
    p.moving = moving;
    p.hydrated = hy != nullptr;
    const int R = hy ? R_aa_ + C_ * C_ + R_aa_ : R_;
    p.R = R;
    if ((int64_t)(n_conf + 1) * R >= std::numeric_limits<int32_t>::max() / (nbins + 1))
        throw std::runtime_error("ausaxs_hip: too many histogram rows in one rigid-body batch");
    const int n_rows = (n_conf + 1) * R;

    // set 0 = waters (none shared in a hydration-modified batch), 1 + b = body b;
    // layout: waters, static bodies (shared), moving bodies (replica 0)
    std::vector<PointSet> sets(1 + B);
    if (hy) sets[0].n_classes = C_;
    else sets[0] = waters_;
    for (int b = 0; b < B; ++b) {
        sets[1 + b].n_classes = C_;
        sets[1 + b].orig.resize(sizes_[b]);
        for (int i = 0; i < sizes_[b]; ++i)
            sets[1 + b].orig[i] = make_float4(0.f, 0.f, 0.f, float_of(point_cls_[body_off_[b] + i]));
    }
    std::vector<SetLayout> L(1 + B);
    std::vector<std::vector<int32_t>> perm(B);
    std::vector<float4> buf;
    std::vector<int32_t> wperm;
    L[0] = layout_set(sets[0], buf, weighted_ ? &wperm : nullptr);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) p.S = (int64_t)buf.size();
        for (int b = 0; b < B; ++b)
            if ((moving[b] != 0) == (pass == 1)) L[1 + b] = layout_set(sets[1 + b], buf, &perm[b]);
    }
    p.M = (int64_t)buf.size() - p.S;
    const int64_t body_total = p.S + (int64_t)n_conf * p.M;
    int64_t water_total = 0;
    if (hy)
        for (const PointSet& w : hy->sets) water_total += w.size();
    const int64_t total = body_total + water_total;
    if (total >= std::numeric_limits<int32_t>::max())
        throw std::runtime_error("ausaxs_hip: too many points in one rigid-body batch");
    // hydration-modified batches: each conformation's waters after the body replicas
    std::vector<SetLayout> LW;
    std::vector<float4> wbuf, worig;
    std::vector<int32_t> wperm_h;
    if (hy) {
        for (int c = 0; c < n_conf; ++c) {
            LW.push_back(layout_set(hy->sets[c], wbuf, weighted_ ? &wperm_h : nullptr));
            shift_layout(LW.back(), body_total);
        }
        worig.resize(wbuf.size());
        for (int c = 0; c < n_conf; ++c)
            std::copy(hy->sets[c].orig.begin(), hy->sets[c].orig.end(), worig.begin() + (LW[c].base - body_total));
    }

    // slots of the transform kernel
    std::vector<int32_t> src, body, dst, dsto, tag, cls;
    std::vector<float> sw;
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < sizes_[b]; ++k) {
            const int o = perm[b][k];
            sw.push_back(w_[body_off_[b] + o]);
            src.push_back(body_off_[b] + o);
            body.push_back(b);
            dst.push_back(L[1 + b].base + k);
            dsto.push_back(L[1 + b].base + o);
            tag.push_back(o);
            cls.push_back(point_cls_[body_off_[b] + o]);
        }
    p.n_slots = (int)src.size();

    // jobs: inter-body blocks (rows body j, columns body i, j < i) and body-water blocks
    std::vector<int32_t> aa_map((size_t)C_ * C_), aw_map((size_t)C_ * C_, -1);
    for (int a = 0; a < C_; ++a)
        for (int b = 0; b < C_; ++b) aa_map[a * C_ + b] = tri(a, b);
    for (size_t k = 0; k < water_classes_.size(); ++k)
        for (int b = 0; b < C_; ++b) aw_map[water_classes_[k] * C_ + b] = R_aa_ + (int)k * C_ + b;
    auto shifted = [&](std::vector<int32_t> m) {
        for (auto& r : m)
            if (r >= 0) r += n_conf * R;
        return m;
    };
    auto make_lists = [&](bool split, WorkList& shared, WorkList& tmpl) {
        for (int i = 0; i < B; ++i) {
            for (int j = 0; j < i; ++j) {
                const bool mv = moving[i] || moving[j];
                append_job(mv ? tmpl : shared, PairJob{1 + j, 1 + i, mv ? aa_map : shifted(aa_map)}, sets, L,
                           n_rows, tile_j, split);
            }
            if (!hy && waters_.size() > 0)
                append_job(moving[i] ? tmpl : shared, PairJob{0, 1 + i, moving[i] ? aw_map : shifted(aw_map)}, sets,
                           L, n_rows, tile_j, split);
        }
    };
    WorkList shared, tmpl, shared_s, tmpl_s;
    make_lists(false, shared, tmpl);
    make_lists(true, shared_s, tmpl_s);
    const int64_t work = padded_work(shared.items) + (int64_t)n_conf * padded_work(tmpl.items);
    const int64_t work_s = padded_work(shared_s.items) + (int64_t)n_conf * padded_work(tmpl_s.items);
    const int rows = std::max(item_rows(shared.items), item_rows(tmpl.items));
    const int rows_s = std::max(item_rows(shared_s.items), item_rows(tmpl_s.items));
    const bool use_split = prefer_split(work, rows, work_s, rows_s);
    if (use_split) {
        shared = std::move(shared_s);
        tmpl = std::move(tmpl_s);
    }
    p.pairs = shared.pairs + (int64_t)n_conf * tmpl.pairs;

    // hydration-modified batches: per conformation, every body against its
    // waters (calc_aw) and the waters' self-correlation (calc_ww), rows absolute
    WorkList hyd;
    std::vector<std::pair<int64_t, int64_t>> hyd_base;  // per strip of hyd: row and column set bases
    if (hy) {
        std::vector<PointSet> hs(1 + B);
        for (int b = 0; b < B; ++b) hs[1 + b] = sets[1 + b];
        std::vector<SetLayout> hl(1 + B);
        std::vector<int32_t> awm((size_t)C_ * C_), wwm((size_t)C_ * C_);
        for (int c = 0; c < n_conf; ++c) {
            for (int a = 0; a < C_; ++a)
                for (int b = 0; b < C_; ++b) {
                    awm[a * C_ + b] = c * R + R_aa_ + a * C_ + b;
                    wwm[a * C_ + b] = c * R + R_aa_ + C_ * C_ + tri(a, b);
                }
            hs[0] = hy->sets[c];
            hl[0] = LW[c];
            for (int b = 0; b < B; ++b) {
                hl[1 + b] = L[1 + b];
                if (moving[b]) shift_layout(hl[1 + b], (int64_t)c * p.M);
            }
            const size_t s0 = hyd.strips.size();
            if (hs[0].size() > 0)
                for (int i = 0; i < B; ++i) append_job(hyd, PairJob{0, 1 + i, awm}, hs, hl, n_rows, tile_j, use_split);
            append_job(hyd, PairJob{0, -1, wwm}, hs, hl, n_rows, tile_j, use_split);
            for (size_t k = s0; k < hyd.strips.size(); ++k)
                hyd_base.emplace_back(hl[hyd.strips[k].set_rows].base, hl[hyd.strips[k].set_cols].base);
        }
        p.pairs += hyd.pairs;
    }

    // items: shared ones, then each template item for every conformation
    sort_items(shared.items);
    sort_items(tmpl.items);
    p.items = shared.items;
    auto shift = [&](int32_t v, int c) { return v >= p.S ? (int32_t)(v + (int64_t)c * p.M) : v; };
    for (const PairWork& t : tmpl.items)
        for (int c = 0; c < n_conf; ++c) {
            PairWork w = t;
            w.i_begin = shift(t.i_begin, c);
            w.i_end = t.i_begin >= p.S ? w.i_begin + (t.i_end - t.i_begin) : t.i_end;
            w.j_begin = shift(t.j_begin, c);
            w.j_end = t.j_begin >= p.S ? w.j_begin + (t.j_end - t.j_begin) : t.j_end;
            w.row0 += c * R;
            w.row1 += c * R;
            p.items.push_back(w);
        }
    p.items.insert(p.items.end(), hyd.items.begin(), hyd.items.end());
    if (hy) sort_items(p.items);
    p.mp = plan_merge(p.items, n_rows);
    p.rows = item_rows(p.items);
    p.copies = pick_copies(nbins, p.rows);

    // device residency
    Context& ctx = Context::get();
    hipStream_t s = ctx.stream();
    p.pts.ensure(total);
    p.orig.ensure(total);
    if (hy && !wbuf.empty()) {
        check(hipMemcpyAsync(p.pts.p + body_total, wbuf.data(), wbuf.size() * sizeof(float4), hipMemcpyHostToDevice,
                             s),
              "H2D waters");
        check(hipMemcpyAsync(p.orig.p + body_total, worig.data(), worig.size() * sizeof(float4),
                             hipMemcpyHostToDevice, s),
              "H2D waters");
    }
    if (!hy && waters_.size()) {
        check(hipMemcpyAsync(p.pts.p, buf.data(), waters_.size() * sizeof(float4), hipMemcpyHostToDevice, s),
              "H2D waters");
        check(hipMemcpyAsync(p.orig.p, waters_.orig.data(), waters_.size() * sizeof(float4), hipMemcpyHostToDevice,
                             s),
              "H2D waters");
    }
    p.d_items.upload(p.items.data(), p.items.size(), s);
    p.row_list.upload(p.mp.row_list.data(), p.mp.row_list.size(), s);
    p.chunk_ptr.upload(p.mp.chunk_ptr.data(), p.mp.chunk_ptr.size(), s);
    p.row_chunk.upload(p.mp.row_chunk.data(), p.mp.row_chunk.size(), s);
    std::vector<int32_t> rowmaps = tmpl.rowmaps;
    rowmaps.insert(rowmaps.end(), shared.rowmaps.begin(), shared.rowmaps.end());
    const int hyd_rowmap_off = (int)rowmaps.size();
    rowmaps.insert(rowmaps.end(), hyd.rowmaps.begin(), hyd.rowmaps.end());
    p.rowmaps.upload(rowmaps.data(), rowmaps.size(), s);
    p.partial.ensure(std::max<size_t>(p.items.size(), 1) * 2 * (nbins + 1));
    p.tmp_cnt.ensure((size_t)p.mp.n_chunks() * (nbins + 1));
    p.tmp_dsum.ensure((size_t)p.mp.n_chunks() * (nbins + 1));
    p.cnt.ensure((size_t)n_rows * (nbins + 1));
    p.dsum.ensure((size_t)n_rows * (nbins + 1));
    p.slot_src.upload(src.data(), src.size(), s);
    p.slot_body.upload(body.data(), body.size(), s);
    p.slot_dst.upload(dst.data(), dst.size(), s);
    p.slot_dsto.upload(dsto.data(), dsto.size(), s);
    p.slot_tag.upload(tag.data(), tag.size(), s);
    p.slot_cls.upload(cls.data(), cls.size(), s);
    p.body_moving.upload(moving.data(), moving.size(), s);
    if (weighted_) {
        // the weight pass's point copies: waters here, bodies by the transform kernel
        p.pts_w.ensure(total);
        p.orig_w.ensure(total);
        p.slot_w.upload(sw.data(), sw.size(), s);
        p.wcnt.ensure((size_t)n_rows * (nbins + 1));
        p.wsum.ensure((size_t)n_rows * (nbins + 1));
        p.wstrip.ensure((size_t)n_rows * (nbins + 1));
        std::vector<float4> hw_pts, hw_orig;
        if (hy && !wbuf.empty()) {
            hw_pts = wbuf;
            hw_orig = worig;
            for (int c = 0; c < n_conf; ++c) {
                const int64_t b0 = LW[c].base - body_total;
                for (int i = 0; i < hy->sets[c].size(); ++i) {
                    hw_pts[b0 + i].w = hy->w[c][wperm_h[b0 + i]];
                    hw_orig[b0 + i].w = hy->w[c][i];
                }
            }
            check(hipMemcpyAsync(p.pts_w.p + body_total, hw_pts.data(), hw_pts.size() * sizeof(float4),
                                 hipMemcpyHostToDevice, s),
                  "H2D waters");
            check(hipMemcpyAsync(p.orig_w.p + body_total, hw_orig.data(), hw_orig.size() * sizeof(float4),
                                 hipMemcpyHostToDevice, s),
                  "H2D waters");
            ctx.sync();  // the host vectors go out of scope
        }
        const int nw = hy ? 0 : waters_.size();
        if (nw) {
            std::vector<float4> a(buf.begin(), buf.begin() + nw), o(waters_.orig);
            for (int i = 0; i < nw; ++i) a[i].w = water_w_[wperm[i]], o[i].w = water_w_[i];
            check(hipMemcpyAsync(p.pts_w.p, a.data(), nw * sizeof(float4), hipMemcpyHostToDevice, s), "H2D waters");
            check(hipMemcpyAsync(p.orig_w.p, o.data(), nw * sizeof(float4), hipMemcpyHostToDevice, s), "H2D waters");
            ctx.sync();  // the host vectors go out of scope
        }
    }
    p.T.ensure((size_t)n_conf * B * 12);
    p.hT.ensure((size_t)n_conf * B * 12);
    p.H.ensure((size_t)n_conf * nbins);
    p.dbar.ensure((size_t)n_conf * nbins);
    p.max_bin.ensure(n_conf);
    p.flags.ensure(n_conf);
    p.hflags.ensure(n_conf);

    // strips: template strips once per conformation (replica strides), shared ones once
    auto strip = [&](const StripSpec& st, int rowmap_off, bool replicated) {
        const bool rm = st.set_rows > 0 && moving[st.set_rows - 1];
        const bool cm = st.set_cols > 0 && moving[st.set_cols - 1];
        StripLaunch sl{p.orig.p + L[st.set_rows].base, p.orig.p + L[st.set_cols].base, st.w,
                       p.rowmaps.p + rowmap_off, p.cnt.p, p.dsum.p, inv_w_, nbins, nbins + 1};
        sl.n_rep = replicated ? n_conf : 1;
        sl.rows_rep = replicated && rm ? p.M : 0;
        sl.cols_rep = replicated && cm ? p.M : 0;
        sl.out_rep = replicated ? R : 0;
        p.strips.push_back(sl);
    };
    for (const StripSpec& st : tmpl.strips) strip(st, st.rowmap_off, true);
    for (const StripSpec& st : shared.strips) strip(st, (int)tmpl.rowmaps.size() + st.rowmap_off, false);
    // the water jobs' strips: one table, launched at once
    std::vector<StripLaunch> table, wtable;
    std::vector<int64_t> first;
    for (size_t k = 0; k < hyd.strips.size(); ++k) {
        const StripSpec& st = hyd.strips[k];
        const int64_t rb = hyd_base[k].first, cb = hyd_base[k].second;
        StripLaunch sl{p.orig.p + rb, p.orig.p + cb, st.w, p.rowmaps.p + hyd_rowmap_off + st.rowmap_off, p.cnt.p,
                       p.dsum.p, inv_w_, nbins, nbins + 1};
        const int64_t nb = strip_table_blocks(sl);
        if (!nb) continue;
        // bounds of every strip, checked here before any launch
        if (rb < 0 || rb + st.w.n_rows > total || cb + st.w.t_begin < 0 || cb + st.w.t_end > total)
            throw std::runtime_error("ausaxs_hip: rigid-body plan out of bounds: water strip");
        const size_t off = (size_t)hyd_rowmap_off + st.rowmap_off, nmap = (size_t)C_ * st.w.n_col_classes;
        if (off + nmap > rowmaps.size())
            throw std::runtime_error("ausaxs_hip: rigid-body plan out of bounds: water strip rowmap");
        for (size_t e = 0; e < nmap; ++e)
            if (rowmaps[off + e] >= n_rows)
                throw std::runtime_error("ausaxs_hip: rigid-body plan out of bounds: water strip row");
        first.push_back(p.table_blocks);
        p.table_blocks += nb;
        table.push_back(sl);
        if (weighted_) {
            sl.rows = p.orig_w.p + rb;
            sl.cols = p.orig_w.p + cb;
            sl.cnt = p.wcnt.p;
            sl.dsum = p.wstrip.p;
            sl.wmode = 1;
            sl.wscale = wscale_;
            wtable.push_back(sl);
        }
    }
    if (p.table_blocks > 0x7fffffff) throw std::runtime_error("ausaxs_hip: too many water strip pairs in one batch");
    p.n_table = (int)table.size();
    if (p.n_table) {
        first.push_back(p.table_blocks);
        p.d_table.upload(table.data(), table.size(), s);
        p.d_first.upload(first.data(), first.size(), s);
        if (weighted_) p.d_wtable.upload(wtable.data(), wtable.size(), s);
    }
    if (hy) {
        // row kinds and weights of this batch's rows
        p.row_kind.assign(R, 0);
        p.row_w.assign(R, 0.0);
        for (int a = 0; a < C_; ++a)
            for (int b = 0; b < C_; ++b) {
                const double w = (double)(class_w_[a] * class_w_[b]);
                p.row_w[tri(a, b)] = w;
                p.row_kind[R_aa_ + a * C_ + b] = 1;
                p.row_w[R_aa_ + a * C_ + b] = w;
                p.row_kind[R_aa_ + C_ * C_ + tri(a, b)] = 2;
                p.row_w[R_aa_ + C_ * C_ + tri(a, b)] = w;
            }
        p.d_row_kind.upload(p.row_kind.data(), p.row_kind.size(), s);
        p.d_row_w.upload(p.row_w.data(), p.row_w.size(), s);
        p.ww_self.assign(2 * (size_t)n_conf, 0.0);
        for (int c = 0; c < n_conf; ++c) {
            double t = 0;  // SimpleCalculator.h:144-160: sum of w^2 at bin 0
            for (float v : hy->w[c]) t += (double)(v * v);
            p.ww_self[2 * c] = t;
            p.ww_self[2 * c + 1] = (double)(uint32_t)t;
        }
        p.d_ww_self.upload(p.ww_self.data(), p.ww_self.size(), s);
    }
    if (weighted_)
        for (StripLaunch sl : p.strips) {
            // the same strip on the weight copies, into the weight sums (one row per job)
            sl.rows = p.orig_w.p + (sl.rows - p.orig.p);
            sl.cols = p.orig_w.p + (sl.cols - p.orig.p);
            sl.cnt = p.wcnt.p;
            sl.dsum = p.wstrip.p;
            sl.wmode = 1;
            sl.wscale = wscale_;
            p.wstrips.push_back(sl);
        }

    // every index the kernels will touch, checked on the host before any launch
    auto fail = [](const char* what) {
        throw std::runtime_error(std::string("ausaxs_hip: rigid-body plan out of bounds: ") + what);
    };
    for (const PairWork& w : p.items) {
        if (w.i_begin < 0 || w.i_begin >= w.i_end || w.i_end > total || w.j_begin < 0 || w.j_begin >= w.j_end ||
            w.j_end > total)
            fail("item point range");
        if (w.row0 < 0 || w.row0 >= n_rows || w.row1 < 0 || w.row1 >= n_rows) fail("item row");
        if ((int64_t)(w.i_end - w.i_begin) * (w.j_end - w.j_begin) > kMaxPairsPerItem) fail("item size");
    }
    for (size_t k = 0; k < p.strips.size(); ++k) {
        const StripLaunch& sl = p.strips[k];
        const StripSpec& st = k < tmpl.strips.size() ? tmpl.strips[k] : shared.strips[k - tmpl.strips.size()];
        const int64_t rb = L[st.set_rows].base, cb = L[st.set_cols].base;
        const int64_t last = sl.n_rep - 1;
        if (rb < 0 || rb + last * sl.rows_rep + sl.w.n_rows > total) fail("strip rows");
        if (cb + sl.w.t_begin < 0 || cb + last * sl.cols_rep + sl.w.t_end > total) fail("strip columns");
        const int off = (int)(sl.rowmap - p.rowmaps.p);
        const size_t nmap = (size_t)C_ * sl.w.n_col_classes;
        if (off < 0 || off + nmap > rowmaps.size()) fail("strip rowmap");
        for (size_t e = 0; e < nmap; ++e)
            if (rowmaps[off + e] >= 0 && rowmaps[off + e] + last * sl.out_rep >= n_rows) fail("strip output row");
    }
    for (int k = 0; k < p.n_slots; ++k) {
        const int64_t reach = moving[body[k]] ? (int64_t)(n_conf - 1) * p.M : 0;
        const int64_t lim = moving[body[k]] ? body_total : p.S;
        if (dst[k] < 0 || dst[k] + reach >= lim || dsto[k] < 0 || dsto[k] + reach >= lim) fail("transform slot");
        if (src[k] < 0 || src[k] >= body_off_.back()) fail("transform source");
    }
    ctx.sync();
    return P;
}

void RigidScan::enqueue(Plan& p, int nq, double cw, bool copies) {
    Context& ctx = Context::get();
    hipStream_t s = ctx.stream();
    const int B = n_bodies();
    if (copies) {
        check(hipMemcpyAsync(p.T.p, p.hT.p, (size_t)p.n_conf * B * 12 * sizeof(double), hipMemcpyHostToDevice, s),
              "H2D transforms");
        if (nq > 0)
            check(hipMemcpyAsync(p.q.p, p.hq.p, (size_t)nq * sizeof(double), hipMemcpyHostToDevice, s), "H2D q");
    }
    TransformLaunch tl{d_xyz_.p, p.T.p, p.slot_src.p, p.slot_body.p, p.slot_dst.p, p.slot_dsto.p, p.slot_tag.p,
                       p.slot_cls.p, p.body_moving.p, p.n_slots, B, p.n_conf, p.M, p.pts.p, p.orig.p};
    if (weighted_) tl.slot_w = p.slot_w.p, tl.pts_w = p.pts_w.p, tl.orig_w = p.orig_w.p;
    check(launch_transform(tl, s), "transform_kernel");
    PairLaunch pl{p.pts.p, nullptr, p.d_items.p, (int)p.items.size(), p.partial.p, inv_w_, p.nbins, 1, p.grid, p.rows,
                  p.copies};
    check(launch_pair_hist(pl, s), "pair_hist_kernel");
    MergeLaunch ml{p.partial.p, p.row_list.p, p.chunk_ptr.p, p.row_chunk.p, p.mp.n_chunks(), (p.n_conf + 1) * p.R,
                   p.nbins, p.tmp_cnt.p, p.tmp_dsum.p, p.cnt.p, p.dsum.p};
    check(launch_merge(ml, s), "merge kernels");
    for (const StripLaunch& sl : p.strips) check(launch_strip(sl, s), "strip_kernel");
    check(launch_strip_table(p.d_table.p, p.d_first.p, p.n_table, p.table_blocks, s), "strip_table_kernel");
    if (weighted_) {
        PairLaunch wl{p.pts_w.p, nullptr, p.d_items.p, (int)p.items.size(), p.partial.p, inv_w_, p.nbins, 1, p.grid, p.rows,
                      p.copies, 1, wscale_};
        check(launch_pair_hist(wl, s), "pair_hist_kernel (weights)");
        MergeLaunch wm{p.partial.p, p.row_list.p, p.chunk_ptr.p, p.row_chunk.p, p.mp.n_chunks(), (p.n_conf + 1) * p.R,
                       p.nbins, p.tmp_cnt.p, p.tmp_dsum.p, p.wcnt.p, p.wsum.p, 1};
        check(launch_merge(wm, s), "merge kernels (weights)");
        check(launch_lattice_zero_u64((unsigned long long*)p.wstrip.p, (size_t)(p.n_conf + 1) * p.R * (p.nbins + 1), s),
              "lattice_zero_u64_kernel");
        for (const StripLaunch& sl : p.wstrips) check(launch_strip(sl, s), "strip_kernel (weights)");
        check(launch_strip_table(p.d_wtable.p, p.d_first.p, p.n_table, p.table_blocks, s),
              "strip_table_kernel (weights)");
    }
    FinalizeLaunch fl{p.cnt.p, p.dsum.p, p.R, p.nbins, p.n_conf,
                      p.hydrated ? p.d_row_kind.p : d_row_kind_.p, p.hydrated ? p.d_row_w.p : d_row_w_.p,
                      p.hydrated ? d_base_aa_.p : d_base_.p, base_nb_,
                      (double)inv_w_, (s_.bin_count * s_.bin_width) / s_.bin_count, cw, s_.weighted_bins ? 1 : 0,
                      p.H.p, p.dbar.p, p.max_bin.p, p.flags.p};
    if (weighted_) fl.wsum = p.wsum.p, fl.wstrip = p.wstrip.p, fl.wscale = wscale_;
    if (p.hydrated) fl.ww_self = p.d_ww_self.p;
    check(launch_finalize(fl, s), "finalize_kernel");
    DebyeBatchLaunch dl{p.H.p, p.dbar.p, p.max_bin.p, p.q.p, p.nbins, p.n_conf, nq, p.I.p};
    check(launch_debye_batch(dl, s), "debye_batch_kernel");
    if (copies) {
        if (nq > 0)
            check(hipMemcpyAsync(p.hI.p, p.I.p, (size_t)p.n_conf * nq * sizeof(double), hipMemcpyDeviceToHost, s),
                  "D2H I");
        check(hipMemcpyAsync(p.hflags.p, p.flags.p, (size_t)p.n_conf * sizeof(int), hipMemcpyDeviceToHost, s),
              "D2H flags");
    }
}

float RigidScan::time_pair(Plan& p, int grid, int copies) {
    Context& ctx = Context::get();
    PairLaunch pl{p.pts.p, nullptr, p.d_items.p, (int)p.items.size(), p.partial.p, inv_w_, p.nbins, 1, grid, p.rows, copies};
    check(launch_pair_hist(pl, ctx.stream()), "pair_hist_kernel");  // warm-up
    check(hipEventRecord(p.ev[0], ctx.stream()), "hipEventRecord");
    check(launch_pair_hist(pl, ctx.stream()), "pair_hist_kernel");
    check(hipEventRecord(p.ev[1], ctx.stream()), "hipEventRecord");
    check(hipEventSynchronize(p.ev[1]), "hipEventSynchronize");
    float ms = 0;
    check(hipEventElapsedTime(&ms, p.ev[0], p.ev[1]), "hipEventElapsedTime");
    return ms;
}

RigidScan::Plan& RigidScan::plan_for(int n_conf, const std::vector<int>& moving, int nbins, const double* T) {
    auto key = std::make_tuple(n_conf, moving, nbins);
    auto it = plans_.find(key);
    if (it != plans_.end()) return *it->second;
    Context& ctx = Context::get();
    std::unique_ptr<Plan> best;
    if (s_.tile_j != 0) {
        best = build_plan(n_conf, moving, nbins, s_.tile_j);
    } else {
        // autotuner: the tile width first (one workgroup per item), then the
        // persistent grid of the winner, each timed on this batch's transforms
        const int cus = ctx.prop.multiProcessorCount;
        float best_ms = std::numeric_limits<float>::max();
        int cand = 0;
        for (int tile : {1024, 2048, 4096, 8000}) {
            auto p = build_plan(n_conf, moving, nbins, tile);
            check(hipMemcpyAsync(p->T.p, T, (size_t)n_conf * n_bodies() * 12 * sizeof(double),
                                 hipMemcpyHostToDevice, ctx.stream()),
                  "H2D transforms");
            TransformLaunch tl{d_xyz_.p, p->T.p, p->slot_src.p, p->slot_body.p, p->slot_dst.p, p->slot_dsto.p,
                               p->slot_tag.p, p->slot_cls.p, p->body_moving.p, p->n_slots, n_bodies(), n_conf,
                               p->M, p->pts.p, p->orig.p};
            check(launch_transform(tl, ctx.stream()), "transform_kernel");
            const float ms = time_pair(*p, 0, p->copies);
            ++cand;
            if (ms < best_ms) {
                best_ms = ms;
                best = std::move(p);
            }
        }
        // LDS histogram copies (up to 64 KB of LDS per workgroup), then the grid
        for (int k : {1, 2, 4, 8}) {
            if (k == best->copies || pair_lds_bytes(nbins, best->rows, k) > 65536) continue;
            const float ms = time_pair(*best, 0, k);
            ++cand;
            if (ms < best_ms) {
                best_ms = ms;
                best->copies = k;
            }
        }
        for (int g : {2 * cus, 4 * cus, 8 * cus}) {
            if (g >= (int)best->items.size()) continue;
            const float ms = time_pair(*best, g, best->copies);
            ++cand;
            if (ms < best_ms) {
                best_ms = ms;
                best->grid = g;
            }
        }
        info_.pair_ms = best_ms;
        info_.tuned_candidates = cand;
    }
    Plan& ref = *best;
    plans_.emplace(key, std::move(best));
    return ref;
}

void RigidScan::run_batch(Plan& p, const double* T, const double* q, int nq, double cw, double* I) {
    Context& ctx = Context::get();
    hipStream_t s = ctx.stream();
    const int B = n_bodies();
    std::memcpy(p.hT.p, T, (size_t)p.n_conf * B * 12 * sizeof(double));
    if (nq > p.nq_cap || !p.q.p) {
        p.drop_graph();
        p.q.ensure(nq);
        p.hq.ensure(nq);
        p.I.ensure((size_t)p.n_conf * nq);
        p.hI.ensure((size_t)p.n_conf * nq);
        p.nq_cap = nq;
    }
    std::memcpy(p.hq.p, q, (size_t)nq * sizeof(double));
    if (p.hydrated) {
        // a plan of one call: launched directly, no graph to capture
        check(hipEventRecord(p.ev[2], s), "hipEventRecord");
        enqueue(p, nq, cw, true);
        check(hipEventRecord(p.ev[3], s), "hipEventRecord");
    } else if (!p.exec || p.graph_nq != nq || p.graph_cw != cw) {
        p.drop_graph();
        // the whole batch as one graph: copies in, kernels, copies out (no
        // memset nodes: every output the graph reads back is written by a kernel)
        check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            enqueue(p, nq, cw, true);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(s, &p.graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        p.graph_nq = nq;
        p.graph_cw = cw;
    }
    if (!p.hydrated) {
        check(hipEventRecord(p.ev[2], s), "hipEventRecord");
        check(hipGraphLaunch(p.exec, s), "hipGraphLaunch");
        check(hipEventRecord(p.ev[3], s), "hipEventRecord");
    }
    ctx.sync();
    float ms = 0;
    check(hipEventElapsedTime(&ms, p.ev[2], p.ev[3]), "hipEventElapsedTime");
    info_.total_ms += ms;
    for (int c = 0; c < p.n_conf; ++c)
        if (p.hflags.p[c])
            throw std::runtime_error("ausaxs_hip: a pair distance falls beyond bin_count*bin_width (" +
                                     std::to_string(s_.bin_count) + " bins of " + std::to_string(s_.bin_width) + " A)");
    std::memcpy(I, p.hI.p, (size_t)p.n_conf * nq * sizeof(double));
}

std::vector<int> RigidScan::moving_bodies(const double* T, int n_conf) const {
    // bodies whose transform is not the same in every conformation of the batch
    const int B = n_bodies();
    std::vector<int> moving(B, 0);
    for (int b = 0; b < B; ++b)
        for (int c = 1; c < n_conf && !moving[b]; ++c)
            moving[b] = std::memcmp(T + ((size_t)c * B + b) * 12, T + (size_t)b * 12, 12 * sizeof(double)) != 0;
    return moving;
}

void RigidScan::record(Plan& p, int c0, int nc) {
    info_.tile_j = p.tile_j;
    info_.grid = p.grid;
    info_.copies = p.copies;
    info_.n_items = (int)p.items.size();
    info_.batch = nc;
    info_.pairs += p.pairs;
    last_ = &p;
    last_conf0_ = c0;
    last_n_conf_ = nc;
}

void RigidScan::scan(const double* T, int n_conf, const double* q, int nq, double cw, double* I) {
    if (n_conf < 0 || nq < 0) throw std::runtime_error("ausaxs_hip: negative conformation or q count");
    const int B = n_bodies();
    info_.total_ms = 0;
    info_.pairs = 0;
    for (int c0 = 0; c0 < n_conf; c0 += kMaxBatch) {
        const int nc = std::min(kMaxBatch, n_conf - c0);
        const double* Tc = T + (size_t)c0 * B * 12;
        const std::vector<int> moving = moving_bodies(Tc, nc);
        const int nbins = bound_bins(Tc, nc);
        Plan& p = plan_for(nc, moving, nbins, Tc);
        run_batch(p, Tc, q, nq, cw, I + (size_t)c0 * nq);
        record(p, c0, nc);
    }
}

void RigidScan::scan_hydrated(const double* T, int n_conf, const double* wxyz, const float* ww,
                              const int32_t* counts, const double* q, int nq, double cw, double* I) {
    if (n_conf < 0 || nq < 0) throw std::runtime_error("ausaxs_hip: negative conformation or q count");
    if (n_conf > 0 && !counts) throw std::runtime_error("ausaxs_hip: NULL water counts");
    int64_t n_w = 0;
    for (int c = 0; c < n_conf; ++c) {
        if (counts[c] < 0) throw std::runtime_error("ausaxs_hip: negative water count");
        n_w += counts[c];
    }
    if (n_w > 0 && (!wxyz || !ww)) throw std::runtime_error("ausaxs_hip: NULL waters");
    for (int64_t i = 0; i < 3 * n_w; ++i)
        if (!std::isfinite(wxyz[i])) throw std::runtime_error("ausaxs_hip: water coordinates must be finite");
    for (int64_t i = 0; i < n_w; ++i)
        if (!std::isfinite(ww[i])) throw std::runtime_error("ausaxs_hip: weights must be finite");
    const int B = n_bodies();
    // weight classes of the waters: those of the scan (unless it runs a weight pass)
    std::map<uint32_t, int> cls;
    for (int c = 0; c < (int)class_w_.size(); ++c) {
        uint32_t key;
        std::memcpy(&key, &class_w_[c], 4);
        cls.emplace(key, c);
    }
    double wmax = 0;
    for (int64_t i = 0; i < n_w; ++i) wmax = std::max(wmax, (double)std::fabs(ww[i]));
    if (weighted_ && weight_scale(wmax) < wscale_)
        throw std::runtime_error("ausaxs_hip: a water weight is larger in magnitude than every weight given at "
                                 "creation");
    if (!d_base_aa_.p) {
        // the cached self terms of the bodies only: the creation waters' ww is replaced
        std::vector<double> base(6 * (size_t)base_nb_, 0.0);
        for (int b = 0; b < base_nb_; ++b) {
            base[b] = base_aa_[b].value;
            base[base_nb_ + b] = base_aa_[b].count;
            base[2 * base_nb_ + b] = base_aa_[b].bin_center;
        }
        Context& ctx = Context::get();
        d_base_aa_.upload(base.data(), base.size(), ctx.stream());
        ctx.sync();
    }
    info_.total_ms = 0;
    info_.pairs = 0;
    int64_t off = 0;
    for (int c0 = 0; c0 < n_conf; c0 += kMaxBatch) {
        const int nc = std::min(kMaxBatch, n_conf - c0);
        const double* Tc = T + (size_t)c0 * B * 12;
        Hydration hy;
        hy.sets.resize(nc);
        hy.w.resize(nc);
        hy.sphere.assign(4 * (size_t)nc, 0.0);
        for (int c = 0; c < nc; ++c) {
            const int m = counts[c0 + c];
            PointSet& ps = hy.sets[c];
            ps.n_classes = C_;
            ps.orig.resize(m);
            hy.w[c].assign(ww + off, ww + off + m);
            std::vector<double> pd(3 * (size_t)m);
            for (int i = 0; i < m; ++i) {
                const double* x = wxyz + 3 * (off + i);
                int k = 0;
                if (!weighted_) {
                    uint32_t key;
                    std::memcpy(&key, &ww[off + i], 4);
                    auto it = cls.find(key);
                    if (it == cls.end())
                        throw std::runtime_error("ausaxs_hip: a water weight that no point of the scan's creation "
                                                 "has (its weight classes are fixed then)");
                    k = it->second;
                }
                // CompactCoordinates: float cast of the FP64 coordinates
                ps.orig[i] = make_float4((float)x[0], (float)x[1], (float)x[2], float_of(k));
                for (int e = 0; e < 3; ++e) pd[3 * i + e] = (&ps.orig[i].x)[e];
            }
            sphere_of(pd.data(), m, 3, &hy.sphere[4 * (size_t)c], hy.sphere[4 * (size_t)c + 3]);
            off += m;
        }
        const std::vector<int> moving = moving_bodies(Tc, nc);
        const int nbins = bound_bins(Tc, nc, &hy);
        // the tile width of the scan's tuned plans, else the default
        const int tile = s_.tile_j ? s_.tile_j : (info_.tile_j ? info_.tile_j : kDefaultTileJ);
        if (last_ == hplan_.get()) last_ = nullptr;
        hplan_.reset();
        hplan_ = build_plan(nc, moving, nbins, tile, &hy);
        run_batch(*hplan_, Tc, q, nq, cw, I + (size_t)c0 * nq);
        record(*hplan_, c0, nc);
    }
}

std::unique_ptr<CompositeDistanceHistogramHIP> RigidScan::histogram(int c) {
    if (!last_ || c < last_conf0_ || c >= last_conf0_ + last_n_conf_)
        throw std::runtime_error("ausaxs_hip: conformation " + std::to_string(c) +
                                 " is not in the last batch of the last scan");
    Plan& p = *last_;
    Context& ctx = Context::get();
    const int stride = p.nbins + 1;
    const int lc = c - last_conf0_;
    const int R = p.R;
    const std::vector<int32_t>& row_kind = p.hydrated ? p.row_kind : row_kind_;
    const std::vector<double>& row_w = p.hydrated ? p.row_w : row_w_;
    std::vector<uint64_t> cnt((size_t)2 * R * stride);
    std::vector<int64_t> dsum((size_t)2 * R * stride), wsum(weighted_ ? (size_t)2 * R * stride : 0),
        wstrip(wsum.size());
    const size_t rows_bytes = (size_t)R * stride;
    for (int part = 0; part < 2; ++part) {
        const size_t row0 = part == 0 ? (size_t)lc * R : (size_t)p.n_conf * R;
        check(hipMemcpyAsync(cnt.data() + part * rows_bytes, p.cnt.p + row0 * stride, rows_bytes * sizeof(uint64_t),
                             hipMemcpyDeviceToHost, ctx.stream()),
              "D2H");
        check(hipMemcpyAsync(dsum.data() + part * rows_bytes, p.dsum.p + row0 * stride,
                             rows_bytes * sizeof(int64_t), hipMemcpyDeviceToHost, ctx.stream()),
              "D2H");
        if (weighted_) {
            check(hipMemcpyAsync(wsum.data() + part * rows_bytes, p.wsum.p + row0 * stride,
                                 rows_bytes * sizeof(int64_t), hipMemcpyDeviceToHost, ctx.stream()),
                  "D2H");
            check(hipMemcpyAsync(wstrip.data() + part * rows_bytes, p.wstrip.p + row0 * stride,
                                 rows_bytes * sizeof(int64_t), hipMemcpyDeviceToHost, ctx.stream()),
                  "D2H");
        }
    }
    ctx.sync();
    const int nb = p.nbins;
    const double unit = delta_unit(inv_w_);
    std::vector<WeightedEntry> aa(nb), aw(nb), ww(nb);
    for (int b = 0; b < std::min(nb, base_nb_); ++b) {
        aa[b] = base_aa_[b];
        if (!p.hydrated) ww[b] = base_ww_[b];
    }
    for (int part = 0; part < 2; ++part)
        for (int r = 0; r < R; ++r)
            for (int b = 0; b < nb; ++b) {
                const size_t cell = (size_t)(part * R + r) * stride + b;
                const uint64_t n = cnt[cell];
                if (!n) continue;
                WeightedEntry e;  // add<2> (SimpleCalculator.h:195-207)
                e.value = weighted_ ? 2.0 * weight_value(wsum[cell], wstrip[cell], wscale_) : 2.0 * (double)n * row_w[r];
                e.count = (uint32_t)(2 * n);
                e.bin_center = 2.0 * distance_sum_u(n, b, dsum[cell], inv_w_, unit);
                entry_add(row_kind[r] == 0 ? aa[b] : row_kind[r] == 1 ? aw[b] : ww[b], e);
            }
    if (p.hydrated)  // this conformation's water set's self term (SimpleCalculator.h:144-160)
        entry_add(ww[0], WeightedEntry{p.ww_self[2 * lc], (uint32_t)p.ww_self[2 * lc + 1], 0});
    std::vector<WeightedEntry> ptot(nb);
    for (int b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(ptot[b], aw[b]);
        entry_add(ptot[b], ww[b]);
    }
    auto res = std::make_unique<CompositeDistanceHistogramHIP>();
    res->mode = Mode::Simple;
    res->bin_width = s_.bin_width;
    res->weighted_bins = s_.weighted_bins;
    res->debye_terms = s_.debye_terms;
    const unsigned mb = std::min<unsigned>(max_bin_of(ptot), (unsigned)nb);
    res->max_bin = mb;
    res->aa.reset(1, 1, (int)mb);
    res->aw.reset(1, 1, (int)mb);
    std::copy_n(aa.begin(), mb, res->aa.touch(0, 0));
    std::copy_n(aw.begin(), mb, res->aw.touch(0, 0));
    res->ww.assign(ww.begin(), ww.begin() + mb);
    res->p_tot.assign(ptot.begin(), ptot.begin() + mb);
    res->d_axis = axis_of(res->p_tot, s_.bin_width, s_.weighted_bins);
    return res;
}

}  // namespace ausaxs::gpu
