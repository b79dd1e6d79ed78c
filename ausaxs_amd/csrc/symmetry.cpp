// Symmetry histograms (see symmetry.hpp).
#include "symmetry.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>

namespace ausaxs::gpu {

std::vector<CopyPair> chain_pair_schedule(int reps, bool closed) {
    if (reps < 1) throw std::runtime_error("ausaxs_hip: a symmetry needs at least one repetition");
    // Copies form a chain one generator step apart, so a pair's distances
    // depend only on the index separation s: (reps + 1 - s) pairs at s; in a
    // closed chain separations s and reps + 1 - s coincide.
    std::vector<CopyPair> out;
    if (closed && reps == 1) return {CopyPair{0, 1, 1}};
    for (int k = 0; k < reps - (closed ? 1 : 0); ++k) out.push_back(CopyPair{0, k + 1, reps - k + (k == 0 && closed)});
    return out;
}

namespace {
using Mat = std::array<double, 9>;
using Vec = std::array<double, 3>;

Mat mat_mul(const Mat& A, const Mat& B) {  // Matrix.h:189-205 (sum from 0 in inner order)
    Mat C{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[3 * r + k] * B[3 * k + c];
            C[3 * r + c] = s;
        }
    return C;
}
Mat transpose(const Mat& A) {
    Mat T{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) T[3 * c + r] = A[3 * r + c];
    return T;
}
Vec mat_vec(const Mat& M, const Vec& v) {  // Vector3.h:169-181
    Vec o{};
    for (int r = 0; r < 3; ++r) o[r] = M[3 * r] * v[0] + M[3 * r + 1] * v[1] + M[3 * r + 2] * v[2];
    return o;
}
using Key = std::array<long long, 12>;
Key make_key(const Mat& R, const Vec& t) {  // PairScheduleUtil.cpp:12-31 (resolution 1e-4)
    Key k{};
    for (int i = 0; i < 9; ++i) k[i] = std::llround(R[i] * 1e4);
    for (int i = 0; i < 3; ++i) k[9 + i] = std::llround(t[i] * 1e4);
    return k;
}
}  // namespace

std::vector<CopyPair> placement_pair_schedule(const std::vector<double>& placements) {
    if (placements.size() % 12) throw std::runtime_error("ausaxs_hip: placements must hold 12 doubles each");
    const int n = (int)(placements.size() / 12);
    auto R_of = [&](int i) {
        Mat R;
        std::copy_n(placements.data() + 12 * (size_t)i, 9, R.begin());
        return R;
    };
    auto t_of = [&](int i) {
        Vec t;
        std::copy_n(placements.data() + 12 * (size_t)i + 9, 3, t.begin());
        return t;
    };
    std::map<Key, CopyPair> buckets;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            // relative map D = placement_j^-1 o placement_i (rigid: inverse = transpose)
            const Mat RjT = transpose(R_of(j));
            const Mat DR = mat_mul(RjT, R_of(i));
            const Vec ti = t_of(i), tj = t_of(j);
            const Vec DT = mat_vec(RjT, Vec{ti[0] - tj[0], ti[1] - tj[1], ti[2] - tj[2]});
            // D and D^-1 describe the same pair of bodies
            const Mat DiR = transpose(DR);
            const Vec m = mat_vec(DiR, DT);
            const Key key = std::min(make_key(DR, DT), make_key(DiR, Vec{-m[0], -m[1], -m[2]}));
            auto [it, fresh] = buckets.try_emplace(key, CopyPair{i, j, 1});
            if (!fresh) ++it->second.scale;
        }
    std::vector<CopyPair> out;
    for (const auto& [k, p] : buckets) out.push_back(p);
    return out;
}

std::vector<double> cyclic_transforms(const double* cm, const double* offset, const double* step, const double* axis,
                                      double angle, int reps, bool* closed) {
    if (reps < 1) throw std::runtime_error("ausaxs_hip: a symmetry needs at least one repetition");
    auto norm = [](Vec v) {
        const double m = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        return Vec{v[0] / m, v[1] / m, v[2] / m};
    };
    // CyclicSymmetry::_make_transform normalises the axis, and
    // rotation_matrix (MatrixUtils.cpp:54-74, Euler-Rodrigues) once more
    const Vec ax = norm(norm(Vec{axis[0], axis[1], axis[2]}));
    const double a = std::cos(angle / 2), h = std::sin(angle / 2);
    const double b = h * ax[0], c = h * ax[1], d = h * ax[2];
    const double aa = a * a, bb = b * b, cc = c * c, dd = d * d;
    const double bc = b * c, ad = a * d, ac = a * c, ab = a * b, bd = b * d, cd = c * d;
    const Mat R{aa + bb - cc - dd, 2 * (bc - ad),   2 * (bd + ac),  //
                2 * (bc + ad),   aa + cc - bb - dd, 2 * (cd - ab),  //
                2 * (bd - ac),   2 * (cd + ab),   aa + dd - bb - cc};
    const Vec vc{cm[0], cm[1], cm[2]}, ti{offset[0], offset[1], offset[2]}, tr{step[0], step[1], step[2]};
    // copy 1: v -> R (v - cm) + R t_i - t_i + t_r + cm; copy k: T_k = R T_{k-1} + T_1, R_k = R^k
    const Vec Rc = mat_vec(R, vc), Rt = mat_vec(R, ti);
    Vec base;
    for (int k = 0; k < 3; ++k) base[k] = (vc[k] - Rc[k]) + ((Rt[k] + tr[k]) - ti[k]);
    std::vector<double> out(12 * (size_t)reps);
    Mat Rk = R;
    Vec Tk = base;
    for (int rep = 1; rep <= reps; ++rep) {
        if (rep > 1) {
            const Vec m = mat_vec(R, Tk);
            for (int k = 0; k < 3; ++k) Tk[k] = m[k] + base[k];
            Rk = mat_mul(R, Rk);
        }
        std::copy(Rk.begin(), Rk.end(), out.begin() + 12 * (size_t)(rep - 1));
        std::copy(Tk.begin(), Tk.end(), out.begin() + 12 * (size_t)(rep - 1) + 9);
    }
    if (closed) {
        // CyclicSymmetry::is_closed (CyclicSymmetry.cpp:109-118)
        const double tm = std::sqrt(step[0] * step[0] + step[1] * step[1] + step[2] * step[2]);
        if (tm != 0) *closed = false;
        else if (std::abs(angle) < 1e-9) *closed = true;
        else *closed = ((int)std::round(100 * std::abs((reps + 1) * angle)) % 628) == 0;
    }
    return out;
}

namespace {
// weight classes of XYZW points (one per distinct weight bit pattern)
void classify(const float* p, int n, std::map<uint32_t, int>& id, std::vector<float>& cw, std::vector<int>& cls) {
    cls.resize(n);
    for (int i = 0; i < n; ++i) {
        const float w = p[4 * (size_t)i + 3];
        uint32_t key;
        std::memcpy(&key, &w, 4);
        auto it = id.find(key);
        if (it == id.end()) {
            it = id.emplace(key, (int)cw.size()).first;
            cw.push_back(w);
        }
        cls[i] = it->second;
    }
    if (cw.size() > 64) throw std::runtime_error("ausaxs_hip: more than 64 distinct weights in one point set");
}

int tri(int a, int b, int C) {
    if (a > b) std::swap(a, b);
    return a * C - a * (a - 1) / 2 + (b - a);
}
}  // namespace

void SymmetryStructure::check_spec(const SymmetrySpec& sp) {
    if (sp.reps < 1) throw std::runtime_error("ausaxs_hip: a symmetry needs at least one repetition");
    if (sp.transforms.size() != 12 * (size_t)sp.reps)
        throw std::runtime_error("ausaxs_hip: a symmetry needs 12 transform values per repetition");
    for (const CopyPair& p : sp.schedule)
        if (p.a < 0 || p.b < 0 || p.a > sp.reps || p.b > sp.reps || p.a == p.b || p.scale < 1)
            throw std::runtime_error("ausaxs_hip: invalid pair schedule entry (" + std::to_string(p.a) + ", " +
                                     std::to_string(p.b) + ", scale " + std::to_string(p.scale) + ")");
}

namespace {
int copies_of(const std::vector<SymmetrySpec>& syms) {
    int total = 0;
    for (const SymmetrySpec& sp : syms) total += sp.reps;
    return total;
}
}  // namespace

SymmetryStructure::SymmetryStructure(const float* atoms, const std::vector<int>& sizes,
                                     std::vector<std::vector<SymmetrySpec>> syms, const float* waters, int n_waters,
                                     const HistSettings& s, bool partial)
    : s_(s), partial_(partial), syms_(std::move(syms)) {
    const int B = (int)sizes.size();
    if ((int)syms_.size() != B) throw std::runtime_error("ausaxs_hip: one symmetry list per body is required");
    if (n_waters < 0) throw std::runtime_error("ausaxs_hip: negative water count");
    size_t off = 0;
    for (int b = 0; b < B; ++b) {
        if (sizes[b] < 0) throw std::runtime_error("ausaxs_hip: negative body size");
        bodies_.emplace_back(atoms + 4 * off, atoms + 4 * (off + sizes[b]));
        off += (size_t)sizes[b];
        for (const SymmetrySpec& sp : syms_[b]) check_spec(sp);
    }
    waters_.assign(waters, waters + 4 * (size_t)n_waters);
    self_bodies_ = bodies_;
    for (int b = 0; b < B; ++b) self_scale_.push_back(1 + copies_of(syms_[b]));
    rebuild();
}

void SymmetryStructure::set_body(int b, const float* pts, bool internal) {
    if (b < 0 || b >= (int)bodies_.size()) throw std::runtime_error("ausaxs_hip: body index out of range");
    bodies_[b].assign(pts, pts + bodies_[b].size());
    // calc_aa_self runs on an internal modification only (PartialSymmetryManagerMT.cpp:265-276)
    if (internal || !partial_) {
        self_bodies_[b] = bodies_[b];
        self_scale_[b] = 1 + copies_of(syms_[b]);
    }
    dirty_ = true;
}

void SymmetryStructure::set_symmetry(int b, int sym, SymmetrySpec spec) {
    if (b < 0 || b >= (int)bodies_.size() || sym < 0 || sym >= (int)syms_[b].size())
        throw std::runtime_error("ausaxs_hip: symmetry index out of range");
    check_spec(spec);
    syms_[b][sym] = std::move(spec);
    if (!partial_) self_scale_[b] = 1 + copies_of(syms_[b]);
    dirty_ = true;
}

void SymmetryStructure::set_waters(const float* pts, int n) {
    if (n < 0) throw std::runtime_error("ausaxs_hip: negative water count");
    waters_.assign(pts, pts + 4 * (size_t)n);
    dirty_ = true;
}

PairEngine& SymmetryStructure::engine() {
    if (dirty_) rebuild();
    return *engine_;
}

void SymmetryStructure::rebuild() {
    const int B = (int)bodies_.size();
    const int n_waters = (int)(waters_.size() / 4);

    // weight classes, shared by all body copies (a copy keeps its weights)
    std::map<uint32_t, int> aid, wid;
    std::vector<float> all;
    for (const auto& v : bodies_) all.insert(all.end(), v.begin(), v.end());
    for (const auto& v : self_bodies_) all.insert(all.end(), v.begin(), v.end());
    std::vector<int> acls, wcls;
    atom_class_w_.clear();
    water_class_w_.clear();
    classify(all.data(), (int)(all.size() / 4), aid, atom_class_w_, acls);
    classify(waters_.data(), n_waters, wid, water_class_w_, wcls);
    const int CA = std::max<int>(1, (int)atom_class_w_.size());
    const int CW = std::max<int>(1, (int)water_class_w_.size());
    if (atom_class_w_.empty()) atom_class_w_.push_back(0.f);
    if (water_class_w_.empty()) water_class_w_.push_back(0.f);
    auto cls_of = [&](float w) {
        uint32_t key;
        std::memcpy(&key, &w, 4);
        return aid.at(key);
    };

    // point sets: waters, then per body its original and every copy (and, in
    // partial mode, the coordinates of its self term)
    std::vector<PointSet> sets;
    auto add_set = [&](const std::vector<float>& pts, bool water) {
        PointSet ps;
        ps.n_classes = water ? CW : CA;
        const int n = (int)(pts.size() / 4);
        for (int i = 0; i < n; ++i)
            ps.orig.push_back(make_float4(pts[4 * i], pts[4 * i + 1], pts[4 * i + 2],
                                          float_of(water ? wcls[i] : cls_of(pts[4 * i + 3]))));
        sets.push_back(std::move(ps));
        return (int)sets.size() - 1;
    };
    const bool has_waters = n_waters > 0;
    const int set_w = has_waters ? add_set(waters_, true) : -1;
    std::vector<int> set_orig(B), set_self(B);
    std::vector<std::vector<int>> set_copy0(B);  // first copy set of each symmetry
    for (int b = 0; b < B; ++b) {
        const std::vector<float>& pb = bodies_[b];
        const int n = (int)(pb.size() / 4);
        set_orig[b] = add_set(pb, false);
        set_self[b] = partial_ ? add_set(self_bodies_[b], false) : set_orig[b];
        for (const SymmetrySpec& sp : syms_[b]) {
            set_copy0[b].push_back((int)sets.size());
            for (int k = 0; k < sp.reps; ++k) {
                // SymmetryHelpers.cpp:26-36: t(v.pos) in FP64 on the float
                // coordinates, rotation*v + translation (Vector3.h:169-181),
                // stored back as float
                const double* T = sp.transforms.data() + 12 * (size_t)k;
                std::vector<float> pts(4 * (size_t)n);
                for (int i = 0; i < n; ++i) {
                    const double x = pb[4 * i], y = pb[4 * i + 1], z = pb[4 * i + 2];
                    for (int c = 0; c < 3; ++c)
                        pts[4 * i + c] = (float)((T[3 * c] * x + T[3 * c + 1] * y + T[3 * c + 2] * z) + T[9 + c]);
                    pts[4 * i + 3] = pb[4 * i + 3];
                }
                add_set(pts, false);
            }
        }
    }
    auto set_at = [&](int b, int sym, int rep) { return rep == 0 ? set_orig[b] : set_copy0[b][sym] + rep - 1; };

    // row blocks: aa per multiplicity, aw, ww
    rows_.clear();
    int n_rows = 0;
    std::map<int, std::vector<int32_t>> aa_maps;
    auto aa_map = [&](int scale) -> const std::vector<int32_t>& {
        auto it = aa_maps.find(scale);
        if (it != aa_maps.end()) return it->second;
        std::vector<int32_t> m((size_t)CA * CA);
        const int base = n_rows;
        for (int a = 0; a < CA; ++a)
            for (int c = a; c < CA; ++c) {
                m[a * CA + c] = m[c * CA + a] = base + tri(a, c, CA);
                rows_.push_back({0, a, c, 2 * scale});
            }
        n_rows += CA * (CA + 1) / 2;
        return aa_maps.emplace(scale, std::move(m)).first->second;
    };
    // aw rows (atom class, water class); the job's row points are the body
    // copy (full mode) or the waters (partial mode)
    std::vector<int32_t> aw_body_rows, aw_water_rows, ww_map;
    if (has_waters) {
        aw_body_rows.resize((size_t)CA * CW);
        aw_water_rows.resize((size_t)CW * CA);
        for (int a = 0; a < CA; ++a)
            for (int w = 0; w < CW; ++w) {
                aw_body_rows[a * CW + w] = aw_water_rows[w * CA + a] = n_rows++;
                rows_.push_back({1, a, w, 2});
            }
        ww_map.resize((size_t)CW * CW);
        for (int a = 0; a < CW; ++a)
            for (int c = a; c < CW; ++c) {
                ww_map[a * CW + c] = ww_map[c * CW + a] = n_rows++;
                rows_.push_back({2, a, c, 2});
            }
    }

    // an enqueue_calculate_cross(d1, d2) runs rows over d2 and columns over d1
    // (SimpleCalculator.h:188-211)
    std::vector<PairJob> jobs;
    auto cross = [&](int d1, int d2, const std::vector<int32_t>& map) {
        if (sets[d1].size() == 0 || sets[d2].size() == 0) return;
        PairJob j;
        j.row_set = d2;
        j.col_set = d1;
        j.rowmap = map;
        jobs.push_back(std::move(j));
    };
    auto self = [&](int d, const std::vector<int32_t>& map) {
        if (sets[d].size() == 0) return;
        PairJob j;
        j.row_set = d;
        j.rowmap = map;
        jobs.push_back(std::move(j));
    };
    body_self_w2_.assign(B, 0.0);
    for (int b = 0; b < B; ++b) {
        // every copy has the original's internal distances: evaluated once and
        // scaled by 1 + copies; bin 0 gets scale * sum w^2
        self(set_self[b], aa_map(self_scale_[b]));
        double t = 0;
        const std::vector<float>& pb = self_bodies_[b];
        for (size_t i = 0; i < pb.size() / 4; ++i) t += (double)(pb[4 * i + 3] * pb[4 * i + 3]);
        body_self_w2_[b] = self_scale_[b] * t;
    }
    if (!partial_) {
        // SymmetryManagerMT.cpp:57-132
        for (int b1 = 0; b1 < B; ++b1) {
            if (has_waters) cross(set_w, set_orig[b1], aw_body_rows);
            for (int s1 = 0; s1 < (int)syms_[b1].size(); ++s1) {
                const SymmetrySpec& sp = syms_[b1][s1];
                for (const CopyPair& p : sp.schedule) cross(set_at(b1, s1, p.a), set_at(b1, s1, p.b), aa_map(p.scale));
                for (int r1 = 1; r1 <= sp.reps; ++r1) {
                    const int c1 = set_at(b1, s1, r1);
                    if (has_waters) cross(set_w, c1, aw_body_rows);
                    for (int b2 = b1 + 1; b2 < B; ++b2) {
                        cross(set_orig[b2], c1, aa_map(1));
                        for (int s2 = 0; s2 < (int)syms_[b2].size(); ++s2)
                            for (int r2 = 1; r2 <= syms_[b2][s2].reps; ++r2) cross(c1, set_at(b2, s2, r2), aa_map(1));
                    }
                    for (int s2 = s1 + 1; s2 < (int)syms_[b1].size(); ++s2)
                        for (int r2 = 1; r2 <= syms_[b1][s2].reps; ++r2) cross(c1, set_at(b1, s2, r2), aa_map(1));
                }
            }
            for (int b2 = b1 + 1; b2 < B; ++b2) {
                cross(set_orig[b1], set_orig[b2], aa_map(1));
                for (int s2 = 0; s2 < (int)syms_[b2].size(); ++s2)
                    for (int r2 = 1; r2 <= syms_[b2][s2].reps; ++r2) cross(set_orig[b1], set_at(b2, s2, r2), aa_map(1));
            }
        }
    } else {
        // PartialSymmetryManagerMT::calc_aa / calc_aw (:596-706): between two
        // bodies (b1 > b2) rows run over b2's coordinates, within a body over
        // the lower symmetry, and body x water over the waters
        auto reps_of = [&](int b, int sym) { return sym == 0 ? 1 : syms_[b][sym - 1].reps; };
        auto coords = [&](int b, int sym, int rep) { return sym == 0 ? set_orig[b] : set_copy0[b][sym - 1] + rep; };
        for (int b1 = 0; b1 < B; ++b1) {
            const int ns1 = (int)syms_[b1].size();
            for (int b2 = 0; b2 < b1; ++b2)
                for (int s1 = 0; s1 <= ns1; ++s1)
                    for (int s2 = 0; s2 <= (int)syms_[b2].size(); ++s2)
                        for (int r1 = 0; r1 < reps_of(b1, s1); ++r1)
                            for (int r2 = 0; r2 < reps_of(b2, s2); ++r2)
                                cross(coords(b1, s1, r1), coords(b2, s2, r2), aa_map(1));
            for (int s1 = 0; s1 < ns1; ++s1) {
                for (const CopyPair& p : syms_[b1][s1].schedule)
                    cross(set_at(b1, s1, p.a), set_at(b1, s1, p.b), aa_map(p.scale));
                for (int s2 = 0; s2 < s1; ++s2)
                    for (int r1 = 0; r1 < syms_[b1][s1].reps; ++r1)
                        for (int r2 = 0; r2 < syms_[b1][s2].reps; ++r2)
                            cross(coords(b1, s1 + 1, r1), coords(b1, s2 + 1, r2), aa_map(1));
            }
            if (has_waters)
                for (int s1 = 0; s1 <= ns1; ++s1)
                    for (int r1 = 0; r1 < reps_of(b1, s1); ++r1) cross(coords(b1, s1, r1), set_w, aw_water_rows);
        }
    }
    water_self_w2_ = 0;
    if (has_waters) {
        self(set_w, ww_map);
        for (int i = 0; i < n_waters; ++i) water_self_w2_ += (double)(waters_[4 * i + 3] * waters_[4 * i + 3]);
    }
    if (n_rows == 0) aa_map(1);  // keeps the row table non-empty

    const float inv_w = inv_width(s_);
    std::vector<const PointSet*> ps;
    for (const auto& x : sets) ps.push_back(&x);
    // at least the 10 bins every truncated histogram keeps (max_bin >= 10)
    const int nbins = (int)std::min<int64_t>((int64_t)s_.bin_count, std::max(bin_bound(ps, inv_w), 10));
    engine_ = std::make_unique<PairEngine>(std::move(sets), std::move(jobs), n_rows, inv_w, nbins, s_.tile_j);
    dirty_ = false;
}

std::unique_ptr<CompositeDistanceHistogramHIP> SymmetryStructure::calculate(bool timed) {
    if (dirty_) rebuild();
    engine_->run(cnt_, dsum_, timed);
    const int nbins = engine_->nbins();
    const int stride = nbins + 1;
    if (engine_->overflowed(cnt_))
        throw std::runtime_error("ausaxs_hip: a pair distance falls beyond bin_count*bin_width (" +
                                 std::to_string(s_.bin_count) + " bins of " + std::to_string(s_.bin_width) +
                                 " A); increase bin_count or bin_width");
    const size_t nb = nbins;
    const float inv_w = engine_->inv_w();
    const double unit = delta_unit(inv_w);
    std::vector<WeightedEntry> aa(nb), aw(nb), ww(nb);
    for (size_t r = 0; r < rows_.size(); ++r) {
        const auto& k = rows_[r];
        double wprod;
        std::vector<WeightedEntry>* dst;
        if (k[0] == 0) wprod = (double)(atom_class_w_[k[1]] * atom_class_w_[k[2]]), dst = &aa;
        else if (k[0] == 1) wprod = (double)(atom_class_w_[k[1]] * water_class_w_[k[2]]), dst = &aw;
        else wprod = (double)(water_class_w_[k[1]] * water_class_w_[k[2]]), dst = &ww;
        const uint64_t factor = (uint64_t)k[3];  // add<2 * scaling> (SimpleCalculator.h:118-134)
        for (int b = 0; b < nbins; ++b) {
            const uint64_t c = cnt_[r * stride + b];
            if (!c) continue;
            const int64_t sd = dsum_[r * stride + b];
            WeightedEntry e;
            e.value = (double)(factor * c) * wprod;
            e.count = (uint32_t)(factor * c);
            e.bin_center = (double)factor * distance_sum_u(c, b, sd, inv_w, unit);
            entry_add((*dst)[b], e);
        }
    }
    // SimpleCalculator.h:144-160 per self enqueue: bin 0 += (t, (unsigned) t, 0)
    for (double t : body_self_w2_) entry_add(aa[0], WeightedEntry{t, (uint32_t)t, 0});
    entry_add(ww[0], WeightedEntry{water_self_w2_, (uint32_t)water_self_w2_, 0});
    std::vector<WeightedEntry> ptot(nb);
    for (size_t b = 0; b < nb; ++b) {
        ptot[b] = aa[b];
        entry_add(ptot[b], ww[b]);
        entry_add(ptot[b], aw[b]);
    }
    const unsigned max_bin = std::min<unsigned>(max_bin_of(ptot), (unsigned)nbins);
    auto res = std::make_unique<CompositeDistanceHistogramHIP>();
    res->mode = Mode::Simple;
    res->bin_width = s_.bin_width;
    res->weighted_bins = s_.weighted_bins;
    res->debye_terms = s_.debye_terms;
    res->max_bin = max_bin;
    res->aa.reset(1, 1, (int)max_bin);
    res->aw.reset(1, 1, (int)max_bin);
    std::copy_n(aa.begin(), max_bin, res->aa.touch(0, 0));
    std::copy_n(aw.begin(), max_bin, res->aw.touch(0, 0));
    res->ww.assign(ww.begin(), ww.begin() + max_bin);
    res->p_tot.assign(ptot.begin(), ptot.begin() + max_bin);
    res->d_axis = axis_of(res->p_tot, s_.bin_width, s_.weighted_bins);
    return res;
}

}  // namespace ausaxs::gpu
