// Host side of the pair-histogram engine (see engine.hpp).
#include "engine.hpp"

#include "host_parallel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace ausaxs::gpu {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

Context::Context() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        throw std::runtime_error("ausaxs_hip: no HIP device available (the GPU path has no CPU fallback)");
    check(hipGetDevice(&device), "hipGetDevice");
    check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    check(hipStreamCreateWithFlags(&default_stream, hipStreamNonBlocking), "hipStreamCreate");
}

Context& Context::get() {
    static Context ctx;
    return ctx;
}

namespace {
thread_local hipStream_t t_stream = nullptr;
}

void Context::set_thread_stream(hipStream_t s) { t_stream = s; }

hipStream_t Context::stream() const { return t_stream ? t_stream : default_stream; }

namespace {
constexpr int kTileI = 512;  // rows per work item: 256 lanes x 2 (pair_hist.hip)
}

int bin_bound(const std::vector<const PointSet*>& sets, float inv_w) {
    double lo[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double hi[3] = {-lo[0], -lo[1], -lo[2]};
    bool any = false;
    for (auto* s : sets)
        for (const auto& p : s->orig) {
            any = true;
            const double v[3] = {p.x, p.y, p.z};
            for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], v[k]), hi[k] = std::max(hi[k], v[k]);
        }
    if (!any) return 1;
    double d2 = 0;
    for (int k = 0; k < 3; ++k) d2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    // no pair is farther apart than twice the largest distance from the box
    // centre (|p - q| <= |p - c| + |q - c|): for a round structure that is
    // close to its diameter, while the box diagonal can exceed it by 40 %
    const double c[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    double r2 = 0;
    for (auto* s : sets)
        for (const auto& p : s->orig) {
            const double dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
            r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
        }
    const double span = std::min(std::sqrt(d2), 2 * std::sqrt(r2));
    const double x = span * (double)inv_w * (1 + 1e-5) + 2;
    if (!(x < 1e9)) throw std::runtime_error("ausaxs_hip: coordinates are not finite or too far apart");
    // the pair kernel's offset codes hold the Newton correction of pairs up
    // to kMaxExtentBins bins apart
    if (x > kMaxExtentBins)
        throw std::runtime_error("ausaxs_hip: the structure spans more than " + std::to_string(kMaxExtentBins) +
                                 " bins; use wider bins");
    return (int)std::ceil(x);
}

void check_tile_j(int tile_j) {
    if (tile_j < 1 || (int64_t)kTileI * tile_j > kMaxPairsPerItem)
        throw std::runtime_error("ausaxs_hip: tile_j must keep a work item within " +
                                 std::to_string(kMaxPairsPerItem) + " pairs (tile_j <= " +
                                 std::to_string(kMaxPairsPerItem / kTileI) + ")");
}

int item_rows(const std::vector<PairWork>& items) {
    for (const PairWork& w : items)
        if (w.flags & (kOrient | kSplit)) return 2;
    return 1;
}

int pick_copies(int nbins, int rows) {
    int k = 8;
    while (k > 1 && pair_lds_bytes(nbins, rows, k) > 65536) k /= 2;
    return k;
}

size_t check_pair_lds(int nbins) {
    if (nbins < 1) throw std::runtime_error("ausaxs_hip: nbins must be positive");
    // two histogram rows + the static column tile of pair_hist_kernel
    const size_t lds = pair_lds_bytes(nbins, 2, 1);
    Context& ctx = Context::get();
    if (lds > ctx.prop.sharedMemPerBlock)
        throw std::runtime_error("ausaxs_hip: " + std::to_string(nbins) + " bins need " + std::to_string(lds) +
                                 " B of LDS per workgroup, device offers " +
                                 std::to_string(ctx.prop.sharedMemPerBlock));
    return lds;
}

SetLayout layout_set(const PointSet& ps, std::vector<float4>& buf, std::vector<int32_t>* perm) {
    const int n = ps.size(), C = ps.n_classes;
    const int t_self = n - std::min(3, n);  // self-loop tail candidates: [t_self, n)
    const int t_col = n - (n % 4);          // cross-loop scalar columns: [t_col, n)
    SetLayout L;
    L.base = (int)buf.size();
    L.seg_start.assign(C, 0);
    L.seg_len.assign(C, 0);
    L.seg_body.assign(C, 0);
    L.seg_colbody.assign(C, 0);
    std::vector<std::vector<int>> body(C), mid(C), tail(C);
    for (int i = 0; i < n; ++i) {
        const int c = bits_of(ps.orig[i].w);
        if (c < 0 || c >= C) throw std::runtime_error("ausaxs_hip: point class out of range");
        if (i < t_self) body[c].push_back(i);
        else if (i < t_col) mid[c].push_back(i);
        else tail[c].push_back(i);
    }
    const size_t base = buf.size(), pbase = perm ? perm->size() : 0;
    buf.resize(base + (size_t)n);
    if (perm) perm->resize(pbase + (size_t)n);
    size_t pos = base;
    for (int c = 0; c < C; ++c) {
        // Scatter the body part: consecutive points of a structure are spatial
        // neighbours (chains), so a row tile in input order is a compact blob
        // whose distances to one column point crowd into few bins, and the
        // lanes of a wave collide on the same LDS cells. A golden-ratio stride
        // through the segment spreads every tile over the whole structure.
        // Order within a segment is free: orientation follows the tags and
        // diagonal blocks mask by buffer position.
        std::vector<int>& v = body[c];
        const size_t nv = v.size();
        if (nv > 2) {
            size_t st = (size_t)(0.6180339887498949 * (double)nv);
            while (std::gcd(st, nv) != 1) ++st;
            std::vector<int> s(nv);
            // the gathers jump through memory: chunks on several threads
            parallel_for((int64_t)nv, 1 << 16, [&](int64_t k0, int64_t k1) {
                size_t j = (size_t)(((unsigned __int128)k0 * st) % nv);
                for (int64_t k = k0; k < k1; ++k) {
                    s[k] = v[j];
                    j += st;
                    if (j >= nv) j -= nv;
                }
            });
            v.swap(s);
        }
        L.seg_start[c] = (int)pos;
        for (auto* v : {&body[c], &mid[c], &tail[c]}) {
            const size_t at = pos;
            parallel_for((int64_t)v->size(), 1 << 16, [&](int64_t k0, int64_t k1) {
                for (int64_t k = k0; k < k1; ++k) {
                    const int i = (*v)[k];
                    float4 p = ps.orig[i];
                    p.w = float_of(i);
                    buf[at + k] = p;
                    if (perm) (*perm)[pbase + (at - base) + k] = i;
                }
            });
            pos += v->size();
        }
        L.seg_len[c] = (int)(body[c].size() + mid[c].size() + tail[c].size());
        L.seg_body[c] = (int)body[c].size();
        L.seg_colbody[c] = (int)(body[c].size() + mid[c].size());
    }
    return L;
}

namespace {
void add_items(std::vector<PairWork>& items, int i0, int i1, int j0, int j1, int row0, int row1, bool diag,
               int tile_j) {
    const bool orient = row0 != row1;
    for (int a = i0; a < i1; a += kTileI) {
        const int a1 = std::min(a + kTileI, i1);
        for (int b = j0; b < j1; b += tile_j) {
            PairWork w{};
            w.i_begin = a;
            w.i_end = a1;
            w.j_begin = b;
            w.j_end = std::min(b + tile_j, j1);
            w.row0 = row0;
            w.row1 = orient ? row1 : row0;
            w.flags = (diag ? kDiag : 0) | (orient ? kOrient : 0);
            items.push_back(w);
        }
    }
}
}  // namespace

bool lattice_runs(const PointSet& ps, int mb, int max_cols, std::vector<float4>& runs) {
    // integer lattice points only (the run kernel steps z by exactly one):
    // |coordinates| below 2^22 that a cast to int and back leaves unchanged
    const float lim = (float)(1 << 22);
    auto lattice = [&](const float4& p) {
        return std::fabs(p.x) < lim && std::fabs(p.y) < lim && std::fabs(p.z) < lim && p.x == (float)(int)p.x &&
               p.y == (float)(int)p.y && p.z == (float)(int)p.z;
    };
    const size_t r0 = runs.size();
    int i = 0;
    while (i < mb) {
        const float4& p = ps.orig[i];
        if (!lattice(p)) {
            runs.resize(r0);
            return false;
        }
        int len = 1;
        while (i + len < mb && len < max_cols && ps.orig[i + len].x == p.x && ps.orig[i + len].y == p.y &&
               ps.orig[i + len].z == p.z + (float)len)
            ++len;  // the points of the run are lattice points too (same x, y; z integer)
        runs.push_back(make_float4(p.x, p.y, p.z, float_of(len)));
        i += len;
    }
    // short runs would cost more per pair than plain columns
    if (mb > 0 && (double)mb / (double)(runs.size() - r0) < 4.0) {
        runs.resize(r0);
        return false;
    }
    return true;
}

void append_job(WorkList& wl, const PairJob& job, const std::vector<PointSet>& sets,
                const std::vector<SetLayout>& layouts, int n_rows, int tile_j, bool split_tiles, bool body_items) {
    const bool self = job.col_set < 0;
    const PointSet& R = sets[job.row_set];
    const PointSet& Cs = self ? R : sets[job.col_set];
    const SetLayout& LR = layouts[job.row_set];
    const SetLayout& LC = self ? LR : layouts[job.col_set];
    const int CR = R.n_classes, CC = Cs.n_classes;
    if ((int)job.rowmap.size() != CR * CC) throw std::runtime_error("ausaxs_hip: rowmap size mismatch");
    for (int32_t r : job.rowmap)
        if (r < -1 || r >= n_rows) throw std::runtime_error("ausaxs_hip: rowmap entry out of range");
    auto row_of = [&](int a, int b) {
        const int32_t r = job.rowmap[a * CC + b];
        if (r < 0) throw std::runtime_error("ausaxs_hip: missing histogram row for a present class pair");
        return r;
    };
    auto add_strip = [&](const StripWork& w) {
        StripSpec st{};
        st.set_rows = job.row_set;
        st.set_cols = self ? job.row_set : job.col_set;
        st.w = w;
        st.rowmap_off = (int)wl.rowmaps.size();
        wl.strips.push_back(st);
        wl.rowmaps.insert(wl.rowmaps.end(), job.rowmap.begin(), job.rowmap.end());
    };
    if (self) {
        const int64_t n = R.size();
        wl.pairs += n * (n - 1) / 2;
        for (int a = 0; a < CR && body_items; ++a) {
            if (LR.seg_body[a] == 0) continue;
            const int s0 = LR.seg_start[a], s1 = s0 + LR.seg_body[a];
            // triangle: diagonal blocks masked to j > i, the rest plain
            const int ra = row_of(a, a);
            for (int i0 = s0; i0 < s1; i0 += kTileI) {
                const int i1 = std::min(i0 + kTileI, s1);
                add_items(wl.items, i0, i1, i0, i1, ra, ra, true, tile_j);
                if (i1 < s1) add_items(wl.items, i0, i1, i1, s1, ra, ra, false, tile_j);
            }
            for (int b = a + 1; b < CR; ++b) {
                if (LR.seg_body[b] == 0) continue;
                // oriented items compare original indices as floats
                if (row_of(a, b) != row_of(b, a) && n > (int64_t(1) << 24))
                    throw std::runtime_error("ausaxs_hip: a point set with per-orientation rows holds at most 2^24 points");
                const int t0 = LR.seg_start[b], t1 = t0 + LR.seg_body[b];
                add_items(wl.items, s0, s1, t0, t1, row_of(a, b), row_of(b, a), false, tile_j);
            }
        }
        const int nn = R.size();
        if (nn >= 2) {
            add_strip(StripWork{nn, nn - std::min(3, nn), nn, nn, 1, CC});
            for (int t = nn - std::min(3, nn); t < nn; ++t) wl.strip_pairs += t;  // rows i < t
        }
    } else {
        const int64_t n = R.size(), m = Cs.size();
        wl.pairs += n * m;
        // Row tiles run through the class segments (contiguous in class order)
        // and may span two classes: a lane's output row then depends only on
        // its own row's class, which is constant over the item (kSplit).
        // Column tiles stay within one class segment.
        std::vector<int> seg_cls;  // non-empty row classes in buffer order
        for (int a = 0; a < CR; ++a)
            if (LR.seg_len[a]) seg_cls.push_back(a);
        for (int b = 0; b < CC && body_items; ++b) {
            if (LC.seg_colbody[b] == 0) continue;
            const int t0 = LC.seg_start[b], t1 = t0 + LC.seg_colbody[b];
            // column chunks of the items: point ranges of the buffer, or, for
            // lattice columns, ranges of z-runs (kRuns) over the same points
            struct Chunk {
                int jb, je, flags;
            };
            std::vector<Chunk> chunks;
            // the runs of a column set, built once per list
            int run_lo = 0, run_hi = -1;
            if (job.col_runs && CC == 1) {
                auto it = std::find_if(wl.set_runs.begin(), wl.set_runs.end(),
                                       [&](const std::array<int, 3>& e) { return e[0] == job.col_set; });
                if (it == wl.set_runs.end()) {
                    const int r0 = (int)wl.runs.size();
                    const bool ok = lattice_runs(Cs, LC.seg_colbody[0], tile_j, wl.runs);
                    wl.set_runs.push_back({job.col_set, r0, ok ? (int)wl.runs.size() : -1});
                    it = wl.set_runs.end() - 1;
                }
                run_lo = (*it)[1];
                run_hi = (*it)[2];
            }
            if (run_hi >= 0) {
                int cols = 0, first = run_lo;
                for (int q = run_lo; q < run_hi; ++q) {
                    const int len = bits_of(wl.runs[q].w);
                    if (cols + len > tile_j) {
                        chunks.push_back({first, q, kRuns | (cols << kRunColsShift)});
                        first = q;
                        cols = 0;
                    }
                    cols += len;
                }
                if (cols) chunks.push_back({first, run_hi, kRuns | (cols << kRunColsShift)});
            } else {
                for (int jb = t0; jb < t1; jb += tile_j) chunks.push_back({jb, std::min(jb + tile_j, t1), 0});
            }
            size_t k = 0;
            int i = seg_cls.empty() ? 0 : LR.seg_start[seg_cls[0]];
            while (k < seg_cls.size()) {
                const int a = seg_cls[k];
                const int a_end = LR.seg_start[a] + LR.seg_len[a];
                int i1 = std::min(i + kTileI, a_end);
                int split = 0, r1 = row_of(a, b);
                size_t next = k;
                if (i1 == a_end) {
                    next = k + 1;
                    // fill the tile with the next class (at most two classes per tile)
                    if (split_tiles && next < seg_cls.size() && i + kTileI > a_end) {
                        const int a2 = seg_cls[next];
                        const int a2_end = LR.seg_start[a2] + LR.seg_len[a2];
                        const int i2 = std::min(i + kTileI, a2_end);
                        split = a_end - i;
                        r1 = row_of(a2, b);
                        i1 = i2;
                        if (i2 == a2_end) next = k + 2;
                    }
                }
                const int r0 = row_of(a, b);
                for (const Chunk& ch : chunks) {
                    PairWork w{};
                    w.i_begin = i;
                    w.i_end = i1;
                    w.j_begin = ch.jb;
                    w.j_end = ch.je;
                    w.row0 = r0;
                    w.row1 = split && r1 != r0 ? r1 : r0;
                    w.flags = ch.flags | (split && r1 != r0 ? (kSplit | (split << kSplitShift)) : 0);
                    wl.items.push_back(w);
                }
                i = i1;
                k = next;
            }
        }
        const int mm = Cs.size();
        if (mm % 4 && R.size() > 0) {
            add_strip(StripWork{R.size(), mm - (mm % 4), mm, mm, 0, CC});
            wl.strip_pairs += (int64_t)R.size() * (mm % 4);
        }
    }
}

bool plan_lattice(const PairJob& job, const std::vector<PointSet>& sets, int mode, LatticeJob& out) {
    if (mode == 0 || !job.lattice) return false;
    const bool self = job.col_set < 0;
    const PointSet& R = sets[job.row_set];
    const PointSet& Cs = self ? R : sets[job.col_set];
    if (R.n_classes != 1 || Cs.n_classes != 1 || job.rowmap.size() != 1 || job.rowmap[0] < 0) return false;
    LatticeJob j;
    j.row_set = job.row_set;
    j.col_set = job.col_set;
    j.out_row = job.rowmap[0];
    // the body pairs: the strips take the self-loop tail candidates (last 3
    // points) and the cross loop's scalar columns (last m mod 4)
    const int n = R.size(), m = Cs.size();
    j.n_rows = self ? n - std::min(3, n) : n;
    j.n_cols = self ? j.n_rows : m - (m % 4);
    j.pairs = self ? (int64_t)j.n_rows * (j.n_rows - 1) / 2 : (int64_t)j.n_rows * j.n_cols;
    if (j.pairs == 0) return false;
    // integer coordinates well inside float's exact range, and their span
    constexpr float kLim = 1 << 20;
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    auto scan = [&](const PointSet& ps, int count) {
        for (int i = 0; i < count; ++i) {
            const float v[3] = {ps.orig[i].x, ps.orig[i].y, ps.orig[i].z};
            for (int k = 0; k < 3; ++k) {
                if (!(std::fabs(v[k]) <= kLim) || v[k] != std::rint(v[k])) return false;
                lo[k] = std::min(lo[k], (int)v[k]);
                hi[k] = std::max(hi[k], (int)v[k]);
            }
        }
        return true;
    };
    if (!scan(R, j.n_rows) || (!self && !scan(Cs, j.n_cols))) return false;
    int64_t cells = 1, s_true = 0;
    for (int k = 0; k < 3; ++k) {
        const int64_t E = (int64_t)hi[k] - lo[k] + 1;
        const int L = E <= kLatticeMaxL ? lattice_length((int)E) : -1;
        if (L < 0) return false;
        j.g.origin[k] = lo[k];
        j.g.L[k] = L;
        j.g.E[k] = (int)E;
        cells *= L;
        s_true += (E - 1) * (E - 1);
    }
    // counts per squared displacement, |D| < E per axis
    j.smax = (int)std::min<int64_t>(s_true, INT32_MAX);
    // d2 = s exact in float: every squared displacement below 2^24
    if (s_true >= (int64_t(1) << 24) || cells > (int64_t(1) << 28)) return false;
    // automatic choice: a transform costs about as much as 50 pair
    // evaluations per cell, and its ~10 launches as 1e8 pairs of the pair kernel
    if (mode == 1 && (j.pairs < 64 * cells * (self ? 1 : 2) || j.pairs < 100000000)) return false;
    out = j;
    return true;
}


void sort_items(std::vector<PairWork>& items) {
    std::stable_sort(items.begin(), items.end(), [](const PairWork& x, const PairWork& y) {
        return (int64_t)(x.i_end - x.i_begin) * item_cols(x) > (int64_t)(y.i_end - y.i_begin) * item_cols(y);
    });
}

int64_t padded_work(const std::vector<PairWork>& items) {
    int64_t w = 0;
    for (const PairWork& x : items) w += (int64_t)kTileI * item_cols(x);
    return w;
}

bool prefer_split(int64_t work_plain, int rows_plain, int64_t work_split, int rows_split) {
    if (rows_split <= rows_plain) return work_split * 50 <= work_plain * 49;
    return work_split * 10 <= work_plain * 9;
}

MergePlan plan_merge(std::vector<PairWork>& items, int n_rows) {
    std::vector<std::vector<int32_t>> per_row(n_rows);
    for (size_t k = 0; k < items.size(); ++k) {
        items[k].slot = (int32_t)k;
        per_row[items[k].row0].push_back((int32_t)(2 * k));
        if (items[k].flags & (kOrient | kSplit)) per_row[items[k].row1].push_back((int32_t)(2 * k + 1));
    }
    // entries grouped by row, cut into merge chunks of at most kMergeChunk
    MergePlan mp;
    mp.chunk_ptr.assign(1, 0);
    mp.row_chunk.assign(1, 0);
    for (auto& v : per_row) {
        for (size_t k = 0; k < v.size(); k += kMergeChunk) {
            const size_t k1 = std::min(v.size(), k + kMergeChunk);
            mp.row_list.insert(mp.row_list.end(), v.begin() + k, v.begin() + k1);
            mp.chunk_ptr.push_back((int32_t)mp.row_list.size());
        }
        mp.row_chunk.push_back((int32_t)mp.chunk_ptr.size() - 1);
    }
    // slack for the merge's batched reads: partial row 0 (read, then dropped)
    mp.row_list.resize(mp.row_list.size() + kMergeBatch - 1, 0);
    return mp;
}

PairEngine::PairEngine(std::vector<PointSet> sets, std::vector<PairJob> jobs, int n_rows, float inv_w, int nbins,
                       int tile_j, int lattice, std::vector<std::vector<float>> weights)
    : PairEngine(EngineSpec{std::move(sets), std::move(jobs), n_rows, inv_w, nbins, tile_j, lattice,
                            std::move(weights), -1}) {}

EngineSpec merge_specs(const std::vector<SpecView>& specs, std::vector<int>& row0) {
    if (specs.empty()) throw std::runtime_error("ausaxs_hip: nothing to merge");
    EngineSpec m;
    m.inv_w = specs[0].inv_w;
    m.tile_j = specs[0].tile_j;
    m.lattice = specs[0].lattice;
    m.nbins = 0;
    m.clamp = 0;
    row0.clear();
    for (const SpecView& v : specs) {
        const SpecView* sp = &v;
        if (sp->inv_w != m.inv_w || sp->tile_j != m.tile_j || sp->lattice != m.lattice || sp->weighted)
            throw std::runtime_error("ausaxs_hip: merged structures must share bin width, tile width and lattice "
                                     "mode and carry no per-point weights");
        const int set0 = (int)m.sets.size();
        row0.push_back(m.n_rows);
        for (const PointSet& ps : *sp->sets) m.sets.push_back(ps);
        for (PairJob j : *sp->jobs) {
            j.row_set += set0;
            if (j.col_set >= 0) j.col_set += set0;
            for (int32_t& r : j.rowmap)
                if (r >= 0) r += m.n_rows;
            m.jobs.push_back(std::move(j));
        }
        m.n_rows += sp->n_rows;
        m.nbins = std::max(m.nbins, sp->nbins);
        // a spec without a known clamp flag is taken as clamped
        m.clamp = m.clamp || sp->clamp != 0;
    }
    return m;
}

PairEngine::~PairEngine() {
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
}

size_t PairEngine::device_bytes() const {
    size_t b = d_pts_.bytes() + d_runs_.bytes() + d_pts_w_.bytes() + d_wcnt_.bytes() + d_wsum_.bytes() + d_wstrip_.bytes() +
               d_items_.bytes() + d_row_list_.bytes() + d_chunk_ptr_.bytes() + d_row_chunk_.bytes() +
               d_rowmaps_.bytes() + d_partial_.bytes() + d_cnt_.bytes() + d_tmp_cnt_.bytes() + d_dsum_.bytes() +
               d_tmp_dsum_.bytes() + d_lat_a_.bytes() + d_lat_b_.bytes() + d_tw_.bytes() + d_scount_.bytes() +
               d_lat_flag_.bytes() + strips_.bytes() + wstrips_.bytes();
    for (const auto& o : d_orig_) b += o.bytes();
    for (const auto& o : d_orig_w_) b += o.bytes();
    return b;
}

PairEngine::PairEngine(EngineSpec spec)
    : sets_(std::move(spec.sets)), jobs_(std::move(spec.jobs)), n_rows_(spec.n_rows), nbins_(spec.nbins),
      inv_w_(spec.inv_w), lattice_mode_(spec.lattice), weights_(std::move(spec.weights)) {
    const int lattice = spec.lattice, tile_j = spec.tile_j, nbins = spec.nbins;
    tile_j_ = tile_j;
    if (lattice < 0 || lattice > 2) throw std::runtime_error("ausaxs_hip: lattice must be 0, 1 or 2");
    weighted_ = !weights_.empty();
    if (weighted_) {
        // the weight pass bins like the count pass and adds weight products:
        // one class per set (no oriented or split items), no lattice jobs
        if (weights_.size() != sets_.size()) throw std::runtime_error("ausaxs_hip: one weight vector per point set");
        double wmax = 0;
        for (size_t k = 0; k < sets_.size(); ++k) {
            if (sets_[k].n_classes != 1 || weights_[k].size() != sets_[k].orig.size())
                throw std::runtime_error("ausaxs_hip: weighted point sets hold one class and one weight per point");
            for (float w : weights_[k]) {
                if (!std::isfinite(w)) throw std::runtime_error("ausaxs_hip: weights must be finite");
                wmax = std::max(wmax, (double)std::fabs(w));
            }
        }
        for (const PairJob& j : jobs_)
            if (j.lattice || j.col_runs)
                throw std::runtime_error("ausaxs_hip: weighted point sets take no lattice jobs");
        wscale_ = weight_scale(wmax);
    }
    check_tile_j(tile_j);
    check_pair_lds(nbins);
    // bins need no clamp in the pair kernel when the bounding-box diagonal
    // already keeps every pair below nbins
    if (spec.clamp < 0) {
        std::vector<const PointSet*> ps;
        for (const auto& x : sets_) ps.push_back(&x);
        clamp_ = nbins_ < bin_bound(ps, inv_w_);
    } else {
        clamp_ = spec.clamp != 0;
    }
    build(tile_j);
    for (auto& e : ev_) check(hipEventCreate(&e), "hipEventCreate");
}

namespace {
// The class segments' sizes of a set that is not placed in the point buffer
// (base -1): its points are read only in original order (lattice jobs, tail
// strips) or as z-runs.
SetLayout layout_counts(const PointSet& ps) {
    const int n = ps.size(), C = ps.n_classes;
    const int t_self = n - std::min(3, n), t_col = n - (n % 4);
    SetLayout L;
    L.base = -1;
    L.seg_start.assign(C, -1);
    L.seg_len.assign(C, 0);
    L.seg_body.assign(C, 0);
    L.seg_colbody.assign(C, 0);
    for (int i = 0; i < n; ++i) {
        const int c = bits_of(ps.orig[i].w);
        if (c < 0 || c >= C) throw std::runtime_error("ausaxs_hip: point class out of range");
        ++L.seg_len[c];
        if (i < t_self) ++L.seg_body[c];
        if (i < t_col) ++L.seg_colbody[c];
    }
    return L;
}
}  // namespace

void PairEngine::build(int tile_j) {
    Context& ctx = Context::get();
    // lattice jobs first: their body pairs need no work items
    lat_.clear();
    std::vector<char> on_lattice(jobs_.size(), 0);
    int64_t lattice_pairs = 0;
    for (size_t k = 0; k < jobs_.size(); ++k) {
        LatticeJob lj;
        if (plan_lattice(jobs_[k], sets_, lattice_mode_, lj)) {
            on_lattice[k] = 1;
            lattice_pairs += lj.pairs;
            lat_.push_back(lj);
        }
    }
    // z-runs of the lattice column sets, once for both work lists
    std::vector<float4> runs;
    std::vector<std::array<int, 3>> set_runs;
    for (size_t k = 0; k < jobs_.size(); ++k) {
        const PairJob& job = jobs_[k];
        if (on_lattice[k] || !job.col_runs || job.col_set < 0 || sets_[job.col_set].n_classes != 1) continue;
        bool seen = false;
        for (const auto& e : set_runs) seen = seen || e[0] == job.col_set;
        if (seen) continue;
        const PointSet& cs = sets_[job.col_set];
        const int r0 = (int)runs.size();
        const bool ok = lattice_runs(cs, cs.size() - cs.size() % 4, tile_j, runs);
        set_runs.push_back({job.col_set, r0, ok ? (int)runs.size() : -1});
    }
    auto runs_of = [&](int set) {
        for (const auto& e : set_runs)
            if (e[0] == set) return e[2] >= 0;
        return false;
    };
    // sets whose points some work item reads from the class-sorted buffer
    std::vector<char> placed(sets_.size(), weighted_ ? 1 : 0);
    for (size_t k = 0; k < jobs_.size(); ++k) {
        const PairJob& job = jobs_[k];
        if (on_lattice[k]) continue;
        placed[job.row_set] = 1;
        if (job.col_set >= 0 && !(job.col_runs && runs_of(job.col_set))) placed[job.col_set] = 1;
    }
    std::vector<float4> buf;
    std::vector<int32_t> perm;
    layout_.clear();
    for (size_t k = 0; k < sets_.size(); ++k)
        layout_.push_back(placed[k] ? layout_set(sets_[k], buf, weighted_ ? &perm : nullptr) : layout_counts(sets_[k]));
    // Narrower items for small structures: a launch of a few full-width
    // items leaves most CUs idle and runs as long as one item takes, so the
    // column width halves (down to kMinTile) until there are two items per
    // CU. The histograms are integer sums and do not depend on the tiling.
    constexpr int kMinTile = 128;
    const int target_items = 2 * ctx.prop.multiProcessorCount;
    for (int tile = tile_j;; tile = std::max(kMinTile, tile / 2)) {
        WorkList plain, split;
        plain.runs = split.runs = runs;
        plain.set_runs = split.set_runs = set_runs;
        plain.lattice_pairs = split.lattice_pairs = lattice_pairs;
        for (size_t k = 0; k < jobs_.size(); ++k) {
            const PairJob& job = jobs_[k];
            append_job(plain, job, sets_, layout_, n_rows_, tile, false, !on_lattice[k]);
            append_job(split, job, sets_, layout_, n_rows_, tile, true, !on_lattice[k]);
        }
        const bool use_split = prefer_split(padded_work(plain.items), item_rows(plain.items),
                                            padded_work(split.items), item_rows(split.items));
        wl_ = use_split ? std::move(split) : std::move(plain);
        if ((int)wl_.items.size() >= target_items || tile <= kMinTile) break;
    }
    sort_items(wl_.items);
    rows_ = item_rows(wl_.items);
    copies_ = pick_copies(nbins_, rows_);
    mp_ = plan_merge(wl_.items, n_rows_);

    // ---- device residency -------------------------------------------------
    hipStream_t s = ctx.stream();
    d_pts_.upload(buf.data(), buf.size(), s);
    d_orig_.clear();
    for (auto& ps : sets_) {
        d_orig_.emplace_back();
        d_orig_.back().upload(ps.orig.data(), ps.orig.size(), s);
    }
    d_items_.upload(wl_.items.data(), wl_.items.size(), s);
    if (!wl_.runs.empty()) d_runs_.upload(wl_.runs.data(), wl_.runs.size(), s);
    d_row_list_.upload(mp_.row_list.data(), mp_.row_list.size(), s);
    d_chunk_ptr_.upload(mp_.chunk_ptr.data(), mp_.chunk_ptr.size(), s);
    d_row_chunk_.upload(mp_.row_chunk.data(), mp_.row_chunk.size(), s);
    d_tmp_cnt_.ensure((size_t)mp_.n_chunks() * (nbins_ + 1));
    d_tmp_dsum_.ensure((size_t)mp_.n_chunks() * (nbins_ + 1));
    d_rowmaps_.upload(wl_.rowmaps.data(), wl_.rowmaps.size(), s);
    if (weighted_) {
        // weight-pass copies of the point buffers: the same positions with each
        // point's weight in the fourth slot
        std::vector<float4> wbuf(buf.size());
        for (size_t k = 0; k < sets_.size(); ++k) {
            const int base = layout_[k].base, n = sets_[k].size();
            for (int i = 0; i < n; ++i) {
                float4 v = buf[(size_t)base + i];
                v.w = weights_[k][perm[(size_t)base + i]];
                wbuf[(size_t)base + i] = v;
            }
        }
        d_pts_w_.upload(wbuf.data(), wbuf.size(), s);
        d_orig_w_.clear();
        for (size_t k = 0; k < sets_.size(); ++k) {
            std::vector<float4> o(sets_[k].orig);
            for (size_t i = 0; i < o.size(); ++i) o[i].w = weights_[k][i];
            d_orig_w_.emplace_back();
            d_orig_w_.back().upload(o.data(), o.size(), s);
        }
        d_wcnt_.ensure((size_t)n_rows_ * (nbins_ + 1));
        d_wsum_.ensure((size_t)n_rows_ * (nbins_ + 1));
        d_wstrip_.ensure((size_t)n_rows_ * (nbins_ + 1));
    }
    d_partial_.ensure(std::max<size_t>(wl_.items.size(), 1) * 2 * (nbins_ + 1));
    d_cnt_.ensure((size_t)n_rows_ * (nbins_ + 1));
    d_dsum_.ensure((size_t)n_rows_ * (nbins_ + 1));
    if (!lat_.empty()) {
        // twiddle tables exp(-2 pi i m / L), m < L, per job and axis
        std::vector<double2> tw;
        size_t cells = 0, cross = 0;
        int smax = 0;
        for (LatticeJob& j : lat_) {
            for (int a = 0; a < 3; ++a) {
                j.tw_off[a] = (int)tw.size();
                const int L = j.g.L[a];
                for (int k = 0; k < L; ++k) {
                    const double ang = -2.0 * 3.14159265358979323846 * (double)k / (double)L;
                    tw.push_back(make_double2(std::cos(ang), std::sin(ang)));
                }
            }
            const size_t c = (size_t)j.g.L[0] * j.g.L[1] * j.g.L[2];
            cells = std::max(cells, c);
            if (j.col_set >= 0) cross = std::max(cross, c);
            smax = std::max(smax, j.smax);
        }
        d_tw_.upload(tw.data(), tw.size(), s);
        d_lat_a_.ensure(cells);
        if (cross) d_lat_b_.ensure(cross);
        d_scount_.ensure((size_t)smax + 1);
        d_lat_flag_.ensure(2);
    }
    // the strips' launch descriptors (the buffers above stay where they are)
    std::vector<StripLaunch> st, wst;
    for (const StripSpec& sp : wl_.strips) {
        StripLaunch sl{d_orig_[sp.set_rows].p, d_orig_[sp.set_cols].p, sp.w, d_rowmaps_.p + sp.rowmap_off,
                       d_cnt_.p, d_dsum_.p, inv_w_, nbins_, nbins_ + 1};
        if (sets_[sp.set_rows].n_classes == 1 && sets_[sp.set_cols].n_classes == 1)
            sl.fixed_row = wl_.rowmaps[sp.rowmap_off];
        st.push_back(sl);
        if (weighted_) {
            // the weight pass: weights in the points' fourth slot, one row per job
            StripLaunch wl{d_orig_w_[sp.set_rows].p, d_orig_w_[sp.set_cols].p, sp.w, d_rowmaps_.p + sp.rowmap_off,
                           d_wcnt_.p, d_wstrip_.p, inv_w_, nbins_, nbins_ + 1};
            wl.fixed_row = wl_.rowmaps[sp.rowmap_off];
            wl.wmode = 1;
            wl.wscale = wscale_;
            wst.push_back(wl);
        }
    }
    build_strip_tables(st, strips_, s);
    build_strip_tables(wst, wstrips_, s);
    ctx.sync();
}

void PairEngine::build_strip_tables(const std::vector<StripLaunch>& all, StripTables& t, hipStream_t s) {
    std::vector<StripLaunch> a, l;
    std::vector<int64_t> fa{0}, fl{0};
    for (const StripLaunch& p : all) {
        if (strip_uses_lds(p)) {
            l.push_back(p);
            fl.push_back(fl.back() + strip_lds_blocks(p));
        } else if (const int64_t b = strip_table_blocks(p)) {
            a.push_back(p);
            fa.push_back(fa.back() + b);
        }
    }
    if (fa.back() > 0x7fffffff || fl.back() > 0x7fffffff)
        throw std::runtime_error("ausaxs_hip: strips beyond one launch's grid");
    t.n_atomic = (int)a.size();
    t.n_lds = (int)l.size();
    t.atomic_blocks = fa.back();
    t.lds_blocks = fl.back();
    if (!a.empty()) {
        t.atomic.upload(a.data(), a.size(), s);
        t.atomic_first.upload(fa.data(), fa.size(), s);
    }
    if (!l.empty()) {
        t.lds.upload(l.data(), l.size(), s);
        t.lds_first.upload(fl.data(), fl.size(), s);
    }
}

void PairEngine::launch_strip_tables(const StripTables& t, hipStream_t s, const char* what) const {
    check(launch_strip_table(t.atomic.p, t.atomic_first.p, t.n_atomic, t.atomic_blocks, s), what);
    check(launch_strip_lds_table(t.lds.p, t.lds_first.p, t.n_lds, t.lds_blocks, nbins_, s), what);
}

void PairEngine::enqueue(bool timed) {
    Context& ctx = Context::get();
    hipStream_t s = ctx.stream();
    if (timed) check(hipEventRecord(ev_[0], s), "hipEventRecord");
    PairLaunch pl{d_pts_.p, d_runs_.p, d_items_.p, (int)wl_.items.size(), d_partial_.p, inv_w_, nbins_, clamp_ ? 1 : 0, grid_,
                  rows_, copies_};
    check(launch_pair_hist(pl, s), "pair_hist_kernel");
    if (timed) check(hipEventRecord(ev_[1], s), "hipEventRecord");
    MergeLaunch ml{d_partial_.p, d_row_list_.p, d_chunk_ptr_.p, d_row_chunk_.p, mp_.n_chunks(),
                   n_rows_, nbins_, d_tmp_cnt_.p, d_tmp_dsum_.p, d_cnt_.p, d_dsum_.p};
    ml.direct = mp_.one_chunk_per_row() ? 1 : 0;  // small structures: one merge launch
    check(launch_merge(ml, s), "merge kernels");
    if (timed) check(hipEventRecord(ev_[3], s), "hipEventRecord");
    // lattice jobs add into the merged rows
    // the flag buffer holds two ints, cleared as one 8-byte word
    if (!lat_.empty())
        check(launch_lattice_zero_u64((unsigned long long*)d_lat_flag_.p, 1, s), "lattice_zero_u64_kernel");
    for (const LatticeJob& j : lat_) {
        const double2* tw[3] = {d_tw_.p + j.tw_off[0], d_tw_.p + j.tw_off[1], d_tw_.p + j.tw_off[2]};
        const bool self = j.col_set < 0;
        if (!self) {
            // the column set's transform first: the row set's x pass multiplies by it
            check(launch_lattice_zero_box(d_lat_b_.p, j.g, s), "lattice_zero_box_kernel");
            check(launch_lattice_fill(d_orig_[j.col_set].p, j.n_cols, j.g, d_lat_b_.p, s), "lattice_fill_kernel");
            check(launch_fft3_forward(d_lat_b_.p, j.g, tw, s), "fft_lines_kernel");
        }
        check(launch_lattice_zero_box(d_lat_a_.p, j.g, s), "lattice_zero_box_kernel");
        check(launch_lattice_fill(d_orig_[j.row_set].p, j.n_rows, j.g, d_lat_a_.p, s), "lattice_fill_kernel");
        check(launch_lattice_zero_u64(d_scount_.p, (size_t)j.smax + 1, s), "lattice_zero_u64_kernel");
        check(launch_fft3_correlate(d_lat_a_.p, j.g, tw, self ? nullptr : d_lat_b_.p, self ? j.n_rows : 0, d_scount_.p,
                                    j.smax, d_lat_flag_.p, s),
              "fft_lines_kernel");
        const size_t row = (size_t)j.out_row * (nbins_ + 1);
        check(launch_lattice_bins(d_scount_.p, j.smax, inv_w_, nbins_, d_cnt_.p + row, d_dsum_.p + row, s),
              "lattice_bins_kernel");
    }
    launch_strip_tables(strips_, s, "strip table kernels");
    if (weighted_) {
        // the weight pass over the same work lists (wmode), its own outputs
        PairLaunch wl{d_pts_w_.p, d_runs_.p, d_items_.p, (int)wl_.items.size(), d_partial_.p, inv_w_, nbins_, clamp_ ? 1 : 0,
                      grid_, rows_, copies_, 1, wscale_};
        check(launch_pair_hist(wl, s), "pair_hist_kernel (weights)");
        MergeLaunch wm{d_partial_.p, d_row_list_.p, d_chunk_ptr_.p, d_row_chunk_.p, mp_.n_chunks(),
                       n_rows_, nbins_, d_tmp_cnt_.p, d_tmp_dsum_.p, d_wcnt_.p, d_wsum_.p, 1};
        wm.direct = ml.direct;
        check(launch_merge(wm, s), "merge kernels (weights)");
        check(launch_lattice_zero_u64((unsigned long long*)d_wstrip_.p, (size_t)n_rows_ * (nbins_ + 1), s),
              "lattice_zero_u64_kernel");
        launch_strip_tables(wstrips_, s, "strip table kernels (weights)");
    }
    if (timed) check(hipEventRecord(ev_[2], s), "hipEventRecord");
    have_timing_ = timed;
}

void PairEngine::download(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum, WeightSums* wsum) {
    Context& ctx = Context::get();
    const size_t n = (size_t)n_rows_ * (nbins_ + 1);
    if (wsum && !weighted_) throw std::runtime_error("ausaxs_hip: weight sums of an engine without weights");
    // counts, delta sums (weight sums, lattice flag) gathered by one kernel
    // into a per-thread pinned buffer, then one wait
    CopyOut c{};
    auto add = [&](const void* src, size_t bytes) {
        c.src[c.n] = static_cast<const uint32_t*>(src);
        c.words[c.n++] = bytes / sizeof(uint32_t);
    };
    add(d_cnt_.p, n * sizeof(uint64_t));
    add(d_dsum_.p, n * sizeof(int64_t));
    if (wsum) {
        add(d_wsum_.p, n * sizeof(int64_t));
        add(d_wstrip_.p, n * sizeof(int64_t));
    }
    if (!lat_.empty()) add(d_lat_flag_.p, sizeof(int));
    size_t bytes = 0;
    for (int b = 0; b < c.n; ++b) bytes += c.words[b] * sizeof(uint32_t);
    thread_local MappedStage stage;
    const char* h = stage.ensure(bytes);
    c.dst = static_cast<uint32_t*>(stage.dev);
    check(launch_copy_out(c, ctx.stream()), "copy_out_kernel");
    ctx.sync();
    cnt.resize(n);
    dsum.resize(n);
    std::memcpy(cnt.data(), h, n * sizeof(uint64_t));
    std::memcpy(dsum.data(), h + n * sizeof(uint64_t), n * sizeof(int64_t));
    size_t off = 2 * n * sizeof(uint64_t);
    if (wsum) {
        wsum->items.resize(n);
        wsum->strips.resize(n);
        std::memcpy(wsum->items.data(), h + off, n * sizeof(int64_t));
        std::memcpy(wsum->strips.data(), h + off + n * sizeof(int64_t), n * sizeof(int64_t));
        off += 2 * n * sizeof(int64_t);
    }
    int flag = 0;
    if (!lat_.empty()) std::memcpy(&flag, h + off, sizeof(int));
    if (flag)
        throw std::runtime_error("ausaxs_hip: lattice correlation left a count off an integer (flag " +
                                 std::to_string(flag) + ")");
}

void PairEngine::run(std::vector<uint64_t>& cnt, std::vector<int64_t>& dsum, bool timed, WeightSums* wsum) {
    enqueue(timed);
    download(cnt, dsum, wsum);
}

EngineTiming PairEngine::timing() {
    EngineTiming t;
    t.pairs = wl_.pairs;
    t.lattice_pairs = wl_.lattice_pairs;
    t.kernel_pairs = kernel_pairs();
    t.n_items = (int)wl_.items.size();
    t.n_lattice = (int)lat_.size();
    const int64_t row_bytes = (int64_t)(nbins_ + 1) * 8;
    const int64_t entries = mp_.chunk_ptr.back();  // partial rows summed (row_list without its slack)
    // partial rows read; chunk sums written and read back (none when each row
    // is one chunk); rows written
    const int64_t chunk_rw = mp_.one_chunk_per_row() ? 0 : 4 * (int64_t)mp_.n_chunks();
    t.merge_bytes = row_bytes * (entries + chunk_rw + 2 * (int64_t)n_rows_);
    if (have_timing_) {
        check(hipEventSynchronize(ev_[2]), "hipEventSynchronize");
        check(hipEventElapsedTime(&t.pair_ms, ev_[0], ev_[1]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&t.merge_ms, ev_[1], ev_[3]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&t.total_ms, ev_[0], ev_[2]), "hipEventElapsedTime");
    }
    return t;
}

bool PairEngine::overflowed(const std::vector<uint64_t>& cnt) const {
    for (int r = 0; r < n_rows_; ++r)
        if (cnt[(size_t)r * (nbins_ + 1) + nbins_]) return true;
    return false;
}

}  // namespace ausaxs::gpu
