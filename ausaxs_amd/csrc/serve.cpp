// Serving loop (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace ausaxs::gpu {

struct Server::Entry {
    uint64_t key = 0;
    Mode mode = Mode::FFExplicit;
    HistSettings s;
    double z_exv = 1;
    std::vector<double> atoms, waters;  // the exact content, compared on every hit
    std::vector<int32_t> ff;
    std::unique_ptr<ResidentStructure> st;
    bool busy = false;
    bool in_cache = false;
    size_t bytes = 0;  // device bytes of its own engine (idle entries count them in cached_bytes)
    std::list<std::shared_ptr<Entry>>::iterator lru;
};

namespace {

// FNV-1a over the bytes of a value sequence
uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 0x100000001b3ull;
    return h;
}

// estimated grid dummies of n atoms in a box (1 A^3 per lattice point)
double dummies_estimate(const double* atoms_xyz, int n) {
    if (n <= 0) return 0;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = std::numeric_limits<double>::max(), hi[k] = -lo[k];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = atoms_xyz[3 * (size_t)i + k];
            lo[k] = std::min(lo[k], v), hi[k] = std::max(hi[k], v);
        }
    double vol = 1;
    for (int k = 0; k < 3; ++k) vol *= (hi[k] - lo[k]) + 2 * 2.15 + 2;
    return std::min(vol, 82.0 * n);
}

}  // namespace

Server::Server(int slots, size_t budget_bytes) {
    if (slots < 1 || slots > 16) throw std::runtime_error("ausaxs_hip: server slots must be in [1, 16]");
    if (budget_bytes == 0) throw std::runtime_error("ausaxs_hip: server memory budget must be positive");
    Context& ctx = Context::get();
    (void)ctx;
    st_.budget = budget_bytes;
    streams_.resize(slots);
    for (auto& s : streams_) check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    for (int k = 0; k < slots; ++k) workers_.emplace_back([this, k] { worker(k); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(m_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : workers_) t.join();
    // the resident structures' buffers go back to the pool before the streams end
    (void)hipDeviceSynchronize();
    lru_.clear();
    cache_.clear();
    for (auto& s : streams_) (void)hipStreamDestroy(s);
}

size_t Server::footprint(Mode mode, const double* atoms_xyz, int n, const double* waters_xyz, int m, int nq,
                         const HistSettings& s) {
    // bounding box of everything (bins) and of the atoms (grid dummies)
    double lo[3], hi[3], alo[3], ahi[3];
    for (int k = 0; k < 3; ++k) lo[k] = alo[k] = std::numeric_limits<double>::max(), hi[k] = ahi[k] = -lo[k];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = atoms_xyz[3 * (size_t)i + k];
            alo[k] = std::min(alo[k], v), ahi[k] = std::max(ahi[k], v);
        }
    for (int k = 0; k < 3; ++k) lo[k] = alo[k], hi[k] = ahi[k];
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = waters_xyz[3 * (size_t)i + k];
            lo[k] = std::min(lo[k], v), hi[k] = std::max(hi[k], v);
        }
    double diag2 = 0;
    for (int k = 0; k < 3; ++k)
        if (n + m > 0) diag2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
    const double nb = std::min<double>(s.bin_count, std::sqrt(diag2) / s.bin_width + 4);
    const double x = mode == Mode::FFGrid ? dummies_estimate(atoms_xyz, n) : 0.0;
    const double P = n + m + x;
    // lattice correlation of the dummies: one transform grid of at least twice
    // the padded atom box per axis, and the squared-distance counts
    double cells = 1, smax = 0;
    bool fits = mode == Mode::FFGrid && s.lattice != 0 && n > 0;
    if (fits)
        for (int k = 0; k < 3; ++k) {
            const double E = std::ceil(ahi[k] - alo[k]) + 2 * 3 + 1;
            const int L = E <= kLatticeMaxL ? lattice_length((int)E) : -1;
            fits = fits && L > 0;
            const double Lk = L > 0 ? L : 2 * E;
            cells *= Lk;
            smax += (E - 1) * (E - 1);
        }
    // pair-kernel rows: atoms and waters, and the dummies unless the lattice
    // takes their block. With the lattice on and the box within its lines it
    // declines only dummy sets below 1e8 pairs (14,143 points) or below 64
    // pairs per transform cell (sqrt(128 cells) points), which then are rows.
    const double dummy_rows = fits ? std::min(x, std::max(14143.0, std::sqrt(128.0 * cells))) : x;
    const double rows_pts = n + m + dummy_rows;
    const double C = 16;
    const double tile = s.tile_j > 0 ? s.tile_j : kDefaultTileJ;
    // columns counted as points (a lattice set's z-runs are fewer)
    const double items = (rows_pts / 512 + C) * (P / tile + C);
    const double rows = 15 * 15 + 32;
    const double cell = 8 * (nb + 1);
    double bytes = 2 * 16 * P                          // class-sorted + original-order points
                   + items * (32 + 2 * cell)           // work items + partial slots
                   + (items / 16 + rows) * 2 * cell    // merge chunk sums
                   + rows * 2 * cell                   // counts + delta sums
                   + rows * nb * (8 + 6) + nq * nb * 6 + rows * nq * 8 + nq * 8;  // Debye operands
    if (fits) bytes += cells * 16 + smax * 8;
    return (size_t)(bytes * 1.1) + (1u << 20);
}

int64_t Server::submit(Mode mode, const double* atoms_xyz, const int32_t* ff, int n, const double* waters_xyz,
                       int m, const double* q, int nq, const HistSettings& s, double z_exv) {
    if (mode != Mode::FFAvg && mode != Mode::FFExplicit && mode != Mode::FFGrid)
        throw std::runtime_error("ausaxs_hip: the server scores FF-avg, FF-explicit and FF-grid requests");
    if (n < 0 || m < 0 || nq < 0) throw std::runtime_error("ausaxs_hip: negative sizes");
    for (int i = 0; i < n; ++i)
        if (ff[i] < 1 || ff[i] >= kMaxFF) throw std::runtime_error("ausaxs_hip: atom form factor type outside [1, 14]");
    auto r = std::make_shared<Request>();
    r->mode = mode;
    r->atoms.assign(atoms_xyz, atoms_xyz + 3 * (size_t)n);
    r->ff.assign(ff, ff + n);
    r->waters.assign(waters_xyz, waters_xyz + 3 * (size_t)m);
    r->q.assign(q, q + nq);
    r->s = s;
    r->z_exv = z_exv;
    r->bytes = footprint(mode, atoms_xyz, n, waters_xyz, m, nq, s);
    // pair-kernel pairs: everything but the grid's dummy-dummy block (lattice)
    const double pw = (double)n + m, x = mode == Mode::FFGrid ? dummies_estimate(atoms_xyz, n) : 0.0;
    r->est_pairs = pw * (pw - 1) / 2 + pw * x;
    uint64_t h = 0xcbf29ce484222325ull;
    const int md = (int)mode;
    h = fnv(h, &md, sizeof md);
    h = fnv(h, &s.bin_width, sizeof s.bin_width);
    h = fnv(h, &s.bin_count, sizeof s.bin_count);
    h = fnv(h, &z_exv, sizeof z_exv);
    h = fnv(h, r->atoms.data(), r->atoms.size() * sizeof(double));
    h = fnv(h, r->ff.data(), r->ff.size() * sizeof(int32_t));
    h = fnv(h, r->waters.data(), r->waters.size() * sizeof(double));
    r->key = h;
    std::lock_guard<std::mutex> lk(m_);
    if (r->bytes > st_.budget)
        throw std::runtime_error("ausaxs_hip: request needs an estimated " + std::to_string(r->bytes >> 20) +
                                 " MiB of device memory, more than the server budget of " +
                                 std::to_string(st_.budget >> 20) + " MiB");
    r->id = next_id_++;
    r->t_submit = clock::now();
    requests_[r->id] = r;
    pending_.push_back(r);
    ++st_.submitted;
    // every idle worker looks: the express slot may take what the others cannot
    cv_work_.notify_all();
    return r->id;
}

bool Server::mergeable(const Request& a, const Request& b) const {
    return a.s.bin_width == b.s.bin_width && a.s.tile_j == b.s.tile_j && a.s.lattice == b.s.lattice;
}

void Server::make_room(size_t need) {
    while (st_.in_use + st_.cached_bytes + need > st_.budget) {
        // least recently used idle entry
        auto it = lru_.end();
        bool found = false;
        while (it != lru_.begin()) {
            --it;
            if (!(*it)->busy) {
                found = true;
                break;
            }
        }
        if (!found) return;
        std::shared_ptr<Entry> e = *it;
        lru_.erase(it);
        auto range = cache_.equal_range(e->key);
        for (auto c = range.first; c != range.second; ++c)
            if (c->second == e) {
                cache_.erase(c);
                break;
            }
        e->in_cache = false;
        st_.cached_bytes -= e->bytes;
        ++st_.evictions;
    }
}

std::shared_ptr<Server::Entry> Server::acquire(const Request& r) {
    {
        std::lock_guard<std::mutex> lk(m_);
        auto range = cache_.equal_range(r.key);
        for (auto c = range.first; c != range.second; ++c) {
            Entry& e = *c->second;
            if (e.busy || e.mode != r.mode || e.s.bin_width != r.s.bin_width || e.s.bin_count != r.s.bin_count ||
                e.s.weighted_bins != r.s.weighted_bins || e.s.tile_j != r.s.tile_j || e.s.lattice != r.s.lattice ||
                e.z_exv != r.z_exv || e.atoms != r.atoms || e.ff != r.ff || e.waters != r.waters)
                continue;
            e.busy = true;
            st_.cached_bytes -= e.bytes;
            lru_.splice(lru_.begin(), lru_, e.lru);
            ++st_.cache_hits;
            return c->second;
        }
        ++st_.cache_misses;
    }
    // build the structure outside the lock (host set-up: packing, grid dummies, work lists)
    auto e = std::make_shared<Entry>();
    e->key = r.key;
    e->mode = r.mode;
    e->s = r.s;
    e->z_exv = r.z_exv;
    e->atoms = r.atoms;
    e->ff = r.ff;
    e->waters = r.waters;
    const int n = (int)r.ff.size(), m = (int)(r.waters.size() / 3);
    if (r.mode == Mode::FFGrid) {
        e->st = make_grid_structure(r.atoms.data(), r.ff.data(), n, r.waters.data(), m, r.s);
    } else {
        std::vector<float> a, w;
        pack_ff_points(r.atoms.data(), r.ff.data(), n, 0, a);
        pack_ff_points(r.waters.data(), nullptr, m, kWaterBin, w);
        e->st = std::make_unique<ResidentStructure>(r.mode, a.data(), n, w.data(), m, r.s, r.z_exv);
    }
    e->busy = true;
    std::lock_guard<std::mutex> lk(m_);
    cache_.emplace(e->key, e);
    lru_.push_front(e);
    e->lru = lru_.begin();
    e->in_cache = true;
    // beyond the entry cap: the least recently used idle entries go
    while (cache_.size() > kCacheEntries) {
        auto it = lru_.end();
        bool found = false;
        while (it != lru_.begin()) {
            --it;
            if (!(*it)->busy) {
                found = true;
                break;
            }
        }
        if (!found) break;
        std::shared_ptr<Entry> old = *it;
        lru_.erase(it);
        auto range = cache_.equal_range(old->key);
        for (auto c = range.first; c != range.second; ++c)
            if (c->second == old) {
                cache_.erase(c);
                break;
            }
        old->in_cache = false;
        st_.cached_bytes -= old->bytes;
        ++st_.evictions;
    }
    return e;
}

void Server::release(const std::shared_ptr<Entry>& e) {
    // the entry's device memory now counts as cached (its last work is done:
    // the request's results were copied back on this thread's stream)
    const size_t bytes = e->st->has_engine() ? e->st->engine().device_bytes() : 0;
    std::lock_guard<std::mutex> lk(m_);
    e->busy = false;
    e->bytes = bytes;
    if (e->in_cache) st_.cached_bytes += bytes;
}

void Server::run(const std::vector<std::shared_ptr<Request>>& batch) {
    // each member's resident structure; a member whose structure cannot be
    // built fails alone
    std::vector<std::shared_ptr<Entry>> es(batch.size());
    for (size_t k = 0; k < batch.size(); ++k) {
        try {
            es[k] = acquire(*batch[k]);
        } catch (const std::exception& ex) {
            batch[k]->error = ex.what();
        }
    }
    auto run_alone = [&](size_t k) {
        Request& r = *batch[k];
        try {
            ResidentStructure& st = *es[k]->st;
            auto h = st.calculate();
            r.pairs = st.pairs();
            r.kernel_pairs = st.engine().kernel_pairs();
            r.I = h->debye_transform(r.q);
        } catch (const std::exception& ex) {
            r.error = ex.what();
        }
    };
    std::vector<size_t> live;
    for (size_t k = 0; k < batch.size(); ++k)
        if (es[k]) live.push_back(k);
    if (live.size() == 1) {
        run_alone(live[0]);
    } else if (live.size() > 1) {
        std::vector<ResidentStructure*> ms;
        for (size_t k : live) ms.push_back(es[k]->st.get());
        bool merged_ok = false;
        try {
            int64_t pairs = 0, kpairs = 0;
            auto hs = ResidentStructure::calculate_merged(ms, &pairs, &kpairs);
            for (size_t i = 0; i < live.size(); ++i) {
                Request& r = *batch[live[i]];
                r.pairs = ms[i]->pairs();
                r.kernel_pairs = i == 0 ? kpairs : 0;  // the launch's pair-kernel pairs, once
                r.I = hs[i]->debye_transform(r.q);
            }
            merged_ok = true;
            std::lock_guard<std::mutex> lk(m_);
            ++st_.batches;
            st_.batched += (int64_t)live.size();
        } catch (const std::exception&) {
        }
        // a failure in the merged launch (one member's distances beyond its
        // bins, say) is found again member by member, so each gets its own error
        if (!merged_ok)
            for (size_t k : live) {
                batch[k]->error.clear();
                run_alone(k);
            }
    }
    for (auto& e : es)
        if (e) release(e);
}

void Server::worker(int slot) {
    Context::set_thread_stream(streams_[slot]);
    // the last of two or more slots admits only small requests
    const bool express = slot > 0 && slot == (int)streams_.size() - 1;
    for (;;) {
        std::vector<std::shared_ptr<Request>> batch;
        {
            std::unique_lock<std::mutex> lk(m_);
            // first come, first admitted: the head waits until its footprint fits
            // (idle resident structures give way to it); the express slot takes
            // the first small request, wherever it is queued
            auto pick = pending_.end();
            cv_work_.wait(lk, [&] {
                if (stop_) return true;
                pick = pending_.begin();
                if (express)
                    while (pick != pending_.end() && (*pick)->est_pairs >= kExpressPairs) ++pick;
                if (pick == pending_.end()) return false;
                const size_t need = (*pick)->bytes;
                make_room(need);
                return st_.in_use + st_.cached_bytes + need <= st_.budget ||
                       (st_.in_use == 0 && st_.in_use + need <= st_.budget);
            });
            if (stop_) {
                for (auto& p : pending_) {
                    p->error = "server stopped";
                    p->done = true;
                }
                pending_.clear();
                cv_done_.notify_all();
                return;
            }
            batch.push_back(*pick);
            auto next = pending_.erase(pick);
            if (express) ++st_.express;
            size_t bytes = batch[0]->bytes;
            double pairs = batch[0]->est_pairs;
            // small requests queued right behind a small first member join its launch
            if (pairs < kSmallPairs)
                while (next != pending_.end() && (int)batch.size() < kBatchMembers) {
                    const Request& nx = **next;
                    if (nx.est_pairs >= kSmallPairs || !mergeable(*batch[0], nx) || pairs + nx.est_pairs > kBatchPairs ||
                        st_.in_use + st_.cached_bytes + bytes + nx.bytes > st_.budget)
                        break;
                    bytes += nx.bytes;
                    pairs += nx.est_pairs;
                    batch.push_back(*next);
                    next = pending_.erase(next);
                }
            const auto now = clock::now();
            for (auto& r : batch) {
                st_.in_use += r->bytes;
                r->t_admit = now;
            }
            st_.peak_in_use = std::max(st_.peak_in_use, st_.in_use);
            st_.in_flight += (int)batch.size();
            st_.peak_in_flight = std::max(st_.peak_in_flight, st_.in_flight);
        }
        // the next head may fit beside this one
        cv_work_.notify_all();
        try {
            run(batch);
        } catch (const std::exception& e) {
            for (auto& r : batch)
                if (r->error.empty() && r->I.empty()) r->error = e.what();
        } catch (...) {
            for (auto& r : batch)
                if (r->error.empty() && r->I.empty()) r->error = "unknown error";
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            const auto now = clock::now();
            for (auto& r : batch) {
                st_.in_use -= r->bytes;
                --st_.in_flight;
                r->t_done = now;
                r->done = true;
                if (r->error.empty()) {
                    ++st_.completed;
                    st_.pairs += r->pairs;
                    st_.kernel_pairs += r->kernel_pairs;
                } else {
                    ++st_.failed;
                }
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

bool Server::wait(int64_t id, double timeout_s, std::vector<double>& I, double& latency_ms, double& queue_ms) {
    std::unique_lock<std::mutex> lk(m_);
    auto it = requests_.find(id);
    if (it == requests_.end()) throw std::runtime_error("ausaxs_hip: unknown request id " + std::to_string(id));
    std::shared_ptr<Request> r = it->second;
    const auto until = clock::now() + std::chrono::duration_cast<clock::duration>(
                                          std::chrono::duration<double>(std::max(0.0, timeout_s)));
    if (!cv_done_.wait_until(lk, until, [&] { return r->done; })) return false;
    requests_.erase(id);
    if (!r->error.empty()) throw std::runtime_error(r->error);
    I = std::move(r->I);
    latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    queue_ms = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
    return true;
}

ServeStats Server::stats() {
    std::lock_guard<std::mutex> lk(m_);
    return st_;
}

}  // namespace ausaxs::gpu
