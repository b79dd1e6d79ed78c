// Serving loop (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace ausaxs::gpu {

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
    double diag2 = 0, vol = 1;
    for (int k = 0; k < 3; ++k) {
        if (n + m > 0) diag2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
        vol *= n > 0 ? (ahi[k] - alo[k]) + 2 * 2.15 + 2 : 0;
    }
    const double nb = std::min<double>(s.bin_count, std::sqrt(diag2) / s.bin_width + 4);
    // dummies: at most the padded atom box (1 A^3 per lattice point)
    const double x = mode == Mode::FFGrid ? std::min(vol, 82.0 * n) : 0.0;
    const double P = n + m + x;
    const double C = 16;
    const double tile = s.tile_j > 0 ? s.tile_j : 8000;
    const double items = (P / 512 + C) * (P / tile + C);
    const double rows = 15 * 15 + 32;
    const double cell = 8 * (nb + 1);
    double bytes = 2 * 16 * P                          // class-sorted + original-order points
                   + items * (32 + 2 * cell)           // work items + partial slots
                   + (items / 16 + rows) * 2 * cell    // merge chunk sums
                   + rows * 2 * cell                   // counts + delta sums
                   + rows * nb * (8 + 6) + nq * nb * 6 + rows * nq * 8 + nq * 8;  // Debye operands
    if (mode == Mode::FFGrid && s.lattice != 0 && n > 0) {
        // lattice correlation of the dummies: one transform grid of at least
        // twice the padded atom box per axis, and the squared-distance counts
        double cells = 1, smax = 0;
        for (int k = 0; k < 3; ++k) {
            const double E = std::ceil(ahi[k] - alo[k]) + 2 * 3 + 1;
            double L = 1;
            while (L < 2 * E - 1) L *= 2;
            cells *= L;
            smax += (L / 2) * (L / 2);
        }
        bytes += cells * 16 + smax * 8;
    }
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
    cv_work_.notify_one();
    return r->id;
}

void Server::run(Request& r) {
    const int n = (int)r.ff.size(), m = (int)(r.waters.size() / 3);
    std::unique_ptr<ResidentStructure> st;
    if (r.mode == Mode::FFGrid) {
        st = make_grid_structure(r.atoms.data(), r.ff.data(), n, r.waters.data(), m, r.s);
    } else {
        std::vector<float> a, w;
        pack_ff_points(r.atoms.data(), r.ff.data(), n, 0, a);
        pack_ff_points(r.waters.data(), nullptr, m, kWaterBin, w);
        st = std::make_unique<ResidentStructure>(r.mode, a.data(), n, w.data(), m, r.s, r.z_exv);
    }
    auto h = st->calculate();
    r.pairs = st->engine().pairs();
    r.kernel_pairs = st->engine().kernel_pairs();
    r.I = h->debye_transform(r.q);
}

void Server::worker(int slot) {
    Context::set_thread_stream(streams_[slot]);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(m_);
            // first come, first admitted: the head waits until its footprint fits
            cv_work_.wait(lk, [&] {
                return stop_ || (!pending_.empty() && st_.in_use + pending_.front()->bytes <= st_.budget);
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
            r = pending_.front();
            pending_.pop_front();
            st_.in_use += r->bytes;
            st_.peak_in_use = std::max(st_.peak_in_use, st_.in_use);
            st_.peak_in_flight = std::max(st_.peak_in_flight, ++st_.in_flight);
            r->t_admit = clock::now();
        }
        // the next head may fit beside this one
        cv_work_.notify_one();
        try {
            run(*r);
        } catch (const std::exception& e) {
            r->error = e.what();
        } catch (...) {
            r->error = "unknown error";
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            st_.in_use -= r->bytes;
            --st_.in_flight;
            r->t_done = clock::now();
            r->done = true;
            if (r->error.empty()) {
                ++st_.completed;
                st_.pairs += r->pairs;
                st_.kernel_pairs += r->kernel_pairs;
            } else {
                ++st_.failed;
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
