// Serving loop for profile requests: continuous batching over device slots.
//
// Each request is one structure (FP64 atoms with form-factor types, waters)
// scored by one of the managers (FF-avg, FF-explicit or FF-grid,
// HistogramManagerMTFF{Avg,Explicit,Grid}::calculate_all) followed by the
// Debye transform on its own q grid. The server owns `slots` worker threads,
// each with its own HIP stream. Requests are admitted in arrival order as soon
// as a slot is idle and their estimated device footprint fits the memory
// budget; admitted requests run concurrently, so the kernels of requests that
// arrived while others were already on the GPU overlap with them. With two or
// more slots the last one is an express slot: it admits only small requests
// (below kExpressPairs pair-kernel pairs, about a millisecond of device time),
// the first one queued even if larger ones are ahead of it, so small requests
// do not wait for a slot while large ones hold all the others (they share the
// device with them once admitted).
//
// Three things keep the per-request cost down:
//   * resident structures: a structure (its float points, grid dummies, work
//     lists and, once it ran alone, its device buffers) stays cached across
//     requests, keyed by its exact content, mode and settings; a repeated
//     request skips the host set-up and the uploads and reruns the kernels on
//     the resident data. Idle entries are evicted, least recently used first,
//     when the budget needs their memory;
//   * merged launches: small requests that are queued together (each below
//     kSmallPairs pairs) run as one engine — their point sets side by side in
//     one buffer, all their work items in one pair launch, one merge — and
//     their results are split again;
//   * pooled memory: every device and pinned buffer comes from the caching
//     pool (pool.hpp), so steady serving makes no hipMalloc/hipFree calls.
// Every result is computed exactly as for the request alone (integer
// histograms, deterministic kernels), so it is the same alone, merged with
// others and in any mix.
#pragma once

#include "hist.hpp"

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace ausaxs::gpu {

struct ServeStats {
    int64_t submitted = 0, completed = 0, failed = 0;
    int in_flight = 0, peak_in_flight = 0;
    size_t in_use = 0, peak_in_use = 0, budget = 0;
    int64_t pairs = 0;         // pairs in the histograms of completed requests
    int64_t kernel_pairs = 0;  // of which evaluated by the pair kernel
    int64_t cache_hits = 0;    // requests served from a resident structure
    int64_t cache_misses = 0;  // requests whose structure was built
    int64_t batches = 0;       // merged launches
    int64_t batched = 0;       // requests that ran in a merged launch
    size_t cached_bytes = 0;   // device bytes held by idle resident structures
    int64_t evictions = 0;
    int64_t express = 0;       // requests admitted by the express slot
};

class Server {
  public:
    using clock = std::chrono::steady_clock;
    // requests below this many pairs may share a merged launch
    static constexpr double kSmallPairs = 3e8;
    // at most this many pairs and members per merged launch
    static constexpr double kBatchPairs = 1.5e9;
    static constexpr int kBatchMembers = 32;
    // requests below this many estimated pair-kernel pairs may take the
    // express slot
    static constexpr double kExpressPairs = 2e9;
    // resident structures kept at most (idle ones beyond are evicted)
    static constexpr size_t kCacheEntries = 256;

    Server(int slots, size_t budget_bytes);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // mode: Mode::FFAvg, FFExplicit or FFGrid. Inputs are copied.
    int64_t submit(Mode mode, const double* atoms_xyz, const int32_t* ff, int n, const double* waters_xyz, int m,
                   const double* q, int nq, const HistSettings& s, double z_exv);
    // Blocks until the request is done or timeout_s passes (false). On
    // success I receives nq values; latency = done - submit, queue = admit -
    // submit. A failed request throws its error. The request is forgotten
    // after a successful wait.
    bool wait(int64_t id, double timeout_s, std::vector<double>& I, double& latency_ms, double& queue_ms);
    ServeStats stats();

    // estimated device bytes of a request (admission)
    static size_t footprint(Mode mode, const double* atoms_xyz, int n, const double* waters_xyz, int m, int nq,
                            const HistSettings& s);

  private:
    struct Entry;  // a resident structure
    struct Request {
        int64_t id = 0;
        Mode mode = Mode::FFExplicit;
        std::vector<double> atoms, waters, q;
        std::vector<int32_t> ff;
        HistSettings s;
        double z_exv = 1;
        size_t bytes = 0;
        double est_pairs = 0;
        uint64_t key = 0;  // content hash of the structure
        clock::time_point t_submit, t_admit, t_done;
        std::vector<double> I;
        std::string error;
        int64_t pairs = 0, kernel_pairs = 0;
        bool done = false;
    };
    std::mutex m_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Request>> pending_;
    std::unordered_map<int64_t, std::shared_ptr<Request>> requests_;
    std::unordered_multimap<uint64_t, std::shared_ptr<Entry>> cache_;
    std::list<std::shared_ptr<Entry>> lru_;  // front: most recently used
    std::vector<std::thread> workers_;
    std::vector<hipStream_t> streams_;
    ServeStats st_;
    int64_t next_id_ = 1;
    bool stop_ = false;

    void worker(int slot);
    // run one request alone, or several small ones in one merged launch
    void run(const std::vector<std::shared_ptr<Request>>& batch);
    // a resident structure for r, marked busy (from the cache or built)
    std::shared_ptr<Entry> acquire(const Request& r);
    void release(const std::shared_ptr<Entry>& e);
    // evict idle entries (least recently used first) until `need` more bytes
    // fit the budget beside what is admitted; caller holds m_
    void make_room(size_t need);
    bool mergeable(const Request& a, const Request& b) const;
};

}  // namespace ausaxs::gpu
