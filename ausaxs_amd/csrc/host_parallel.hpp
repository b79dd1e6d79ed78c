// Static split of an index range over host threads, for the O(points) passes
// of structure set-up (layouts, gathers, grid marking).
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace ausaxs::gpu {

// Number of host threads the set-up passes use: the hardware's, at most 16.
inline int host_threads() { return (int)std::clamp<unsigned>(std::thread::hardware_concurrency(), 1u, 16u); }

// Runs fn(begin, end) on contiguous chunks of [0, n), at least min_chunk long,
// one chunk per thread (the calling thread takes the first). fn must be safe
// to run concurrently on disjoint chunks.
template <class F>
void parallel_for(int64_t n, int64_t min_chunk, F&& fn) {
    if (n <= 0) return;
    const int t = (int)std::clamp<int64_t>(n / std::max<int64_t>(min_chunk, 1), 1, host_threads());
    if (t == 1) {
        fn(int64_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    for (int k = 1; k < t; ++k) pool.emplace_back([&, k] { fn(n * k / t, n * (k + 1) / t); });
    fn(int64_t(0), n / t);
    for (auto& th : pool) th.join();
}

}  // namespace ausaxs::gpu
