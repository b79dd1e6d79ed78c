// Caching allocator (see pool.hpp).
#include "pool.hpp"

#include "engine.hpp"

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace ausaxs::gpu {

size_t pool_size_class(size_t bytes) {
    size_t b = std::max<size_t>(bytes, 256);
    size_t p = 1;
    while (p < b) p <<= 1;  // smallest power of two >= b
    if (p <= (size_t(1) << 20)) return p;
    const size_t step = p / 8;  // b lies in (p/2, p]: quarter steps of p/2
    return (b + step - 1) / step * step;
}

namespace {

struct Block {
    void* p = nullptr;
    hipEvent_t ev = nullptr;       // recorded at release; null: nothing to wait for
    hipStream_t stream = nullptr;  // the stream ev was recorded on
};

// the block's last work is done (a query that finds it pending leaves
// hipErrorNotReady as the thread's last error, which is cleared here so the
// next kernel launch's error check does not see it)
bool block_done(const Block& b) {
    if (!b.ev) return true;
    const hipError_t e = hipEventQuery(b.ev);
    if (e == hipSuccess) return true;
    (void)hipGetLastError();
    return false;
}

struct Arena {
    bool host = false;
    std::unordered_map<void*, size_t> live;              // block -> size class
    std::unordered_map<size_t, std::vector<Block>> free;  // size class -> free blocks
    std::vector<std::pair<void*, size_t>> parked;         // released during graph capture (block, class)
    size_t held = 0, in_use = 0;
    int64_t mallocs = 0;
};

class Pool {
  public:
    static Pool& get() {
        // never destroyed: blocks and events stay valid until the process ends
        static Pool* pool = new Pool;
        return *pool;
    }

    // Reuse order: a free block whose last work is done, then one released on
    // this thread's stream (device memory: stream order alone keeps the new
    // work behind the old), then a new block; a block whose release is still
    // pending on another stream is taken (with a wait for that stream's event)
    // only when HIP has no memory left. Waiting on another stream would put
    // this request behind whatever that stream had queued: a small serving
    // request behind a large one's pair kernel.
    void* alloc(Arena& a, size_t bytes) {
        const size_t cls = pool_size_class(bytes);
        const hipStream_t s = Context::get().stream();
        Block blk;
        if (take(a, cls, s, false, blk)) return ready(a, s, blk);
        void* p = raw_alloc(a, cls);
        if (!p && take(a, cls, s, true, blk)) return ready(a, s, blk);
        if (!p) {
            trim();
            p = raw_alloc(a, cls);
        }
        if (!p)
            throw std::runtime_error(std::string("ausaxs_hip: out of ") + (a.host ? "pinned host" : "device") +
                                     " memory for " + std::to_string(cls) + " bytes");
        std::lock_guard<std::mutex> lk(m_);
        a.live[p] = cls;
        a.held += cls;
        a.in_use += cls;
        ++a.mallocs;
        return p;
    }

    // a free block of class cls: done or released on stream s (any = false),
    // or any free block (any = true); marks it handed out
    bool take(Arena& a, size_t cls, hipStream_t s, bool any, Block& blk) {
        std::lock_guard<std::mutex> lk(m_);
        auto it = a.free.find(cls);
        if (it == a.free.end() || it->second.empty()) return false;
        std::vector<Block>& v = it->second;
        int pick = -1;
        if (any) pick = (int)v.size() - 1;
        for (int i = (int)v.size() - 1; i >= 0 && pick < 0; --i)
            if (block_done(v[i])) pick = i;
        for (int i = (int)v.size() - 1; i >= 0 && pick < 0; --i)
            if (v[i].stream == s) pick = i;
        if (pick < 0) return false;
        blk = v[pick];
        v.erase(v.begin() + pick);
        a.live[blk.p] = cls;
        a.in_use += cls;
        ++reuses_;
        return true;
    }

    // makes a reused block safe for new work on stream s
    void* ready(Arena& a, hipStream_t s, Block& blk) {
        if (blk.ev) {
            if (a.host) check(hipEventSynchronize(blk.ev), "hipEventSynchronize");  // the host writes it
            else if (blk.stream != s) check(hipStreamWaitEvent(s, blk.ev, 0), "hipStreamWaitEvent");
            give_event(blk.ev);
        }
        return blk.p;
    }

    void release(Arena& a, void* p) {
        if (!p) return;
        {
            std::lock_guard<std::mutex> lk(m_);
            if (!a.live.count(p)) return;  // not a pool block
        }
        hipStream_t s = Context::get().stream();
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        const bool capturing = hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone;
        hipEvent_t ev = nullptr;
        if (!capturing) {
            ev = take_event();
            if (hipEventRecord(ev, s) != hipSuccess) {
                give_event(ev);
                ev = nullptr;
                (void)hipStreamSynchronize(s);
            }
        }
        std::lock_guard<std::mutex> lk(m_);
        auto it = a.live.find(p);
        if (it == a.live.end()) {  // released twice meanwhile
            if (ev) events_.push_back(ev);
            return;
        }
        const size_t cls = it->second;
        a.live.erase(it);
        a.in_use -= cls;
        if (capturing) {
            // an event record would join the captured graph: keep the block
            // out of circulation until the next trim
            a.parked.emplace_back(p, cls);
            return;
        }
        a.free[cls].push_back(Block{p, ev, ev ? s : nullptr});
    }

    void trim() {
        // every free block's last work must be done before HIP takes it back
        (void)hipDeviceSynchronize();
        std::lock_guard<std::mutex> lk(m_);
        for (Arena* a : {&dev_, &host_}) {
            for (auto& [cls, blocks] : a->free) {
                for (Block& b : blocks) {
                    if (a->host) (void)hipHostFree(b.p);
                    else (void)hipFree(b.p);
                    a->held -= cls;
                    if (b.ev) events_.push_back(b.ev);
                }
                blocks.clear();
            }
            for (auto [p, cls] : a->parked) {
                if (a->host) (void)hipHostFree(p);
                else (void)hipFree(p);
                a->held -= cls;
            }
            a->parked.clear();
        }
    }

    PoolStats stats() {
        std::lock_guard<std::mutex> lk(m_);
        PoolStats s;
        s.device_held = dev_.held;
        s.device_in_use = dev_.in_use;
        s.host_held = host_.held;
        s.host_in_use = host_.in_use;
        s.device_mallocs = dev_.mallocs;
        s.host_mallocs = host_.mallocs;
        s.reuses = reuses_;
        return s;
    }

    Arena dev_, host_;

  private:
    Pool() { host_.host = true; }
    std::mutex m_;
    std::vector<hipEvent_t> events_;
    int64_t reuses_ = 0;

    void* raw_alloc(Arena& a, size_t cls) {
        void* p = nullptr;
        const hipError_t e = a.host ? hipHostMalloc(&p, cls, hipHostMallocDefault) : hipMalloc(&p, cls);
        if (e != hipSuccess) {
            (void)hipGetLastError();  // clear the sticky error of the failed allocation
            return nullptr;
        }
        return p;
    }
    hipEvent_t take_event() {
        {
            std::lock_guard<std::mutex> lk(m_);
            if (!events_.empty()) {
                hipEvent_t e = events_.back();
                events_.pop_back();
                return e;
            }
        }
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        return e;
    }
    void give_event(hipEvent_t e) {
        std::lock_guard<std::mutex> lk(m_);
        events_.push_back(e);
    }
};

}  // namespace

void* pool_alloc_device(size_t bytes) { return Pool::get().alloc(Pool::get().dev_, bytes); }
void pool_free_device(void* p) { Pool::get().release(Pool::get().dev_, p); }
void* pool_alloc_host(size_t bytes) { return Pool::get().alloc(Pool::get().host_, bytes); }
void pool_free_host(void* p) { Pool::get().release(Pool::get().host_, p); }
void pool_trim() { Pool::get().trim(); }
PoolStats pool_stats() { return Pool::get().stats(); }

}  // namespace ausaxs::gpu
