// Caching allocator for the library's device and pinned host memory.
//
// hipFree and hipHostFree wait for the whole device, which serialises
// requests that run side by side on separate streams and finish at different
// times. Here a freed block goes back to a free list of its size class
// instead, with an event recorded on the freeing thread's stream (the stream
// its last work was queued on). An allocation of that class takes back a block
// whose event has completed, or else one freed on the allocating thread's own
// stream (stream order keeps the new work behind the old; pinned host memory,
// which the host writes directly, waits for the event on the host), and
// otherwise allocates a new block: a block whose event is still pending on
// another stream is reused (the allocating stream then waits for that event)
// only when HIP is out of memory, since the wait would queue this thread's
// work behind everything the other stream had queued. Blocks go back to HIP
// only in pool_trim, or when an allocation fails (the pool trims and tries
// once more).
//
// A block freed from another stream than the one its last work ran on is safe
// only if that work has finished: the library frees a structure's buffers
// after its results were copied back (stream synchronised), never while a
// kernel on another stream may still use them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ausaxs::gpu {

struct PoolStats {
    size_t device_held = 0;    // bytes of device blocks (in use + free)
    size_t device_in_use = 0;  // of which handed out
    size_t host_held = 0, host_in_use = 0;
    int64_t device_mallocs = 0;  // hipMalloc calls so far
    int64_t host_mallocs = 0;    // hipHostMalloc calls so far
    int64_t reuses = 0;          // allocations served from a free list
};

// size class of a request: powers of two up to 1 MiB, then quarter steps
// between powers of two (at most 25 % over the request)
size_t pool_size_class(size_t bytes);

void* pool_alloc_device(size_t bytes);
void pool_free_device(void* p);
void* pool_alloc_host(size_t bytes);
void pool_free_host(void* p);
// returns every free block to HIP (waits for the device)
void pool_trim();
PoolStats pool_stats();

}  // namespace ausaxs::gpu
