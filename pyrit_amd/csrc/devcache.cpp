// devcache.cpp - the library's device-memory cache: every build buffer
// (scratch and caller-owned outputs) and every resident server generation is
// taken from it and given back to it.
//
// A freed block stays mapped and is handed to the next request of its size
// class, so a steady stream of builds allocates nothing from the driver after
// the first ones (a fresh multi-GB allocation costs the driver's clearing of
// the memory, about a second for the 1024-chunk batch). Reuse is stream
// ordered: a block records an event on the stream that frees it, and a
// request on another stream waits for that event on the device before the
// block is used. Sizes are rounded to classes (powers of two up to 1 MiB,
// then eighths of a power of two), so a block serves any request of its
// class and the waste stays under 1/8.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "internal.h"

namespace pyr {
namespace {

struct Block {
    void* ptr = nullptr;
    uint64_t bytes = 0;          // class size
    int device = 0;
    hipStream_t freed_on = nullptr;
    hipEvent_t ready = nullptr;  // recorded on freed_on when the block was freed
};

struct Cache {
    std::mutex m;
    std::multimap<std::pair<int, uint64_t>, Block> free_blocks;  // (device, class) -> block
    std::unordered_map<void*, Block> live;
    uint64_t free_bytes = 0, live_bytes = 0, limit = 0;
};

Cache& cache() {
    static Cache* c = new Cache();  // lives until the process ends (blocks may be freed from atexit paths)
    return *c;
}

uint64_t size_class(uint64_t bytes) {
    if (bytes <= 256) return 256;
    if (bytes <= (1ull << 20)) {
        uint64_t c = 256;
        while (c < bytes) c <<= 1;
        return c;
    }
    uint64_t p = 1ull << 20;
    while ((p << 1) <= bytes) p <<= 1;
    const uint64_t step = p / 8;
    return (bytes + step - 1) / step * step;
}

// frees cached blocks, largest first, until at most `keep` free bytes remain
void trim_locked(Cache& c, uint64_t keep) {
    while (c.free_bytes > keep && !c.free_blocks.empty()) {
        auto it = c.free_blocks.end();
        --it;
        Block b = it->second;
        c.free_blocks.erase(it);
        c.free_bytes -= b.bytes;
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != b.device) (void)hipSetDevice(b.device);
        if (b.ready) (void)hipEventSynchronize(b.ready);
        if (b.ready) (void)hipEventDestroy(b.ready);
        (void)hipFree(b.ptr);
        if (cur != b.device) (void)hipSetDevice(cur);
    }
}

}  // namespace

hipError_t pool_alloc(void** p, uint64_t bytes, hipStream_t s) {
    *p = nullptr;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const uint64_t cls = size_class(bytes ? bytes : 1);
    Cache& c = cache();
    {
        std::lock_guard<std::mutex> lk(c.m);
        auto it = c.free_blocks.find({dev, cls});
        if (it != c.free_blocks.end()) {
            Block b = it->second;
            c.free_blocks.erase(it);
            c.free_bytes -= b.bytes;
            // work queued before the free on another stream must finish first
            if (b.ready && b.freed_on != s) {
                const hipError_t e = hipStreamWaitEvent(s, b.ready, 0);
                if (e != hipSuccess) {
                    c.free_blocks.emplace(std::make_pair(dev, cls), b);
                    c.free_bytes += b.bytes;
                    return e;
                }
            }
            c.live[b.ptr] = b;
            c.live_bytes += b.bytes;
            *p = b.ptr;
            return hipSuccess;
        }
    }
    Block b;
    b.bytes = cls;
    b.device = dev;
    hipError_t e = hipMalloc(&b.ptr, cls);
    if (e != hipSuccess) {
        // give the cached blocks back to the driver and try once more
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(c.m);
            trim_locked(c, 0);
        }
        e = hipMalloc(&b.ptr, cls);
        if (e != hipSuccess) return e;
    }
    std::lock_guard<std::mutex> lk(c.m);
    c.live[b.ptr] = b;
    c.live_bytes += b.bytes;
    *p = b.ptr;
    return hipSuccess;
}

void pool_free(void* p, hipStream_t s) {
    if (!p) return;
    Cache& c = cache();
    std::lock_guard<std::mutex> lk(c.m);
    auto it = c.live.find(p);
    if (it == c.live.end()) return;  // not ours
    Block b = it->second;
    c.live.erase(it);
    c.live_bytes -= b.bytes;
    if (!b.ready && hipEventCreateWithFlags(&b.ready, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        b.ready = nullptr;
    }
    b.freed_on = s;
    if (b.ready && hipEventRecord(b.ready, s) != hipSuccess) {
        // cannot order the reuse: wait here instead
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        b.freed_on = nullptr;
    }
    c.free_blocks.emplace(std::make_pair(b.device, b.bytes), b);
    c.free_bytes += b.bytes;
    if (!c.limit) {
        size_t fr = 0, total = 0;
        c.limit = hipMemGetInfo(&fr, &total) == hipSuccess ? total / 2 : (64ull << 30);
    }
    if (c.free_bytes > c.limit) trim_locked(c, c.limit);
}

}  // namespace pyr

extern "C" {

PYR_API void pyrhip_trim(void) {
    pyr::Cache& c = pyr::cache();
    std::lock_guard<std::mutex> lk(c.m);
    pyr::trim_locked(c, 0);
}

PYR_API PyrResult pyrhip_pool_stats(uint64_t out[4]) {
    pyr::diag_clear();
    if (!out) return pyr::fail(PYR_ERROR_INVALID_ARGUMENT, "null argument");
    pyr::Cache& c = pyr::cache();
    std::lock_guard<std::mutex> lk(c.m);
    out[0] = c.limit;
    out[1] = c.free_bytes + c.live_bytes;
    out[2] = c.live_bytes;
    out[3] = c.free_bytes;
    return PYR_OK;
}

}  // extern "C"
