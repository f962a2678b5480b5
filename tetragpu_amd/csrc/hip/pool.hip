// Stream-ordered caching allocator for the library's device buffers, and a
// small pinned staging slab.
//
// hipMalloc/hipFree per mesh would dominate short jobs (the serving loop),
// and hipFree synchronises the whole device, serialising every stream.
// Here a freed block is recycled only after the stream that last used it
// has passed the point of the free (an event recorded at free time), so a
// block can move between meshes on different streams safely.  Blocks are
// size-classed (powers of two up to 1 MiB, then 2 MiB multiples) and reused
// when no more than twice the request.  Cached memory above a limit
// (TG_POOL_LIMIT_MB, default 64 GiB) is returned to the runtime; a failed
// hipMalloc first waits for pending frees and drops the cache.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

namespace {

struct Pending {
    void *p;
    size_t bytes;
    hipEvent_t ev;
};

struct Pool {
    std::mutex mu;
    std::multimap<size_t, void *> free_blocks;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> events;  // idle events for reuse
    size_t cached = 0;               // bytes in free_blocks + pending
    size_t limit = 0;

    Pool() {
        const char *env = std::getenv("TG_POOL_LIMIT_MB");
        limit = (env ? (size_t)std::strtoull(env, nullptr, 10) : (size_t)64 << 10) << 20;
    }

    void reap(bool wait) {
        size_t keep = 0;
        for (Pending &x : pending) {
            const hipError_t q = wait ? hipEventSynchronize(x.ev) : hipEventQuery(x.ev);
            if (q == hipSuccess) {
                free_blocks.emplace(x.bytes, x.p);
                events.push_back(x.ev);
            } else {
                (void)hipGetLastError();  // hipErrorNotReady is not an error here
                pending[keep++] = x;
            }
        }
        pending.resize(keep);
    }

    void drop_free() {
        for (auto &kv : free_blocks) {
            (void)hipFree(kv.second);
            cached -= kv.first;
        }
        free_blocks.clear();
    }
};

Pool &pool() {
    static Pool *p = new Pool();  // process lifetime (blocks live until exit)
    return *p;
}

size_t size_class(size_t bytes) {
    if (bytes <= (size_t)1 << 20) {
        size_t c = 512;
        while (c < bytes) c <<= 1;
        return c;
    }
    const size_t g = (size_t)2 << 20;
    return (bytes + g - 1) / g * g;
}

}  // namespace

namespace tg {

void *pool_alloc(size_t bytes, size_t *granted) {
    Pool &P = pool();
    const size_t want = size_class(std::max<size_t>(bytes, 1));
    std::lock_guard<std::mutex> g(P.mu);
    P.reap(false);
    auto it = P.free_blocks.lower_bound(want);
    if (it != P.free_blocks.end() && it->first <= 2 * want) {
        void *p = it->second;
        *granted = it->first;
        P.cached -= it->first;
        P.free_blocks.erase(it);
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        P.reap(true);
        P.drop_free();
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *granted = want;
    return p;
}

void pool_free(void *p, size_t bytes, hipStream_t stream) {
    if (!p) return;
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    hipEvent_t ev = nullptr;
    if (!P.events.empty()) {
        ev = P.events.back();
        P.events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        ev = nullptr;
    }
    if (ev && hipEventRecord(ev, stream) == hipSuccess) {
        P.pending.push_back({p, bytes, ev});
    } else {
        // no event: wait for the stream, then the block is free now
        (void)hipGetLastError();
        if (ev) P.events.push_back(ev);
        (void)hipStreamSynchronize(stream);
        P.free_blocks.emplace(bytes, p);
    }
    P.cached += bytes;
    if (P.cached > P.limit) {
        P.reap(true);
        // largest blocks first
        while (P.cached > P.limit / 2 && !P.free_blocks.empty()) {
            auto last = std::prev(P.free_blocks.end());
            (void)hipFree(last->second);
            P.cached -= last->first;
            P.free_blocks.erase(last);
        }
    }
}

void pool_trim() {
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    P.reap(true);
    P.drop_free();
}

size_t pool_cached_bytes() {
    Pool &P = pool();
    std::lock_guard<std::mutex> g(P.mu);
    return P.cached;
}

// ---- pinned host blocks for results (size-classed cache) ---------------------------
namespace {
struct HostPool {
    std::mutex mu;
    std::multimap<size_t, void *> free_blocks;
    std::map<void *, size_t> sizes;  // every live or cached block
    size_t cached = 0, limit = 0;
    HostPool() {
        const char *env = std::getenv("TG_HOST_POOL_LIMIT_MB");
        limit = (env ? (size_t)std::strtoull(env, nullptr, 10) : (size_t)16 << 10) << 20;
    }
};
HostPool &host_pool() {
    static HostPool *p = new HostPool();
    return *p;
}
}  // namespace

void *host_alloc(size_t bytes) {
    HostPool &P = host_pool();
    const size_t want = size_class(std::max<size_t>(bytes, 1));
    std::lock_guard<std::mutex> g(P.mu);
    auto it = P.free_blocks.lower_bound(want);
    if (it != P.free_blocks.end() && it->first <= 2 * want) {
        void *p = it->second;
        P.cached -= it->first;
        P.free_blocks.erase(it);
        return p;
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        for (auto &kv : P.free_blocks) {
            (void)hipHostFree(kv.second);
            P.sizes.erase(kv.second);
        }
        P.free_blocks.clear();
        P.cached = 0;
        if (hipHostMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    P.sizes[p] = want;
    return p;
}

void host_free(void *p) {
    if (!p) return;
    HostPool &P = host_pool();
    std::lock_guard<std::mutex> g(P.mu);
    auto it = P.sizes.find(p);
    if (it == P.sizes.end()) return;
    P.free_blocks.emplace(it->second, p);
    P.cached += it->second;
    while (P.cached > P.limit && !P.free_blocks.empty()) {
        auto last = std::prev(P.free_blocks.end());
        (void)hipHostFree(last->second);
        P.sizes.erase(last->second);
        P.cached -= last->first;
        P.free_blocks.erase(last);
    }
}

// ---- pinned staging: 512-byte slots of one pinned slab per 64 slots ---------------
namespace {
std::mutex g_pin_mu;
std::vector<void *> g_pin_free;
}  // namespace

void *pinned_slot_alloc() {
    std::lock_guard<std::mutex> g(g_pin_mu);
    if (g_pin_free.empty()) {
        char *slab = nullptr;
        if (hipHostMalloc((void **)&slab, 512 * 64) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        for (int i = 63; i >= 0; i--) g_pin_free.push_back(slab + 512 * i);
    }
    void *p = g_pin_free.back();
    g_pin_free.pop_back();
    return p;
}

void pinned_slot_free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(g_pin_mu);
    g_pin_free.push_back(p);
}

}  // namespace tg
