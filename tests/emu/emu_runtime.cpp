// Fiber scheduler and runtime API of the CPU kernel emulator (see
// include/hip/hip_runtime.h), plus exact-size replacements of the library's
// device pool (pool.hip is not part of the emulated build): every device
// allocation is its own heap block, freed at once, so AddressSanitizer
// reports out-of-bounds and use-after-free accesses at the faulting line.
#include <hip/hip_runtime.h>
#include <ucontext.h>

#include <sys/mman.h>

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define EMU_ASAN 1
#include <sanitizer/common_interface_defs.h>
#endif
#endif

namespace emu {

thread_local dim3 g_threadIdx, g_blockIdx, g_blockDim, g_gridDim;

namespace {

enum State { RUN, WAIT_BLOCK, WAIT_WAVE, DONE };

constexpr size_t STACK = 256 << 10;
constexpr int MAX_ITEMS = 1024;

struct Fiber {
    ucontext_t ctx;
    char *stack = nullptr;
    State st = DONE;
    uint64_t slot = 0;
};

// Scheduler state is per host thread: a thread runs one workgroup at a time.
struct Fibers {
    Fiber f[MAX_ITEMS];
    ~Fibers() {
        for (Fiber &x : f)
            if (x.stack) munmap(x.stack, STACK);
    }
};
thread_local Fibers g_fibers;
thread_local Fiber *const g_f = g_fibers.f;
thread_local ucontext_t g_sched;
thread_local int g_cur = -1, g_n = 0;
thread_local uint64_t g_part[MAX_ITEMS / 64];  // lanes released together by the last wave barrier
thread_local const std::function<void()> *g_body = nullptr;
thread_local const char *g_name = "";
#ifdef EMU_ASAN
thread_local void *g_fake = nullptr;
thread_local const void *g_main_bottom = nullptr;  // the scheduler's stack, learned on the first switch
thread_local size_t g_main_size = 0;
#endif

void die(const char *what) {
    std::fprintf(stderr, "[emu] %s in kernel %s (block %u, item %d)\n", what, g_name,
                 g_blockIdx.x, g_cur);
    std::fflush(stderr);
    std::abort();
}

// EMU_YIELD=k: every k-th atomic operation hands the CPU to the next work-item
// of the block, so lock-free code sees other lanes' updates between its own
// read and write (lost compare-and-swaps, retried probes).
thread_local uint64_t g_atomics = 0;

void to_sched(State st) {
    g_f[g_cur].st = st;
#ifdef EMU_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, g_main_bottom, g_main_size);
#endif
    swapcontext(&g_f[g_cur].ctx, &g_sched);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(fake, &g_main_bottom, &g_main_size);
#endif
}

void entry() {
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &g_main_bottom, &g_main_size);
#endif
    (*g_body)();
    g_f[g_cur].st = DONE;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, g_main_bottom, g_main_size);
#endif
    setcontext(&g_sched);
}

void run_block() {
    const int n = g_n;
    for (int i = 0; i < n; i++) {
        Fiber &f = g_f[i];
        if (!f.stack) {
            f.stack = (char *)mmap(nullptr, STACK, PROT_READ | PROT_WRITE,
                                   MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
            if (f.stack == MAP_FAILED) die("fiber stack allocation failed");
        }
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = STACK;
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, entry, 0);
        f.st = RUN;
        f.slot = 0;
    }
    for (;;) {
        for (int i = 0; i < n; i++) {
            if (g_f[i].st != RUN) continue;
            g_cur = i;
            g_threadIdx = dim3((unsigned)i, 0, 0);
#ifdef EMU_ASAN
            __sanitizer_start_switch_fiber(&g_fake, g_f[i].stack, STACK);
#endif
            swapcontext(&g_sched, &g_f[i].ctx);
#ifdef EMU_ASAN
            __sanitizer_finish_switch_fiber(g_fake, nullptr, nullptr);
#endif
        }
        int live = 0, at_block = 0, runnable = 0;
        for (int i = 0; i < n; i++) {
            live += g_f[i].st != DONE;
            at_block += g_f[i].st == WAIT_BLOCK;
            runnable += g_f[i].st == RUN;
        }
        if (!live) break;
        if (runnable) continue;  // work-items that yielded run before any barrier opens
        bool released = false;
        if (at_block == live) {
            for (int i = 0; i < n; i++)
                if (g_f[i].st == WAIT_BLOCK) g_f[i].st = RUN;
            released = true;
        } else {
            for (int w = 0; w * 64 < n; w++) {
                int wl = 0, ww = 0;
                uint64_t mask = 0;
                for (int l = 0; l < 64 && w * 64 + l < n; l++) {
                    const State s = g_f[w * 64 + l].st;
                    wl += s != DONE;
                    if (s == WAIT_WAVE) {
                        ww++;
                        mask |= 1ull << l;
                    }
                }
                // Lanes of the wave that wait at a workgroup barrier have left the
                // (divergent) code the others are in: the op runs over the lanes
                // that reached it, as a wave with those lanes masked off would.
                if (wl && ww) {
                    for (int l = 0; l < 64 && w * 64 + l < n; l++)
                        if (g_f[w * 64 + l].st == WAIT_WAVE) g_f[w * 64 + l].st = RUN;
                    g_part[w] = mask;
                    released = true;
                }
            }
        }
        if (!released) {
            g_cur = -1;
            die("deadlock: a barrier or wave operation that not every live work-item reaches");
        }
    }
    g_cur = -1;
}

// ---- stream capture ----------------------------------------------------------
struct Capture {
    bool on = false;
    std::vector<std::function<void()>> ops;
};
Capture g_cap;

int env_int(const char *name, int dflt) {
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// Workgroups of the grid taken from a shared counter (in order, or from the
// last one down) until none is left.
void run_blocks(const char *name, dim3 grid, dim3 block, const std::function<void()> &fn,
                std::atomic<int64_t> &next, bool reverse) {
    g_name = name;
    g_body = &fn;
    g_blockDim = block;
    g_gridDim = grid;
    g_n = (int)block.x;
    const int64_t total = (int64_t)grid.x * grid.y * grid.z;
    for (;;) {
        int64_t k = next.fetch_add(1);
        if (k >= total) break;
        if (reverse) k = total - 1 - k;
        g_blockIdx = dim3((unsigned)(k % grid.x), (unsigned)(k / grid.x % grid.y),
                          (unsigned)(k / ((int64_t)grid.x * grid.y)));
        run_block();
    }
    g_body = nullptr;
}

// Persistent host threads for EMU_THREADS > 1 (each keeps its fibers).
struct Workers {
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::vector<std::thread> threads;
    uint64_t gen = 0;
    int busy = 0;
    bool quit = false;
    const char *name = nullptr;
    dim3 grid, block;
    const std::function<void()> *fn = nullptr;
    std::atomic<int64_t> next{0};
    bool reverse = false;

    explicit Workers(int n) {
        for (int i = 0; i < n; i++) threads.emplace_back([this]() { loop(); });
    }
    ~Workers() {
        {
            std::lock_guard<std::mutex> g(mu);
            quit = true;
        }
        cv.notify_all();
        for (auto &t : threads) t.join();
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&]() { return quit || gen != seen; });
            if (quit) return;
            seen = gen;
            g.unlock();
            run_blocks(name, grid, block, *fn, next, reverse);
            g.lock();
            if (--busy == 0) done_cv.notify_all();
        }
    }
    void run(const char *nm, dim3 gr, dim3 bl, const std::function<void()> &f, bool rev) {
        std::unique_lock<std::mutex> g(mu);
        name = nm;
        grid = gr;
        block = bl;
        fn = &f;
        reverse = rev;
        next = 0;
        busy = (int)threads.size();
        gen++;
        cv.notify_all();
        done_cv.wait(g, [&]() { return busy == 0; });
    }
};

void exec_grid(const char *name, dim3 grid, dim3 block, const std::function<void()> &fn) {
    if (block.y != 1 || block.z != 1 || block.x == 0 || block.x > (unsigned)MAX_ITEMS)
        die("unsupported workgroup shape");
    static const int threads = env_int("EMU_THREADS", 1);
    static const bool reverse = [] {
        const char *e = std::getenv("EMU_ORDER");
        return e && std::string(e) == "reverse";
    }();
    if (threads <= 1) {
        std::atomic<int64_t> next{0};
        run_blocks(name, grid, block, fn, next, reverse);
        return;
    }
    // one launch at a time (host threads of the library may launch together)
    static std::mutex launch_mu;
    static Workers *w = new Workers(threads);  // process lifetime
    std::lock_guard<std::mutex> g(launch_mu);
    w->run(name, grid, block, fn, reverse);
}

}  // namespace

int lane() { return g_cur & 63; }

void atomic_point() {
    static const int every = env_int("EMU_YIELD", 0);
    if (every > 0 && g_cur >= 0 && ++g_atomics % (uint64_t)every == 0) to_sched(RUN);
}

void syncthreads() { to_sched(WAIT_BLOCK); }

int syncthreads_and(int p) {
    g_f[g_cur].slot = p != 0;
    to_sched(WAIT_BLOCK);
    int r = 1;
    for (int i = 0; i < g_n; i++)
        if (g_f[i].st != DONE && !g_f[i].slot) r = 0;
    to_sched(WAIT_BLOCK);
    return r;
}

uint64_t wave_ballot(int p) {
    const int w = g_cur >> 6;
    g_f[g_cur].slot = p != 0;
    to_sched(WAIT_WAVE);
    uint64_t m = 0;
    const uint64_t part = g_part[w];
    for (int l = 0; l < 64; l++)
        if (((part >> l) & 1) && g_f[w * 64 + l].slot) m |= 1ull << l;
    to_sched(WAIT_WAVE);
    return m;
}

uint64_t wave_shfl_bits(uint64_t bits, int src) {
    const int w = g_cur >> 6;
    g_f[g_cur].slot = bits;
    to_sched(WAIT_WAVE);
    const uint64_t part = g_part[w];
    const uint64_t r = ((part >> src) & 1) ? g_f[w * 64 + src].slot : bits;
    to_sched(WAIT_WAVE);
    return r;
}

void submit(const char *name, dim3 grid, dim3 block, std::function<void()> fn) {
    if (g_cap.on) {
        g_cap.ops.push_back([name, grid, block, fn]() { exec_grid(name, grid, block, fn); });
        return;
    }
    exec_grid(name, grid, block, fn);
}

}  // namespace emu

// ---- runtime API -----------------------------------------------------------------
struct emu_stream { int id; };
struct emu_event { int id; };
struct emu_graph { std::vector<std::function<void()>> ops; };

namespace {
hipError_t g_last = hipSuccess;
// uninitialised device memory reads as this byte, not as zero
constexpr int POISON = 0xA5;

void run_or_record(std::function<void()> op) {
    if (emu::g_cap.on) emu::g_cap.ops.push_back(std::move(op));
    else op();
}
}  // namespace

const char *hipGetErrorString(hipError_t e) {
    switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotReady: return "hipErrorNotReady";
    default: return "hipErrorInvalidValue";
    }
}
hipError_t hipGetLastError() { hipError_t e = g_last; g_last = hipSuccess; return e; }
hipError_t hipPeekAtLastError() { return g_last; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipMalloc(void **p, size_t bytes) {
    *p = std::malloc(bytes ? bytes : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::memset(*p, POISON, bytes);
    return hipSuccess;
}
hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
    if (bytes) std::memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind,
                          hipStream_t) {
    run_or_record([=]() {
        if (bytes) std::memmove(dst, src, bytes);
    });
    return hipSuccess;
}
hipError_t hipMemset(void *p, int v, size_t bytes) { std::memset(p, v, bytes); return hipSuccess; }
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t) {
    run_or_record([=]() { std::memset(p, v, bytes); });
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { *s = new emu_stream{1}; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
// a fixed 16 GiB "device": budgets in tests are passed explicitly
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes) {
    *free_bytes = *total_bytes = (size_t)16 << 30;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    emu::g_cap.on = true;
    emu::g_cap.ops.clear();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    emu::g_cap.on = false;
    *g = new emu_graph{std::move(emu::g_cap.ops)};
    emu::g_cap.ops.clear();
    return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *st) {
    *st = emu::g_cap.on ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *,
                               size_t) {
    *e = new emu_graph{g->ops};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &op : e->ops) op();
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event{1}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}

// ---- the library's device pool, exact-size ------------------------------------------
namespace tg {

void *pool_alloc(size_t bytes, size_t *granted) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    *granted = bytes;
    return p;
}
void pool_free(void *p, size_t, hipStream_t) { std::free(p); }
void pool_trim() {}
size_t pool_cached_bytes() { return 0; }
void *host_alloc(size_t bytes) { return std::malloc(bytes ? bytes : 1); }
void host_free(void *p) { std::free(p); }
void *pinned_slot_alloc() { return std::malloc(512); }
void pinned_slot_free(void *p) { std::free(p); }

}  // namespace tg
