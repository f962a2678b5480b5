// CPU stand-in for <hip/hip_runtime.h>, used only by the kernel emulator
// under tests/emu.  The library's HIP sources are compiled as host C++
// against this header: every workgroup runs on the CPU with one fiber per
// work-item, __syncthreads and the wave64 cross-lane operations are
// barriers between fibers, device memory is plain heap memory of the exact
// requested size (so AddressSanitizer sees every out-of-bounds access), and
// streams, events and graphs run their work synchronously, in order.
//
// Workgroups run one after another, or, with EMU_THREADS=k (k > 1), on k
// host threads at once, so that workgroups of one launch interleave their
// global-memory accesses as on the GPU; each thread has its own copy of the
// LDS arrays (__shared__ is thread_local) and its own fibers.  EMU_ORDER=
// reverse hands out the workgroups from the last one down.
//
// Limits: one-dimensional workgroups of at most 1024 work-items; a wave op
// runs over the lanes of the wave that reached it while the others wait at a
// workgroup barrier or have exited (wave ops in both arms of one divergent
// branch would merge).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#define __HIPCC__ 1
#define __global__
#define __device__
#define __host__
#define __shared__ static thread_local
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) ((const void *)&(x))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

struct alignas(16) int4 { int x, y, z, w; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(8) int2 { int x, y; };
struct alignas(8) uint2 { unsigned x, y; };
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(8) float2 { float x, y; };
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline float2 make_float2(float x, float y) { return float2{x, y}; }

namespace emu {
extern thread_local dim3 g_threadIdx, g_blockIdx, g_blockDim, g_gridDim;
void syncthreads();
int syncthreads_and(int p);
uint64_t wave_ballot(int p);
uint64_t wave_shfl_bits(uint64_t bits, int src_lane);
int lane();
void atomic_point();  // may hand the CPU to another work-item (EMU_YIELD)
}  // namespace emu

#define threadIdx (::emu::g_threadIdx)
#define blockIdx (::emu::g_blockIdx)
#define blockDim (::emu::g_blockDim)
#define gridDim (::emu::g_gridDim)

inline void __syncthreads() { emu::syncthreads(); }
inline int __syncthreads_and(int p) { return emu::syncthreads_and(p); }
inline int __lane_id() { return emu::lane(); }
inline unsigned long long __ballot(int p) { return emu::wave_ballot(p); }
inline int __any(int p) { return emu::wave_ballot(p) != 0; }
inline int __all(int p);

template <typename T>
inline T emu_from_bits(uint64_t b) {
    static_assert(sizeof(T) <= 8, "shuffle of a value wider than 64 bits");
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}
template <typename T>
inline uint64_t emu_to_bits(T v) {
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <typename T>
inline T __shfl(T v, int src, int width = 64) {
    const int l = emu::lane();
    const int s = (l & ~(width - 1)) | (src & (width - 1));
    return emu_from_bits<T>(emu::wave_shfl_bits(emu_to_bits(v), s));
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
    const int l = emu::lane();
    const int s = (l & ~(width - 1)) | ((l ^ mask) & (width - 1));
    return emu_from_bits<T>(emu::wave_shfl_bits(emu_to_bits(v), s));
}
template <typename T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
    const int l = emu::lane();
    const int base = l & ~(width - 1);
    const int s = (l - base) >= (int)d ? l - (int)d : l;
    (void)base;
    return emu_from_bits<T>(emu::wave_shfl_bits(emu_to_bits(v), s));
}
template <typename T>
inline T __shfl_down(T v, unsigned d, int width = 64) {
    const int l = emu::lane();
    const int base = l & ~(width - 1);
    const int s = (l - base) + (int)d < width ? l + (int)d : l;
    return emu_from_bits<T>(emu::wave_shfl_bits(emu_to_bits(v), s));
}
inline int __all(int p) {
    return emu::wave_ballot(!p) == 0;
}

inline int __ffs(int x) { return __builtin_ffs(x); }
inline int __ffs(unsigned x) { return __builtin_ffs((int)x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __clz(int x) { return x ? __builtin_clz((unsigned)x) : 32; }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }

template <typename A, typename B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    typedef typename std::common_type<A, B>::type C;
    return (C)a < (C)b ? (C)a : (C)b;
}
template <typename A, typename B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    typedef typename std::common_type<A, B>::type C;
    return (C)a > (C)b ? (C)a : (C)b;
}

// Atomics are real read-modify-writes (workgroups may run on several host
// threads); each is also a point where EMU_YIELD may switch work-items.
template <typename T, typename F>
inline T emu_rmw(T *p, F f) {
    emu::atomic_point();
    T o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (!__atomic_compare_exchange_n(p, &o, f(o), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return o;
}
template <typename T, typename U>
inline T atomicAdd(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)(o + (T)v); }); }
template <typename T, typename U>
inline T atomicSub(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)(o - (T)v); }); }
template <typename T, typename U>
inline T atomicOr(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)(o | (T)v); }); }
template <typename T, typename U>
inline T atomicAnd(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)(o & (T)v); }); }
template <typename T, typename U>
inline T atomicMax(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)v > o ? (T)v : o; }); }
template <typename T, typename U>
inline T atomicMin(T *p, U v) { return emu_rmw(p, [&](T o) { return (T)v < o ? (T)v : o; }); }
template <typename T, typename U>
inline T atomicExch(T *p, U v) { return emu_rmw(p, [&](T) { return (T)v; }); }
template <typename T, typename U, typename W>
inline T atomicCAS(T *p, U cmp, W v) {
    return emu_rmw(p, [&](T o) { return o == (T)cmp ? (T)v : o; });
}

// ---- runtime API ----------------------------------------------------------
typedef int hipError_t;
enum : int {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
};
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph *hipGraphExec_t;
typedef struct emu_graph_node *hipGraphNode_t;
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
};
enum hipStreamCaptureStatus {
    hipStreamCaptureStatusNone = 0,
    hipStreamCaptureStatusActive = 1,
    hipStreamCaptureStatusInvalidated = 2
};
enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0,
    hipStreamCaptureModeThreadLocal = 1,
    hipStreamCaptureModeRelaxed = 2
};
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u

const char *hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k,
                          hipStream_t s = nullptr);
hipError_t hipMemset(void *p, int v, size_t bytes);
hipError_t hipMemsetAsync(void *p, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags = 0);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err_node,
                               char *log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);

namespace emu {
// Runs (or, under stream capture, records) grid x block work-items of fn.
void submit(const char *name, dim3 grid, dim3 block, std::function<void()> fn);

template <typename... P, typename... A>
void launch(const char *name, dim3 grid, dim3 block, void (*k)(P...), A &&...a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    auto args = std::make_tuple((typename std::decay<P>::type)(std::forward<A>(a))...);
    submit(name, grid, block, [k, args]() { std::apply(k, args); });
}
}  // namespace emu

#define hipLaunchKernelGGL(K, G, B, SHM, STREAM, ...) \
    ::emu::launch(#K, ::dim3(G), ::dim3(B), K, ##__VA_ARGS__)
