// Shared helpers for the gfx950 kernels and their host launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace tg {

// Status codes of the C ABI (mirrored in include/tetra_c.h).
enum Status : int {
    OK = 0,
    ERR_ARG = -1,
    ERR_HIP = -2,
    ERR_ALLOC = -3,
    ERR_MESH = -4,      // invalid cell connectivity
    ERR_CAPACITY = -5,  // a row exceeded a kernel's supported size
    ERR_STATE = -6,     // relation requested before its inputs exist
    ERR_NO_DEVICE = -7,
};

void set_error(const std::string &msg);
const char *last_error();

// Record a HIP failure with its call site; returns ERR_HIP.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define TG_HIP(call)                                                        \
    do {                                                                    \
        hipError_t tg_e_ = (call);                                          \
        if (tg_e_ != hipSuccess)                                            \
            return ::tg::hip_fail(tg_e_, #call, __FILE__, __LINE__);        \
    } while (0)

#define TG_LAUNCH_CHECK(name)                                               \
    do {                                                                    \
        hipError_t tg_e_ = hipGetLastError();                               \
        if (tg_e_ != hipSuccess)                                            \
            return ::tg::hip_fail(tg_e_, name, __FILE__, __LINE__);         \
    } while (0)

constexpr int WAVE = 64;

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// Device-wide exclusive scan of n counts into out[0..n] (out[n] = total).
// `in` and `out` may alias.  `tmp` needs scan_tmp_bytes(n) bytes.
size_t scan_tmp_bytes(int64_t n);
int scan_exclusive_i32(const int32_t *in, int32_t *out, int64_t n, void *tmp,
                       hipStream_t s);
int scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp,
                       hipStream_t s);
// Widening scan: 32-bit counts into 64-bit offsets.
int scan_exclusive_i32_to_i64(const int32_t *in, int64_t *out, int64_t n,
                              void *tmp, hipStream_t s);

// Stream-ordered caching allocator (pool.hip): a freed block is reused only
// after `stream` has passed the free.  pool_alloc returns nullptr on failure
// and the usable size in *granted.
void *pool_alloc(size_t bytes, size_t *granted);
void pool_free(void *p, size_t bytes, hipStream_t stream);
void pool_trim();
size_t pool_cached_bytes();
// Pinned host blocks from a size-classed cache (results handed to callers).
void *host_alloc(size_t bytes);
void host_free(void *p);
// 512-byte pinned host slots for small device->host reads.
void *pinned_slot_alloc();
void pinned_slot_free(void *p);

// Fill kernels (used instead of hipMemsetAsync so launch sequences can be
// captured into graphs).
int fill_i32(int32_t *p, int32_t value, int64_t n, hipStream_t s);
int fill_u8(uint8_t *p, uint8_t value, int64_t n, hipStream_t s);

}  // namespace tg

// Device helpers --------------------------------------------------------
#ifdef __HIPCC__
namespace tgd {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive wave64 prefix sum.
template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

}  // namespace tgd
#endif
