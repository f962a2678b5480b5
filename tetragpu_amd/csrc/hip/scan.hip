// Device-wide exclusive scan (reduce -> scan partials -> downsweep) and fill
// kernels.  Tiles of 4096 items per 256-thread workgroup; every tile is read
// twice and written once.
#include "common.h"

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename TIn, typename TAcc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(
    const TIn *__restrict__ in, int64_t n, TAcc *__restrict__ partials) {
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    TAcc sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int64_t g = base + k * SCAN_BLOCK + threadIdx.x;
        if (g < n) sum += (TAcc)in[g];
    }
    sum = tgd::wave_sum(sum);
    __shared__ TAcc ws[SCAN_BLOCK / 64];
    if (__lane_id() == 0) ws[threadIdx.x / 64] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        TAcc t = 0;
#pragma unroll
        for (int w = 0; w < SCAN_BLOCK / 64; w++) t += ws[w];
        partials[blockIdx.x] = t;
    }
}

// Single workgroup: exclusive scan of the per-tile sums, total to *total.
template <typename TAcc, typename TOut>
__global__ __launch_bounds__(1024) void k_scan_partials(
    TAcc *__restrict__ partials, int64_t nb, TOut *__restrict__ total) {
    __shared__ TAcc wsum[16];
    const int lane = __lane_id(), wid = threadIdx.x / 64;
    TAcc carry = 0;
    for (int64_t c = 0; c < nb; c += 1024) {
        const int64_t i = c + threadIdx.x;
        const TAcc x = i < nb ? partials[i] : (TAcc)0;
        const TAcc inc = tgd::wave_inclusive_sum(x);
        if (lane == 63) wsum[wid] = inc;
        __syncthreads();
        if (wid == 0) {
            TAcc w = lane < 16 ? wsum[lane] : (TAcc)0;
            w = tgd::wave_inclusive_sum(w);
            if (lane < 16) wsum[lane] = w;
        }
        __syncthreads();
        const TAcc before = wid > 0 ? wsum[wid - 1] : (TAcc)0;
        if (i < nb) partials[i] = carry + before + inc - x;
        const TAcc tile_total = wsum[15];
        __syncthreads();
        carry += tile_total;
    }
    if (threadIdx.x == 0) *total = (TOut)carry;
}

__device__ __forceinline__ int pad(int i) { return i + (i >> 5); }

template <typename TIn, typename TOut, typename TAcc>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_down(
    const TIn *in, TOut *out, int64_t n, const TAcc *__restrict__ partials) {
    __shared__ TAcc lds[SCAN_TILE + SCAN_TILE / 32];
    __shared__ TAcc wsum[SCAN_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    const int t = threadIdx.x, lane = __lane_id(), wid = t / 64;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const int i = k * SCAN_BLOCK + t;
        const int64_t g = base + i;
        lds[pad(i)] = g < n ? (TAcc)in[g] : (TAcc)0;
    }
    __syncthreads();
    TAcc v[SCAN_ITEMS];
    TAcc s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = lds[pad(t * SCAN_ITEMS + k)];
        s += v[k];
    }
    const TAcc inc = tgd::wave_inclusive_sum(s);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    TAcc run = partials[blockIdx.x] + inc - s;
    for (int w = 0; w < wid; w++) run += wsum[w];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        lds[pad(t * SCAN_ITEMS + k)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        const int i = k * SCAN_BLOCK + t;
        const int64_t g = base + i;
        if (g < n) out[g] = (TOut)lds[pad(i)];
    }
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, T value, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

template <typename TIn, typename TOut, typename TAcc>
int scan_impl(const TIn *in, TOut *out, int64_t n, void *tmp, hipStream_t s) {
    if (n < 0) return tg::ERR_ARG;
    if (n == 0) {
        hipLaunchKernelGGL((k_fill<TOut>), dim3(1), dim3(64), 0, s, out, (TOut)0, (int64_t)1);
        TG_LAUNCH_CHECK("k_fill");
        return tg::OK;
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    TAcc *partials = (TAcc *)tmp;
    hipLaunchKernelGGL((k_scan_reduce<TIn, TAcc>), dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, s,
                       in, n, partials);
    TG_LAUNCH_CHECK("k_scan_reduce");
    hipLaunchKernelGGL((k_scan_partials<TAcc, TOut>), dim3(1), dim3(1024), 0, s, partials, nb,
                       out + n);
    TG_LAUNCH_CHECK("k_scan_partials");
    hipLaunchKernelGGL((k_scan_down<TIn, TOut, TAcc>), dim3((unsigned)nb), dim3(SCAN_BLOCK), 0,
                       s, in, out, n, partials);
    TG_LAUNCH_CHECK("k_scan_down");
    return tg::OK;
}

}  // namespace

namespace tg {

size_t scan_tmp_bytes(int64_t n) {
    int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(nb < 1 ? 1 : nb) * sizeof(int64_t) + 256;
}

int scan_exclusive_i32(const int32_t *in, int32_t *out, int64_t n, void *tmp,
                       hipStream_t s) {
    return scan_impl<int32_t, int32_t, int32_t>(in, out, n, tmp, s);
}
int scan_exclusive_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp,
                       hipStream_t s) {
    return scan_impl<int64_t, int64_t, int64_t>(in, out, n, tmp, s);
}
int scan_exclusive_i32_to_i64(const int32_t *in, int64_t *out, int64_t n,
                              void *tmp, hipStream_t s) {
    return scan_impl<int32_t, int64_t, int64_t>(in, out, n, tmp, s);
}

int fill_i32(int32_t *p, int32_t value, int64_t n, hipStream_t s) {
    if (n <= 0) return OK;
    unsigned g = grid_for(n, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL((k_fill<int32_t>), dim3(g), dim3(256), 0, s, p, value, n);
    TG_LAUNCH_CHECK("k_fill<i32>");
    return OK;
}
int fill_u8(uint8_t *p, uint8_t value, int64_t n, hipStream_t s) {
    if (n <= 0) return OK;
    unsigned g = grid_for(n, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL((k_fill<uint8_t>), dim3(g), dim3(256), 0, s, p, value, n);
    TG_LAUNCH_CHECK("k_fill<u8>");
    return OK;
}

}  // namespace tg
