// VV + critical points, bucketed form (the default hot path on gfx950).
//
// Vertices are grouped in buckets of BV = 2^BV_LOG consecutive ids.  Every
// tet corner travels once, as a 12-byte record {v, o0, o1, o2} (four 24-bit
// fields: the corner's vertex and the tet's other three vertices), to v's
// bucket; then one workgroup per bucket builds the neighbour sets and link
// components of all its vertices in LDS.
//
// Routing is two-level so that every workgroup writes to a few hundred
// output lines at a time (which L2 merges) instead of to all buckets:
//   L1  k_l1_count / k_l1_scatter : chunks of tets -> <= 256 coarse bins
//       (a coarse bin = 2^FS consecutive buckets); per-(bin, chunk) counts
//       h1[bin][chunk], exclusive scan, LDS-atomic ranks.
//   L2  k_l2_count / k_l2_scatter : slices of each coarse bin -> buckets;
//       per-(bucket, slice) counts h2[bucket][slice], exclusive scan (which
//       also gives each bucket's record range), LDS-atomic ranks.
//   B   k_bkt_vv_cp : per bucket, a lock-free linear-probing hash set of 32
//       slots per vertex (neighbour ids), a 32-bit adjacency mask per slot
//       from the shared-tet link triangles, lower/upper masks from the
//       scalars (ballots), link components by bit-parallel BFS, rows sorted
//       by a 32-register sorting network and staged in LDS for coalesced
//       stores.  Vertices with more than 32 neighbours go to k_bkt_heavy
//       (one workgroup per vertex: LDS hash set, lock-free union-find).
// Then: scan of degrees, row compaction into the CSR, and the row-membership
// link pass for vertices whose shared-tet link has >= 2 components in a class
// (the shared-tet link graph is a subgraph of the reference's, so counts of 0
// or 1 are already final; critPoints.cu:206-246).
#include <algorithm>

#include "bucket.h"

namespace {

using tg::C_ERR;
using tg::C_FIX;
using tg::C_FIXBIG;
using tg::C_HEAVY;
using tg::C_MAXDEG;
using tg::HEAVY_CAP;

constexpr int SLOTS = 32;
constexpr uint32_t M24 = 0xFFFFFFu;

template <typename S>
__device__ __forceinline__ bool is_lower(S sa, S sv, int32_t a, int32_t v) {
    return (sa == sv) ? (a < v) : (sa < sv);
}

__device__ __forceinline__ uint32_t verdict(uint32_t nl, uint32_t nu) {
    if (nl >= 1 && nu == 0) return 2u;
    if (nl == 0 && nu >= 1) return 1u;
    if (nl == 1 && nu == 1) return 3u;
    return 4u;
}

__device__ __forceinline__ void wave_append(int32_t *list, int32_t *counter,
                                            int32_t v, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (!mask) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (int32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = v;
}

// ---- corner records ------------------------------------------------------------
struct Rec {
    uint32_t *w0, *w1, *w2;  // structure of arrays, n records each
};

__device__ __forceinline__ void rec_store(const Rec &r, int64_t i, uint32_t v, uint32_t a,
                                          uint32_t b, uint32_t c) {
    r.w0[i] = v | (a << 24);
    r.w1[i] = (a >> 8) | (b << 16);
    r.w2[i] = (b >> 16) | (c << 8);
}

__device__ __forceinline__ void rec_unpack(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t &v,
                                           uint32_t &a, uint32_t &b, uint32_t &c) {
    v = w0 & M24;
    a = (w0 >> 24) | ((w1 & 0xFFFFu) << 8);
    b = (w1 >> 16) | ((w2 & 0xFFu) << 16);
    c = w2 >> 8;
}

// ---- L1: chunks of tets -> coarse bins -------------------------------------
constexpr int L1_THREADS = 1024;
constexpr int MAX_COARSE = 256;
constexpr int MAX_FINE = 512;

__global__ __launch_bounds__(L1_THREADS) void k_l1_count(const int4 *__restrict__ tv4, int64_t T,
                                                          int64_t tpc, int cshift, int32_t nc,
                                                          int32_t nch, int32_t *__restrict__ h1) {
    constexpr int NW = L1_THREADS / 64;
    __shared__ int32_t h[NW][MAX_COARSE];  // one sub-histogram per wave
    int32_t *hw = h[threadIdx.x / 64];
    for (int i = threadIdx.x; i < NW * MAX_COARSE; i += L1_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x * tpc, t1 = min(T, t0 + tpc);
    for (int64_t t = t0 + threadIdx.x; t < t1; t += L1_THREADS) {
        const int4 q = tv4[t];
        atomicAdd(&hw[q.x >> cshift], 1);
        atomicAdd(&hw[q.y >> cshift], 1);
        atomicAdd(&hw[q.z >> cshift], 1);
        atomicAdd(&hw[q.w >> cshift], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += L1_THREADS) {
        int32_t sum = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) sum += h[w][i];
        h1[(int64_t)i * nch + blockIdx.x] = sum;
    }
}

// Staged scatter shared by both routing levels: a tile of records is
// counting-sorted by destination bin in LDS, then each bin's run is written
// with consecutive lanes to consecutive addresses.  A record's bin is
// recomputed from its vertex field, so a tile costs 12 bytes of LDS per
// record; larger tiles give longer runs per bin and fewer partial lines.
constexpr int STAGE1 = 8192;   // level-1 records per tile (2048 tets)
constexpr int STAGE2 = 12288;  // level-2 records per tile

template <int N>
struct Stage {
    uint32_t w0[N], w1[N], w2[N];
    int32_t cnt[MAX_FINE], toff[MAX_FINE], gbase[MAX_FINE];
    int32_t total;
};

// Exclusive scan of cnt[0..nbin) into toff, total into st.total (block-wide).
template <typename ST>
__device__ __forceinline__ void stage_scan(ST &st, int nbin) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int32_t carry = 0;
        for (int b0 = 0; b0 < nbin; b0 += 64) {
            const int i = b0 + lane;
            const int32_t x = i < nbin ? st.cnt[i] : 0;
            const int32_t inc = tgd::wave_inclusive_sum(x);
            if (i < nbin) st.toff[i] = carry + inc - x;
            carry += __shfl(inc, 63, 64);
        }
        if (lane == 0) st.total = carry;
    }
}

// Write the staged tile: consecutive slots of one bin go to consecutive
// addresses of that bin's range; the bin of a slot is (v >> shift) - base.
template <typename ST>
__device__ __forceinline__ void stage_flush(ST &st, Rec out, int nthreads, int shift,
                                            int32_t base) {
    const int32_t total = st.total;
    for (int slot = threadIdx.x; slot < total; slot += nthreads) {
        const uint32_t w0 = st.w0[slot];
        const int b = (int)((w0 & M24) >> shift) - base;
        const int64_t pos = (int64_t)st.gbase[b] + (slot - st.toff[b]);
        out.w0[pos] = w0;
        out.w1[pos] = st.w1[slot];
        out.w2[pos] = st.w2[slot];
    }
}

template <bool VTREC>
__global__ __launch_bounds__(L1_THREADS) void k_l1_scatter(const int4 *__restrict__ tv4, int64_t T,
                                                            int64_t tpc, int cshift, int32_t nc,
                                                            int32_t nch,
                                                            const int32_t *__restrict__ h1s,
                                                            Rec out) {
    constexpr int TILE = STAGE1 / 4;          // tets per tile
    constexpr int TPT = TILE / L1_THREADS;    // tets per thread
    __shared__ Stage<STAGE1> st;
    const int tid = threadIdx.x;
    for (int i = tid; i < nc; i += L1_THREADS) st.gbase[i] = h1s[(int64_t)i * nch + blockIdx.x];
    const int64_t t0 = (int64_t)blockIdx.x * tpc, t1 = min(T, t0 + tpc);
    int4 qn[TPT];
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int64_t t = t0 + k * L1_THREADS + tid;
        qn[k] = t < t1 ? tv4[t] : make_int4(0, 0, 0, 0);
    }
    for (int64_t tb = t0; tb < t1; tb += TILE) {
        for (int i = tid; i < nc; i += L1_THREADS) st.cnt[i] = 0;
        __syncthreads();
        int4 q[TPT];
        bool on[TPT];
        int r[TPT][4];
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            q[k] = qn[k];
            on[k] = tb + k * L1_THREADS + tid < t1;
            if (on[k]) {
                r[k][0] = atomicAdd(&st.cnt[q[k].x >> cshift], 1);
                r[k][1] = atomicAdd(&st.cnt[q[k].y >> cshift], 1);
                r[k][2] = atomicAdd(&st.cnt[q[k].z >> cshift], 1);
                r[k][3] = atomicAdd(&st.cnt[q[k].w >> cshift], 1);
            }
        }
        __syncthreads();
        stage_scan(st, nc);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            if (!on[k]) continue;
            const uint32_t x = q[k].x, y = q[k].y, z = q[k].z, w = q[k].w;
            auto put = [&](int bin, int rank, uint32_t v, uint32_t a, uint32_t b, uint32_t c) {
                const int slot = st.toff[bin] + rank;
                st.w0[slot] = v | (a << 24);
                st.w1[slot] = (a >> 8) | (b << 16);
                st.w2[slot] = (b >> 16) | (c << 8);
            };
            if constexpr (VTREC) {
                const uint32_t c0 = 4u * (uint32_t)(tb + k * L1_THREADS + tid);
                put(q[k].x >> cshift, r[k][0], x, c0 & M24, c0 >> 24, 0u);
                put(q[k].y >> cshift, r[k][1], y, (c0 + 1) & M24, (c0 + 1) >> 24, 0u);
                put(q[k].z >> cshift, r[k][2], z, (c0 + 2) & M24, (c0 + 2) >> 24, 0u);
                put(q[k].w >> cshift, r[k][3], w, (c0 + 3) & M24, (c0 + 3) >> 24, 0u);
            } else {
                put(q[k].x >> cshift, r[k][0], x, y, z, w);
                put(q[k].y >> cshift, r[k][1], y, x, z, w);
                put(q[k].z >> cshift, r[k][2], z, x, y, w);
                put(q[k].w >> cshift, r[k][3], w, x, y, z);
            }
        }
        __syncthreads();
        // next tile's tets are in flight during the flush
#pragma unroll
        for (int k = 0; k < TPT; k++) {
            const int64_t t = tb + TILE + k * L1_THREADS + tid;
            if (t < t1) qn[k] = tv4[t];
        }
        stage_flush(st, out, L1_THREADS, cshift, 0);
        __syncthreads();
        for (int i = tid; i < nc; i += L1_THREADS) st.gbase[i] += st.cnt[i];
    }
}

// ---- L2: slices of a coarse bin -> buckets ---------------------------------
constexpr int L2_THREADS = 1024;

__device__ __forceinline__ void slice_range(const int32_t *h1s, int32_t nch, int32_t c,
                                            int32_t s, int32_t S, int32_t &lo, int32_t &hi) {
    const int64_t rb = h1s[(int64_t)c * nch], re = h1s[(int64_t)(c + 1) * nch];
    lo = (int32_t)(rb + (re - rb) * s / S);
    hi = (int32_t)(rb + (re - rb) * (s + 1) / S);
}

__global__ __launch_bounds__(L2_THREADS) void k_l2_count(const uint32_t *__restrict__ w0,
                                                          const int32_t *__restrict__ h1s,
                                                          int32_t nch, int bv_log, int fs,
                                                          int32_t nb, int32_t *__restrict__ h2) {
    constexpr int NSUB = 8;  // sub-histograms, two waves each
    __shared__ int32_t h[NSUB][MAX_FINE];
    const int32_t c = blockIdx.x, s = blockIdx.y, S = gridDim.y;
    const int F = 1 << fs;
    int32_t *hw = h[(threadIdx.x / 64) % NSUB];
    for (int i = threadIdx.x; i < NSUB * MAX_FINE; i += L2_THREADS) (&h[0][0])[i] = 0;
    __syncthreads();
    int32_t lo, hi;
    slice_range(h1s, nch, c, s, S, lo, hi);
    const int32_t fb0 = c << fs;
    constexpr int U = 8;
    for (int32_t i0 = lo + threadIdx.x; i0 < hi; i0 += U * L2_THREADS) {
        uint32_t w[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int32_t i = i0 + k * L2_THREADS;
            w[k] = i < hi ? w0[i] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int k = 0; k < U; k++)
            if (w[k] != 0xFFFFFFFFu) atomicAdd(&hw[(int32_t)((w[k] & M24) >> bv_log) - fb0], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < F && fb0 + i < nb; i += L2_THREADS) {
        int32_t sum = 0;
#pragma unroll
        for (int w = 0; w < NSUB; w++) sum += h[w][i];
        h2[(int64_t)(fb0 + i) * S + s] = sum;
    }
}

__global__ __launch_bounds__(L2_THREADS) void k_l2_scatter(const Rec in,
                                                            const int32_t *__restrict__ h1s,
                                                            int32_t nch, int bv_log, int fs,
                                                            int32_t nb,
                                                            const int32_t *__restrict__ h2s,
                                                            Rec out) {
    __shared__ Stage<STAGE2> st;
    const int tid = threadIdx.x;
    const int32_t c = blockIdx.x, s = blockIdx.y, S = gridDim.y;
    const int32_t fb0 = c << fs;
    const int F = min(1 << fs, nb - fb0);
    for (int i = tid; i < F; i += L2_THREADS) st.gbase[i] = h2s[(int64_t)(fb0 + i) * S + s];
    int32_t lo, hi;
    slice_range(h1s, nch, c, s, S, lo, hi);
    constexpr int PER = STAGE2 / L2_THREADS;
    uint32_t a[PER], b[PER], d[PER];
    auto load = [&](int32_t ib) {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int32_t i = ib + k * L2_THREADS + tid;
            a[k] = 0xFFFFFFFFu;
            if (i < hi) {
                a[k] = in.w0[i];
                b[k] = in.w1[i];
                d[k] = in.w2[i];
            }
        }
    };
    load(lo);
    for (int32_t ib = lo; ib < hi; ib += STAGE2) {
        for (int i = tid; i < F; i += L2_THREADS) st.cnt[i] = 0;
        __syncthreads();
        int f[PER], r[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            f[k] = a[k] != 0xFFFFFFFFu ? (int)((a[k] & M24) >> bv_log) - fb0 : -1;
            if (f[k] >= 0) r[k] = atomicAdd(&st.cnt[f[k]], 1);
        }
        __syncthreads();
        stage_scan(st, F);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; k++)
            if (f[k] >= 0) {
                const int slot = st.toff[f[k]] + r[k];
                st.w0[slot] = a[k];
                st.w1[slot] = b[k];
                st.w2[slot] = d[k];
            }
        __syncthreads();
        load(ib + STAGE2);  // next tile's records are in flight during the flush
        stage_flush(st, out, L2_THREADS, bv_log, fb0);
        __syncthreads();
        for (int i = tid; i < F; i += L2_THREADS) st.gbase[i] += st.cnt[i];
    }
}

// ---- one-level routing: tiles of tets -> buckets through a window -----------
// When consecutive tets reference nearby vertex ids (meshes written in a
// spatial order, structured grids), the corners of one tile of S1_TILE tets
// fall into fewer than S1_WIN consecutive buckets.  Such a tile is counting-
// sorted in LDS by (bucket - first bucket of the tile) and every non-empty
// run takes its place in the bucket's range with one global atomic per
// (tile, bucket); the order of records inside a bucket does not matter to
// the bucket kernels.  A wider tile places each corner with its own global
// atomic.  k_s1_count counts the wide tiles; a mesh where they are more than
// 1/64 of all tiles is routed in two levels instead.
constexpr int S1_THREADS = 1024;     // one tet per thread
constexpr int S1_TILE = S1_THREADS;  // tets per tile
constexpr int S1_WIN = S1_THREADS;   // buckets per window (one per thread)
constexpr int S1_NSUB = 4;           // count sub-histograms
constexpr int S1_NW = S1_THREADS / 64;

// Smallest and largest bucket of the tile (threads without a tet pass
// INT32_MAX / -1).  Contains a __syncthreads.
__device__ __forceinline__ void tile_window(int32_t lo, int32_t hi, int32_t (&red)[2][S1_NW],
                                            int32_t &bmin, int32_t &bmax) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        lo = min(lo, __shfl_xor(lo, s, 64));
        hi = max(hi, __shfl_xor(hi, s, 64));
    }
    if (__lane_id() == 0) {
        red[0][threadIdx.x / 64] = lo;
        red[1][threadIdx.x / 64] = hi;
    }
    __syncthreads();
    bmin = red[0][0];
    bmax = red[1][0];
#pragma unroll
    for (int i = 1; i < S1_NW; i++) {
        bmin = min(bmin, red[0][i]);
        bmax = max(bmax, red[1][i]);
    }
}

__global__ __launch_bounds__(S1_THREADS) void k_s1_count(const int4 *__restrict__ tv4, int64_t T,
                                                          int64_t tpc, int bv_log,
                                                          int32_t *__restrict__ btot,
                                                          int32_t *__restrict__ counters) {
    __shared__ int32_t h[S1_NSUB][S1_WIN];
    __shared__ int32_t red[2][S1_NW];
    const int tid = threadIdx.x;
    int32_t *hs = h[(tid / 64) % S1_NSUB];
    const int64_t t0 = (int64_t)blockIdx.x * tpc, t1 = min(T, t0 + tpc);
    int4 qn = make_int4(0, 0, 0, 0);
    if (t0 + tid < t1) qn = tv4[t0 + tid];
    for (int64_t tb = t0; tb < t1; tb += S1_TILE) {
        const bool on = tb + tid < t1;
        const int4 q = qn;
        if (tb + S1_TILE + tid < t1) qn = tv4[tb + S1_TILE + tid];
        const int32_t b0 = q.x >> bv_log, b1 = q.y >> bv_log, b2 = q.z >> bv_log,
                      b3 = q.w >> bv_log;
#pragma unroll
        for (int k = 0; k < S1_NSUB; k++) h[k][tid] = 0;
        int32_t bmin, bmax;
        tile_window(on ? min(min(b0, b1), min(b2, b3)) : INT32_MAX,
                    on ? max(max(b0, b1), max(b2, b3)) : -1, red, bmin, bmax);
        if (bmax - bmin < S1_WIN) {
            if (on) {
                atomicAdd(&hs[b0 - bmin], 1);
                atomicAdd(&hs[b1 - bmin], 1);
                atomicAdd(&hs[b2 - bmin], 1);
                atomicAdd(&hs[b3 - bmin], 1);
            }
            __syncthreads();
            int32_t c = 0;
#pragma unroll
            for (int k = 0; k < S1_NSUB; k++) c += h[k][tid];
            if (c) atomicAdd(&btot[bmin + tid], c);
        } else {
            if (on) {
                atomicAdd(&btot[b0], 1);
                atomicAdd(&btot[b1], 1);
                atomicAdd(&btot[b2], 1);
                atomicAdd(&btot[b3], 1);
            }
            if (tid == 0) atomicAdd(counters + tg::C_WIDE, 1);
        }
        __syncthreads();  // h and red are reused by the next tile
    }
}

struct S1Stage {
    uint32_t w0[4 * S1_TILE], w1[4 * S1_TILE], w2[4 * S1_TILE];
    uint16_t bin[4 * S1_TILE];
    int32_t cnt[S1_WIN], toff[S1_WIN], gb[S1_WIN];
    int32_t red[2][S1_NW];
    int32_t wsum[S1_NW];
};

// bstart: exclusive scan of the bucket counts; bused: zeroed fill counters.
template <bool VTREC>
__global__ __launch_bounds__(S1_THREADS) void k_s1_scatter(const int4 *__restrict__ tv4,
                                                            int64_t T, int64_t tpc, int bv_log,
                                                            const int32_t *__restrict__ bstart,
                                                            int32_t *__restrict__ bused,
                                                            Rec out) {
    __shared__ S1Stage st;
    const int tid = threadIdx.x, lane = __lane_id(), wid = tid / 64;
    const int64_t t0 = (int64_t)blockIdx.x * tpc, t1 = min(T, t0 + tpc);
    int4 qn = make_int4(0, 0, 0, 0);
    if (t0 + tid < t1) qn = tv4[t0 + tid];
    for (int64_t tb = t0; tb < t1; tb += S1_TILE) {
        const bool on = tb + tid < t1;
        const int4 q = qn;
        if (tb + S1_TILE + tid < t1) qn = tv4[tb + S1_TILE + tid];
        const uint32_t x = (uint32_t)q.x, y = (uint32_t)q.y, z = (uint32_t)q.z, w = (uint32_t)q.w;
        const int32_t b0 = q.x >> bv_log, b1 = q.y >> bv_log, b2 = q.z >> bv_log,
                      b3 = q.w >> bv_log;
        st.cnt[tid] = 0;
        int32_t bmin, bmax;
        tile_window(on ? min(min(b0, b1), min(b2, b3)) : INT32_MAX,
                    on ? max(max(b0, b1), max(b2, b3)) : -1, st.red, bmin, bmax);
        if (bmax - bmin < S1_WIN) {
            int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
            if (on) {
                r0 = atomicAdd(&st.cnt[b0 - bmin], 1);
                r1 = atomicAdd(&st.cnt[b1 - bmin], 1);
                r2 = atomicAdd(&st.cnt[b2 - bmin], 1);
                r3 = atomicAdd(&st.cnt[b3 - bmin], 1);
            }
            __syncthreads();
            const int32_t c = st.cnt[tid];
            // this tile's run of bucket bmin + tid: one global atomic
            const int32_t base = c ? bstart[bmin + tid] + atomicAdd(&bused[bmin + tid], c) : 0;
            const int32_t inc = tgd::wave_inclusive_sum(c);
            if (lane == 63) st.wsum[wid] = inc;
            __syncthreads();
            int32_t pre = 0, total = 0;
#pragma unroll
            for (int i = 0; i < S1_NW; i++) {
                if (i < wid) pre += st.wsum[i];
                total += st.wsum[i];
            }
            st.toff[tid] = pre + inc - c;
            st.gb[tid] = base;
            __syncthreads();
            if (on) {
                auto put = [&](int j, int rank, uint32_t v, uint32_t a, uint32_t b, uint32_t cc) {
                    const int slot = st.toff[j] + rank;
                    st.w0[slot] = v | (a << 24);
                    st.w1[slot] = (a >> 8) | (b << 16);
                    st.w2[slot] = (b >> 16) | (cc << 8);
                    st.bin[slot] = (uint16_t)j;
                };
                if constexpr (VTREC) {
                    const uint32_t c0 = 4u * (uint32_t)(tb + tid);
                    put(b0 - bmin, r0, x, c0 & M24, c0 >> 24, 0u);
                    put(b1 - bmin, r1, y, (c0 + 1) & M24, (c0 + 1) >> 24, 0u);
                    put(b2 - bmin, r2, z, (c0 + 2) & M24, (c0 + 2) >> 24, 0u);
                    put(b3 - bmin, r3, w, (c0 + 3) & M24, (c0 + 3) >> 24, 0u);
                } else {
                    put(b0 - bmin, r0, x, y, z, w);
                    put(b1 - bmin, r1, y, x, z, w);
                    put(b2 - bmin, r2, z, x, y, w);
                    put(b3 - bmin, r3, w, x, y, z);
                }
            }
            __syncthreads();
            for (int slot = tid; slot < total; slot += S1_THREADS) {
                const int j = st.bin[slot];
                const int64_t pos = (int64_t)st.gb[j] + (slot - st.toff[j]);
                out.w0[pos] = st.w0[slot];
                out.w1[pos] = st.w1[slot];
                out.w2[pos] = st.w2[slot];
            }
        } else if (on) {
            if constexpr (VTREC) {
                const uint32_t c0 = 4u * (uint32_t)(tb + tid);
                rec_store(out, (int64_t)bstart[b0] + atomicAdd(&bused[b0], 1), x, c0 & M24, c0 >> 24, 0u);
                rec_store(out, (int64_t)bstart[b1] + atomicAdd(&bused[b1], 1), y, (c0 + 1) & M24,
                          (c0 + 1) >> 24, 0u);
                rec_store(out, (int64_t)bstart[b2] + atomicAdd(&bused[b2], 1), z, (c0 + 2) & M24,
                          (c0 + 2) >> 24, 0u);
                rec_store(out, (int64_t)bstart[b3] + atomicAdd(&bused[b3], 1), w, (c0 + 3) & M24,
                          (c0 + 3) >> 24, 0u);
            } else {
                rec_store(out, (int64_t)bstart[b0] + atomicAdd(&bused[b0], 1), x, y, z, w);
                rec_store(out, (int64_t)bstart[b1] + atomicAdd(&bused[b1], 1), y, x, z, w);
                rec_store(out, (int64_t)bstart[b2] + atomicAdd(&bused[b2], 1), z, x, y, w);
                rec_store(out, (int64_t)bstart[b3] + atomicAdd(&bused[b3], 1), w, x, y, z);
            }
        }
        __syncthreads();  // the stage is reused by the next tile
    }
}

// ---- B: per-bucket neighbour sets, link components, sorted rows --------------
// Probe order of key u in a 32-slot table: 8 groups of 4 slots (one 16-byte
// LDS read per group), starting at the home group from a multiplicative
// hash; inside a group a key claims the first empty slot at or after its
// hint slot (cyclic), so two new keys of one group seldom race for a slot.
__device__ __forceinline__ uint32_t key_hash(int32_t u) {
    return ((uint32_t)u * 2654435769u) >> 27;  // 5 bits: group (3) | slot hint (2)
}

// First empty slot of a group snapshot at or after slot p (cyclic), or -1.
__device__ __forceinline__ int first_empty_from(const int4 q, int p) {
    const uint32_t m = (uint32_t)(q.x == -1) | ((uint32_t)(q.y == -1) << 1) |
                       ((uint32_t)(q.z == -1) << 2) | ((uint32_t)(q.w == -1) << 3);
    const uint32_t r = ((m | (m << 4)) >> p) & 15u;  // rotate so slot p is bit 0
    return r ? (int)((__ffs(r) - 1 + p) & 3) : -1;
}

// Insert u into a vertex's table (lock-free: slots only go from empty (-1)
// to a key by CAS, and every thread inserting u follows the same probe
// order, so u lands in exactly one slot).  Returns the slot or -1 if full.
// HINT: start each group at the key's hint slot (else at slot 0); every
// insert into one table must use the same setting.
template <bool HINT>
__device__ __forceinline__ int hash_insert(int32_t *tab, int32_t u) {
    const uint32_t h = key_hash(u);
#pragma unroll 1
    for (int step = 0; step < SLOTS / 4; step++) {
        const int g = (int)(((h >> 2) + (uint32_t)step) & (SLOTS / 4 - 1));
        int32_t *grp = tab + 4 * g;
        int4 q = *reinterpret_cast<const int4 *>(grp);
#pragma unroll 1
        for (int tries = 0; tries < 4; tries++) {
            if (q.x == u) return 4 * g;
            if (q.y == u) return 4 * g + 1;
            if (q.z == u) return 4 * g + 2;
            if (q.w == u) return 4 * g + 3;
            const int e = first_empty_from(q, HINT ? (int)(h & 3u) : 0);
            if (e < 0) break;
            const int32_t old = atomicCAS(&grp[e], -1, u);
            if (old == -1 || old == u) return 4 * g + e;
            q = *reinterpret_cast<const int4 *>(grp);  // lost to another key
        }
    }
    return -1;
}

// Components of `mask` under the slot adjacency, numbered 0.. in order of
// their lowest slot.  Every slot of `mask` is visited once, and its key is
// rewritten in place as (key << 6) | tag | component number: the row's sort
// key and its descriptor (lower bit 5, component bits 0-4) in one word.
__device__ __forceinline__ uint32_t label_components(const uint32_t *adj, int32_t *key,
                                                     uint32_t mask, uint32_t tag) {
    uint32_t cnt = 0, rem = mask;
    while (rem) {
        uint32_t comp = rem & (0u - rem);
        uint32_t front = comp;
        while (front) {
            const int f = __ffs(front) - 1;
            front &= front - 1u;
            const uint32_t nb = adj[f] & mask & ~comp;
            key[f] = (int32_t)(((uint32_t)key[f] << 6) | tag | cnt);
            comp |= nb;
            front |= nb;
        }
        rem &= ~comp;
        cnt++;
    }
    return cnt;
}

// Bitonic sorting network over 32 registers (ascending).
template <typename K>
__device__ __forceinline__ void sort32(K (&r)[32]) {
#pragma unroll
    for (int k = 2; k <= 32; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < 32; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const K a = r[i], b = r[l];
                    const bool up = (i & k) == 0;
                    r[i] = up ? min(a, b) : max(a, b);
                    r[l] = up ? max(a, b) : min(a, b);
                }
            }
}

// Lean insert of one record's three keys into the vertex's 32-slot table:
// each key probes linearly from its hash slot, claiming a slot by CAS; a
// slot found holding the key (or claimed by another thread for the same key)
// places it.  The three probe chains advance together, so a lane costs the
// longest of its chains, not their sum.  Slots go only from empty (-1) to a
// key, and every thread follows the same probe sequence for a key, so each
// key lands in exactly one slot.  A key that meets 32 occupied slots leaves
// its slot at -1 (the table is full: the vertex goes to the heavy kernel).
__device__ __forceinline__ void lean_insert3(int32_t *tab, int32_t x, int32_t y, int32_t z,
                                             int &sx, int &sy, int &sz) {
    int px = (int)key_hash(x), py = (int)key_hash(y), pz = (int)key_hash(z);
    bool dx = false, dy = false, dz = false;
#pragma unroll 1
    for (int step = 0; step < SLOTS; step++) {
        if (!dx) {
            const int32_t o = atomicCAS(tab + px, -1, x);
            dx = o == -1 || o == x;
            if (!dx) px = (px + 1) & (SLOTS - 1);
        }
        if (!dy) {
            const int32_t o = atomicCAS(tab + py, -1, y);
            dy = o == -1 || o == y;
            if (!dy) py = (py + 1) & (SLOTS - 1);
        }
        if (!dz) {
            const int32_t o = atomicCAS(tab + pz, -1, z);
            dz = o == -1 || o == z;
            if (!dz) pz = (pz + 1) & (SLOTS - 1);
        }
        if (dx && dy && dz) break;
    }
    sx = dx ? px : -1;
    sy = dy ? py : -1;
    sz = dz ? pz : -1;
}

// INS selects phase 1: 0 batches of four records per thread with group
// snapshots; 1 record by record with lean_insert3, four records per thread
// in flight, the wave in lock-step per record (about 40 % fewer VALU
// instructions per key; same time on MI355X, where the insert is bound by the
// dependent LDS round trips of the probe chains, not by instruction count).
template <int BV_LOG, typename S, int INS>
__global__ __launch_bounds__(2 << BV_LOG) void k_bkt_vv_cp(
    const Rec rec, const int32_t *__restrict__ h2s, int32_t nslice,
    const S *__restrict__ scal, int32_t V, int classify, int32_t *__restrict__ scratch,
    int32_t *__restrict__ deg, int32_t *__restrict__ rowsrc, uint32_t *__restrict__ classes,
    int32_t *__restrict__ heavy, int32_t *__restrict__ fixd, uint8_t *__restrict__ rowdesc,
    int32_t *__restrict__ counters) {
    constexpr int BV = 1 << BV_LOG;
    constexpr int NT = 2 * BV;
    constexpr int BATCH = 4;
    constexpr int PER_T = BV * SLOTS / NT;  // slots per thread in phase 2 (16)
    // table strides: a vertex's slots start 36 (keys, 16-byte groups) or 33
    // (adjacency) words after the previous vertex's, so accesses of many
    // vertices at one slot index spread over the LDS banks
    constexpr int KS = SLOTS + 4, AS = SLOTS + 1;
    // hint slots cut CAS conflicts but cost registers: at 512 threads they
    // would halve the resident blocks (measured on MI355X, Kuhn 128^3)
    constexpr bool HINT = BV_LOG == 7;
    __shared__ __align__(16) int32_t key[BV * KS];
    __shared__ uint32_t adj[BV * AS];
    __shared__ uint32_t occ[BV], lowm[BV];
    __shared__ int32_t ccount[BV];
    __shared__ int32_t ovf[BV];
    __shared__ S sv[BV];
    __shared__ int32_t wsum[2][NT / 64];

    const int tid = threadIdx.x;
    const int32_t b = blockIdx.x;
    const int32_t v0 = b << BV_LOG;
    const int nv = min(BV, V - v0);
    const int32_t cbeg = h2s[(int64_t)b * nslice], cend = h2s[(int64_t)(b + 1) * nslice];

    for (int i = tid; i < BV * KS; i += NT) key[i] = -1;
    for (int i = tid; i < BV * AS; i += NT) adj[i] = 0u;
    if (tid < BV) {
        occ[tid] = 0u;
        lowm[tid] = 0u;
        ccount[tid] = 0;
        ovf[tid] = 0;
        sv[tid] = (classify && tid < nv) ? scal[v0 + tid] : S(0);
    }
    __syncthreads();

    if constexpr (INS == 1) {
        // four records per thread in flight: record i + k * NT sits in ring
        // slot k; a slot is refilled with the record four strides ahead as
        // soon as its record is taken
        constexpr int DEPTH = 4;
        uint32_t q[DEPTH][3];
#pragma unroll
        for (int k = 0; k < DEPTH; k++) {
            const int32_t i = cbeg + tid + k * NT;
            q[k][0] = 0xFFFFFFFFu;
            if (i < cend) {
                q[k][0] = rec.w0[i];
                q[k][1] = rec.w1[i];
                q[k][2] = rec.w2[i];
            }
        }
#pragma unroll 1
        for (int32_t i0 = cbeg + tid; i0 < cend; i0 += DEPTH * NT) {
#pragma unroll
            for (int k = 0; k < DEPTH; k++) {
                const uint32_t w0 = q[k][0], w1 = q[k][1], w2 = q[k][2];
                const int32_t nx = i0 + (k + DEPTH) * NT;
                q[k][0] = 0xFFFFFFFFu;
                if (nx < cend) {
                    q[k][0] = rec.w0[nx];
                    q[k][1] = rec.w1[nx];
                    q[k][2] = rec.w2[nx];
                }
                if (w0 == 0xFFFFFFFFu) continue;
                uint32_t v, x, y, z;
                rec_unpack(w0, w1, w2, v, x, y, z);
                const int vl = (int)(v & (BV - 1));
                int a, b2, c;
                lean_insert3(key + vl * KS, (int32_t)x, (int32_t)y, (int32_t)z, a, b2, c);
                atomicAdd(&ccount[vl], 1);
                if ((a | b2 | c) < 0) {
                    ovf[vl] = 1;
                } else {
                    uint32_t *av = adj + vl * AS;
                    atomicOr(&av[a], (1u << b2) | (1u << c));
                    atomicOr(&av[b2], (1u << a) | (1u << c));
                    atomicOr(&av[c], (1u << a) | (1u << b2));
                }
            }
        }
    } else
    // phase 1: neighbour sets and shared-tet link adjacency.  For a batch of
    // keys: all group reads are issued together; a key found in its group
    // snapshot is placed, otherwise it claims the snapshot's first empty
    // slot by CAS, and all those CAS are issued before any result is used,
    // so a batch costs about two LDS round trips.  Keys whose CAS lost to a
    // different key (or whose group was full) are queued in a per-lane mask
    // and resolved by one shared probing loop.
    for (int32_t i0 = cbeg + tid; i0 < cend; i0 += BATCH * NT) {
        constexpr int NK = 3 * BATCH;
        uint32_t w[BATCH][3];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int32_t i = i0 + k * NT;
            w[k][0] = 0xFFFFFFFFu;
            if (i < cend) {
                w[k][0] = rec.w0[i];
                w[k][1] = rec.w1[i];
                w[k][2] = rec.w2[i];
            }
        }
        int32_t u[NK];
        int vls[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            uint32_t v = 0, x = 0, y = 0, z = 0;
            if (w[k][0] != 0xFFFFFFFFu) rec_unpack(w[k][0], w[k][1], w[k][2], v, x, y, z);
            vls[k] = (int)(v & (BV - 1));
            u[3 * k] = (int32_t)x;
            u[3 * k + 1] = (int32_t)y;
            u[3 * k + 2] = (int32_t)z;
        }
        int4 grp[NK];
#pragma unroll
        for (int j = 0; j < NK; j++) {
            const uint32_t g = key_hash(u[j]) >> 2;
            grp[j] = *reinterpret_cast<const int4 *>(key + vls[j / 3] * KS + 4 * g);
        }
        int sl[NK], cs[NK];
#pragma unroll
        for (int j = 0; j < NK; j++) {
            const int32_t x = u[j];
            const uint32_t h = key_hash(x);
            const int g = (int)(h >> 2);
            const int4 q = grp[j];
            const int f = q.x == x ? 0 : q.y == x ? 1 : q.z == x ? 2 : q.w == x ? 3 : -1;
            const int e = first_empty_from(q, HINT ? (int)(h & 3u) : 0);
            const bool valid = w[j / 3][0] != 0xFFFFFFFFu;
            sl[j] = f >= 0 ? 4 * g + f : -1;
            cs[j] = (valid && f < 0 && e >= 0) ? 4 * g + e : -1;  // slot to claim
            if (valid && f < 0 && e < 0) sl[j] = -2;            // group full
        }
        int32_t old[NK];
#pragma unroll
        for (int j = 0; j < NK; j++)
            old[j] = cs[j] >= 0 ? atomicCAS(key + vls[j / 3] * KS + cs[j], -1, u[j]) : 0;
        uint32_t pending = 0;
#pragma unroll
        for (int j = 0; j < NK; j++) {
            if (cs[j] >= 0) sl[j] = (old[j] == -1 || old[j] == u[j]) ? cs[j] : -2;
            if (sl[j] == -2) pending |= 1u << j;
        }
        while (__any(pending != 0)) {
            if (pending) {
                const int j = __ffs(pending) - 1;
                pending &= pending - 1u;
                int32_t x = 0;
                int vl = 0;
#pragma unroll
                for (int t = 0; t < NK; t++)
                    if (t == j) {
                        x = u[t];
                        vl = vls[t / 3];
                    }
                const int r = hash_insert<HINT>(key + vl * KS, x);
#pragma unroll
                for (int t = 0; t < NK; t++)
                    if (t == j) sl[t] = r;
            }
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            if (w[k][0] == 0xFFFFFFFFu) continue;
            const int vl = vls[k];
            atomicAdd(&ccount[vl], 1);
            const int a = sl[3 * k], b2 = sl[3 * k + 1], c = sl[3 * k + 2];
            if ((a | b2 | c) < 0) {
                ovf[vl] = 1;
            } else {
                uint32_t *av = adj + vl * AS;
                atomicOr(&av[a], (1u << b2) | (1u << c));
                atomicOr(&av[b2], (1u << a) | (1u << c));
                atomicOr(&av[c], (1u << a) | (1u << b2));
            }
        }
    }
    __syncthreads();
    // occupancy masks from the final tables
    for (int vl = tid / SLOTS; vl < BV; vl += NT / SLOTS) {
        const uint64_t bal = __ballot(key[vl * KS + (tid & (SLOTS - 1))] >= 0);
        if ((tid & (SLOTS - 1)) == 0) occ[vl] = (uint32_t)(bal >> (tid & 32));
    }
    __syncthreads();

    // phase 2: lower/upper mask of every slot; a half-wave covers one vertex's
    // 32 slots, so the mask is a ballot
    if (classify) {
        int32_t u[PER_T];
        S su[PER_T];
#pragma unroll
        for (int k = 0; k < PER_T; k++) {
            const int idx = tid + k * NT;
            u[k] = (idx / SLOTS < nv) ? key[(idx / SLOTS) * KS + idx % SLOTS] : -1;
        }
#pragma unroll
        for (int k = 0; k < PER_T; k++) su[k] = u[k] >= 0 ? scal[u[k]] : S(0);
#pragma unroll
        for (int k = 0; k < PER_T; k++) {
            const int idx = tid + k * NT;
            const int vl = idx / SLOTS;
            const bool lo = u[k] >= 0 && is_lower(su[k], sv[vl], u[k], v0 + vl);
            const uint64_t bal = __ballot(lo);
            if ((tid & 31) == 0)
                lowm[vl] = (uint32_t)(bal >> (tid & 32));
        }
    }
    __syncthreads();

    // phase 3: per vertex (lanes < nv): components, row registers.  Sort
    // keys are unsigned, so empty slots (-1) sort last.
    const int vl = tid;
    const bool act = vl < nv;
    const bool hv = act && ovf[vl] != 0;
    const bool ok = act && !hv;
    uint32_t r[32];
    uint32_t nl = 0, nu = 0;
    int d = 0;
    if (ok) {
        const uint32_t om = occ[vl];
        d = __popc(om);
        if (classify) {
            const uint32_t lm = lowm[vl] & om;
            nl = label_components(adj + vl * AS, key + vl * KS, lm, 32u);
            nu = label_components(adj + vl * AS, key + vl * KS, om & ~lm, 0u);
        }
        const uint4 *kv = reinterpret_cast<const uint4 *>(key + vl * KS);
#pragma unroll
        for (int i = 0; i < SLOTS / 4; i++) {
            const uint4 q = kv[i];
            r[4 * i] = q.x;
            r[4 * i + 1] = q.y;
            r[4 * i + 2] = q.z;
            r[4 * i + 3] = q.w;
        }
    }
    const int32_t hres = hv ? 3 * ccount[vl] : 0;
    __syncthreads();  // every lane has read its keys; key[] becomes row staging

    // block exclusive scan of (row lengths, heavy reservations)
    const int32_t x = ok ? d : 0, y = hres;
    const int32_t xi = tgd::wave_inclusive_sum(x), yi = tgd::wave_inclusive_sum(y);
    const int lane = __lane_id(), wid = tid / 64;
    if (lane == 63) {
        wsum[0][wid] = xi;
        wsum[1][wid] = yi;
    }
    __syncthreads();
    int32_t xb = 0, yb = 0, xt = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        if (w < wid) {
            xb += wsum[0][w];
            yb += wsum[1][w];
        }
        xt += wsum[0][w];
    }
    const int32_t loff = hv ? xt + yb + yi - y : xb + xi - x;
    if (ok) {
        sort32(r);
#pragma unroll
        for (int i = 0; i < 32; i++)
            if (i < d) key[loff + i] = (int32_t)r[i];
        const int32_t v = v0 + vl;
        deg[v] = d;
        if (classify) {
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
    }
    if (act) rowsrc[v0 + vl] = 3 * cbeg + loff;
    wave_append(heavy, counters + C_HEAVY, v0 + vl, hv);
    wave_append(fixd, counters + tg::C_FIXD, v0 + vl, ok && classify && (nl >= 2 || nu >= 2));
    int dm = ok ? d : 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) dm = max(dm, __shfl_xor(dm, s, 64));
    if (lane == 0 && dm > 0) atomicMax(counters + C_MAXDEG, dm);
    __syncthreads();

    // phase 4: coalesced store of the bucket's rows and row descriptors
    int32_t *dst = scratch + 3 * (int64_t)cbeg;
    uint8_t *ddst = rowdesc + 3 * (int64_t)cbeg;
    if (classify) {
        for (int i = tid; i < xt; i += NT) {
            const int32_t pk = key[i];
            dst[i] = pk >> 6;
            ddst[i] = (uint8_t)(pk & 63);
        }
    } else {
        for (int i = tid; i < xt; i += NT) dst[i] = key[i];
    }
}

// Components of `mask` as label_components, over a lane-private adjacency
// table stored slot-major (slot f of lane at adj[f * stride]).
__device__ __forceinline__ uint32_t label_components_strided(const uint32_t *adj, int stride,
                                                             uint32_t mask, uint32_t (&pl)[5]) {
    uint32_t cnt = 0, rem = mask;
    while (rem) {
        uint32_t comp = rem & (0u - rem);
        uint32_t front = comp;
        while (front) {
            const int f = __ffs(front) - 1;
            front &= front - 1u;
            const uint32_t nb = adj[f * stride] & mask & ~comp;
            comp |= nb;
            front |= nb;
        }
#pragma unroll
        for (int k = 0; k < 5; k++)
            if ((cnt >> k) & 1u) pl[k] |= comp;
        rem &= ~comp;
        cnt++;
    }
    return cnt;
}

// B, owner form: the bucket's records are first grouped by vertex with an
// LDS counting sort (a list of record indices), then each lane owns one
// vertex and builds its neighbour table and link adjacency alone, with plain
// LDS reads and writes instead of contended compare-and-swap.  Tables are
// slot-major (slot s of lane l at s * BV + l), so a wave's table accesses
// fall on distinct banks.  Rows, descriptors and the rest are as in
// k_bkt_vv_cp.  A bucket with more than OWN_CAP records sends all its
// vertices to the serial kernel.
constexpr int OWN_CAP = 8192;

__device__ __forceinline__ uint32_t own_hash(int32_t u) { return ((uint32_t)u * 2654435769u) >> 27; }

template <int BV_LOG, typename S>
__global__ __launch_bounds__(2 << BV_LOG) void k_bkt_own(
    const Rec rec, const int32_t *__restrict__ h2s, int32_t nslice,
    const S *__restrict__ scal, int32_t V, int classify, int32_t *__restrict__ scratch,
    int32_t *__restrict__ deg, int32_t *__restrict__ rowsrc, uint32_t *__restrict__ classes,
    int32_t *__restrict__ heavy, int32_t *__restrict__ fixd, uint8_t *__restrict__ rowdesc,
    int32_t *__restrict__ counters) {
    constexpr int BV = 1 << BV_LOG;
    constexpr int NT = 2 * BV;
    __shared__ int32_t key[SLOTS * BV];   // slot-major; reused as row staging
    __shared__ uint32_t adj[SLOTS * BV];
    __shared__ uint16_t idx[OWN_CAP];
    __shared__ int32_t cnt[BV + 1], cur[BV];
    __shared__ int32_t wsum[2][NT / 64];

    const int tid = threadIdx.x;
    const int32_t b = blockIdx.x;
    const int32_t v0 = b << BV_LOG;
    const int nv = min(BV, V - v0);
    const int32_t cbeg = h2s[(int64_t)b * nslice], cend = h2s[(int64_t)(b + 1) * nslice];
    const int32_t nrec = cend - cbeg;
    const bool fits = nrec <= OWN_CAP;

    for (int i = tid; i < SLOTS * BV; i += NT) {
        key[i] = -1;
        adj[i] = 0u;
    }
    if (tid < BV) cnt[tid] = 0;
    __syncthreads();
    // counting sort of the bucket's records by vertex (indices into the run);
    // the counts are needed even when the indices do not fit
    for (int32_t i = tid; i < nrec; i += NT) atomicAdd(&cnt[rec.w0[cbeg + i] & (BV - 1)], 1);
    __syncthreads();
    if (tid < 64) {  // exclusive scan of cnt[0..BV) by the first wave
        int32_t carry = 0;
        for (int i0 = 0; i0 < BV; i0 += 64) {
            const int32_t x = cnt[i0 + tid];
            const int32_t inc = tgd::wave_inclusive_sum(x);
            cur[i0 + tid] = carry + inc - x;
            carry += __shfl(inc, 63, 64);
        }
        if (tid == 0) cnt[BV] = carry;
    }
    __syncthreads();
    if (tid < BV) cnt[tid] = cur[tid];  // cnt = row starts, cnt[BV] = total
    __syncthreads();
    if (fits) {
        for (int32_t i = tid; i < nrec; i += NT) {
            const int p = atomicAdd(&cur[rec.w0[cbeg + i] & (BV - 1)], 1);
            idx[p] = (uint16_t)i;
        }
    }
    __syncthreads();

    // phase 1 (owner): lane vl inserts the neighbours of vertex v0 + vl
    const int vl = tid;
    const bool act = vl < nv;
    bool over = !fits;
    uint32_t om = 0;
    int d = 0;
    int32_t nr = 0;
    if (act) nr = cnt[vl + 1] - cnt[vl];
    if (act && fits) {
        const int p0 = cnt[vl], p1 = cnt[vl + 1];
        auto insert = [&](int32_t u) -> int {
            int s = (int)own_hash(u);
#pragma unroll 1
            for (int step = 0; step < SLOTS; step++, s = (s + 1) & (SLOTS - 1)) {
                const int32_t k = key[s * BV + vl];
                if (k == u) return s;
                if (k < 0) {
                    key[s * BV + vl] = u;
                    om |= 1u << s;
                    d++;
                    return s;
                }
            }
            return -1;
        };
        uint32_t a0 = 0, a1 = 0, a2 = 0;
        if (p0 < p1) {
            const int32_t i = cbeg + idx[p0];
            a0 = rec.w0[i];
            a1 = rec.w1[i];
            a2 = rec.w2[i];
        }
        for (int p = p0; p < p1 && !over; p++) {
            uint32_t rv, x, y, z;
            rec_unpack(a0, a1, a2, rv, x, y, z);
            if (p + 1 < p1) {  // next record in flight
                const int32_t i = cbeg + idx[p + 1];
                a0 = rec.w0[i];
                a1 = rec.w1[i];
                a2 = rec.w2[i];
            }
            const int s0 = insert((int32_t)x), s1 = insert((int32_t)y), s2 = insert((int32_t)z);
            if ((s0 | s1 | s2) < 0) {
                over = true;
                break;
            }
            adj[s0 * BV + vl] |= (1u << s1) | (1u << s2);
            adj[s1 * BV + vl] |= (1u << s0) | (1u << s2);
            adj[s2 * BV + vl] |= (1u << s0) | (1u << s1);
        }
    }
    const bool hv = act && over;
    const bool ok = act && !over;

    // phase 2-3 (owner): lower mask from the scalars, components, row keys
    int32_t r[32];
    uint32_t nl = 0, nu = 0;
    if (ok) {
        uint32_t lm = 0, pl[5] = {0, 0, 0, 0, 0};
        if (classify) {
            const int32_t v = v0 + vl;
            const S sv = scal[v];
            uint32_t rem = om;
            while (rem) {
                const int s = __ffs(rem) - 1;
                rem &= rem - 1u;
                const int32_t u = key[s * BV + vl];
                if (is_lower(scal[u], sv, u, v)) lm |= 1u << s;
            }
            nl = label_components_strided(adj + vl, BV, lm, pl);
            nu = label_components_strided(adj + vl, BV, om & ~lm, pl);
        }
#pragma unroll
        for (int i = 0; i < 32; i++) {
            const int32_t k = key[i * BV + vl];
            const uint32_t desc = (((lm >> i) & 1u) << 5) | ((pl[0] >> i) & 1u) |
                                  (((pl[1] >> i) & 1u) << 1) | (((pl[2] >> i) & 1u) << 2) |
                                  (((pl[3] >> i) & 1u) << 3) | (((pl[4] >> i) & 1u) << 4);
            r[i] = k < 0 ? INT32_MAX : (int32_t)(((uint32_t)k << 6) | desc);
        }
    }
    // a heavy vertex reserves room for 3 neighbours per record (in a bucket
    // whose index list did not fit, every vertex is heavy)
    const int32_t hres = hv ? 3 * nr : 0;
    __syncthreads();  // every lane has read its keys; key[] becomes row staging

    const int32_t x = ok ? d : 0, y = hres;
    const int32_t xi = tgd::wave_inclusive_sum(x), yi = tgd::wave_inclusive_sum(y);
    const int lane = __lane_id(), wid = tid / 64;
    if (lane == 63) {
        wsum[0][wid] = xi;
        wsum[1][wid] = yi;
    }
    __syncthreads();
    int32_t xb = 0, yb = 0, xt = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
        if (w < wid) {
            xb += wsum[0][w];
            yb += wsum[1][w];
        }
        xt += wsum[0][w];
    }
    // rows and heavy reservations are laid out at 3 * cbeg; they fit in the
    // bucket's 3 x nrec words of scratch (a row has at most 3 entries per
    // record, a heavy reservation exactly that)
    const int32_t loff = hv ? xt + yb + yi - y : xb + xi - x;
    if (ok) {
        sort32(r);
#pragma unroll
        for (int i = 0; i < 32; i++)
            if (i < d) key[loff + i] = r[i];
        const int32_t v = v0 + vl;
        deg[v] = d;
        if (classify) {
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
    }
    if (act) rowsrc[v0 + vl] = 3 * cbeg + loff;
    wave_append(heavy, counters + C_HEAVY, v0 + vl, hv);
    wave_append(fixd, counters + tg::C_FIXD, v0 + vl, ok && classify && (nl >= 2 || nu >= 2));
    int dm = ok ? d : 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) dm = max(dm, __shfl_xor(dm, s, 64));
    if (lane == 0 && dm > 0) atomicMax(counters + C_MAXDEG, dm);
    __syncthreads();

    int32_t *dst = scratch + 3 * (int64_t)cbeg;
    uint8_t *ddst = rowdesc + 3 * (int64_t)cbeg;
    for (int i = tid; i < xt; i += NT) {
        const int32_t pk = key[i];
        dst[i] = pk >> 6;
        if (classify) ddst[i] = (uint8_t)(pk & 63);
    }
}

// Vertices with more than 32 neighbours (or whose bucket table overflowed):
// one workgroup per vertex scans its bucket's corner records.  Neighbours go
// to an LDS open-addressing set of 2 * HEAVY_CAP slots (CAS inserts), the
// shared-tet link edges between same-class slots to a lock-free union-find
// (a root is only ever linked under a smaller root, by CAS), and the row is
// sorted by a bitonic network in LDS.
constexpr int HV_THREADS = 256;
constexpr int HV_SLOTS = 2 * HEAVY_CAP;
constexpr int HV_BITS = 13;  // log2(HV_SLOTS)
static_assert((1 << HV_BITS) == HV_SLOTS, "heavy set size");

// Slot of u in the set (inserted when absent), or -1 when the set is full.
__device__ __forceinline__ int hv_insert(int32_t *set, int32_t u, int32_t *count) {
    uint32_t i = ((uint32_t)u * 2654435769u) >> (32 - HV_BITS);
    for (int probe = 0; probe < HV_SLOTS; probe++, i = (i + 1) & (HV_SLOTS - 1)) {
        const int32_t k = set[i];
        if (k == u) return (int)i;
        if (k == -1) {
            const int32_t old = atomicCAS(&set[i], -1, u);
            if (old == -1) {
                atomicAdd(count, 1);
                return (int)i;
            }
            if (old == u) return (int)i;
        }
    }
    return -1;
}

// Slot of u, which is in the set.
__device__ __forceinline__ int hv_find(const int32_t *set, int32_t u) {
    uint32_t i = ((uint32_t)u * 2654435769u) >> (32 - HV_BITS);
    for (int probe = 0; probe < HV_SLOTS; probe++, i = (i + 1) & (HV_SLOTS - 1))
        if (set[i] == u) return (int)i;
    return -1;
}

// Lock-free union-find over par[] (par[x] <= x always; path halving).
__device__ __forceinline__ int uf_root(int32_t *par, int x) {
    while (true) {
        const int p = par[x];
        if (p == x) return x;
        const int g = par[p];
        if (g != p) par[x] = g;
        x = g;
    }
}

__device__ __forceinline__ void uf_unite(int32_t *par, int a, int b) {
    while (true) {
        a = uf_root(par, a);
        b = uf_root(par, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b) return;
    }
}

// Ascending bitonic sort of list[0..P) in LDS (P a power of two), whole block.
__device__ __forceinline__ void block_bitonic(int32_t *list, int P) {
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const int32_t a = list[i], b = list[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        list[i] = b;
                        list[l] = a;
                    }
                }
            }
            __syncthreads();
        }
}

template <int BV_LOG, typename S>
__global__ __launch_bounds__(HV_THREADS) void k_bkt_heavy(
    const Rec rec, const int32_t *__restrict__ h2s, int32_t nslice,
    const S *__restrict__ scal, int classify, int32_t *__restrict__ scratch,
    int32_t *__restrict__ deg, const int32_t *__restrict__ rowsrc,
    uint32_t *__restrict__ classes, const int32_t *__restrict__ heavy,
    int32_t *__restrict__ fix, int32_t *__restrict__ fixbig, int32_t *__restrict__ counters) {
    __shared__ int32_t set[HV_SLOTS];
    __shared__ int32_t par[HV_SLOTS];
    __shared__ uint8_t lowf[HV_SLOTS];
    __shared__ int32_t list[HEAVY_CAP];
    __shared__ int32_t misc[5];  // distinct neighbours, full flag, lower roots, upper roots, fill
    const int tid = threadIdx.x;
    const int32_t n = counters[C_HEAVY];
    for (int32_t h = blockIdx.x; h < n; h += gridDim.x) {
        const int32_t v = heavy[h];
        const int32_t bkt = v >> BV_LOG;
        const int32_t cbeg = h2s[(int64_t)bkt * nslice], cend = h2s[(int64_t)(bkt + 1) * nslice];
        const S sv = classify ? scal[v] : S(0);
        for (int i = tid; i < HV_SLOTS; i += HV_THREADS) {
            set[i] = -1;
            par[i] = i;
        }
        if (tid < 5) misc[tid] = 0;
        __syncthreads();
        // neighbour set
        for (int32_t i = cbeg + tid; i < cend; i += HV_THREADS) {
            uint32_t rv, x, y, z;
            rec_unpack(rec.w0[i], rec.w1[i], rec.w2[i], rv, x, y, z);
            if ((int32_t)rv != v) continue;
            if ((hv_insert(set, (int32_t)x, &misc[0]) | hv_insert(set, (int32_t)y, &misc[0]) |
                 hv_insert(set, (int32_t)z, &misc[0])) < 0)
                misc[1] = 1;
        }
        __syncthreads();
        const int m = misc[0];
        if (misc[1] || m > HEAVY_CAP) {  // block-uniform
            if (tid == 0) {
                atomicOr(counters + C_ERR, tg::ERRBIT_CAPACITY);
                deg[v] = 0;
                if (classify) {
                    classes[3 * (int64_t)v] = 0;
                    classes[3 * (int64_t)v + 1] = 0;
                    classes[3 * (int64_t)v + 2] = 0;
                }
            }
            __syncthreads();
            continue;
        }
        if (classify) {
            for (int i = tid; i < HV_SLOTS; i += HV_THREADS) {
                const int32_t u = set[i];
                lowf[i] = u >= 0 ? (uint8_t)is_lower(scal[u], sv, u, v) : (uint8_t)0;
            }
            __syncthreads();
            // shared-tet link edges between neighbours of one class
            for (int32_t i = cbeg + tid; i < cend; i += HV_THREADS) {
                uint32_t rv, x, y, z;
                rec_unpack(rec.w0[i], rec.w1[i], rec.w2[i], rv, x, y, z);
                if ((int32_t)rv != v) continue;
                const int a = hv_find(set, (int32_t)x), b = hv_find(set, (int32_t)y),
                          c = hv_find(set, (int32_t)z);
                if ((a | b | c) < 0) continue;  // not reached: every key was inserted above
                if (lowf[a] == lowf[b]) uf_unite(par, a, b);
                if (lowf[a] == lowf[c]) uf_unite(par, a, c);
                if (lowf[b] == lowf[c]) uf_unite(par, b, c);
            }
            __syncthreads();
        }
        // row (compacted) and component roots per class
        for (int i = tid; i < HV_SLOTS; i += HV_THREADS) {
            const int32_t u = set[i];
            if (u < 0) continue;
            list[atomicAdd(&misc[4], 1)] = u;
            if (classify && par[i] == i) atomicAdd(&misc[lowf[i] ? 2 : 3], 1);
        }
        int P = 1;
        while (P < m) P <<= 1;
        for (int i = m + tid; i < P; i += HV_THREADS) list[i] = INT32_MAX;
        __syncthreads();
        block_bitonic(list, P);
        int32_t *dst = scratch + rowsrc[v];
        for (int i = tid; i < m; i += HV_THREADS) dst[i] = list[i];
        if (tid == 0) {
            const uint32_t nl = (uint32_t)misc[2], nu = (uint32_t)misc[3];
            deg[v] = m;
            atomicMax(counters + C_MAXDEG, m);
            if (classify) {
                classes[3 * (int64_t)v] = nl;
                classes[3 * (int64_t)v + 1] = nu;
                classes[3 * (int64_t)v + 2] = verdict(nl, nu);
                if (nl >= 2 || nu >= 2) {
                    if (m <= 64) fix[atomicAdd(counters + C_FIX, 1)] = v;
                    else fixbig[atomicAdd(counters + C_FIXBIG, 1)] = v;
                }
            }
        }
        __syncthreads();  // the set, list and counters are reused by the next vertex
    }
}

// ---- VT rows from the routed corner records -------------------------------------
// VT[v] lists the cells holding v in ascending cell order, as the reference's
// elective_make_VT appends them (src/cpu_extraction.cpp:291-342); here as
// corner codes 4t + j, so the row order is the cell order.
// One workgroup per bucket.  The bucket's records (v, 4t + j) are counted per
// vertex and placed per row (LDS, or the dead level-1 array `stage` for
// buckets with more than VT_CAP records).  Rows of at most 32 entries are
// sorted in registers, one thread per vertex; every entry of a longer row
// takes its place by counting the smaller codes of that row (codes in a row
// are distinct).  The bucket's rows fill the same range of the CSR that its
// records fill in bucket order, so row offsets need no global scan.
constexpr int VT_THREADS = 256, VT_CAP = 8192;

template <int BV, typename P>
__device__ __forceinline__ void vt_place_rank(const Rec &rec, int32_t cbeg, int32_t cend,
                                              int32_t *cur, const int32_t *off, P st,
                                              int32_t *__restrict__ vtc) {
    const int tid = threadIdx.x;
    for (int32_t i = cbeg + tid; i < cend; i += VT_THREADS) {
        uint32_t v, lo, hi, unused;
        rec_unpack(rec.w0[i], rec.w1[i], rec.w2[i], v, lo, hi, unused);
        const int32_t pos = atomicAdd(&cur[v & (uint32_t)(BV - 1)], 1);
        st[pos] = (int32_t)(lo | (hi << 24));
    }
    // the workgroup barrier orders the staged writes for its own reads, LDS
    // or global; a device-scope fence here would write back L2 per block
    __syncthreads();
    // rows of at most 32 entries: one thread per vertex sorts its row in
    // registers; longer rows are listed and ranked by the whole block
    __shared__ int32_t nlong;
    __shared__ int32_t lng[BV];
    if (tid == 0) nlong = 0;
    __syncthreads();
    if (tid < BV) {
        const int32_t rb = off[tid], d = off[tid + 1] - rb;
        if (d > 32) {
            lng[atomicAdd(&nlong, 1)] = tid;
        } else if (d > 0) {
            uint32_t r[32];
#pragma unroll
            for (int i = 0; i < 32; i++) r[i] = i < d ? (uint32_t)st[rb + i] : 0xFFFFFFFFu;
            sort32(r);
#pragma unroll
            for (int i = 0; i < 32; i++)
                if (i < d) vtc[cbeg + rb + i] = (int32_t)r[i];
        }
    }
    __syncthreads();
    for (int k = 0; k < nlong; k++) {
        const int vl = lng[k];
        const int32_t rb = off[vl], re = off[vl + 1];
        for (int32_t p = rb + tid; p < re; p += VT_THREADS) {
            const int32_t e = st[p];
            int32_t r = 0;
#pragma unroll 4
            for (int32_t q = rb; q < re; q++) r += st[q] < e;
            vtc[cbeg + rb + r] = e;
        }
    }
}

template <int BV_LOG>
__global__ __launch_bounds__(VT_THREADS) void k_bkt_vt(const Rec rec,
                                                        const int32_t *__restrict__ h2s,
                                                        int32_t nslice, int32_t V, int32_t nb,
                                                        int32_t *__restrict__ stage,
                                                        int32_t *__restrict__ vt_off,
                                                        int32_t *__restrict__ vtc,
                                                        int32_t *__restrict__ vt_cnt) {
    constexpr int BV = 1 << BV_LOG;
    static_assert(BV <= VT_THREADS, "one thread per vertex in the scan");
    __shared__ int32_t cnt[BV], off[BV + 1], cur[BV];
    __shared__ int32_t buf[VT_CAP];
    __shared__ int32_t wsum[VT_THREADS / 64];
    const int tid = threadIdx.x, lane = __lane_id(), wid = tid / 64;
    const int32_t b = blockIdx.x, v0 = b << BV_LOG;
    const int nv = min(BV, V - v0);
    const int32_t cbeg = h2s[(int64_t)b * nslice], cend = h2s[(int64_t)(b + 1) * nslice];
    const int32_t n = cend - cbeg;
    for (int i = tid; i < BV; i += VT_THREADS) cnt[i] = 0;
    __syncthreads();
    for (int32_t i = cbeg + tid; i < cend; i += VT_THREADS)
        atomicAdd(&cnt[rec.w0[i] & (uint32_t)(BV - 1)], 1);
    __syncthreads();
    const int32_t x = tid < BV ? cnt[tid] : 0;
    const int32_t inc = tgd::wave_inclusive_sum(x);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    int32_t pre = 0;
#pragma unroll
    for (int w = 0; w < VT_THREADS / 64; w++)
        if (w < wid) pre += wsum[w];
    if (tid < BV) {
        off[tid] = pre + inc - x;
        cur[tid] = pre + inc - x;
    }
    if (tid == 0) off[BV] = n;
    __syncthreads();
    if (tid < nv) {
        vt_off[v0 + tid] = cbeg + off[tid];
        vt_cnt[v0 + tid] = x;
    }
    if (b == nb - 1 && tid == 0) vt_off[V] = cend;
    // LDS staging when the bucket fits, else the dead level-1 array (a
    // separate instantiation keeps the LDS accesses on LDS instructions)
    if (n <= VT_CAP) vt_place_rank<BV>(rec, cbeg, cend, cur, off, buf, vtc);
    else vt_place_rank<BV>(rec, cbeg, cend, cur, off, stage + cbeg, vtc);
}

// Route the corner records of every cell to the buckets of their vertex:
// r2 holds them in bucket order, bucket b in [ranges[b * nsl], ranges[(b + 1)
// * nsl]).  One level (ranges = h2[0..nb]) or two levels (h2 is
// [bucket][slice]); the choice is made on the mesh's first run.  VTREC:
// records carry the corner code 4t + j (low 24 bits, high bits) instead of
// the other three corners.
template <bool VTREC>
int route_records(tg::BucketBuffers &k, hipStream_t s, const int32_t *&ranges, int32_t &nsl) {
    const tg::BucketPlan &p = k.plan;
    tg::VvCpBuffers &b = k.base;
    const int64_t T = b.n_cells;
    const int4 *tv4 = (const int4 *)b.tv;
    tg::StageMarks *mk = b.marks;
    const int64_t n = 4 * T;
    const Rec r1{k.rec1, k.rec1 + n, k.rec1 + 2 * n};
    const Rec r2{k.rec2, k.rec2 + n, k.rec2 + 2 * n};
    const int cshift = p.bv_log + p.fs;
    int rc = tg::OK;
    int route = k.route ? *k.route : tg::ROUTE_TWO;
    ranges = k.h2;
    nsl = p.slices;
    if (route != tg::ROUTE_TWO) {
        int32_t *bstart = k.h2, *bused = k.h1;
        if ((rc = tg::fill_i32(bstart, 0, p.nb, s)) || (rc = tg::fill_i32(bused, 0, p.nb, s)))
            return rc;
        hipLaunchKernelGGL(k_s1_count, dim3(p.nchunks), dim3(S1_THREADS), 0, s, tv4, T, p.tpc,
                           p.bv_log, bstart, b.counters);
        TG_LAUNCH_CHECK("k_s1_count");
        if (route == tg::ROUTE_AUTO) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            TG_HIP(hipStreamIsCapturing(s, &cs));
            if (cs != hipStreamCaptureStatusNone) {
                route = tg::ROUTE_TWO;  // no waiting under capture; decide on an eager run
            } else {
                int32_t wide = 0;
                TG_HIP(hipMemcpyAsync(&wide, b.counters + tg::C_WIDE, sizeof wide,
                                      hipMemcpyDeviceToHost, s));
                TG_HIP(hipStreamSynchronize(s));
                const int64_t per = (p.tpc + S1_TILE - 1) / S1_TILE;
                const int64_t last = T - (int64_t)(p.nchunks - 1) * p.tpc;
                const int64_t tiles = (int64_t)(p.nchunks - 1) * per + (last + S1_TILE - 1) / S1_TILE;
                route = (int64_t)wide * 64 <= tiles ? tg::ROUTE_ONE : tg::ROUTE_TWO;
                *k.route = route;
            }
        }
        if (route == tg::ROUTE_ONE) {
            if ((rc = tg::scan_exclusive_i32(bstart, bstart, p.nb, b.scan_tmp, s))) return rc;
            if (mk) mk->mark("s1_count", s);
            hipLaunchKernelGGL(k_s1_scatter<VTREC>, dim3(p.nchunks), dim3(S1_THREADS), 0, s, tv4, T,
                               p.tpc, p.bv_log, bstart, bused, r2);
            TG_LAUNCH_CHECK("k_s1_scatter");
            if (mk) mk->mark("s1_scatter", s);
            ranges = bstart;
            nsl = 1;
        }
    }
    if (route == tg::ROUTE_TWO) {
        hipLaunchKernelGGL(k_l1_count, dim3(p.nchunks), dim3(L1_THREADS), 0, s, tv4, T, p.tpc,
                           cshift, p.nc, p.nchunks, k.h1);
        TG_LAUNCH_CHECK("k_l1_count");
        if ((rc = tg::scan_exclusive_i32(k.h1, k.h1, (int64_t)p.nc * p.nchunks, b.scan_tmp, s)))
            return rc;
        if (mk) mk->mark("l1_count", s);
        hipLaunchKernelGGL(k_l1_scatter<VTREC>, dim3(p.nchunks), dim3(L1_THREADS), 0, s, tv4, T, p.tpc,
                           cshift, p.nc, p.nchunks, k.h1, r1);
        TG_LAUNCH_CHECK("k_l1_scatter");
        if (mk) mk->mark("l1_scatter", s);
        const dim3 g2(p.nc, p.slices);
        hipLaunchKernelGGL(k_l2_count, g2, dim3(L2_THREADS), 0, s, r1.w0, k.h1, p.nchunks,
                           p.bv_log, p.fs, p.nb, k.h2);
        TG_LAUNCH_CHECK("k_l2_count");
        if ((rc = tg::scan_exclusive_i32(k.h2, k.h2, (int64_t)p.nb * p.slices, b.scan_tmp, s)))
            return rc;
        if (mk) mk->mark("l2_count", s);
        hipLaunchKernelGGL(k_l2_scatter, g2, dim3(L2_THREADS), 0, s, r1, k.h1, p.nchunks,
                           p.bv_log, p.fs, p.nb, k.h2, r2);
        TG_LAUNCH_CHECK("k_l2_scatter");
        if (mk) mk->mark("l2_scatter", s);
    }
    return tg::OK;
}

template <typename S>
int launch_bucketed_typed(tg::BucketBuffers &k, bool classify, hipStream_t s) {
    const tg::BucketPlan &p = k.plan;
    tg::VvCpBuffers &b = k.base;
    const int32_t V = (int32_t)b.n_points;
    const int64_t T = b.n_cells;
    const int cls = classify ? 1 : 0;
    const S *scal = (const S *)b.scalars;
    tg::StageMarks *mk = b.marks;
    const int64_t n = 4 * T;
    const Rec r2{k.rec2, k.rec2 + n, k.rec2 + 2 * n};
    const int32_t *ranges = nullptr;
    int32_t nsl = 1;
    int rc = route_records<false>(k, s, ranges, nsl);
    if (rc) return rc;
#define TG_BKT(KERNEL)                                                                       \
    hipLaunchKernelGGL(KERNEL, dim3(p.nb), dim3(p.bv_log == 7 ? 256 : 512), 0, s, r2, ranges, nsl, \
                       scal, V, cls, b.scratch, b.deg, k.rowsrc, b.classes, b.heavy, b.fixd,   \
                       b.rowdesc, b.counters)
    if (p.owner) {
        if (p.bv_log == 7) TG_BKT((k_bkt_own<7, S>));
        else TG_BKT((k_bkt_own<8, S>));
    } else if (p.ins == 1) {
        if (p.bv_log == 7) TG_BKT((k_bkt_vv_cp<7, S, 1>));
        else TG_BKT((k_bkt_vv_cp<8, S, 1>));
    } else {
        if (p.bv_log == 7) TG_BKT((k_bkt_vv_cp<7, S, 0>));
        else TG_BKT((k_bkt_vv_cp<8, S, 0>));
    }
#undef TG_BKT
    TG_LAUNCH_CHECK("k_bkt_vv_cp");
    if (mk) mk->mark("bkt_vv_cp", s);
    if (p.bv_log == 7)
        hipLaunchKernelGGL((k_bkt_heavy<7, S>), dim3(256), dim3(HV_THREADS), 0, s, r2, ranges, nsl,
                           scal, cls, b.scratch, b.deg, k.rowsrc, b.classes, b.heavy, b.fix,
                           b.fixbig, b.counters);
    else
        hipLaunchKernelGGL((k_bkt_heavy<8, S>), dim3(256), dim3(HV_THREADS), 0, s, r2, ranges, nsl,
                           scal, cls, b.scratch, b.deg, k.rowsrc, b.classes, b.heavy, b.fix,
                           b.fixbig, b.counters);
    TG_LAUNCH_CHECK("k_bkt_heavy");
    if (mk) mk->mark("bkt_heavy", s);
    return tg::OK;
}

int ceil_log2(int64_t x) {
    int l = 0;
    while (((int64_t)1 << l) < x) l++;
    return l;
}

}  // namespace

namespace tg {

BucketPlan plan_buckets(int64_t n_cells, int64_t n_points, int bv_log, int64_t tpc) {
    BucketPlan p;
    const int64_t V = n_points, T = n_cells;
    p.bv_log = bv_log;
    const int64_t nb = std::max<int64_t>(1, (V + (1 << bv_log) - 1) >> bv_log);
    p.nb = (int32_t)nb;
    p.fs = std::max(0, ceil_log2(nb) - 8);  // coarse bins <= 256
    p.nc = (int32_t)((nb + (1 << p.fs) - 1) >> p.fs);
    p.ok = V > 0 && V <= (int64_t)M24 && (1 << p.fs) <= MAX_FINE && p.nc <= MAX_COARSE;
    p.tpc = std::max<int64_t>(1024, tpc);
    p.nchunks = (int32_t)std::max<int64_t>(1, (T + p.tpc - 1) / p.tpc);
    // about 2048 level-2 workgroups: several rounds of one resident block per CU
    p.slices = std::max(1, std::min(64, (2048 + p.nc - 1) / p.nc));
    return p;
}

int launch_bucketed_vv_cp(BucketBuffers &k, bool classify, hipStream_t s) {
    return k.base.scalars_f64 ? launch_bucketed_typed<double>(k, classify, s)
                              : launch_bucketed_typed<float>(k, classify, s);
}

int run_vt_bucketed(BucketBuffers &k, int32_t *vt_off, int32_t *vtc, int32_t *vt_cnt,
                    hipStream_t s) {
    const BucketPlan &p = k.plan;
    VvCpBuffers &b = k.base;
    const int32_t V = (int32_t)b.n_points;
    if (V <= 0) return fill_i32(vt_off, 0, 1, s);
    if (4 * b.n_cells > (int64_t)INT32_MAX) return ERR_CAPACITY;
    int rc = fill_i32(b.counters + C_WIDE, 0, 1, s);  // the routing choice counts wide tiles
    if (rc) return rc;
    const int32_t *ranges = nullptr;
    int32_t nsl = 1;
    if ((rc = route_records<true>(k, s, ranges, nsl))) return rc;
    const int64_t n = 4 * b.n_cells;
    const Rec r2{k.rec2, k.rec2 + n, k.rec2 + 2 * n};
    int32_t *stage = (int32_t *)k.rec1;
    if (p.bv_log == 7)
        hipLaunchKernelGGL((k_bkt_vt<7>), dim3(p.nb), dim3(VT_THREADS), 0, s, r2, ranges, nsl, V,
                           p.nb, stage, vt_off, vtc, vt_cnt);
    else
        hipLaunchKernelGGL((k_bkt_vt<8>), dim3(p.nb), dim3(VT_THREADS), 0, s, r2, ranges, nsl, V,
                           p.nb, stage, vt_off, vtc, vt_cnt);
    TG_LAUNCH_CHECK("k_bkt_vt");
    return OK;
}

}  // namespace tg
