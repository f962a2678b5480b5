// VV relation (vertex -> distinct co-vertices, ascending rows) and critical-
// point classification on gfx950.
//
// Reference: the padded, atomically appended VV_kernel
// (src/cuda_extraction.cu:960-1057) and the block-per-vertex critPoints
// kernel (src/alg/critPoints.cu:49-338).  Here:
//   1. vertex -> corner incidence by counting (atomics on V counters), an
//      exclusive scan and an atomic-cursor scatter;
//   2. one lane per vertex walks its tets, keeps its distinct neighbours in
//      an LDS list laid out lane-interleaved (element i of lane l at
//      i*BLOCK + l, so concurrent lanes never share a bank), and unions
//      same-class neighbours that share one of its tets with a union-find in
//      LDS;  vertices with more neighbours than the lane list holds go to a
//      serial per-workgroup kernel with a 4096-entry list;
//   3. the sorted row goes to a padded scratch row at 3*vt_off[v] (a vertex
//      in d tets has at most 3d neighbours), rows are compacted after a scan
//      of the degrees;
//   4. the shared-tet link graph is a subgraph of the reference's link graph
//      (link edge (a,c) iff c is in VV(a), critPoints.cu:206-246), so a class
//      whose shared-tet components number 0 or 1 already has the reference
//      count; only vertices with >= 2 components in a class are recomputed
//      with the row-membership rule on the finished VV rows.
#include "bucket.h"

#include <algorithm>

namespace {

using tg::C_ERR;
using tg::C_FIX;
using tg::C_FIXBIG;
using tg::C_HEAVY;
using tg::C_MAXDEG;
using tg::HEAVY_CAP;

// critPoints.cu:136-140: equal scalars fall back to the id comparison.
template <typename S>
__device__ __forceinline__ bool is_lower(S sa, S sv, int32_t a, int32_t v) {
    return (sa == sv) ? (a < v) : (sa < sv);
}

// Numeric verdict written by critPoints.cu:303-330 given the lower and upper
// component counts: 1 minimum, 2 maximum, 3 regular, 4 saddle.
__device__ __forceinline__ uint32_t verdict(uint32_t nl, uint32_t nu) {
    if (nl >= 1 && nu == 0) return 2u;
    if (nl == 0 && nu >= 1) return 1u;
    if (nl == 1 && nu == 1) return 3u;
    return 4u;
}

// Append v to a device list with one atomic per wave.
__device__ __forceinline__ void wave_append(int32_t *list, int32_t *counter,
                                            int32_t v, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (!mask) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (int32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) {
        const uint64_t below = mask & ((1ull << lane) - 1ull);
        list[base + __popcll(below)] = v;
    }
}

// ---- vertex -> corner incidence -------------------------------------------
__global__ void k_count_corners(const int4 *__restrict__ tv4, int64_t T,
                                int32_t *__restrict__ cnt) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv4[t];
        atomicAdd(&cnt[q.x], 1);
        atomicAdd(&cnt[q.y], 1);
        atomicAdd(&cnt[q.z], 1);
        atomicAdd(&cnt[q.w], 1);
    }
}

__global__ void k_fill_corners(const int4 *__restrict__ tv4, int64_t T,
                               const int32_t *__restrict__ off,
                               int32_t *__restrict__ cur,
                               int32_t *__restrict__ vtc) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv4[t];
        const int32_t p = (int32_t)(4 * t);
        vtc[off[q.x] + atomicAdd(&cur[q.x], 1)] = p;
        vtc[off[q.y] + atomicAdd(&cur[q.y], 1)] = p + 1;
        vtc[off[q.z] + atomicAdd(&cur[q.z], 1)] = p + 2;
        vtc[off[q.w] + atomicAdd(&cur[q.w], 1)] = p + 3;
    }
}

// Lane-serial sort of one row: insertion sort for short rows, heapsort
// otherwise.
__device__ void sort_row(int32_t *a, int32_t n) {
    if (n <= 32) {
        for (int32_t i = 1; i < n; i++) {
            const int32_t x = a[i];
            int32_t j = i - 1;
            while (j >= 0 && a[j] > x) {
                a[j + 1] = a[j];
                j--;
            }
            a[j + 1] = x;
        }
        return;
    }
    auto sift = [&](int32_t root, int32_t end) {
        while (2 * root + 1 < end) {
            int32_t child = 2 * root + 1;
            if (child + 1 < end && a[child] < a[child + 1]) child++;
            if (a[root] >= a[child]) return;
            const int32_t t = a[root];
            a[root] = a[child];
            a[child] = t;
            root = child;
        }
    };
    for (int32_t i = n / 2 - 1; i >= 0; i--) sift(i, n);
    for (int32_t end = n - 1; end > 0; end--) {
        const int32_t t = a[0];
        a[0] = a[end];
        a[end] = t;
        sift(0, end);
    }
}

__global__ void k_sort_rows(const int32_t *__restrict__ off, int32_t *vals,
                            int64_t rows) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
         r += (int64_t)gridDim.x * blockDim.x)
        sort_row(vals + off[r], off[r + 1] - off[r]);
}

// ---- lane-per-vertex VV + shared-tet link components ----------------------
template <int CAP, int BLOCK, typename S>
__global__ __launch_bounds__(BLOCK) void k_vv_cp_lane(
    const int4 *__restrict__ tv4, const int32_t *__restrict__ vt_off,
    const int32_t *__restrict__ vtc, const S *__restrict__ scal, int32_t V,
    int32_t *__restrict__ scratch, int32_t *__restrict__ deg,
    uint32_t *__restrict__ classes, int32_t *__restrict__ heavy,
    int32_t *__restrict__ fix, int32_t *__restrict__ counters, int classify) {
    static_assert(CAP <= 64, "lane list is indexed by a 64-bit class mask");
    __shared__ int32_t s_list[CAP * BLOCK];
    __shared__ uint8_t s_par[CAP * BLOCK];
    const int tid = threadIdx.x;
    const int32_t v = (int32_t)blockIdx.x * BLOCK + tid;
    const bool active = v < V;
    int32_t *list = s_list + tid;
    uint8_t *par = s_par + tid;
    int m = 0;
    bool over = false;
    uint64_t low = 0;
    uint32_t nl = 0, nu = 0;
    int32_t beg = 0;
    if (active) {
        beg = vt_off[v];
        const int32_t end = vt_off[v + 1];
        const S sv = classify ? scal[v] : S(0);
        auto find_or_insert = [&](int32_t u) -> int {
            for (int i = 0; i < m; i++)
                if (list[i * BLOCK] == u) return i;
            if (m == CAP) return -1;
            list[m * BLOCK] = u;
            par[m * BLOCK] = (uint8_t)m;
            if (classify && is_lower(scal[u], sv, u, v)) low |= 1ull << m;
            return m++;
        };
        auto root = [&](int x) {
            int p;
            while ((p = par[x * BLOCK]) != x) {
                const int g = par[p * BLOCK];
                par[x * BLOCK] = (uint8_t)g;
                x = g;
            }
            return x;
        };
        auto unite = [&](int a, int b) {
            if (((low >> a) ^ (low >> b)) & 1ull) return;
            a = root(a);
            b = root(b);
            if (a == b) return;
            if (a < b) par[b * BLOCK] = (uint8_t)a;
            else par[a * BLOCK] = (uint8_t)b;
        };
        for (int32_t k = beg; k < end; k++) {
            const int32_t p = vtc[k];
            const int4 q = tv4[p >> 2];
            const int jv = p & 3;
            const int32_t o0 = jv == 0 ? q.y : q.x;
            const int32_t o1 = jv <= 1 ? q.z : q.y;
            const int32_t o2 = jv <= 2 ? q.w : q.z;
            const int i0 = find_or_insert(o0);
            const int i1 = find_or_insert(o1);
            const int i2 = find_or_insert(o2);
            if ((i0 | i1 | i2) < 0) {
                over = true;
                break;
            }
            if (classify) {
                unite(i0, i1);
                unite(i0, i2);
                unite(i1, i2);
            }
        }
        if (!over) {
            if (classify) {
                for (int i = 0; i < m; i++)
                    if (par[i * BLOCK] == i) {
                        if ((low >> i) & 1ull) nl++;
                        else nu++;
                    }
                classes[3 * (int64_t)v] = nl;
                classes[3 * (int64_t)v + 1] = nu;
                classes[3 * (int64_t)v + 2] = verdict(nl, nu);
            }
            deg[v] = m;
            int32_t *dst = scratch + 3 * (int64_t)beg;
            for (int i = 0; i < m; i++) {
                const int32_t x = list[i * BLOCK];
                int r = 0;
                for (int j = 0; j < m; j++) r += list[j * BLOCK] < x;
                dst[r] = x;
            }
        }
    }
    wave_append(heavy, counters + C_HEAVY, v, active && over);
    wave_append(fix, counters + C_FIX, v,
                active && !over && classify && (nl >= 2 || nu >= 2));
    int wmax = (active && !over) ? m : 0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) wmax = max(wmax, __shfl_xor(wmax, d, 64));
    if (__lane_id() == 0 && wmax > 0) atomicMax(counters + C_MAXDEG, wmax);
}

// Serial fallback for vertices whose neighbour list overflows the lane list:
// lane 0 of each workgroup handles one vertex with a 4096-entry LDS list.
template <typename S>
__global__ __launch_bounds__(64) void k_vv_cp_heavy(
    const int4 *__restrict__ tv4, const int32_t *__restrict__ vt_off,
    const int32_t *__restrict__ vtc, const S *__restrict__ scal,
    int32_t *__restrict__ scratch, int32_t *__restrict__ deg,
    uint32_t *__restrict__ classes, const int32_t *__restrict__ heavy,
    int32_t *__restrict__ fix, int32_t *__restrict__ fixbig,
    int32_t *__restrict__ counters, int classify) {
    __shared__ int32_t list[HEAVY_CAP];
    __shared__ uint16_t par[HEAVY_CAP];
    __shared__ uint8_t lowf[HEAVY_CAP];
    if (threadIdx.x != 0) return;
    const int32_t n = counters[C_HEAVY];
    for (int32_t h = blockIdx.x; h < n; h += gridDim.x) {
        const int32_t v = heavy[h];
        const int32_t beg = vt_off[v], end = vt_off[v + 1];
        const S sv = classify ? scal[v] : S(0);
        int m = 0;
        bool over = false;
        auto find_or_insert = [&](int32_t u) -> int {
            for (int i = 0; i < m; i++)
                if (list[i] == u) return i;
            if (m == HEAVY_CAP) return -1;
            list[m] = u;
            par[m] = (uint16_t)m;
            lowf[m] = classify ? (uint8_t)is_lower(scal[u], sv, u, v) : 0;
            return m++;
        };
        auto root = [&](int x) {
            while (par[x] != x) {
                par[x] = par[par[x]];
                x = par[x];
            }
            return x;
        };
        auto unite = [&](int a, int b) {
            if (lowf[a] != lowf[b]) return;
            a = root(a);
            b = root(b);
            if (a == b) return;
            if (a < b) par[b] = (uint16_t)a;
            else par[a] = (uint16_t)b;
        };
        for (int32_t k = beg; k < end && !over; k++) {
            const int32_t p = vtc[k];
            const int4 q = tv4[p >> 2];
            const int jv = p & 3;
            const int i0 = find_or_insert(jv == 0 ? q.y : q.x);
            const int i1 = find_or_insert(jv <= 1 ? q.z : q.y);
            const int i2 = find_or_insert(jv <= 2 ? q.w : q.z);
            if ((i0 | i1 | i2) < 0) {
                over = true;
                break;
            }
            if (classify) {
                unite(i0, i1);
                unite(i0, i2);
                unite(i1, i2);
            }
        }
        if (over) {
            atomicOr(counters + C_ERR, tg::ERRBIT_CAPACITY);
            deg[v] = 0;
            if (classify) {
                classes[3 * (int64_t)v] = 0;
                classes[3 * (int64_t)v + 1] = 0;
                classes[3 * (int64_t)v + 2] = 0;
            }
            continue;
        }
        uint32_t nl = 0, nu = 0;
        if (classify) {
            for (int i = 0; i < m; i++)
                if (par[i] == i) {
                    if (lowf[i]) nl++;
                    else nu++;
                }
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
        deg[v] = m;
        atomicMax(counters + C_MAXDEG, m);
        // ascending row via insertion sort in LDS
        for (int i = 1; i < m; i++) {
            const int32_t x = list[i];
            int j = i - 1;
            while (j >= 0 && list[j] > x) {
                list[j + 1] = list[j];
                j--;
            }
            list[j + 1] = x;
        }
        int32_t *dst = scratch + 3 * (int64_t)beg;
        for (int i = 0; i < m; i++) dst[i] = list[i];
        if (classify && (nl >= 2 || nu >= 2)) {
            if (m <= 64) fix[atomicAdd(counters + C_FIX, 1)] = v;
            else fixbig[atomicAdd(counters + C_FIXBIG, 1)] = v;
        }
    }
}

// ---- compaction of padded rows into the CSR ---------------------------------
__global__ __launch_bounds__(256) void k_vv_compact(
    const int32_t *__restrict__ vt_off, const int32_t *__restrict__ vv_off,
    const int32_t *__restrict__ scratch, int32_t V, int32_t *__restrict__ cols) {
    __shared__ int32_t s_off[257];
    __shared__ int64_t s_src[256];
    const int32_t v0 = (int32_t)blockIdx.x * 256;
    const int nv = min(256, V - v0);
    for (int i = threadIdx.x; i <= nv; i += 256) {
        s_off[i] = vv_off[v0 + i];
        if (i < nv) s_src[i] = 3 * (int64_t)vt_off[v0 + i];
    }
    __syncthreads();
    const int32_t ob = s_off[0], oe = s_off[nv];
    for (int32_t o = ob + threadIdx.x; o < oe; o += 256) {
        int lo = 0, hi = nv - 1;  // last row r with s_off[r] <= o
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= o) lo = mid;
            else hi = mid - 1;
        }
        cols[o] = scratch[s_src[lo] + (o - s_off[lo])];
    }
}

__global__ __launch_bounds__(256) void k_rows_compact(
    const int32_t *__restrict__ rowsrc, const int32_t *__restrict__ vv_off,
    const int32_t *__restrict__ scratch, int32_t V, int32_t *__restrict__ cols) {
    __shared__ int32_t s_off[257];
    __shared__ int32_t s_src[256];
    const int32_t v0 = (int32_t)blockIdx.x * 256;
    const int nv = min(256, V - v0);
    for (int i = threadIdx.x; i <= nv; i += 256) {
        s_off[i] = vv_off[v0 + i];
        if (i < nv) s_src[i] = rowsrc[v0 + i];
    }
    __syncthreads();
    const int32_t ob = s_off[0], oe = s_off[nv];
    // rows that sit back to back in scratch (a bucket of 256 vertices with
    // no heavy vertex) move as one run
    const int32_t shift = s_src[0] - ob;
    bool run = true;
    for (int i = threadIdx.x; i < nv; i += 256) run = run && s_src[i] - s_off[i] == shift;
    if (__syncthreads_and(run)) {
        for (int32_t o = ob + threadIdx.x; o < oe; o += 256) cols[o] = scratch[(int64_t)o + shift];
        return;
    }
    for (int32_t o = ob + threadIdx.x; o < oe; o += 256) {
        int lo = 0, hi = nv - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= o) lo = mid;
            else hi = mid - 1;
        }
        cols[o] = scratch[(int64_t)s_src[lo] + (o - s_off[lo])];
    }
}

// ---- row-membership link components for flagged vertices -------------------
// Link edge (a, c) between two neighbours a < c of v exists iff c is in
// VV(a) (critPoints.cu:231-233).

__device__ __forceinline__ uint32_t n_components32(const uint32_t *adj, uint32_t mask) {
    uint32_t cnt = 0, rem = mask;
    while (rem) {
        uint32_t comp = rem & (0u - rem);
        uint32_t front = comp;
        while (front) {
            const int f = __ffs(front) - 1;
            front &= front - 1u;
            const uint32_t nb = adj[f] & mask & ~comp;
            comp |= nb;
            front |= nb;
        }
        rem &= ~comp;
        cnt++;
    }
    return cnt;
}

__device__ __forceinline__ uint32_t n_components64(const uint64_t *adj, uint64_t mask) {
    uint32_t cnt = 0;
    uint64_t rem = mask;
    while (rem) {
        uint64_t comp = rem & (0ull - rem);
        uint64_t front = comp;
        while (front) {
            const int f = __ffsll((long long)front) - 1;
            front &= front - 1ull;
            const uint64_t nb = adj[f] & mask & ~comp;
            comp |= nb;
            front |= nb;
        }
        rem &= ~comp;
        cnt++;
    }
    return cnt;
}

// One wave per queued vertex v (degree <= 64): lane i holds neighbour
// a = VV(v)[i], walks VV(a) and marks each member of VV(v) it meets, giving
// a 64-bit adjacency row of the link graph under the row-membership rule;
// lane 0 then counts lower and upper components by bit-parallel BFS.
template <typename S>
__global__ __launch_bounds__(64) void k_cp_fix_wave(
    const int32_t *__restrict__ off, const int32_t *__restrict__ cols,
    const S *__restrict__ scal, const int32_t *__restrict__ fix,
    int32_t *__restrict__ fixbig, int32_t *__restrict__ counters,
    uint32_t *__restrict__ classes) {
    __shared__ int32_t nbs[64];
    __shared__ uint64_t adjm[64];
    const int lane = threadIdx.x;
    const int32_t n = counters[C_FIX];
    for (int32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t v = fix[i];
        const int32_t b = off[v];
        const int d = off[v + 1] - b;
        if (d > 64) {
            if (lane == 0) fixbig[atomicAdd(counters + C_FIXBIG, 1)] = v;
            continue;
        }
        const bool on = lane < d;
        const int32_t a = on ? cols[b + lane] : INT32_MAX;
        const bool lo = on && is_lower(scal[a], scal[v], a, v);
        const uint64_t lowmask = __ballot(lo);
        const uint64_t full = d == 64 ? ~0ull : ((1ull << d) - 1ull);
        nbs[lane] = a;
        __syncthreads();
        uint64_t row = 0;
        if (on) {
            const int32_t rb = off[a], re = off[a + 1];
            for (int32_t j0 = rb; j0 < re; j0 += 8) {
                int32_t c[8];
#pragma unroll
                for (int k = 0; k < 8; k++) c[k] = j0 + k < re ? cols[j0 + k] : INT32_MAX;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    int lo2 = 0, hi2 = d;  // first index with nbs >= c
                    while (lo2 < hi2) {
                        const int mid = (lo2 + hi2) >> 1;
                        if (nbs[mid] < c[k]) lo2 = mid + 1;
                        else hi2 = mid;
                    }
                    if (lo2 < d && nbs[lo2] == c[k]) row |= 1ull << lo2;
                }
            }
        }
        adjm[lane] = row;
        __syncthreads();
        if (lane == 0) {
            const uint32_t nl = n_components64(adjm, lowmask & full);
            const uint32_t nu = n_components64(adjm, ~lowmask & full);
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
        __syncthreads();
    }
}

// Link pass driven by the row descriptors of the bucketed form: for a queued
// vertex v (degree <= 32) each neighbour carries its class and the number of
// its shared-tet component.  Only neighbours outside the largest component of
// a multi-component class walk their rows; every row-membership edge between
// two components of one class joins them, and the classes' counts are the
// components of that component graph.  Two vertices per wave, 32 lanes each.
__global__ __launch_bounds__(64) void k_cp_fix_desc(
    const int32_t *__restrict__ off, const int32_t *__restrict__ cols,
    const int32_t *__restrict__ rowsrc, const uint8_t *__restrict__ rowdesc,
    const int32_t *__restrict__ fixd, int32_t *__restrict__ fix,
    int32_t *__restrict__ counters, uint32_t *__restrict__ classes) {
    __shared__ int32_t nbs[2][32];
    __shared__ uint8_t dsc[2][32];
    __shared__ uint32_t cadj[2][2][32];
    const int lane = threadIdx.x & 31, half = threadIdx.x >> 5;
    const int32_t n = counters[tg::C_FIXD];
    const uint64_t hmask = half ? 0xFFFFFFFF00000000ull : 0xFFFFFFFFull;
    // the next vertex's id and row bounds are loaded during the current one
    // (the pass is bound by chains of dependent loads)
    auto head = [&](int32_t i, int32_t &v, int32_t &b, int32_t &e, int32_t &rs) {
        v = i < n ? fixd[i] : 0;
        b = i < n ? off[v] : 0;
        e = i < n ? off[v + 1] : 0;
        rs = i < n ? rowsrc[v] : 0;
    };
    int32_t nv_, nb_, ne_, nrs_;
    head(2 * blockIdx.x + half, nv_, nb_, ne_, nrs_);
    for (int32_t i0 = 2 * blockIdx.x; i0 < n; i0 += 2 * gridDim.x) {
        const int32_t i = i0 + half;
        const bool live = i < n;
        const int32_t v = nv_, b = nb_, rs = nrs_;
        int d = live ? ne_ - b : 0;
        head(i + 2 * gridDim.x, nv_, nb_, ne_, nrs_);
        if (d > 32) {  // not from the bucketed kernel: generic pass
            if (lane == 0) fix[atomicAdd(counters + C_FIX, 1)] = v;
            d = 0;
        }
        const bool on = lane < d;
        const int32_t a = on ? cols[b + lane] : INT32_MAX;
        const uint32_t ds = on ? rowdesc[rs + lane] : 0u;
        const int cls = (int)(ds >> 5), comp = (int)(ds & 31u);
        nbs[half][lane] = a;
        dsc[half][lane] = (uint8_t)ds;
        cadj[half][0][lane] = 0u;
        cadj[half][1][lane] = 0u;
        int nc1 = on && cls == 1 ? comp + 1 : 0, nc0 = on && cls == 0 ? comp + 1 : 0;
#pragma unroll
        for (int s = 16; s >= 1; s >>= 1) {
            nc1 = max(nc1, __shfl_xor(nc1, s, 32));
            nc0 = max(nc0, __shfl_xor(nc0, s, 32));
        }
        // largest component of each multi-component class (lowest number on
        // ties); the ballots run over the wave, each half reads its 32 bits
        int big0 = 0, big1 = 0;
        const int ncmax = max(__shfl(max(nc0, nc1), 0, 64), __shfl(max(nc0, nc1), 32, 64));
        int best0 = -1, best1 = -1;
        for (int j = 0; j < ncmax; j++) {
            const int s0 = __popcll(__ballot(on && cls == 0 && comp == j) & hmask);
            const int s1 = __popcll(__ballot(on && cls == 1 && comp == j) & hmask);
            if (j < nc0 && s0 > best0) { best0 = s0; big0 = j; }
            if (j < nc1 && s1 > best1) { best1 = s1; big1 = j; }
        }
        __syncthreads();
        const int nck = cls ? nc1 : nc0;
        const int big = cls ? big1 : big0;
        if (on && nck >= 2 && comp != big) {
            const int32_t rb = off[a], re = off[a + 1];
            for (int32_t j0 = rb; j0 < re; j0 += 16) {
                int32_t c[16];
#pragma unroll
                for (int k = 0; k < 16; k++) c[k] = j0 + k < re ? cols[j0 + k] : INT32_MAX;
#pragma unroll
                for (int k = 0; k < 16; k++) {
                    // lower bound in the 32 sorted entries (INT32_MAX past d):
                    // fixed steps, so the 16 searches overlap their LDS reads
                    int lo = 0;
#pragma unroll
                    for (int step = 16; step >= 1; step >>= 1)
                        if (nbs[half][lo + step - 1] < c[k]) lo += step;
                    if (lo < d && nbs[half][lo] == c[k]) {
                        const int dk = dsc[half][lo];
                        const int ck = dk & 31;
                        if ((dk >> 5) == cls && ck != comp) {
                            atomicOr(&cadj[half][cls][comp], 1u << ck);
                            atomicOr(&cadj[half][cls][ck], 1u << comp);
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (lane == 0 && live && d > 0) {
            const uint32_t m1 = nc1 >= 32 ? ~0u : ((1u << nc1) - 1u);
            const uint32_t m0 = nc0 >= 32 ? ~0u : ((1u << nc0) - 1u);
            const uint32_t nl = nc1 >= 2 ? n_components32(cadj[half][1], m1) : (uint32_t)nc1;
            const uint32_t nu = nc0 >= 2 ? n_components32(cadj[half][0], m0) : (uint32_t)nc0;
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
        __syncthreads();
    }
}

// Row-membership link pass for vertices with more than 64 neighbours: one
// workgroup per vertex.  Each thread walks the rows of some neighbours a and
// looks every later entry c up in v's row (binary search in LDS); pairs of
// one class are joined in a lock-free union-find (a root is only linked
// under a smaller root, by CAS).  Same edge set as link_components.
constexpr int FB_THREADS = 256;

__device__ __forceinline__ int fb_root(int32_t *par, int x) {
    while (true) {
        const int p = par[x];
        if (p == x) return x;
        const int g = par[p];
        if (g != p) par[x] = g;
        x = g;
    }
}

__device__ __forceinline__ void fb_unite(int32_t *par, int a, int b) {
    while (true) {
        a = fb_root(par, a);
        b = fb_root(par, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b) return;
    }
}

template <typename S>
__global__ __launch_bounds__(FB_THREADS) void k_cp_fix_big(
    const int32_t *__restrict__ off, const int32_t *__restrict__ cols,
    const S *__restrict__ scal, const int32_t *__restrict__ fixbig,
    int32_t *__restrict__ counters, uint32_t *__restrict__ classes) {
    __shared__ int32_t nb[HEAVY_CAP];
    __shared__ int32_t par[HEAVY_CAP];
    __shared__ uint8_t lowf[HEAVY_CAP];
    __shared__ int32_t cnt[2];
    const int tid = threadIdx.x;
    const int32_t n = counters[C_FIXBIG];
    for (int32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int32_t v = fixbig[i];
        const int32_t b = off[v];
        const int d = off[v + 1] - b;
        if (d > HEAVY_CAP) {  // block-uniform
            if (tid == 0) atomicOr(counters + C_ERR, tg::ERRBIT_CAPACITY);
            continue;
        }
        const S sv = scal[v];
        for (int k = tid; k < d; k += FB_THREADS) {
            const int32_t a = cols[b + k];
            nb[k] = a;
            par[k] = k;
            lowf[k] = (uint8_t)is_lower(scal[a], sv, a, v);
        }
        if (tid < 2) cnt[tid] = 0;
        __syncthreads();
        for (int ia = tid; ia < d; ia += FB_THREADS) {
            const int32_t a = nb[ia];
            const uint8_t la = lowf[ia];
            const int32_t je = off[a + 1];
            for (int32_t j = off[a]; j < je; j++) {
                const int32_t c = cols[j];
                if (c <= a) continue;
                int lo = 0, hi = d;  // first entry of v's row >= c
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (nb[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < d && nb[lo] == c && lowf[lo] == la) fb_unite(par, ia, lo);
            }
        }
        __syncthreads();
        for (int k = tid; k < d; k += FB_THREADS)
            if (par[k] == k) atomicAdd(&cnt[lowf[k] ? 0 : 1], 1);
        __syncthreads();
        if (tid == 0) {
            const uint32_t nl = (uint32_t)cnt[0], nu = (uint32_t)cnt[1];
            classes[3 * (int64_t)v] = nl;
            classes[3 * (int64_t)v + 1] = nu;
            classes[3 * (int64_t)v + 2] = verdict(nl, nu);
        }
        __syncthreads();  // nb, par, lowf and cnt are reused by the next vertex
    }
}

template <typename S>
int launch_cp_fix_typed(tg::VvCpBuffers &b, hipStream_t s) {
    const int32_t V = (int32_t)b.n_points;
    const S *scal = (const S *)b.scalars;
    const unsigned g = std::min<unsigned>(tg::grid_for(V, 1), 16384u);
    if (b.fixd) {
        hipLaunchKernelGGL(k_cp_fix_desc, dim3(std::max(1u, g / 2)), dim3(64), 0, s, b.vv_off,
                           b.vv_cols, b.rowsrc,
                           b.rowdesc, b.fixd, b.fix, b.counters, b.classes);
        TG_LAUNCH_CHECK("k_cp_fix_desc");
    }
    hipLaunchKernelGGL((k_cp_fix_wave<S>), dim3(g), dim3(64), 0, s, b.vv_off, b.vv_cols, scal,
                       b.fix, b.fixbig, b.counters, b.classes);
    TG_LAUNCH_CHECK("k_cp_fix_wave");
    hipLaunchKernelGGL((k_cp_fix_big<S>), dim3(256), dim3(FB_THREADS), 0, s, b.vv_off, b.vv_cols, scal,
                       b.fixbig, b.counters, b.classes);
    TG_LAUNCH_CHECK("k_cp_fix_big");
    if (b.marks) b.marks->mark("cp_fix", s);
    return tg::OK;
}

template <typename S>
int launch_vv_cp_typed(tg::VvCpBuffers &b, const tg::VvCpConfig &cfg,
                       bool classify, hipStream_t s) {
    const int32_t V = (int32_t)b.n_points;
    const S *scal = (const S *)b.scalars;
    const int4 *tv4 = (const int4 *)b.tv;
    const int cls = classify ? 1 : 0;
    if (V > 0) {
#define TG_LANE(CAP, BLK)                                                          \
    hipLaunchKernelGGL((k_vv_cp_lane<CAP, BLK, S>), dim3(tg::grid_for(V, BLK)),      \
                       dim3(BLK), 0, s, tv4, b.vt_off, b.vtc, scal, V, b.scratch,    \
                       b.deg, b.classes, b.heavy, b.fix, b.counters, cls)
        if (cfg.lane_cap <= 32) {
            if (cfg.lane_block <= 64) TG_LANE(32, 64);
            else if (cfg.lane_block <= 128) TG_LANE(32, 128);
            else TG_LANE(32, 256);
        } else {
            if (cfg.lane_block <= 64) TG_LANE(64, 64);
            else TG_LANE(64, 128);
        }
#undef TG_LANE
        TG_LAUNCH_CHECK("k_vv_cp_lane");
        if (b.marks) b.marks->mark("vv_lane", s);
        hipLaunchKernelGGL((k_vv_cp_heavy<S>), dim3(256), dim3(64), 0, s, tv4, b.vt_off,
                           b.vtc, scal, b.scratch, b.deg, b.classes, b.heavy, b.fix,
                           b.fixbig, b.counters, cls);
        TG_LAUNCH_CHECK("k_vv_cp_heavy");
        if (b.marks) b.marks->mark("vv_heavy", s);
    }
    int rc = tg::scan_exclusive_i32(b.deg, b.vv_off, V, b.scan_tmp, s);
    if (rc) return rc;
    if (b.marks) b.marks->mark("vv_scan", s);
    if (V > 0) {
        hipLaunchKernelGGL(k_vv_compact, dim3(tg::grid_for(V, 256)), dim3(256), 0, s,
                           b.vt_off, b.vv_off, b.scratch, V, b.vv_cols);
        TG_LAUNCH_CHECK("k_vv_compact");
        if (b.marks) b.marks->mark("vv_compact", s);
        if (classify) {
            int rc2 = tg::launch_cp_fix(b, s);
            if (rc2) return rc2;
        }
    }
    return tg::OK;
}

}  // namespace

namespace tg {

void StageMarks::begin(hipStream_t s) {
    n = 0;
    if (!on) return;
    mark("start", s);
}

void StageMarks::mark(const char *stage, hipStream_t s) {
    if (!on || n >= MAX) return;
    if (!ev[n] && hipEventCreate(&ev[n]) != hipSuccess) return;
    name[n] = stage;
    (void)hipEventRecord(ev[n], s);
    n++;
}

int launch_vt_corners(VvCpBuffers &b, int block, hipStream_t s) {
    const int64_t V = b.n_points, T = b.n_cells;
    int rc = fill_i32(b.vt_cnt, 0, V, s);
    if (!rc) rc = fill_i32(b.vt_cur, 0, V, s);
    if (rc) return rc;
    if (b.marks) b.marks->mark("vt_fill0", s);
    const int4 *tv4 = (const int4 *)b.tv;
    unsigned g = grid_for(T, block);
    if (g > 65536) g = 65536;
    if (T > 0) {
        hipLaunchKernelGGL(k_count_corners, dim3(g), dim3(block), 0, s, tv4, T, b.vt_cnt);
        TG_LAUNCH_CHECK("k_count_corners");
        if (b.marks) b.marks->mark("vt_count", s);
    }
    rc = scan_exclusive_i32(b.vt_cnt, b.vt_off, V, b.scan_tmp, s);
    if (rc) return rc;
    if (b.marks) b.marks->mark("vt_scan", s);
    if (T > 0) {
        hipLaunchKernelGGL(k_fill_corners, dim3(g), dim3(block), 0, s, tv4, T, b.vt_off,
                           b.vt_cur, b.vtc);
        TG_LAUNCH_CHECK("k_fill_corners");
        if (b.marks) b.marks->mark("vt_scatter", s);
    }
    return OK;
}

int launch_rows_compact(const int32_t *rowsrc, const int32_t *vv_off, const int32_t *scratch,
                        int32_t V, int32_t *cols, hipStream_t s) {
    if (V <= 0) return OK;
    hipLaunchKernelGGL(k_rows_compact, dim3(grid_for(V, 256)), dim3(256), 0, s, rowsrc, vv_off,
                       scratch, V, cols);
    TG_LAUNCH_CHECK("k_rows_compact");
    return OK;
}

int launch_cp_fix(VvCpBuffers &b, hipStream_t s) {
    if (b.n_points <= 0) return OK;
    return b.scalars_f64 ? launch_cp_fix_typed<double>(b, s) : launch_cp_fix_typed<float>(b, s);
}

int run_vv_cp_bucketed(BucketBuffers &k, bool classify, hipStream_t s) {
    VvCpBuffers &b = k.base;
    if (b.marks) b.marks->begin(s);
    int rc = fill_i32(b.counters, 0, N_COUNTERS, s);
    if (rc) return rc;
    const int32_t V = (int32_t)b.n_points;
    if (V > 0 && (rc = launch_bucketed_vv_cp(k, classify, s))) return rc;
    if ((rc = scan_exclusive_i32(b.deg, b.vv_off, V, b.scan_tmp, s))) return rc;
    if (b.marks) b.marks->mark("vv_scan", s);
    if ((rc = launch_rows_compact(k.rowsrc, b.vv_off, b.scratch, V, b.vv_cols, s))) return rc;
    if (b.marks) b.marks->mark("vv_compact", s);
    if (classify && (rc = launch_cp_fix(b, s))) return rc;
    return OK;
}

int launch_sort_rows_i32(const int32_t *off, int32_t *vals, int64_t rows,
                         hipStream_t s) {
    if (rows <= 0) return OK;
    unsigned g = grid_for(rows, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_sort_rows, dim3(g), dim3(256), 0, s, off, vals, rows);
    TG_LAUNCH_CHECK("k_sort_rows");
    return OK;
}

int launch_vv_cp(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
                 hipStream_t s) {
    return b.scalars_f64 ? launch_vv_cp_typed<double>(b, cfg, classify, s)
                         : launch_vv_cp_typed<float>(b, cfg, classify, s);
}

int run_vv_cp(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
              hipStream_t s) {
    if (b.marks) b.marks->begin(s);
    int rc = fill_i32(b.counters, 0, N_COUNTERS, s);
    if (!rc) rc = launch_vt_corners(b, cfg.corner_block, s);
    if (!rc) rc = launch_vv_cp(b, cfg, classify, s);
    return rc;
}

int run_vv_cp_on_corners(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
                         hipStream_t s) {
    if (b.marks) b.marks->begin(s);
    int rc = fill_i32(b.counters, 0, N_COUNTERS, s);
    if (!rc) rc = launch_vv_cp(b, cfg, classify, s);
    return rc;
}

}  // namespace tg
