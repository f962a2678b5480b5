// Edge and face ids in the reference's first-occurrence order, and the
// relations derived from them (include/cuda_extraction.h relation set).
//
// Edges (src/cpu_extraction.cpp:28-92): the edge {lo, hi} gets the rank of
// its first position 6c + k (cells ascending, pair k over (0,1) (0,2) (0,3)
// (1,2) (1,3) (2,3)) among all edges' first positions.  One lane per vertex
// v walks v's tets in ascending order and records, for every upper
// neighbour u > v (an entry of the ascending VV row), the first position it
// meets; first positions are flagged in a byte array, packed to a bitmap
// whose word popcounts are scanned, and an edge's id is the rank of its
// first position.  No atomics.
//
// Faces (src/cpu_extraction.cpp:122-222): slot s of cell c is the positional
// triple (cv[lo], cv[mid], cv[hi]) with (lo, mid, hi) = (0,1,2) (0,1,3)
// (0,2,3) (1,2,3), position 4c + s; the key is not sorted.  The cells holding
// all three vertices are ET(lo,mid) n ET(lo,hi); the first occurrence of a
// key is the smallest position among those cells' slots with the same
// positional triple, and face ids are ranks of first occurrences.
//
// The remaining relations are two-pass CSR builds (count, scan, fill) over
// sorted rows: ET(e) = VT(lo) n VT(hi); EF, TT, FF, EE by merging sorted
// lists; FT from the face matches.
#include <algorithm>

#include "ids.h"
#include "tetra_c.h"

namespace {

constexpr int INF = INT32_MAX;

// pair index of corner positions i < j: (0,1)0 (0,2)1 (0,3)2 (1,2)3 (1,3)4 (2,3)5
__device__ __forceinline__ int pair_index(int i, int j) {
    return i == 0 ? j - 1 : (i == 1 ? j + 1 : 5);
}

__device__ __forceinline__ int32_t corner(const int4 &q, int j) {
    return j == 0 ? q.x : j == 1 ? q.y : j == 2 ? q.z : q.w;
}

// positions (lo, mid, hi) of face slot s
__device__ __forceinline__ void face_pos(int s, int &lo, int &mid, int &hi) {
    lo = s == 3 ? 1 : 0;
    mid = s <= 1 ? 1 : 2;
    hi = s == 0 ? 2 : 3;
}

// index of x in the ascending array a[0..n), or -1.  Rows shorter than 128
// take seven fixed steps (no data-dependent loop), so several searches of
// one thread overlap their loads; longer rows bisect.
__device__ __forceinline__ int32_t bsearch_i32(const int32_t *a, int32_t n, int32_t x) {
    int32_t lo = 0;
    if (n < 128) {
#pragma unroll
        for (int step = 64; step >= 1; step >>= 1) {
            const int32_t probe = lo + step - 1;
            if (probe < n && a[probe] < x) lo += step;
        }
    } else {
        int32_t hi = n;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (a[mid] < x) lo = mid + 1;
            else hi = mid;
        }
    }
    return (lo < n && a[lo] == x) ? lo : -1;
}

// Unaligned 16- and 8-byte loads: dword alignment is enough for
// global_load_dwordx4/x2 on gfx950.
typedef int32_t i32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef int32_t i32x2u __attribute__((ext_vector_type(2), aligned(4)));

// Merges of a few short sorted rows (ET, TT, EE) stage the rows in LDS
// first: one column per thread (consecutive threads in consecutive banks),
// filled by independent 16-byte loads, so a row costs one memory latency
// instead of one per element; the merge then reads LDS.  Whole quads are
// read, up to three words past the row (the source array carries that
// slack), and all four words go to the column: the words past the row land
// where the next row is staged, or in the column's three spare slots after
// the last row.  A thread whose rows exceed the column merges from global
// memory.
template <int W>
__device__ __forceinline__ void stage_quads(int32_t (*buf)[W], int at,
                                            const int32_t *__restrict__ src, int32_t n,
                                            int shift) {
#pragma unroll 4
    for (int32_t k = 0; k < n; k += 4) {
        const i32x4u v = *reinterpret_cast<const i32x4u *>(src + k);
        buf[at + k][threadIdx.x] = v.x >> shift;
        buf[at + k + 1][threadIdx.x] = v.y >> shift;
        buf[at + k + 2][threadIdx.x] = v.z >> shift;
        buf[at + k + 3][threadIdx.x] = v.w >> shift;
    }
}

// A thread's output row written in 16-byte stores: words gather in a
// register quad, stored when full; the last, partial quad is stored whole,
// so the row's slot must start 16-byte aligned and span a multiple of four
// words.  One store per four words instead of one per word: scattered
// per-thread rows cost the vector memory path per store instruction.
struct QuadOut {
    int32_t *row;
    int64_t n = 0;
    int4 q = make_int4(0, 0, 0, 0);
    __device__ __forceinline__ void put(int32_t v) {
        const int k = (int)(n & 3);
        q.x = k == 0 ? v : q.x;
        q.y = k == 1 ? v : q.y;
        q.z = k == 2 ? v : q.z;
        q.w = k == 3 ? v : q.w;
        if (k == 3) *reinterpret_cast<int4 *>(row + n - 3) = q;
        n++;
    }
    __device__ __forceinline__ void flush() {
        if (n & 3) *reinterpret_cast<int4 *>(row + (n & ~(int64_t)3)) = q;
    }
};

// QuadOut for a row at its final, exact position: full quads in unaligned
// 16-byte stores, the last one to three words one by one (the words after
// the row belong to the next row).
struct QuadOutExact {
    int32_t *row;
    int64_t n = 0;
    int4 q = make_int4(0, 0, 0, 0);
    __device__ __forceinline__ void put(int32_t v) {
        const int k = (int)(n & 3);
        q.x = k == 0 ? v : q.x;
        q.y = k == 1 ? v : q.y;
        q.z = k == 2 ? v : q.z;
        q.w = k == 3 ? v : q.w;
        if (k == 3) *reinterpret_cast<i32x4u *>(row + n - 3) = i32x4u{q.x, q.y, q.z, q.w};
        n++;
    }
    __device__ __forceinline__ void flush() {
        const int k = (int)(n & 3);
        int32_t *tail = row + (n & ~(int64_t)3);
        if (k > 0) tail[0] = q.x;
        if (k > 1) tail[1] = q.y;
        if (k > 2) tail[2] = q.z;
    }
};

// ---- edges --------------------------------------------------------------------
// Lane per vertex v: first position of every upper edge (v, u>v), flag it.
// Every pair of v's cells whose low vertex is v also records the VV index of
// its edge in te (replaced by the edge id once ids exist): the row search is
// done here, on v's own row, instead of once per cell on random rows.
// Rows up to EDGE_CAP entries are staged in LDS (one column per thread: the
// VV row and the first positions), and the vertex's cells are fetched eight
// at a time with independent loads; longer rows work in global memory.
constexpr int EDGE_THREADS = 64, EDGE_CAP = 48;

__global__ __launch_bounds__(EDGE_THREADS) void k_edge_first(
    const int4 *__restrict__ tv4, const int32_t *__restrict__ vt_off,
    const int32_t *__restrict__ vtc, const int32_t *__restrict__ vv_off,
    const int32_t *__restrict__ vv, int64_t V, int32_t *__restrict__ efirst,
    uint8_t *__restrict__ flags, int32_t *__restrict__ te_idx) {
    __shared__ int32_t srow[EDGE_CAP + 3][EDGE_THREADS];  // three spare slots for stage_quads
    __shared__ int32_t sef[EDGE_CAP][EDGE_THREADS];
    const int tid = threadIdx.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + tid; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t rb = vv_off[v], rn = vv_off[v + 1] - rb;
        const int32_t *row = vv + rb;
        int32_t *ef = efirst + rb;
        const int32_t kb = vt_off[v], ke = vt_off[v + 1];
        if (rn <= EDGE_CAP) {
            stage_quads(srow, 0, row, rn, 0);
            for (int32_t i = 0; i < rn; i++) sef[i][tid] = INF;
            for (int32_t k0 = kb; k0 < ke; k0 += 8) {
                int32_t pc[8];
                int4 q[8];
#pragma unroll
                for (int c = 0; c < 8; c++) pc[c] = k0 + c < ke ? vtc[k0 + c] : -1;
#pragma unroll
                for (int c = 0; c < 8; c++) q[c] = pc[c] >= 0 ? tv4[pc[c] >> 2] : make_int4(0, 0, 0, 0);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if (pc[c] < 0) continue;
                    const int32_t t = pc[c] >> 2, jv = pc[c] & 3;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        if (j == jv) continue;
                        const int32_t u = corner(q[c], j);
                        if (u < v) continue;
                        int32_t lo = 0;  // lower bound of u in the staged row (rn < 64)
#pragma unroll
                        for (int step = 32; step >= 1; step >>= 1)
                            if (lo + step - 1 < rn && srow[lo + step - 1][tid] < u) lo += step;
                        const int32_t idx = lo < rn && srow[lo][tid] == u ? lo : -1;
                        const int32_t pos = 6 * t + pair_index(min(j, (int)jv), max(j, (int)jv));
                        te_idx[pos] = idx >= 0 ? rb + idx : -1;
                        if (idx >= 0 && sef[idx][tid] == INF) sef[idx][tid] = pos;
                    }
                }
            }
            QuadOutExact o;
            o.row = ef;
            for (int32_t i = 0; i < rn; i++) {
                const int32_t e = sef[i][tid];
                o.put(e);
                if (srow[i][tid] > v && e != INF) flags[e] = 1;
            }
            o.flush();
            continue;
        }
        for (int32_t k = kb; k < ke; k++) {
            const int32_t p = vtc[k];
            const int32_t t = p >> 2, jv = p & 3;
            const int4 q = tv4[t];
            for (int j = 0; j < 4; j++) {
                if (j == jv) continue;
                const int32_t u = corner(q, j);
                if (u < v) continue;
                const int32_t idx = bsearch_i32(row, rn, u);
                const int32_t pos = 6 * t + pair_index(min(j, jv), max(j, jv));
                te_idx[pos] = idx >= 0 ? rb + idx : -1;
                if (idx >= 0 && ef[idx] == INF) ef[idx] = pos;
            }
        }
        for (int32_t i = 0; i < rn; i++)
            if (row[i] > v && ef[i] != INF) flags[ef[i]] = 1;
    }
}

// bytes -> bitmap words and their popcounts
__global__ void k_pack_flags(const uint8_t *__restrict__ flags, int64_t n,
                             uint32_t *__restrict__ bits, int32_t *__restrict__ pop) {
    const int64_t nw = (n + 31) / 32;
    for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
         w += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = 0;
        const int64_t b = 32 * w;
        for (int i = 0; i < 32 && b + i < n; i++) m |= (uint32_t)(flags[b + i] != 0) << i;
        bits[w] = m;
        pop[w] = __popc(m);
    }
}

__device__ __forceinline__ int32_t rank_of(const uint32_t *bits, const int32_t *wpre, int32_t p) {
    const int32_t w = p >> 5;
    return wpre[w] + __popc(bits[w] & ((1u << (p & 31)) - 1u));
}

// Lane per vertex v: ids of v's upper edges (VV-aligned) and EV.
__global__ void k_edge_ids(const int32_t *__restrict__ vv_off, const int32_t *__restrict__ vv,
                           int64_t V, const int32_t *__restrict__ efirst,
                           const uint32_t *__restrict__ bits, const int32_t *__restrict__ wpre,
                           int32_t *__restrict__ vv_eid, int32_t *__restrict__ ev) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t rb = vv_off[v], re = vv_off[v + 1];
        // eight entries at a time: unconditional loads from clamped indices
        // (the rank words of an unused lane are those of the row's first
        // entry), then the writes of the upper entries
        for (int32_t i0 = rb; i0 < re; i0 += 8) {
            int32_t u[8], p[8], id[8];
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int32_t i = min(i0 + c, re - 1);
                u[c] = vv[i];
                p[c] = efirst[i];
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int on = (int)(i0 + c < re) & (int)(u[c] > v) & (int)(p[c] != INF);
                const int32_t q = on ? p[c] : 0;
                const uint32_t w = bits[q >> 5];
                const int32_t pre = wpre[q >> 5];
                id[c] = on ? pre + __popc(w & ((1u << (q & 31)) - 1u)) : -1;
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                if (id[c] < 0) continue;
                vv_eid[i0 + c] = id[c];
                ev[2 * (int64_t)id[c]] = (int32_t)v;
                ev[2 * (int64_t)id[c] + 1] = u[c];
            }
        }
    }
}

// VE rows {hi, id}: the upper entries (u > v) of v's VV row, ordered by edge
// id.  An edge's id ranks its first position, and the reference mints the
// ids of v's row in the order v's cells first meet its upper neighbours, so
// ordering by id is the reference's row order.  Rows are sorted in an LDS
// column (insertion sort) when they fit, else in place in global memory.
constexpr int VE_THREADS = 64, VE_CAP = 64;

__global__ void k_ve_count(const int32_t *__restrict__ vv_off, const int32_t *__restrict__ vv,
                           int64_t V, int32_t *__restrict__ cnt) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t rb = vv_off[v], rn = vv_off[v + 1] - rb;
        // entries > v: the row is ascending and never holds v
        int32_t lo = 0, hi = rn;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (vv[rb + mid] < (int32_t)v) lo = mid + 1;
            else hi = mid;
        }
        cnt[v] = rn - lo;
    }
}

__global__ __launch_bounds__(VE_THREADS) void k_ve_fill(const int32_t *__restrict__ vv_off,
                                                        const int32_t *__restrict__ vv,
                                                        const int32_t *__restrict__ vv_eid,
                                                        int64_t V, const int32_t *__restrict__ off,
                                                        int32_t *__restrict__ out) {
    __shared__ int64_t col[VE_CAP][VE_THREADS];  // (id << 32) | hi
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t re = vv_off[v + 1];
        const int32_t n = off[v + 1] - off[v];
        int32_t *row = out + 2 * (int64_t)off[v];
        const int32_t ub = re - n;  // first upper entry
        if (n <= VE_CAP) {
            for (int32_t k = 0; k < n; k++) {
                const int64_t key = ((int64_t)vv_eid[ub + k] << 32) | (uint32_t)vv[ub + k];
                int32_t p = k;
                while (p > 0 && col[p - 1][threadIdx.x] > key) {
                    col[p][threadIdx.x] = col[p - 1][threadIdx.x];
                    p--;
                }
                col[p][threadIdx.x] = key;
            }
            for (int32_t k = 0; k < n; k++) {  // (hi, id) pairs, one 8-byte store each
                const int64_t key = col[k][threadIdx.x];
                reinterpret_cast<int2 *>(row)[k] = make_int2((int32_t)(uint32_t)key, (int32_t)(key >> 32));
            }
        } else {
            for (int32_t k = 0; k < n; k++) {
                const int32_t id = vv_eid[ub + k], u = vv[ub + k];
                int32_t p = k;
                while (p > 0 && row[2 * (p - 1) + 1] > id) {
                    row[2 * p] = row[2 * (p - 1)];
                    row[2 * p + 1] = row[2 * (p - 1) + 1];
                    p--;
                }
                row[2 * p] = u;
                row[2 * p + 1] = id;
            }
        }
    }
}

// lower entries (u < v) take the id stored in u's row
__global__ void k_edge_lower(const int32_t *__restrict__ vv_off, const int32_t *__restrict__ vv,
                             int64_t V, int32_t *__restrict__ vv_eid) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t i = vv_off[v]; i < vv_off[v + 1]; i++) {
            const int32_t u = vv[i];
            if (u > v) continue;
            const int32_t ub = vv_off[u];
            const int32_t j = bsearch_i32(vv + ub, vv_off[u + 1] - ub, (int32_t)v);
            vv_eid[i] = j >= 0 ? vv_eid[ub + j] : -1;
        }
    }
}

// TE: the VV indices recorded by k_edge_first become edge ids
__global__ void k_te(int32_t *__restrict__ te, int64_t n, const int32_t *__restrict__ vv_eid) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = te[i];
        te[i] = k >= 0 ? vv_eid[k] : -1;
    }
}

// ---- ET = VT(lo) n VT(hi) -------------------------------------------------------
// Both rows are staged in LDS when they fit (stage_quads), then merged in
// one pass into a padded slot at pad[e] of min(|VT(lo)|, |VT(hi)|) words
// rounded up to a multiple of four (QuadOut), with the row length in
// cnt[e]; k_pack_rows packs the rows afterwards.
constexpr int ET_THREADS = 64, ET_CAP = 125;  // with stage_quads' three spare slots: 32 KB

__global__ void k_et_bound(const int32_t *__restrict__ ev, int64_t E, int64_t V,
                           const int32_t *__restrict__ vt_off, int64_t *__restrict__ ub) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = ev[2 * e], b = ev[2 * e + 1];
        const bool ok = a >= 0 && a < V && b >= 0 && b < V;  // always, for a built EV
        const int64_t n = ok ? min(vt_off[a + 1] - vt_off[a], vt_off[b + 1] - vt_off[b]) : 0;
        ub[e] = (n + 3) & ~(int64_t)3;
    }
}

__global__ __launch_bounds__(ET_THREADS) void k_et(const int32_t *__restrict__ ev, int64_t E,
                                                   int64_t V, const int32_t *__restrict__ vt_off,
                                                   const int32_t *__restrict__ vtc,
                                                   const int64_t *__restrict__ pad,
                                                   int64_t *__restrict__ cnt,
                                                   int32_t *__restrict__ out) {
    __shared__ int32_t buf[ET_CAP + 3][ET_THREADS];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = ev[2 * e], b = ev[2 * e + 1];
        const bool ok = a >= 0 && a < V && b >= 0 && b < V;  // always, for a built EV
        int32_t i = ok ? vt_off[a] : 0, ie = ok ? vt_off[a + 1] : 0;
        int32_t j = ok ? vt_off[b] : 0, je = ok ? vt_off[b + 1] : 0;
        QuadOut o;
        o.row = out + pad[e];
        const int32_t na = ie - i, nb = je - j;
        if (na + nb <= ET_CAP) {
            stage_quads(buf, 0, vtc + i, na, 2);
            stage_quads(buf, na, vtc + j, nb, 2);
            i = 0, ie = na, j = na, je = na + nb;
            int32_t x = i < ie ? buf[i][threadIdx.x] : INF, y = j < je ? buf[j][threadIdx.x] : INF;
            const int32_t last = max(je - 1, 0);
            while (x != INF && y != INF) {  // branch-free advance (two reads per step)
                if (x == y) o.put(x);
                const int32_t m = min(x, y);
                i += x == m;
                j += y == m;
                const int32_t xa = buf[min(i, last)][threadIdx.x], yb = buf[min(j, last)][threadIdx.x];
                x = i < ie ? xa : INF;
                y = j < je ? yb : INF;
            }
        } else {
            int32_t x = i < ie ? vtc[i] >> 2 : INF, y = j < je ? vtc[j] >> 2 : INF;
            while (x != INF && y != INF) {
                if (x == y) o.put(x);
                const int32_t m = min(x, y);
                if (x == m) x = ++i < ie ? vtc[i] >> 2 : INF;
                if (y == m) y = ++j < je ? vtc[j] >> 2 : INF;
            }
        }
        o.flush();
        cnt[e] = o.n;
    }
}

// ---- faces ----------------------------------------------------------------------
// tets holding face slot s of tet t (ascending) via ET(lo,mid) n ET(lo,hi);
// returns the first matching position (smallest 4t'+s' with the same
// positional triple) and counts all matching tets.
// Ids read from TE and ET are checked against E and T: an inconsistent
// relation yields INF (reported by the callers) instead of a stray access.
constexpr int FF_ROW = 8;  // ET rows up to this length take the register path

__device__ __forceinline__ int32_t face_first(const int4 *tv4, int64_t T, const int32_t *te,
                                              int64_t E, const int64_t *et_off,
                                              const int32_t *et, int32_t t, int s,
                                              int32_t &nmatch, int32_t *tets2) {
    const int4 q = tv4[t];
    int lo, mid, hi;
    face_pos(s, lo, mid, hi);
    const int32_t a = corner(q, lo), b = corner(q, mid), c = corner(q, hi);
    const int32_t e1 = te[6 * (int64_t)t + pair_index(lo, mid)];
    const int32_t e2 = te[6 * (int64_t)t + pair_index(lo, hi)];
    int32_t best = INF;
    nmatch = 0;
    if (e1 < 0 || e1 >= E || e2 < 0 || e2 >= E) return best;
    int64_t i = et_off[e1], ie = et_off[e1 + 1], j = et_off[e2], je = et_off[e2 + 1];
    if (ie - i <= FF_ROW && je - j <= FF_ROW) {
        // short rows: both are loaded at once and intersected in registers,
        // and the cells they share are gathered together, so the slot costs
        // two dependent loads instead of one per row element
        const int na = (int)(ie - i), nb = (int)(je - j);
        int32_t ra[FF_ROW], rb[FF_ROW];
#pragma unroll
        for (int k = 0; k < FF_ROW; k++) {
            ra[k] = k < na ? et[i + k] : 0;
            rb[k] = k < nb ? et[j + k] : 0;
        }
        bool hit[FF_ROW];
        int4 r[FF_ROW];
#pragma unroll
        for (int k = 0; k < FF_ROW; k++) {
            bool h = false;
#pragma unroll
            for (int l = 0; l < FF_ROW; l++) h = h || (l < nb && rb[l] == ra[k]);
            hit[k] = k < na && h;
            if (hit[k] && (ra[k] < 0 || ra[k] >= T)) return INF;
        }
#pragma unroll
        for (int k = 0; k < FF_ROW; k++) r[k] = hit[k] ? tv4[ra[k]] : make_int4(-1, -1, -1, -1);
#pragma unroll
        for (int k = 0; k < FF_ROW; k++) {
            if (!hit[k]) continue;
            for (int s2 = 0; s2 < 4; s2++) {
                int l2, m2, h2;
                face_pos(s2, l2, m2, h2);
                if (corner(r[k], l2) == a && corner(r[k], m2) == b && corner(r[k], h2) == c) {
                    if (tets2 && nmatch == 0) tets2[0] = ra[k];
                    if (tets2 && nmatch == 1) tets2[1] = ra[k];
                    nmatch++;
                    best = min(best, 4 * ra[k] + s2);
                    break;
                }
            }
        }
        return best;
    }
    while (i < ie && j < je) {
        const int32_t x = et[i], y = et[j];
        if (x < y) i++;
        else if (y < x) j++;
        else if (x < 0 || x >= T) return INF;
        else {
            const int4 r = tv4[x];
            for (int s2 = 0; s2 < 4; s2++) {
                int l2, m2, h2;
                face_pos(s2, l2, m2, h2);
                if (corner(r, l2) == a && corner(r, m2) == b && corner(r, h2) == c) {
                    if (nmatch < 2 && tets2) tets2[nmatch] = x;
                    nmatch++;
                    best = min(best, 4 * x + s2);
                    break;
                }
            }
            i++;
            j++;
        }
    }
    return best;
}

// Per slot position p = 4t + s: the first position of its face, and the
// other cell holding it (-1 if none), so FT needs no second match.  A face
// held by more than two cells sets fterr.
__global__ void k_face_first(const int4 *__restrict__ tv4, int64_t T, const int32_t *__restrict__ te,
                             int64_t E, const int64_t *__restrict__ et_off, const int32_t *__restrict__ et,
                             int32_t *__restrict__ ffirst, uint8_t *__restrict__ flags,
                             int32_t *__restrict__ other, int32_t *__restrict__ err,
                             int32_t *__restrict__ fterr) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < 4 * T;
         p += (int64_t)gridDim.x * blockDim.x) {
        int32_t nm, two[2] = {-1, -1};
        const int32_t t = (int32_t)(p >> 2);
        const int32_t f = face_first(tv4, T, te, E, et_off, et, t, (int)(p & 3), nm, two);
        // the cell itself always matches; no match means ET and TE disagree
        const bool bad = f > (int32_t)p;
        if (bad) *err = 1;
        if (nm > 2) *fterr = 1;  // every writer stores the same value
        ffirst[p] = bad ? (int32_t)p : f;
        flags[p] = bad || f == (int32_t)p;
        other[p] = nm >= 2 ? (two[0] == t ? two[1] : two[0]) : -1;
    }
}

// TF; FV as (mid, hi, lo) -- the fields (middleVert, highVert, low vertex);
// FT (<= 2 tets, ascending, -1 pad)
__global__ void k_face_ids(const int4 *__restrict__ tv4, int64_t T,
                           const int32_t *__restrict__ ffirst, const int32_t *__restrict__ other,
                           const uint32_t *__restrict__ bits,
                           const int32_t *__restrict__ wpre, int32_t *__restrict__ tf,
                           int32_t *__restrict__ fv, int32_t *__restrict__ ft) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < 4 * T;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t f = ffirst[p];
        const int32_t id = rank_of(bits, wpre, f);
        tf[p] = id;
        if (f != (int32_t)p) continue;
        const int32_t t = (int32_t)(p >> 2);
        const int s = (int)(p & 3);
        const int4 q = tv4[t];
        int lo, mid, hi;
        face_pos(s, lo, mid, hi);
        fv[3 * (int64_t)id] = corner(q, mid);
        fv[3 * (int64_t)id + 1] = corner(q, hi);
        fv[3 * (int64_t)id + 2] = corner(q, lo);
        // the first position's cell is the smallest holding the face
        ft[2 * (int64_t)id] = t;
        ft[2 * (int64_t)id + 1] = other[p];
    }
}

// VF rows {mid, hi, id}: the row of a holds the faces whose positional low
// vertex is a.  Lane per vertex walks its tets ascending and takes the slots
// where a sits at the low position and which are their face's first
// occurrence; positions ascend, so ids ascend.
template <bool FILL>
__global__ void k_vf(const int4 *__restrict__ tv4, const int32_t *__restrict__ vt_off,
                     const int32_t *__restrict__ vtc, int64_t V, const int32_t *__restrict__ ffirst,
                     const int32_t *__restrict__ tf, int32_t *__restrict__ cnt_or_off,
                     int32_t *__restrict__ out) {
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < V;
         a += (int64_t)gridDim.x * blockDim.x) {
        int32_t *row = FILL ? out + 3 * (int64_t)cnt_or_off[a] : nullptr;
        int32_t n = 0;
        const int32_t kb = vt_off[a], ke = vt_off[a + 1];
        // eight cells at a time: their slot flags, ids and corners are
        // loaded together, then the row is written in cell order
        for (int32_t k0 = kb; k0 < ke; k0 += 8) {
            int32_t pc[8], ff[8][3], id[8][3];
            int4 q[8];
#pragma unroll
            for (int c = 0; c < 8; c++) pc[c] = k0 + c < ke ? vtc[k0 + c] : -1;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int32_t t = pc[c] >> 2, j = pc[c] & 3;
                // position 0 is low in slots 0-2, position 1 in slot 3
                const bool on = pc[c] >= 0 && j <= 1;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const bool use = on && (j == 0 || m == 0);
                    const int64_t pos = 4 * (int64_t)t + (j == 0 ? m : 3);
                    ff[c][m] = use ? ffirst[pos] : -1;
                    id[c][m] = use && FILL ? tf[pos] : 0;
                }
                q[c] = on && FILL ? tv4[t] : make_int4(0, 0, 0, 0);
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int32_t t = pc[c] >> 2, j = pc[c] & 3;
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const int sl = j == 0 ? m : 3;
                    if (ff[c][m] != 4 * t + sl || ff[c][m] < 0) continue;
                    if (FILL) {
                        int lo, mid, hi;
                        face_pos(sl, lo, mid, hi);
                        // one 12-byte store per entry
                        const int32_t ent[3] = {corner(q[c], mid), corner(q[c], hi), id[c][m]};
                        __builtin_memcpy(row + 3 * n, ent, sizeof ent);
                    }
                    n++;
                }
            }
        }
        if (!FILL) cnt_or_off[a] = n;
    }
}

// FE[f] = {max vertex, edge id} for the edges (lo,mid), (lo,hi), (mid,hi),
// read from TE of the face's first cell (FT[f][0]) at the slot whose
// positional triple is the face's.
__global__ void k_fe(const int4 *__restrict__ tv4, const int32_t *__restrict__ fv, int64_t F,
                     const int32_t *__restrict__ ft, const int32_t *__restrict__ te,
                     int32_t *__restrict__ fe) {
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < F;
         f += (int64_t)gridDim.x * blockDim.x) {
        const int32_t mid = fv[3 * f], hi = fv[3 * f + 1], lo = fv[3 * f + 2];
        const int32_t t = ft[2 * f];
        if (t < 0) {  // not reached for a built FT: every face has a first cell
            for (int k = 0; k < 6; k++) fe[6 * f + k] = -1;
            continue;
        }
        const int4 q = tv4[t];
        int s = 0;
        for (int k = 0; k < 4; k++) {
            int l, m, h;
            face_pos(k, l, m, h);
            if (corner(q, l) == lo && corner(q, m) == mid && corner(q, h) == hi) {
                s = k;
                break;
            }
        }
        int pl, pm, ph;
        face_pos(s, pl, pm, ph);
        const int32_t *tr = te + 6 * (int64_t)t;
        fe[6 * f] = max(lo, mid);
        fe[6 * f + 1] = tr[pair_index(pl, pm)];
        fe[6 * f + 2] = max(lo, hi);
        fe[6 * f + 3] = tr[pair_index(pl, ph)];
        fe[6 * f + 4] = max(mid, hi);
        fe[6 * f + 5] = tr[pair_index(pm, ph)];
    }
}

// insert x into the ascending row[0..m) unless present
__device__ __forceinline__ void insert_sorted(int32_t *row, int64_t &m, int32_t x) {
    int64_t i = m;
    while (i > 0 && row[i - 1] > x) i--;
    if (i > 0 && row[i - 1] == x) return;
    for (int64_t k = m; k > i; k--) row[k] = row[k - 1];
    row[i] = x;
    m++;
}

// EF(e): faces holding e, from the two faces of e in each cell of ET(e).
// One pass builds each sorted, duplicate-free row in a padded slot of
// 2|ET(e)| words (tmp + 2 et_off[e]; in an LDS column while it fits) and
// counts it; k_ef_copy then packs the rows after the scan of the counts.
constexpr int EF_CAP = 64;

__global__ __launch_bounds__(ET_THREADS) void k_ef(const int64_t *__restrict__ et_off,
                                                   int64_t E, const int32_t *__restrict__ et,
                                                   const int32_t *__restrict__ te,
                                                   const int32_t *__restrict__ tf,
                                                   int64_t *__restrict__ cnt,
                                                   int32_t *__restrict__ tmp) {
    __shared__ int32_t col[EF_CAP][ET_THREADS];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b0 = et_off[e], b1 = et_off[e + 1];
        const bool lds = 2 * (b1 - b0) <= EF_CAP;
        int32_t *grow = tmp + 2 * b0;
        int64_t m = 0;
        // four cells at a time: their TE and TF rows are loaded together
        // (unconditional loads from a clamped cell index)
        for (int64_t i0 = b0; i0 < b1; i0 += 4) {
            int32_t tc[4];
            int2 r01[4], r23[4], r45[4];
            int4 fr[4];
#pragma unroll
            for (int c = 0; c < 4; c++) tc[c] = et[min(i0 + c, b1 - 1)];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int2 *tr = reinterpret_cast<const int2 *>(te + 6 * (int64_t)tc[c]);
                r01[c] = tr[0];
                r23[c] = tr[1];
                r45[c] = tr[2];
                fr[c] = reinterpret_cast<const int4 *>(tf)[tc[c]];
            }
#pragma unroll
            for (int c = 0; c < 4; c++) {
                if (i0 + c >= b1) break;
                const int32_t ee = (int32_t)e;
                const int k = r01[c].x == ee ? 0 : r01[c].y == ee ? 1 : r23[c].x == ee ? 2
                            : r23[c].y == ee ? 3 : r45[c].x == ee ? 4 : r45[c].y == ee ? 5 : 6;
                if (k == 6) continue;  // TE and ET disagree: no face of this cell
                // corner pair of edge slot k; faces holding both corners
                const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {1, 2, 3, 2, 3, 3};
                for (int s = 0; s < 4; s++) {
                    int lo, mid, hi;
                    face_pos(s, lo, mid, hi);
                    const bool has_i = lo == pi[k] || mid == pi[k] || hi == pi[k];
                    const bool has_j = lo == pj[k] || mid == pj[k] || hi == pj[k];
                    if (!(has_i && has_j)) continue;
                    const int32_t x = s == 0 ? fr[c].x : s == 1 ? fr[c].y : s == 2 ? fr[c].z : fr[c].w;
                    if (lds) {  // insert into the ascending column unless present
                        int64_t p = m;
                        while (p > 0 && col[p - 1][threadIdx.x] > x) p--;
                        if (p > 0 && col[p - 1][threadIdx.x] == x) continue;
                        for (int64_t q = m; q > p; q--) col[q][threadIdx.x] = col[q - 1][threadIdx.x];
                        col[p][threadIdx.x] = x;
                        m++;
                    } else {
                        insert_sorted(grow, m, x);
                    }
                }
            }
        }
        if (lds) {
            QuadOutExact o;
            o.row = grow;
            for (int64_t q = 0; q < m; q++) o.put(col[q][threadIdx.x]);
            o.flush();
        }
        cnt[e] = m;
    }
}

// packed rows from the padded ones (off: exclusive scan of the counts)
__global__ void k_ef_copy(const int64_t *__restrict__ et_off, int64_t E,
                          const int64_t *__restrict__ off, const int32_t *__restrict__ tmp,
                          int32_t *__restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        // 16-byte loads (the padded buffer carries four words of slack) and
        // stores, the last one to three words one by one
        const int32_t *src = tmp + 2 * et_off[e];
        const int64_t o = off[e], n = off[e + 1] - o;
        QuadOutExact w;
        w.row = out + o;
        for (int64_t q = 0; q < n; q += 4) {
            const i32x4u v = *reinterpret_cast<const i32x4u *>(src + q);
            w.put(v.x);
            if (q + 1 < n) w.put(v.y);
            if (q + 2 < n) w.put(v.z);
            if (q + 3 < n) w.put(v.w);
        }
        w.flush();
    }
}

// TT and EE merge columns (stage_quads): ROW_CAP words per thread.
constexpr int ROW_THREADS = 64;
constexpr int ROW_CAP = 100;  // 26 KB per 64 threads (with three spare slots): six waves per CU

// Cells in the order of their corner-0 entries in VT (rows by vertex, cells
// ascending within a row), for k_tt on meshes with rows beyond its LDS
// column: neighbouring lanes then take cells around one vertex and share
// its VT row, where the mesh's own cell order may scatter them (on a mesh
// of uniform short rows the pass costs more than the sharing saves).  One pass counts the corner-0 codes per chunk, a scan of
// the chunk counts gives each chunk's start, a second pass fills the order
// stably (block-wide ballot scans, one per round of 256 codes).
constexpr int C0_THREADS = 256, C0_PER = 16, C0_CHUNK = C0_THREADS * C0_PER;

__global__ __launch_bounds__(C0_THREADS) void k_c0_count(const int32_t *__restrict__ vtc,
                                                         int64_t n, int64_t *__restrict__ bcnt) {
    __shared__ int32_t total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const int64_t b0 = (int64_t)blockIdx.x * C0_CHUNK;
    int32_t c = 0;
    for (int i = 0; i < C0_PER; i++) {
        const int64_t k = b0 + (int64_t)i * C0_THREADS + threadIdx.x;
        c += k < n && (vtc[k] & 3) == 0;
    }
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) atomicAdd(&total, c);
    __syncthreads();
    if (threadIdx.x == 0) bcnt[blockIdx.x] = total;
}

__global__ __launch_bounds__(C0_THREADS) void k_c0_fill(const int32_t *__restrict__ vtc, int64_t n,
                                                        const int64_t *__restrict__ bbase,
                                                        int32_t *__restrict__ order) {
    __shared__ int32_t wsum[C0_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t b0 = (int64_t)blockIdx.x * C0_CHUNK;
    int64_t base = bbase[blockIdx.x];
    for (int i = 0; i < C0_PER; i++) {
        const int64_t k = b0 + (int64_t)i * C0_THREADS + threadIdx.x;
        const int32_t code = k < n ? vtc[k] : 1;
        const bool f = (code & 3) == 0;
        const unsigned long long m = __ballot(f);
        const int before = __popcll(m & ((1ull << lane) - 1));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int32_t off = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < C0_THREADS / 64; j++) {
            const int32_t x = wsum[j];
            off += j < w ? x : 0;
            tot += x;
        }
        if (f) order[base + off + before] = code >> 2;
        base += tot;
        __syncthreads();  // wsum is rewritten next round
    }
}

// TT(t): union of VT rows of t's corners, minus t (4-way merge), written in
// one pass to a padded slot at pad[t] of |VT rows| - 4 words rounded up to a
// multiple of four, with its length in cnt[t]; the rows are packed after
// the scan of cnt.
__global__ void k_tt_bound(const int4 *__restrict__ tv4, int64_t T,
                           const int32_t *__restrict__ vt_off, int64_t *__restrict__ ub,
                           int32_t *__restrict__ long_rows) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv4[t];
        const i32x2u r0 = *reinterpret_cast<const i32x2u *>(vt_off + q.x);
        const i32x2u r1 = *reinterpret_cast<const i32x2u *>(vt_off + q.y);
        const i32x2u r2 = *reinterpret_cast<const i32x2u *>(vt_off + q.z);
        const i32x2u r3 = *reinterpret_cast<const i32x2u *>(vt_off + q.w);
        const int64_t n = (int64_t)(r0.y - r0.x) + (r1.y - r1.x) + (r2.y - r2.x) + (r3.y - r3.x) - 4;
        ub[t] = (n + 3) & ~(int64_t)3;
        if (n + 4 > ROW_CAP) *long_rows = 1;  // every writer stores the same value
    }
}

__global__ __launch_bounds__(ROW_THREADS) void k_tt(const int4 *__restrict__ tv4, int64_t T,
                                                    const int32_t *__restrict__ vt_off,
                                                    const int32_t *__restrict__ vtc,
                                                    const int32_t *__restrict__ order,
                                                    const int64_t *__restrict__ pad,
                                                    int64_t *__restrict__ cnt,
                                                    int32_t *__restrict__ out) {
    __shared__ int32_t buf[ROW_CAP + 3][ROW_THREADS];  // three spare slots for stage_quads
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < T;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = order ? order[i] : i;  // cells in VT order when given
        const int4 q = tv4[t];
        const i32x2u r0 = *reinterpret_cast<const i32x2u *>(vt_off + q.x);
        const i32x2u r1 = *reinterpret_cast<const i32x2u *>(vt_off + q.y);
        const i32x2u r2 = *reinterpret_cast<const i32x2u *>(vt_off + q.z);
        const i32x2u r3 = *reinterpret_cast<const i32x2u *>(vt_off + q.w);
        int32_t p0 = r0.x, e0 = r0.y, p1 = r1.x, e1 = r1.y;
        int32_t p2 = r2.x, e2 = r2.y, p3 = r3.x, e3 = r3.y;
        QuadOut o;
        o.row = out + pad[t];
        const int32_t n0 = e0 - p0, n1 = e1 - p1, n2 = e2 - p2, n3 = e3 - p3;
        if (n0 + n1 + n2 + n3 <= ROW_CAP) {
            // heads as LDS positions
            int32_t a0 = 0, a1 = n0, a2 = n0 + n1, a3 = n0 + n1 + n2;
            const int32_t z0 = n0, z1 = n0 + n1, z2 = n0 + n1 + n2, z3 = n0 + n1 + n2 + n3;
            stage_quads(buf, 0, vtc + p0, n0, 2);
            stage_quads(buf, z0, vtc + p1, n1, 2);
            stage_quads(buf, z1, vtc + p2, n2, 2);
            stage_quads(buf, z2, vtc + p3, n3, 2);
            int32_t h0 = a0 < z0 ? buf[a0][threadIdx.x] : INF;
            int32_t h1 = a1 < z1 ? buf[a1][threadIdx.x] : INF;
            int32_t h2 = a2 < z2 ? buf[a2][threadIdx.x] : INF;
            int32_t h3 = a3 < z3 ? buf[a3][threadIdx.x] : INF;
            // branch-free advance: every head re-reads its (clamped) position,
            // so the four LDS reads of an iteration issue together
            const int32_t last = max(z3 - 1, 0);
            while (true) {
                const int32_t mn = min(min(h0, h1), min(h2, h3));
                if (mn == INF) break;
                a0 += h0 == mn;
                a1 += h1 == mn;
                a2 += h2 == mn;
                a3 += h3 == mn;
                const int32_t y0 = buf[min(a0, last)][threadIdx.x];
                const int32_t y1 = buf[min(a1, last)][threadIdx.x];
                const int32_t y2 = buf[min(a2, last)][threadIdx.x];
                const int32_t y3 = buf[min(a3, last)][threadIdx.x];
                h0 = a0 < z0 ? y0 : INF;
                h1 = a1 < z1 ? y1 : INF;
                h2 = a2 < z2 ? y2 : INF;
                h3 = a3 < z3 ? y3 : INF;
                if (mn == (int32_t)t) continue;
                o.put(mn);
            }
        } else {
            int32_t h0 = p0 < e0 ? vtc[p0] >> 2 : INF, h1 = p1 < e1 ? vtc[p1] >> 2 : INF;
            int32_t h2 = p2 < e2 ? vtc[p2] >> 2 : INF, h3 = p3 < e3 ? vtc[p3] >> 2 : INF;
            while (true) {
                const int32_t mn = min(min(h0, h1), min(h2, h3));
                if (mn == INF) break;
                if (h0 == mn) h0 = ++p0 < e0 ? vtc[p0] >> 2 : INF;
                if (h1 == mn) h1 = ++p1 < e1 ? vtc[p1] >> 2 : INF;
                if (h2 == mn) h2 = ++p2 < e2 ? vtc[p2] >> 2 : INF;
                if (h3 == mn) h3 = ++p3 < e3 ? vtc[p3] >> 2 : INF;
                if (mn == (int32_t)t) continue;
                o.put(mn);
            }
        }
        o.flush();
        cnt[t] = o.n;
    }
}

// Pack padded rows: row r of length off[r+1] - off[r] moves from src +
// pad[r] to dst + off[r].  A block takes 256 consecutive rows and copies its
// contiguous output range element by element (a row found by bisection of
// the block's row starts in LDS), so reads and writes are coalesced.
__global__ __launch_bounds__(256) void k_pack_rows(const int64_t *__restrict__ pad,
                                                   const int64_t *__restrict__ off, int64_t rows,
                                                   const int32_t *__restrict__ src,
                                                   int32_t *__restrict__ dst) {
    __shared__ int64_t s_off[257], s_pad[256];
    const int64_t r0 = (int64_t)blockIdx.x * 256;
    const int nr = (int)min<int64_t>(256, rows - r0);
    for (int i = threadIdx.x; i <= nr; i += 256) {
        s_off[i] = off[r0 + i];
        if (i < nr) s_pad[i] = pad[r0 + i];
    }
    __syncthreads();
    const int64_t ob = s_off[0], oe = s_off[nr];
    for (int64_t o = ob + threadIdx.x; o < oe; o += 256) {
        int lo = 0, hi = nr - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_off[mid] <= o) lo = mid;
            else hi = mid - 1;
        }
        dst[o] = src[s_pad[lo] + (o - s_off[lo])];
    }
}

// Pack of long padded rows with a lane per output word: one wave per 64
// consecutive rows copies their contiguous output range 64 words at a time
// (full, aligned stores; reads nearly contiguous, the slots being
// consecutive), each word finding its row by a six-step search of the
// rows' output starts held in LDS.  Four chunks per lane are in flight.
__global__ __launch_bounds__(256) void k_pack_rows_wave(const int64_t *__restrict__ pad,
                                                        const int64_t *__restrict__ off,
                                                        int64_t rows,
                                                        const int32_t *__restrict__ src,
                                                        int32_t *__restrict__ dst) {
    __shared__ int32_t s_start[4][65];  // output start of each row, relative to the wave's
    __shared__ int64_t s_pad[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + w) * 64;
    const int nr = r0 < rows ? (int)min<int64_t>(64, rows - r0) : 0;
    const int64_t o0 = nr > 0 ? off[r0] : 0;
    // rows past the end start at INT32_MAX and are never chosen
    s_start[w][lane] = lane < nr ? (int32_t)(off[r0 + lane] - o0) : INT32_MAX;
    if (lane == 0) s_start[w][64] = nr > 0 ? (int32_t)(off[r0 + nr] - o0) : 0;
    s_pad[w][lane] = lane < nr ? pad[r0 + lane] : 0;
    __syncthreads();
    const int32_t total = s_start[w][64];
    for (int32_t k0 = 0; k0 < total; k0 += 4 * 64) {
        int32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int32_t k = min(k0 + 64 * u + lane, total - 1);
            int lo = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1)
                if (s_start[w][lo + step] <= k) lo += step;
            v[u] = src[s_pad[w][lo] + (k - s_start[w][lo])];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k0 + 64 * u + lane < total) dst[o0 + k0 + 64 * u + lane] = v[u];
    }
}

// FF(f): faces of the cells holding f, minus f
template <bool FILL>
__global__ void k_ff(const int32_t *__restrict__ ft, int64_t F, const int32_t *__restrict__ tf,
                     int64_t *__restrict__ cnt_or_off, int32_t *__restrict__ out) {
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < F;
         f += (int64_t)gridDim.x * blockDim.x) {
        // the faces of f's (at most two) cells other than f: both TF rows are
        // loaded at once, the six candidates sorted by a network, then the
        // distinct ones counted or written in order
        const int2 cells = reinterpret_cast<const int2 *>(ft)[f];
        const int4 r0 = reinterpret_cast<const int4 *>(tf)[cells.x >= 0 ? cells.x : 0];
        const int4 r1 = reinterpret_cast<const int4 *>(tf)[cells.y >= 0 ? cells.y : 0];
        const int32_t me = (int32_t)f;
        int32_t c[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const bool on = i < 4 ? cells.x >= 0 : cells.y >= 0;
            if (!on || c[i] == me) c[i] = INT32_MAX;
        }
        // bitonic network over eight registers (ascending)
#pragma unroll
        for (int k = 2; k <= 8; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const int l = i ^ j;
                    if (l > i) {
                        const int32_t a = c[i], b = c[l];
                        const bool up = (i & k) == 0;
                        c[i] = up ? min(a, b) : max(a, b);
                        c[l] = up ? max(a, b) : min(a, b);
                    }
                }
        int64_t o = FILL ? cnt_or_off[f] : 0;
        int n = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const bool keep = c[i] != INT32_MAX && (i == 0 || c[i] != c[i - 1]);
            if (FILL && keep) out[o + n] = c[i];
            n += keep;
        }
        if (!FILL) cnt_or_off[f] = n;
    }
}

// EE(e): edges at lo or hi, minus e, ascending: merge of the two vertices'
// edge-id lists (vv_eid rows sorted by id, in eids).  The lists share only
// e itself, so the row length is deg(lo) + deg(hi) - 2.
__global__ __launch_bounds__(ROW_THREADS) void k_ee(const int32_t *__restrict__ ev, int64_t E,
                                                    const int32_t *__restrict__ vv_off,
                                                    const int32_t *__restrict__ eids,
                                                    const int64_t *__restrict__ off,
                                                    int32_t *__restrict__ out) {
    __shared__ int32_t buf[ROW_CAP + 3][ROW_THREADS];  // three spare slots for stage_quads
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = ev[2 * e], b = ev[2 * e + 1];
        int32_t i = vv_off[a], ie = vv_off[a + 1], j = vv_off[b], je = vv_off[b + 1];
        QuadOutExact o;
        o.row = out + off[e];
        const int32_t na = ie - i, nb = je - j;
        if (na + nb <= ROW_CAP) {
            stage_quads(buf, 0, eids + i, na, 0);
            stage_quads(buf, na, eids + j, nb, 0);
            i = 0, ie = na, j = na, je = na + nb;
            int32_t x = i < ie ? buf[i][threadIdx.x] : INF, y = j < je ? buf[j][threadIdx.x] : INF;
            const int32_t last = max(je - 1, 0);
            while (x != INF || y != INF) {  // branch-free advance (two reads per step)
                const int32_t mn = min(x, y);
                i += x == mn;
                j += y == mn;
                const int32_t xa = buf[min(i, last)][threadIdx.x], yb = buf[min(j, last)][threadIdx.x];
                x = i < ie ? xa : INF;
                y = j < je ? yb : INF;
                if (mn != (int32_t)e) o.put(mn);
            }
        } else {
            int32_t x = i < ie ? eids[i] : INF, y = j < je ? eids[j] : INF;
            while (x != INF || y != INF) {
                const int32_t mn = min(x, y);
                if (x == mn) x = ++i < ie ? eids[i] : INF;
                if (y == mn) y = ++j < je ? eids[j] : INF;
                if (mn != (int32_t)e) o.put(mn);
            }
        }
        o.flush();
    }
}

// row lengths of EE: deg(lo) + deg(hi) - 2
__global__ void k_ee_count(const int32_t *__restrict__ ev, int64_t E,
                           const int32_t *__restrict__ vv_off, int64_t *__restrict__ cnt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = ev[2 * e], b = ev[2 * e + 1];
        cnt[e] = (int64_t)(vv_off[a + 1] - vv_off[a]) + (vv_off[b + 1] - vv_off[b]) - 2;
    }
}

__global__ void k_fill_i32v(int32_t *p, int32_t v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_widen(const int32_t *__restrict__ in, int64_t n, int64_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[i];
}

unsigned grid(int64_t n, int block = 256) {
    return std::min<unsigned>(tg::grid_for(n, block), 65536u);
}

int fail(int code, const char *msg) {
    tg::set_error(msg);
    return code;
}

int read_i32(tg_mesh &m, const void *src, int64_t index, int32_t &out) {
    int32_t *h = (int32_t *)m.h_pinned;
    TG_HIP(hipMemcpyAsync(h, (const int32_t *)src + index, sizeof(int32_t),
                          hipMemcpyDeviceToHost, m.stream));
    TG_HIP(hipStreamSynchronize(m.stream));
    out = *h;
    return tg::OK;
}

int read_i64(tg_mesh &m, const void *src, int64_t index, int64_t &out) {
    int64_t *h = (int64_t *)m.h_pinned;
    TG_HIP(hipMemcpyAsync(h, (const int64_t *)src + index, sizeof(int64_t),
                          hipMemcpyDeviceToHost, m.stream));
    TG_HIP(hipStreamSynchronize(m.stream));
    out = *h;
    return tg::OK;
}

// rank structure over n flags: bits + exclusive prefix of word popcounts
int build_ranks(tg_mesh &m, const uint8_t *flags, int64_t n, int32_t *&bits, int32_t *&wpre,
                int64_t &total) {
    const int64_t nw = (n + 31) / 32;
    int rc;
    if ((rc = m.bitmap.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(nw, 1))) ||
        (rc = m.rank_tmp.ensure(sizeof(int32_t) * (size_t)(nw + 1))) ||
        (rc = m.scan_tmp.ensure(tg::scan_tmp_bytes(nw))))
        return rc;
    bits = m.bitmap.as<int32_t>();
    wpre = m.rank_tmp.as<int32_t>();
    if (nw > 0) {
        hipLaunchKernelGGL(k_pack_flags, dim3(grid(nw)), dim3(256), 0, m.stream, flags, n,
                           (uint32_t *)bits, wpre);
        TG_LAUNCH_CHECK("k_pack_flags");
    }
    if ((rc = tg::scan_exclusive_i32(wpre, wpre, nw, m.scan_tmp.p, m.stream))) return rc;
    int32_t t32 = 0;
    if ((rc = read_i32(m, wpre, nw, t32))) return rc;
    total = t32;
    return tg::OK;
}

// count -> exclusive scan (int64) -> offsets; returns nnz
int scan_counts64(tg_mesh &m, int64_t *cnt_off, int64_t n, int64_t &nnz) {
    int rc = m.scan_tmp.ensure(tg::scan_tmp_bytes(n));
    if (rc) return rc;
    if ((rc = tg::scan_exclusive_i64(cnt_off, cnt_off, n, m.scan_tmp.p, m.stream))) return rc;
    return read_i64(m, cnt_off, n, nnz);
}

}  // namespace

namespace tg {

int ensure_edges(tg_mesh &m) {
    if (m.edges_built) return OK;
    int rc;
    if ((rc = ensure_vv(m)) || (rc = ensure_vt_sorted(m)) || (rc = sync_and_check(m))) return rc;
    const int64_t V = m.n_points, T = m.n_cells;
    int32_t nnz = 0;
    if ((rc = read_i32(m, m.vv_off.p, V, nnz))) return rc;
    const size_t I = sizeof(int32_t);
    if ((rc = m.edge_first.ensure(I * (size_t)std::max(nnz, 1))) ||
        (rc = m.vv_eid.ensure(I * (size_t)std::max(nnz, 1))) ||
        (rc = m.te.ensure(I * 6 * (size_t)std::max<int64_t>(T, 1))) ||
        (rc = m.tmp_a.ensure(std::max<size_t>(6 * (size_t)T, 1))) ||  // flags
        (rc = m.ve_off.ensure(I * (size_t)(V + 1))))
        return rc;
    uint8_t *flags = m.tmp_a.as<uint8_t>();
    if ((rc = fill_u8(flags, 0, 6 * T, m.stream))) return rc;
    hipLaunchKernelGGL(k_fill_i32v, dim3(grid(nnz)), dim3(256), 0, m.stream,
                       m.edge_first.as<int32_t>(), INF, (int64_t)nnz);
    TG_LAUNCH_CHECK("k_fill_i32v");
    if (V > 0) {
        hipLaunchKernelGGL(k_edge_first, dim3(grid(V, EDGE_THREADS)), dim3(EDGE_THREADS), 0, m.stream,
                           m.tv.as<const int4>(), m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(),
                           m.vv_off.as<int32_t>(), m.vv_cols.as<int32_t>(), V,
                           m.edge_first.as<int32_t>(), flags, m.te.as<int32_t>());
        TG_LAUNCH_CHECK("k_edge_first");
    }
    int32_t *bits, *wpre;
    int64_t E = 0;
    if ((rc = build_ranks(m, flags, 6 * T, bits, wpre, E))) return rc;
    m.n_edges = E;
    if ((rc = m.ev.ensure(I * 2 * (size_t)std::max<int64_t>(E, 1))) ||
        (rc = m.ve_vals.ensure(I * 2 * (size_t)std::max<int64_t>(E, 1))))
        return rc;
    int32_t *voff = m.ve_off.as<int32_t>();
    if (V > 0) {
        hipLaunchKernelGGL(k_ve_count, dim3(grid(V)), dim3(256), 0, m.stream,
                           m.vv_off.as<int32_t>(), m.vv_cols.as<int32_t>(), V, voff);
        TG_LAUNCH_CHECK("k_ve_count");
    }
    if ((rc = scan_exclusive_i32(voff, voff, V, m.scan_tmp.p, m.stream))) return rc;
    if (V > 0) {
        hipLaunchKernelGGL(k_edge_ids, dim3(grid(V)), dim3(256), 0, m.stream,
                           m.vv_off.as<int32_t>(), m.vv_cols.as<int32_t>(), V,
                           m.edge_first.as<int32_t>(), (const uint32_t *)bits, wpre,
                           m.vv_eid.as<int32_t>(), m.ev.as<int32_t>());
        TG_LAUNCH_CHECK("k_edge_ids");
        hipLaunchKernelGGL(k_ve_fill, dim3(grid(V, VE_THREADS)), dim3(VE_THREADS), 0, m.stream,
                           m.vv_off.as<int32_t>(), m.vv_cols.as<int32_t>(),
                           m.vv_eid.as<int32_t>(), V, voff, m.ve_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_ve_fill");
        hipLaunchKernelGGL(k_edge_lower, dim3(grid(V)), dim3(256), 0, m.stream,
                           m.vv_off.as<int32_t>(), m.vv_cols.as<int32_t>(), V,
                           m.vv_eid.as<int32_t>());
        TG_LAUNCH_CHECK("k_edge_lower");
    }
    if (T > 0) {
        hipLaunchKernelGGL(k_te, dim3(grid(6 * T)), dim3(256), 0, m.stream, m.te.as<int32_t>(),
                           6 * T, m.vv_eid.as<int32_t>());
        TG_LAUNCH_CHECK("k_te");
    }
    m.edges_built = true;
    return OK;
}

int ensure_et(tg_mesh &m) {
    if (m.et_built) return OK;
    int rc;
    // the vertex-lane VV form rewrites the corner lists unsorted: sort again
    if ((rc = ensure_edges(m)) || (rc = ensure_vt_sorted(m))) return rc;
    const int64_t E = m.n_edges;
    if ((rc = m.et_off.ensure(sizeof(int64_t) * (size_t)(E + 1))) ||
        (rc = m.tmp_b.ensure(sizeof(int64_t) * (size_t)(E + 1))))
        return rc;
    int64_t *off = m.et_off.as<int64_t>(), *pad = m.tmp_b.as<int64_t>();
    if (E > 0) {
        hipLaunchKernelGGL(k_et_bound, dim3(grid(E)), dim3(256), 0, m.stream, m.ev.as<int32_t>(),
                           E, m.n_points, m.vt_off.as<int32_t>(), pad);
        TG_LAUNCH_CHECK("k_et_bound");
    }
    int64_t padded = 0;
    if ((rc = scan_counts64(m, pad, E, padded))) return rc;
    if ((rc = m.tmp_c.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(padded, 1)))) return rc;
    if (E > 0) {
        hipLaunchKernelGGL(k_et, dim3(grid(E, ET_THREADS)), dim3(ET_THREADS), 0, m.stream,
                           m.ev.as<int32_t>(), E, m.n_points, m.vt_off.as<int32_t>(),
                           m.vtc.as<int32_t>(), pad, off, m.tmp_c.as<int32_t>());
        TG_LAUNCH_CHECK("k_et");
    }
    int64_t nnz = 0;
    if ((rc = scan_counts64(m, off, E, nnz))) return rc;
    if ((rc = m.et_vals.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (E > 0) {
        hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)((E + 255) / 256)), dim3(256), 0,
                           m.stream, pad, off, E, m.tmp_c.as<int32_t>(), m.et_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_pack_rows");
    }
    m.et_built = true;
    return OK;
}

int ensure_faces(tg_mesh &m) {
    if (m.faces_built) return OK;
    int rc;
    if ((rc = ensure_et(m)) || (rc = ensure_vt_sorted(m))) return rc;
    const int64_t V = m.n_points, T = m.n_cells;
    const size_t I = sizeof(int32_t);
    if ((rc = m.tmp_b.ensure(I * std::max<size_t>(4 * (size_t)T, 1))) ||  // ffirst
        (rc = m.tmp_a.ensure(std::max<size_t>(4 * (size_t)T, 1))) ||      // flags
        (rc = m.tf.ensure(I * std::max<size_t>(4 * (size_t)T, 1))))
        return rc;
    int32_t *ffirst = m.tmp_b.as<int32_t>();
    uint8_t *flags = m.tmp_a.as<uint8_t>();
    int32_t *err = m.counters.as<int32_t>() + C_FTERR;
    int32_t *iderr = m.counters.as<int32_t>() + C_IDERR;
    if ((rc = fill_i32(err, 0, 2, m.stream))) return rc;
    if ((rc = m.tmp_c.ensure(I * std::max<size_t>(4 * (size_t)T, 1)))) return rc;  // other cell
    int32_t *other = m.tmp_c.as<int32_t>();
    if (T > 0) {
        hipLaunchKernelGGL(k_face_first, dim3(grid(4 * T)), dim3(256), 0, m.stream,
                           m.tv.as<const int4>(), T, m.te.as<int32_t>(), m.n_edges,
                           m.et_off.as<int64_t>(), m.et_vals.as<int32_t>(), ffirst, flags, other,
                           iderr, err);
        TG_LAUNCH_CHECK("k_face_first");
    }
    int32_t *bits, *wpre;
    int64_t F = 0;
    if ((rc = build_ranks(m, flags, 4 * T, bits, wpre, F))) return rc;
    m.n_faces = F;
    if ((rc = m.fv.ensure(I * 3 * (size_t)std::max<int64_t>(F, 1))) ||
        (rc = m.ft.ensure(I * 2 * (size_t)std::max<int64_t>(F, 1))) ||
        (rc = m.face_off.ensure(I * (size_t)(V + 1))))
        return rc;
    if (T > 0) {
        hipLaunchKernelGGL(k_face_ids, dim3(grid(4 * T)), dim3(256), 0, m.stream,
                           m.tv.as<const int4>(), T, ffirst, other, (const uint32_t *)bits, wpre,
                           m.tf.as<int32_t>(), m.fv.as<int32_t>(), m.ft.as<int32_t>());
        TG_LAUNCH_CHECK("k_face_ids");
    }
    int32_t ferr = 0, ierr = 0;
    if ((rc = read_i32(m, err, 0, ferr)) || (rc = read_i32(m, iderr, 0, ierr))) return rc;
    if (ierr) return fail(ERR_STATE, "face ids: TE and ET disagree (a cell is missing from ET)");
    m.ft_ok = ferr == 0;
    int32_t *voff = m.face_off.as<int32_t>();
    if (V > 0) {
        hipLaunchKernelGGL((k_vf<false>), dim3(grid(V)), dim3(256), 0, m.stream,
                           m.tv.as<const int4>(), m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(), V,
                           ffirst, m.tf.as<int32_t>(), voff, nullptr);
        TG_LAUNCH_CHECK("k_vf count");
    }
    if ((rc = scan_exclusive_i32(voff, voff, V, m.scan_tmp.p, m.stream))) return rc;
    if ((rc = m.vf_vals.ensure(I * 3 * (size_t)std::max<int64_t>(F, 1)))) return rc;
    if (V > 0) {
        hipLaunchKernelGGL((k_vf<true>), dim3(grid(V)), dim3(256), 0, m.stream,
                           m.tv.as<const int4>(), m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(), V,
                           ffirst, m.tf.as<int32_t>(), voff, m.vf_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_vf fill");
    }
    m.faces_built = true;
    return OK;
}

int ensure_fe(tg_mesh &m) {
    if (m.fe_built) return OK;
    int rc;
    if ((rc = ensure_edges(m)) || (rc = ensure_faces(m))) return rc;
    const int64_t F = m.n_faces;
    if ((rc = m.fe.ensure(sizeof(int32_t) * 6 * (size_t)std::max<int64_t>(F, 1)))) return rc;
    if (F > 0) {
        hipLaunchKernelGGL(k_fe, dim3(grid(F)), dim3(256), 0, m.stream, m.tv.as<const int4>(),
                           m.fv.as<int32_t>(), F, m.ft.as<int32_t>(), m.te.as<int32_t>(),
                           m.fe.as<int32_t>());
        TG_LAUNCH_CHECK("k_fe");
    }
    m.fe_built = true;
    return OK;
}

int ensure_ft(tg_mesh &m) {
    int rc = ensure_faces(m);
    if (rc) return rc;
    if (!m.ft_ok) return fail(ERR_MESH, "a face is held by more than two cells; FT needs <= 2");
    return OK;
}

int ensure_ef(tg_mesh &m) {
    if (m.ef_built) return OK;
    int rc;
    if ((rc = ensure_et(m)) || (rc = ensure_faces(m))) return rc;
    const int64_t E = m.n_edges;
    int64_t et_nnz = 0;
    if ((rc = read_i64(m, m.et_off.p, E, et_nnz))) return rc;
    if ((rc = m.ef_off.ensure(sizeof(int64_t) * (size_t)(E + 1))) ||
        (rc = m.tmp_c.ensure(sizeof(int32_t) * (2 * (size_t)std::max<int64_t>(et_nnz, 1) + 4))))
        return rc;
    int64_t *off = m.ef_off.as<int64_t>();
    if (E > 0) {
        hipLaunchKernelGGL(k_ef, dim3(grid(E, ET_THREADS)), dim3(ET_THREADS), 0, m.stream,
                           m.et_off.as<int64_t>(), E, m.et_vals.as<int32_t>(),
                           m.te.as<int32_t>(), m.tf.as<int32_t>(), off, m.tmp_c.as<int32_t>());
        TG_LAUNCH_CHECK("k_ef");
    }
    int64_t nnz = 0;
    if ((rc = scan_counts64(m, off, E, nnz))) return rc;
    if ((rc = m.ef_vals.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (E > 0) {
        hipLaunchKernelGGL(k_ef_copy, dim3(grid(E)), dim3(256), 0, m.stream,
                           m.et_off.as<int64_t>(), E, off, m.tmp_c.as<int32_t>(),
                           m.ef_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_ef_copy");
    }
    m.ef_nnz = nnz;
    m.ef_built = true;
    return OK;
}

int ensure_tt(tg_mesh &m) {
    if (m.tt_built) return OK;
    int rc;
    if ((rc = ensure_vt_sorted(m))) return rc;
    const int64_t T = m.n_cells;
    if ((rc = m.tt_off.ensure(sizeof(int64_t) * (size_t)(T + 1))) ||
        (rc = m.tmp_b.ensure(sizeof(int64_t) * (size_t)(T + 2))))
        return rc;
    int64_t *off = m.tt_off.as<int64_t>(), *pad = m.tmp_b.as<int64_t>();
    // pad[T + 1]: set when a cell's rows exceed the LDS column (they merge
    // from global memory, where lanes that share rows gain most)
    int32_t *long_rows = reinterpret_cast<int32_t *>(pad + T + 1);
    if (T > 0) {
        hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(64), 0, m.stream, long_rows, 0, (int64_t)2);
        hipLaunchKernelGGL(k_tt_bound, dim3(grid(T)), dim3(256), 0, m.stream,
                           m.tv.as<const int4>(), T, m.vt_off.as<int32_t>(), pad, long_rows);
        TG_LAUNCH_CHECK("k_tt_bound");
    }
    int64_t padded = 0, has_long = 0;
    if ((rc = scan_counts64(m, pad, T, padded))) return rc;
    if (T > 0 && (rc = read_i64(m, pad, T + 1, has_long))) return rc;
    // tmp_c: the padded rows, then the cell order and its chunk counts
    const int64_t nchunk = (4 * T + C0_CHUNK - 1) / C0_CHUNK;
    const int64_t at_order = (padded + 1) & ~(int64_t)1;  // 8-byte aligned chunk counts
    const int64_t at_cnt = at_order + ((T + 1) & ~(int64_t)1);
    if ((rc = m.tmp_c.ensure(sizeof(int32_t) * (size_t)(at_cnt + 2 * (nchunk + 1))))) return rc;
    const int32_t *order = nullptr;
    if (T > 0 && has_long) {
        int32_t *ord = m.tmp_c.as<int32_t>() + at_order;
        int64_t *bcnt = reinterpret_cast<int64_t *>(m.tmp_c.as<int32_t>() + at_cnt);
        hipLaunchKernelGGL(k_c0_count, dim3((unsigned)nchunk), dim3(C0_THREADS), 0, m.stream,
                           m.vtc.as<int32_t>(), 4 * T, bcnt);
        TG_LAUNCH_CHECK("k_c0_count");
        int64_t found = 0;
        if ((rc = scan_counts64(m, bcnt, nchunk, found))) return rc;
        if (found == T) {  // one corner-0 code per cell
            hipLaunchKernelGGL(k_c0_fill, dim3((unsigned)nchunk), dim3(C0_THREADS), 0, m.stream,
                               m.vtc.as<int32_t>(), 4 * T, bcnt, ord);
            TG_LAUNCH_CHECK("k_c0_fill");
            order = ord;
        }
    }
    if (T > 0) {
        hipLaunchKernelGGL(k_tt, dim3(grid(T, ROW_THREADS)), dim3(ROW_THREADS), 0, m.stream,
                           m.tv.as<const int4>(), T, m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(),
                           order, pad, off, m.tmp_c.as<int32_t>());
        TG_LAUNCH_CHECK("k_tt");
    }
    int64_t nnz = 0;
    if ((rc = scan_counts64(m, off, T, nnz))) return rc;
    if ((rc = m.tt_vals.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (T > 0) {
        hipLaunchKernelGGL(k_pack_rows_wave, dim3((unsigned)((T + 255) / 256)), dim3(256), 0,
                           m.stream, pad, off, T, m.tmp_c.as<int32_t>(),
                           m.tt_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_pack_rows_wave");
    }
    m.tt_nnz = nnz;
    m.tt_built = true;
    return OK;
}

int ensure_ff(tg_mesh &m) {
    if (m.ff_built) return OK;
    int rc;
    if ((rc = ensure_ft(m))) return rc;
    const int64_t F = m.n_faces;
    if ((rc = m.ff_off.ensure(sizeof(int64_t) * (size_t)(F + 1)))) return rc;
    int64_t *off = m.ff_off.as<int64_t>();
    if (F > 0) {
        hipLaunchKernelGGL((k_ff<false>), dim3(grid(F)), dim3(256), 0, m.stream,
                           m.ft.as<int32_t>(), F, m.tf.as<int32_t>(), off, nullptr);
        TG_LAUNCH_CHECK("k_ff count");
    }
    int64_t nnz = 0;
    if ((rc = scan_counts64(m, off, F, nnz))) return rc;
    if ((rc = m.ff_vals.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (F > 0) {
        hipLaunchKernelGGL((k_ff<true>), dim3(grid(F)), dim3(256), 0, m.stream,
                           m.ft.as<int32_t>(), F, m.tf.as<int32_t>(), off, m.ff_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_ff fill");
    }
    m.ff_nnz = nnz;
    m.ff_built = true;
    return OK;
}

int ensure_ee(tg_mesh &m) {
    if (m.ee_built) return OK;
    int rc;
    if ((rc = ensure_edges(m))) return rc;
    const int64_t E = m.n_edges, V = m.n_points;
    if ((rc = m.ee_off.ensure(sizeof(int64_t) * (size_t)(E + 1)))) return rc;
    int64_t *off = m.ee_off.as<int64_t>();
    if (E > 0) {
        hipLaunchKernelGGL(k_ee_count, dim3(grid(E)), dim3(256), 0, m.stream, m.ev.as<int32_t>(),
                           E, m.vv_off.as<int32_t>(), off);
        TG_LAUNCH_CHECK("k_ee_count");
    }
    int64_t nnz = 0;
    if ((rc = scan_counts64(m, off, E, nnz))) return rc;
    if ((rc = m.ee_vals.ensure(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (E > 0) {
        // each vertex's edge ids (the vv_eid row) sorted ascending
        const int64_t n2 = 2 * E;
        // four words of slack: k_ee stages rows by whole 16-byte quads
        if ((rc = m.tmp_c.ensure(sizeof(int32_t) * (size_t)(n2 + 4)))) return rc;
        TG_HIP(hipMemcpyAsync(m.tmp_c.p, m.vv_eid.p, sizeof(int32_t) * (size_t)n2,
                              hipMemcpyDeviceToDevice, m.stream));
        if ((rc = launch_sort_rows_i32(m.vv_off.as<int32_t>(), m.tmp_c.as<int32_t>(), V,
                                       m.stream)))
            return rc;
        hipLaunchKernelGGL(k_ee, dim3(grid(E, ROW_THREADS)), dim3(ROW_THREADS), 0, m.stream,
                           m.ev.as<int32_t>(), E, m.vv_off.as<int32_t>(), m.tmp_c.as<int32_t>(),
                           off, m.ee_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_ee");
    }
    m.ee_nnz = nnz;
    m.ee_built = true;
    return OK;
}

int id_relation_view(tg_mesh &m, int32_t rel, RelView &r) {
    const int64_t V = m.n_points, T = m.n_cells, E = m.n_edges, F = m.n_faces;
    auto need = [&](bool ok, const char *what) { return ok ? OK : fail(ERR_STATE, what); };
    int rc;
    int64_t nnz = 0;
    switch (rel) {
    case TG_REL_TE:
        if ((rc = need(m.edges_built, "TE not built"))) return rc;
        r = {T, T, 6, nullptr, 0, m.te.as<int32_t>()};
        return OK;
    case TG_REL_EV:
        if ((rc = need(m.edges_built, "EV not built"))) return rc;
        r = {E, E, 2, nullptr, 0, m.ev.as<int32_t>()};
        return OK;
    case TG_REL_VE:
        if ((rc = need(m.edges_built, "VE not built"))) return rc;
        r = {V, E, 2, m.ve_off.p, 4, m.ve_vals.as<int32_t>()};
        return OK;
    case TG_REL_ET:
        if ((rc = need(m.et_built, "ET not built"))) return rc;
        if ((rc = read_i64(m, m.et_off.p, E, nnz))) return rc;
        r = {E, nnz, 1, m.et_off.p, 8, m.et_vals.as<int32_t>()};
        return OK;
    case TG_REL_TF:
        if ((rc = need(m.faces_built, "TF not built"))) return rc;
        r = {T, T, 4, nullptr, 0, m.tf.as<int32_t>()};
        return OK;
    case TG_REL_FV:
        if ((rc = need(m.faces_built, "FV not built"))) return rc;
        r = {F, F, 3, nullptr, 0, m.fv.as<int32_t>()};
        return OK;
    case TG_REL_VF:
        if ((rc = need(m.faces_built, "VF not built"))) return rc;
        r = {V, F, 3, m.face_off.p, 4, m.vf_vals.as<int32_t>()};
        return OK;
    case TG_REL_FE:
        if ((rc = need(m.fe_built, "FE not built"))) return rc;
        r = {F, F, 6, nullptr, 0, m.fe.as<int32_t>()};
        return OK;
    case TG_REL_FT:
        if ((rc = need(m.faces_built, "FT not built"))) return rc;
        if (!m.ft_ok) return fail(ERR_MESH, "a face is held by more than two cells");
        r = {F, F, 2, nullptr, 0, m.ft.as<int32_t>()};
        return OK;
    case TG_REL_EF:
        if ((rc = need(m.ef_built, "EF not built"))) return rc;
        r = {E, m.ef_nnz, 1, m.ef_off.p, 8, m.ef_vals.as<int32_t>()};
        return OK;
    case TG_REL_TT:
        if ((rc = need(m.tt_built, "TT not built"))) return rc;
        r = {T, m.tt_nnz, 1, m.tt_off.p, 8, m.tt_vals.as<int32_t>()};
        return OK;
    case TG_REL_FF:
        if ((rc = need(m.ff_built, "FF not built"))) return rc;
        r = {F, m.ff_nnz, 1, m.ff_off.p, 8, m.ff_vals.as<int32_t>()};
        return OK;
    case TG_REL_EE:
        if ((rc = need(m.ee_built, "EE not built"))) return rc;
        r = {E, m.ee_nnz, 1, m.ee_off.p, 8, m.ee_vals.as<int32_t>()};
        return OK;
    default:
        return fail(ERR_ARG, "unknown relation id");
    }
}

}  // namespace tg
