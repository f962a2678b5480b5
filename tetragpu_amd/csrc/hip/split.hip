// Critical points of a mesh in parts, each part localised on the device.
//
// Two uses:
//  - caller partitions (one id per vertex, optionally an explicit cell row per
//    partition): the reference's partition loop (src/alg/critPoints.cu:654-1003,
//    results merged as at :1438-1491).  Partition p is classified on the cells
//    that touch one of its vertices (or on its row), largest partition first,
//    and only its own vertices are kept.  Localisation (:846-913: included
//    cells, their vertices renumbered in ascending global order, scalars
//    gathered) runs here as device kernels instead of host loops.
//  - automatic split: when the hot path's device footprint for the whole mesh
//    exceeds a memory budget, or the mesh exceeds one mesh handle's 32-bit
//    limits, the vertices are cut into id ranges.  A range's part holds every
//    cell that touches a vertex of the range or one of its neighbours, so the
//    neighbours' rows are complete and every owned vertex gets exactly the
//    class it gets in the whole mesh.  Range sizes adapt: a range whose part
//    does not fit is halved, and after a part that fits the next range is
//    tried at twice the size (up to the first guess).
// TV and the scalars stay resident; every part's local mesh is built from
// them on the device and released before the next one (its buffers return
// to the library's stream-ordered cache and are reused).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "common.h"
#include "tetra_c.h"

namespace {

using tg::grid_for;

constexpr int SPLIT_BLOCK = 256;
constexpr unsigned SPLIT_GRID_CAP = 16384;

// Ownership of a part: a vertex-id range, or a partition id per vertex.
struct Owned {
    int64_t lo, hi;
    const int32_t *part;
    int32_t p;
    __device__ __forceinline__ bool operator()(int32_t v) const {
        return part ? part[v] == p : ((int64_t)v >= lo && (int64_t)v < hi);
    }
};

// ring[u] = 1 for every vertex u of a cell that has an owned vertex.
__global__ void k_ring_mark(const int4 *__restrict__ tv, int64_t T, Owned own,
                            uint8_t *__restrict__ ring) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv[c];
        if (own(q.x) || own(q.y) || own(q.z) || own(q.w)) {
            ring[q.x] = 1;
            ring[q.y] = 1;
            ring[q.z] = 1;
            ring[q.w] = 1;
        }
    }
}

// cflag[c] = 1 for the part's cells: row[c] == 1 with an explicit row, else a
// cell with a ring vertex (ring given) or an owned vertex.  incl[u] = 1 for
// their vertices; *n_cells counts the cells (64-bit).
__global__ void k_cell_flag(const int4 *__restrict__ tv, int64_t T, Owned own,
                            const uint8_t *__restrict__ ring, const int32_t *__restrict__ row,
                            int32_t *__restrict__ cflag, int32_t *__restrict__ incl,
                            unsigned long long *__restrict__ n_cells) {
    unsigned long long mine = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv[c];
        bool in;
        if (row)
            in = row[c] == 1;
        else if (ring)
            in = (ring[q.x] | ring[q.y] | ring[q.z] | ring[q.w]) != 0;
        else
            in = own(q.x) || own(q.y) || own(q.z) || own(q.w);
        cflag[c] = in ? 1 : 0;
        if (in) {
            incl[q.x] = 1;
            incl[q.y] = 1;
            incl[q.z] = 1;
            incl[q.w] = 1;
            mine++;
        }
    }
    mine = tgd::wave_sum(mine);
    if (__lane_id() == 0 && mine) atomicAdd(n_cells, mine);
}

// After the exclusive scans, cpos[c + 1] > cpos[c] marks an included cell
// and vpos[v + 1] > vpos[v] an included vertex; the positions are the local
// ids (vertices in ascending global order).
__global__ void k_local_cells(const int4 *__restrict__ tv, int64_t T,
                              const int32_t *__restrict__ cpos, const int32_t *__restrict__ vpos,
                              int4 *__restrict__ lcells) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t at = cpos[c];
        if (cpos[c + 1] == at) continue;
        const int4 q = tv[c];
        lcells[at] = make_int4(vpos[q.x], vpos[q.y], vpos[q.z], vpos[q.w]);
    }
}

template <typename S>
__global__ void k_local_verts(int64_t V, const int32_t *__restrict__ vpos,
                              const S *__restrict__ s, int32_t *__restrict__ l2g,
                              S *__restrict__ ls) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t i = vpos[v];
        if (vpos[v + 1] == i) continue;
        l2g[i] = (int32_t)v;
        ls[i] = s[v];
    }
}

// Owned vertices' classes into the whole mesh's array.
__global__ void k_scatter_owned(int64_t nv, const int32_t *__restrict__ l2g, Owned own,
                                const uint32_t *__restrict__ local, uint32_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = l2g[i];
        if (!own(g)) continue;
        out[3 * (int64_t)g] = local[3 * i];
        out[3 * (int64_t)g + 1] = local[3 * i + 1];
        out[3 * (int64_t)g + 2] = local[3 * i + 2];
    }
}

// *bad = 1 + the first cell (lowest index seen) holding an id outside [0, V).
__global__ void k_check_ids(const int4 *__restrict__ tv, int64_t T, int64_t V,
                            unsigned long long *__restrict__ bad) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv[c];
        const bool out = q.x < 0 || q.y < 0 || q.z < 0 || q.w < 0 || q.x >= V || q.y >= V ||
                         q.z >= V || q.w >= V;
        if (out) atomicMin(bad, (unsigned long long)c + 1);
    }
}

// used[u] = 1 for every vertex of a cell.
__global__ void k_used(const int4 *__restrict__ tv, int64_t T, uint8_t *__restrict__ used) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < T;
         c += (int64_t)gridDim.x * blockDim.x) {
        const int4 q = tv[c];
        used[q.x] = 1;
        used[q.y] = 1;
        used[q.z] = 1;
        used[q.w] = 1;
    }
}

// A vertex in no cell has no neighbours: the whole-mesh verdict for 0 lower
// and 0 upper components is class 4 with counts {0, 0}.
__global__ void k_isolated(int64_t V, const uint8_t *__restrict__ used,
                           uint32_t *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x)
        if (!used[v]) out[3 * v + 2] = 4u;
}

unsigned grid(int64_t n) { return std::min<unsigned>(grid_for(n, SPLIT_BLOCK), SPLIT_GRID_CAP); }

int fail_msg(int code, const std::string &msg) {
    tg::set_error(msg);
    return code;
}

// A device buffer from the library's cache, returned at scope exit.
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t s = nullptr;
    int get(size_t n, hipStream_t st) {
        s = st;
        size_t granted = 0;
        p = tg::pool_alloc(std::max<size_t>(n, 16), &granted);
        bytes = granted;
        return p ? 0
                 : fail_msg(TG_ERR_ALLOC,
                            "device allocation of " + std::to_string(n) + " bytes failed");
    }
    template <typename T>
    T *as() const { return (T *)p; }
    ~Buf() {
        if (p) tg::pool_free(p, bytes, s);
    }
};

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

// tg_mesh_create's limits for one handle.
bool handle_fits(int64_t T, int64_t V) {
    return 4 * T < (int64_t)INT32_MAX / 3 && V < (int64_t)INT32_MAX;
}

int64_t hot_footprint(int64_t T, int64_t V) {
    // what tg_mesh_create and the hot path request: TV 16T, record scratch
    // 48T, rows 48T, row descriptors 12T; per vertex degrees, offsets,
    // classes (12), heavy and fix lists (12), row sources and fix flags (8),
    // scalars (up to 8); routing histograms (at most T/8 + V/2 bytes); scan
    // scratch; counters.  Requested bytes: the cache's size classes round
    // each buffer up (to a power of two below 1 MiB, to 2 MiB above)
    return 124 * T + T / 8 + 52 * V + V / 2 +
           (int64_t)tg::scan_tmp_bytes(std::max<int64_t>(V, 12 * T)) + (16 << 10);
}

struct Splitter {
    const int32_t *cells;
    int64_t T, V;
    const void *scalars;
    int f64;
    int64_t avail = 0;  // bytes a part's mesh may use
    hipStream_t s = nullptr;
    Buf tv, sc, out, ring, cpos, vpos, counter, scan_tmp;
    tg_split_stats st{};

    size_t ssize() const { return f64 ? 8 : 4; }

    // Resident bytes besides the parts (upper bound, before allocating).
    int64_t resident(bool halo) const {
        return 16 * T + (int64_t)ssize() * V + 12 * V + 4 * (T + 1) + 4 * (V + 1) +
               (halo ? V : 0) + (int64_t)tg::scan_tmp_bytes(std::max(T, V)) + (16 << 10);
    }

    int setup(bool halo) {
        int rc;
        if ((rc = tv.get(16 * (size_t)T, s)) || (rc = sc.get(ssize() * (size_t)V, s)) ||
            (rc = out.get(12 * (size_t)V, s)) || (rc = cpos.get(4 * ((size_t)T + 1), s)) ||
            (rc = vpos.get(4 * ((size_t)V + 1), s)) || (rc = counter.get(64, s)) ||
            (rc = scan_tmp.get(tg::scan_tmp_bytes(std::max(T, V)), s)))
            return rc;
        if (halo && (rc = ring.get((size_t)V, s))) return rc;
        if (T > 0) TG_HIP(hipMemcpyAsync(tv.p, cells, 16 * (size_t)T, hipMemcpyHostToDevice, s));
        if (V > 0)
            TG_HIP(hipMemcpyAsync(sc.p, scalars, ssize() * (size_t)V, hipMemcpyHostToDevice, s));
        // every later kernel indexes per-vertex arrays with the cells' ids
        unsigned long long bad = ~0ull;
        TG_HIP(hipMemcpyAsync(counter.p, &bad, sizeof bad, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_check_ids, dim3(grid(T)), dim3(SPLIT_BLOCK), 0, s,
                           tv.as<const int4>(), T, V, counter.as<unsigned long long>());
        TG_LAUNCH_CHECK("k_check_ids");
        TG_HIP(hipMemcpyAsync(&bad, counter.p, sizeof bad, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        if (bad != ~0ull)
            return fail_msg(TG_ERR_MESH, "cell " + std::to_string(bad - 1) +
                                             " holds a vertex id outside [0, n_points)");
        return tg::fill_i32(out.as<int32_t>(), 0, 3 * V, s);
    }

    // Flags and counts of one part; *lt cells, *lv vertices.
    int measure(const Owned &own, const int32_t *row, bool halo, int64_t *lt, int64_t *lv) {
        int rc;
        const auto t0 = Clock::now();
        if (halo) {
            if ((rc = tg::fill_u8(ring.as<uint8_t>(), 0, V, s))) return rc;
            hipLaunchKernelGGL(k_ring_mark, dim3(grid(T)), dim3(SPLIT_BLOCK), 0, s,
                               tv.as<const int4>(), T, own, ring.as<uint8_t>());
            TG_LAUNCH_CHECK("k_ring_mark");
        }
        if ((rc = tg::fill_i32(vpos.as<int32_t>(), 0, V + 1, s)) ||
            (rc = tg::fill_i32(counter.as<int32_t>(), 0, 2, s)))
            return rc;
        hipLaunchKernelGGL(k_cell_flag, dim3(grid(T)), dim3(SPLIT_BLOCK), 0, s,
                           tv.as<const int4>(), T, own, halo ? ring.as<const uint8_t>() : nullptr,
                           row, cpos.as<int32_t>(), vpos.as<int32_t>(),
                           counter.as<unsigned long long>());
        TG_LAUNCH_CHECK("k_cell_flag");
        // V < 2^31: the vertex scan cannot overflow; the cell scan waits for
        // the count
        if ((rc = tg::scan_exclusive_i32(vpos.as<int32_t>(), vpos.as<int32_t>(), V, scan_tmp.p, s)))
            return rc;
        unsigned long long nc = 0;
        int32_t nv = 0;
        TG_HIP(hipMemcpyAsync(&nc, counter.p, sizeof nc, hipMemcpyDeviceToHost, s));
        TG_HIP(hipMemcpyAsync(&nv, vpos.as<int32_t>() + V, sizeof nv, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        *lt = (int64_t)nc;
        *lv = nv;
        st.localize_s += secs(t0, Clock::now());
        return 0;
    }

    bool fits(int64_t lt, int64_t lv) const {
        const int64_t local = 16 * lt + (4 + (int64_t)ssize()) * lv;  // lcells, l2g, scalars
        return handle_fits(lt, lv) && hot_footprint(lt, lv) + local <= avail;
    }

    // Localise, classify and merge the part measured last.
    int run_part(const Owned &own, int64_t lt, int64_t lv) {
        int rc;
        auto t0 = Clock::now();
        if ((rc = tg::scan_exclusive_i32(cpos.as<int32_t>(), cpos.as<int32_t>(), T, scan_tmp.p, s)))
            return rc;
        Buf lcells, l2g, ls;
        if ((rc = lcells.get(16 * (size_t)lt, s)) || (rc = l2g.get(4 * (size_t)lv, s)) ||
            (rc = ls.get(ssize() * (size_t)lv, s)))
            return rc;
        hipLaunchKernelGGL(k_local_cells, dim3(grid(T)), dim3(SPLIT_BLOCK), 0, s,
                           tv.as<const int4>(), T, cpos.as<const int32_t>(),
                           vpos.as<const int32_t>(), lcells.as<int4>());
        TG_LAUNCH_CHECK("k_local_cells");
        if (f64)
            hipLaunchKernelGGL(k_local_verts<double>, dim3(grid(V)), dim3(SPLIT_BLOCK), 0, s, V,
                               vpos.as<const int32_t>(), sc.as<const double>(), l2g.as<int32_t>(),
                               ls.as<double>());
        else
            hipLaunchKernelGGL(k_local_verts<float>, dim3(grid(V)), dim3(SPLIT_BLOCK), 0, s, V,
                               vpos.as<const int32_t>(), sc.as<const float>(), l2g.as<int32_t>(),
                               ls.as<float>());
        TG_LAUNCH_CHECK("k_local_verts");
        TG_HIP(hipStreamSynchronize(s));
        auto t1 = Clock::now();
        st.localize_s += secs(t0, t1);
        // the part's mesh on the same stream (inputs copied on the device)
        tg_mesh *m = nullptr;
        rc = tg_mesh_create(&m, lcells.as<const int32_t>(), lt, lv, ls.p, f64, 1, (void *)s);
        const uint32_t *dcls = nullptr;
        if (rc == TG_OK) rc = tg_vv_critpoints(m);
        if (rc == TG_OK) rc = tg_classes_device(m, &dcls);
        if (rc == TG_OK) {
            hipLaunchKernelGGL(k_scatter_owned, dim3(grid(lv)), dim3(SPLIT_BLOCK), 0, s, lv,
                               l2g.as<const int32_t>(), own, dcls, out.as<uint32_t>());
            rc = hipGetLastError() == hipSuccess ? tg_sync(m)
                                                 : fail_msg(TG_ERR_HIP, "k_scatter_owned launch");
        }
        if (m) {
            const std::string err = rc == TG_OK ? "" : tg_last_error();
            tg_mesh_destroy(m);
            if (rc != TG_OK) return fail_msg(rc, err);
        } else if (rc != TG_OK) {
            return rc;
        }
        st.classify_s += secs(t1, Clock::now());
        st.parts++;
        st.max_cells = std::max(st.max_cells, lt);
        st.max_points = std::max(st.max_points, lv);
        st.cells_total += lt;
        return TG_OK;
    }

    int isolated() {
        Buf used;
        int rc;
        if ((rc = used.get((size_t)V, s)) || (rc = tg::fill_u8(used.as<uint8_t>(), 0, V, s)))
            return rc;
        hipLaunchKernelGGL(k_used, dim3(grid(T)), dim3(SPLIT_BLOCK), 0, s, tv.as<const int4>(), T,
                           used.as<uint8_t>());
        TG_LAUNCH_CHECK("k_used");
        hipLaunchKernelGGL(k_isolated, dim3(grid(V)), dim3(SPLIT_BLOCK), 0, s, V,
                           used.as<const uint8_t>(), out.as<uint32_t>());
        TG_LAUNCH_CHECK("k_isolated");
        return TG_OK;
    }

    int finish(uint32_t *classes3) {
        if (V > 0)
            TG_HIP(hipMemcpyAsync(classes3, out.p, 12 * (size_t)V, hipMemcpyDeviceToHost, s));
        TG_HIP(hipStreamSynchronize(s));
        return TG_OK;
    }
};

int whole_mesh(const int32_t *cells, int64_t T, int64_t V, const void *scalars, int f64,
               uint32_t *classes3, hipStream_t s) {
    tg_mesh *m = nullptr;
    int rc = tg_mesh_create(&m, cells, T, V, scalars, f64, 0, (void *)s);
    if (rc == TG_OK) rc = tg_vv_critpoints(m);
    if (rc == TG_OK) rc = tg_classes_copy(m, classes3);
    if (m) {
        const std::string err = rc == TG_OK ? "" : tg_last_error();
        tg_mesh_destroy(m);
        if (rc != TG_OK) return fail_msg(rc, err);
    }
    return rc;
}

}  // namespace

extern "C" int tg_device_budget(int64_t *bytes) {
    if (!bytes) return fail_msg(TG_ERR_ARG, "bytes is NULL");
    size_t free_b = 0, total_b = 0;
    TG_HIP(hipMemGetInfo(&free_b, &total_b));
    *bytes = (int64_t)((double)(free_b + tg::pool_cached_bytes()) * 0.9);
    return TG_OK;
}

extern "C" int tg_vv_footprint(int64_t n_cells, int64_t n_points, int64_t *bytes) {
    if (n_cells < 0 || n_points < 0 || !bytes) return fail_msg(TG_ERR_ARG, "bad arguments");
    *bytes = hot_footprint(n_cells, n_points);
    return TG_OK;
}

extern "C" int tg_critpoints_split(const int32_t *cells, int64_t n_cells, int64_t n_points,
                                   const void *scalars, int scalar_f64, const int32_t *parts,
                                   int32_t n_parts, const int32_t *part_cells,
                                   int64_t budget_bytes, uint32_t *classes3,
                                   tg_split_stats *stats) {
    const auto t_begin = Clock::now();
    if (n_cells < 0 || n_points < 0 || (n_cells > 0 && !cells) ||
        (n_points > 0 && (!scalars || !classes3)) || (parts && n_parts < 1) ||
        (part_cells && !parts))
        return fail_msg(TG_ERR_ARG, "bad arguments");
    if (n_points >= (int64_t)INT32_MAX) return fail_msg(TG_ERR_ARG, "too many points for int32 ids");
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    int rc = tg_init(dev);
    if (rc) return rc;
    // the stream outlives every buffer freed on it (declared first, destroyed last)
    struct OwnedStream {
        hipStream_t s = nullptr;
        ~OwnedStream() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        }
    } os;
    TG_HIP(hipStreamCreateWithFlags(&os.s, hipStreamNonBlocking));
    Splitter sp;
    sp.s = os.s;
    sp.cells = cells;
    sp.T = n_cells;
    sp.V = n_points;
    sp.scalars = scalars;
    sp.f64 = scalar_f64 ? 1 : 0;
    sp.st.footprint_bytes = hot_footprint(n_cells, n_points);
    if (budget_bytes <= 0 && (rc = tg_device_budget(&budget_bytes))) return rc;
    sp.st.budget_bytes = budget_bytes;
    std::vector<int32_t> hparts;
    if (parts) {
        for (int64_t v = 0; v < n_points; v++)
            if (parts[v] < 0 || parts[v] >= n_parts)
                return fail_msg(TG_ERR_ARG, "vertex " + std::to_string(v) + " has partition id " +
                                                std::to_string(parts[v]) + " outside 0.." +
                                                std::to_string(n_parts - 1));
    } else if (handle_fits(n_cells, n_points) && sp.st.footprint_bytes <= budget_bytes) {
        // the whole mesh fits: one handle, no localisation
        if ((rc = whole_mesh(cells, n_cells, n_points, scalars, sp.f64, classes3, sp.s))) return rc;
        sp.st.parts = 1;
        sp.st.max_cells = sp.st.cells_total = n_cells;
        sp.st.max_points = n_points;
        sp.st.total_s = sp.st.classify_s = secs(t_begin, Clock::now());
        if (stats) *stats = sp.st;
        return TG_OK;
    }
    const bool halo = parts == nullptr;
    sp.avail = budget_bytes - sp.resident(halo);
    if (sp.avail <= 0)
        return fail_msg(TG_ERR_ALLOC, "cells and scalars alone (" +
                                          std::to_string(sp.resident(halo)) +
                                          " bytes resident) exceed the device budget of " +
                                          std::to_string(budget_bytes) + " bytes");
    {
        Splitter &S = sp;
        if ((rc = S.setup(halo))) return rc;
        if (parts) {
            // largest partition first (critPoints.cu:658-666)
            std::vector<int64_t> count((size_t)n_parts, 0);
            for (int64_t v = 0; v < n_points; v++) count[(size_t)parts[v]]++;
            std::vector<int32_t> order((size_t)n_parts);
            for (int32_t p = 0; p < n_parts; p++) order[(size_t)p] = p;
            std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
                return count[(size_t)a] > count[(size_t)b];
            });
            Buf dparts, row;
            if ((rc = dparts.get(4 * (size_t)n_points, S.s))) return rc;
            TG_HIP(hipMemcpyAsync(dparts.p, parts, 4 * (size_t)n_points, hipMemcpyHostToDevice,
                                  S.s));
            if (part_cells && (rc = row.get(4 * (size_t)n_cells, S.s))) return rc;
            for (int32_t p : order) {
                const int32_t *drow = nullptr;
                if (part_cells) {
                    TG_HIP(hipMemcpyAsync(row.p, part_cells + (size_t)p * (size_t)n_cells,
                                          4 * (size_t)n_cells, hipMemcpyHostToDevice, S.s));
                    drow = row.as<const int32_t>();
                }
                const Owned own{0, 0, dparts.as<const int32_t>(), p};
                int64_t lt = 0, lv = 0;
                if ((rc = S.measure(own, drow, false, &lt, &lv))) return rc;
                if (lt == 0) continue;
                if (!S.fits(lt, lv))
                    return fail_msg(TG_ERR_ALLOC,
                                    "partition " + std::to_string(p) + " (" + std::to_string(lt) +
                                        " cells, " + std::to_string(lv) +
                                        " points) does not fit the device budget");
                if ((rc = S.run_part(own, lt, lv))) return rc;
            }
        } else {
            const int64_t k0 = std::max<int64_t>(
                2, (sp.st.footprint_bytes + S.avail - 1) / std::max<int64_t>(S.avail, 1));
            const int64_t r0 = std::max<int64_t>(1, (n_points + k0 - 1) / k0);
            int64_t r = r0;
            for (int64_t lo = 0; lo < n_points;) {
                const int64_t hi = std::min(n_points, lo + r);
                const Owned own{lo, hi, nullptr, 0};
                int64_t lt = 0, lv = 0;
                if ((rc = S.measure(own, nullptr, true, &lt, &lv))) return rc;
                if (lt > 0 && !S.fits(lt, lv)) {
                    if (r == 1)
                        return fail_msg(TG_ERR_ALLOC,
                                        "the part of vertex " + std::to_string(lo) + " (" +
                                            std::to_string(lt) +
                                            " cells) does not fit the device budget");
                    r = (r + 1) / 2;
                    continue;
                }
                if (lt > 0 && (rc = S.run_part(own, lt, lv))) return rc;
                lo = hi;
                r = std::min(2 * r, r0);  // halos vary along the mesh: try larger again
            }
            if ((rc = S.isolated())) return rc;
        }
        if ((rc = S.finish(classes3))) return rc;
    }
    sp.st.total_s = secs(t_begin, Clock::now());
    if (stats) *stats = sp.st;
    return TG_OK;
}
