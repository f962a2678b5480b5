// The reference's GPU relation builders (include/cuda_extraction.h:10-138)
// over the C ABI of libtetra_hip.so.  The reference keeps TV/VE/VF in
// file-scope device globals; here each calling thread has a current mesh
// handle (created by make_TV_for_GPU, or by any builder that receives a TV
// it has not seen), so worker threads do not share state.
#include "cuda_extraction.h"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tetra_c.h"

namespace {

// The caller's edge or face numbering, as a map from the device's ids (the
// first-occurrence order of make_TE_and_VE / make_TF_and_VF) to the ids in
// the caller's VE or VF table.  The reference uploads the caller's table and
// computes every id relation from it (src/cuda_extraction.cu:85-252, used by
// EV/TF/TE/FV/FE at :271, :409, :632, :770), so its results carry the
// caller's ids; here the device's results are renumbered through this map.
struct IdMap {
    const void *table = nullptr;  // the VE_Data / VF_Data it was made from
    size_t rows = 0;
    std::vector<int64_t> to_caller;
    bool identity = true;
};

struct Session {
    tg_mesh *mesh = nullptr;
    const TV_Data *tv = nullptr;  // the TV the mesh was made from
    vtkIdType n_points = 0, n_cells = 0;
    bool initialised = false;
    IdMap edges, faces;
};

thread_local Session g_session;

[[noreturn]] void die(const char *what, int status) {
    std::fprintf(stderr, "tetra: %s failed [%d]: %s\n", what, status, tg_last_error());
    std::exit(EXIT_FAILURE);
}

void ok(int status, const char *what) {
    if (status != TG_OK) die(what, status);
}

void release_session() {
    if (g_session.mesh) tg_mesh_destroy(g_session.mesh);
    g_session.mesh = nullptr;
    g_session.tv = nullptr;
    g_session.edges = IdMap();
    g_session.faces = IdMap();
}

// (Re)create the thread's mesh from TV: cells narrowed to int32 with a range
// check, vertexAttributes (if any) as float64 scalars.
void upload(const TV_Data &tv) {
    if (!g_session.initialised) {
        ok(tg_init(0), "tg_init");
        g_session.initialised = true;
    }
    release_session();
    if (tv.nPoints > INT32_MAX || 4 * tv.nCells > INT32_MAX) {
        std::fprintf(stderr, "tetra: mesh too large for 32-bit device ids\n");
        std::exit(EXIT_FAILURE);
    }
    std::vector<int32_t> cells((size_t)(4 * tv.nCells));
    for (size_t i = 0; i < cells.size(); i++) {
        const vtkIdType x = tv.cells[i];
        if (x < 0 || x >= tv.nPoints) {
            std::fprintf(stderr, "tetra: cell %zu references vertex %lld outside [0, %lld)\n",
                         i / 4, x, tv.nPoints);
            std::exit(EXIT_FAILURE);
        }
        cells[i] = (int32_t)x;
    }
    ok(tg_mesh_create(&g_session.mesh, cells.data(), tv.nCells, tv.nPoints, tv.vertexAttributes,
                      1, 0, nullptr),
       "tg_mesh_create");
    g_session.tv = &tv;
    g_session.n_points = tv.nPoints;
    g_session.n_cells = tv.nCells;
}

tg_mesh *mesh_for(const TV_Data &tv) {
    if (!g_session.mesh || g_session.tv != &tv || g_session.n_cells != tv.nCells ||
        g_session.n_points != tv.nPoints)
        upload(tv);
    return g_session.mesh;
}

tg_mesh *current(const char *who) {
    if (!g_session.mesh) {
        std::fprintf(stderr, "tetra: %s needs make_TV_for_GPU first\n", who);
        std::exit(EXIT_FAILURE);
    }
    return g_session.mesh;
}

void build(tg_mesh *m, int rel) { ok(tg_build(m, TG_FLAG(rel)), "tg_build"); }

// host copy of a built relation: int64 offsets (empty for fixed) + values
struct HostRel {
    int64_t rows = 0, nnz = 0;
    int32_t width = 1;
    bool fixed = true;
    std::vector<int64_t> off;
    std::vector<int32_t> val;
};

HostRel fetch(tg_mesh *m, int rel) {
    HostRel h;
    int32_t fixed = 1;
    ok(tg_relation_shape(m, rel, &h.rows, &h.nnz, &h.width, &fixed), "tg_relation_shape");
    h.fixed = fixed != 0;
    h.val.resize((size_t)(h.nnz * h.width));
    if (!h.fixed) h.off.resize((size_t)(h.rows + 1));
    ok(tg_relation_copy(m, rel, h.fixed ? nullptr : h.off.data(), h.val.data()),
       "tg_relation_copy");
    return h;
}

std::vector<std::vector<vtkIdType>> to_rows(const HostRel &h) {
    std::vector<std::vector<vtkIdType>> out((size_t)h.rows);
    for (int64_t r = 0; r < h.rows; r++)
        out[r].assign(h.val.begin() + h.off[r], h.val.begin() + h.off[r + 1]);
    return out;
}

template <size_t W>
std::vector<std::array<vtkIdType, W>> to_fixed(const HostRel &h) {
    std::vector<std::array<vtkIdType, W>> out((size_t)h.rows);
    for (int64_t r = 0; r < h.rows; r++)
        for (size_t k = 0; k < W; k++) out[r][k] = h.val[(size_t)r * W + k];
    return out;
}

vtkIdType *export_values(tg_mesh *m, int rel) {
    void *vals = nullptr;
    ok(tg_relation_export(m, rel, 8, nullptr, 8, &vals), "tg_relation_export");
    return (vtkIdType *)vals;
}

void expect_count(const char *what, vtkIdType got, vtkIdType want) {
    if (got != want) {
        std::fprintf(stderr, "tetra: %s: the device finds %lld, the caller's table has %lld\n",
                     what, got, want);
        std::exit(EXIT_FAILURE);
    }
}

int64_t n_edges(tg_mesh *m) {
    int64_t v, t, e, f;
    build(m, TG_REL_EV);
    ok(tg_mesh_counts(m, &v, &t, &e, &f), "tg_mesh_counts");
    return e;
}

int64_t n_faces(tg_mesh *m) {
    int64_t v, t, e, f;
    build(m, TG_REL_FV);
    ok(tg_mesh_counts(m, &v, &t, &e, &f), "tg_mesh_counts");
    return f;
}

[[noreturn]] void table_mismatch(const char *what, vtkIdType v, const char *detail) {
    std::fprintf(stderr, "tetra: the caller's %s does not describe this mesh (vertex %lld: %s)\n",
                 what, (long long)v, detail);
    std::exit(EXIT_FAILURE);
}

// Map device edge ids to the ids of the caller's VE: row v of VE must hold
// exactly the device's edges (v, hi) with hi > v, in any order, and the ids
// must be a permutation of [0, E).
const IdMap &edge_map(tg_mesh *m, const VE_Data &ve) {
    IdMap &M = g_session.edges;
    if (M.table == &ve && M.rows == ve.size()) return M;
    build(m, TG_REL_VE);
    const HostRel d = fetch(m, TG_REL_VE);  // rows ascending by id: {hi, id}
    if ((int64_t)ve.size() != d.rows) table_mismatch("VE", -1, "row count differs");
    const int64_t E = d.nnz;
    std::vector<int64_t> map((size_t)E, -1);
    std::vector<char> seen((size_t)E, 0);
    bool identity = true;
    for (int64_t v = 0; v < d.rows; v++) {
        const auto &row = ve[(size_t)v];
        if ((int64_t)row.size() != d.off[v + 1] - d.off[v]) table_mismatch("VE", v, "row length differs");
        for (int64_t i = d.off[v]; i < d.off[v + 1]; i++) {
            const int64_t hi = d.val[2 * i], id = d.val[2 * i + 1];
            int64_t cid = -1;
            for (const EdgeData &e : row)
                if (e.highVert == hi) {
                    cid = e.id;
                    break;
                }
            if (cid < 0) table_mismatch("VE", v, "an edge of the mesh is missing");
            if (cid >= E || seen[(size_t)cid]) table_mismatch("VE", v, "ids are not a permutation of [0, E)");
            seen[(size_t)cid] = 1;
            map[(size_t)id] = cid;
            identity = identity && cid == id;
        }
    }
    M.table = &ve;
    M.rows = ve.size();
    M.to_caller = std::move(map);
    M.identity = identity;
    return M;
}

// Map device face ids to the caller's VF ids: row v of VF must hold exactly
// the device's faces whose positional low vertex is v, matched by (mid, hi).
const IdMap &face_map(tg_mesh *m, const VF_Data &vf) {
    IdMap &M = g_session.faces;
    if (M.table == &vf && M.rows == vf.size()) return M;
    build(m, TG_REL_VF);
    const HostRel d = fetch(m, TG_REL_VF);  // {mid, hi, id}
    if ((int64_t)vf.size() != d.rows) table_mismatch("VF", -1, "row count differs");
    const int64_t F = d.nnz;
    std::vector<int64_t> map((size_t)F, -1);
    std::vector<char> seen((size_t)F, 0);
    bool identity = true;
    for (int64_t v = 0; v < d.rows; v++) {
        const auto &row = vf[(size_t)v];
        if ((int64_t)row.size() != d.off[v + 1] - d.off[v]) table_mismatch("VF", v, "row length differs");
        for (int64_t i = d.off[v]; i < d.off[v + 1]; i++) {
            const int64_t mid = d.val[3 * i], hi = d.val[3 * i + 1], id = d.val[3 * i + 2];
            int64_t cid = -1;
            for (const FaceData &f : row)
                if (f.middleVert == mid && f.highVert == hi) {
                    cid = f.id;
                    break;
                }
            if (cid < 0) table_mismatch("VF", v, "a face of the mesh is missing");
            if (cid >= F || seen[(size_t)cid]) table_mismatch("VF", v, "ids are not a permutation of [0, F)");
            seen[(size_t)cid] = 1;
            map[(size_t)id] = cid;
            identity = identity && cid == id;
        }
    }
    M.table = &vf;
    M.rows = vf.size();
    M.to_caller = std::move(map);
    M.identity = identity;
    return M;
}

// Device relation in the caller's numbering: rows indexed by a renumbered id
// move to their new index (row_map), and value columns holding ids are
// renumbered (val_map; for FE only the id column of each {hi, id} pair).
std::vector<int64_t> renumbered(const HostRel &h, const IdMap *row_map, const IdMap *val_map,
                                int val_stride = 1, int val_col = 0) {
    const size_t W = (size_t)h.width;
    std::vector<int64_t> out(h.val.size());
    for (int64_t r = 0; r < h.rows; r++) {
        const size_t dst = row_map ? (size_t)row_map->to_caller[(size_t)r] : (size_t)r;
        for (size_t k = 0; k < W; k++) {
            int64_t x = h.val[(size_t)r * W + k];
            if (val_map && (int)(k % val_stride) == val_col) x = val_map->to_caller[(size_t)x];
            out[dst * W + k] = x;
        }
    }
    return out;
}

vtkIdType *upload_values(const std::vector<int64_t> &v) {
    void *d = nullptr;
    ok(tg_device_upload(v.data(), v.size() * sizeof(int64_t), &d), "tg_device_upload");
    return (vtkIdType *)d;
}

void finish(bool free_transients) {
    if (free_transients) release_session();
}

int max_incidence_bound(const TV_Data &TV, const vtkIdType n_points, const vtkIdType *ivvi,
                        int debug) {
    std::vector<int> seen((size_t)n_points, 0);
    int most = 0;
    for (vtkIdType i = 0; i < 4 * TV.nCells; i++) {
        const vtkIdType v = ivvi ? ivvi[TV.cells[i]] : TV.cells[i];
        most = std::max(most, ++seen[v]);
    }
    const int bound = ((3 * most + 31) / 32) * 32;
    if (debug > NO_DEBUG) std::fprintf(stderr, "Approximated max VV adjacency: %d\n", bound);
    return bound;
}

}  // namespace

void free_transients_GPU() { release_session(); }

// The calling thread's current mesh handle (used by the critical-point
// driver to run the C ABI on the mesh make_TV_for_GPU uploaded).
tg_mesh *tetra_current_mesh() { return g_session.mesh; }

void make_TV_for_GPU(const TV_Data &tv_relationship) { upload(tv_relationship); }

// Transfer and translation intervals of the host-returning builders, in the
// reference timer's line format (src/metrics.cpp:104-113: seconds, then
// microseconds zero-padded to six digits).
namespace {
std::atomic<bool> g_report_times{false};
thread_local std::vector<HostTiming> g_times;
constexpr const char *kTransfer = "GPU Device->Host transfer";
constexpr const char *kTranslation = "GPU Device->Host translation";

struct Stopwatch {
    const char *timer;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit Stopwatch(const char *name) : timer(name) {}
    void lap(const char *label) {
        const auto t1 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
        // a caller that never takes them keeps the latest 1024 intervals
        if (g_times.size() >= 1024) g_times.erase(g_times.begin(), g_times.begin() + 512);
        g_times.push_back(HostTiming{timer, label, s});
        if (g_report_times.load(std::memory_order_relaxed)) {
            const long long us = (long long)(s * 1e6);
            std::printf("\u23f3Timer[%s] %s: %lld.%06lld\n", timer, label, us / 1000000,
                        us % 1000000);
            std::fflush(stdout);
        }
    }
};
}  // namespace

// The reference declares (table, n_edges, n_verts) (include/cuda_extraction.h:
// 15-17) but defines and calls it as (table, n_verts, n_edges)
// (src/cuda_extraction.cu:85-88, :271); the count equal to the table's row
// count is taken as the vertex count, the other as the edge count.
void make_VE_for_GPU(const VE_Data &ve, const vtkIdType a, const vtkIdType b) {
    tg_mesh *m = current("make_VE_for_GPU");
    const vtkIdType n_verts = (vtkIdType)ve.size();
    if (a != n_verts && b != n_verts) expect_count("vertices of VE", n_verts, b);
    expect_count("edges", n_edges(m), a == n_verts ? b : a);
    edge_map(m, ve);
}

void make_VF_for_GPU(const VF_Data &vf, const vtkIdType n_verts, const vtkIdType n_faces_) {
    tg_mesh *m = current("make_VF_for_GPU");
    expect_count("vertices of VF", (vtkIdType)vf.size(), n_verts);
    expect_count("faces", n_faces(m), n_faces_);
    face_map(m, vf);
}

vtkIdType *make_EV_GPU_return(const VE_Data &VE, const vtkIdType, const vtkIdType n_edges_,
                              const bool free_transients) {
    tg_mesh *m = current("make_EV_GPU_return");
    expect_count("edges", n_edges(m), n_edges_);
    const IdMap &em = edge_map(m, VE);
    vtkIdType *out = em.identity ? export_values(m, TG_REL_EV)
                                 : upload_values(renumbered(fetch(m, TG_REL_EV), &em, nullptr));
    finish(free_transients);
    return out;
}

std::unique_ptr<EV_Data> make_EV_GPU(const VE_Data &VE, const vtkIdType, const vtkIdType n_edges_,
                                     const bool free_transients) {
    tg_mesh *m = current("make_EV_GPU");
    expect_count("edges", n_edges(m), n_edges_);
    const IdMap &em = edge_map(m, VE);
    Stopwatch sw("EV_GPU");
    HostRel h = fetch(m, TG_REL_EV);
    sw.lap(kTransfer);
    if (!em.identity) {
        const auto v = renumbered(h, &em, nullptr);
        h.val.assign(v.begin(), v.end());
    }
    auto out = std::make_unique<EV_Data>(to_fixed<2>(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

vtkIdType *make_TF_GPU_return(const TV_Data &TV, const VF_Data &VF, const vtkIdType,
                              const vtkIdType n_faces_, const vtkIdType,
                              const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &fm = face_map(m, VF);
    build(m, TG_REL_TF);
    vtkIdType *out = fm.identity ? export_values(m, TG_REL_TF)
                                 : upload_values(renumbered(fetch(m, TG_REL_TF), nullptr, &fm));
    finish(free_transients);
    return out;
}

std::unique_ptr<TF_Data> make_TF_GPU(const TV_Data &TV, const VF_Data &VF, const vtkIdType,
                                     const vtkIdType n_faces_, const vtkIdType,
                                     const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &fm = face_map(m, VF);
    build(m, TG_REL_TF);
    Stopwatch sw("TF_GPU");
    HostRel h = fetch(m, TG_REL_TF);
    sw.lap(kTransfer);
    if (!fm.identity) {
        const auto v = renumbered(h, nullptr, &fm);
        h.val.assign(v.begin(), v.end());
    }
    auto out = std::make_unique<TF_Data>(to_fixed<4>(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

vtkIdType *make_TE_GPU_return(const TV_Data &TV, const VE_Data &VE, const vtkIdType,
                              const vtkIdType n_edges_, const vtkIdType,
                              const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    expect_count("edges", n_edges(m), n_edges_);
    const IdMap &em = edge_map(m, VE);
    build(m, TG_REL_TE);
    vtkIdType *out = em.identity ? export_values(m, TG_REL_TE)
                                 : upload_values(renumbered(fetch(m, TG_REL_TE), nullptr, &em));
    finish(free_transients);
    return out;
}

std::unique_ptr<TE_Data> make_TE_GPU(const TV_Data &TV, const VE_Data &VE, const vtkIdType,
                                     const vtkIdType n_edges_, const vtkIdType,
                                     const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    expect_count("edges", n_edges(m), n_edges_);
    const IdMap &em = edge_map(m, VE);
    build(m, TG_REL_TE);
    Stopwatch sw("TE_GPU");
    HostRel h = fetch(m, TG_REL_TE);
    sw.lap(kTransfer);
    if (!em.identity) {
        const auto v = renumbered(h, nullptr, &em);
        h.val.assign(v.begin(), v.end());
    }
    auto out = std::make_unique<TE_Data>(to_fixed<6>(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

vtkIdType *make_FV_GPU_return(const VF_Data &VF, const vtkIdType, const vtkIdType n_faces_,
                              const bool free_transients) {
    tg_mesh *m = current("make_FV_GPU_return");
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &fm = face_map(m, VF);
    vtkIdType *out = fm.identity ? export_values(m, TG_REL_FV)
                                 : upload_values(renumbered(fetch(m, TG_REL_FV), &fm, nullptr));
    finish(free_transients);
    return out;
}

std::unique_ptr<FV_Data> make_FV_GPU(const VF_Data &VF, const vtkIdType, const vtkIdType n_faces_,
                                     const bool free_transients) {
    tg_mesh *m = current("make_FV_GPU");
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &fm = face_map(m, VF);
    Stopwatch sw("FV_GPU");
    HostRel h = fetch(m, TG_REL_FV);
    sw.lap(kTransfer);
    if (!fm.identity) {
        const auto v = renumbered(h, &fm, nullptr);
        h.val.assign(v.begin(), v.end());
    }
    auto out = std::make_unique<FV_Data>((size_t)h.rows);
    for (int64_t f = 0; f < h.rows; f++)
        (*out)[f] = FaceData(h.val[3 * f], h.val[3 * f + 1], h.val[3 * f + 2]);
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

vtkIdType *make_FE_GPU_return(const VF_Data &VF, const VE_Data &VE, const vtkIdType,
                              const vtkIdType n_edges_, const vtkIdType n_faces_,
                              const bool free_transients) {
    tg_mesh *m = current("make_FE_GPU_return");
    expect_count("edges", n_edges(m), n_edges_);
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &em = edge_map(m, VE);
    const IdMap &fm = face_map(m, VF);
    build(m, TG_REL_FE);
    vtkIdType *out = em.identity && fm.identity
                         ? export_values(m, TG_REL_FE)
                         : upload_values(renumbered(fetch(m, TG_REL_FE), &fm, &em, 2, 1));
    finish(free_transients);
    return out;
}

std::unique_ptr<FE_Data> make_FE_GPU(const VF_Data &VF, const VE_Data &VE, const vtkIdType,
                                     const vtkIdType n_edges_, const vtkIdType n_faces_,
                                     const bool free_transients) {
    tg_mesh *m = current("make_FE_GPU");
    expect_count("edges", n_edges(m), n_edges_);
    expect_count("faces", n_faces(m), n_faces_);
    const IdMap &em = edge_map(m, VE);
    const IdMap &fm = face_map(m, VF);
    build(m, TG_REL_FE);
    Stopwatch sw("FE_GPU");
    HostRel h = fetch(m, TG_REL_FE);
    sw.lap(kTransfer);
    if (!em.identity || !fm.identity) {
        const auto v = renumbered(h, &fm, &em, 2, 1);
        h.val.assign(v.begin(), v.end());
    }
    auto out = std::make_unique<FE_Data>((size_t)h.rows);
    for (int64_t f = 0; f < h.rows; f++)
        for (int k = 0; k < 3; k++)
            (*out)[f][k] = EdgeData(h.val[6 * f + 2 * k], h.val[6 * f + 2 * k + 1]);
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

int get_approx_max_VV_partitioned(const TV_Data &TV, const vtkIdType n_points,
                                  const vtkIdType *ivvi, const int debug) {
    return max_incidence_bound(TV, n_points, ivvi, debug);
}

int get_approx_max_VV(const TV_Data &TV, const vtkIdType n_points, const int debug) {
    return max_incidence_bound(TV, n_points, nullptr, debug);
}

device_VV *make_VV_GPU_return(const TV_Data &TV, const int, const int, const int,
                              const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    build(m, TG_REL_VV);
    void *off = nullptr, *val = nullptr;
    ok(tg_relation_export(m, TG_REL_VV, 4, &off, 8, &val), "tg_relation_export");
    finish(free_transients);
    return new device_VV((vtkIdType *)val, (unsigned int *)off);
}

std::unique_ptr<VV_Data> make_VV_GPU(const TV_Data &TV, const int, const int,
                                     const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    build(m, TG_REL_VV);
    Stopwatch sw("VV_GPU");
    const HostRel h = fetch(m, TG_REL_VV);
    sw.lap(kTransfer);
    auto out = std::make_unique<VV_Data>(to_rows(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

device_VT *make_VT_GPU_return(const TV_Data &TV) {
    tg_mesh *m = mesh_for(TV);
    build(m, TG_REL_VT);
    void *off = nullptr, *val = nullptr;
    ok(tg_relation_export(m, TG_REL_VT, 8, &off, 8, &val), "tg_relation_export");
    return new device_VT((vtkIdType *)val, (unsigned long long *)off);
}

device_VT *make_VT_GPU_return(const VT_Data &VT) {
    std::vector<unsigned long long> off(VT.size() + 1, 0);
    for (size_t v = 0; v < VT.size(); v++) off[v + 1] = off[v] + VT[v].size();
    std::vector<vtkIdType> val;
    val.reserve(off.back());
    for (const auto &row : VT) val.insert(val.end(), row.begin(), row.end());
    if (!g_session.initialised) {
        ok(tg_init(0), "tg_init");
        g_session.initialised = true;
    }
    void *d_off = nullptr, *d_val = nullptr;
    ok(tg_device_upload(off.data(), off.size() * sizeof(off[0]), &d_off), "tg_device_upload");
    ok(tg_device_upload(val.data(), val.size() * sizeof(val[0]), &d_val), "tg_device_upload");
    return new device_VT((vtkIdType *)d_val, (unsigned long long *)d_off);
}

namespace {
template <int REL>
std::unique_ptr<std::vector<std::vector<vtkIdType>>> csr_relation(const TV_Data &TV,
                                                                   bool free_transients,
                                                                   const char *timer) {
    tg_mesh *m = mesh_for(TV);
    build(m, REL);
    Stopwatch sw(timer);
    const HostRel h = fetch(m, REL);
    sw.lap(kTransfer);
    auto out = std::make_unique<std::vector<std::vector<vtkIdType>>>(to_rows(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}
}  // namespace

std::unique_ptr<ET_Data> make_ET_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_ET>(TV, f, "ET_GPU");
}
std::unique_ptr<EF_Data> make_EF_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_EF>(TV, f, "EF_GPU");
}
std::unique_ptr<TT_Data> make_TT_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_TT>(TV, f, "TT_GPU");
}
std::unique_ptr<FF_Data> make_FF_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_FF>(TV, f, "FF_GPU");
}
std::unique_ptr<EE_Data> make_EE_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_EE>(TV, f, "EE_GPU");
}
std::unique_ptr<VT_Data> make_VT_GPU(const TV_Data &TV, const bool f) {
    return csr_relation<TG_REL_VT>(TV, f, "VT_GPU");
}

std::unique_ptr<FT_Data> make_FT_GPU(const TV_Data &TV, const bool free_transients) {
    tg_mesh *m = mesh_for(TV);
    build(m, TG_REL_FT);
    Stopwatch sw("FT_GPU");
    const HostRel h = fetch(m, TG_REL_FT);
    sw.lap(kTransfer);
    auto out = std::make_unique<FT_Data>(to_fixed<2>(h));
    sw.lap(kTranslation);
    finish(free_transients);
    return out;
}

void set_host_timing_report(bool on) { g_report_times.store(on, std::memory_order_relaxed); }

std::vector<HostTiming> take_host_timings() {
    std::vector<HostTiming> out;
    out.swap(g_times);
    return out;
}
