// Many meshes, one launch sequence: the VV + critical-point path over a batch
// of independent meshes of uneven size (no reference counterpart: the
// reference runs one mesh per process, src/alg/critPoints.cu:1359-1514).
//
// VV rows and classes are vertex-local, so a disjoint union of meshes is
// itself a mesh whose results are the meshes' results side by side.  A batch
// is split into waves of consecutive meshes (each wave within the 32-bit
// index range of one pipeline); each wave is one union mesh built on the
// device from per-mesh cell and point offset tables, and runs the hot path
// as one launch sequence.  With graphs on, a wave's launch sequence is
// captured once into a hipGraph and replayed on every later call.
#include <algorithm>
#include <vector>

#include "relations.h"
#include "tetra_c.h"

namespace {

using tg::set_error;

int fail(int code, const char *msg) {
    set_error(msg);
    return code;
}

// cells of the wave, local ids -> union ids; a local id outside its mesh's
// vertex range raises err
__global__ void k_offset_cells(int32_t *__restrict__ cells, int64_t n_cells,
                               const int64_t *__restrict__ coff, const int64_t *__restrict__ poff,
                               int32_t n_meshes, int32_t *__restrict__ err) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_cells;
         c += (int64_t)gridDim.x * blockDim.x) {
        // mesh = last i with coff[i] <= c
        int32_t lo = 0, hi = n_meshes - 1;
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (coff[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const int64_t base = poff[lo], nv = poff[lo + 1] - poff[lo];
        int4 q = reinterpret_cast<int4 *>(cells)[c];
        const bool ok = q.x >= 0 && q.x < nv && q.y >= 0 && q.y < nv && q.z >= 0 && q.z < nv &&
                        q.w >= 0 && q.w < nv;
        if (!ok) *err = 1;  // every writer stores the same value
        q.x += (int32_t)base;
        q.y += (int32_t)base;
        q.z += (int32_t)base;
        q.w += (int32_t)base;
        reinterpret_cast<int4 *>(cells)[c] = q;
    }
}

struct Wave {
    int32_t first = 0, count = 0;  // meshes [first, first + count)
    int64_t cell0 = 0, point0 = 0, n_cells = 0, n_points = 0;
    tg_mesh *mesh = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool warmed = false;
};

}  // namespace

struct tg_batch {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int32_t n_meshes = 0;
    std::vector<int64_t> coff, poff;  // [n_meshes + 1] global cell / point offsets
    std::vector<Wave> waves;
    bool use_graph = false;
};

namespace {

int build_wave(tg_batch &b, Wave &w, const int32_t *cells, const void *scalars, int f64) {
    const int64_t nc = w.n_cells, np = w.n_points;
    int32_t *d_cells = nullptr;
    int64_t *d_coff = nullptr, *d_poff = nullptr;
    int32_t *d_err = nullptr;
    size_t g1 = 0, g2 = 0, g3 = 0, g4 = 0;
    auto release = [&](int code) {
        // the stream has been synchronised (or nothing was enqueued) by now
        tg::pool_free(d_cells, g1, b.stream);
        tg::pool_free(d_coff, g2, b.stream);
        tg::pool_free(d_poff, g3, b.stream);
        tg::pool_free(d_err, g4, b.stream);
        return code;
    };
    std::vector<int64_t> coff(w.count + 1), poff(w.count + 1);
    for (int32_t i = 0; i <= w.count; i++) {
        coff[i] = b.coff[w.first + i] - w.cell0;
        poff[i] = b.poff[w.first + i] - w.point0;
    }
    d_cells = (int32_t *)tg::pool_alloc(sizeof(int32_t) * 4 * (size_t)std::max<int64_t>(nc, 1), &g1);
    d_coff = (int64_t *)tg::pool_alloc(sizeof(int64_t) * coff.size(), &g2);
    d_poff = (int64_t *)tg::pool_alloc(sizeof(int64_t) * poff.size(), &g3);
    d_err = (int32_t *)tg::pool_alloc(sizeof(int32_t), &g4);
    if (!d_cells || !d_coff || !d_poff || !d_err)
        return release(fail(tg::ERR_ALLOC, "batch: device allocation failed"));
    hipError_t e = hipMemcpyAsync(d_cells, cells + 4 * w.cell0, sizeof(int32_t) * 4 * (size_t)nc,
                                  hipMemcpyHostToDevice, b.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_coff, coff.data(), sizeof(int64_t) * coff.size(),
                           hipMemcpyHostToDevice, b.stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_poff, poff.data(), sizeof(int64_t) * poff.size(),
                           hipMemcpyHostToDevice, b.stream);
    if (e != hipSuccess) return release(tg::hip_fail(e, "batch upload", __FILE__, __LINE__));
    if (int rc = tg::fill_i32(d_err, 0, 1, b.stream)) return release(rc);
    if (nc > 0) {
        const unsigned g = std::min<unsigned>(tg::grid_for(nc, 256), 65536u);
        hipLaunchKernelGGL(k_offset_cells, dim3(g), dim3(256), 0, b.stream, d_cells, nc, d_coff,
                           d_poff, w.count, d_err);
        if ((e = hipGetLastError()) != hipSuccess)
            return release(tg::hip_fail(e, "k_offset_cells", __FILE__, __LINE__));
    }
    int32_t herr = 0;
    e = hipMemcpyAsync(&herr, d_err, sizeof herr, hipMemcpyDeviceToHost, b.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b.stream);
    if (e != hipSuccess) return release(tg::hip_fail(e, "batch offsets", __FILE__, __LINE__));
    if (herr)
        return release(fail(tg::ERR_MESH, "batch: a cell references a vertex outside its mesh"));
    const size_t sb = f64 ? 8 : 4;
    const void *sc = scalars ? (const char *)scalars + sb * (size_t)w.point0 : nullptr;
    // the union mesh copies the offset cells from d_cells (device to device)
    int rc = tg_mesh_create(&w.mesh, d_cells, nc, np, nullptr, 0, 1, b.stream);
    if (rc == TG_OK && sc) rc = tg_mesh_set_scalars(w.mesh, sc, f64, 0);
    if (rc == TG_OK) {
        e = hipStreamSynchronize(b.stream);
        if (e != hipSuccess) rc = tg::hip_fail(e, "batch create", __FILE__, __LINE__);
    }
    return release(rc);
}

void destroy_waves(tg_batch &b) {
    for (Wave &w : b.waves) {
        if (w.exec) (void)hipGraphExecDestroy(w.exec);
        if (w.graph) (void)hipGraphDestroy(w.graph);
        if (w.mesh) tg_mesh_destroy(w.mesh);
        w = Wave();
    }
}

// One run of a wave.  With graphs on, the first run is eager (it allocates
// every buffer the sequence needs, so nothing allocates under capture), the
// second captures the launch sequence and launches the graph, later runs
// replay it.
int run_wave(tg_batch &b, Wave &w) {
    if (!b.use_graph) return tg_vv_critpoints(w.mesh);
    if (!w.warmed) {
        // first run eagerly: it allocates every buffer the sequence needs
        int rc = tg_vv_critpoints(w.mesh);
        if (rc) return rc;
        TG_HIP(hipStreamSynchronize(b.stream));
        w.warmed = true;
        return tg::OK;
    }
    if (!w.exec) {
        TG_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
        int rc = tg::run_hot_path(*w.mesh);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b.stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return tg::hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        w.graph = g;
        TG_HIP(hipGraphInstantiate(&w.exec, g, nullptr, nullptr, 0));
    }
    TG_HIP(hipGraphLaunch(w.exec, b.stream));
    return tg::OK;
}

}  // namespace

extern "C" {

int tg_batch_create(tg_batch **out, int32_t n_meshes, const int64_t *cell_offsets,
                    const int32_t *cells, const int64_t *point_offsets, const void *scalars,
                    int scalar_f64, int64_t max_wave_cells, void *stream) {
    if (!out) return fail(TG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_meshes <= 0 || !cell_offsets || !point_offsets)
        return fail(TG_ERR_ARG, "batch needs n_meshes > 0 and offset tables");
    if (cell_offsets[0] != 0 || point_offsets[0] != 0)
        return fail(TG_ERR_ARG, "offset tables must start at 0");
    for (int32_t i = 0; i < n_meshes; i++)
        if (cell_offsets[i + 1] < cell_offsets[i] || point_offsets[i + 1] < point_offsets[i])
            return fail(TG_ERR_ARG, "offset tables must be non-decreasing");
    if (cell_offsets[n_meshes] > 0 && !cells) return fail(TG_ERR_ARG, "cells is NULL");
    // one wave must fit the 32-bit index range of the pipeline (12 T < 2^31)
    const int64_t cap = (int64_t)(INT32_MAX / 3) / 4 - 1;
    if (max_wave_cells <= 0 || max_wave_cells > cap) max_wave_cells = std::min<int64_t>(cap, 64 << 20);
    for (int32_t i = 0; i < n_meshes; i++)
        if (cell_offsets[i + 1] - cell_offsets[i] > max_wave_cells)
            return fail(TG_ERR_ARG, "a mesh is larger than one wave");
    auto *b = new tg_batch();
    b->n_meshes = n_meshes;
    b->coff.assign(cell_offsets, cell_offsets + n_meshes + 1);
    b->poff.assign(point_offsets, point_offsets + n_meshes + 1);
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return fail(TG_ERR_HIP, "hipStreamCreate failed");
        }
        b->own_stream = true;
    }
    for (int32_t i = 0; i < n_meshes;) {
        Wave w;
        w.first = i;
        w.cell0 = b->coff[i];
        w.point0 = b->poff[i];
        while (i < n_meshes && b->coff[i + 1] - w.cell0 <= max_wave_cells &&
               b->poff[i + 1] - w.point0 < INT32_MAX)
            i++;
        w.count = i - w.first;
        w.n_cells = b->coff[i] - w.cell0;
        w.n_points = b->poff[i] - w.point0;
        b->waves.push_back(w);
        int rc = build_wave(*b, b->waves.back(), cells, scalars, scalar_f64);
        if (rc) {
            tg_batch_destroy(b);
            return rc;
        }
    }
    *out = b;
    return TG_OK;
}

int tg_batch_destroy(tg_batch *b) {
    if (!b) return TG_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_waves(*b);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TG_OK;
}

int tg_batch_counts(tg_batch *b, int32_t *n_waves, int64_t *n_cells, int64_t *n_points) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    if (n_waves) *n_waves = (int32_t)b->waves.size();
    if (n_cells) *n_cells = b->coff.back();
    if (n_points) *n_points = b->poff.back();
    return TG_OK;
}

int tg_batch_use_graph(tg_batch *b, int on) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    b->use_graph = on != 0;
    return TG_OK;
}

int tg_batch_set_config(tg_batch *b, int32_t variant, int32_t lane_cap, int32_t lane_block,
                        int32_t bv_log, int64_t tets_per_chunk) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    for (Wave &w : b->waves) {
        int rc = tg_mesh_set_config(w.mesh, variant, lane_cap, lane_block);
        if (!rc && variant >= 2) rc = tg_mesh_set_bucket(w.mesh, bv_log, tets_per_chunk);
        if (rc) return rc;
        if (w.exec) (void)hipGraphExecDestroy(w.exec);
        if (w.graph) (void)hipGraphDestroy(w.graph);
        w.exec = nullptr;
        w.graph = nullptr;
        w.warmed = false;
    }
    return TG_OK;
}

int tg_batch_vv_critpoints(tg_batch *b) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    for (Wave &w : b->waves) {
        int rc = run_wave(*b, w);
        if (rc) return rc;
    }
    return TG_OK;
}

int tg_batch_sync(tg_batch *b) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    for (Wave &w : b->waves) {
        int rc = tg_sync(w.mesh);
        if (rc) return rc;
    }
    return TG_OK;
}

int tg_batch_classes_copy(tg_batch *b, uint32_t *classes3) {
    if (!b || !classes3) return fail(TG_ERR_ARG, "NULL argument");
    for (Wave &w : b->waves) {
        int rc = tg_classes_copy(w.mesh, classes3 + 3 * w.point0);
        if (rc) return rc;
    }
    return TG_OK;
}

int tg_batch_build(tg_batch *b, uint32_t flags) {
    if (!b) return fail(TG_ERR_ARG, "NULL batch");
    if (flags & ~(TG_FLAG(TG_REL_VV) | TG_FLAG(TG_REL_VT) | TG_FLAG(TG_REL_TT)))
        return fail(TG_ERR_ARG, "batches build VV, VT and TT");
    for (Wave &w : b->waves) {
        int rc = tg_build(w.mesh, flags);
        if (rc) return rc;
    }
    return TG_OK;
}

int tg_batch_relation_copy(tg_batch *b, int32_t mesh, int32_t rel, int64_t *offsets,
                           int32_t *values, int64_t *nnz) {
    if (!b || mesh < 0 || mesh >= b->n_meshes) return fail(TG_ERR_ARG, "bad batch or mesh index");
    if (rel != TG_REL_VV && rel != TG_REL_VT && rel != TG_REL_TT)
        return fail(TG_ERR_ARG, "batches hand out VV, VT and TT");
    const Wave *w = nullptr;
    for (const Wave &x : b->waves)
        if (mesh >= x.first && mesh < x.first + x.count) w = &x;
    // rows: the mesh's vertices (VV, VT) or cells (TT); values: vertex ids
    // (VV) or cell ids (VT, TT), made local by subtracting the mesh's base
    const int64_t pb = b->poff[mesh] - w->point0, cb = b->coff[mesh] - w->cell0;
    const bool cell_rows = rel == TG_REL_TT;
    const int64_t r0 = cell_rows ? cb : pb;
    const int64_t r1 = cell_rows ? b->coff[mesh + 1] - w->cell0 : b->poff[mesh + 1] - w->point0;
    const int32_t base = (int32_t)(rel == TG_REL_VV ? pb : cb);
    const void *d_off = nullptr;
    const int32_t *d_val = nullptr;
    int32_t ob = 0;
    int rc = tg_relation_device(w->mesh, rel, &d_off, &ob, &d_val);
    if (rc) return rc;
    std::vector<int64_t> off((size_t)(r1 - r0 + 1));
    if (ob == 8) {
        TG_HIP(hipMemcpy(off.data(), (const int64_t *)d_off + r0, sizeof(int64_t) * off.size(),
                         hipMemcpyDeviceToHost));
    } else {
        std::vector<int32_t> o32(off.size());
        TG_HIP(hipMemcpy(o32.data(), (const int32_t *)d_off + r0, sizeof(int32_t) * o32.size(),
                         hipMemcpyDeviceToHost));
        for (size_t i = 0; i < off.size(); i++) off[i] = o32[i];
    }
    if (nnz) *nnz = off.back() - off.front();
    if (offsets)
        for (size_t i = 0; i < off.size(); i++) offsets[i] = off[i] - off[0];
    if (values && off.back() > off.front()) {
        const size_t n = (size_t)(off.back() - off.front());
        TG_HIP(hipMemcpy(values, d_val + off.front(), sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) values[i] -= base;
    }
    return TG_OK;
}

int tg_batch_vv_copy(tg_batch *b, int32_t mesh, int64_t *offsets, int32_t *values, int64_t *nnz) {
    if (!b || mesh < 0 || mesh >= b->n_meshes) return fail(TG_ERR_ARG, "bad batch or mesh index");
    const Wave *w = nullptr;
    for (const Wave &x : b->waves)
        if (mesh >= x.first && mesh < x.first + x.count) w = &x;
    const int64_t p0 = b->poff[mesh] - w->point0, p1 = b->poff[mesh + 1] - w->point0;
    const void *d_off = nullptr;
    const int32_t *d_val = nullptr;
    int32_t ob = 0;
    int rc = tg_relation_device(w->mesh, TG_REL_VV, &d_off, &ob, &d_val);
    if (rc) return rc;
    std::vector<int32_t> off((size_t)(p1 - p0 + 1));
    TG_HIP(hipMemcpy(off.data(), (const int32_t *)d_off + p0, sizeof(int32_t) * off.size(),
                     hipMemcpyDeviceToHost));
    if (nnz) *nnz = off.back() - off.front();
    if (offsets)
        for (size_t i = 0; i < off.size(); i++) offsets[i] = off[i] - off[0];
    if (values && off.back() > off.front()) {
        const size_t n = (size_t)(off.back() - off.front());
        TG_HIP(hipMemcpy(values, d_val + off.front(), sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) values[i] -= (int32_t)p0;
    }
    return TG_OK;
}

}  // extern "C"
