// C ABI of libtetra_hip.so (include/tetra_c.h): the mesh handle, device
// buffers it owns, and the dispatch of relation builds.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "pipeline.h"
#include "relations.h"
#include "tetra_c.h"

namespace tg {

int noalloc_init_device();  // noalloc.hip

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e),
                  file, line);
    set_error(buf);
    return ERR_HIP;
}

int DevArr::ensure(size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (p && cap >= bytes) return OK;
    release();
    size_t granted = 0;
    p = pool_alloc(bytes, &granted);
    if (!p) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "device allocation of %zu bytes failed", bytes);
        set_error(buf);
        return ERR_ALLOC;
    }
    cap = granted;
    return OK;
}

void DevArr::release() {
    if (p) pool_free(p, cap, stream);
    p = nullptr;
    cap = 0;
}

}  // namespace tg

namespace {

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ in, int64_t n, B *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (B)in[i];
}

// Connectivity check (err[0] |= 1 for an id outside [0, V) or a vertex
// repeated in a cell) and, per tile of 1024 consecutive cells, whether its
// vertices span 1024 or more buckets of 2^7 (err[1] += 1) or 2^8 (err[2] += 1)
// vertices: the input to the bucketed form's routing choice.
constexpr int VAL_TILE = 1024, VAL_THREADS = 256;
__global__ __launch_bounds__(VAL_THREADS) void k_validate(const int4 *__restrict__ tv4, int64_t T,
                                                          int64_t V, int32_t *__restrict__ err) {
    __shared__ int32_t red[2][VAL_THREADS / 64];
    const int64_t ntiles = (T + VAL_TILE - 1) / VAL_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int32_t lo = INT32_MAX, hi = -1;
        bool bad = false;
        for (int k = 0; k < VAL_TILE / VAL_THREADS; k++) {
            const int64_t t = tile * VAL_TILE + k * VAL_THREADS + threadIdx.x;
            if (t >= T) break;
            const int4 q = tv4[t];
            const bool range = q.x >= 0 && q.y >= 0 && q.z >= 0 && q.w >= 0 && q.x < V &&
                               q.y < V && q.z < V && q.w < V;
            const bool distinct = q.x != q.y && q.x != q.z && q.x != q.w && q.y != q.z &&
                                  q.y != q.w && q.z != q.w;
            bad |= !(range && distinct);
            lo = min(lo, min(min(q.x, q.y), min(q.z, q.w)));
            hi = max(hi, max(max(q.x, q.y), max(q.z, q.w)));
        }
        if (bad) atomicOr(err, 1);
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
        if (threadIdx.x == 0) {
            for (int w = 1; w < VAL_THREADS / 64; w++) {
                lo = min(lo, red[0][w]);
                hi = max(hi, red[1][w]);
            }
            if (lo >= 0 && hi >= lo) {
                if ((hi >> 7) - (lo >> 7) >= 1024) atomicAdd(err + 1, 1);
                if ((hi >> 8) - (lo >> 8) >= 1024) atomicAdd(err + 2, 1);
            }
        }
        __syncthreads();  // red is reused by the next tile
    }
}

__global__ void k_label6(const uint32_t *__restrict__ classes, int64_t V,
                         uint8_t *__restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t l = classes[3 * v], u = classes[3 * v + 1];
        uint8_t k;
        if (l == 0 && u >= 1) k = 0;
        else if (l >= 1 && u == 0) k = 3;
        else if (l == 1 && u == 1) k = 5;
        else if (l == 2 && u == 1) k = 1;
        else if (l == 1 && u == 2) k = 2;
        else k = 4;
        out[v] = k;
    }
}

__global__ void k_max_i32(const int32_t *__restrict__ a, int64_t n,
                          int32_t *__restrict__ out) {
    int32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, a[i]);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if (__lane_id() == 0) atomicMax(out, m);
}

int fail(int code, const std::string &msg) {
    tg::set_error(msg);
    return code;
}

}  // namespace

using tg::DevArr;

extern "C" {

const char *tg_last_error(void) { return tg::last_error(); }

int tg_device_count(int *count) {
    if (!count) return fail(TG_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return fail(TG_ERR_NO_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    }
    *count = n;
    return TG_OK;
}

int tg_init(int device) {
    int n = 0;
    int rc = tg_device_count(&n);
    if (rc) return rc;
    if (n == 0) return fail(TG_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n) return fail(TG_ERR_ARG, "device index out of range");
    TG_HIP(hipSetDevice(device));
    TG_HIP(hipFree(nullptr));
    return tg::noalloc_init_device();
}

int tg_mesh_create(tg_mesh **out, const int32_t *cells, int64_t n_cells,
                   int64_t n_points, const void *scalars, int scalar_f64,
                   int inputs_on_device, void *stream) {
    if (!out) return fail(TG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_cells < 0 || n_points < 0 || (n_cells > 0 && !cells))
        return fail(TG_ERR_ARG, "bad mesh sizes or NULL cells");
    if (4 * n_cells >= (int64_t)INT32_MAX / 3 || n_points >= (int64_t)INT32_MAX)
        return fail(TG_ERR_ARG, "mesh too large for 32-bit ids (needs 12*cells < 2^31)");
    int ndev = 0;
    int rc = tg_device_count(&ndev);
    if (rc) return rc;
    if (ndev == 0) return fail(TG_ERR_NO_DEVICE, "no HIP device visible");
    auto *m = new tg_mesh();
    TG_HIP(hipGetDevice(&m->device));
    m->n_cells = n_cells;
    m->n_points = n_points;
    if (stream) {
        m->stream = (hipStream_t)stream;
        m->own_stream = false;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete m;
            return tg::hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
        }
        m->own_stream = true;
    }
    m->bind_stream(m->stream);
    auto bail = [&](int code) {
        tg_mesh_destroy(m);
        return code;
    };
    if ((rc = m->tv.ensure(sizeof(int32_t) * 4 * (size_t)n_cells))) return bail(rc);
    if ((rc = m->counters.ensure(sizeof(int32_t) * tg::N_COUNTERS))) return bail(rc);
    if (!(m->h_pinned = tg::pinned_slot_alloc()))
        return bail(fail(TG_ERR_ALLOC, "pinned staging allocation failed"));
    if (n_cells > 0) {
        hipError_t e = hipMemcpyAsync(m->tv.p, cells, sizeof(int32_t) * 4 * (size_t)n_cells,
                                      inputs_on_device ? hipMemcpyDeviceToDevice
                                                       : hipMemcpyHostToDevice,
                                      m->stream);
        if (e != hipSuccess) return bail(tg::hip_fail(e, "hipMemcpyAsync(cells)", __FILE__, __LINE__));
        // validate connectivity once
        int32_t *err = m->counters.as<int32_t>();
        if ((rc = tg::fill_i32(err, 0, 3, m->stream))) return bail(rc);
        const int64_t tiles = (n_cells + VAL_TILE - 1) / VAL_TILE;
        const unsigned g = (unsigned)std::min<int64_t>(tiles, 16384);
        hipLaunchKernelGGL(k_validate, dim3(g), dim3(VAL_THREADS), 0, m->stream,
                           (const int4 *)m->tv.p, n_cells, n_points, err);
        if (hipGetLastError() != hipSuccess) return bail(fail(TG_ERR_HIP, "k_validate launch"));
        int32_t hv[3] = {0, 0, 0};
        if (hipMemcpyAsync(hv, err, sizeof hv, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess)
            return bail(fail(TG_ERR_HIP, "validation readback failed"));
        m->span_tiles = tiles;
        m->span_wide[0] = hv[1];
        m->span_wide[1] = hv[2];
        const int32_t herr = hv[0];
        if (herr) return bail(fail(TG_ERR_MESH, "cells hold a vertex id outside [0, n_points) "
                                                 "or a vertex repeated within a cell"));
    }
    if (scalars) {
        if ((rc = tg_mesh_set_scalars(m, scalars, scalar_f64, inputs_on_device))) return bail(rc);
    }
    *out = m;
    return TG_OK;
}

int tg_mesh_destroy(tg_mesh *m) {
    if (!m) return TG_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->release_all();
    if (m->h_pinned) tg::pinned_slot_free(m->h_pinned);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return TG_OK;
}

int tg_mesh_set_scalars(tg_mesh *m, const void *scalars, int scalar_f64, int on_device) {
    if (!m || !scalars) return fail(TG_ERR_ARG, "NULL mesh or scalars");
    const size_t bytes = (scalar_f64 ? 8 : 4) * (size_t)m->n_points;
    int rc = m->scal.ensure(bytes);
    if (rc) return rc;
    if (bytes)
        TG_HIP(hipMemcpyAsync(m->scal.p, scalars, bytes,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              m->stream));
    m->scal_f64 = scalar_f64 != 0;
    m->has_scalars = true;
    m->cp_built = false;
    return TG_OK;
}

int tg_mesh_stream(tg_mesh *m, void **stream) {
    if (!m || !stream) return fail(TG_ERR_ARG, "NULL argument");
    *stream = (void *)m->stream;
    return TG_OK;
}

int tg_mesh_set_config(tg_mesh *m, int32_t variant, int32_t lane_cap, int32_t lane_block) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    if (variant < 0 || variant > 4) return fail(TG_ERR_ARG, "variant must be 0 to 4");
    if (lane_cap != 32 && lane_cap != 64) return fail(TG_ERR_ARG, "lane_cap must be 32 or 64");
    if (lane_block != 64 && lane_block != 128 && lane_block != 256)
        return fail(TG_ERR_ARG, "lane_block must be 64, 128 or 256");
    if (lane_cap == 64 && lane_block == 256)
        return fail(TG_ERR_ARG, "lane_cap 64 supports lane_block 64 or 128");
    m->cfg.variant = variant;
    m->cfg.lane_cap = lane_cap;
    m->cfg.lane_block = lane_block;
    m->cfg.tuned = false;
    m->hot_ready = false;
    m->bk_ready = false;
    return TG_OK;
}

int tg_mesh_set_bucket(tg_mesh *m, int32_t bv_log, int64_t tets_per_chunk) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    if (bv_log != 7 && bv_log != 8) return fail(TG_ERR_ARG, "bv_log must be 7 or 8");
    if (tets_per_chunk < 1024) return fail(TG_ERR_ARG, "tets_per_chunk must be >= 1024");
    m->cfg.bv_log = bv_log;
    m->cfg.tpc = tets_per_chunk;
    m->hot_ready = false;  // re-plan the buckets on the next run
    m->bk_ready = false;
    return TG_OK;
}

int tg_mesh_set_routing(tg_mesh *m, int32_t route) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    if (route < tg::ROUTE_AUTO || route > tg::ROUTE_TWO)
        return fail(TG_ERR_ARG, "route must be 0 (auto), 1 (one level) or 2 (two levels)");
    m->cfg.route = route;
    m->hot_ready = false;  // re-plan (and re-decide in auto mode) on the next run
    m->bk_ready = false;
    return TG_OK;
}

int tg_mesh_routing(tg_mesh *m, int32_t *route) {
    if (!m || !route) return fail(TG_ERR_ARG, "NULL argument");
    *route = m->hot_ready ? m->bk_route : m->cfg.route;
    return TG_OK;
}

int tg_set_tuning(const tg_tune_entry *entries, int32_t n) {
    if (n < 0 || (n > 0 && !entries)) return fail(TG_ERR_ARG, "bad tuning table");
    return tg::set_tuning(entries, n);
}

int tg_get_tuning(tg_tune_entry *entries, int32_t capacity, int32_t *n) {
    if (!n || (capacity > 0 && !entries)) return fail(TG_ERR_ARG, "NULL argument");
    return tg::get_tuning(entries, capacity, n);
}

int tg_build(tg_mesh *m, uint32_t flags) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    return tg::build_relations(*m, flags);
}

int tg_critpoints(tg_mesh *m) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    if (!m->has_scalars) return fail(TG_ERR_STATE, "mesh has no scalar field");
    if (m->cp_built) return TG_OK;
    return tg::run_hot_path(*m);
}

int tg_vv_critpoints(tg_mesh *m) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    if (!m->has_scalars) return fail(TG_ERR_STATE, "mesh has no scalar field");
    return tg::run_hot_path(*m);
}

int tg_sync(tg_mesh *m) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    return tg::sync_and_check(*m);
}

int tg_mesh_counts(tg_mesh *m, int64_t *n_points, int64_t *n_cells, int64_t *n_edges,
                   int64_t *n_faces) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    int rc = tg::sync_and_check(*m);
    if (rc) return rc;
    if (n_points) *n_points = m->n_points;
    if (n_cells) *n_cells = m->n_cells;
    if (n_edges) *n_edges = m->edges_built ? m->n_edges : -1;
    if (n_faces) *n_faces = m->faces_built ? m->n_faces : -1;
    return TG_OK;
}

int tg_relation_shape(tg_mesh *m, int32_t rel, int64_t *rows, int64_t *nnz,
                      int32_t *width, int32_t *fixed) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    tg::RelView r;
    int rc = tg::relation_view(*m, rel, r);
    if (rc) return rc;
    if (rows) *rows = r.rows;
    if (nnz) *nnz = r.nnz;
    if (width) *width = r.width;
    if (fixed) *fixed = r.offsets ? 0 : 1;
    return TG_OK;
}

int tg_relation_copy(tg_mesh *m, int32_t rel, int64_t *offsets, int32_t *values) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    tg::RelView r;
    int rc = tg::relation_view(*m, rel, r);
    if (rc) return rc;
    if (r.nnz * r.width > 0) {
        if (!values) return fail(TG_ERR_ARG, "values is NULL");
        TG_HIP(hipMemcpyAsync(values, r.values, sizeof(int32_t) * (size_t)(r.nnz * r.width),
                              hipMemcpyDeviceToHost, m->stream));
    }
    if (r.offsets && offsets) {
        if (r.offset_bytes == 8) {
            TG_HIP(hipMemcpyAsync(offsets, r.offsets, sizeof(int64_t) * (size_t)(r.rows + 1),
                                  hipMemcpyDeviceToHost, m->stream));
            TG_HIP(hipStreamSynchronize(m->stream));
        } else {
            int32_t *tmp = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(r.rows + 1));
            if (!tmp) return fail(TG_ERR_ALLOC, "host allocation failed");
            hipError_t e = hipMemcpyAsync(tmp, r.offsets, sizeof(int32_t) * (size_t)(r.rows + 1),
                                          hipMemcpyDeviceToHost, m->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
            if (e != hipSuccess) {
                std::free(tmp);
                return tg::hip_fail(e, "offset copy", __FILE__, __LINE__);
            }
            for (int64_t i = 0; i <= r.rows; i++) offsets[i] = tmp[i];
            std::free(tmp);
        }
    }
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

int tg_relation_device(tg_mesh *m, int32_t rel, const void **d_offsets, int32_t *offset_bytes,
                       const int32_t **d_values) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    tg::RelView r;
    int rc = tg::relation_view(*m, rel, r);
    if (rc) return rc;
    if (d_offsets) *d_offsets = r.offsets;
    if (offset_bytes) *offset_bytes = r.offsets ? r.offset_bytes : 0;
    if (d_values) *d_values = r.values;
    return TG_OK;
}

int tg_classes_device(tg_mesh *m, const uint32_t **d_classes3) {
    if (!m || !d_classes3) return fail(TG_ERR_ARG, "NULL argument");
    if (!m->cp_built) return fail(TG_ERR_STATE, "critical points not computed");
    *d_classes3 = (const uint32_t *)m->classes.p;
    return TG_OK;
}

int tg_classes_copy(tg_mesh *m, uint32_t *classes3) {
    if (!m || (!classes3 && m->n_points > 0)) return fail(TG_ERR_ARG, "NULL argument");
    if (!m->cp_built) return fail(TG_ERR_STATE, "critical points not computed");
    int rc = tg::sync_and_check(*m);
    if (rc) return rc;
    if (m->n_points > 0)
        TG_HIP(hipMemcpyAsync(classes3, m->classes.p, sizeof(uint32_t) * 3 * (size_t)m->n_points,
                              hipMemcpyDeviceToHost, m->stream));
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

int tg_labels6_copy(tg_mesh *m, uint8_t *labels) {
    if (!m || (!labels && m->n_points > 0)) return fail(TG_ERR_ARG, "NULL argument");
    if (!m->cp_built) return fail(TG_ERR_STATE, "critical points not computed");
    if (m->n_points == 0) return TG_OK;
    int rc = m->labels.ensure((size_t)m->n_points);
    if (rc) return rc;
    unsigned g = std::min<unsigned>(tg::grid_for(m->n_points, 256), 16384u);
    hipLaunchKernelGGL(k_label6, dim3(g), dim3(256), 0, m->stream,
                       (const uint32_t *)m->classes.p, m->n_points, (uint8_t *)m->labels.p);
    TG_LAUNCH_CHECK("k_label6");
    rc = tg::sync_and_check(*m);
    if (rc) return rc;
    TG_HIP(hipMemcpyAsync(labels, m->labels.p, (size_t)m->n_points, hipMemcpyDeviceToHost,
                          m->stream));
    TG_HIP(hipStreamSynchronize(m->stream));
    return TG_OK;
}

int tg_approx_max_vv(tg_mesh *m, int32_t *max_vv) {
    if (!m || !max_vv) return fail(TG_ERR_ARG, "NULL argument");
    int rc = tg::ensure_vt_corners(*m);
    if (rc) return rc;
    int32_t *d = m->counters.as<int32_t>() + tg::C_SCRATCH;
    if ((rc = tg::fill_i32(d, 0, 1, m->stream))) return rc;
    if (m->n_points > 0) {
        unsigned g = std::min<unsigned>(tg::grid_for(m->n_points, 256), 4096u);
        hipLaunchKernelGGL(k_max_i32, dim3(g), dim3(256), 0, m->stream,
                           (const int32_t *)m->vt_cnt.p, m->n_points, d);
        TG_LAUNCH_CHECK("k_max_i32");
    }
    int32_t mx = 0;
    TG_HIP(hipMemcpyAsync(&mx, d, sizeof mx, hipMemcpyDeviceToHost, m->stream));
    TG_HIP(hipStreamSynchronize(m->stream));
    mx *= 3;
    *max_vv = ((mx + 31) / 32) * 32;
    return TG_OK;
}

int tg_profile_enable(tg_mesh *m, int on) {
    if (!m) return fail(TG_ERR_ARG, "NULL mesh");
    m->marks.on = on != 0;
    m->marks.n = 0;
    return TG_OK;
}

int tg_profile_read(tg_mesh *m, int32_t max_stages, int32_t *n_stages, float *ms,
                    char *names) {
    if (!m || !n_stages) return fail(TG_ERR_ARG, "NULL argument");
    int rc = tg::sync_and_check(*m);
    if (rc) return rc;
    const int n = m->marks.n > 0 ? m->marks.n - 1 : 0;
    *n_stages = n;
    for (int i = 0; i < n && i < max_stages; i++) {
        float t = 0;
        TG_HIP(hipEventElapsedTime(&t, m->marks.ev[i], m->marks.ev[i + 1]));
        if (ms) ms[i] = t;
        if (names) {
            std::strncpy(names + 32 * i, m->marks.name[i + 1], 31);
            names[32 * i + 31] = 0;
        }
    }
    return TG_OK;
}

int tg_relation_export(tg_mesh *m, int32_t rel, int32_t offset_bytes, void **d_offsets,
                       int32_t value_bytes, void **d_values) {
    if (!m || !d_values) return fail(TG_ERR_ARG, "NULL mesh or output");
    if ((offset_bytes != 4 && offset_bytes != 8) || (value_bytes != 4 && value_bytes != 8))
        return fail(TG_ERR_ARG, "offset_bytes and value_bytes must be 4 or 8");
    *d_values = nullptr;
    if (d_offsets) *d_offsets = nullptr;
    tg::RelView r;
    int rc = tg::relation_view(*m, rel, r);
    if (rc) return rc;
    const int64_t nv = r.nnz * r.width;
    void *vals = nullptr, *offs = nullptr;
    auto release = [&](int code) {
        if (vals) (void)hipFree(vals);
        if (offs) (void)hipFree(offs);
        return code;
    };
    if (hipMalloc(&vals, (size_t)value_bytes * (size_t)std::max<int64_t>(nv, 1)) != hipSuccess)
        return release(fail(TG_ERR_ALLOC, "hipMalloc of relation values failed"));
    if (nv > 0) {
        const unsigned g = std::min<unsigned>(tg::grid_for(nv, 256), 65536u);
        if (value_bytes == 8)
            hipLaunchKernelGGL((k_convert<int32_t, int64_t>), dim3(g), dim3(256), 0, m->stream,
                               r.values, nv, (int64_t *)vals);
        else
            hipLaunchKernelGGL((k_convert<int32_t, int32_t>), dim3(g), dim3(256), 0, m->stream,
                               r.values, nv, (int32_t *)vals);
        if (hipGetLastError() != hipSuccess) return release(fail(TG_ERR_HIP, "export launch"));
    }
    if (d_offsets && r.offsets) {
        const int64_t n = r.rows + 1;
        if (hipMalloc(&offs, (size_t)offset_bytes * (size_t)n) != hipSuccess)
            return release(fail(TG_ERR_ALLOC, "hipMalloc of relation offsets failed"));
        const unsigned g = std::min<unsigned>(tg::grid_for(n, 256), 65536u);
        if (r.offset_bytes == 8 && offset_bytes == 8)
            hipLaunchKernelGGL((k_convert<int64_t, int64_t>), dim3(g), dim3(256), 0, m->stream,
                               (const int64_t *)r.offsets, n, (int64_t *)offs);
        else if (r.offset_bytes == 8)
            hipLaunchKernelGGL((k_convert<int64_t, uint32_t>), dim3(g), dim3(256), 0, m->stream,
                               (const int64_t *)r.offsets, n, (uint32_t *)offs);
        else if (offset_bytes == 8)
            hipLaunchKernelGGL((k_convert<int32_t, int64_t>), dim3(g), dim3(256), 0, m->stream,
                               (const int32_t *)r.offsets, n, (int64_t *)offs);
        else
            hipLaunchKernelGGL((k_convert<int32_t, uint32_t>), dim3(g), dim3(256), 0, m->stream,
                               (const int32_t *)r.offsets, n, (uint32_t *)offs);
        if (hipGetLastError() != hipSuccess) return release(fail(TG_ERR_HIP, "export launch"));
        if (r.offset_bytes == 8 && offset_bytes == 4 && r.nnz > (int64_t)UINT32_MAX)
            return release(fail(TG_ERR_CAPACITY, "relation too large for 32-bit offsets"));
    }
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return release(fail(TG_ERR_HIP, "export synchronisation failed"));
    *d_values = vals;
    if (d_offsets) *d_offsets = offs;
    return TG_OK;
}

int tg_device_upload(const void *host, size_t bytes, void **d_out) {
    if (!d_out || (bytes && !host)) return fail(TG_ERR_ARG, "NULL argument");
    *d_out = nullptr;
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
        return fail(TG_ERR_ALLOC, "hipMalloc failed");
    if (bytes && hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return fail(TG_ERR_HIP, "upload failed");
    }
    *d_out = p;
    return TG_OK;
}

int tg_host_alloc(size_t bytes, void **out) {
    if (!out) return fail(TG_ERR_ARG, "NULL argument");
    *out = tg::host_alloc(bytes);
    return *out ? TG_OK : fail(TG_ERR_ALLOC, "pinned host allocation failed");
}

int tg_host_free(void *p) {
    tg::host_free(p);
    return TG_OK;
}

int tg_pool_trim(void) {
    tg::pool_trim();
    return TG_OK;
}

int tg_pool_cached(int64_t *bytes) {
    if (!bytes) return fail(TG_ERR_ARG, "NULL argument");
    *bytes = (int64_t)tg::pool_cached_bytes();
    return TG_OK;
}

int tg_stream_create(void **stream) {
    if (!stream) return fail(TG_ERR_ARG, "NULL argument");
    hipStream_t s = nullptr;
    TG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return TG_OK;
}

int tg_stream_destroy(void *stream) {
    if (stream) TG_HIP(hipStreamDestroy((hipStream_t)stream));
    return TG_OK;
}

int tg_device_free(void *p) {
    if (p) TG_HIP(hipFree(p));
    return TG_OK;
}

}  // extern "C"
