// Relation dispatch behind tg_build / tg_relation_*: buffer management of the
// mesh handle, the VV + critical point hot path, VT, the tuning table, and
// (in ids.hip) the id-assigning and derived relations.
#include <algorithm>
#include <mutex>
#include <vector>
#include <cstdio>

#include "ids.h"
#include "relations.h"
#include "tetra_c.h"

namespace {

__global__ void k_corner_to_tet(const int32_t *__restrict__ vtc, int64_t n,
                                int32_t *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = vtc[i] >> 2;
}

int fail(int code, const char *msg) {
    tg::set_error(msg);
    return code;
}

// Read one int32 from device memory (synchronises the stream).
int read_i32(tg_mesh &m, const void *src, int64_t index, int32_t &out) {
    int32_t *h = (int32_t *)m.h_pinned;
    TG_HIP(hipMemcpyAsync(h, (const int32_t *)src + index, sizeof(int32_t),
                          hipMemcpyDeviceToHost, m.stream));
    TG_HIP(hipStreamSynchronize(m.stream));
    out = *h;
    return tg::OK;
}

}  // namespace

static void release_all_impl(tg_mesh *m, hipStream_t s, bool release);

void tg_mesh::bind_stream(hipStream_t s) {
    release_all_impl(this, s, false);
}

void tg_mesh::release_all() { release_all_impl(this, stream, true); }

static void release_all_impl(tg_mesh *m, hipStream_t s, bool release) {
    tg_mesh &me = *m;
    tg::DevArr *all[] = {&me.tv,       &me.scal,     &me.vt_cnt,   &me.vt_off,   &me.vt_cur,   &me.vtc,
                         &me.scratch,  &me.deg,      &me.vv_off,   &me.vv_cols,  &me.classes,  &me.heavy,
                         &me.fix,      &me.fixbig,   &me.counters, &me.scan_tmp, &me.labels,   &me.vt_vals,
                         &me.edge_first, &me.ve_off, &me.ve_vals,  &me.ev,       &me.te,       &me.vv_eid,
                         &me.et_off,   &me.et_vals,  &me.face_off, &me.vf_vals,  &me.fv,       &me.tf,
                         &me.fe,       &me.ft,       &me.ef_off,   &me.ef_vals,  &me.tt_off,   &me.tt_vals,
                         &me.ff_off,   &me.ff_vals,  &me.ee_off,   &me.ee_vals,  &me.bitmap,   &me.rank_tmp,
                         &me.tmp_a,    &me.tmp_b,    &me.tmp_c,    &me.bk_h1,    &me.bk_h2,    &me.bk_rowsrc,
                         &me.bk_fixd,  &me.bk_rowdesc, &me.bk_rec};
    for (auto *a : all) {
        a->stream = s;
        if (release) a->release();
    }
    if (!release) return;
    for (auto &e : me.marks.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

namespace tg {

VvCpBuffers hot_buffers(tg_mesh &m) {
    VvCpBuffers b;
    b.tv = m.tv.as<int32_t>();
    b.scalars = m.scal.p;
    b.scalars_f64 = m.scal_f64;
    b.n_cells = m.n_cells;
    b.n_points = m.n_points;
    b.vt_cnt = m.vt_cnt.as<int32_t>();
    b.vt_off = m.vt_off.as<int32_t>();
    b.vt_cur = m.vt_cur.as<int32_t>();
    b.vtc = m.vtc.as<int32_t>();
    b.scratch = m.scratch.as<int32_t>();
    b.deg = m.deg.as<int32_t>();
    b.vv_off = m.vv_off.as<int32_t>();
    b.vv_cols = m.vv_cols.as<int32_t>();
    b.classes = m.classes.as<uint32_t>();
    b.heavy = m.heavy.as<int32_t>();
    b.fix = m.fix.as<int32_t>();
    b.fixbig = m.fixbig.as<int32_t>();
    b.counters = m.counters.as<int32_t>();
    b.scan_tmp = m.scan_tmp.p;
    b.marks = &m.marks;
    return b;
}

void resolve_tuning(tg_mesh &m);

int prepare_hot_path(tg_mesh &m) {
    if (m.hot_ready) return OK;
    const size_t V = (size_t)m.n_points, T = (size_t)m.n_cells, I = sizeof(int32_t);
    int rc;
    if ((rc = m.scratch.ensure(I * 12 * T)) || (rc = m.deg.ensure(I * V)) ||
        (rc = m.vv_off.ensure(I * (V + 1))) || (rc = m.vv_cols.ensure(I * (12 * T + 4))) ||
        (rc = m.classes.ensure(I * 3 * V)) || (rc = m.heavy.ensure(I * V)) ||
        (rc = m.fix.ensure(I * V)) || (rc = m.fixbig.ensure(I * V)) ||
        (rc = m.counters.ensure(I * N_COUNTERS)) ||
        (rc = m.scan_tmp.ensure(scan_tmp_bytes(std::max<int64_t>(m.n_points, 12 * m.n_cells)))))
        return rc;
    resolve_tuning(m);
    m.bplan = plan_buckets(m.n_cells, m.n_points, m.cfg.bv_log, m.cfg.tpc);
    m.bk_route = m.cfg.route;
    if (m.bk_route == ROUTE_AUTO && m.span_tiles >= 0)  // decided from the creation-time spans
        m.bk_route = (int64_t)m.span_wide[m.cfg.bv_log == 8 ? 1 : 0] * 64 <= m.span_tiles
                         ? ROUTE_ONE
                         : ROUTE_TWO;
    m.hot_ready = true;
    return OK;
}

// ---- tuning table (process-wide, set by the autotuner) -----------------------------
namespace {
std::mutex g_tune_mu;
std::vector<tg_tune_entry> g_tune;  // ascending max_cells
}  // namespace

int set_tuning(const tg_tune_entry *e, int32_t n) {
    std::vector<tg_tune_entry> t(e, e + n);
    for (const auto &x : t)
        if (x.variant < 1 || x.variant > 4 || (x.lane_cap != 32 && x.lane_cap != 64) ||
            (x.lane_block != 64 && x.lane_block != 128 && x.lane_block != 256) ||
            (x.lane_cap == 64 && x.lane_block == 256) || (x.bv_log != 7 && x.bv_log != 8) ||
            x.tets_per_chunk < 1024)
            return fail(ERR_ARG, "invalid tuning entry");
    std::sort(t.begin(), t.end(),
              [](const tg_tune_entry &a, const tg_tune_entry &b) { return a.max_cells < b.max_cells; });
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune = std::move(t);
    return OK;
}

int get_tuning(tg_tune_entry *e, int32_t cap, int32_t *n) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    *n = (int32_t)g_tune.size();
    for (int32_t i = 0; i < cap && i < *n; i++) e[i] = g_tune[i];
    return OK;
}

// variant 0: the first table entry whose max_cells covers the mesh
void resolve_tuning(tg_mesh &m) {
    if (m.cfg.variant != 0 && !m.cfg.tuned) return;
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const auto &x : g_tune)
        if (m.n_cells <= x.max_cells) {
            m.cfg.variant = x.variant;
            m.cfg.lane_cap = x.lane_cap;
            m.cfg.lane_block = x.lane_block;
            m.cfg.bv_log = x.bv_log;
            m.cfg.tpc = x.tets_per_chunk;
            m.cfg.tuned = true;
            return;
        }
}

int prepare_vt(tg_mesh &m) {
    if (m.vt_ready) return OK;
    int rc = prepare_hot_path(m);
    if (rc) return rc;
    const size_t V = (size_t)m.n_points, T = (size_t)m.n_cells, I = sizeof(int32_t);
    // vtc carries four words of slack: row staging reads whole 16-byte quads
    if ((rc = m.vt_cnt.ensure(I * V)) || (rc = m.vt_off.ensure(I * (V + 1))) ||
        (rc = m.vt_cur.ensure(I * V)) || (rc = m.vtc.ensure(I * (4 * T + 4))))
        return rc;
    m.vt_ready = true;
    return OK;
}

int prepare_buckets(tg_mesh &m) {
    if (m.bk_ready) return OK;
    int rc = prepare_hot_path(m);
    if (rc) return rc;
    const BucketPlan &p = m.bplan;
    const size_t I = sizeof(int32_t);
    const size_t n1 = std::max((size_t)p.nc * (size_t)p.nchunks, (size_t)p.nb) + 1,
                 n2 = (size_t)p.nb * p.slices + 1;
    if ((rc = m.bk_h1.ensure(I * n1)) || (rc = m.bk_h2.ensure(I * n2)) ||
        (rc = m.bk_rowsrc.ensure(I * (size_t)m.n_points)) ||
        (rc = m.bk_fixd.ensure(I * (size_t)m.n_points)) ||
        (rc = m.bk_rowdesc.ensure(12 * (size_t)m.n_cells)) ||
        (rc = m.scan_tmp.ensure(scan_tmp_bytes((int64_t)std::max(
             {n1, n2, (size_t)m.n_points, (size_t)(12 * m.n_cells)})))))
        return rc;
    m.bk_ready = true;
    return OK;
}

int hot_variant(const tg_mesh &m) {
    if (m.cfg.variant == 1 || !m.bplan.ok) return 1;
    return m.cfg.variant == 3 || m.cfg.variant == 4 ? m.cfg.variant : 2;
}

static BucketBuffers bucket_buffers(tg_mesh &m) {
    BucketBuffers k;
    k.base = hot_buffers(m);
    k.plan = m.bplan;
    k.route = &m.bk_route;
    k.h1 = m.bk_h1.as<int32_t>();
    k.h2 = m.bk_h2.as<int32_t>();
    k.rec1 = m.scratch.as<uint32_t>();
    k.rec2 = m.vv_cols.as<uint32_t>();
    k.rowsrc = m.bk_rowsrc.as<int32_t>();
    k.base.rowsrc = k.rowsrc;
    k.base.fixd = m.bk_fixd.as<int32_t>();
    k.base.rowdesc = m.bk_rowdesc.as<uint8_t>();
    return k;
}

// Enqueue VV (and classes when classify) in the mesh's chosen form.
static int run_vv(tg_mesh &m, bool classify) {
    int rc = prepare_hot_path(m);
    if (rc) return rc;
    if (hot_variant(m) >= 2) {
        if ((rc = prepare_buckets(m))) return rc;
        BucketBuffers k = bucket_buffers(m);
        k.plan.owner = hot_variant(m) == 3;
        k.plan.ins = hot_variant(m) == 4 ? 1 : 0;
        if ((rc = run_vv_cp_bucketed(k, classify, m.stream))) return rc;
    } else {
        if ((rc = prepare_vt(m))) return rc;
        VvCpBuffers b = hot_buffers(m);
        if (m.vt_corners_built) {
            // the corner lists are already there (rows in any order serve the
            // lane kernel): keep them as they are, sorted rows included, so
            // relations built before and after this run read one state
            if ((rc = run_vv_cp_on_corners(b, m.cfg, classify, m.stream))) return rc;
        } else {
            if ((rc = run_vv_cp(b, m.cfg, classify, m.stream))) return rc;
            m.vt_corners_built = true;
            m.vt_sorted = false;
        }
    }
    m.vv_built = true;
    if (classify) m.cp_built = true;
    m.counters_pending = true;
    return OK;
}

int run_hot_path(tg_mesh &m) { return run_vv(m, true); }

// VT through the bucketed routing when the mesh fits it: rows come out
// ascending, with no global atomics and no row sort.  Level-1 records use
// the hot path's row scratch (dead between runs), level-2 records a buffer
// of their own that is released afterwards.
static int vt_bucketed(tg_mesh &m, bool &done) {
    done = false;
    int rc = prepare_hot_path(m);
    if (rc) return rc;
    if (!m.bplan.ok || 4 * m.n_cells > (int64_t)INT32_MAX) return OK;
    if ((rc = prepare_buckets(m))) return rc;
    if ((rc = m.bk_rec.ensure(12 * sizeof(int32_t) * (size_t)std::max<int64_t>(m.n_cells, 1))))
        return rc;
    BucketBuffers k = bucket_buffers(m);
    k.rec2 = m.bk_rec.as<uint32_t>();
    if ((rc = run_vt_bucketed(k, m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(),
                              m.vt_cnt.as<int32_t>(), m.stream)))
        return rc;
    m.bk_rec.release();
    done = true;
    return OK;
}

int ensure_vt_corners(tg_mesh &m) {
    if (m.vt_corners_built) return OK;
    int rc = prepare_vt(m);
    if (rc) return rc;
    bool done = false;
    if ((rc = vt_bucketed(m, done))) return rc;
    if (done) {
        m.vt_corners_built = true;
        m.vt_sorted = true;
        return OK;
    }
    VvCpBuffers b = hot_buffers(m);
    if ((rc = launch_vt_corners(b, m.cfg.corner_block, m.stream))) return rc;
    m.vt_corners_built = true;
    m.vt_sorted = false;
    return OK;
}

int ensure_vt_sorted(tg_mesh &m) {
    int rc = ensure_vt_corners(m);
    if (rc) return rc;
    if (m.vt_sorted) return OK;
    if ((rc = launch_sort_rows_i32(m.vt_off.as<int32_t>(), m.vtc.as<int32_t>(), m.n_points,
                                   m.stream)))
        return rc;
    m.vt_sorted = true;
    return OK;
}

int ensure_vv(tg_mesh &m) {
    if (m.vv_built) return OK;
    return run_vv(m, false);
}

static int ensure_vt(tg_mesh &m) {
    if (m.vt_built) return OK;
    int rc = ensure_vt_sorted(m);
    if (rc) return rc;
    const int64_t n = 4 * m.n_cells;
    if ((rc = m.vt_vals.ensure(sizeof(int32_t) * (size_t)n))) return rc;
    if (n > 0) {
        unsigned g = std::min<unsigned>(grid_for(n, 256), 65536u);
        hipLaunchKernelGGL(k_corner_to_tet, dim3(g), dim3(256), 0, m.stream,
                           m.vtc.as<int32_t>(), n, m.vt_vals.as<int32_t>());
        TG_LAUNCH_CHECK("k_corner_to_tet");
    }
    m.vt_built = true;
    return OK;
}

int build_relations(tg_mesh &m, uint32_t flags) {
    int rc = OK;
    auto want = [&](int rel) { return (flags >> rel) & 1u; };
    if (want(TG_REL_VV) && (rc = ensure_vv(m))) return rc;
    if (want(TG_REL_VT) && (rc = ensure_vt(m))) return rc;
    // id-assigning relations and those derived from them
    const bool need_edges = want(TG_REL_TE) || want(TG_REL_EV) || want(TG_REL_ET) ||
                            want(TG_REL_FE) || want(TG_REL_EF) || want(TG_REL_EE) ||
                            want(TG_REL_VE);
    const bool need_faces = want(TG_REL_TF) || want(TG_REL_FV) || want(TG_REL_FE) ||
                            want(TG_REL_FT) || want(TG_REL_EF) || want(TG_REL_FF) ||
                            want(TG_REL_VF);
    if (need_edges && (rc = ensure_edges(m))) return rc;
    if (need_faces && (rc = ensure_faces(m))) return rc;
    if (want(TG_REL_ET) && (rc = ensure_et(m))) return rc;
    if ((want(TG_REL_FE) || want(TG_REL_EF)) && (rc = ensure_fe(m))) return rc;
    if ((want(TG_REL_FT) || want(TG_REL_FF)) && (rc = ensure_ft(m))) return rc;
    if (want(TG_REL_EF) && (rc = ensure_ef(m))) return rc;
    if (want(TG_REL_TT) && (rc = ensure_tt(m))) return rc;
    if (want(TG_REL_FF) && (rc = ensure_ff(m))) return rc;
    if (want(TG_REL_EE) && (rc = ensure_ee(m))) return rc;
    return OK;
}

int sync_and_check(tg_mesh &m) {
    TG_HIP(hipStreamSynchronize(m.stream));
    if (!m.counters_pending) return OK;
    m.counters_pending = false;
    int32_t err = 0;
    int rc = read_i32(m, m.counters.p, C_ERR, err);
    if (rc) return rc;
    if (err & ERRBIT_CAPACITY) {
        char buf[160];
        std::snprintf(buf, sizeof buf,
                      "a vertex has more than %d distinct neighbours; its row and class "
                      "were not computed",
                      HEAVY_CAP);
        set_error(buf);
        return ERR_CAPACITY;
    }
    return OK;
}

int relation_view(tg_mesh &m, int32_t rel, RelView &r) {
    int rc = sync_and_check(m);
    if (rc) return rc;
    r = RelView();
    int32_t nnz = 0;
    switch (rel) {
    case TG_REL_VV:
        if (!m.vv_built) return fail(ERR_STATE, "VV not built");
        if ((rc = read_i32(m, m.vv_off.p, m.n_points, nnz))) return rc;
        r.rows = m.n_points;
        r.nnz = nnz;
        r.offsets = m.vv_off.p;
        r.offset_bytes = 4;
        r.values = m.vv_cols.as<int32_t>();
        return OK;
    case TG_REL_VT:
        if (!m.vt_built) return fail(ERR_STATE, "VT not built");
        r.rows = m.n_points;
        r.nnz = 4 * m.n_cells;
        r.offsets = m.vt_off.p;
        r.offset_bytes = 4;
        r.values = m.vt_vals.as<int32_t>();
        return OK;
    default:
        return id_relation_view(m, rel, r);
    }
}

}  // namespace tg
