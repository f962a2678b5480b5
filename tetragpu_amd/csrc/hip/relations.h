// The mesh handle behind tg_mesh* and the relation builders it dispatches.
#pragma once

#include "bucket.h"
#include "tetra_c.h"

namespace tg {

// Grow-only device allocation owned by a mesh.
struct DevArr {
    void *p = nullptr;
    size_t cap = 0;
    hipStream_t stream = nullptr;  // the stream that uses it (orders its reuse)
    int ensure(size_t bytes);
    void release();
    template <typename T> T *as() const { return (T *)p; }
};

// A relation as stored on the device.  Fixed-width relations have
// offsets == nullptr and rows*width values; CSR relations have rows+1
// offsets (int32 or int64) and nnz*width values.
struct RelView {
    int64_t rows = 0, nnz = 0;
    int32_t width = 1;
    const void *offsets = nullptr;
    int32_t offset_bytes = 4;
    const int32_t *values = nullptr;
};

}  // namespace tg

struct tg_mesh {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int64_t n_cells = 0, n_points = 0;
    int64_t n_edges = 0, n_faces = 0;
    tg::DevArr tv, scal;
    bool scal_f64 = false, has_scalars = false;
    tg::VvCpConfig cfg;

    // VV + critical point pipeline
    tg::DevArr vt_cnt, vt_off, vt_cur, vtc, scratch, deg, vv_off, vv_cols, classes,
        heavy, fix, fixbig, counters, scan_tmp, labels;
    tg::DevArr bk_h1, bk_h2, bk_rowsrc, bk_fixd, bk_rowdesc;
    tg::DevArr bk_rec;  // level-2 records of the bucketed VT build (released after it)
    tg::BucketPlan bplan;
    int32_t bk_route = 0;  // tg::Route of the bucketed pipeline (AUTO until the first run)
    // measured at creation: tiles of 1024 cells, and those whose vertices span
    // >= 1024 buckets of 2^7 / 2^8 vertices (span_tiles < 0: not measured)
    int64_t span_tiles = -1;
    int32_t span_wide[2] = {0, 0};
    bool vt_ready = false, bk_ready = false;
    bool hot_ready = false;         // buffers sized for the hot path
    bool vt_corners_built = false;  // vt_off / vtc valid
    bool vt_sorted = false;         // vtc rows ascending
    bool vv_built = false, cp_built = false;
    bool counters_pending = false;  // a pipeline ran since the last error check

    // id-assigning and derived relations
    tg::DevArr vt_vals;
    tg::DevArr edge_first, ve_off, ve_vals, ev, te, vv_eid, et_off, et_vals;
    tg::DevArr face_off, vf_vals, fv, tf, fe, ft, ef_off, ef_vals;
    tg::DevArr tt_off, tt_vals, ff_off, ff_vals, ee_off, ee_vals;
    tg::DevArr bitmap, rank_tmp, tmp_a, tmp_b, tmp_c;
    bool edges_built = false, faces_built = false, vt_built = false, et_built = false,
         fe_built = false, ft_built = false, ef_built = false, tt_built = false,
         ff_built = false, ee_built = false;
    bool ft_ok = true;  // every face is held by at most two cells
    int64_t tt_nnz = 0, ff_nnz = 0, ee_nnz = 0, ef_nnz = 0;

    void *h_pinned = nullptr;  // small pinned staging area (64 int64)
    tg::StageMarks marks;      // per-stage events of the last pipeline run

    void release_all();
    void bind_stream(hipStream_t s);  // every DevArr of the mesh uses stream s
};

namespace tg {

// Allocate (once) the buffers shared by both VV forms, the vertex -> corner
// arrays, and the bucketed form's arrays.
int prepare_hot_path(tg_mesh &m);
int prepare_vt(tg_mesh &m);
int prepare_buckets(tg_mesh &m);
// Process-wide tuning table (tg_set_tuning / tg_get_tuning).
int set_tuning(const tg_tune_entry *e, int32_t n);
int get_tuning(tg_tune_entry *e, int32_t cap, int32_t *n);
// Variant the hot path runs for this mesh (1 vertex-lane, 2 bucketed).
int hot_variant(const tg_mesh &m);
// Enqueue VV + classes (recomputed from TV and scalars).
int run_hot_path(tg_mesh &m);
int ensure_vt_corners(tg_mesh &m);
int ensure_vv(tg_mesh &m);
// Enqueue the relations selected by `flags` (reference flag bits).
int build_relations(tg_mesh &m, uint32_t flags);
// Synchronise the mesh stream and report device-side errors.
int sync_and_check(tg_mesh &m);
// Describe a built relation (synchronises to learn its size).
int relation_view(tg_mesh &m, int32_t rel, RelView &out);
// Fill a VvCpBuffers view over the mesh's arrays.
VvCpBuffers hot_buffers(tg_mesh &m);

}  // namespace tg
