// Bucketed VV + critical point pipeline (vv_cp_bucket.hip).
#pragma once

#include "pipeline.h"

namespace tg {

struct BucketPlan {
    bool ok = false;      // mesh fits the bucketed form (V <= 2^24)
    int bv_log = 7;       // vertices per bucket = 2^bv_log (7 or 8)
    int32_t nb = 1;       // buckets
    int fs = 0;           // buckets per coarse bin = 2^fs
    int32_t nc = 1;       // coarse bins (<= 256)
    int32_t nchunks = 1;  // tet chunks of the first routing level
    int64_t tpc = 8192;   // tets per chunk
    int32_t slices = 1;   // slices per coarse bin in the second level
    bool owner = false;   // bucket kernel: owner form (records grouped by vertex)
    int ins = 0;          // bucket kernel insert: 0 batched snapshots, 1 lean lock-step
                          // (shared tables in both)
};

// Routing of the corner records to buckets.
enum Route : int {
    ROUTE_AUTO = 0,  // decide on the first run from the tiles' bucket spans
    ROUTE_ONE = 1,   // tiles of tets -> buckets through an LDS window
    ROUTE_TWO = 2,   // chunks -> coarse bins -> buckets
};

// Record arrays are 3 x 4T uint32 (structure of arrays).  rec1 aliases the
// padded row scratch and rec2 the CSR columns: both are dead while the
// records live.
struct BucketBuffers {
    VvCpBuffers base;
    BucketPlan plan;
    int32_t *route = nullptr;   // the mesh's routing (Route); AUTO is replaced on first run
    int32_t *h1 = nullptr;      // [max(nc*nchunks, nb) + 1] level-1 counts -> offsets
                                //   (one level: per-bucket fill counters)
    int32_t *h2 = nullptr;      // [nb*slices + 1] level-2 counts -> offsets
                                //   (one level: bucket starts, [nb + 1])
    uint32_t *rec1 = nullptr;   // level-1 output
    uint32_t *rec2 = nullptr;   // level-2 output (bucket order)
    int32_t *rowsrc = nullptr;  // [V] row start in scratch
};

BucketPlan plan_buckets(int64_t n_cells, int64_t n_points, int bv_log = 7, int64_t tpc = 8192);
int launch_bucketed_vv_cp(BucketBuffers &k, bool classify, hipStream_t s);
// Whole bucketed pipeline: counters, routing, VV rows, CSR, link pass.
int run_vv_cp_bucketed(BucketBuffers &k, bool classify, hipStream_t s);
// VT corner lists (codes 4t + j, rows ascending) through the same routing:
// vt_off [V + 1], vtc [4T], vt_cnt [V] (row lengths).  rec1 and rec2 are
// scratch here: they must not alias live data (the mesh passes its own
// buffers, not the VV columns).  Needs 4T <= 2^31 - 1.
int run_vt_bucketed(BucketBuffers &k, int32_t *vt_off, int32_t *vtc, int32_t *vt_cnt,
                    hipStream_t s);

}  // namespace tg
