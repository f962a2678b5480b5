// Host launchers of the relation and critical-point pipelines.  Every
// launcher only enqueues work on the given stream (no host synchronisation,
// no allocation, no hipMemset), so a launch sequence can be captured into a
// hipGraph and replayed.  One exception: the first bucketed run of a mesh
// waits once for the routing decision (one-level or two-level), which is
// then kept with the mesh; a run under capture never waits.
#pragma once

#include "common.h"

namespace tg {

// Device counters shared by one pipeline run.
enum Counter : int {
    C_HEAVY = 0,   // vertices routed to the serial VV kernel
    C_FIX = 1,     // vertices needing the row-membership link pass (deg<=64)
    C_FIXBIG = 2,  // same, deg > 64
    C_ERR = 3,     // error bits (ERRBIT_*)
    C_MAXDEG = 4,  // largest VV degree seen
    C_FIXD = 5,    // bucketed vertices needing the link pass (row descriptors)
    C_FTERR = 6,   // a face held by more than two cells
    C_IDERR = 7,   // face ids: a relation read back inconsistent
    C_WIDE = 8,    // one-level routing: tiles whose corners span too many buckets
    C_SCRATCH = 9, // one-off reductions of the host API
    N_COUNTERS = 10
};
enum ErrBit : int { ERRBIT_CAPACITY = 1 };

// Largest vertex degree the serial per-vertex kernels hold in LDS.
constexpr int HEAVY_CAP = 4096;

// Optional per-stage HIP events recorded between the launches of a pipeline
// (enabled by tg_profile_enable); mark() is a no-op when disabled.
struct StageMarks {
    static constexpr int MAX = 32;
    bool on = false;
    int n = 0;
    hipEvent_t ev[MAX] = {};
    const char *name[MAX] = {};
    void begin(hipStream_t s);
    void mark(const char *stage, hipStream_t s);
};

// Tunables of the VV + critical point pipeline (picked by the autotuner).
struct VvCpConfig {
    int variant = 0;       // 0 auto, 1 vertex-lane form, 2 bucketed, 3 bucketed owner form,
                           // 4 bucketed with the lean insert
    int lane_cap = 32;     // neighbours held per lane (32 or 64)
    int lane_block = 128;  // lanes (vertices) per workgroup
    int corner_block = 256;
    int bv_log = 7;        // bucketed form: 2^bv_log vertices per bucket (7 or 8)
    int64_t tpc = 8192;    // bucketed form: tets per first-level chunk
    int route = 0;         // bucketed form: 0 auto, 1 one-level, 2 two-level routing
    bool tuned = false;    // variant 0 resolved from the tuning table
};

// Device buffers of one VV + critical point run.  Sizes:
//   vt_cnt, vt_cur, deg : V      vt_off, vv_off : V+1
//   vtc : 4T                     scratch : 12T (padded rows at 3*vt_off[v])
//   vv_cols : 12T (capacity)     classes : 3V
//   heavy, fix, fixbig : V       counters : N_COUNTERS
//   scan_tmp : scan_tmp_bytes(V)
struct VvCpBuffers {
    const int32_t *tv = nullptr;
    const void *scalars = nullptr;
    bool scalars_f64 = false;
    int64_t n_cells = 0, n_points = 0;
    int32_t *vt_cnt = nullptr, *vt_off = nullptr, *vt_cur = nullptr,
            *vtc = nullptr, *scratch = nullptr, *deg = nullptr,
            *vv_off = nullptr, *vv_cols = nullptr;
    uint32_t *classes = nullptr;
    int32_t *heavy = nullptr, *fix = nullptr, *fixbig = nullptr,
            *counters = nullptr;
    // bucketed form: per-entry (lower bit, shared-tet component) of each row,
    // in the padded row layout, and the vertices that use it
    int32_t *fixd = nullptr, *rowsrc = nullptr;
    uint8_t *rowdesc = nullptr;
    void *scan_tmp = nullptr;
    StageMarks *marks = nullptr;
};

// vertex -> corner incidence (p = 4t + j) grouped by vertex, unordered rows.
int launch_vt_corners(VvCpBuffers &b, int block, hipStream_t s);
// Sort every vt row ascending (needed by the id-assigning builders).
int launch_sort_rows_i32(const int32_t *off, int32_t *vals, int64_t rows,
                         hipStream_t s);
// VV CSR (ascending rows) + classes[3V]; assumes launch_vt_corners ran.
int launch_vv_cp(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
                 hipStream_t s);
// Gather rows from padded scratch (row of v starts at rowsrc[v]) into the CSR.
int launch_rows_compact(const int32_t *rowsrc, const int32_t *vv_off, const int32_t *scratch,
                        int32_t V, int32_t *cols, hipStream_t s);
// Row-membership link components for the vertices queued in fix/fixbig.
int launch_cp_fix(VvCpBuffers &b, hipStream_t s);
// Whole pipeline: counters reset, VT corners, VV, critical points.
int run_vv_cp(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
              hipStream_t s);
// The same on corner lists that are already built (vt_off / vtc untouched).
int run_vv_cp_on_corners(VvCpBuffers &b, const VvCpConfig &cfg, bool classify,
                         hipStream_t s);

}  // namespace tg
