/*
 * serve.h -- a request loop for VV + critical-point jobs on one GPU, with
 * continuous batching (no counterpart in the reference, which runs one mesh
 * per process).
 *
 * Requests (a mesh + scalars) are admitted in arrival order as GPU memory
 * (an HBM budget, by a per-request footprint estimate) and slots free up;
 * each slot is a worker thread with its own HIP stream, so groups on
 * different slots overlap on the device.  A slot that takes a small request
 * also takes the small requests waiting behind it, up to a cell and count
 * limit, and runs them as one batch launch (tg_batch_*): results are per
 * request and identical to running the request alone.  A finished request's
 * relations stay resident until it is released, for follow-up VT/TT queries.
 *
 * Input buffers must stay valid until the request's first result is ready.
 */
#ifndef TETRA_SERVE_H
#define TETRA_SERVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tg_server tg_server;

typedef struct tg_server_config {
    int32_t slots;             /* concurrent groups (worker threads / streams), >= 1 */
    int64_t hbm_budget_bytes;  /* admission limit for resident requests (> 0) */
    int64_t fuse_max_cells;    /* fuse waiting requests while their cells sum to at most this */
    int32_t fuse_max_meshes;   /* and they are at most this many (<= 1: no fusion) */
} tg_server_config;

typedef struct tg_request_result {
    int32_t status;             /* TG_OK or the failing call's status */
    int64_t n_points, n_cells;
    const uint32_t *classes;    /* [n_points][3] */
    const int64_t *vt_off;      /* [n_points + 1] when VT was asked for, else NULL */
    const int32_t *vt_val;
    int64_t vt_nnz;
    const int64_t *tt_off;      /* [n_cells + 1] when TT was asked for, else NULL */
    const int32_t *tt_val;
    int64_t tt_nnz;
    double queue_ms;            /* submit -> start on a slot */
    double run_ms;              /* start -> first result (incl. submit-time follow-ups) */
    double total_ms;            /* submit -> first result */
    int32_t group_size;         /* requests fused into the launch that served it */
    /* phases of run_ms (for a fused group, the group's): */
    double upload_ms;           /* mesh creation: host->device copy, connectivity check */
    double compute_ms;          /* VV + critical points and the classes' copy back */
    /* the last follow-up's phases (also submit-time ones): */
    double followup_ms;         /* building VT / TT on the device */
    double copy_ms;             /* their copy to host memory (0 without TG_SERVE_COPY) */
} tg_request_result;

/* Follow-up flag: also copy the relations to host memory (vt_off/vt_val,
 * tt_off/tt_val).  Without it they are built and stay on the device (counts
 * in vt_nnz / tt_nnz), ready for the next query. */
#define TG_SERVE_COPY (1u << 31)

int tg_server_create(tg_server **out, const tg_server_config *cfg);
/* followups: TG_FLAG(TG_REL_VT) | TG_FLAG(TG_REL_TT) to compute with the
 * first result (counted in its latency), optionally | TG_SERVE_COPY.  cells
 * and scalars may be pageable; they are read (into the slot's pinned
 * staging buffer) when the request starts, and must stay valid until
 * tg_server_wait returns. */
int tg_server_submit(tg_server *s, const int32_t *cells, int64_t n_cells, int64_t n_points,
                     const float *scalars, uint32_t followups, uint64_t *id);
/* Block until the request's first result; pointers stay valid until release. */
int tg_server_wait(tg_server *s, uint64_t id, tg_request_result *out);
/* Follow-up query on the request's resident relations (VT and/or TT,
 * optionally | TG_SERVE_COPY). */
int tg_server_followup(tg_server *s, uint64_t id, uint32_t flags, tg_request_result *out);
/* Drop the request; its device memory returns to the budget when every
 * request of its group is released. */
int tg_server_release(tg_server *s, uint64_t id);
/* Requests admitted and groups launched so far; bytes resident now. */
int tg_server_stats(tg_server *s, int64_t *requests, int64_t *groups, int64_t *resident_bytes);
int tg_server_destroy(tg_server *s);

#ifdef __cplusplus
}
#endif
#endif
