/*
 * tetra_c.h -- C ABI of libtetra_hip.so: TetraGPU's relation builders and
 * critical-point classifier on one MI355X (gfx950).
 *
 * Every function returns TG_OK (0) or a negative TG_ERR_* status; the
 * message of the last failure on the calling thread is tg_last_error().
 * Work is enqueued on the mesh's HIP stream and runs asynchronously;
 * functions that hand results to the host synchronise that stream.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *  - make_TV_for_GPU / make_VE_for_GPU / make_VF_for_GPU
 *      (include/cuda_extraction.h:11-23): tg_mesh_create uploads TV once;
 *      VE and VF are built on the device, not flattened on the host.
 *  - make_{EV,TF,TE,FV,FE}_GPU_return / make_{...}_GPU
 *      (include/cuda_extraction.h:27-114) and make_VV_GPU_return /
 *      make_VV_GPU (:119-127), make_VT_GPU_return (:137-138):
 *      tg_build(flags) + tg_relation_device / tg_relation_copy.
 *  - get_approx_max_VV (include/cuda_extraction.h:118): tg_approx_max_vv.
 *  - the critPoints kernel and its driver (src/alg/critPoints.cu:49-61,
 *      1092-1269): tg_critpoints / tg_vv_critpoints + tg_classes_copy,
 *      same classes[3*nPoints] layout.
 *  - the relation flag bits of runtime_arguments
 *      (include/datatypes.h.in:158-178): TG_REL_* are the same bit
 *      positions, TG_FLAG(rel) the mask.
 */
#ifndef TETRA_C_H
#define TETRA_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TG_OK             0
#define TG_ERR_ARG       -1
#define TG_ERR_HIP       -2
#define TG_ERR_ALLOC     -3
#define TG_ERR_MESH      -4  /* vertex id out of range or repeated in a cell */
#define TG_ERR_CAPACITY  -5  /* a vertex exceeds the supported degree        */
#define TG_ERR_STATE     -6  /* relation not built / inputs missing         */
#define TG_ERR_NO_DEVICE -7

/* relation ids = bit positions of the reference's arg_flags */
#define TG_REL_TE 1
#define TG_REL_EV 2
#define TG_REL_ET 3
#define TG_REL_TF 4
#define TG_REL_FV 5
#define TG_REL_FE 6
#define TG_REL_FT 7
#define TG_REL_EF 8
#define TG_REL_VT 9
#define TG_REL_TT 10
#define TG_REL_FF 11
#define TG_REL_EE 12
#define TG_REL_VV 13
/* always-initialised relations of the reference (not flag-selectable) */
#define TG_REL_VE 14
#define TG_REL_VF 15
#define TG_FLAG(rel) (1u << (rel))
#define TG_ALL_RELATIONS 0x3FFEu /* bits 1..13 */

typedef struct tg_mesh tg_mesh;

/* Select and warm up a device (hipSetDevice + context creation). */
int tg_init(int device);
const char *tg_last_error(void);
int tg_device_count(int *count);

/* Create a mesh from int32 cells [n_cells][4] and optional per-vertex
 * scalars (float32 when scalar_f64 == 0, else float64; NULL for none).
 * inputs_on_device: the pointers are device pointers (copied on device).
 * stream: hipStream_t to run on, or NULL for a stream owned by the mesh.
 * Connectivity is validated once here (ids in range, no repeated vertex in
 * a cell).  Mirrors make_TV_for_GPU (include/cuda_extraction.h:11). */
int tg_mesh_create(tg_mesh **out, const int32_t *cells, int64_t n_cells,
                   int64_t n_points, const void *scalars, int scalar_f64,
                   int inputs_on_device, void *stream);
int tg_mesh_destroy(tg_mesh *m);
/* Replace the scalar field (resets critical-point results). */
int tg_mesh_set_scalars(tg_mesh *m, const void *scalars, int scalar_f64,
                        int on_device);
int tg_mesh_stream(tg_mesh *m, void **stream);
int tg_mesh_counts(tg_mesh *m, int64_t *n_points, int64_t *n_cells,
                   int64_t *n_edges, int64_t *n_faces);

/* Kernel form and shape of the VV + critical point pipeline.  variant:
 * 0 automatic, 1 vertex-lane form (vertex -> tet lists, one lane per
 * vertex), 2 bucketed form (corners routed to 128/256-vertex buckets, one
 * workgroup per bucket, shared hash tables), 3 bucketed owner form (the
 * bucket's records grouped by vertex in LDS, one lane per vertex), 4
 * bucketed form with the lean insert (record by record, linear CAS probing).
 * lane_cap (32 or 64) and lane_block (64, 128 or 256) shape the vertex-lane
 * form.  Chosen by the autotuner per size class;
 * every supported setting gives identical results. */
int tg_mesh_set_config(tg_mesh *m, int32_t variant, int32_t lane_cap,
                       int32_t lane_block);

/* Shape of the bucketed form: 2^bv_log vertices per bucket (7 or 8) and
 * tets per first-level routing chunk (>= 1024).  Results do not depend on
 * it.  Chosen by the autotuner per size class. */
int tg_mesh_set_bucket(tg_mesh *m, int32_t bv_log, int64_t tets_per_chunk);

/* Routing of the bucketed form's corner records: 0 automatic (one level
 * when the corners of 1024 consecutive tets span fewer than 1024 buckets in
 * at least 63 of 64 such tiles, else two levels; the spans are measured by
 * the connectivity check of tg_mesh_create), 1 one level (tiles ->
 * buckets), 2 two levels (chunks -> coarse bins -> buckets).  Results do
 * not depend on it.  tg_mesh_routing reports the routing in use. */
int tg_mesh_set_routing(tg_mesh *m, int32_t route);
int tg_mesh_routing(tg_mesh *m, int32_t *route);

/* Process-wide tuning table used by variant 0 (automatic): the first entry
 * (ascending max_cells) with n_cells <= max_cells gives the kernel form and
 * its shape.  Written by the autotuner (include/tetra/autotune.h); no
 * counterpart in the reference, which has fixed launch shapes. */
typedef struct tg_tune_entry {
    int64_t max_cells;
    int32_t variant;     /* 1 vertex-lane, 2 bucketed, 3 bucketed owner form, 4 lean */
    int32_t lane_cap;    /* 32 or 64 */
    int32_t lane_block;  /* 64, 128 or 256 */
    int32_t bv_log;      /* 7 or 8 */
    int64_t tets_per_chunk;
} tg_tune_entry;
int tg_set_tuning(const tg_tune_entry *entries, int32_t n);
int tg_get_tuning(tg_tune_entry *entries, int32_t capacity, int32_t *n);

/* Build the relations whose TG_FLAG bits are set, plus what they need.
 * Bit 0 (validate) is ignored.  Asynchronous. */
int tg_build(tg_mesh *m, uint32_t flags);
/* VV (if not built) and classes[3V]. Asynchronous. */
int tg_critpoints(tg_mesh *m);
/* The measured hot path: VV CSR + classes recomputed from TV and scalars
 * resident in HBM, no host synchronisation, no allocation after the first
 * call on a mesh (graph-capturable). */
int tg_vv_critpoints(tg_mesh *m);
/* Wait for the mesh's stream; reports device-side capacity errors. */
int tg_sync(tg_mesh *m);

/* Shape of a built relation: rows, entries, int32 words per entry, and
 * fixed = 1 for fixed-width relations (no offsets). Synchronises. */
int tg_relation_shape(tg_mesh *m, int32_t rel, int64_t *rows, int64_t *nnz,
                      int32_t *width, int32_t *fixed);
/* Copy a built relation to host memory: offsets[rows+1] (ignored for fixed
 * relations, may be NULL) and values[nnz*width]. Synchronises. */
int tg_relation_copy(tg_mesh *m, int32_t rel, int64_t *offsets,
                     int32_t *values);
/* Device pointers of a built relation (owned by the mesh). offset_bytes is
 * 4 or 8 (0 for fixed relations). */
int tg_relation_device(tg_mesh *m, int32_t rel, const void **d_offsets,
                       int32_t *offset_bytes, const int32_t **d_values);

/* classes[3v..3v+2] = (lower link components, upper link components,
 * class 1 minimum / 2 maximum / 3 regular / 4 saddle). Synchronises. */
int tg_classes_copy(tg_mesh *m, uint32_t *classes3);
int tg_classes_device(tg_mesh *m, const uint32_t **d_classes3);
/* Six-way labels numbered as TTK's CriticalType (0 minimum, 1 1-saddle,
 * 2 2-saddle, 3 maximum, 4 multi-saddle, 5 regular). Synchronises. */
int tg_labels6_copy(tg_mesh *m, uint8_t *labels);

/* round_up_32(3 * largest number of cells on one vertex), the reference's
 * padded VV width (src/cuda_extraction.cu:1154-1178), computed on the
 * device for a built mesh. */
int tg_approx_max_vv(tg_mesh *m, int32_t *max_vv);

/* Stage timing of the last pipeline run: when enabled, HIP events are
 * recorded between the pipeline's launches on the mesh stream.
 * tg_profile_read synchronises and returns up to max_stages durations in ms
 * with their stage names (32 chars each) in launch order. */
int tg_profile_enable(tg_mesh *m, int on);
int tg_profile_read(tg_mesh *m, int32_t max_stages, int32_t *n_stages,
                    float *ms, char *names);

/* A built relation converted into new device buffers that the caller owns
 * and releases with tg_device_free: values[nnz*width] as int32 or int64
 * (value_bytes 4 or 8) and, for CSR relations when d_offsets is not NULL,
 * offsets[rows+1] as uint32 or int64 (offset_bytes 4 or 8; NULL for fixed
 * relations).  Backs the vtkIdType* / device_VV / device_VT results of
 * make_X_GPU_return (include/cuda_extraction.h:27-138).  Synchronises. */
int tg_relation_export(tg_mesh *m, int32_t rel, int32_t offset_bytes, void **d_offsets,
                       int32_t value_bytes, void **d_values);
/* Copy `bytes` of host memory into a new device buffer owned by the caller
 * (released with tg_device_free).  Backs make_VT_GPU_return(const VT_Data&)
 * (include/cuda_extraction.h:138), which uploads a host relation. */
int tg_device_upload(const void *host, size_t bytes, void **d_out);

/* ---- batches of independent meshes ----------------------------------------
 * No counterpart in the reference (one mesh per process).  n_meshes meshes
 * given as concatenated int32 cells with LOCAL vertex ids, cell_offsets and
 * point_offsets [n_meshes + 1] (prefix sums, starting at 0), and optional
 * concatenated scalars.  Meshes are grouped into waves of at most
 * max_wave_cells cells (<= 0: the default); each wave runs the VV +
 * critical-point path as one launch sequence over the union of its meshes
 * (results are per mesh: rows and classes are vertex-local).  With
 * tg_batch_use_graph(b, 1) each wave's sequence is captured into a
 * hipGraph on its second run and replayed after.  A cell naming a vertex
 * outside its own mesh is rejected (TG_ERR_MESH). */
typedef struct tg_batch tg_batch;
int tg_batch_create(tg_batch **out, int32_t n_meshes, const int64_t *cell_offsets,
                    const int32_t *cells, const int64_t *point_offsets,
                    const void *scalars, int scalar_f64, int64_t max_wave_cells,
                    void *stream);
int tg_batch_destroy(tg_batch *b);
int tg_batch_counts(tg_batch *b, int32_t *n_waves, int64_t *n_cells,
                    int64_t *n_points);
int tg_batch_use_graph(tg_batch *b, int on);
/* Kernel form for every wave (see tg_mesh_set_config / tg_mesh_set_bucket;
 * bv_log and tets_per_chunk apply to variant 2). Drops captured graphs. */
int tg_batch_set_config(tg_batch *b, int32_t variant, int32_t lane_cap,
                        int32_t lane_block, int32_t bv_log,
                        int64_t tets_per_chunk);
/* VV + classes of every mesh. Asynchronous (graph replay when enabled). */
int tg_batch_vv_critpoints(tg_batch *b);
int tg_batch_sync(tg_batch *b);
/* classes of all meshes, concatenated in mesh order: uint32 [sum V][3]. */
int tg_batch_classes_copy(tg_batch *b, uint32_t *classes3);
/* VV of one mesh with local ids: offsets [V_mesh + 1] (from 0) and values;
 * either may be NULL; *nnz receives the entry count. Synchronises. */
int tg_batch_vv_copy(tg_batch *b, int32_t mesh, int64_t *offsets, int32_t *values,
                     int64_t *nnz);

/* Follow-up relations on the resident batch: VV, VT and/or TT (TG_FLAG
 * bits) of every wave, then one mesh's rows with local ids (VT/TT values
 * are the mesh's own cell ids).  Copy synchronises. */
int tg_batch_build(tg_batch *b, uint32_t flags);
int tg_batch_relation_copy(tg_batch *b, int32_t mesh, int32_t rel, int64_t *offsets,
                           int32_t *values, int64_t *nnz);

/* A HIP stream at the default priority, for callers without a HIP runtime
 * of their own (the serving loop's slots). */
int tg_stream_create(void **stream);
int tg_stream_destroy(void *stream);

/* The library keeps freed device buffers in a stream-ordered cache for
 * reuse by later meshes (capped by TG_POOL_LIMIT_MB, default 64 GiB).
 * tg_pool_trim returns every cached block to the runtime. */
int tg_pool_trim(void);
int tg_pool_cached(int64_t *bytes);

/* Pinned host memory from a reusing cache (fast device->host copies of
 * large relations without page-locking cost per call); TG_HOST_POOL_LIMIT_MB
 * caps what stays cached (default 16 GiB). */
int tg_host_alloc(size_t bytes, void **out);
int tg_host_free(void *p);

/* Free a device pointer handed out by the library to the caller. */
int tg_device_free(void *p);

/* ---- critical points in parts ----------------------------------------
 * Replaces the device side of the reference's partition loop
 * (src/alg/critPoints.cu:654-1003, merged at :1438-1491) and adds the split
 * that its memory handling (:1038-1084, buffers sized per partition) leaves
 * to the caller.
 * tg_vv_footprint: device bytes the VV + critical-point path allocates for a
 * mesh of this size (TV included).
 * tg_critpoints_split: classes [n_points][3] (host) of cells [n_cells][4]
 * int32 and scalars [n_points] float32 / float64 (host).
 *   parts == NULL: the whole mesh on one handle when it fits budget_bytes
 *   and the handle's 32-bit limits; else vertex-id ranges, each classified
 *   on the cells that touch the range or its neighbours (results identical
 *   to the whole mesh; a range whose part does not fit is halved).
 *   parts [n_points] in 0..n_parts-1: the reference's partitions, largest
 *   first; partition p uses row p of part_cells [n_parts][n_cells] (1 =
 *   included) when given, else every cell touching one of its vertices; a
 *   vertex keeps the classes of its own partition ({0,0,0} if no cell of
 *   that partition holds it).
 *   budget_bytes <= 0: 90 % of the device memory free now (the library's
 *   cached blocks counted as free).
 * Localisation runs on the device; TV and scalars stay resident.  Returns
 * TG_ERR_ALLOC when TV and scalars alone, a caller partition or a
 * one-vertex range do not fit the budget. */
typedef struct tg_split_stats {
    int64_t parts;            /* parts classified (with at least one cell) */
    int64_t max_cells;        /* largest part */
    int64_t max_points;
    int64_t cells_total;      /* over all parts (shared cells counted again) */
    int64_t footprint_bytes;  /* tg_vv_footprint of the whole mesh */
    int64_t budget_bytes;     /* the budget used */
    double total_s;           /* wall time of the call */
    double localize_s;        /* device localisation (flags, scans, local ids) */
    double classify_s;        /* part meshes: create, VV + classes, merge */
} tg_split_stats;
int tg_vv_footprint(int64_t n_cells, int64_t n_points, int64_t *bytes);
/* The default budget: 90 % of the device memory free now (cached blocks
 * counted as free). */
int tg_device_budget(int64_t *bytes);
int tg_critpoints_split(const int32_t *cells, int64_t n_cells, int64_t n_points,
                        const void *scalars, int scalar_f64, const int32_t *parts,
                        int32_t n_parts, const int32_t *part_cells, int64_t budget_bytes,
                        uint32_t *classes3, tg_split_stats *stats);

/* ---- no-allocation entry point ---------------------------------------
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as tg_vv_critpoints +
 * tg_relation_device(VV) + tg_classes_device on a mesh, but every pointer is
 * a caller-owned device pointer and the call allocates nothing: its
 * intermediates live in one static device array of 128 MiB per GPU, held
 * until the process ends (tg_noalloc_limits reports its size and what a
 * mesh needs; about 1.1 million cells at most).
 *   cells [n_cells][4] int32, scalars [n_points] float32 (scalar_f64 == 0)
 *   or float64; outputs vv_off [n_points + 1], vv_cols [vv_cap] with
 *   vv_cap >= min(12*n_cells, n_points*(n_points-1)), classes [n_points][3];
 *   status (optional, one int32): the device-side error bits of the run
 *   (1: a vertex over 4096 neighbours), written at its end.
 * Work is enqueued on `stream` (NULL: the null stream) with no host wait, so
 * it can be captured into a graph.  Calls on one GPU share the static array
 * and keep the last call's contents; calls that overlap on one GPU (on
 * unordered streams) race.  Cell connectivity is not validated: ids must be
 * in [0, n_points) and distinct within a cell.  Sizes over the limit return
 * TG_ERR_CAPACITY before any work is enqueued. */
int tg_vv_critpoints_noalloc(const int32_t *cells, int64_t n_cells, int64_t n_points,
                             const void *scalars, int scalar_f64, int32_t *vv_off,
                             int32_t *vv_cols, int64_t vv_cap, uint32_t *classes,
                             int32_t *status, void *stream);
/* Scratch bytes a mesh of this size needs, and the static array's size. */
int tg_noalloc_limits(int64_t n_cells, int64_t n_points, int64_t *need_bytes,
                      int64_t *max_bytes);

#ifdef __cplusplus
}
#endif
#endif
