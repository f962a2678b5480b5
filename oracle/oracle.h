/*
 * oracle.h -- CPU restatement of TetraGPU's relation builders and critical-
 * point classifier.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (tetragpu_amd/, the C-ABI
 * library, the CLI) links, loads or calls this code.  It is imported by
 * tests/ (as the parity checker), by __graft_entry__.smoke() (as the checker
 * of one small run) and by bench.py's cpu_baseline leg (as the timed CPU
 * baseline).  The serial builders follow the reference's algorithms step by
 * step (linear find over per-vertex lists, first-occurrence ID minting); the
 * *_omp entry points compute the same outputs with OpenMP for timing.
 *
 * Conventions: vertex/cell/edge/face ids are int64 (the reference's
 * vtkIdType).  A CSR result has `rows`, `nnz` (entries), `width` (int64 words
 * per entry) and arrays off[rows+1], val[nnz*width], allocated with malloc
 * and released by orc_csr_free().
 *
 * Parity pinning of every function is listed in oracle.c's header.
 */
#ifndef TETRA_ORACLE_H
#define TETRA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t rows;
    int64_t nnz;
    int32_t width;
    int64_t *off;  /* rows + 1 */
    int64_t *val;  /* nnz * width */
} orc_csr;

void orc_csr_free(orc_csr *c);

/* Edge ids by first occurrence; VE rows {highVert, id} in ascending id.
 * TE may be NULL (then this is make_VE).  Returns E, or -1 on bad input. */
int64_t orc_make_TE_and_VE(const int64_t *cells, int64_t n_cells,
                           int64_t n_points, int64_t *TE, orc_csr *VE);

/* Face ids by first occurrence of the positional key; VF rows
 * {middleVert, highVert, id} in ascending id.  TF may be NULL. */
int64_t orc_make_TF_and_VF(const int64_t *cells, int64_t n_cells,
                           int64_t n_points, int64_t *TF, orc_csr *VF);

/* EV[id] = {lo, hi}: out[2*E]. */
void orc_make_EV(const orc_csr *VE, int64_t n_edges, int64_t *EV);
/* ET: edge -> tets in ascending tet order. */
void orc_make_ET(const int64_t *TE, int64_t n_cells, int64_t n_edges,
                 orc_csr *ET);
/* FV[id] = FaceData(middleVert, highVert, id := low vertex): out[3*F]
 * in field order (middleVert, highVert, lowVertex). */
void orc_make_FV(const orc_csr *VF, int64_t n_faces, int64_t *FV);
/* VV in the reference's insertion order (cells ascending, corners 0..3). */
void orc_make_VV(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *VV);
/* VT: vertex -> tets in ascending order. */
void orc_make_VT(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *VT);

/* Relations the reference names by flag only (definitions in oracle.c). */
int orc_make_FE(const int64_t *FV, int64_t n_faces, const orc_csr *VE,
                int64_t *FE /* [F][3][2] = {highVert, edge id} */);
int orc_make_FT(const int64_t *TF, int64_t n_cells, int64_t n_faces,
                int64_t *FT /* [F][2], -1 padded */);
void orc_make_EF(const int64_t *FE, int64_t n_faces, int64_t n_edges,
                 orc_csr *EF);
void orc_make_TT(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *TT);
void orc_make_FF(const int64_t *TF, int64_t n_cells, const int64_t *FT,
                 int64_t n_faces, orc_csr *FF);
void orc_make_EE(const int64_t *EV, int64_t n_edges, int64_t n_points,
                 orc_csr *EE);

/* Critical points: classes[3v..3v+2] = (lower components, upper components,
 * class) with class 1 = minimum, 2 = maximum, 3 = regular, 4 = saddle (the
 * numeric values the reference kernel writes).  VV is any CSR whose rows
 * hold each vertex's distinct neighbours (order free). */
void orc_critpoints(const orc_csr *VV, const double *scalars,
                    int64_t n_points, uint32_t *classes);
/* Six-way label from (cL, cU), numbered as TTK's CriticalType: 0 minimum,
 * 1 1-saddle, 2 2-saddle, 3 maximum, 4 multi-saddle, 5 regular. */
void orc_label6(const uint32_t *classes, int64_t n_points, uint8_t *label);

/* round_up_32(3 * max #tets per vertex) as the reference's max-VV estimate. */
int32_t orc_approx_max_vv(const int64_t *cells, int64_t n_cells,
                          int64_t n_points);

/* OpenMP form used as the timed CPU baseline: VV with ascending rows plus
 * critical points, same outputs as the serial builders (as row sets) and
 * orc_critpoints.  cells are int32 as resident on the device.  Returns the
 * wall time in seconds spent in the computation. */
double orc_vv_cp_omp(const int32_t *cells, int64_t n_cells, int64_t n_points,
                     const float *scalars, int threads, orc_csr *VV,
                     uint32_t *classes);

/* Widen int32 cells to int64 (helper for callers holding device-layout TV). */
void orc_widen(const int32_t *in, int64_t n, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
