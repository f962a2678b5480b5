/*
 * oracle.c -- CPU restatement of TetraGPU's relation builders and critical-
 * point rule.  TEST INFRASTRUCTURE ONLY (see oracle.h): the product never
 * links or calls it.
 *
 * Which reference code each function follows, and what pins it (the tests
 * that check the restatement itself, not just the GPU against it):
 *
 *  orc_make_TE_and_VE  src/cpu_extraction.cpp:28-62 (make_TE_and_VE) and
 *                      :64-92 (make_VE).  Pinned by the Kuhn closed form
 *                      E = 7n^3+9n^2+3n, the hand-derived single-cube
 *                      fixture, a dictionary-based first-occurrence
 *                      re-derivation on random tiny meshes, EV endpoints
 *                      ordered lo<hi, and V-E+F-T = 1 on ball meshes.
 *  orc_make_TF_and_VF  src/cpu_extraction.cpp:122-174 / :176-222 (slot loop
 *                      skip_id = 3..0, TF[c][3-skip_id]).  Pinned by
 *                      F = 12n^3+6n^2 for ascending cells, F >= that for
 *                      unsorted cells (positional keys), the single-cube
 *                      fixture and the Euler characteristic.
 *  orc_make_EV         src/cpu_extraction.cpp:95-107.  Pinned through EV
 *                      endpoint checks against TE/TV.
 *  orc_make_ET         src/cpu_extraction.cpp:109-120.  Pinned by nnz = 6T
 *                      and by membership checks against TV.
 *  orc_make_FV         src/cpu_extraction.cpp:225-243.  Pinned by the
 *                      single-cube fixture and TV membership.
 *  orc_make_VV         src/cpu_extraction.cpp:259-289.  Pinned by brute
 *                      force (pairs sharing a tet), sum of degrees = 2E,
 *                      Kuhn interior degree 14.
 *  orc_make_VT         src/cpu_extraction.cpp:291-342.  Pinned by nnz = 4T,
 *                      Kuhn max incidence 24, TV membership.
 *  orc_approx_max_vv   src/cuda_extraction.cu:1154-1178.  Pinned by the Kuhn
 *                      value round32(3*24) = 96.
 *  orc_critpoints      src/alg/critPoints.cu:49-338 (kernel critPoints):
 *                      lower neighbour test :136-140, link edges by row
 *                      membership :206-246, root counting :275-284 into
 *                      classes[3v+0] (lower) / [3v+1] (upper), verdict
 *                      :292-331 with the numeric values 1..4.  Pinned by:
 *                      a linear field on an unpermuted Kuhn grid (exactly one
 *                      class-1 and one class-2 vertex, at the known grid
 *                      corners), a constant field (ties resolved by id), a
 *                      hand-built vertex with a two-component lower link, an
 *                      isolated vertex (0,0,4), a non-flag link where the
 *                      row-membership rule and the shared-tet rule differ,
 *                      and a BFS re-derivation of component counts.
 *  orc_label6          no reference counterpart (the reference stops at the
 *                      numeric saddle class 4); the six-way split follows
 *                      Morse indices: (2,1) 1-saddle, (1,2) 2-saddle.
 *                      Parity unpinned against an external tool.
 *  orc_make_FE/FT/EF/TT/FF/EE  the reference names these by flag only
 *                      (include/datatypes.h.in:173-178, "not implemented" in
 *                      src/alg/test_driver.cu:302-343).  Definitions:
 *                        FE[f] = edges (lo,mid),(lo,hi),(mid,hi) of the
 *                                positional face, each {max vertex, edge id}
 *                        FT[f] = the <=2 tets holding f, ascending, -1 pad
 *                        EF[e] = faces holding e, ascending
 *                        TT[t] = tets sharing >=1 vertex with t, ascending
 *                        FF[f] = faces sharing a tet with f, ascending
 *                        EE[e] = edges sharing a vertex with e, ascending
 *                      Pinned only by brute force on tiny meshes and the
 *                      Kuhn counts (FF nnz = 12T on closed-face counts etc.);
 *                      parity against the reference is unpinned because the
 *                      reference has no implementation.
 *  orc_vv_cp_omp       same outputs as orc_make_VV (row sets, here ascending)
 *                      plus orc_critpoints; checked against them in tests.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- growable per-row lists ------------------------------------------ */
typedef struct {
    int64_t *d;
    int64_t n;    /* words used */
    int64_t cap;  /* words allocated */
} vec;

static void vpush(vec *v, int64_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 8;
        v->d = (int64_t *)realloc(v->d, (size_t)v->cap * sizeof(int64_t));
    }
    v->d[v->n++] = x;
}

/* Flatten an array of lists (each holding `width`-word entries) into CSR. */
static void flatten(vec *rows, int64_t n_rows, int32_t width, orc_csr *out) {
    out->rows = n_rows;
    out->width = width;
    out->off = (int64_t *)malloc((size_t)(n_rows + 1) * sizeof(int64_t));
    int64_t total = 0;
    for (int64_t r = 0; r < n_rows; r++) {
        out->off[r] = total;
        total += rows[r].n / width;
    }
    out->off[n_rows] = total;
    out->nnz = total;
    out->val = (int64_t *)malloc((size_t)(total * width + 1) * sizeof(int64_t));
    for (int64_t r = 0; r < n_rows; r++) {
        if (rows[r].n)
            memcpy(out->val + out->off[r] * width, rows[r].d,
                   (size_t)rows[r].n * sizeof(int64_t));
        free(rows[r].d);
    }
    free(rows);
}

void orc_csr_free(orc_csr *c) {
    if (!c) return;
    free(c->off);
    free(c->val);
    c->off = c->val = NULL;
    c->rows = c->nnz = 0;
}

/* CSR from per-row counts: cnt[0..rows) becomes the offsets (rows + 1
 * entries, cnt must hold rows + 1); values are allocated, not filled. */
static void csr_from_counts(int64_t *cnt, int64_t rows, int32_t width, orc_csr *out) {
    int64_t total = 0;
    for (int64_t r = 0; r < rows; r++) {
        const int64_t n = cnt[r];
        cnt[r] = total;
        total += n;
    }
    cnt[rows] = total;
    out->rows = rows;
    out->width = width;
    out->nnz = total;
    out->off = cnt;
    out->val = (int64_t *)malloc((size_t)(total * width + 1) * sizeof(int64_t));
}

/* Rows r <- items i for every (r, i) pair given by key(i) over i ascending,
 * so each row lists its items ascending (a counting sort by row). */
static void csr_by_key(const int64_t *key, int64_t n_items, int64_t per_item_div,
                       int64_t rows, orc_csr *out) {
    int64_t *cnt = (int64_t *)calloc((size_t)rows + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n_items; i++) cnt[key[i]]++;
    csr_from_counts(cnt, rows, 1, out);
    int64_t *cur = (int64_t *)malloc(((size_t)rows + 1) * sizeof(int64_t));
    memcpy(cur, out->off, ((size_t)rows + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n_items; i++) out->val[cur[key[i]]++] = i / per_item_div;
    free(cur);
}

static int cells_valid(const int64_t *cells, int64_t n_cells, int64_t n_points) {
    for (int64_t i = 0; i < 4 * n_cells; i++)
        if (cells[i] < 0 || cells[i] >= n_points) return 0;
    return 1;
}

/* ---- edges (cpu_extraction.cpp:28-92) --------------------------------- */
int64_t orc_make_TE_and_VE(const int64_t *cells, int64_t n_cells,
                           int64_t n_points, int64_t *TE, orc_csr *VE) {
    if (!cells_valid(cells, n_cells, n_points)) return -1;
    vec *table = (vec *)calloc((size_t)n_points + 1, sizeof(vec));
    int64_t count = 0;
    for (int64_t c = 0; c < n_cells; c++) {
        int k = 0;
        for (int j = 0; j < 3; j++) {
            for (int l = j + 1; l < 4; l++, k++) {
                int64_t lo = cells[4 * c + j], hi = cells[4 * c + l];
                if (lo > hi) { int64_t t = lo; lo = hi; hi = t; }
                vec *L = &table[lo];
                int64_t id = -1;
                for (int64_t i = 0; i < L->n; i += 2)
                    if (L->d[i] == hi) { id = L->d[i + 1]; break; }
                if (id < 0) {
                    id = count++;
                    vpush(L, hi);
                    vpush(L, id);
                }
                if (TE) TE[6 * c + k] = id;
            }
        }
    }
    flatten(table, n_points, 2, VE);
    return count;
}

/* ---- faces (cpu_extraction.cpp:122-222) ------------------------------- */
/* Slot s = 3 - skip_id; corner positions (lo, mid, hi) of each slot. */
static const int FACE_POS[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

int64_t orc_make_TF_and_VF(const int64_t *cells, int64_t n_cells,
                           int64_t n_points, int64_t *TF, orc_csr *VF) {
    if (!cells_valid(cells, n_cells, n_points)) return -1;
    vec *table = (vec *)calloc((size_t)n_points + 1, sizeof(vec));
    int64_t count = 0;
    for (int64_t c = 0; c < n_cells; c++) {
        const int64_t *cv = cells + 4 * c;
        for (int s = 0; s < 4; s++) {
            int64_t a = cv[FACE_POS[s][0]], b = cv[FACE_POS[s][1]],
                    d = cv[FACE_POS[s][2]];
            vec *L = &table[a];
            int64_t id = -1;
            for (int64_t i = 0; i < L->n; i += 3)
                if (L->d[i] == b && L->d[i + 1] == d) { id = L->d[i + 2]; break; }
            if (id < 0) {
                id = count++;
                vpush(L, b);
                vpush(L, d);
                vpush(L, id);
            }
            if (TF) TF[4 * c + s] = id;
        }
    }
    flatten(table, n_points, 3, VF);
    return count;
}

/* ---- transposes and elective builders --------------------------------- */
void orc_make_EV(const orc_csr *VE, int64_t n_edges, int64_t *EV) {
    (void)n_edges;
    for (int64_t v = 0; v < VE->rows; v++)
        for (int64_t i = VE->off[v]; i < VE->off[v + 1]; i++) {
            int64_t hi = VE->val[2 * i], id = VE->val[2 * i + 1];
            EV[2 * id] = v;
            EV[2 * id + 1] = hi;
        }
}

void orc_make_ET(const int64_t *TE, int64_t n_cells, int64_t n_edges,
                 orc_csr *ET) {
    /* cell c is appended to the rows of its six edges, cells ascending */
    csr_by_key(TE, 6 * n_cells, 6, n_edges, ET);
}

void orc_make_FV(const orc_csr *VF, int64_t n_faces, int64_t *FV) {
    (void)n_faces;
    for (int64_t v = 0; v < VF->rows; v++)
        for (int64_t i = VF->off[v]; i < VF->off[v + 1]; i++) {
            int64_t mid = VF->val[3 * i], hi = VF->val[3 * i + 1],
                    id = VF->val[3 * i + 2];
            FV[3 * id] = mid;
            FV[3 * id + 1] = hi;
            FV[3 * id + 2] = v;
        }
}

void orc_make_VV(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *VV) {
    vec *rows = (vec *)calloc((size_t)n_points + 1, sizeof(vec));
    for (int64_t c = 0; c < n_cells; c++)
        for (int i = 0; i < 4; i++) {
            int64_t src = cells[4 * c + i];
            for (int j = 0; j < 4; j++) {
                if (i == j) continue;
                int64_t dst = cells[4 * c + j];
                vec *L = &rows[src];
                int found = 0;
                for (int64_t k = 0; k < L->n; k++)
                    if (L->d[k] == dst) { found = 1; break; }
                if (!found) vpush(L, dst);
            }
        }
    flatten(rows, n_points, 1, VV);
}

void orc_make_VT(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *VT) {
    /* cell c is appended to the rows of its four vertices, cells ascending */
    csr_by_key(cells, 4 * n_cells, 4, n_points, VT);
}

/* ---- relations defined here (no reference implementation) ------------ */
static int64_t ve_lookup(const orc_csr *VE, int64_t a, int64_t b) {
    if (a > b) { int64_t t = a; a = b; b = t; }
    for (int64_t i = VE->off[a]; i < VE->off[a + 1]; i++)
        if (VE->val[2 * i] == b) return VE->val[2 * i + 1];
    return -1;
}

int orc_make_FE(const int64_t *FV, int64_t n_faces, const orc_csr *VE,
                int64_t *FE) {
    int missing = 0;
#pragma omp parallel for schedule(static) reduction(| : missing)
    for (int64_t f = 0; f < n_faces; f++) {
        int64_t mid = FV[3 * f], hi = FV[3 * f + 1], lo = FV[3 * f + 2];
        int64_t p[3][2] = {{lo, mid}, {lo, hi}, {mid, hi}};
        for (int e = 0; e < 3; e++) {
            int64_t id = ve_lookup(VE, p[e][0], p[e][1]);
            missing |= id < 0;
            FE[6 * f + 2 * e] = p[e][0] > p[e][1] ? p[e][0] : p[e][1];
            FE[6 * f + 2 * e + 1] = id;
        }
    }
    return missing ? -1 : 0;
}

int orc_make_FT(const int64_t *TF, int64_t n_cells, int64_t n_faces,
                int64_t *FT) {
    for (int64_t f = 0; f < 2 * n_faces; f++) FT[f] = -1;
    for (int64_t c = 0; c < n_cells; c++)
        for (int s = 0; s < 4; s++) {
            int64_t f = TF[4 * c + s];
            if (FT[2 * f] == c || FT[2 * f + 1] == c) continue;
            if (FT[2 * f] < 0) FT[2 * f] = c;
            else if (FT[2 * f + 1] < 0) FT[2 * f + 1] = c;
            else return -1; /* a face held by more than two tets */
        }
    return 0;
}

void orc_make_EF(const int64_t *FE, int64_t n_faces, int64_t n_edges,
                 orc_csr *EF) {
    /* face f is appended once to the row of each distinct edge it holds,
     * faces ascending */
    int64_t *cnt = (int64_t *)calloc((size_t)n_edges + 1, sizeof(int64_t));
    for (int64_t f = 0; f < n_faces; f++)
        for (int e = 0; e < 3; e++) {
            const int64_t id = FE[6 * f + 2 * e + 1];
            int dup = 0;
            for (int x = 0; x < e; x++) dup |= FE[6 * f + 2 * x + 1] == id;
            if (!dup) cnt[id]++;
        }
    csr_from_counts(cnt, n_edges, 1, EF);
    int64_t *cur = (int64_t *)malloc(((size_t)n_edges + 1) * sizeof(int64_t));
    memcpy(cur, EF->off, ((size_t)n_edges + 1) * sizeof(int64_t));
    for (int64_t f = 0; f < n_faces; f++)
        for (int e = 0; e < 3; e++) {
            const int64_t id = FE[6 * f + 2 * e + 1];
            int dup = 0;
            for (int x = 0; x < e; x++) dup |= FE[6 * f + 2 * x + 1] == id;
            if (!dup) EF->val[cur[id]++] = f;
        }
    free(cur);
}

/* Ascending sort of a short array: quicksort (middle pivot) down to runs of
 * 16, then insertion sort. */
static void sort_i64(int64_t *a, int64_t n) {
    while (n > 16) {
        const int64_t p = a[n / 2];
        int64_t i = 0, j = n - 1;
        while (i <= j) {
            while (a[i] < p) i++;
            while (a[j] > p) j--;
            if (i <= j) {
                const int64_t t = a[i];
                a[i] = a[j];
                a[j] = t;
                i++;
                j--;
            }
        }
        /* recurse into the smaller side, loop on the larger */
        if (j + 1 < n - i) {
            sort_i64(a, j + 1);
            a += i;
            n -= i;
        } else {
            sort_i64(a + i, n - i);
            n = j + 1;
        }
    }
    for (int64_t i = 1; i < n; i++) {
        const int64_t x = a[i];
        int64_t j = i - 1;
        while (j >= 0 && a[j] > x) {
            a[j + 1] = a[j];
            j--;
        }
        a[j + 1] = x;
    }
}

/* Sort, drop duplicates and drop `self`; returns the new length. */
static int64_t sort_unique_drop(int64_t *buf, int64_t n, int64_t self) {
    sort_i64(buf, n);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) {
        if (buf[i] == self) continue;
        if (m > 0 && buf[m - 1] == buf[i]) continue;
        buf[m++] = buf[i];
    }
    return m;
}

/* Row-parallel CSR build: row(r, buf) writes row r (sorted, unique) into a
 * thread's buffer and returns its length; rows are computed twice, once to
 * count and once to fill, so every row is the same computation. */
typedef int64_t (*row_fn)(const void *ctx, int64_t r, vec *buf);

static void csr_by_rows(const void *ctx, row_fn row, int64_t rows, orc_csr *out) {
    int64_t *cnt = (int64_t *)calloc((size_t)rows + 1, sizeof(int64_t));
#pragma omp parallel
    {
        vec buf = {0, 0, 0};
#pragma omp for schedule(dynamic, 4096)
        for (int64_t r = 0; r < rows; r++) cnt[r] = row(ctx, r, &buf);
        free(buf.d);
    }
    csr_from_counts(cnt, rows, 1, out);
#pragma omp parallel
    {
        vec buf = {0, 0, 0};
#pragma omp for schedule(dynamic, 4096)
        for (int64_t r = 0; r < rows; r++) {
            const int64_t m = row(ctx, r, &buf);
            if (m) memcpy(out->val + out->off[r], buf.d, (size_t)m * sizeof(int64_t));
        }
        free(buf.d);
    }
}

typedef struct {
    const int64_t *cells;
    orc_csr VT;
} tt_ctx;

/* TT[t]: the cells of t's four VT rows, sorted, unique, without t */
static int64_t tt_row(const void *p, int64_t t, vec *buf) {
    const tt_ctx *c = (const tt_ctx *)p;
    buf->n = 0;
    for (int i = 0; i < 4; i++) {
        const int64_t v = c->cells[4 * t + i];
        for (int64_t k = c->VT.off[v]; k < c->VT.off[v + 1]; k++) vpush(buf, c->VT.val[k]);
    }
    return sort_unique_drop(buf->d, buf->n, t);
}

void orc_make_TT(const int64_t *cells, int64_t n_cells, int64_t n_points,
                 orc_csr *TT) {
    tt_ctx c;
    c.cells = cells;
    orc_make_VT(cells, n_cells, n_points, &c.VT);
    csr_by_rows(&c, tt_row, n_cells, TT);
    orc_csr_free(&c.VT);
}

typedef struct {
    const int64_t *TF, *FT;
} ff_ctx;

/* FF[f]: the faces of f's (at most two) cells, sorted, unique, without f */
static int64_t ff_row(const void *p, int64_t f, vec *buf) {
    const ff_ctx *c = (const ff_ctx *)p;
    buf->n = 0;
    for (int j = 0; j < 2; j++) {
        const int64_t t = c->FT[2 * f + j];
        if (t < 0) continue;
        for (int s = 0; s < 4; s++) vpush(buf, c->TF[4 * t + s]);
    }
    return sort_unique_drop(buf->d, buf->n, f);
}

void orc_make_FF(const int64_t *TF, int64_t n_cells, const int64_t *FT,
                 int64_t n_faces, orc_csr *FF) {
    (void)n_cells;
    ff_ctx c = {TF, FT};
    csr_by_rows(&c, ff_row, n_faces, FF);
}

typedef struct {
    const int64_t *EV;
    orc_csr inc;  /* vertex -> incident edges, ascending */
} ee_ctx;

/* EE[e]: the edges at either endpoint of e, sorted, unique, without e */
static int64_t ee_row(const void *p, int64_t e, vec *buf) {
    const ee_ctx *c = (const ee_ctx *)p;
    buf->n = 0;
    for (int j = 0; j < 2; j++) {
        const int64_t v = c->EV[2 * e + j];
        for (int64_t k = c->inc.off[v]; k < c->inc.off[v + 1]; k++) vpush(buf, c->inc.val[k]);
    }
    return sort_unique_drop(buf->d, buf->n, e);
}

void orc_make_EE(const int64_t *EV, int64_t n_edges, int64_t n_points,
                 orc_csr *EE) {
    ee_ctx c;
    c.EV = EV;
    csr_by_key(EV, 2 * n_edges, 2, n_points, &c.inc);
    csr_by_rows(&c, ee_row, n_edges, EE);
    orc_csr_free(&c.inc);
}

/* ---- critical points (critPoints.cu:49-338) --------------------------- */
static int64_t uf_find(int64_t *p, int64_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

/* Neighbour a of v is "lower" exactly as critPoints.cu:136-140 decides it:
 * equal scalars fall back to the id comparison, otherwise a strict less-than
 * (so a NaN on either side makes a an upper neighbour). */
static int is_lower(double sa, double sv, int64_t a, int64_t v) {
    return (sa == sv) ? (a < v) : (sa < sv);
}

static uint32_t verdict(uint32_t lower, uint32_t upper) {
    /* critPoints.cu:303-330 reads classes[3v] (lower components) as its
     * "upper" and classes[3v+1] as its "lower"; the numbers it writes are: */
    if (lower >= 1 && upper == 0) return 2; /* every neighbour below: maximum */
    if (lower == 0 && upper >= 1) return 1; /* every neighbour above: minimum */
    if (lower == 1 && upper == 1) return 3; /* regular */
    return 4;                               /* saddle, including (0,0) */
}

void orc_critpoints(const orc_csr *VV, const double *scalars,
                    int64_t n_points, uint32_t *classes) {
    int64_t *pos = (int64_t *)malloc((size_t)(n_points + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n_points; i++) pos[i] = -1;
    int64_t cap = 16;
    int64_t *nb = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    int64_t *parent = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    char *low = (char *)malloc((size_t)cap);
    for (int64_t v = 0; v < n_points; v++) {
        int64_t deg = 0;
        int64_t row = VV->off[v + 1] - VV->off[v];
        if (row > cap) {
            cap = row;
            nb = (int64_t *)realloc(nb, (size_t)cap * sizeof(int64_t));
            parent = (int64_t *)realloc(parent, (size_t)cap * sizeof(int64_t));
            low = (char *)realloc(low, (size_t)cap);
        }
        /* distinct neighbours, first occurrence kept (critPoints.cu:100-111) */
        for (int64_t i = VV->off[v]; i < VV->off[v + 1]; i++) {
            int64_t a = VV->val[i];
            if (a < 0 || a >= n_points || pos[a] >= 0) continue;
            pos[a] = deg;
            nb[deg] = a;
            parent[deg] = deg;
            low[deg] = (char)is_lower(scalars[a], scalars[v], a, v);
            deg++;
        }
        /* link edge (a,c) iff c is in row(a); union same-class ends */
        for (int64_t ia = 0; ia < deg; ia++) {
            int64_t a = nb[ia];
            for (int64_t j = VV->off[a]; j < VV->off[a + 1]; j++) {
                int64_t c = VV->val[j];
                if (c < 0 || c >= n_points) continue;
                int64_t ic = pos[c];
                if (ic < 0 || ic == ia || low[ic] != low[ia]) continue;
                int64_t ra = uf_find(parent, ia), rc = uf_find(parent, ic);
                if (ra != rc) parent[ra > rc ? ra : rc] = ra < rc ? ra : rc;
            }
        }
        uint32_t nl = 0, nu = 0;
        for (int64_t i = 0; i < deg; i++) {
            if (uf_find(parent, i) == i) {
                if (low[i]) nl++;
                else nu++;
            }
            pos[nb[i]] = -1;
        }
        classes[3 * v] = nl;
        classes[3 * v + 1] = nu;
        classes[3 * v + 2] = verdict(nl, nu);
    }
    free(pos);
    free(nb);
    free(parent);
    free(low);
}

/* Six-way label, numbered as TTK's CriticalType: 0 minimum, 1 1-saddle,
 * 2 2-saddle, 3 maximum, 4 multi-saddle (degenerate), 5 regular. */
void orc_label6(const uint32_t *classes, int64_t n_points, uint8_t *label) {
    for (int64_t v = 0; v < n_points; v++) {
        uint32_t l = classes[3 * v], u = classes[3 * v + 1];
        uint8_t k;
        if (l == 0 && u >= 1) k = 0;
        else if (l >= 1 && u == 0) k = 3;
        else if (l == 1 && u == 1) k = 5;
        else if (l == 2 && u == 1) k = 1;
        else if (l == 1 && u == 2) k = 2;
        else k = 4;
        label[v] = k;
    }
}

int32_t orc_approx_max_vv(const int64_t *cells, int64_t n_cells,
                          int64_t n_points) {
    int32_t *appears = (int32_t *)calloc((size_t)n_points + 1, sizeof(int32_t));
    int32_t mx = 0;
    for (int64_t i = 0; i < 4 * n_cells; i++)
        if (++appears[cells[i]] > mx) mx++;
    free(appears);
    mx *= 3;
    return ((mx + 31) / 32) * 32;
}

void orc_widen(const int32_t *in, int64_t n, int64_t *out) {
    for (int64_t i = 0; i < n; i++) out[i] = in[i];
}

/* ---- OpenMP baseline: VV (ascending rows) + critical points ---------- */
static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

static int64_t bsearch_i64(const int64_t *a, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && a[lo] == key) ? lo : -1;
}

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

double orc_vv_cp_omp(const int32_t *cells, int64_t n_cells, int64_t n_points,
                     const float *scalars, int threads, orc_csr *VV,
                     uint32_t *classes) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    double t0 = now_s();
    int64_t V = n_points, T = n_cells;
    /* vertex -> tets (unordered rows are enough for VV) */
    int64_t *vt_off = (int64_t *)calloc((size_t)V + 1, sizeof(int64_t));
    int64_t *cursor = (int64_t *)calloc((size_t)V + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 4 * T; i++) {
#pragma omp atomic
        vt_off[cells[i] + 1]++;
    }
    for (int64_t v = 0; v < V; v++) vt_off[v + 1] += vt_off[v];
    int32_t *vt = (int32_t *)malloc((size_t)(4 * T + 1) * sizeof(int32_t));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 4 * T; i++) {
        int64_t v = cells[i], slot;
#pragma omp atomic capture
        slot = cursor[v]++;
        vt[vt_off[v] + slot] = (int32_t)(i >> 2);
    }
    free(cursor);
    /* pass 1: degrees */
    int32_t max_vt = 0;
    for (int64_t v = 0; v < V; v++) {
        int64_t d = vt_off[v + 1] - vt_off[v];
        if (d > max_vt) max_vt = (int32_t)d;
    }
    int64_t *deg = (int64_t *)calloc((size_t)V + 1, sizeof(int64_t));
#pragma omp parallel
    {
        int32_t *buf = (int32_t *)malloc((size_t)(3 * max_vt + 1) * sizeof(int32_t));
#pragma omp for schedule(dynamic, 1024)
        for (int64_t v = 0; v < V; v++) {
            int64_t n = 0;
            for (int64_t k = vt_off[v]; k < vt_off[v + 1]; k++) {
                const int32_t *c = cells + 4 * (int64_t)vt[k];
                for (int j = 0; j < 4; j++)
                    if (c[j] != v) buf[n++] = c[j];
            }
            qsort(buf, (size_t)n, sizeof(int32_t), cmp_i32);
            int64_t m = 0;
            for (int64_t i = 0; i < n; i++)
                if (m == 0 || buf[m - 1] != buf[i]) buf[m++] = buf[i];
            deg[v + 1] = m;
        }
        free(buf);
    }
    for (int64_t v = 0; v < V; v++) deg[v + 1] += deg[v];
    VV->rows = V;
    VV->width = 1;
    VV->nnz = deg[V];
    VV->off = deg;
    VV->val = (int64_t *)malloc((size_t)(deg[V] + 1) * sizeof(int64_t));
    /* pass 2: rows */
#pragma omp parallel
    {
        int32_t *buf = (int32_t *)malloc((size_t)(3 * max_vt + 1) * sizeof(int32_t));
#pragma omp for schedule(dynamic, 1024)
        for (int64_t v = 0; v < V; v++) {
            int64_t n = 0;
            for (int64_t k = vt_off[v]; k < vt_off[v + 1]; k++) {
                const int32_t *c = cells + 4 * (int64_t)vt[k];
                for (int j = 0; j < 4; j++)
                    if (c[j] != v) buf[n++] = c[j];
            }
            qsort(buf, (size_t)n, sizeof(int32_t), cmp_i32);
            int64_t m = 0, base = deg[v];
            for (int64_t i = 0; i < n; i++)
                if (m == 0 || buf[m - 1] != buf[i]) buf[m++] = buf[i];
            for (int64_t i = 0; i < m; i++) VV->val[base + i] = buf[i];
        }
        free(buf);
    }
    free(vt);
    free(vt_off);
    /* critical points on the ascending rows */
    int64_t max_deg = 0;
    for (int64_t v = 0; v < V; v++)
        if (deg[v + 1] - deg[v] > max_deg) max_deg = deg[v + 1] - deg[v];
#pragma omp parallel
    {
        int64_t *parent = (int64_t *)malloc((size_t)(max_deg + 1) * sizeof(int64_t));
        char *low = (char *)malloc((size_t)max_deg + 1);
#pragma omp for schedule(dynamic, 1024)
        for (int64_t v = 0; v < V; v++) {
            const int64_t *nb = VV->val + deg[v];
            int64_t d = deg[v + 1] - deg[v];
            double sv = scalars[v];
            for (int64_t i = 0; i < d; i++) {
                parent[i] = i;
                low[i] = (char)is_lower((double)scalars[nb[i]], sv, nb[i], v);
            }
            for (int64_t ia = 0; ia < d; ia++) {
                int64_t a = nb[ia];
                for (int64_t j = deg[a]; j < deg[a + 1]; j++) {
                    int64_t c = VV->val[j];
                    if (c <= a) continue; /* each undirected pair once */
                    int64_t ic = bsearch_i64(nb, d, c);
                    if (ic < 0 || low[ic] != low[ia]) continue;
                    int64_t ra = uf_find(parent, ia), rc = uf_find(parent, ic);
                    if (ra != rc) parent[ra > rc ? ra : rc] = ra < rc ? ra : rc;
                }
            }
            uint32_t nl = 0, nu = 0;
            for (int64_t i = 0; i < d; i++)
                if (uf_find(parent, i) == i) {
                    if (low[i]) nl++;
                    else nu++;
                }
            classes[3 * v] = nl;
            classes[3 * v + 1] = nu;
            classes[3 * v + 2] = verdict(nl, nu);
        }
        free(parent);
        free(low);
    }
    return now_s() - t0;
}
