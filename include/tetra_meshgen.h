/*
 * tetra_meshgen.h -- deterministic synthetic tetrahedral meshes (C ABI of
 * libtetra_host.so).
 *
 * No reference counterpart: the reference reads .vtu files
 * (src/vtk_load.cu:46) and ships a Python fan generator (make_mesh.py:41-93).
 * These generators produce the workloads the relation and critical-point
 * paths are measured on: Kuhn (Freudenthal) n^3 cube grids with 6 tets per
 * cube, seeded vertex/tet permutations, jittered positions, optional stellar
 * 1->4 splits and several scalar fields.  All randomness comes from
 * xoshiro256** streams seeded through SplitMix64 (S. Vigna, "Further
 * scramblings of Marsaglia's xorshift generators", 2017), so the same seed
 * gives the same mesh on every host.
 */
#ifndef TETRA_MESHGEN_H
#define TETRA_MESHGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* option bits for tg_gen_kuhn */
#define TG_GEN_PERMUTE_VERTICES 0x1u
#define TG_GEN_PERMUTE_CELLS    0x2u
#define TG_GEN_JITTER           0x4u  /* uniform +-0.2 cell position jitter */
#define TG_GEN_ORDER_SHIFT      4      /* within-cell corner order, 2 bits:  */
#define TG_GEN_ORDER_SORTED     0u     /*   ascending vertex id (default)    */
#define TG_GEN_ORDER_PATH       1u     /*   Kuhn path order, unsorted        */
#define TG_GEN_ORDER_SHUFFLED   2u     /*   random per cell                  */
#define TG_GEN_STANDARD (TG_GEN_PERMUTE_VERTICES | TG_GEN_PERMUTE_CELLS | TG_GEN_JITTER)

/* scalar fields */
#define TG_SCALAR_UNIFORM   0  /* k * 2^-24, k uniform in [0, 2^24): exact f32 */
#define TG_SCALAR_GAUSSIANS 1  /* sum of 32 isotropic Gaussians + N(0, 1e-3)  */
#define TG_SCALAR_SINES     2  /* 8 sinusoidal modes, |k| <= 6                */
#define TG_SCALAR_LINEAR    3  /* unpermuted grid index i + (n+1)(j + (n+1)k) */
#define TG_SCALAR_CONSTANT  4  /* 0.5 everywhere (all ties)                   */

typedef struct tg_gen_mesh {
    int64_t n_points;
    int64_t n_cells;
    int32_t *cells;     /* [n_cells][4] vertex ids                        */
    float *positions;   /* [n_points][3]                                  */
    float *scalars;     /* [n_points]                                     */
    int32_t max_vt;     /* largest number of tets on one vertex           */
} tg_gen_mesh;

/* Kuhn n^3 grid, V = (n+1)^3, T = 6 n^3 before splits.
 * split_fraction: share of the original tets split 1->4 at the centroid
 *   (each chosen once).
 * n_hubs / hub_splits: additionally pick n_hubs vertices and split
 *   hub_splits tets around each (re-choosing sub-tets), which raises their
 *   vertex degree by hub_splits.
 * Returns 0, or a negative status on bad arguments / allocation failure. */
int tg_gen_kuhn(int32_t n, uint64_t seed, uint32_t options,
                double split_fraction, int32_t n_hubs, int32_t hub_splits,
                int32_t scalar_kind, tg_gen_mesh **out);

/* Recompute the scalar field of an existing mesh (same seed rules). */
int tg_gen_scalars(tg_gen_mesh *m, int32_t scalar_kind, uint64_t seed);

void tg_gen_free(tg_gen_mesh *m);

#ifdef __cplusplus
}
#endif
#endif
