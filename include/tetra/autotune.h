/*
 * autotune.h -- picks the kernel form and launch shape of the VV +
 * critical-point path per mesh size class (no counterpart in the reference,
 * whose launch shapes are fixed).  For each size class it generates a Kuhn
 * grid of that size (seeded, permuted), times every candidate
 * (vertex-lane form: lane_cap x lane_block; bucketed form: bucket size x
 * tets per routing chunk) over `reps` runs after a warm-up, and keeps the
 * fastest.  Every candidate computes identical results, so only time is
 * compared.  The table is installed with tg_set_tuning and can be saved to
 * and loaded from a small JSON file.
 */
#ifndef TETRA_AUTOTUNE_H
#define TETRA_AUTOTUNE_H

#include <string>
#include <vector>

#include "tetra_c.h"

struct TuneCandidateTime {
    int64_t n_cells;
    tg_tune_entry config;
    double ms;  // median over reps
};

// sides: Kuhn grid sides, one size class each (ascending).  The last class
// covers every larger mesh.  Returns the table (ascending max_cells); all
// timings go to `log` when not null.
std::vector<tg_tune_entry> autotune(const std::vector<int> &sides, int reps,
                                    std::vector<TuneCandidateTime> *log = nullptr);
bool save_tuning(const std::string &path, const std::vector<tg_tune_entry> &table);
std::vector<tg_tune_entry> load_tuning(const std::string &path);

extern "C" {
/* C entry points of the above (libtetra_host.so).  tg_autotune writes at
 * most `capacity` entries to `out` and installs the table. */
int tg_autotune(const int32_t *sides, int32_t n_sides, int32_t reps, tg_tune_entry *out,
                int32_t capacity, int32_t *n_out);
int tg_tuning_load(const char *path);
int tg_tuning_save(const char *path);
}

#endif
