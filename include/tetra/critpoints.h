/*
 * critpoints.h -- the critical-point driver of the reference
 * (src/alg/critPoints.cu:340-453 export_classes, :1359-1514 main) for one
 * GPU, with the reference's partition loop (:654-1003) run one partition
 * after another on that GPU.
 */
#ifndef TETRA_CRITPOINTS_H
#define TETRA_CRITPOINTS_H

#include <cstdint>
#include <vector>

#include "datatypes.h"

// Totals per numeric class, printed with the reference's labels: its
// "minima" line counts class 2 and its "maxima" line class 1 (the macros in
// critPoints.cu:18-21 name the classes the other way round from what the
// kernel computes).  Writes to args.export_ or stdout; with debug level 2
// also the per-partition lines (partition ids from TV.partitionIDs unless
// args.no_partitioning).
void export_classes(unsigned int *classes, TV_Data &TV, runtime_arguments &args);

// VV + classes[3 * nPoints] for TV on device 0 (TV.vertexAttributes are the
// scalars).  With partitions (TV.partitionIDs set and not
// args.no_partitioning) this is partitioned_critical_points on the GPU.
// The returned buffer is owned by the caller (delete[]).
unsigned int *gpu_critical_points(const TV_Data &TV, const runtime_arguments &args);

// One partition's mesh: the cells it includes, with vertices renumbered in
// ascending global order (critPoints.cu:846-913).
struct LocalMesh {
    std::vector<vtkIdType> vertices;  // local -> global id, ascending
    std::vector<int32_t> cells;       // [4 * cells], local ids
    std::vector<double> scalars;      // [vertices]
};

// Fill lm from the given cells of TV (g2l: scratch, global -> local id).
// Returns -1 if the partition has more than 2^31 - 1 vertices.
int localize_partition(const TV_Data &TV, const std::vector<int64_t> &cells, LocalMesh &lm,
                       std::vector<int64_t> &g2l);

// The reference's partition loop (critPoints.cu:654-1003, 1438-1491), one
// partition at a time, largest first.
// - Partition p is classified on its cells: TV.partitionCells row p when
//   present, else every cell with at least one vertex in p.
// - Only the vertices of p are kept. Classes come from that local mesh's VV
//   rows, so a link edge between two neighbours counts only if a cell of
//   the partition holds it.
// - A vertex in no cell of its partition stays {0, 0, 0} (void).
// - The device holds one partition's mesh at a time, so a mesh larger than
//   HBM runs in pieces.
// - on_gpu = false runs the same loop with the host builders (validation).
// The returned buffer [3 * nPoints] is owned by the caller (delete[]).
struct PartitionStats {
    int64_t partitions_run = 0;            // partitions with at least one cell
    int64_t max_cells = 0, max_points = 0; // largest local mesh
    int64_t cells_total = 0;               // cells over all local meshes (halos counted again)
    double total_s = 0, localize_s = 0;    // wall time: whole loop, host localization
    double classify_s = 0;                 // upload + VV + classes + copy back, per partition
};
unsigned int *partitioned_critical_points(const TV_Data &TV, const runtime_arguments &args,
                                          bool on_gpu, PartitionStats *stats = nullptr);

// The same loop on the device for callers without the C++ types: int32
// cells, float64 scalars, partition ids 0.. per point.  stats (optional,
// 7 doubles) receives partitions_run, total_s, localize_s, classify_s,
// max_cells, max_points, cells_total.  Returns 0; -1 for bad arguments or
// ids; else the failing tg_* status.  tetra_partition_error() gives the
// message of the thread's last failure.
extern "C" int tetra_partitioned_classes(const int32_t *cells, int64_t n_cells, int64_t n_points,
                                         const double *scalars, const int32_t *parts,
                                         uint32_t *classes, double *stats);
extern "C" const char *tetra_partition_error();

#endif
