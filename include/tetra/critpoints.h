/*
 * critpoints.h -- the critical-point driver of the reference
 * (src/alg/critPoints.cu:340-453 export_classes, :1359-1514 main) for one
 * GPU and no partitioning.
 */
#ifndef TETRA_CRITPOINTS_H
#define TETRA_CRITPOINTS_H

#include "datatypes.h"

// Totals per numeric class, printed with the reference's labels: its
// "minima" line counts class 2 and its "maxima" line class 1 (the macros in
// critPoints.cu:18-21 name the classes the other way round from what the
// kernel computes).  Writes to args.export_ or stdout; with debug level 2
// also the per-partition lines.
void export_classes(unsigned int *classes, TV_Data &TV, runtime_arguments &args);

// VV + classes[3 * nPoints] for TV on device 0 (TV.vertexAttributes are the
// scalars).  The returned buffer is owned by the caller (delete[]).
unsigned int *gpu_critical_points(const TV_Data &TV, const runtime_arguments &args);

#endif
