/*
 * vtk_load.h -- mesh input of the driver (src/vtk_load.cu:46-298 of the
 * reference, which reads .vtu through VTK).  Here without VTK:
 *   - "*.vtu": XML UnstructuredGrid, ASCII or inline base64 binary data
 *     (raw or zlib-compressed); tetrahedral cells; the point array named by
 *     --arrayname (default: the first point array) becomes the scalars;
 *     --partitioningname names the point array of partition ids, re-indexed
 *     as the reference does (src/vtk_load.cu:171-204), and cell arrays
 *     partition_cells_<p> (all of them or none) give per-partition cell
 *     lists; without it every point is in partition 0;
 *   - "kuhn:N[:seed[:scalar]]": a generated Kuhn N^3 grid with seeded
 *     vertex/cell permutations (scalar = uniform, gaussians, sines, linear
 *     or constant).
 * Errors print a message and exit(EXIT_FAILURE), as the reference does.
 */
#ifndef TETRA_VTK_LOAD_H
#define TETRA_VTK_LOAD_H

#include "datatypes.h"

std::shared_ptr<TV_Data> get_TV_from_VTK(const runtime_arguments args);

#endif
