/*
 * cpu_extraction.h -- CPU relation builders with the reference's names and
 * results (include/cpu_extraction.h:8-35 of the reference), used by
 * --validate and by callers that build on the host.
 *
 * Results are identical to the reference's serial builders
 * (src/cpu_extraction.cpp): edge ids by first occurrence over cells and the
 * pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3); face ids by first occurrence of
 * the positional triple of slots (0,1,2),(0,1,3),(0,2,3),(1,2,3); VE/VF rows
 * in ascending id; VV rows in insertion order; VT rows ascending.  The
 * implementation differs: per-vertex hash-free lookups over short rows and
 * OpenMP where the result does not depend on order.
 */
#ifndef TETRA_CPU_EXTRACTION_H
#define TETRA_CPU_EXTRACTION_H

#include "datatypes.h"

vtkIdType make_TE_and_VE(const TV_Data &tv_relationship, TE_Data &cellEdgeList,
                         VE_Data &edgeTable);
vtkIdType make_VE(const TV_Data &tv_relationship, VE_Data &edgeTable);
std::unique_ptr<EV_Data> elective_make_EV(const VE_Data &edgeTable, const vtkIdType n_points,
                                          const vtkIdType n_edges);
std::unique_ptr<ET_Data> elective_make_ET(const TE_Data &cellEdgeList, const vtkIdType n_edges);
std::unique_ptr<FV_Data> elective_make_FV(const VF_Data &VF, const vtkIdType n_faces);
std::unique_ptr<VV_Data> elective_make_VV(const TV_Data &TV, const vtkIdType n_points);
vtkIdType make_TF_and_VF(const TV_Data &tv_relationship, TF_Data &cellFaceList,
                         VF_Data &faceTable);
vtkIdType make_VF(const TV_Data &tv_relationship, VF_Data &faceTable);
std::unique_ptr<VT_Data> elective_make_VT(const TV_Data &tv_relationship);

// Relations the reference names by flag only; this library's definitions
// (rows ascending, self excluded).
std::unique_ptr<FE_Data> elective_make_FE(const FV_Data &FV, const VE_Data &VE);
std::unique_ptr<FT_Data> elective_make_FT(const TF_Data &TF, const vtkIdType n_faces);
std::unique_ptr<EF_Data> elective_make_EF(const FE_Data &FE, const vtkIdType n_edges);
std::unique_ptr<TT_Data> elective_make_TT(const TV_Data &TV);
std::unique_ptr<FF_Data> elective_make_FF(const TF_Data &TF, const FT_Data &FT);
std::unique_ptr<EE_Data> elective_make_EE(const EV_Data &EV, const vtkIdType n_points);

// Critical points on the host (classes[3v..3v+2] = lower components, upper
// components, class 1 min / 2 max / 3 regular / 4 saddle), the numeric
// output of src/alg/critPoints.cu:49-338.  threads <= 0: all cores.
void cpu_critpoints(const VV_Data &VV, const double *scalars, vtkIdType n_points,
                    unsigned int *classes, int threads = 0);

#endif
