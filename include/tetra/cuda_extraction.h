/*
 * cuda_extraction.h -- the reference's GPU relation builders, implemented on
 * MI355X over the C ABI of libtetra_hip.so (include/tetra_c.h).  The header
 * keeps the reference's name so that callers compile unchanged; nothing in
 * it is CUDA.
 *
 * Declarations follow include/cuda_extraction.h:10-138 of the reference
 * (kernels excluded: they are not host API).  Semantics:
 *
 *  - make_TV_for_GPU uploads TV (ids narrowed to int32 on the device) into
 *    the calling thread's current mesh; make_VE_for_GPU / make_VF_for_GPU
 *    record the host tables' sizes (the device derives the same ids from
 *    TV, in the same first-occurrence order) and fail if they disagree.
 *    A second make_TV_for_GPU without free_transients replaces the mesh.
 *  - make_X_GPU returns host structures the caller owns; make_X_GPU_return
 *    returns a device vtkIdType* array the caller owns and releases with
 *    tg_device_free (the reference used cudaFree).  Edge and face ids are
 *    those of the CPU builders (src/cpu_extraction.cpp), TE/TF rows in the
 *    CPU slot order.
 *  - make_VV_GPU(_return): VV as CSR with ascending rows (device_VV.index
 *    holds nPoints+1 offsets); make_VT_GPU_return: CSR with uint64 offsets.
 *  - free_transients = true releases the thread's current mesh afterwards.
 *  - errors print the library's message and exit(EXIT_FAILURE), as the
 *    reference's CUDA_ASSERT does (include/cuda_safety.h:19-23).
 */
#ifndef TETRA_CUDA_EXTRACTION_H
#define TETRA_CUDA_EXTRACTION_H

#include "cpu_extraction.h"
#include "datatypes.h"
#include "tetra_c.h"  // tg_device_free for the _return results

void make_TV_for_GPU(const TV_Data &tv_relationship);
void make_VE_for_GPU(const VE_Data &ve_relationship, const vtkIdType n_edges,
                     const vtkIdType n_verts);
void make_VF_for_GPU(const VF_Data &vf_relationship, const vtkIdType n_verts,
                     const vtkIdType n_faces);

// EV = VE'
vtkIdType *make_EV_GPU_return(const VE_Data &edgeTable, const vtkIdType n_points,
                              const vtkIdType n_edges, const bool free_transients);
std::unique_ptr<EV_Data> make_EV_GPU(const VE_Data &edgeTable, const vtkIdType n_points,
                                     const vtkIdType n_edges, const bool free_transients);
// TF = TV x VF
vtkIdType *make_TF_GPU_return(const TV_Data &TV, const VF_Data &VF, const vtkIdType n_points,
                              const vtkIdType n_faces, const vtkIdType n_cells,
                              const bool free_transients);
std::unique_ptr<TF_Data> make_TF_GPU(const TV_Data &TV, const VF_Data &VF,
                                     const vtkIdType n_points, const vtkIdType n_faces,
                                     const vtkIdType n_cells, const bool free_transients);
// TE = TV x VE
vtkIdType *make_TE_GPU_return(const TV_Data &TV, const VE_Data &VE, const vtkIdType n_points,
                              const vtkIdType n_edges, const vtkIdType n_cells,
                              const bool free_transients);
std::unique_ptr<TE_Data> make_TE_GPU(const TV_Data &TV, const VE_Data &VE,
                                     const vtkIdType n_points, const vtkIdType n_edges,
                                     const vtkIdType n_cells, const bool free_transients);
// FV = VF'
vtkIdType *make_FV_GPU_return(const VF_Data &VF, const vtkIdType n_points,
                              const vtkIdType n_faces, const bool free_transients);
std::unique_ptr<FV_Data> make_FV_GPU(const VF_Data &VF, const vtkIdType n_points,
                                     const vtkIdType n_faces, const bool free_transients);
// FE = VF' x VE
vtkIdType *make_FE_GPU_return(const VF_Data &VF, const VE_Data &VE, const vtkIdType n_points,
                              const vtkIdType n_edges, const vtkIdType n_faces,
                              const bool free_transients);
std::unique_ptr<FE_Data> make_FE_GPU(const VF_Data &VF, const VE_Data &VE,
                                     const vtkIdType n_points, const vtkIdType n_edges,
                                     const vtkIdType n_faces, const bool free_transients);
// VV = TV' x TV
int get_approx_max_VV_partitioned(const TV_Data &TV, const vtkIdType n_points,
                                  const vtkIdType *ivvi, const int debug = NO_DEBUG);
int get_approx_max_VV(const TV_Data &TV, const vtkIdType n_points, const int debug = NO_DEBUG);
device_VV *make_VV_GPU_return(const TV_Data &TV, const int n_cells, const int n_points,
                              const int max_VV_guess, const bool free_transients);
std::unique_ptr<VV_Data> make_VV_GPU(const TV_Data &TV, const int n_cells, const int n_points,
                                     const bool free_transients);
// VT = TV'
device_VT *make_VT_GPU_return(const TV_Data &TV);
device_VT *make_VT_GPU_return(const VT_Data &VT);

// Relations the reference names by flag only (datatypes.h.in:173-178):
// same conventions, definitions of this library (rows ascending, self
// excluded; FT padded with -1).
std::unique_ptr<ET_Data> make_ET_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<FT_Data> make_FT_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<EF_Data> make_EF_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<TT_Data> make_TT_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<FF_Data> make_FF_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<EE_Data> make_EE_GPU(const TV_Data &TV, const bool free_transients);
std::unique_ptr<VT_Data> make_VT_GPU(const TV_Data &TV, const bool free_transients);

// Release the calling thread's current mesh (what free_transients does).
void free_transients_GPU();

// Device->host transfer and host translation times of the builders that
// return host structures.  The reference's builders time the same two
// intervals with their own timers and print them (src/cuda_extraction.cu:
// 326-336, 1250-1265: "GPU Device->Host transfer", "GPU Device->Host
// translation"); here they are recorded per calling thread and printed to
// stdout in that form only when reporting is on (off by default).
struct HostTiming {
    std::string timer;  // "VV_GPU", "TE_GPU", ...
    std::string label;  // "GPU Device->Host transfer" / "... translation"
    double seconds;
};
void set_host_timing_report(bool on);
std::vector<HostTiming> take_host_timings();  // this thread's, since the last call (latest 1024 at most)

#endif
