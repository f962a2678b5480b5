/*
 * datatypes.h -- host-side data types of the TetraGPU API, for code that
 * switches from the reference to this library without source changes.
 *
 * Same type names, field names and layouts as the reference's
 * include/datatypes.h.in:
 *   - vtkIdType: VTK's default 64-bit id (VTK itself is not used here);
 *   - EdgeData / FaceData (datatypes.h.in:44-62);
 *   - TV_Data with cells[4*nCells], vertexAttributes and the partition
 *     arrays (datatypes.h.in:63-101);
 *   - device_VV / device_VT handles (datatypes.h.in:103-127): here both are
 *     CSR on the device -- device_VV.index holds nPoints+1 uint32 row
 *     offsets instead of per-row counts of a padded [nPoints][max_VV]
 *     array, device_VT.index nPoints+1 uint64 offsets as in the reference;
 *   - the *_Data relation typedefs (datatypes.h.in:129-151);
 *   - runtime_arguments with the same 14-bit flag layout and default
 *     "10000000000001" (datatypes.h.in:154-212).
 */
#ifndef TETRA_DATATYPES_H
#define TETRA_DATATYPES_H

#include <array>
#include <bitset>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

typedef long long vtkIdType;

const int nbVertsInCell = 4;
const int nbVertsInFace = 3;
const int nbVertsInEdge = 2;
const int nbEdgesInCell = 6;
const int nbEdgesInFace = 3;
const int nbFacesInCell = 4;

#define DEBUG_MAX 2
#define DEBUG_MIN 1
#define NO_DEBUG 0

struct EdgeData {
    vtkIdType highVert = 0;  // the edge's larger vertex id
    vtkIdType id = 0;        // edge id
    EdgeData(vtkIdType hv, vtkIdType i) : highVert{hv}, id{i} {}
    EdgeData() = default;
};

struct FaceData {
    vtkIdType middleVert = 0;  // second vertex (by cell position)
    vtkIdType highVert = 0;    // third vertex (by cell position)
    vtkIdType id = 0;          // face id; in FV_Data the first (low) vertex
    FaceData(vtkIdType mv, vtkIdType hv, vtkIdType i) : middleVert{mv}, highVert{hv}, id{i} {}
    FaceData() = default;
    bool operator==(const FaceData &o) const {
        return id == o.id && middleVert == o.middleVert && highVert == o.highVert;
    }
};

struct TV_Data {
    vtkIdType nPoints, nCells;
    int n_partitions = 1;
    vtkIdType *cells;                    // [nCells * 4]
    double *vertexAttributes = nullptr;  // [nPoints] or null
    int *partitionIDs = nullptr;
    int *n_per_partition = nullptr;
    int *partitionCells = nullptr;

    TV_Data(vtkIdType points, vtkIdType n_cells)
        : nPoints(points), nCells(n_cells), cells(new vtkIdType[n_cells * 4]) {}
    ~TV_Data() {
        delete[] cells;
        delete[] vertexAttributes;
        delete[] partitionIDs;
        delete[] n_per_partition;
        delete[] partitionCells;
    }
    TV_Data(const TV_Data &) = delete;
    TV_Data &operator=(const TV_Data &) = delete;
};

typedef struct deviceVV {
    vtkIdType *computed = nullptr;  // CSR values (neighbour ids, ascending rows)
    unsigned int *index = nullptr;  // CSR row offsets [nPoints + 1]
    deviceVV() = default;
    deviceVV(vtkIdType *c, unsigned int *i) : computed(c), index(i) {}
} device_VV;

typedef struct deviceVT {
    vtkIdType *computed = nullptr;            // CSR values (cell ids, ascending rows)
    unsigned long long int *index = nullptr;  // CSR row offsets [nPoints + 1]
    deviceVT() = default;
    deviceVT(vtkIdType *c, unsigned long long int *i) : computed(c), index(i) {}
} device_VT;

typedef std::vector<std::array<vtkIdType, nbEdgesInCell>> TE_Data;
typedef std::vector<std::vector<EdgeData>> VE_Data;
typedef std::vector<std::array<vtkIdType, nbFacesInCell>> TF_Data;
typedef std::vector<std::vector<FaceData>> VF_Data;

typedef std::vector<std::array<vtkIdType, nbVertsInEdge>> EV_Data;
typedef std::vector<std::vector<vtkIdType>> ET_Data;
typedef std::vector<FaceData> FV_Data;
typedef std::vector<std::array<EdgeData, nbEdgesInFace>> FE_Data;
typedef std::vector<std::array<vtkIdType, 2>> FT_Data;
typedef std::vector<std::vector<vtkIdType>> EF_Data;
typedef std::vector<std::vector<vtkIdType>> VT_Data;

typedef std::vector<std::vector<vtkIdType>> VV_Data;
typedef std::vector<std::vector<vtkIdType>> EE_Data;
typedef std::vector<std::vector<vtkIdType>> FF_Data;
typedef std::vector<std::vector<vtkIdType>> TT_Data;

typedef struct argstruct {
    int threadNumber = 1, max_VV = -1, n_GPUS = 0, debug = DEBUG_MAX;
    bool no_partitioning = false;
    // export_classes also writes one "A <vertex> <class>" line per classified
    // vertex (the short form that the reference's pyclassify.py:51-59 reads;
    // the reference's own line for it is commented out, critPoints.cu:396-401)
    bool export_vertices = false;
    std::string fileName = "tiny.vtu", export_ = "", arrayname = "", partitioningname = "",
                thread_logname = "";
    // bit 0 validate, bits 1..13 TE EV ET TF FV FE FT EF VT TT FF EE VV
    std::bitset<14> arg_flags{"10000000000001"};
    std::map<int, std::string> flag_names = {
        {0, "validate"}, {1, "TE"}, {2, "EV"},  {3, "ET"},  {4, "TF"},
        {5, "FV"},       {6, "FE"}, {7, "FT"},  {8, "EF"},  {9, "VT"},
        {10, "TT"},      {11, "FF"}, {12, "EE"}, {13, "VV"}};
    bool validate() const { return arg_flags[0]; }
    bool build_TE() const { return arg_flags[1]; }
    bool build_EV() const { return arg_flags[2]; }
    bool build_ET() const { return arg_flags[3]; }
    bool build_TF() const { return arg_flags[4]; }
    bool build_FV() const { return arg_flags[5]; }
    bool build_FE() const { return arg_flags[6]; }
    bool build_FT() const { return arg_flags[7]; }
    bool build_EF() const { return arg_flags[8]; }
    bool build_VT() const { return arg_flags[9]; }
    bool build_TT() const { return arg_flags[10]; }
    bool build_FF() const { return arg_flags[11]; }
    bool build_EE() const { return arg_flags[12]; }
    bool build_VV() const { return arg_flags[13]; }
} runtime_arguments;

typedef std::map<const char *, std::string> option_map;

#endif
