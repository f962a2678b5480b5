// Test harness for the C++ API (include/tetra/*.h) on the host: prints what
// the CPU builders, argparse, export_classes and the mesh loader produce, as
// JSON or text, for the pytest suite to compare against the oracle.
//
//   dropin_probe cpu <cells.i64> <n_points> [<scalars.f64>]
//   dropin_probe parse <args...>
//   dropin_probe export <classes.u32> <n_points> <debug> [parts=<ids.i32>:<n>] [vertices]
//   dropin_probe load <input> [<arrayname> [<partitioningname>]]
//   dropin_probe part <cells.i64> <n_points> <scalars.f64> <parts.i32> <cpu|gpu>
//                [<partcells.i32>]
//   dropin_probe gpuids <cells.i64> <n_points> <edge_perm.i64> <face_perm.i64> [bad]
//                (GPU builders given host VE/VF renumbered by the permutations;
//                 "bad" moves one VE entry to an edge the mesh does not have)
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "argparse.h"
#include "cpu_extraction.h"
#include "cuda_extraction.h"
#include "critpoints.h"
#include "vtk_load.h"

// HIP runtime entry used to read back the _return builders' device arrays
// (a reference caller does the same with cudaMemcpy); linked from libamdhip64.
extern "C" int hipMemcpy(void *dst, const void *src, size_t bytes, int kind);

namespace {

template <typename T>
std::vector<T> read_file(const char *path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<T> v(b.size() / sizeof(T));
    std::memcpy(v.data(), b.data(), v.size() * sizeof(T));
    return v;
}

void rows(const char *name, const std::vector<std::vector<vtkIdType>> &r, bool comma = true) {
    std::cout << "\"" << name << "\": [";
    for (size_t i = 0; i < r.size(); i++) {
        std::cout << (i ? "," : "") << "[";
        for (size_t j = 0; j < r[i].size(); j++) std::cout << (j ? "," : "") << r[i][j];
        std::cout << "]";
    }
    std::cout << "]" << (comma ? ",\n" : "\n");
}

template <size_t W>
void fixed(const char *name, const std::vector<std::array<vtkIdType, W>> &r) {
    std::cout << "\"" << name << "\": [";
    for (size_t i = 0; i < r.size(); i++) {
        std::cout << (i ? "," : "") << "[";
        for (size_t j = 0; j < W; j++) std::cout << (j ? "," : "") << r[i][j];
        std::cout << "]";
    }
    std::cout << "],\n";
}

int cpu(int argc, char **argv) {
    auto cells = read_file<int64_t>(argv[2]);
    const vtkIdType V = std::atoll(argv[3]);
    TV_Data tv(V, (vtkIdType)cells.size() / 4);
    std::memcpy(tv.cells, cells.data(), cells.size() * sizeof(int64_t));
    TE_Data TE;
    VE_Data VE;
    TF_Data TF;
    VF_Data VF;
    const vtkIdType E = make_TE_and_VE(tv, TE, VE);
    const vtkIdType F = make_TF_and_VF(tv, TF, VF);
    VE_Data VE2;
    VF_Data VF2;
    const bool same = make_VE(tv, VE2) == E && make_VF(tv, VF2) == F;
    std::cout << "{\n\"E\": " << E << ", \"F\": " << F << ", \"same_counts\": " << same << ",\n";
    fixed("TE", TE);
    fixed("TF", TF);
    std::vector<std::vector<vtkIdType>> ve(VE.size()), vf(VF.size());
    for (size_t v = 0; v < VE.size(); v++)
        for (auto &e : VE[v]) ve[v].insert(ve[v].end(), {e.highVert, e.id});
    for (size_t v = 0; v < VF.size(); v++)
        for (auto &f : VF[v]) vf[v].insert(vf[v].end(), {f.middleVert, f.highVert, f.id});
    rows("VE", ve);
    rows("VF", vf);
    auto EV = elective_make_EV(VE, V, E);
    fixed("EV", *EV);
    rows("ET", *elective_make_ET(TE, E));
    auto FV = elective_make_FV(VF, F);
    std::vector<std::array<vtkIdType, 3>> fv(FV->size());
    for (size_t f = 0; f < FV->size(); f++) fv[f] = {(*FV)[f].middleVert, (*FV)[f].highVert, (*FV)[f].id};
    fixed("FV", fv);
    auto FE = elective_make_FE(*FV, VE);
    std::vector<std::array<vtkIdType, 6>> fe(FE->size());
    for (size_t f = 0; f < FE->size(); f++)
        for (int k = 0; k < 3; k++) fe[f][2 * k] = (*FE)[f][k].highVert, fe[f][2 * k + 1] = (*FE)[f][k].id;
    fixed("FE", fe);
    auto FT = elective_make_FT(TF, F);
    if (FT) {
        fixed("FT", *FT);
        rows("FF", *elective_make_FF(TF, *FT));
    }
    rows("EF", *elective_make_EF(*FE, E));
    rows("VV", *elective_make_VV(tv, V));
    rows("VT", *elective_make_VT(tv));
    rows("TT", *elective_make_TT(tv));
    rows("EE", *elective_make_EE(*EV, V));
    std::cout << "\"approx\": 0";
    if (argc > 4) {
        auto s = read_file<double>(argv[4]);
        auto VV = elective_make_VV(tv, V);
        std::vector<unsigned> cls(3 * (size_t)V);
        cpu_critpoints(*VV, s.data(), V, cls.data(), 2);
        std::cout << ",\n\"classes\": [";
        for (size_t i = 0; i < cls.size(); i++) std::cout << (i ? "," : "") << cls[i];
        std::cout << "]";
    }
    std::cout << "\n}\n";
    return 0;
}

int parse_cmd(int argc, char **argv) {
    runtime_arguments args;
    args.debug = NO_DEBUG;
    std::vector<char *> av;
    av.push_back(argv[0]);
    for (int i = 2; i < argc; i++) av.push_back(argv[i]);
    parse((int)av.size(), av.data(), args);
    std::cout << "{\"fileName\": \"" << args.fileName << "\", \"threads\": " << args.threadNumber
              << ", \"max_VV\": " << args.max_VV << ", \"gpus\": " << args.n_GPUS
              << ", \"debug\": " << args.debug << ", \"export\": \"" << args.export_
              << "\", \"arrayname\": \"" << args.arrayname << "\", \"no_partitioning\": "
              << args.no_partitioning << ", \"flags\": \"" << args.arg_flags.to_string()
              << "\"}\n";
    return 0;
}

// export <classes.u32> <n_points> <debug> [parts=<ids.i32>:<n_parts>] [vertices]
int export_cmd(int argc, char **argv) {
    auto cls = read_file<uint32_t>(argv[2]);
    const vtkIdType V = std::atoll(argv[3]);
    TV_Data tv(V, 0);
    runtime_arguments args;
    args.debug = std::atoi(argv[4]);
    for (int i = 5; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "vertices") {
            args.export_vertices = true;
        } else if (a.rfind("parts=", 0) == 0) {
            const size_t colon = a.rfind(':');
            const auto ids = read_file<int32_t>(a.substr(6, colon - 6).c_str());
            tv.n_partitions = std::atoi(a.c_str() + colon + 1);
            tv.partitionIDs = new int[(size_t)V];
            for (vtkIdType v = 0; v < V; v++) tv.partitionIDs[v] = ids[(size_t)v];
        }
    }
    export_classes(cls.data(), tv, args);
    return 0;
}

int load_cmd(int argc, char **argv) {
    runtime_arguments args;
    args.debug = NO_DEBUG;
    args.fileName = argv[2];
    if (argc > 3) args.arrayname = argv[3];
    if (argc > 4) args.partitioningname = argv[4];
    auto tv = get_TV_from_VTK(args);
    std::cout << "{\"n_points\": " << tv->nPoints << ", \"n_cells\": " << tv->nCells
              << ", \"cells\": [";
    for (vtkIdType i = 0; i < 4 * tv->nCells; i++) std::cout << (i ? "," : "") << tv->cells[i];
    std::cout << "], \"scalars\": [";
    if (tv->vertexAttributes)
        for (vtkIdType v = 0; v < tv->nPoints; v++) {
            char buf[40];
            std::snprintf(buf, sizeof buf, "%.17g", tv->vertexAttributes[v]);
            std::cout << (v ? "," : "") << buf;
        }
    std::cout << "], \"n_partitions\": " << tv->n_partitions << ", \"partition_ids\": [";
    for (vtkIdType v = 0; v < tv->nPoints; v++) std::cout << (v ? "," : "") << tv->partitionIDs[v];
    std::cout << "], \"n_per_partition\": [";
    for (int p = 0; p < tv->n_partitions; p++) std::cout << (p ? "," : "") << tv->n_per_partition[p];
    std::cout << "], \"partition_cells\": [";
    if (tv->partitionCells)
        for (vtkIdType i = 0; i < (vtkIdType)tv->n_partitions * tv->nCells; i++)
            std::cout << (i ? "," : "") << tv->partitionCells[i];
    std::cout << "]}\n";
    return 0;
}

// partitioned critical points through the C++ API (host or device builders)
int part_cmd(int argc, char **argv) {
    auto cells = read_file<int64_t>(argv[2]);
    const vtkIdType V = std::atoll(argv[3]);
    auto sc = read_file<double>(argv[4]);
    auto parts = read_file<int32_t>(argv[5]);
    const bool gpu = std::string(argv[6]) == "gpu";
    TV_Data tv(V, (vtkIdType)cells.size() / 4);
    std::memcpy(tv.cells, cells.data(), cells.size() * sizeof(int64_t));
    tv.vertexAttributes = new double[V];
    std::memcpy(tv.vertexAttributes, sc.data(), V * sizeof(double));
    int np = 1;
    for (int32_t p : parts) np = std::max(np, p + 1);
    tv.n_partitions = np;
    tv.partitionIDs = new int[V];
    tv.n_per_partition = new int[np]();
    for (vtkIdType v = 0; v < V; v++) {
        tv.partitionIDs[v] = parts[v];
        tv.n_per_partition[parts[v]]++;
    }
    if (argc > 7) {
        auto pc = read_file<int32_t>(argv[7]);
        tv.partitionCells = new int[pc.size()];
        std::memcpy(tv.partitionCells, pc.data(), pc.size() * sizeof(int32_t));
    }
    runtime_arguments args;
    args.debug = NO_DEBUG;
    unsigned int *cls = partitioned_critical_points(tv, args, gpu);
    std::cout << "{\"classes\": [";
    for (vtkIdType i = 0; i < 3 * V; i++) std::cout << (i ? "," : "") << cls[i];
    std::cout << "]}\n";
    delete[] cls;
    return 0;
}

template <size_t W>
void fixed_ptr(const char *name, const vtkIdType *d, size_t rows) {
    std::vector<std::array<vtkIdType, W>> h(rows);
    if (rows && hipMemcpy(h.data(), d, rows * W * sizeof(vtkIdType), 2 /* device to host */)) {
        std::fprintf(stderr, "hipMemcpy failed\n");
        std::exit(3);
    }
    tg_device_free((void *)d);
    fixed(name, h);
}

int gpuids_cmd(int argc, char **argv) {
    auto cells = read_file<int64_t>(argv[2]);
    const vtkIdType V = std::atoll(argv[3]);
    auto eperm = read_file<int64_t>(argv[4]);
    auto fperm = read_file<int64_t>(argv[5]);
    const bool bad = argc > 6 && std::string(argv[6]) == "bad";
    TV_Data tv(V, (vtkIdType)cells.size() / 4);
    std::memcpy(tv.cells, cells.data(), cells.size() * sizeof(int64_t));
    TE_Data TE0;
    VE_Data VE;
    TF_Data TF0;
    VF_Data VF;
    const vtkIdType E = make_TE_and_VE(tv, TE0, VE);
    const vtkIdType F = make_TF_and_VF(tv, TF0, VF);
    for (auto &row : VE)
        for (auto &e : row) e.id = eperm[e.id];
    for (auto &row : VF)
        for (auto &f : row) f.id = fperm[f.id];
    if (bad)
        for (auto &row : VE)
            if (!row.empty()) {
                row[0].highVert = V + 7;  // an edge no cell holds
                break;
            }
    const vtkIdType T = tv.nCells;
    make_TV_for_GPU(tv);
    make_VE_for_GPU(VE, V, E);  // the reference's call order (src/cuda_extraction.cu:271)
    make_VF_for_GPU(VF, V, F);
    std::cout << "{\n";
    fixed("EV", *make_EV_GPU(VE, V, E, false));
    fixed("TE", *make_TE_GPU(tv, VE, V, E, T, false));
    fixed("TF", *make_TF_GPU(tv, VF, V, F, T, false));
    auto FV = make_FV_GPU(VF, V, F, false);
    std::vector<std::array<vtkIdType, 3>> fv(FV->size());
    for (size_t f = 0; f < FV->size(); f++) fv[f] = {(*FV)[f].middleVert, (*FV)[f].highVert, (*FV)[f].id};
    fixed("FV", fv);
    auto FE = make_FE_GPU(VF, VE, V, E, F, false);
    std::vector<std::array<vtkIdType, 6>> fe(FE->size());
    for (size_t f = 0; f < FE->size(); f++)
        for (int k = 0; k < 3; k++) fe[f][2 * k] = (*FE)[f][k].highVert, fe[f][2 * k + 1] = (*FE)[f][k].id;
    fixed("FE", fe);
    fixed_ptr<2>("EV_dev", make_EV_GPU_return(VE, V, E, false), (size_t)E);
    fixed_ptr<6>("TE_dev", make_TE_GPU_return(tv, VE, V, E, T, false), (size_t)T);
    fixed_ptr<4>("TF_dev", make_TF_GPU_return(tv, VF, V, F, T, false), (size_t)T);
    fixed_ptr<3>("FV_dev", make_FV_GPU_return(VF, V, F, false), (size_t)F);
    fixed_ptr<6>("FE_dev", make_FE_GPU_return(VF, VE, V, E, F, true), (size_t)F);
    std::cout << "\"E\": " << E << ", \"F\": " << F << "\n}\n";
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "cpu") return cpu(argc, argv);
    if (cmd == "parse") return parse_cmd(argc, argv);
    if (cmd == "export") return export_cmd(argc, argv);
    if (cmd == "load") return load_cmd(argc, argv);
    if (cmd == "part") return part_cmd(argc, argv);
    if (cmd == "gpuids") return gpuids_cmd(argc, argv);
    return 2;
}
