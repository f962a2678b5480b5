// Critical-point driver pieces of the reference (src/alg/critPoints.cu):
// export_classes (:340-453), the single-GPU classification run and the
// partition loop (:654-1003, results merged as at :1438-1491).
#include "critpoints.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <string>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <omp.h>
#include <vector>

#include "cpu_extraction.h"
#include "cuda_extraction.h"
#include "tetra_c.h"

void export_classes(unsigned int *classes, TV_Data &TV, runtime_arguments &args) {
    std::ofstream file;
    std::streambuf *buf = std::cout.rdbuf();
    if (args.export_.empty()) {
        std::cerr << "No export file; outputting classes to stdout\n";
    } else {
        file.open(args.export_);
        buf = file.rdbuf();
        std::cout << "Outputting classes to " << args.export_ << "\n";
    }
    std::ostream out(buf);
    // per partition and numeric class; the reference's labels: "minima" <-
    // class 2, "maxima" <- class 1 (its macro names, critPoints.cu:18-21)
    const bool parts = !args.no_partitioning && TV.partitionIDs;
    const int np = parts ? std::max(TV.n_partitions, 1) : 1;
    // names by class number, as the reference prints them (critPoints.cu:362)
    static const char *const names[5] = {"NULL", "max", "min", "regular", "saddle"};
    // the partition whose non-regular vertices get a line of their own
    // (critPoints.cu:27, DETAILED_PARTITION_INFORMATION; lines at :402-409)
    constexpr int kDetailPartition = 1;
    std::vector<std::array<long long, 5>> n_class((size_t)np, std::array<long long, 5>{});
    long long insane = 0;
    for (vtkIdType v = 0; v < TV.nPoints; v++) {
        const unsigned a = classes[3 * v], b = classes[3 * v + 1], c = classes[3 * v + 2];
        const int p = parts ? TV.partitionIDs[v] : 0;
        // self-consistency of counts and class (critPoints.cu:379-394)
        if ((a >= 1 && b == 0 && c != 2) || (a == 0 && b >= 1 && c != 1) ||
            (a == 1 && b == 1 && c != 3) || (a > 1 && b > 1 && c != 4))
            insane++;
        if (args.export_vertices && c > 0) out << "A " << v << " " << c << "\n";
        if (p == kDetailPartition && c > 0 && c != 3)
            out << "Partition " << kDetailPartition << " has a " << (c <= 4 ? names[c] : "?")
                << " for global vertex ID " << v << " (upper: " << a << ", lower: " << b
                << ")\n";
        if (c <= 4 && p >= 0 && p < np) n_class[(size_t)p][c]++;
    }
    if (insane)
        std::cerr << "Insanity detected; GPU did not agree on its own answers for " << insane
                  << " points.\n";
    std::array<long long, 5> tot{};
    for (const auto &k : n_class)
        for (int c = 0; c < 5; c++) tot[c] += k[c];
    out << "Total number of minima:  " << tot[2] << "\n"
        << "Total number of maxima:  " << tot[1] << "\n"
        << "Total number of saddles: " << tot[4] << "\n"
        << "Total number of regular: " << tot[3] << "\n"
        << "Total number of voids:   " << tot[0] << "\n";
    if (args.debug == DEBUG_MAX) {
        for (int p = 0; p < np; p++) {
            const auto &k = n_class[(size_t)p];
            out << "Partition " << p << " number of minima:  " << k[2] << "\n"
                << "Partition " << p << " number of maxima:  " << k[1] << "\n"
                << "Partition " << p << " number of saddles: " << k[4] << "\n"
                << "Partition " << p << " number of regular: " << k[3] << "\n"
                << "Partition " << p << " number of voids:   " << k[0] << "\n";
        }
    }
}

int64_t device_budget_bytes() {
    const char *env = std::getenv("TETRA_HBM_BUDGET");
    if (env && *env) return std::strtoll(env, nullptr, 10);
    return 0;  // tg_critpoints_split: 90 % of the free device memory
}

namespace {

// int32 copies of TV's cells for the C ABI; *bad = the first entry whose
// id is outside [0, nPoints), else -1.
std::vector<int32_t> cells_i32(const TV_Data &TV, int64_t *bad) {
    std::vector<int32_t> c((size_t)(4 * TV.nCells));
    *bad = -1;
    for (size_t i = 0; i < c.size(); i++) {
        const vtkIdType x = TV.cells[i];
        if (x < 0 || x >= TV.nPoints) {
            *bad = (int64_t)i;
            break;
        }
        c[i] = (int32_t)x;
    }
    return c;
}

std::string bad_cell_message(const TV_Data &TV, int64_t i) {
    return "cell " + std::to_string(i / 4) + " references vertex " +
           std::to_string((long long)TV.cells[i]) + " outside the mesh";
}

}  // namespace

unsigned int *gpu_critical_points(const TV_Data &TV, const runtime_arguments &args) {
    if (!TV.vertexAttributes) {
        std::fprintf(stderr, "tetra: critical points need a scalar point array\n");
        std::exit(EXIT_FAILURE);
    }
    if (!args.no_partitioning && TV.partitionIDs)
        return partitioned_critical_points(TV, args, true);
    // a mesh whose hot-path footprint exceeds the device budget (or one
    // handle's 32-bit limits) runs in parts
    int64_t need = 0, budget = device_budget_bytes();
    if (budget <= 0 && (tg_init(0) != TG_OK || tg_device_budget(&budget) != TG_OK)) {
        std::fprintf(stderr, "tetra: device memory query failed: %s\n", tg_last_error());
        std::exit(EXIT_FAILURE);
    }
    tg_vv_footprint(TV.nCells, TV.nPoints, &need);
    const bool one_handle = 4 * TV.nCells < (int64_t)INT32_MAX / 3;
    if (need > budget || !one_handle) {
        int64_t bad = -1;
        const std::vector<int32_t> cells = cells_i32(TV, &bad);
        if (bad >= 0) {
            std::fprintf(stderr, "tetra: %s\n", bad_cell_message(TV, bad).c_str());
            std::exit(EXIT_FAILURE);
        }
        unsigned int *out = new unsigned int[3 * (size_t)std::max<vtkIdType>(TV.nPoints, 1)];
        tg_split_stats st{};
        const int rc = tg_critpoints_split(cells.data(), TV.nCells, TV.nPoints,
                                           TV.vertexAttributes, 1, nullptr, 0, nullptr, budget,
                                           out, &st);
        if (rc != TG_OK) {
            std::fprintf(stderr, "tetra: critical points in parts failed [%d]: %s\n", rc,
                         tg_last_error());
            std::exit(EXIT_FAILURE);
        }
        if (args.debug > NO_DEBUG)
            std::cout << "Mesh footprint " << need << " bytes exceeds the device budget of "
                      << budget << " bytes: classified in " << st.parts
                      << " parts (largest " << st.max_cells << " cells, " << st.cells_total
                      << " cells in all; localisation " << st.localize_s << " s, classification "
                      << st.classify_s << " s)\n";
        return out;
    }
    make_TV_for_GPU(TV);
    // the mesh made by make_TV_for_GPU is the thread's current handle; the
    // C++ layer exposes it through a VV build, then classes come from the
    // C ABI on the same handle
    extern tg_mesh *tetra_current_mesh();
    tg_mesh *m = tetra_current_mesh();
    int rc = tg_vv_critpoints(m);
    unsigned int *out = new unsigned int[3 * (size_t)TV.nPoints];
    if (rc == TG_OK) rc = tg_classes_copy(m, out);
    if (rc != TG_OK) {
        std::fprintf(stderr, "tetra: critical points failed [%d]: %s\n", rc, tg_last_error());
        std::exit(EXIT_FAILURE);
    }
    return out;
}

// ---- vertex partitions ----------------------------------------------------------
namespace {

struct PartitionError {
    int code;
    std::string msg;
};

[[noreturn]] void part_fail(const std::string &msg, int code = -1) {
    throw PartitionError{code, msg};
}

thread_local std::string g_partition_error;

// Cells of every partition, ascending.  With TV.partitionCells, row p lists
// them (1 = included, src/alg/critPoints.cu:731-748).  Otherwise a cell
// belongs to each partition that owns one of its vertices
// (critPoints.cu:814-838, and the same rule that src/vtk_load.cu:223-234
// precomputes); one pass over the cells.
std::vector<std::vector<int64_t>> cells_per_partition(const TV_Data &TV) {
    const int np = std::max(TV.n_partitions, 1);
    std::vector<std::vector<int64_t>> cells((size_t)np);
    if (TV.partitionCells) {
        for (int p = 0; p < np; p++) {
            const int *row = TV.partitionCells + (size_t)p * (size_t)TV.nCells;
            for (vtkIdType c = 0; c < TV.nCells; c++)
                if (row[c] == 1) cells[(size_t)p].push_back(c);
        }
        return cells;
    }
    // vertex rule: count per (thread chunk, partition), then fill; chunks
    // are contiguous and in thread order, so every list stays ascending
    const int nth = std::max(1, omp_get_max_threads());
    const vtkIdType nc = TV.nCells;
    std::vector<int64_t> cnt((size_t)nth * np, 0);
    int64_t bad_cell = -1;
    auto parts_of = [&](vtkIdType c, int (&seen)[4]) {
        int ns = 0;
        for (int k = 0; k < 4; k++) {
            const vtkIdType v = TV.cells[4 * c + k];
            if (v < 0 || v >= TV.nPoints) return -1;
            const int p = TV.partitionIDs[v];
            if (p < 0 || p >= np) return -2;
            bool dup = false;
            for (int j = 0; j < ns; j++) dup = dup || seen[j] == p;
            if (!dup) seen[ns++] = p;
        }
        return ns;
    };
#pragma omp parallel num_threads(nth)
    {
        const int t = omp_get_thread_num();
        const vtkIdType c0 = nc * t / nth, c1 = nc * (t + 1) / nth;
        int64_t *my = cnt.data() + (size_t)t * np;
        for (vtkIdType c = c0; c < c1; c++) {
            int seen[4];
            const int ns = parts_of(c, seen);
            if (ns < 0) {
#pragma omp critical
                bad_cell = bad_cell < 0 ? c : std::min<int64_t>(bad_cell, c);
                break;
            }
            for (int j = 0; j < ns; j++) my[seen[j]]++;
        }
    }
    if (bad_cell >= 0) {
        for (int k = 0; k < 4; k++) {
            const vtkIdType v = TV.cells[4 * bad_cell + k];
            if (v < 0 || v >= TV.nPoints)
                part_fail("cell " + std::to_string(bad_cell) + " references vertex " +
                          std::to_string(v) + " outside the mesh");
            const int p = TV.partitionIDs[v];
            if (p < 0 || p >= np)
                part_fail("vertex " + std::to_string(v) + " has partition id " +
                          std::to_string(p) + " outside 0.." + std::to_string(np - 1));
        }
    }
    std::vector<int64_t> start((size_t)nth * np);
    for (int p = 0; p < np; p++) {
        int64_t run = 0;
        for (int t = 0; t < nth; t++) {
            start[(size_t)t * np + p] = run;
            run += cnt[(size_t)t * np + p];
        }
        cells[(size_t)p].resize((size_t)run);
    }
#pragma omp parallel num_threads(nth)
    {
        const int t = omp_get_thread_num();
        const vtkIdType c0 = nc * t / nth, c1 = nc * (t + 1) / nth;
        int64_t *pos = start.data() + (size_t)t * np;
        for (vtkIdType c = c0; c < c1; c++) {
            int seen[4];
            const int ns = parts_of(c, seen);
            for (int j = 0; j < ns; j++) cells[(size_t)seen[j]][(size_t)pos[seen[j]]++] = c;
        }
    }
    return cells;
}

}  // namespace

int localize_partition(const TV_Data &TV, const std::vector<int64_t> &cells,
                       LocalMesh &lm, std::vector<int64_t> &g2l) {
    // local ids in ascending global order keep the id tie-break of the lower
    // test unchanged (critPoints.cu:889-905).  Small partitions sort their
    // vertex list; large ones mark the vertices and collect them in one
    // ascending scan of the points.
    const size_t n = (size_t)TV.nPoints;
    lm.vertices.clear();
    if (4 * cells.size() * 24 < n) {
        lm.vertices.reserve(4 * cells.size());
        for (int64_t c : cells)
            for (int k = 0; k < 4; k++) lm.vertices.push_back(TV.cells[4 * c + k]);
        std::sort(lm.vertices.begin(), lm.vertices.end());
        lm.vertices.erase(std::unique(lm.vertices.begin(), lm.vertices.end()),
                          lm.vertices.end());
        if (g2l.size() < n) g2l.resize(n);
    } else {
        std::vector<uint8_t> mark(n, 0);
        const int64_t nc = (int64_t)cells.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nc; i++)
            for (int k = 0; k < 4; k++)
                __atomic_store_n(&mark[(size_t)TV.cells[4 * cells[(size_t)i] + k]], (uint8_t)1,
                                 __ATOMIC_RELAXED);
        // ascending collection: per-block counts, scan, fill
        const int nth = std::max(1, omp_get_max_threads());
        std::vector<size_t> blk((size_t)nth + 1, 0);
#pragma omp parallel num_threads(nth)
        {
            const int t = omp_get_thread_num();
            const size_t v0 = n * t / nth, v1 = n * (t + 1) / nth;
            size_t k = 0;
            for (size_t v = v0; v < v1; v++) k += mark[v];
            blk[(size_t)t + 1] = k;
        }
        for (int t = 0; t < nth; t++) blk[(size_t)t + 1] += blk[(size_t)t];
        lm.vertices.resize(blk[(size_t)nth]);
#pragma omp parallel num_threads(nth)
        {
            const int t = omp_get_thread_num();
            const size_t v0 = n * t / nth, v1 = n * (t + 1) / nth;
            size_t k = blk[(size_t)t];
            for (size_t v = v0; v < v1; v++)
                if (mark[v]) lm.vertices[k++] = (vtkIdType)v;
        }
        if (g2l.size() < n) g2l.resize(n);
    }
    if (lm.vertices.size() > (size_t)INT32_MAX) return -1;
    const int64_t nv = (int64_t)lm.vertices.size(), nc = (int64_t)cells.size();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nv; i++) g2l[(size_t)lm.vertices[(size_t)i]] = i;
    lm.cells.resize(4 * cells.size());
    lm.scalars.resize(lm.vertices.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nc; i++)
        for (int k = 0; k < 4; k++)
            lm.cells[4 * (size_t)i + k] =
                (int32_t)g2l[(size_t)TV.cells[4 * cells[(size_t)i] + k]];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nv; i++)
        lm.scalars[(size_t)i] = TV.vertexAttributes[lm.vertices[(size_t)i]];
    return 0;
}

namespace {

// The loop on the host (the validation path): localisation and the host
// builders per partition; failures throw PartitionError.
std::unique_ptr<unsigned int[]> partition_loop(const TV_Data &TV, const runtime_arguments &args,
                                               PartitionStats *stats) {
    using clock = std::chrono::steady_clock;
    auto secs = [](clock::time_point a, clock::time_point b) {
        return std::chrono::duration<double>(b - a).count();
    };
    PartitionStats st;
    const auto t_begin = clock::now();
    if (!TV.vertexAttributes) part_fail("critical points need a scalar point array");
    if (!TV.partitionIDs) part_fail("partitioned critical points need partition ids");
    const int np = std::max(TV.n_partitions, 1);
    const std::vector<std::vector<int64_t>> cells = cells_per_partition(TV);
    st.localize_s += secs(t_begin, clock::now());
    // largest partitions first (critPoints.cu:658-666): the first one sizes
    // the device buffers that the later ones reuse
    std::vector<std::pair<int64_t, int>> order;
    for (int p = 0; p < np; p++)
        order.emplace_back(TV.n_per_partition ? (int64_t)TV.n_per_partition[p] : 0, p);
    std::sort(order.rbegin(), order.rend());

    std::unique_ptr<unsigned int[]> out(
        new unsigned int[3 * (size_t)std::max<vtkIdType>(TV.nPoints, 1)]());
    LocalMesh lm;
    std::vector<int64_t> g2l;
    std::vector<unsigned int> local;
    for (const auto &e : order) {
        const int p = e.second;
        if (cells[(size_t)p].empty()) continue;
        const auto t0 = clock::now();
        if (localize_partition(TV, cells[(size_t)p], lm, g2l))
            part_fail("partition " + std::to_string(p) + " is too large for 32-bit ids");
        const vtkIdType nv = (vtkIdType)lm.vertices.size(), nc = (vtkIdType)cells[(size_t)p].size();
        const auto t1 = clock::now();
        st.localize_s += secs(t0, t1);
        st.partitions_run++;
        st.max_cells = std::max<int64_t>(st.max_cells, nc);
        st.max_points = std::max<int64_t>(st.max_points, nv);
        st.cells_total += nc;
        if (args.debug > NO_DEBUG)
            std::cout << "works on partition " << p << " with " << nv << " points and " << nc
                      << " cells (" << e.first << " points are within partition)\n";
        local.assign(3 * (size_t)nv, 0u);
        {
            TV_Data sub(nv, nc);
            for (size_t i = 0; i < lm.cells.size(); i++) sub.cells[i] = lm.cells[i];
            auto VV = elective_make_VV(sub, nv);
            cpu_critpoints(*VV, lm.scalars.data(), nv, local.data(), args.threadNumber);
        }
        st.classify_s += secs(t1, clock::now());
        // owned vertices only; a vertex written twice with different values
        // is an error (critPoints.cu:1462-1488)
        for (vtkIdType i = 0; i < nv; i++) {
            const vtkIdType g = lm.vertices[(size_t)i];
            if (TV.partitionIDs[g] != p) continue;
            for (int k = 0; k < 3; k++) {
                const unsigned x = local[3 * (size_t)i + k];
                unsigned &y = out[3 * (size_t)g + k];
                if (x == 0) continue;
                if (y != 0 && y != x)
                    part_fail("double write to point " + std::to_string(g) + " disagrees");
                y = x;
            }
        }
    }
    st.total_s = secs(t_begin, clock::now());
    if (stats) *stats = st;
    return out;
}

// The loop on the device: localisation and classification per partition in
// libtetra_hip (tg_critpoints_split with the partition ids), largest first.
std::unique_ptr<unsigned int[]> partition_loop_device(const TV_Data &TV,
                                                      const runtime_arguments &args,
                                                      PartitionStats *stats) {
    if (!TV.vertexAttributes) part_fail("critical points need a scalar point array");
    if (!TV.partitionIDs) part_fail("partitioned critical points need partition ids");
    const int np = std::max(TV.n_partitions, 1);
    int64_t bad = -1;
    const std::vector<int32_t> cells = cells_i32(TV, &bad);
    if (bad >= 0) part_fail(bad_cell_message(TV, bad));
    std::unique_ptr<unsigned int[]> out(
        new unsigned int[3 * (size_t)std::max<vtkIdType>(TV.nPoints, 1)]());
    tg_split_stats st{};
    const int rc = tg_critpoints_split(cells.data(), TV.nCells, TV.nPoints, TV.vertexAttributes,
                                       1, TV.partitionIDs, np, TV.partitionCells,
                                       device_budget_bytes(), out.get(), &st);
    if (rc != TG_OK)
        part_fail(std::string("partitioned critical points failed: ") + tg_last_error(),
                  rc == TG_ERR_ARG ? -1 : rc);
    if (args.debug > NO_DEBUG)
        std::cout << "classified " << st.parts << " partitions on the device (largest "
                  << st.max_cells << " cells; localisation " << st.localize_s
                  << " s, classification " << st.classify_s << " s)\n";
    if (stats) {
        stats->partitions_run = st.parts;
        stats->max_cells = st.max_cells;
        stats->max_points = st.max_points;
        stats->cells_total = st.cells_total;
        stats->total_s = st.total_s;
        stats->localize_s = st.localize_s;
        stats->classify_s = st.classify_s;
    }
    return out;
}

}  // namespace

unsigned int *partitioned_critical_points(const TV_Data &TV, const runtime_arguments &args,
                                          bool on_gpu, PartitionStats *stats) {
    try {
        return (on_gpu ? partition_loop_device(TV, args, stats)
                       : partition_loop(TV, args, stats))
            .release();
    } catch (const PartitionError &e) {
        std::fprintf(stderr, "tetra: %s\n", e.msg.c_str());
        std::exit(EXIT_FAILURE);
    }
}

extern "C" const char *tetra_partition_error() { return g_partition_error.c_str(); }

extern "C" int tetra_partitioned_classes(const int32_t *cells, int64_t n_cells, int64_t n_points,
                                         const double *scalars, const int32_t *parts,
                                         uint32_t *classes, double *stats) {
    g_partition_error.clear();
    if (n_cells < 0 || n_points < 0 || (n_cells && !cells) || (n_points && (!scalars || !parts)) ||
        !classes) {
        g_partition_error = "bad arguments";
        return -1;
    }
    TV_Data tv(n_points, n_cells);
    for (int64_t i = 0; i < 4 * n_cells; i++) tv.cells[i] = cells[i];
    tv.vertexAttributes = new double[(size_t)std::max<int64_t>(n_points, 1)];
    int np = 1;
    for (int64_t v = 0; v < n_points; v++) {
        if (parts[v] < 0) {
            g_partition_error = "negative partition id at point " + std::to_string(v);
            return -1;
        }
        tv.vertexAttributes[v] = scalars[v];
        np = std::max(np, parts[v] + 1);
    }
    tv.n_partitions = np;
    tv.partitionIDs = new int[(size_t)std::max<int64_t>(n_points, 1)];
    tv.n_per_partition = new int[(size_t)np]();
    for (int64_t v = 0; v < n_points; v++) {
        tv.partitionIDs[v] = parts[v];
        tv.n_per_partition[parts[v]]++;
    }
    runtime_arguments args;
    args.debug = NO_DEBUG;
    PartitionStats st;
    try {
        std::unique_ptr<unsigned int[]> out = partition_loop_device(tv, args, &st);
        std::copy(out.get(), out.get() + 3 * n_points, classes);
    } catch (const PartitionError &e) {
        g_partition_error = e.msg;
        return e.code;
    }
    if (stats) {
        stats[0] = (double)st.partitions_run;
        stats[1] = st.total_s;
        stats[2] = st.localize_s;
        stats[3] = st.classify_s;
        stats[4] = (double)st.max_cells;
        stats[5] = (double)st.max_points;
        stats[6] = (double)st.cells_total;
    }
    return 0;
}
