// Host relation builders with the reference's names and results
// (src/cpu_extraction.cpp), used by --validate and as the host path of the
// C++ API.  Id-assigning builders are sequential because ids are minted in
// cell order; the transposes and adjacencies run rows in parallel.
#include "cpu_extraction.h"

#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <numeric>

namespace {

// corner position pairs in edge order, and face slots (lo, mid, hi)
constexpr int EDGE_PAIR[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr int FACE_SLOT[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

inline const vtkIdType *cell_of(const TV_Data &tv, vtkIdType c) { return tv.cells + 4 * c; }

// Look up `hi` in the low vertex's edge row; mint the next id when absent.
inline vtkIdType edge_lookup_or_add(std::vector<EdgeData> &row, vtkIdType hi, vtkIdType &count) {
    for (const EdgeData &e : row)
        if (e.highVert == hi) return e.id;
    row.emplace_back(hi, count);
    return count++;
}

inline vtkIdType face_lookup_or_add(std::vector<FaceData> &row, vtkIdType mid, vtkIdType hi,
                                    vtkIdType &count) {
    for (const FaceData &f : row)
        if (f.middleVert == mid && f.highVert == hi) return f.id;
    row.emplace_back(mid, hi, count);
    return count++;
}

std::vector<std::vector<vtkIdType>> sorted_copy(const std::vector<std::vector<vtkIdType>> &rows) {
    std::vector<std::vector<vtkIdType>> out(rows);
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long i = 0; i < (long long)out.size(); i++) std::sort(out[i].begin(), out[i].end());
    return out;
}

}  // namespace

vtkIdType make_TE_and_VE(const TV_Data &tv, TE_Data &cellEdgeList, VE_Data &edgeTable) {
    edgeTable.assign(tv.nPoints, {});
    cellEdgeList.assign(tv.nCells, {});
    vtkIdType count = 0;
    for (vtkIdType c = 0; c < tv.nCells; c++) {
        const vtkIdType *q = cell_of(tv, c);
        for (int k = 0; k < 6; k++) {
            vtkIdType a = q[EDGE_PAIR[k][0]], b = q[EDGE_PAIR[k][1]];
            if (a > b) std::swap(a, b);
            cellEdgeList[c][k] = edge_lookup_or_add(edgeTable[a], b, count);
        }
    }
    return count;
}

vtkIdType make_VE(const TV_Data &tv, VE_Data &edgeTable) {
    edgeTable.assign(tv.nPoints, {});
    vtkIdType count = 0;
    for (vtkIdType c = 0; c < tv.nCells; c++) {
        const vtkIdType *q = cell_of(tv, c);
        for (int k = 0; k < 6; k++) {
            vtkIdType a = q[EDGE_PAIR[k][0]], b = q[EDGE_PAIR[k][1]];
            if (a > b) std::swap(a, b);
            edge_lookup_or_add(edgeTable[a], b, count);
        }
    }
    return count;
}

std::unique_ptr<EV_Data> elective_make_EV(const VE_Data &edgeTable, const vtkIdType n_points,
                                          const vtkIdType n_edges) {
    auto ev = std::make_unique<EV_Data>(n_edges);
    for (vtkIdType v = 0; v < n_points && v < (vtkIdType)edgeTable.size(); v++)
        for (const EdgeData &e : edgeTable[v]) (*ev)[e.id] = {v, e.highVert};
    return ev;
}

std::unique_ptr<ET_Data> elective_make_ET(const TE_Data &cellEdgeList, const vtkIdType n_edges) {
    auto et = std::make_unique<ET_Data>(n_edges);
    for (size_t c = 0; c < cellEdgeList.size(); c++)
        for (vtkIdType e : cellEdgeList[c]) (*et)[e].push_back((vtkIdType)c);
    return et;
}

std::unique_ptr<FV_Data> elective_make_FV(const VF_Data &VF, const vtkIdType n_faces) {
    auto fv = std::make_unique<FV_Data>(n_faces);
    for (size_t v = 0; v < VF.size(); v++)
        for (const FaceData &f : VF[v]) (*fv)[f.id] = FaceData(f.middleVert, f.highVert, (vtkIdType)v);
    return fv;
}

std::unique_ptr<VV_Data> elective_make_VV(const TV_Data &tv, const vtkIdType n_points) {
    // insertion order: cells ascending, other corners in position order
    auto vv = std::make_unique<VV_Data>(n_points);
    for (vtkIdType c = 0; c < tv.nCells; c++) {
        const vtkIdType *q = cell_of(tv, c);
        for (int i = 0; i < 4; i++) {
            std::vector<vtkIdType> &row = (*vv)[q[i]];
            for (int j = 0; j < 4; j++) {
                if (i == j) continue;
                if (std::find(row.begin(), row.end(), q[j]) == row.end()) row.push_back(q[j]);
            }
        }
    }
    return vv;
}

vtkIdType make_TF_and_VF(const TV_Data &tv, TF_Data &cellFaceList, VF_Data &faceTable) {
    faceTable.assign(tv.nPoints, {});
    cellFaceList.assign(tv.nCells, {});
    vtkIdType count = 0;
    for (vtkIdType c = 0; c < tv.nCells; c++) {
        const vtkIdType *q = cell_of(tv, c);
        for (int s = 0; s < 4; s++)
            cellFaceList[c][s] = face_lookup_or_add(faceTable[q[FACE_SLOT[s][0]]],
                                                    q[FACE_SLOT[s][1]], q[FACE_SLOT[s][2]], count);
    }
    return count;
}

vtkIdType make_VF(const TV_Data &tv, VF_Data &faceTable) {
    TF_Data unused;
    return make_TF_and_VF(tv, unused, faceTable);
}

std::unique_ptr<VT_Data> elective_make_VT(const TV_Data &tv) {
    auto vt = std::make_unique<VT_Data>(tv.nPoints);
    for (vtkIdType c = 0; c < tv.nCells; c++) {
        const vtkIdType *q = cell_of(tv, c);
        for (int i = 0; i < 4; i++) (*vt)[q[i]].push_back(c);
    }
    return vt;
}

std::unique_ptr<FE_Data> elective_make_FE(const FV_Data &FV, const VE_Data &VE) {
    auto fe = std::make_unique<FE_Data>(FV.size());
    auto edge_id = [&](vtkIdType a, vtkIdType b) -> vtkIdType {
        if (a > b) std::swap(a, b);
        for (const EdgeData &e : VE[a])
            if (e.highVert == b) return e.id;
        return -1;
    };
#pragma omp parallel for schedule(static)
    for (long long f = 0; f < (long long)FV.size(); f++) {
        const vtkIdType lo = FV[f].id, mid = FV[f].middleVert, hi = FV[f].highVert;
        const vtkIdType p[3][2] = {{lo, mid}, {lo, hi}, {mid, hi}};
        for (int k = 0; k < 3; k++)
            (*fe)[f][k] = EdgeData(std::max(p[k][0], p[k][1]), edge_id(p[k][0], p[k][1]));
    }
    return fe;
}

std::unique_ptr<FT_Data> elective_make_FT(const TF_Data &TF, const vtkIdType n_faces) {
    auto ft = std::make_unique<FT_Data>(n_faces, std::array<vtkIdType, 2>{-1, -1});
    for (size_t c = 0; c < TF.size(); c++)
        for (vtkIdType f : TF[c]) {
            auto &h = (*ft)[f];
            if (h[0] == (vtkIdType)c || h[1] == (vtkIdType)c) continue;
            if (h[0] < 0) h[0] = (vtkIdType)c;
            else if (h[1] < 0) h[1] = (vtkIdType)c;
            else return nullptr;  // a face held by more than two cells
        }
    return ft;
}

std::unique_ptr<EF_Data> elective_make_EF(const FE_Data &FE, const vtkIdType n_edges) {
    auto ef = std::make_unique<EF_Data>(n_edges);
    for (size_t f = 0; f < FE.size(); f++)
        for (const EdgeData &e : FE[f]) {
            auto &row = (*ef)[e.id];
            if (row.empty() || row.back() != (vtkIdType)f) row.push_back((vtkIdType)f);
        }
    return ef;
}

std::unique_ptr<TT_Data> elective_make_TT(const TV_Data &tv) {
    auto vt = elective_make_VT(tv);
    auto tt = std::make_unique<TT_Data>(tv.nCells);
#pragma omp parallel for schedule(dynamic, 4096)
    for (long long c = 0; c < (long long)tv.nCells; c++) {
        std::vector<vtkIdType> &row = (*tt)[c];
        const vtkIdType *q = cell_of(tv, c);
        for (int i = 0; i < 4; i++) row.insert(row.end(), (*vt)[q[i]].begin(), (*vt)[q[i]].end());
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        row.erase(std::remove(row.begin(), row.end(), (vtkIdType)c), row.end());
    }
    return tt;
}

std::unique_ptr<FF_Data> elective_make_FF(const TF_Data &TF, const FT_Data &FT) {
    auto ff = std::make_unique<FF_Data>(FT.size());
#pragma omp parallel for schedule(dynamic, 4096)
    for (long long f = 0; f < (long long)FT.size(); f++) {
        std::vector<vtkIdType> &row = (*ff)[f];
        for (vtkIdType c : FT[f])
            if (c >= 0) row.insert(row.end(), TF[c].begin(), TF[c].end());
        std::sort(row.begin(), row.end());
        row.erase(std::unique(row.begin(), row.end()), row.end());
        row.erase(std::remove(row.begin(), row.end(), (vtkIdType)f), row.end());
    }
    return ff;
}

std::unique_ptr<EE_Data> elective_make_EE(const EV_Data &EV, const vtkIdType n_points) {
    std::vector<std::vector<vtkIdType>> inc(n_points);
    for (size_t e = 0; e < EV.size(); e++) {
        inc[EV[e][0]].push_back((vtkIdType)e);
        inc[EV[e][1]].push_back((vtkIdType)e);
    }
    auto ee = std::make_unique<EE_Data>(EV.size());
#pragma omp parallel for schedule(dynamic, 4096)
    for (long long e = 0; e < (long long)EV.size(); e++) {
        const auto &a = inc[EV[e][0]], &b = inc[EV[e][1]];
        std::vector<vtkIdType> &row = (*ee)[e];
        row.reserve(a.size() + b.size());
        std::merge(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(row));
        row.erase(std::remove(row.begin(), row.end(), (vtkIdType)e), row.end());
    }
    return ee;
}

void cpu_critpoints(const VV_Data &VV, const double *s, vtkIdType n_points,
                    unsigned int *classes, int threads) {
    const auto rows = sorted_copy(VV);
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
    {
        std::vector<int> parent;
        std::vector<char> lower;
#pragma omp for schedule(dynamic, 1024)
        for (long long v = 0; v < (long long)n_points; v++) {
            const std::vector<vtkIdType> &nb = rows[v];
            const int d = (int)nb.size();
            parent.resize(d);
            lower.resize(d);
            for (int i = 0; i < d; i++) {
                const vtkIdType a = nb[i];
                parent[i] = i;
                // lower iff s[a] < s[v], ties broken by id; NaN compares as upper
                lower[i] = (s[a] == s[v]) ? (a < v) : (s[a] < s[v]);
            }
            auto find = [&](int x) {
                while (parent[x] != x) x = parent[x] = parent[parent[x]];
                return x;
            };
            for (int i = 0; i < d; i++) {
                // link edge (nb[i], c) for every c of row(nb[i]) inside row(v)
                const std::vector<vtkIdType> &r2 = rows[nb[i]];
                size_t p = 0, q = 0;
                while (p < r2.size() && q < (size_t)d) {
                    if (r2[p] < nb[q]) p++;
                    else if (nb[q] < r2[p]) q++;
                    else {
                        if (lower[i] == lower[q]) {
                            const int x = find(i), y = find((int)q);
                            if (x != y) parent[std::max(x, y)] = std::min(x, y);
                        }
                        p++;
                        q++;
                    }
                }
            }
            unsigned nl = 0, nu = 0;
            for (int i = 0; i < d; i++)
                if (find(i) == i) (lower[i] ? nl : nu)++;
            unsigned cls = (nl >= 1 && nu == 0) ? 2u : (nl == 0 && nu >= 1) ? 1u
                         : (nl == 1 && nu == 1) ? 3u : 4u;
            classes[3 * v] = nl;
            classes[3 * v + 1] = nu;
            classes[3 * v + 2] = cls;
        }
    }
}
