// Host-vs-device relation checks (src/validate.cpp:40-613 of the reference):
// the same acceptance rules, found with hash lookups instead of linear scans
// over the whole device relation.  Acceptance per relation:
//   EV  a host edge matches a device edge anywhere, in either orientation;
//   TF  a host row matches a device row anywhere after sorting both rows;
//   TE  a host row matches a device row anywhere, or the device row at the
//       same index holds the same six ids in any order;
//   FV  a host triple matches a device triple anywhere, both sorted;
//   TT/FF/EE/VV  every host entry of row i appears in device row i.
// Each prints how many entries matched and failed (at most 10 misses in
// detail) and returns true when nothing failed.
#include "validate.h"

#include <algorithm>
#include <cstdio>
#include <unordered_set>

namespace {

constexpr int MAX_PRINT = 10;

struct ArrHash {
    template <size_t N>
    size_t operator()(const std::array<vtkIdType, N> &a) const {
        size_t h = 1469598103934665603ull;
        for (vtkIdType x : a) h = (h ^ (size_t)x) * 1099511628211ull;
        return h;
    }
};

bool report(const char *what, long long found, long long failed) {
    std::fprintf(stderr, "[validate] %s: matched %lld", what, found);
    if (failed) std::fprintf(stderr, ", FAILED %lld", failed);
    std::fprintf(stderr, "\n");
    return failed == 0;
}

bool size_check(const char *what, size_t host, size_t device) {
    if (host == device) return true;
    std::fprintf(stderr, "[validate] %s: device size %zu != host size %zu\n", what, device, host);
    return false;
}

template <typename Rows>
bool rows_contained(const char *what, const Rows &host, const Rows &device) {
    if (!size_check(what, host.size(), device.size())) return false;
    long long found = 0, failed = 0;
    for (size_t i = 0; i < host.size(); i++) {
        std::unordered_set<vtkIdType> row(device[i].begin(), device[i].end());
        for (vtkIdType x : host[i]) {
            if (row.count(x)) {
                found++;
            } else if (++failed <= MAX_PRINT) {
                std::fprintf(stderr, "[validate] %s: %lld missing from device row %zu\n", what, x,
                             i);
            }
        }
    }
    return report(what, found, failed);
}

}  // namespace

bool check_host_vs_device_EV(const EV_Data &host, const EV_Data &device) {
    if (!size_check("EV", host.size(), device.size())) return false;
    std::unordered_set<std::array<vtkIdType, 2>, ArrHash> dev(device.begin(), device.end());
    long long found = 0, failed = 0, inverted = 0;
    for (const auto &e : host) {
        if (dev.count(e)) {
            found++;
        } else if (dev.count({e[1], e[0]})) {
            found++;
            inverted++;
        } else if (++failed <= MAX_PRINT) {
            std::fprintf(stderr, "[validate] EV: edge (%lld, %lld) not on the device\n", e[0], e[1]);
        }
    }
    if (inverted) std::fprintf(stderr, "[validate] EV: %lld edges matched reversed\n", inverted);
    return report("EV", found, failed);
}

bool check_host_vs_device_TF(TF_Data &host, TF_Data &device) {
    if (!size_check("TF", host.size(), device.size())) return false;
    for (auto &r : host) std::sort(r.begin(), r.end());
    for (auto &r : device) std::sort(r.begin(), r.end());
    std::unordered_set<std::array<vtkIdType, 4>, ArrHash> dev(device.begin(), device.end());
    long long found = 0, failed = 0;
    for (size_t i = 0; i < host.size(); i++) {
        if (dev.count(host[i])) found++;
        else if (++failed <= MAX_PRINT)
            std::fprintf(stderr, "[validate] TF: cell %zu faces not on the device\n", i);
    }
    return report("TF", found, failed);
}

bool check_host_vs_device_TE(const TE_Data &host, const TE_Data &device) {
    if (!size_check("TE", host.size(), device.size())) return false;
    std::unordered_set<std::array<vtkIdType, 6>, ArrHash> dev(device.begin(), device.end());
    long long found = 0, failed = 0;
    for (size_t i = 0; i < host.size(); i++) {
        bool ok = dev.count(host[i]) > 0;
        if (!ok) {
            auto a = host[i], b = device[i];
            std::sort(a.begin(), a.end());
            std::sort(b.begin(), b.end());
            ok = a == b;
        }
        if (ok) found++;
        else if (++failed <= MAX_PRINT)
            std::fprintf(stderr, "[validate] TE: cell %zu edges differ\n", i);
    }
    return report("TE", found, failed);
}

bool check_host_vs_device_FV(FV_Data &host, FV_Data &device) {
    if (!size_check("FV", host.size(), device.size())) return false;
    auto key = [](const FaceData &f) {
        std::array<vtkIdType, 3> k{f.id, f.middleVert, f.highVert};
        std::sort(k.begin(), k.end());
        return k;
    };
    std::unordered_set<std::array<vtkIdType, 3>, ArrHash> dev;
    for (const auto &f : device) dev.insert(key(f));
    long long found = 0, failed = 0;
    for (size_t i = 0; i < host.size(); i++) {
        if (dev.count(key(host[i]))) found++;
        else if (++failed <= MAX_PRINT)
            std::fprintf(stderr, "[validate] FV: face %zu not on the device\n", i);
    }
    return report("FV", found, failed);
}

bool check_host_vs_device_TT(const TT_Data &h, const TT_Data &d) { return rows_contained("TT", h, d); }
bool check_host_vs_device_FF(const FF_Data &h, const FF_Data &d) { return rows_contained("FF", h, d); }
bool check_host_vs_device_EE(const EE_Data &h, const EE_Data &d) { return rows_contained("EE", h, d); }
bool check_host_vs_device_VV(const VV_Data &h, const VV_Data &d) { return rows_contained("VV", h, d); }

template <typename T>
bool check_exact(const char *name, const T &host, const T &device) {
    if (!size_check(name, host.size(), device.size())) return false;
    long long failed = 0;
    for (size_t i = 0; i < host.size(); i++)
        if (!(host[i] == device[i]) && ++failed <= MAX_PRINT)
            std::fprintf(stderr, "[validate] %s: row %zu differs\n", name, i);
    return report(name, (long long)host.size() - failed, failed);
}

template bool check_exact<TE_Data>(const char *, const TE_Data &, const TE_Data &);
template bool check_exact<TF_Data>(const char *, const TF_Data &, const TF_Data &);
template bool check_exact<EV_Data>(const char *, const EV_Data &, const EV_Data &);  // also FT
template bool check_exact<FV_Data>(const char *, const FV_Data &, const FV_Data &);
template bool check_exact<VT_Data>(const char *, const VT_Data &, const VT_Data &);  // all CSR

bool check_exact_VV(const VV_Data &host, const VV_Data &device) {
    VV_Data a(host), b(device);
    for (auto &r : a) std::sort(r.begin(), r.end());
    for (auto &r : b) std::sort(r.begin(), r.end());
    return check_exact("VV", a, b);
}

bool check_exact_classes(const unsigned int *host, const unsigned int *device, vtkIdType n) {
    long long failed = 0;
    for (vtkIdType v = 0; v < n; v++) {
        const unsigned *a = host + 3 * v, *b = device + 3 * v;
        if ((a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) && ++failed <= MAX_PRINT)
            std::fprintf(stderr, "[validate] classes: vertex %lld host (%u,%u,%u) device (%u,%u,%u)\n",
                         v, a[0], a[1], a[2], b[0], b[1], b[2]);
    }
    return report("classes", n - failed, failed);
}
