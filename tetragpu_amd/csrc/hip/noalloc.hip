// tg_vv_critpoints_noalloc: VV + critical points for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It runs the same bucketed pipeline as
// tg_vv_critpoints on caller-owned inputs and outputs; every intermediate
// (routing histograms, corner records, padded rows, row descriptors, queues,
// counters, scan workspace) lives in one file-scope device array, `scratch`,
// of NOALLOC_SCRATCH_BYTES per device, held until the process ends.
//
// The array's name is only a host-side stand-in; its device address comes
// from hipGetSymbolAddress, once per device with that device current (at
// tg_init, else on the first call), and only that address reaches kernels.
// Calls on one device share the array and keep the last call's contents:
// everything that must start at zero is cleared by fill kernels inside the
// call, and calls that overlap on one device (different streams, no
// ordering between them) race.
#include <algorithm>
#include <cstdio>
#include <mutex>

#include "bucket.h"
#include "tetra_c.h"

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};

int fail(int code, const char *msg) {
    tg::set_error(msg);
    return code;
}

__global__ void k_status(const int32_t *__restrict__ counters, int32_t *__restrict__ status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) status[0] = counters[tg::C_ERR];
}

// Carves the scratch layout of one call; `at` ends as the bytes it needs.
struct Carve {
    size_t at = 0;
    unsigned char *base = nullptr;
    template <typename T>
    T *take(size_t count) {
        T *p = reinterpret_cast<T *>(base ? base + at : nullptr);
        at += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

struct Layout {
    tg::BucketBuffers k;
    size_t bytes = 0;
};

Layout carve(unsigned char *base, const tg::BucketPlan &p, int64_t T, int64_t V) {
    Layout L;
    Carve c;
    c.base = base;
    const size_t n1 = std::max((size_t)p.nc * (size_t)p.nchunks, (size_t)p.nb) + 1,
                 n2 = (size_t)p.nb * p.slices + 1;
    tg::VvCpBuffers &b = L.k.base;
    b.counters = c.take<int32_t>(tg::N_COUNTERS);
    L.k.h1 = c.take<int32_t>(n1);
    L.k.h2 = c.take<int32_t>(n2);
    L.k.rowsrc = c.take<int32_t>((size_t)V);
    b.rowsrc = L.k.rowsrc;
    b.fixd = c.take<int32_t>((size_t)V);
    b.deg = c.take<int32_t>((size_t)V);
    b.heavy = c.take<int32_t>((size_t)V);
    b.fix = c.take<int32_t>((size_t)V);
    b.fixbig = c.take<int32_t>((size_t)V);
    b.rowdesc = c.take<uint8_t>(12 * (size_t)T);
    b.scan_tmp = c.take<unsigned char>(
        tg::scan_tmp_bytes((int64_t)std::max({n1, n2, (size_t)V, (size_t)(12 * T)})));
    // level-1 records; the bucket kernel's padded rows reuse them
    L.k.rec1 = c.take<uint32_t>(12 * (size_t)T);
    b.scratch = reinterpret_cast<int32_t *>(L.k.rec1);
    // level-2 records (the mesh path keeps them in its VV column array)
    L.k.rec2 = c.take<uint32_t>(12 * (size_t)T + 4);
    L.bytes = c.at;
    return L;
}

int device_scratch(void **out) {
    int dev = 0;
    TG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(TG_ERR_ARG, "device index beyond 64");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return tg::hip_fail(e == hipSuccess ? hipErrorInvalidValue : e,
                                "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return TG_OK;
}

}  // namespace

namespace tg {
// Resolves the no-allocation scratch of the current device (tg_init).
int noalloc_init_device() {
    void *p = nullptr;
    return device_scratch(&p);
}
}  // namespace tg

extern "C" {

int tg_noalloc_limits(int64_t n_cells, int64_t n_points, int64_t *need_bytes,
                      int64_t *max_bytes) {
    if (n_cells < 0 || n_points < 0) return fail(TG_ERR_ARG, "negative mesh size");
    const tg::BucketPlan p = tg::plan_buckets(n_cells, std::max<int64_t>(n_points, 1), 7, 4096);
    if (need_bytes) *need_bytes = (int64_t)carve(nullptr, p, n_cells, n_points).bytes;
    if (max_bytes) *max_bytes = (int64_t)NOALLOC_SCRATCH_BYTES;
    return TG_OK;
}

int tg_vv_critpoints_noalloc(const int32_t *cells, int64_t n_cells, int64_t n_points,
                             const void *scalars, int scalar_f64, int32_t *vv_off,
                             int32_t *vv_cols, int64_t vv_cap, uint32_t *classes,
                             int32_t *status, void *stream) {
    if (n_cells < 0 || n_points < 0) return fail(TG_ERR_ARG, "negative mesh size");
    if ((n_cells > 0 && !cells) || !vv_off || (n_points > 0 && (!scalars || !classes)))
        return fail(TG_ERR_ARG, "NULL cells, scalars, vv_off or classes");
    if (4 * n_cells >= (int64_t)INT32_MAX / 3 || n_points >= ((int64_t)1 << 24))
        return fail(TG_ERR_CAPACITY,
                    "tg_vv_critpoints_noalloc takes meshes with 12*cells < 2^31 and fewer "
                    "than 2^24 points");
    // VV holds at most three entries per corner and at most V(V-1) pairs
    const int64_t bound = std::min<int64_t>(12 * n_cells, n_points * (n_points - 1));
    if (vv_cap < bound || (bound > 0 && !vv_cols))
        return fail(TG_ERR_ARG, "vv_cols must hold min(12*n_cells, n_points*(n_points-1)) "
                                "entries (vv_cap too small)");
    hipStream_t s = (hipStream_t)stream;
    void *base = nullptr;
    int rc = device_scratch(&base);
    if (rc) return rc;
    if (n_points == 0) {
        if ((rc = tg::fill_i32(vv_off, 0, 1, s))) return rc;
        if (status) rc = tg::fill_i32(status, 0, 1, s);
        return rc;
    }
    const tg::BucketPlan p = tg::plan_buckets(n_cells, n_points, 7, 4096);
    if (!p.ok) return fail(TG_ERR_CAPACITY, "mesh does not fit the bucketed form");
    Layout L = carve((unsigned char *)base, p, n_cells, n_points);
    if (L.bytes > NOALLOC_SCRATCH_BYTES) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      "tg_vv_critpoints_noalloc: this mesh needs %zu bytes of scratch, the "
                      "entry point holds %zu (about 1.1 million cells at most)",
                      L.bytes, NOALLOC_SCRATCH_BYTES);
        return fail(TG_ERR_CAPACITY, buf);
    }
    tg::BucketBuffers &k = L.k;
    k.plan = p;
    int32_t route = tg::ROUTE_TWO;  // no host wait: graph capture and tight loops
    k.route = &route;
    tg::VvCpBuffers &b = k.base;
    b.tv = cells;
    b.scalars = scalars;
    b.scalars_f64 = scalar_f64 != 0;
    b.n_cells = n_cells;
    b.n_points = n_points;
    b.vv_off = vv_off;
    b.vv_cols = vv_cols;
    b.classes = classes;
    b.marks = nullptr;
    if ((rc = tg::run_vv_cp_bucketed(k, true, s))) return rc;
    if (status) {
        hipLaunchKernelGGL(k_status, dim3(1), dim3(64), 0, s, b.counters, status);
        TG_LAUNCH_CHECK("k_status");
    }
    return TG_OK;
}

}  // extern "C"
