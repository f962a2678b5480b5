// Request loop with continuous batching for VV + critical-point jobs
// (include/tetra/serve.h).  Host C++ over the C ABI: worker threads are the
// slots, each with its own HIP stream; admission is FIFO by an HBM budget.
#include "serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "tetra_c.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a request can hold at most: TV, scalars, the hot path's
// record/row buffers (12T words twice, 12T descriptor bytes), classes and
// per-vertex arrays; VT corners and values when VT/TT follow; TT at <= 80
// neighbours per cell.
int64_t footprint(int64_t T, int64_t V, uint32_t followups) {
    int64_t b = 16 * T + 8 * V + 96 * T + 12 * T + 12 * V + 40 * V + (1 << 20);
    if (followups) b += 36 * T + 8 * V;
    if (followups & TG_FLAG(TG_REL_TT)) b += 8 * T + 4 * 80 * T;
    return b;
}

struct Group;

// Result storage in pinned host memory from the library's reusing cache, so
// large relations come back at full PCIe rate without page-locking per call.
template <typename T>
struct Pinned {
    T *p = nullptr;
    size_t n = 0, cap = 0;
    Pinned() = default;
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
    ~Pinned() { tg_host_free(p); }
    bool resize(size_t k) {
        n = k;
        if (k <= cap) return true;
        tg_host_free(p);
        void *q = nullptr;
        if (tg_host_alloc(std::max<size_t>(k, 1) * sizeof(T), &q) != TG_OK) {
            p = nullptr;
            n = cap = 0;
            return false;
        }
        p = (T *)q;
        cap = std::max<size_t>(k, 1);
        return true;
    }
    T *data() { return p; }
    const T *data() const { return p; }
    T back() const { return p[n - 1]; }
};

struct Request {
    uint64_t id = 0;
    const int32_t *cells = nullptr;
    const float *scalars = nullptr;
    int64_t T = 0, V = 0;
    uint32_t followups = 0;
    int64_t bytes = 0;
    Clock::time_point submitted, started, finished;
    double upload_ms = 0, compute_ms = 0, followup_ms = 0, copy_ms = 0;
    bool done = false;
    int status = TG_OK;
    std::shared_ptr<Group> group;
    int32_t index = 0;       // position in the group
    int32_t group_size = 1;  // requests served by the same launch
    Pinned<uint32_t> classes;
    Pinned<int64_t> vt_off, tt_off;
    Pinned<int32_t> vt_val, tt_val;
    int64_t vt_nnz = 0, tt_nnz = 0;
    bool has_vt = false, has_tt = false;
};

// Device state shared by the requests served in one launch: a single mesh
// or a batch (union) of meshes.
struct Group {
    std::mutex mu;  // serialises follow-ups on the device state
    tg_mesh *mesh = nullptr;
    tg_batch *batch = nullptr;
    void *stream = nullptr;
    int64_t bytes = 0;
    int refs = 0;
    uint32_t built = 0;  // relations built for follow-ups
    ~Group() {
        if (batch) tg_batch_destroy(batch);
        if (mesh) tg_mesh_destroy(mesh);
    }
};

constexpr uint32_t kRelFlags = TG_FLAG(TG_REL_VT) | TG_FLAG(TG_REL_TT);

int copy_csr(tg_mesh *m, int rel, Pinned<int64_t> &off, Pinned<int32_t> &val) {
    int64_t rows = 0, nnz = 0;
    int32_t width = 0, fixed = 0;
    int rc = tg_relation_shape(m, rel, &rows, &nnz, &width, &fixed);
    if (rc) return rc;
    if (!off.resize((size_t)rows + 1) || !val.resize((size_t)(nnz * width))) return TG_ERR_ALLOC;
    return tg_relation_copy(m, rel, off.data(), val.data());
}

int copy_batch_csr(tg_batch *b, int32_t mesh, int rel, Pinned<int64_t> &off,
                   Pinned<int32_t> &val, int64_t rows) {
    int64_t nnz = 0;
    int rc = tg_batch_relation_copy(b, mesh, rel, nullptr, nullptr, &nnz);
    if (rc) return rc;
    if (!off.resize((size_t)rows + 1) || !val.resize((size_t)nnz)) return TG_ERR_ALLOC;
    return tg_batch_relation_copy(b, mesh, rel, off.data(), val.data(), &nnz);
}

// Entries of a resident relation for one request.
int resident_nnz(Group &g, const Request &r, int rel, int64_t *nnz) {
    if (g.batch) return tg_batch_relation_copy(g.batch, r.index, rel, nullptr, nullptr, nnz);
    int64_t rows = 0;
    int32_t width = 0, fixed = 0;
    return tg_relation_shape(g.mesh, rel, &rows, nnz, &width, &fixed);
}

// VT / TT of one request on its group's resident state (group mutex held):
// built on the device once per group; copied to the request's host arrays
// only with TG_SERVE_COPY.
int followup_locked(Group &g, Request &r, uint32_t flags) {
    const auto t0 = Clock::now();
    const uint32_t rels = flags & kRelFlags;
    const uint32_t need = rels & ~g.built;
    int rc = TG_OK;
    if (need) rc = g.batch ? tg_batch_build(g.batch, need) : tg_build(g.mesh, need);
    if (rc) return rc;
    g.built |= need;
    // counts (they synchronise the group's stream, so the build time is whole)
    if ((rels & TG_FLAG(TG_REL_VT)) && (rc = resident_nnz(g, r, TG_REL_VT, &r.vt_nnz))) return rc;
    if ((rels & TG_FLAG(TG_REL_TT)) && (rc = resident_nnz(g, r, TG_REL_TT, &r.tt_nnz))) return rc;
    const auto t1 = Clock::now();
    r.followup_ms = ms_between(t0, t1);
    r.copy_ms = 0;
    if (!(flags & TG_SERVE_COPY)) return TG_OK;
    if ((rels & TG_FLAG(TG_REL_VT)) && !r.has_vt) {
        rc = g.batch ? copy_batch_csr(g.batch, r.index, TG_REL_VT, r.vt_off, r.vt_val, r.V)
                     : copy_csr(g.mesh, TG_REL_VT, r.vt_off, r.vt_val);
        if (rc) return rc;
        r.has_vt = true;
    }
    if ((rels & TG_FLAG(TG_REL_TT)) && !r.has_tt) {
        rc = g.batch ? copy_batch_csr(g.batch, r.index, TG_REL_TT, r.tt_off, r.tt_val, r.T)
                     : copy_csr(g.mesh, TG_REL_TT, r.tt_off, r.tt_val);
        if (rc) return rc;
        r.has_tt = true;
    }
    r.copy_ms = ms_between(t1, Clock::now());
    return TG_OK;
}

void fill_result(const Request &r, tg_request_result *o) {
    o->status = r.status;
    o->n_points = r.V;
    o->n_cells = r.T;
    o->classes = r.classes.data();
    o->vt_off = r.has_vt ? r.vt_off.data() : nullptr;
    o->vt_val = r.has_vt ? r.vt_val.data() : nullptr;
    o->vt_nnz = r.vt_nnz;
    o->tt_off = r.has_tt ? r.tt_off.data() : nullptr;
    o->tt_val = r.has_tt ? r.tt_val.data() : nullptr;
    o->tt_nnz = r.tt_nnz;
    o->queue_ms = ms_between(r.submitted, r.started);
    o->run_ms = ms_between(r.started, r.finished);
    o->total_ms = ms_between(r.submitted, r.finished);
    o->group_size = r.group_size;
    o->upload_ms = r.upload_ms;
    o->compute_ms = r.compute_ms;
    o->followup_ms = r.followup_ms;
    o->copy_ms = r.copy_ms;
}

}  // namespace

struct tg_server {
    tg_server_config cfg{};
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request *> pending;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_id = 1;
    int64_t resident = 0, n_requests = 0, n_groups = 0;
    bool stop = false;
    std::vector<void *> streams;  // one per slot, owned by the server
    std::vector<std::thread> workers;
};

namespace {

// Take the head request (when it fits the budget) and, if it is small, the
// small requests queued behind it.  Called with the server mutex held.
std::vector<Request *> take_group(tg_server &s) {
    std::vector<Request *> g;
    if (s.pending.empty()) return g;
    Request *head = s.pending.front();
    if (s.resident > 0 && s.resident + head->bytes > s.cfg.hbm_budget_bytes) return g;
    s.pending.pop_front();
    g.push_back(head);
    int64_t cells = head->T, bytes = head->bytes;
    const bool fuse = s.cfg.fuse_max_meshes > 1 && head->T <= s.cfg.fuse_max_cells;
    while (fuse && !s.pending.empty() && (int32_t)g.size() < s.cfg.fuse_max_meshes) {
        Request *n = s.pending.front();
        if (n->followups != head->followups || cells + n->T > s.cfg.fuse_max_cells ||
            s.resident + bytes + n->bytes > s.cfg.hbm_budget_bytes)
            break;
        s.pending.pop_front();
        g.push_back(n);
        cells += n->T;
        bytes += n->bytes;
    }
    s.resident += bytes;
    s.n_groups++;
    return g;
}

// Request inputs are staged in pinned host memory by the slot's own thread
// (one reused buffer per slot): a copy from pageable memory goes through the
// runtime's bounce buffer, one transfer at a time across all slots, while a
// copy from pinned memory runs on the DMA engines, and the slots' host-side
// copies proceed in parallel.  The device copies out of the buffer are done
// once the group's classes are back on the host (same stream), or after the
// sync on an error, before the slot stages its next group.
thread_local Pinned<char> t_stage;

bool stage_inputs(const std::vector<Request *> &reqs, int64_t T, int64_t V, int32_t *&cells,
                  float *&sc) {
    if (!t_stage.resize(std::max<size_t>(16 * (size_t)T + 4 * (size_t)V, 16))) return false;
    cells = (int32_t *)t_stage.data();
    sc = (float *)(t_stage.data() + 16 * (size_t)T);  // 16-byte aligned
    int64_t c = 0, p = 0;
    for (const Request *r : reqs) {
        std::memcpy(cells + 4 * c, r->cells, 16 * (size_t)r->T);
        if (r->scalars) std::memcpy(sc + p, r->scalars, 4 * (size_t)r->V);
        c += r->T;
        p += r->V;
    }
    return true;
}

int run_group(std::vector<Request *> &reqs, Group &g) {
    const auto t0 = Clock::now();
    for (Request *r : reqs) r->started = t0;
    int rc = TG_OK;
    Clock::time_point t1 = t0;
    int32_t *cells = nullptr;
    float *sc = nullptr;
    if (reqs.size() == 1) {
        Request &r = *reqs[0];
        if (!stage_inputs(reqs, r.T, r.V, cells, sc)) rc = TG_ERR_ALLOC;
        if (!rc)
            rc = tg_mesh_create(&g.mesh, cells, r.T, r.V, r.scalars ? sc : nullptr, 0, 0,
                                g.stream);
        t1 = Clock::now();
        if (!rc) rc = tg_vv_critpoints(g.mesh);
        if (!rc && !r.classes.resize(3 * (size_t)std::max<int64_t>(r.V, 1))) rc = TG_ERR_ALLOC;
        if (!rc) rc = tg_classes_copy(g.mesh, r.classes.data());
        if (rc && g.mesh) (void)tg_sync(g.mesh);
    } else {
        const int32_t n = (int32_t)reqs.size();
        std::vector<int64_t> coff(n + 1, 0), poff(n + 1, 0);
        for (int32_t i = 0; i < n; i++) {
            coff[i + 1] = coff[i] + reqs[i]->T;
            poff[i + 1] = poff[i] + reqs[i]->V;
        }
        if (!stage_inputs(reqs, coff[n], poff[n], cells, sc)) rc = TG_ERR_ALLOC;
        if (!rc)
            rc = tg_batch_create(&g.batch, n, coff.data(), cells, poff.data(), sc, 0, 0,
                                 g.stream);
        t1 = Clock::now();
        if (!rc) rc = tg_batch_vv_critpoints(g.batch);
        std::vector<uint32_t> all(3 * (size_t)std::max<int64_t>(poff[n], 1));
        if (!rc) rc = tg_batch_classes_copy(g.batch, all.data());
        for (int32_t i = 0; i < n && !rc; i++) {
            reqs[i]->index = i;
            const size_t k = 3 * (size_t)(poff[i + 1] - poff[i]);
            if (!reqs[i]->classes.resize(std::max<size_t>(k, 3))) {
                rc = TG_ERR_ALLOC;
                break;
            }
            std::copy(all.begin() + 3 * poff[i], all.begin() + 3 * poff[i] + k,
                      reqs[i]->classes.data());
        }
        if (rc && g.batch) (void)tg_batch_sync(g.batch);
    }
    const auto t2 = Clock::now();
    for (Request *r : reqs) {
        r->upload_ms = ms_between(t0, t1);
        r->compute_ms = ms_between(t1, t2);
    }
    if (!rc && (reqs[0]->followups & kRelFlags)) {
        std::lock_guard<std::mutex> lk(g.mu);
        for (Request *r : reqs)
            if (!rc) rc = followup_locked(g, *r, r->followups);
    }
    return rc;
}

void worker(tg_server *s, void *stream) {
    if (tg_init(0) != TG_OK) return;
    while (true) {
        std::vector<Request *> reqs;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->work_cv.wait(lk, [&] {
                if (s->stop) return true;
                reqs = take_group(*s);
                return !reqs.empty();
            });
            if (reqs.empty()) break;  // stopping
        }
        auto g = std::make_shared<Group>();
        g->stream = stream;
        for (Request *r : reqs) g->bytes += r->bytes;
        g->refs = (int)reqs.size();
        const int rc = run_group(reqs, *g);
        const auto t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request *r : reqs) {
                r->status = rc;
                r->finished = t1;
                r->group = g;
                r->group_size = (int32_t)reqs.size();
                r->done = true;
            }
        }
        s->done_cv.notify_all();
    }
}

}  // namespace

extern "C" {

int tg_server_create(tg_server **out, const tg_server_config *cfg) {
    if (!out || !cfg || cfg->slots < 1 || cfg->hbm_budget_bytes <= 0) return TG_ERR_ARG;
    int rc = tg_init(0);
    if (rc) return rc;
    auto *s = new tg_server();
    s->cfg = *cfg;
    for (int i = 0; i < cfg->slots; i++) {
        void *st = nullptr;
        if ((rc = tg_stream_create(&st)) != TG_OK) {
            for (void *x : s->streams) tg_stream_destroy(x);
            delete s;
            return rc;
        }
        s->streams.push_back(st);
    }
    for (int i = 0; i < cfg->slots; i++) s->workers.emplace_back(worker, s, s->streams[i]);
    *out = s;
    return TG_OK;
}

int tg_server_submit(tg_server *s, const int32_t *cells, int64_t n_cells, int64_t n_points,
                     const float *scalars, uint32_t followups, uint64_t *id) {
    if (!s || !id || n_cells < 0 || n_points < 0 || (n_cells && !cells) || (n_points && !scalars))
        return TG_ERR_ARG;
    if (followups & ~(kRelFlags | TG_SERVE_COPY)) return TG_ERR_ARG;
    auto r = std::make_unique<Request>();
    r->cells = cells;
    r->scalars = scalars;
    r->T = n_cells;
    r->V = n_points;
    r->followups = followups;
    r->bytes = footprint(n_cells, n_points, followups);
    r->submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->bytes > s->cfg.hbm_budget_bytes) return TG_ERR_CAPACITY;
        r->id = s->next_id++;
        *id = r->id;
        s->pending.push_back(r.get());
        s->requests[r->id] = std::move(r);
        s->n_requests++;
    }
    s->work_cv.notify_all();
    return TG_OK;
}

int tg_server_wait(tg_server *s, uint64_t id, tg_request_result *out) {
    if (!s || !out) return TG_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return TG_ERR_ARG;
    Request *r = it->second.get();
    s->done_cv.wait(lk, [&] { return r->done; });
    fill_result(*r, out);
    return r->status;
}

int tg_server_followup(tg_server *s, uint64_t id, uint32_t flags, tg_request_result *out) {
    if (!s || (flags & ~(kRelFlags | TG_SERVE_COPY))) return TG_ERR_ARG;
    Request *r = nullptr;
    std::shared_ptr<Group> g;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        if (it == s->requests.end()) return TG_ERR_ARG;
        r = it->second.get();
        s->done_cv.wait(lk, [&] { return r->done; });
        if (r->status) return r->status;
        g = r->group;
    }
    int rc;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        rc = followup_locked(*g, *r, flags);
    }
    if (out) fill_result(*r, out);
    return rc;
}

int tg_server_release(tg_server *s, uint64_t id) {
    if (!s) return TG_ERR_ARG;
    std::shared_ptr<Group> drop;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->requests.find(id);
        if (it == s->requests.end() || !it->second->done) return TG_ERR_ARG;
        std::shared_ptr<Group> g = it->second->group;
        s->requests.erase(it);
        if (g && --g->refs == 0) {
            s->resident -= g->bytes;
            drop = g;  // destroyed below, outside the lock
        }
    }
    drop.reset();
    s->work_cv.notify_all();
    return TG_OK;
}

int tg_server_stats(tg_server *s, int64_t *requests, int64_t *groups, int64_t *resident) {
    if (!s) return TG_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (requests) *requests = s->n_requests;
    if (groups) *groups = s->n_groups;
    if (resident) *resident = s->resident;
    return TG_OK;
}

int tg_server_destroy(tg_server *s) {
    if (!s) return TG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->work_cv.notify_all();
    for (auto &t : s->workers) t.join();
    s->requests.clear();  // drops every group (device state) while the streams live
    for (void *st : s->streams) tg_stream_destroy(st);
    delete s;
    return TG_OK;
}

}  // extern "C"
