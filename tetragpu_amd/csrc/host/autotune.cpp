// Autotuner for the VV + critical-point path: times every kernel form and
// shape per size class on generated meshes and installs the fastest.
#include "autotune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>

#include "tetra_meshgen.h"

namespace {

std::vector<tg_tune_entry> candidates() {
    std::vector<tg_tune_entry> c;
    for (int cap : {32, 64})
        for (int blk : {64, 128, 256}) {
            if (cap == 64 && blk == 256) continue;
            c.push_back({0, 1, cap, blk, 7, 8192});
        }
    for (int variant : {2, 3})
        for (int bv : {7, 8})
            for (int64_t tpc : {4096, 8192, 16384}) c.push_back({0, variant, 32, 128, bv, tpc});
    return c;
}

double time_config(tg_mesh *m, const tg_tune_entry &c, int reps) {
    if (tg_mesh_set_config(m, c.variant, c.lane_cap, c.lane_block) ||
        (c.variant >= 2 && tg_mesh_set_bucket(m, c.bv_log, c.tets_per_chunk)))
        return -1;
    if (tg_vv_critpoints(m) || tg_sync(m)) return -1;  // warm-up (allocations)
    std::vector<double> ms;
    for (int r = 0; r < reps; r++) {
        const auto t0 = std::chrono::steady_clock::now();
        if (tg_vv_critpoints(m) || tg_sync(m)) return -1;
        ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                         .count());
    }
    std::sort(ms.begin(), ms.end());
    return ms[ms.size() / 2];
}

}  // namespace

std::vector<tg_tune_entry> autotune(const std::vector<int> &sides, int reps,
                                    std::vector<TuneCandidateTime> *log) {
    std::vector<tg_tune_entry> table;
    if (tg_init(0)) return table;
    for (size_t k = 0; k < sides.size(); k++) {
        tg_gen_mesh *g = nullptr;
        if (tg_gen_kuhn(sides[k], 1234 + k, TG_GEN_STANDARD, 0.0, 0, 0, TG_SCALAR_GAUSSIANS, &g))
            continue;
        tg_mesh *m = nullptr;
        if (tg_mesh_create(&m, g->cells, g->n_cells, g->n_points, g->scalars, 0, 0, nullptr)) {
            tg_gen_free(g);
            continue;
        }
        tg_tune_entry best{};
        double best_ms = std::numeric_limits<double>::infinity();
        for (tg_tune_entry c : candidates()) {
            const double ms = time_config(m, c, std::max(reps, 1));
            if (ms < 0) continue;
            c.max_cells = g->n_cells;
            if (log) log->push_back({g->n_cells, c, ms});
            std::fprintf(stderr, "tune cells=%lld variant=%d lane=%dx%d bucket=2^%d chunk=%lld: %.4f ms\n",
                         (long long)g->n_cells, c.variant, c.lane_cap, c.lane_block, c.bv_log,
                         (long long)c.tets_per_chunk, ms);
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        if (std::isfinite(best_ms)) {
            best.max_cells = k + 1 == sides.size() ? std::numeric_limits<int64_t>::max()
                                                   : g->n_cells;
            table.push_back(best);
        }
        tg_mesh_destroy(m);
        tg_gen_free(g);
    }
    tg_set_tuning(table.data(), (int32_t)table.size());
    return table;
}

bool save_tuning(const std::string &path, const std::vector<tg_tune_entry> &t) {
    std::ofstream f(path);
    if (!f) return false;
    f << "[\n";
    for (size_t i = 0; i < t.size(); i++)
        f << "  {\"max_cells\": " << t[i].max_cells << ", \"variant\": " << t[i].variant
          << ", \"lane_cap\": " << t[i].lane_cap << ", \"lane_block\": " << t[i].lane_block
          << ", \"bv_log\": " << t[i].bv_log << ", \"tets_per_chunk\": " << t[i].tets_per_chunk
          << "}" << (i + 1 < t.size() ? "," : "") << "\n";
    f << "]\n";
    return bool(f);
}

std::vector<tg_tune_entry> load_tuning(const std::string &path) {
    std::vector<tg_tune_entry> t;
    std::ifstream f(path);
    if (!f) return t;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    auto field = [&](size_t from, size_t to, const char *key) -> long long {
        const std::string k = std::string("\"") + key + "\"";
        const size_t p = s.find(k, from);
        if (p == std::string::npos || p > to) return -1;
        return std::strtoll(s.c_str() + s.find(':', p) + 1, nullptr, 10);
    };
    for (size_t p = s.find('{'); p != std::string::npos; p = s.find('{', p + 1)) {
        const size_t q = s.find('}', p);
        tg_tune_entry e{};
        e.max_cells = field(p, q, "max_cells");
        e.variant = (int32_t)field(p, q, "variant");
        e.lane_cap = (int32_t)field(p, q, "lane_cap");
        e.lane_block = (int32_t)field(p, q, "lane_block");
        e.bv_log = (int32_t)field(p, q, "bv_log");
        e.tets_per_chunk = field(p, q, "tets_per_chunk");
        t.push_back(e);
    }
    return t;
}

extern "C" int tg_autotune(const int32_t *sides, int32_t n_sides, int32_t reps, tg_tune_entry *out,
                           int32_t capacity, int32_t *n_out) {
    if (!sides || n_sides <= 0 || !n_out) return TG_ERR_ARG;
    const auto t = autotune(std::vector<int>(sides, sides + n_sides), reps);
    *n_out = (int32_t)t.size();
    for (int32_t i = 0; i < capacity && i < *n_out && out; i++) out[i] = t[i];
    return t.empty() ? TG_ERR_STATE : TG_OK;
}

extern "C" int tg_tuning_load(const char *path) {
    const auto t = load_tuning(path ? path : "");
    if (t.empty()) return TG_ERR_ARG;
    return tg_set_tuning(t.data(), (int32_t)t.size());
}

extern "C" int tg_tuning_save(const char *path) {
    tg_tune_entry e[64];
    int32_t n = 0;
    if (tg_get_tuning(e, 64, &n)) return TG_ERR_STATE;
    return save_tuning(path ? path : "", std::vector<tg_tune_entry>(e, e + std::min(n, 64))) ? TG_OK
                                                                                             : TG_ERR_ARG;
}
