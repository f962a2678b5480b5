// Deterministic Kuhn-grid tetrahedral meshes with seeded permutations,
// stellar splits and scalar fields.  See include/tetra_meshgen.h.
#include "tetra_meshgen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace {

// SplitMix64 (Vigna): used only to expand a seed into generator state.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
};

// xoshiro256** (Blackman & Vigna).
struct Rng {
    uint64_t s[4];
    Rng(uint64_t seed, uint64_t stream) {
        SplitMix64 sm(seed ^ (0xd1b54a32d192ed03ULL * (stream + 1)));
        for (auto &w : s) w = sm.next();
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // integer in [0, n): Lemire's multiply-high mapping
    uint64_t below(uint64_t n) {
        return (uint64_t)(((unsigned __int128)next() * n) >> 64);
    }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

enum Stream : uint64_t {
    S_VPERM = 1, S_TPERM = 2, S_JITTER = 3, S_SCALAR = 4, S_SPLIT = 5,
    S_ORDER = 6, S_HUB = 7, S_NOISE = 8
};

template <typename T>
void shuffle(std::vector<T> &a, Rng &r) {
    for (size_t i = a.size(); i > 1; i--) {
        size_t j = (size_t)r.below(i);
        std::swap(a[i - 1], a[j]);
    }
}

struct Work {
    int32_t n = 0;
    std::vector<int32_t> cells;     // [T][4], unpermuted ids
    std::vector<double> pos;        // [V][3]
    std::vector<int64_t> grid_idx;  // per vertex, -1 for split vertices
};

void build_grid(Work &w, int32_t n, bool jitter, uint64_t seed) {
    const int64_t s = n + 1;
    const int64_t V = s * s * s;
    w.n = n;
    w.pos.resize(3 * V);
    w.grid_idx.resize(V);
    Rng rj(seed, S_JITTER);
    for (int64_t k = 0; k < s; k++)
        for (int64_t j = 0; j < s; j++)
            for (int64_t i = 0; i < s; i++) {
                int64_t g = i + s * (j + s * k);
                w.grid_idx[g] = g;
                double h = 1.0 / n;
                double jx = 0, jy = 0, jz = 0;
                if (jitter) {
                    jx = (rj.uniform() * 0.4 - 0.2) * h;
                    jy = (rj.uniform() * 0.4 - 0.2) * h;
                    jz = (rj.uniform() * 0.4 - 0.2) * h;
                }
                w.pos[3 * g] = i * h + jx;
                w.pos[3 * g + 1] = j * h + jy;
                w.pos[3 * g + 2] = k * h + jz;
            }
    // 6 tets per cube: one per axis order of the path 000 -> 111.
    static const int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                   {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int64_t step[3] = {1, s, s * s};
    w.cells.resize((size_t)24 * n * n * n);
    size_t t = 0;
    for (int64_t k = 0; k < n; k++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < n; i++) {
                int64_t base = i + s * (j + s * k);
                for (auto &p : PERM) {
                    int64_t g = base;
                    w.cells[4 * t] = (int32_t)g;
                    for (int a = 0; a < 3; a++) {
                        g += step[p[a]];
                        w.cells[4 * t + 1 + a] = (int32_t)g;
                    }
                    t++;
                }
            }
}

// Split tet t 1->4 at its centroid: the new vertex x replaces each corner in
// turn; sub-tet 0 reuses slot t, the others are appended.
int32_t split_tet(Work &w, size_t t) {
    int32_t q[4];
    std::memcpy(q, &w.cells[4 * t], sizeof q);
    int32_t x = (int32_t)w.grid_idx.size();
    double c[3] = {0, 0, 0};
    for (int k = 0; k < 4; k++)
        for (int d = 0; d < 3; d++) c[d] += 0.25 * w.pos[3 * (size_t)q[k] + d];
    for (int d = 0; d < 3; d++) w.pos.push_back(c[d]);
    w.grid_idx.push_back(-1);
    for (int r = 0; r < 4; r++) {
        int32_t sub[4] = {q[0], q[1], q[2], q[3]};
        sub[r] = x;
        if (r == 0) std::memcpy(&w.cells[4 * t], sub, sizeof sub);
        else w.cells.insert(w.cells.end(), sub, sub + 4);
    }
    return x;
}

void apply_splits(Work &w, double frac, int32_t n_hubs, int32_t hub_splits,
                  uint64_t seed) {
    const size_t T0 = w.cells.size() / 4;
    // hubs: repeatedly split tets around a few vertices (sub-tets may be
    // chosen again), raising each hub's degree by hub_splits
    if (n_hubs > 0 && hub_splits > 0) {
        Rng rh(seed, S_HUB);
        const int64_t V0 = (int64_t)w.grid_idx.size();
        for (int32_t h = 0; h < n_hubs; h++) {
            int32_t hub = (int32_t)rh.below((uint64_t)V0);
            std::vector<size_t> around;
            for (size_t t = 0; t < w.cells.size() / 4; t++)
                for (int k = 0; k < 4; k++)
                    if (w.cells[4 * t + k] == hub) { around.push_back(t); break; }
            for (int32_t s = 0; s < hub_splits && !around.empty(); s++) {
                size_t pick = (size_t)rh.below(around.size());
                size_t t = around[pick];
                size_t first_new = w.cells.size() / 4;
                split_tet(w, t);
                // sub-tets holding the hub: slot t (if it kept the hub) and
                // the appended ones that kept it
                around.erase(around.begin() + (long)pick);
                for (size_t u : {t, first_new, first_new + 1, first_new + 2}) {
                    bool has = false;
                    for (int k = 0; k < 4; k++) has |= (w.cells[4 * u + k] == hub);
                    if (has) around.push_back(u);
                }
            }
        }
    }
    // single-level random splits of original tets
    if (frac > 0) {
        Rng rs(seed, S_SPLIT);
        std::vector<uint32_t> ids(T0);
        for (size_t i = 0; i < T0; i++) ids[i] = (uint32_t)i;
        size_t want = (size_t)std::llround(frac * (double)T0);
        // partial Fisher-Yates: the first `want` entries are the choice
        for (size_t i = 0; i < want && i < T0; i++) {
            size_t j = i + (size_t)rs.below(T0 - i);
            std::swap(ids[i], ids[j]);
        }
        std::sort(ids.begin(), ids.begin() + (long)std::min(want, T0));
        for (size_t i = 0; i < want && i < T0; i++) split_tet(w, ids[i]);
    }
}

void fill_scalars(tg_gen_mesh *m, const std::vector<double> &pos_by_id,
                  const std::vector<int64_t> &grid_by_id, int kind,
                  uint64_t seed) {
    const int64_t V = m->n_points;
    Rng r(seed, S_SCALAR);
    switch (kind) {
    case TG_SCALAR_UNIFORM:
        for (int64_t v = 0; v < V; v++)
            m->scalars[v] = (float)((double)(r.next() >> 40) * 0x1.0p-24);
        break;
    case TG_SCALAR_GAUSSIANS: {
        double cx[32], cy[32], cz[32], inv2s2[32];
        for (int g = 0; g < 32; g++) {
            cx[g] = r.uniform();
            cy[g] = r.uniform();
            cz[g] = r.uniform();
            double s = 0.02 + 0.08 * r.uniform();
            inv2s2[g] = 1.0 / (2 * s * s);
        }
        Rng rn(seed, S_NOISE);
        for (int64_t v = 0; v < V; v++) {
            double x = pos_by_id[3 * v], y = pos_by_id[3 * v + 1],
                   z = pos_by_id[3 * v + 2], f = 0;
            for (int g = 0; g < 32; g++) {
                double dx = x - cx[g], dy = y - cy[g], dz = z - cz[g];
                f += std::exp(-(dx * dx + dy * dy + dz * dz) * inv2s2[g]);
            }
            f += 1e-3 * rn.normal();
            m->scalars[v] = (float)f;
        }
        break;
    }
    case TG_SCALAR_SINES: {
        double kx[8], ky[8], kz[8], amp[8], ph[8];
        for (int q = 0; q < 8; q++) {
            // direction uniform on the sphere, |k| uniform in [1, 6]
            double zc = 2 * r.uniform() - 1, phi = 6.283185307179586 * r.uniform();
            double rr = std::sqrt(std::max(0.0, 1 - zc * zc));
            double mag = 1 + 5 * r.uniform();
            kx[q] = mag * rr * std::cos(phi);
            ky[q] = mag * rr * std::sin(phi);
            kz[q] = mag * zc;
            amp[q] = 0.5 + 0.5 * r.uniform();
            ph[q] = 6.283185307179586 * r.uniform();
        }
        for (int64_t v = 0; v < V; v++) {
            double x = pos_by_id[3 * v], y = pos_by_id[3 * v + 1],
                   z = pos_by_id[3 * v + 2], f = 0;
            for (int q = 0; q < 8; q++)
                f += amp[q] * std::sin(6.283185307179586 *
                                       (kx[q] * x + ky[q] * y + kz[q] * z) + ph[q]);
            m->scalars[v] = (float)f;
        }
        break;
    }
    case TG_SCALAR_LINEAR:
        for (int64_t v = 0; v < V; v++) m->scalars[v] = (float)grid_by_id[v];
        break;
    default:
        for (int64_t v = 0; v < V; v++) m->scalars[v] = 0.5f;
        break;
    }
}

}  // namespace

extern "C" int tg_gen_kuhn(int32_t n, uint64_t seed, uint32_t options,
                           double split_fraction, int32_t n_hubs,
                           int32_t hub_splits, int32_t scalar_kind,
                           tg_gen_mesh **out) {
    if (!out || n < 1 || n > 1024 || split_fraction < 0 || split_fraction > 1 ||
        n_hubs < 0 || hub_splits < 0 || scalar_kind < 0 || scalar_kind > 4)
        return -1;
    const bool split = split_fraction > 0 || (n_hubs > 0 && hub_splits > 0);
    if (split && scalar_kind == TG_SCALAR_LINEAR) return -1;
    *out = nullptr;
    try {
        Work w;
        build_grid(w, n, (options & TG_GEN_JITTER) != 0, seed);
        apply_splits(w, split_fraction, n_hubs, hub_splits, seed);
        const int64_t V = (int64_t)w.grid_idx.size();
        const int64_t T = (int64_t)w.cells.size() / 4;
        if (4 * T >= (int64_t)INT32_MAX || V >= (int64_t)INT32_MAX) return -2;
        // vertex permutation: old id -> new id
        std::vector<int32_t> vmap(V);
        for (int64_t i = 0; i < V; i++) vmap[i] = (int32_t)i;
        if (options & TG_GEN_PERMUTE_VERTICES) {
            Rng r(seed, S_VPERM);
            shuffle(vmap, r);
        }
        std::vector<int64_t> torder(T);
        for (int64_t i = 0; i < T; i++) torder[i] = i;
        if (options & TG_GEN_PERMUTE_CELLS) {
            Rng r(seed, S_TPERM);
            shuffle(torder, r);
        }
        auto *m = new tg_gen_mesh;
        m->n_points = V;
        m->n_cells = T;
        m->cells = (int32_t *)std::malloc(sizeof(int32_t) * 4 * (size_t)std::max<int64_t>(T, 1));
        m->positions = (float *)std::malloc(sizeof(float) * 3 * (size_t)V);
        m->scalars = (float *)std::malloc(sizeof(float) * (size_t)V);
        if (!m->cells || !m->positions || !m->scalars) {
            tg_gen_free(m);
            return -2;
        }
        const uint32_t order = (options >> TG_GEN_ORDER_SHIFT) & 3u;
        Rng ro(seed, S_ORDER);
        std::vector<int32_t> deg(V, 0);
        for (int64_t t = 0; t < T; t++) {
            int32_t q[4];
            for (int k = 0; k < 4; k++) q[k] = vmap[w.cells[4 * torder[t] + k]];
            if (order == TG_GEN_ORDER_SORTED) std::sort(q, q + 4);
            else if (order == TG_GEN_ORDER_SHUFFLED)
                for (int i = 3; i > 0; i--) std::swap(q[i], q[ro.below((uint64_t)i + 1)]);
            for (int k = 0; k < 4; k++) {
                m->cells[4 * t + k] = q[k];
                deg[q[k]]++;
            }
        }
        m->max_vt = V ? *std::max_element(deg.begin(), deg.end()) : 0;
        std::vector<double> pos_by_id(3 * V);
        std::vector<int64_t> grid_by_id(V);
        for (int64_t i = 0; i < V; i++) {
            int32_t id = vmap[i];
            for (int d = 0; d < 3; d++) {
                pos_by_id[3 * (size_t)id + d] = w.pos[3 * i + d];
                m->positions[3 * (size_t)id + d] = (float)w.pos[3 * i + d];
            }
            grid_by_id[id] = w.grid_idx[i];
        }
        fill_scalars(m, pos_by_id, grid_by_id, scalar_kind, seed);
        // keep positions to recompute scalars later
        *out = m;
        return 0;
    } catch (const std::bad_alloc &) {
        return -2;
    }
}

extern "C" int tg_gen_scalars(tg_gen_mesh *m, int32_t scalar_kind, uint64_t seed) {
    if (!m || scalar_kind < 0 || scalar_kind > 4) return -1;
    if (scalar_kind == TG_SCALAR_LINEAR) return -1;  // needs the grid index
    std::vector<double> pos(3 * (size_t)m->n_points);
    for (size_t i = 0; i < pos.size(); i++) pos[i] = m->positions[i];
    std::vector<int64_t> grid;
    fill_scalars(m, pos, grid, scalar_kind, seed);
    return 0;
}

extern "C" void tg_gen_free(tg_gen_mesh *m) {
    if (!m) return;
    std::free(m->cells);
    std::free(m->positions);
    std::free(m->scalars);
    delete m;
}
