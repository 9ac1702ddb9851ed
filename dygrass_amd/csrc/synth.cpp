// synth.cpp -- seeded synthetic workloads for tests and benchmarks
// (libdygrass_synth.so).  Not part of the update path: it produces the
// graphs, sparsifiers and update streams the path is run on.
//
// Generators (all deterministic in their seed):
//   grid2d       nx*ny 4-neighbour grid, weights log-uniform [wlo,whi]
//   rgg_knn      n points uniform in [0,1]^2, symmetrised k-nearest-neighbour
//                edges, weight 1/distance; components joined to the largest
//   trimesh      jittered nx*ny grid with one random diagonal per cell
//   grid3d_sc    n^3 6-neighbour grid plus uniform random shortcuts
//   sparsifier   maximum-weight spanning tree (Kruskal, ties by edge index)
//                plus a uniform sample of off-tree edges
//   insertions   random source, target at BFS hop distance d ~ U{dlo..dhi},
//                pairs already in G (or already drawn) rejected
//   deletions    uniform edges of G without replacement, skipping any that
//                would leave a vertex with degree 0
// Weights written to .mtx use %.9g so atof-then-float restores them exactly.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <queue>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Rng {  // SplitMix64-seeded xoshiro256**
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            s[i] = x ^ (x >> 31);
        }
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
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    float loguni(float lo, float hi) {
        return (float)std::exp(std::log((double)lo) + uni() * (std::log((double)hi) - std::log((double)lo)));
    }
};

struct EdgeList {
    std::vector<uint32_t> u, v;
    std::vector<float> w;
    void add(uint32_t a, uint32_t b, float x) {
        u.push_back(a);
        v.push_back(b);
        w.push_back(x);
    }
};

// Result handle passed through the C ABI.
struct Out {
    uint32_t V = 0;
    EdgeList e;
    std::vector<float> xy;  // point coordinates when the generator has them
};

struct DSU {
    std::vector<uint32_t> p;
    explicit DSU(uint32_t n) : p(n) { std::iota(p.begin(), p.end(), 0u); }
    uint32_t find(uint32_t x) {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    bool unite(uint32_t a, uint32_t b) {
        a = find(a);
        b = find(b);
        if (a == b) return false;
        if (a < b) std::swap(a, b);
        p[a] = b;
        return true;
    }
};

inline uint64_t pkey(uint32_t a, uint32_t b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)a << 32) | b;
}

void csr_of(uint32_t V, const EdgeList& e, std::vector<uint64_t>& rp, std::vector<uint32_t>& col) {
    rp.assign((size_t)V + 1, 0);
    for (size_t i = 0; i < e.u.size(); ++i) {
        rp[e.u[i] + 1]++;
        rp[e.v[i] + 1]++;
    }
    for (uint32_t v = 0; v < V; ++v) rp[v + 1] += rp[v];
    col.assign(rp[V], 0);
    std::vector<uint64_t> pos(rp.begin(), rp.end() - 1);
    for (size_t i = 0; i < e.u.size(); ++i) {
        col[pos[e.u[i]]++] = e.v[i];
        col[pos[e.v[i]]++] = e.u[i];
    }
}

// Grid bucketing for 2D points.
struct Grid2 {
    int g;
    std::vector<uint32_t> start, idx;
    Grid2(const std::vector<float>& xy, uint32_t n, double per_cell) {
        g = std::max(1, (int)std::sqrt(n / per_cell));
        start.assign((size_t)g * g + 1, 0);
        std::vector<uint32_t> cell(n);
        for (uint32_t i = 0; i < n; ++i) {
            int cx = std::min(g - 1, (int)(xy[2 * i] * g)), cy = std::min(g - 1, (int)(xy[2 * i + 1] * g));
            cell[i] = (uint32_t)(cy * g + cx);
            start[cell[i] + 1]++;
        }
        for (size_t c = 0; c < (size_t)g * g; ++c) start[c + 1] += start[c];
        idx.resize(n);
        std::vector<uint32_t> pos(start.begin(), start.end() - 1);
        for (uint32_t i = 0; i < n; ++i) idx[pos[cell[i]]++] = i;
    }
};

void knn_edges(const std::vector<float>& xy, uint32_t n, uint32_t k, EdgeList& out) {
    Grid2 G(xy, n, 2.0);
    const int g = G.g;
    std::vector<uint32_t> nb((size_t)n * k);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float x = xy[2 * i], y = xy[2 * i + 1];
        const int cx = std::min(g - 1, (int)(x * g)), cy = std::min(g - 1, (int)(y * g));
        std::vector<std::pair<float, uint32_t>> cand;
        for (int r = 0;; ++r) {
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx) {
                    if (std::max(std::abs(dx), std::abs(dy)) != r) continue;
                    const int qx = cx + dx, qy = cy + dy;
                    if (qx < 0 || qy < 0 || qx >= g || qy >= g) continue;
                    const size_t c = (size_t)qy * g + qx;
                    for (uint32_t p = G.start[c]; p < G.start[c + 1]; ++p) {
                        const uint32_t j = G.idx[p];
                        if (j == i) continue;
                        const float ddx = xy[2 * j] - x, ddy = xy[2 * j + 1] - y;
                        cand.push_back({ddx * ddx + ddy * ddy, j});
                    }
                }
            if (cand.size() >= k) {
                std::nth_element(cand.begin(), cand.begin() + (k - 1), cand.end());
                const float kth = cand[k - 1].first;
                const float reach = (float)r / g;  // every unvisited point is farther than this
                if (kth <= reach * reach || r > 2 * g) break;
            }
            if (r > 2 * g) break;
        }
        std::sort(cand.begin(), cand.end());
        for (uint32_t q = 0; q < k; ++q) nb[(size_t)i * k + q] = q < cand.size() ? cand[q].second : i;
    }
    std::vector<uint64_t> keys;
    keys.reserve((size_t)n * k);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t q = 0; q < k; ++q) {
            const uint32_t j = nb[(size_t)i * k + q];
            if (j != i) keys.push_back(pkey(i, j));
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (uint64_t key : keys) {
        const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
        const float dx = xy[2 * a] - xy[2 * b], dy = xy[2 * a + 1] - xy[2 * b + 1];
        const float d = std::sqrt(dx * dx + dy * dy);
        out.add(a, b, 1.0f / std::max(d, 1e-12f));
    }
}

// Join every component to the largest through its closest outside point.
void connect_components(const std::vector<float>& xy, uint32_t n, EdgeList& e) {
    DSU d(n);
    for (size_t i = 0; i < e.u.size(); ++i) d.unite(e.u[i], e.v[i]);
    std::vector<uint32_t> sz(n, 0);
    for (uint32_t i = 0; i < n; ++i) sz[d.find(i)]++;
    uint32_t big = 0;
    for (uint32_t i = 0; i < n; ++i)
        if (sz[i] > sz[big]) big = i;
    std::vector<std::vector<uint32_t>> comps;
    std::vector<int64_t> cid(n, -1);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t r = d.find(i);
        if (r == big) continue;
        if (cid[r] < 0) {
            cid[r] = (int64_t)comps.size();
            comps.emplace_back();
        }
        comps[cid[r]].push_back(i);
    }
    for (auto& c : comps) {
        float best = 1e30f;
        uint32_t ba = c[0], bb = big;
        for (uint32_t a : c)
            for (uint32_t b = 0; b < n; b += 1) {
                if (d.find(b) != big) continue;
                const float dx = xy[2 * a] - xy[2 * b], dy = xy[2 * a + 1] - xy[2 * b + 1];
                const float dd = dx * dx + dy * dy;
                if (dd < best) {
                    best = dd;
                    ba = a;
                    bb = b;
                }
            }
        e.add(std::min(ba, bb), std::max(ba, bb), 1.0f / std::max(std::sqrt(best), 1e-12f));
    }
}

}  // namespace

extern "C" {

typedef struct Out* dyg_synth;

void dyg_synth_free(dyg_synth s) { delete s; }
uint32_t dyg_synth_nv(dyg_synth s) { return s->V; }
uint64_t dyg_synth_ne(dyg_synth s) { return s->e.u.size(); }
const uint32_t* dyg_synth_u(dyg_synth s) { return s->e.u.data(); }
const uint32_t* dyg_synth_v(dyg_synth s) { return s->e.v.data(); }
const float* dyg_synth_w(dyg_synth s) { return s->e.w.data(); }
const float* dyg_synth_xy(dyg_synth s) { return s->xy.empty() ? nullptr : s->xy.data(); }

dyg_synth dyg_synth_grid2d(uint32_t nx, uint32_t ny, float wlo, float whi, uint64_t seed) {
    auto* o = new Out;
    o->V = nx * ny;
    Rng r(seed);
    for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t x = 0; x < nx; ++x) {
            const uint32_t i = y * nx + x;
            if (x + 1 < nx) o->e.add(i, i + 1, r.loguni(wlo, whi));
            if (y + 1 < ny) o->e.add(i, i + nx, r.loguni(wlo, whi));
        }
    return o;
}

dyg_synth dyg_synth_rgg_knn(uint32_t n, uint32_t k, uint64_t seed) {
    auto* o = new Out;
    o->V = n;
    Rng r(seed);
    o->xy.resize(2 * (size_t)n);
    for (auto& c : o->xy) c = (float)r.uni();
    knn_edges(o->xy, n, k, o->e);
    connect_components(o->xy, n, o->e);
    return o;
}

dyg_synth dyg_synth_trimesh(uint32_t nx, uint32_t ny, float wlo, float whi, uint64_t seed) {
    auto* o = new Out;
    o->V = nx * ny;
    Rng r(seed);
    o->xy.resize(2 * (size_t)o->V);
    for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t x = 0; x < nx; ++x) {
            const uint32_t i = y * nx + x;
            o->xy[2 * i] = (float)((x + 0.3 * (r.uni() - 0.5)) / nx);
            o->xy[2 * i + 1] = (float)((y + 0.3 * (r.uni() - 0.5)) / ny);
        }
    for (uint32_t y = 0; y < ny; ++y)
        for (uint32_t x = 0; x < nx; ++x) {
            const uint32_t i = y * nx + x;
            if (x + 1 < nx) o->e.add(i, i + 1, r.loguni(wlo, whi));
            if (y + 1 < ny) o->e.add(i, i + nx, r.loguni(wlo, whi));
            if (x + 1 < nx && y + 1 < ny) {
                if (r.next() & 1) o->e.add(i, i + nx + 1, r.loguni(wlo, whi));
                else o->e.add(i + 1, i + nx, r.loguni(wlo, whi));
            }
        }
    return o;
}

dyg_synth dyg_synth_grid3d_sc(uint32_t n, uint64_t shortcuts, float wlo, float whi, uint64_t seed) {
    auto* o = new Out;
    const uint64_t V = (uint64_t)n * n * n;
    o->V = (uint32_t)V;
    Rng r(seed);
    for (uint32_t z = 0; z < n; ++z)
        for (uint32_t y = 0; y < n; ++y)
            for (uint32_t x = 0; x < n; ++x) {
                const uint32_t i = (z * n + y) * n + x;
                if (x + 1 < n) o->e.add(i, i + 1, r.loguni(wlo, whi));
                if (y + 1 < n) o->e.add(i, i + n, r.loguni(wlo, whi));
                if (z + 1 < n) o->e.add(i, i + n * n, r.loguni(wlo, whi));
            }
    std::unordered_set<uint64_t> seen;
    seen.reserve(shortcuts * 2);
    uint64_t made = 0;
    while (made < shortcuts) {
        uint32_t a = (uint32_t)r.below(V), b = (uint32_t)r.below(V);
        if (a == b) continue;
        if (a > b) std::swap(a, b);
        // grid neighbours differ by 1, n or n*n; skip those to keep G simple
        const uint32_t d = b - a;
        if (d == 1 || d == n || d == n * n) continue;
        if (!seen.insert(pkey(a, b)).second) continue;
        o->e.add(a, b, r.loguni(wlo, whi));
        ++made;
    }
    return o;
}

// Maximum-weight spanning tree of (V, u, v, w) plus `extra` off-tree edges
// drawn uniformly without replacement.  Edges keep G's orientation and weight.
dyg_synth dyg_synth_sparsifier(uint32_t V, const uint32_t* u, const uint32_t* v, const float* w,
                               uint64_t m, uint64_t extra, uint64_t seed) {
    auto* o = new Out;
    o->V = V;
    std::vector<uint64_t> ord(m);
    std::iota(ord.begin(), ord.end(), 0ull);
    std::stable_sort(ord.begin(), ord.end(), [&](uint64_t a, uint64_t b) { return w[a] > w[b]; });
    DSU d(V);
    std::vector<char> in_tree(m, 0);
    for (uint64_t i : ord)
        if (d.unite(u[i], v[i])) in_tree[i] = 1;
    std::vector<uint64_t> off;
    off.reserve(m);
    for (uint64_t i = 0; i < m; ++i)
        if (!in_tree[i]) off.push_back(i);
    Rng r(seed);
    extra = std::min<uint64_t>(extra, off.size());
    for (uint64_t q = 0; q < extra; ++q) {  // partial Fisher-Yates
        const uint64_t j = q + r.below(off.size() - q);
        std::swap(off[q], off[j]);
        in_tree[off[q]] = 2;
    }
    for (uint64_t i = 0; i < m; ++i)
        if (in_tree[i]) o->e.add(u[i], v[i], w[i]);
    return o;
}

// Insertions: source uniform, target uniform among vertices at BFS distance d
// from it (d ~ U{dlo..dhi}).  Weight: log-uniform [wlo,whi], or 1/distance of
// the points when xy is given.
dyg_synth dyg_synth_insertions(uint32_t V, const uint32_t* u, const uint32_t* v, uint64_t m,
                               const float* xy, uint64_t count, uint32_t dlo, uint32_t dhi,
                               float wlo, float whi, uint64_t seed) {
    auto* o = new Out;
    o->V = V;
    EdgeList g;
    g.u.assign(u, u + m);
    g.v.assign(v, v + m);
    std::vector<uint64_t> rp;
    std::vector<uint32_t> col;
    csr_of(V, g, rp, col);
    std::unordered_set<uint64_t> used;
    used.reserve(count * 2);
    Rng r(seed);
    std::vector<uint32_t> dist(V, 0xFFFFFFFFu), touched, frontier, next, ring;
    uint64_t tries = 0;
    while (o->e.u.size() < count && tries < count * 50) {
        ++tries;
        const uint32_t s = (uint32_t)r.below(V);
        const uint32_t d = dlo + (uint32_t)r.below(dhi - dlo + 1);
        for (uint32_t x : touched) dist[x] = 0xFFFFFFFFu;
        touched.clear();
        frontier.assign(1, s);
        dist[s] = 0;
        touched.push_back(s);
        for (uint32_t lvl = 0; lvl < d && !frontier.empty(); ++lvl) {
            next.clear();
            for (uint32_t x : frontier)
                for (uint64_t p = rp[x]; p < rp[x + 1]; ++p) {
                    const uint32_t y = col[p];
                    if (dist[y] != 0xFFFFFFFFu) continue;
                    dist[y] = lvl + 1;
                    touched.push_back(y);
                    next.push_back(y);
                }
            frontier.swap(next);
        }
        if (frontier.empty()) continue;
        std::sort(frontier.begin(), frontier.end());
        const uint32_t t = frontier[r.below(frontier.size())];
        if (t == s || !used.insert(pkey(s, t)).second) continue;
        float wt;
        if (xy) {
            const float dx = xy[2 * s] - xy[2 * t], dy = xy[2 * s + 1] - xy[2 * t + 1];
            wt = 1.0f / std::max(std::sqrt(dx * dx + dy * dy), 1e-12f);
        } else {
            wt = r.loguni(wlo, whi);
        }
        o->e.add(s, t, wt);
    }
    return o;
}

// Insertions for graphs whose balls grow too fast for the BFS generator
// (shortcut graphs): the endpoint is where a non-backtracking random walk of
// d in [dlo,dhi] hops from a uniform source stops, so it lies within d hops;
// pairs that are already edges of G, or already drawn, are rejected.
dyg_synth dyg_synth_insertions_walk(uint32_t V, const uint32_t* u, const uint32_t* v, uint64_t m,
                                    uint64_t count, uint32_t dlo, uint32_t dhi, float wlo,
                                    float whi, uint64_t seed) {
    auto* o = new Out;
    o->V = V;
    EdgeList g;
    g.u.assign(u, u + m);
    g.v.assign(v, v + m);
    std::vector<uint64_t> rp;
    std::vector<uint32_t> col;
    csr_of(V, g, rp, col);
    std::unordered_set<uint64_t> used;
    used.reserve(count * 2);
    Rng r(seed);
    uint64_t tries = 0;
    while (o->e.u.size() < count && tries < count * 50) {
        ++tries;
        const uint32_t s = (uint32_t)r.below(V);
        const uint32_t d = dlo + (uint32_t)r.below(dhi - dlo + 1);
        uint32_t cur = s, prev = 0xFFFFFFFFu;
        bool ok = true;
        for (uint32_t k = 0; k < d && ok; ++k) {
            const uint64_t deg = rp[cur + 1] - rp[cur];
            if (deg == 0 || (deg == 1 && col[rp[cur]] == prev)) {
                ok = false;
                break;
            }
            uint32_t nxt;
            do {
                nxt = col[rp[cur] + r.below(deg)];
            } while (nxt == prev);
            prev = cur;
            cur = nxt;
        }
        if (!ok || cur == s) continue;
        bool adjacent = false;
        for (uint64_t p = rp[s]; p < rp[s + 1] && !adjacent; ++p) adjacent = col[p] == cur;
        if (adjacent || !used.insert(pkey(s, cur)).second) continue;
        o->e.add(s, cur, r.loguni(wlo, whi));
    }
    return o;
}

// Deletions: uniform edges of G without replacement, skipping those that
// would leave an endpoint with degree 0.  Returned edges keep G's weight.
dyg_synth dyg_synth_deletions(uint32_t V, const uint32_t* u, const uint32_t* v, const float* w,
                              uint64_t m, uint64_t count, uint64_t seed) {
    auto* o = new Out;
    o->V = V;
    std::vector<uint32_t> deg(V, 0);
    for (uint64_t i = 0; i < m; ++i) {
        deg[u[i]]++;
        deg[v[i]]++;
    }
    std::vector<uint64_t> ord(m);
    std::iota(ord.begin(), ord.end(), 0ull);
    Rng r(seed);
    for (uint64_t q = 0; q < m && o->e.u.size() < count; ++q) {
        const uint64_t j = q + r.below(m - q);
        std::swap(ord[q], ord[j]);
        const uint64_t i = ord[q];
        if (deg[u[i]] <= 1 || deg[v[i]] <= 1) continue;
        deg[u[i]]--;
        deg[v[i]]--;
        o->e.add(u[i], v[i], w ? w[i] : 1.0f);
    }
    return o;
}

// Write "u+base v+base [w]" lines; weights as %.9g.
int dyg_synth_write_mtx(const char* path, const uint32_t* u, const uint32_t* v, const float* w,
                        uint64_t m, int base) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf(1 << 20);
    size_t used = 0;
    for (uint64_t i = 0; i < m; ++i) {
        if (used + 96 > buf.size()) {
            std::fwrite(buf.data(), 1, used, f);
            used = 0;
        }
        int k = w ? std::snprintf(buf.data() + used, 96, "%u %u %.9g\n", u[i] + base, v[i] + base,
                                  (double)w[i])
                  : std::snprintf(buf.data() + used, 96, "%u %u\n", u[i] + base, v[i] + base);
        used += (size_t)k;
    }
    std::fwrite(buf.data(), 1, used, f);
    return std::fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"
