/*
 * oracle.cpp -- CPU restatement of dyGRASS's random-walk updaters.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product.  It may
 * be loaded by the tests, the smoke check and the benchmark's CPU-baseline
 * leg, and by nothing else.  The product never links it.
 *
 * Each function follows one reference unit (paths relative to the reference
 * repository root; citations beside each function).  Two replacements are
 * made, both also made by the product:
 *   - cuRAND XORWOW (closed library) -> Philox4x32-10 keyed per
 *     (seed, tag, edge id, walker, draw index);
 *   - defects D1 (decremental target stride), D2 (decremental vertex count
 *     v_max instead of v_max+1) and D3 (heuristic hop two indexing the source
 *     row) are resolved: targets use the batch stride, V = v_max+1, hop two
 *     reads n1's row.
 * Parity status of each part (how it is pinned) is stated in the tests that
 * exercise it; the header comment of each function says what it restates.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------ */
/* Philox4x32-10, written from Salmon, Moraes, Dror, Shaw, "Parallel   */
/* Random Numbers: As Easy as 1, 2, 3" (SC'11), section 4.            */
/* ------------------------------------------------------------------ */
const uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
const uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

inline void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PH_M0 * c0;
        uint64_t p1 = (uint64_t)PH_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PH_W0; k1 += PH_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Draw stream of one walker: counter = (d/4, walker, id_lo, tag<<24 | id_hi&0xffffff),
 * key = seed; draw d is word d%4 of that block. */
struct Stream {
    uint32_t key[2];
    uint32_t c1, c2, c3;
    uint32_t d = 0;
    uint32_t buf[4];
    Stream(uint64_t seed, uint32_t tag, uint64_t id, uint32_t walker) {
        key[0] = (uint32_t)seed;
        key[1] = (uint32_t)(seed >> 32);
        c1 = walker;
        c2 = (uint32_t)id;
        c3 = (tag << 24) | ((uint32_t)(id >> 32) & 0xFFFFFFu);
    }
    uint32_t next() {
        if ((d & 3u) == 0) {
            uint32_t ctr[4] = {d >> 2, c1, c2, c3};
            philox(ctr, key, buf);
        }
        return buf[d++ & 3u];
    }
};

struct Slot {
    uint32_t col;
    float w;
};

}  // namespace

/* Per-row slot lists.  Appends land at slot d-1 and relocation preserves
 * order (random_walk_incremental/gpu_graph.cuh:357-384); deletions swap the
 * last slot into the hole (random_walk_decremental/gpu_graph.cuh:368-399). */
struct orc_graph {
    std::vector<std::vector<Slot>> rows;
};

namespace {

/* ------------------------------------------------------------------ */
/* Walks                                                              */
/* ------------------------------------------------------------------ */

/* One insertion walker: random_walk_incremental/main.cu:79-163.
 * Returns targetFoundAt (-1 when the target is not reached). */
int walk_inc_one(const orc_graph* H, uint32_t s, uint32_t t, float w_e, float thr,
                 uint32_t max_steps, Stream& rng, float& R, uint64_t& hops,
                 std::vector<uint32_t>* traj) {
    uint32_t cur = s, prev = 0;
    int found = -1;
    int step = 1;
    R = 0.0f;
    if (traj) traj->push_back(s);
    const std::vector<Slot>* row = &H->rows[cur];
    uint32_t deg = (uint32_t)row->size();
    if (deg == 0) return -1; /* reference: r % 0 is undefined; an isolated source finds nothing */
    uint32_t i = rng.next() % deg;
    ++hops;
    uint32_t nxt = (*row)[i].col;
    float w = (*row)[i].w;
    R += 1.0f / w;
    if (traj) traj->push_back(nxt);
    if (nxt != t) {
        prev = cur;
        cur = nxt;
        step++;
        while (step < (int)max_steps) {
            row = &H->rows[cur];
            deg = (uint32_t)row->size();
            if (deg == 1) break;
            i = rng.next() % (deg - 1);
            ++hops;
            nxt = (*row)[i].col;
            if (nxt == prev) {
                nxt = (*row)[deg - 1].col;
                w = (*row)[deg - 1].w;
            } else {
                w = (*row)[i].w;
            }
            R += 1.0f / w;
            if (traj) traj->push_back(nxt);
            if (R * w_e >= thr) break;
            if (nxt == t) {
                found = step;
                break;
            }
            prev = cur;
            cur = nxt;
            step++;
        }
    } else {
        found = step;
    }
    return found;
}

/* One deletion walker: random_walk_decremental/main.cu:86-170 (no
 * distortion test; current degree; path recorded). */
int walk_dec_one(const orc_graph* G, uint32_t s, uint32_t t, uint32_t max_steps, Stream& rng,
                 float& R, uint64_t& hops, std::vector<uint32_t>* traj) {
    uint32_t cur = s, prev = 0;
    int found = -1;
    int step = 1;
    R = 0.0f;
    if (traj) traj->push_back(s);
    const std::vector<Slot>* row = &G->rows[cur];
    uint32_t deg = (uint32_t)row->size();
    if (deg == 0) return -1;
    uint32_t i = rng.next() % deg;
    ++hops;
    uint32_t nxt = (*row)[i].col;
    float w = (*row)[i].w;
    R += 1.0f / w;
    if (traj) traj->push_back(nxt);
    if (nxt != t) {
        prev = cur;
        cur = nxt;
        step++;
        while (step < (int)max_steps) {
            row = &G->rows[cur];
            deg = (uint32_t)row->size();
            if (deg == 1) break;
            if (deg == 0) {
                /* the reference prints and then indexes out of range
                 * (main.cu:130-139); deletions never leave a zero degree
                 * (gpu_graph.cuh:362-363), so this is unreachable */
                std::fprintf(stderr, "oracle: degree 0 reached in deletion walk\n");
                std::abort();
            }
            i = rng.next() % (deg - 1);
            ++hops;
            nxt = (*row)[i].col;
            if (nxt == prev) {
                nxt = (*row)[deg - 1].col;
                w = (*row)[deg - 1].w;
            } else {
                w = (*row)[i].w;
            }
            if (traj) traj->push_back(nxt);
            R += 1.0f / w;
            if (nxt == t) {
                found = step;
                break;
            }
            prev = cur;
            cur = nxt;
            step++;
        }
    } else {
        found = step;
    }
    return found;
}

/* Block reduction, written as the reference's loop
 * (random_walk_incremental/main.cu:179-201). */
uint32_t tree_reduce(std::vector<float> c) {
    uint32_t W = (uint32_t)c.size();
    std::vector<uint32_t> idx(W);
    for (uint32_t k = 0; k < W; ++k) idx[k] = k;
    for (uint32_t s = W / 2; s > 0; s >>= 1) {
        for (uint32_t k = 0; k < s; ++k) {
            float left = c[k], right = c[k + s];
            bool isSmall = left < right;
            c[k] = isSmall ? right : left;
            idx[k] = isSmall ? idx[k + s] : idx[k];
        }
    }
    return idx[0];
}

void inc_edge(const orc_graph* H, uint32_t s, uint32_t t, float w_e, float thr, uint32_t W,
              uint32_t max_steps, uint64_t seed, uint64_t eid, int32_t* found_at, float* cond,
              uint32_t* winner, uint64_t* hops) {
    std::vector<float> c(W);
    std::vector<int> f(W);
    std::vector<float> Rv(W);
    uint64_t h = 0;
    for (uint32_t j = 0; j < W; ++j) {
        Stream rng(seed, ORC_TAG_INSERT, eid, j);
        float R;
        f[j] = walk_inc_one(H, s, t, w_e, thr, max_steps, rng, R, h, nullptr);
        Rv[j] = R;
        c[j] = (f[j] != -1) ? 1.0f / R : -1.0f;
    }
    uint32_t win = tree_reduce(c);
    *found_at = f[win];
    *cond = (f[win] > 0) ? 1.0f / Rv[win] : -1.0f;
    if (winner) *winner = win;
    if (hops) *hops = h;
}

void dec_edge(const orc_graph* G, uint32_t s, uint32_t t, uint32_t W, uint32_t max_steps,
              uint64_t seed, uint64_t eid, int32_t* found_at, float* cond, uint32_t* winner,
              uint64_t* hops, uint32_t* path) {
    std::vector<float> c(W);
    std::vector<int> f(W);
    std::vector<float> Rv(W);
    uint64_t h = 0;
    for (uint32_t j = 0; j < W; ++j) {
        Stream rng(seed, ORC_TAG_DELETE, eid, j);
        float R;
        f[j] = walk_dec_one(G, s, t, max_steps, rng, R, h, nullptr);
        Rv[j] = R;
        c[j] = (f[j] != -1) ? 1.0f / R : -1.0f;
    }
    uint32_t win = tree_reduce(c);
    *found_at = f[win];
    *cond = (f[win] > 0) ? 1.0f / Rv[win] : -1.0f;
    if (winner) *winner = win;
    if (hops) *hops = h;
    if (path && f[win] > 0) {
        /* the winner writes path[0..found_at] (main.cu:210-214) */
        Stream rng(seed, ORC_TAG_DELETE, eid, win);
        std::vector<uint32_t> traj;
        float R;
        uint64_t hh = 0;
        walk_dec_one(G, s, t, max_steps, rng, R, hh, &traj);
        for (int k = 0; k <= f[win]; ++k) path[k] = traj[k];
    }
}

void heuristic_one(const orc_graph* G, uint32_t u, uint32_t nwalk, uint64_t seed, uint64_t key,
                   uint32_t* path3) {
    for (uint32_t j = 0; j < nwalk; ++j) {
        Stream rng(seed, ORC_TAG_HEUR, key, j);
        const std::vector<Slot>& ru = G->rows[u];
        uint32_t du = (uint32_t)ru.size();
        if (du == 0) { /* r % 0 in the reference; an isolated endpoint yields (u,u,u) */
            path3[j * 3 + 0] = path3[j * 3 + 1] = path3[j * 3 + 2] = u;
            continue;
        }
        uint32_t n1 = ru[rng.next() % du].col;
        const std::vector<Slot>& r1 = G->rows[n1];
        uint32_t d1 = (uint32_t)r1.size();
        uint32_t n2 = r1[rng.next() % d1].col;
        if (n2 == u) n2 = r1[d1 - 1].col;
        path3[j * 3 + 0] = u;
        path3[j * 3 + 1] = n1;
        path3[j * 3 + 2] = n2;
    }
}

/* ------------------------------------------------------------------ */
/* MTX parsing, restating the reference's three-pass mmap parsers.     */
/* Files are read into memory with a terminating NUL; a field scan     */
/* stops there (the reference would run past the end of a file that    */
/* lacks a trailing delimiter).                                       */
/* ------------------------------------------------------------------ */
inline bool is_delim(char c) { return c == ' ' || c == '\n' || c == '\t'; }
inline void skip_field(const char* ss, size_t& next) {
    while (ss[next] != 0 && !is_delim(ss[next])) next++;
    while (is_delim(ss[next])) next++;
}

bool slurp(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.assign((size_t)n, '\0');
    size_t got = n > 0 ? std::fread(&out[0], 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

struct CSR {
    uint64_t V = 0, lines = 0;
    std::vector<uint64_t> begin;
    std::vector<uint32_t> degree, adj;
    std::vector<float> weight;
    std::vector<std::tuple<uint32_t, uint32_t, float>> mtx;
    /* decremental parser only */
    long multiplier = 0;
    std::unordered_map<long, std::pair<uint32_t, uint32_t>> edge_map;
};

int find_digits(long n) {
    int c = 0;
    while (n != 0) {
        n /= 10;
        c++;
    }
    return c;
}

/* base 0: random_walk_incremental/functions.cpp:58-239 (read_graph, is_reverse=1).
 * base 1: random_walk_decremental/functions.cpp:69-281 (CSRGraph ctor) with
 *         D2 resolved and the edge map built in pass 3 (:243-258). */
bool parse_graph(const char* path, int base, bool weighted, bool build_map, CSR& g) {
    std::string buf;
    if (!slurp(path, buf)) return false;
    const char* ss = buf.c_str();
    size_t file_size = buf.size();
    size_t curr = 0, next = 0, edge_count = 0;
    uint32_t v_max = 0, v_min = (uint32_t)(1 << 30);
    int checkwt = 0;
    /* pass 1 (functions.cpp:89-106 / :100-116) */
    while (next < file_size) {
        uint32_t src = (uint32_t)(atoi(ss + curr) - base);
        if (weighted) checkwt++;
        if (checkwt != 3) {
            v_max = std::max(v_max, src);
            v_min = std::min(v_min, src);
        }
        skip_field(ss, next);
        curr = next;
        if (checkwt != 3) edge_count++;
        if (checkwt == 3) checkwt = 0;
    }
    g.lines = edge_count >> 1;
    g.V = (uint64_t)v_max + 1;
    g.begin.assign(g.V + 1, 0);
    g.degree.assign(g.V, 0);
    g.adj.assign(2 * g.lines, 0);
    g.weight.assign(2 * g.lines, 0.0f);
    g.mtx.resize(g.lines);
    if (build_map) {
        g.multiplier = (long)std::pow(10, find_digits((long)g.V));
        g.edge_map.reserve((size_t)((unsigned)g.lines * 1.5));
    }
    /* pass 2: degrees (functions.cpp:130-152 / :149-171) */
    curr = next = 0;
    while (next < file_size) {
        uint32_t src = (uint32_t)(atoi(ss + curr) - base);
        skip_field(ss, next);
        curr = next;
        uint32_t dst = (uint32_t)(atoi(ss + curr) - base);
        skip_field(ss, next);
        if (weighted) skip_field(ss, next);
        curr = next;
        if (src >= g.V || dst >= g.V) return false;
        g.degree[src]++;
        g.degree[dst]++;
    }
    for (uint64_t i = 1; i <= g.V; i++) g.begin[i] = g.begin[i - 1] + g.degree[i - 1];
    std::fill(g.degree.begin(), g.degree.end(), 0u);
    /* pass 3: slots in file order (functions.cpp:168-232 / :188-270) */
    curr = next = 0;
    uint64_t offset = 0;
    while (offset < g.lines) {
        uint32_t src = (uint32_t)(atoi(ss + curr) - base);
        skip_field(ss, next);
        curr = next;
        uint32_t dst = (uint32_t)(atoi(ss + curr) - base);
        uint64_t pos = g.begin[src] + g.degree[src];
        g.adj[pos] = dst;
        if (dst == src) {
            /* self-loop quirk (functions.cpp:186-191): second copy one slot
             * further, its weight slot left at zero */
            uint64_t p2 = g.begin[dst] + g.degree[dst] + 1;
            if (p2 < g.adj.size()) g.adj[p2] = src;
        } else {
            g.adj[g.begin[dst] + g.degree[dst]] = src;
        }
        float wt = 1.0f;
        if (weighted) {
            skip_field(ss, next);
            curr = next;
            wt = (float)atof(ss + curr);
        }
        g.weight[g.begin[src] + g.degree[src]] = wt;
        g.weight[g.begin[dst] + g.degree[dst]] = wt;
        g.mtx[offset] = std::make_tuple(src, dst, wt);
        if (build_map) {
            long a = src, b = dst;
            uint32_t da = g.degree[a], db = g.degree[b];
            if (a > b) {
                std::swap(a, b);
                std::swap(da, db);
            }
            long key = a * g.multiplier + b;
            g.edge_map.insert({key, {da, db}});
        }
        skip_field(ss, next);
        curr = next;
        g.degree[src]++;
        g.degree[dst]++;
        offset++;
    }
    return true;
}

/* random_walk_incremental/functions.cpp:261-370 and
 * random_walk_decremental/functions.cpp:303-420 (same layout logic). */
bool parse_targets(const char* path, bool weighted, uint32_t divisor, std::vector<uint32_t>& nodes,
                   std::vector<float>& weights, uint64_t& count) {
    std::string buf;
    if (!slurp(path, buf)) return false;
    const char* ss = buf.c_str();
    size_t file_size = buf.size();
    size_t curr = 0, next = 0, edge_count = 0;
    int checkwt = 0;
    while (next < file_size) {
        if (weighted) checkwt++;
        skip_field(ss, next);
        curr = next;
        if (checkwt != 3) edge_count++;
        if (checkwt == 3) checkwt = 0;
    }
    edge_count >>= 1;
    nodes.assign(edge_count * 2, 0);
    weights.assign(weighted ? edge_count : 0, 0.0f);
    size_t offset = 0;
    curr = next = 0;
    int interval = (int)((edge_count + divisor - 1) / divisor);
    int initial_interval = interval;
    int write_location = 0;
    while (offset < edge_count) {
        uint32_t source = (uint32_t)(atoi(ss + curr) - 1);
        skip_field(ss, next);
        curr = next;
        uint32_t dest = (uint32_t)(atoi(ss + curr) - 1);
        skip_field(ss, next);
        if (weighted) {
            curr = next;
            weights[offset] = (float)atof(ss + curr);
            skip_field(ss, next);
        }
        curr = next;
        nodes[write_location] = source;
        nodes[write_location + interval] = dest;
        offset++;
        write_location++;
        if ((write_location % interval == 0) && (interval == initial_interval)) {
            write_location += interval;
            if ((uint64_t)write_location + (uint64_t)interval * 2 > edge_count * 2)
                interval = (int)(edge_count - write_location / 2);
        }
    }
    count = edge_count;
    return true;
}

/* Passes 2-3 of the parsers on lines already in memory (0-based ids):
 * random_walk_decremental/functions.cpp:149-270 (degrees, slots in line
 * order with the self-loop layout, edge map with the first line winning).
 * V = largest id + 1. */
void build_from_lines(const uint32_t* u, const uint32_t* v, const float* w, uint64_t m,
                      bool build_map, CSR& g) {
    uint32_t v_max = 0;
    for (uint64_t l = 0; l < m; ++l) v_max = std::max(v_max, std::max(u[l], v[l]));
    g.lines = m;
    g.V = m ? (uint64_t)v_max + 1 : 0;
    g.begin.assign(g.V + 1, 0);
    g.degree.assign(g.V, 0);
    g.adj.assign(2 * m, 0);
    g.weight.assign(2 * m, 0.0f);
    g.mtx.resize(m);
    if (build_map) {
        g.multiplier = (long)std::pow(10, find_digits((long)g.V));
        g.edge_map.reserve((size_t)((unsigned)m * 1.5));
    }
    for (uint64_t l = 0; l < m; ++l) {
        g.degree[u[l]]++;
        g.degree[v[l]]++;
    }
    for (uint64_t i = 1; i <= g.V; i++) g.begin[i] = g.begin[i - 1] + g.degree[i - 1];
    std::fill(g.degree.begin(), g.degree.end(), 0u);
    for (uint64_t l = 0; l < m; ++l) {
        const uint32_t src = u[l], dst = v[l];
        const float wt = w ? w[l] : 1.0f;
        uint64_t pos = g.begin[src] + g.degree[src];
        g.adj[pos] = dst;
        if (dst == src) {
            uint64_t p2 = g.begin[dst] + g.degree[dst] + 1;
            if (p2 < g.adj.size()) g.adj[p2] = src;
        } else {
            g.adj[g.begin[dst] + g.degree[dst]] = src;
        }
        g.weight[g.begin[src] + g.degree[src]] = wt;
        g.weight[g.begin[dst] + g.degree[dst]] = wt;
        g.mtx[l] = std::make_tuple(src, dst, wt);
        if (build_map) {
            long a = src, b = dst;
            uint32_t da = g.degree[a], db = g.degree[b];
            if (a > b) {
                std::swap(a, b);
                std::swap(da, db);
            }
            g.edge_map.insert({a * g.multiplier + b, {da, db}});
        }
        g.degree[src]++;
        g.degree[dst]++;
    }
}

orc_graph* graph_from_csr(const CSR& c) {
    orc_graph* g = new orc_graph;
    g->rows.resize(c.V);
    for (uint64_t v = 0; v < c.V; ++v) {
        auto& r = g->rows[v];
        r.reserve(c.begin[v + 1] - c.begin[v]);
        for (uint64_t p = c.begin[v]; p < c.begin[v + 1]; ++p) r.push_back({c.adj[p], c.weight[p]});
    }
    return g;
}

/* Snapshot batches: first batch ceil(M/div), later ones the same, last
 * shorter (random_walk_incremental/gpu_graph.cuh:135,395-413). */
std::vector<uint32_t> batch_starts(uint32_t m, uint32_t divisor) {
    std::vector<uint32_t> st;
    uint32_t I = (m + divisor - 1) / divisor;
    uint32_t b = 0;
    st.push_back(0);
    if (m == 0) return st;
    while (b < m) {
        b = std::min(m, b + I);
        st.push_back(b);
    }
    return st;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

typedef std::unordered_map<long, std::pair<uint32_t, uint32_t>> EdgeMap;

/* Dense half of removeAndUpdateDelEdges (random_walk_decremental/gpu_graph.cuh:340-400):
 * a<b.  Returns -1 when the key is absent, -2 when a degree would reach 0
 * (the reference asserts; nothing is changed), 1 when removed. */
int dense_remove(orc_graph* G, EdgeMap& dense_map, long mult, long a, long b) {
    long key = a * mult + b;
    auto it = dense_map.find(key);
    if (it == dense_map.end()) return -1;
    auto& ra = G->rows[a];
    auto& rb = G->rows[b];
    if (ra.size() <= 1 || rb.size() <= 1) return -2;
    uint32_t index_b_a = it->second.first, index_a_b = it->second.second;
    dense_map.erase(it);
    uint32_t degree_a = (uint32_t)ra.size() - 1, degree_b = (uint32_t)rb.size() - 1;
    if (degree_a != index_b_a) {
        uint32_t a_last = ra[degree_a].col;
        ra[index_b_a] = ra[degree_a];
        if (a < (long)a_last) dense_map.at(a * mult + a_last).first = index_b_a;
        else dense_map.at((long)a_last * mult + a).second = index_b_a;
    }
    ra.pop_back();
    if (degree_b != index_a_b) {
        uint32_t b_last = rb[degree_b].col;
        rb[index_a_b] = rb[degree_b];
        if (b < (long)b_last) dense_map.at(b * mult + b_last).first = index_a_b;
        else dense_map.at((long)b_last * mult + b).second = index_a_b;
    }
    rb.pop_back();
    return 1;
}

}  // namespace

/* Dense graph with the reference's edge map (test access to the deletion
 * semantics on their own). */
struct orc_dense {
    orc_graph* G;
    long mult;
    EdgeMap map;
};

namespace {

/* The decremental batch loop (random_walk_decremental/main.cu:363-444):
 * deletions (du[i], dv[i]) in file order, snapshot batches of ceil(M/10);
 * per batch removeAndUpdateDelEdges, the walks on G, incrementalUpdateSparse
 * and, for no-path edges, the two-hop recovery.  Writes the reference's
 * stdout lines to `log`. */
struct DecTotals {
    uint64_t hops = 0, no_path = 0, filtered = 0, batches = 0;
    double walk_s = 0, total_s = 0;
};
int dec_loop(orc_graph* G, EdgeMap& dense_map, long mult, std::unordered_set<long>& sparse_map,
             std::unordered_set<long>& added_edges, const uint32_t* du, const uint32_t* dv,
             uint32_t M, uint64_t seed, int threads, std::ostream& log, DecTotals& tot) {
    const uint32_t n_steps = 100, n_walkers = 512, divisor = 10, n_heur = 16;
    std::vector<uint32_t> starts = batch_starts((uint32_t)M, divisor);
    uint32_t nb = (uint32_t)starts.size() - 1;
    uint64_t hops_total = 0, no_path_total = 0, filtered_total = 0;
    double t_walk = 0, t0 = now_s();
    for (uint32_t k = 0; k < nb; ++k) {
        uint32_t b0 = starts[k], n = starts[k + 1] - starts[k];
        /* removeAndUpdateDelEdges */
        std::vector<uint32_t> fa, fb;
        std::vector<uint64_t> fid;
        for (uint32_t i = 0; i < n; ++i) {
            long a = du[b0 + i];
            long b = dv[b0 + i];
            if (a > b) std::swap(a, b);
            long key = a * mult + b;
            int r = dense_remove(G, dense_map, mult, a, b);
            if (r == -1) {
                log << "dense_map not found: a: " << a << " b: " << b << " key: " << key << "\n";
                continue;
            }
            if (r == -2) {
                std::fprintf(stderr, "oracle: deletion leaves a zero degree\n");
                return -5; /* the reference asserts (gpu_graph.cuh:362-363) */
            }
            if (sparse_map.count(key) == 0) {
                log << "sparse_map not found: a: " << a << " b: " << b << " key: " << key << "\n";
            } else {
                sparse_map.erase(key);
                fa.push_back((uint32_t)a);
                fb.push_back((uint32_t)b);
                fid.push_back(b0 + i);
            }
        }
        if (k == 0) {
            log << "Graph copied to GPU\n";
            t0 = now_s();
        }
        uint32_t nf = (uint32_t)fa.size();
        filtered_total += nf;
        std::vector<int32_t> f(nf);
        std::vector<float> c(nf);
        std::vector<uint64_t> h(nf);
        std::vector<uint32_t> path((uint64_t)nf * n_steps);
        double tw0 = now_s();
        orc_walk_dec_batch(G, nf, fa.data(), fb.data(), fid.data(), n_walkers, n_steps, seed,
                           threads, f.data(), c.data(), nullptr, h.data(), path.data());
        t_walk += now_s() - tw0;
        log << " result copied\n";
        /* incrementalUpdateSparse */
        std::vector<uint32_t> heur;
        std::vector<uint64_t> heur_key;
        uint64_t add_count = 0;
        for (uint32_t i = 0; i < nf; ++i) {
            hops_total += h[i];
            if (f[i] != -1) {
                long from = path[(uint64_t)i * n_steps];
                for (int j = 0; j < f[i]; ++j) {
                    long to = path[(uint64_t)i * n_steps + j + 1];
                    long a = from, b = to;
                    if (a > b) std::swap(a, b);
                    long key = a * mult + b;
                    if (sparse_map.count(key) == 0) {
                        add_count++;
                        sparse_map.insert(key);
                        added_edges.insert(key);
                    }
                    from = to;
                }
            } else {
                heur.push_back(fa[i]);
                heur.push_back(fb[i]);
                heur_key.push_back(fid[i] << 1);
                heur_key.push_back((fid[i] << 1) | 1u);
            }
        }
        no_path_total += heur.size() / 2;
        log << "No path count: " << heur.size() / 2 << "\n";
        log << "Added edges: " << add_count << "\n";
        if (!heur.empty()) {
            std::vector<uint32_t> p3((uint64_t)heur.size() * n_heur * 3);
            for (size_t e = 0; e < heur.size(); ++e)
                heuristic_one(G, heur[e], n_heur, seed, heur_key[e], p3.data() + e * n_heur * 3);
            log << " result copied\n";
            for (size_t e = 0; e < heur.size(); ++e)
                for (uint32_t j = 0; j < n_heur; ++j) {
                    long from = p3[(e * n_heur + j) * 3];
                    for (int q = 1; q < 3; ++q) {
                        long to = p3[(e * n_heur + j) * 3 + q];
                        long a = from, b = to;
                        if (a > b) std::swap(a, b);
                        long key = a * mult + b;
                        if (sparse_map.count(key) == 0) sparse_map.insert(key);
                        from = to;
                    }
                }
        }
    }
    if (nb == 0) /* an empty first batch still launches (main.cu:363-397) */
        log << "Graph copied to GPU\n result copied\nNo path count: 0\nAdded edges: 0\n";
    tot.hops = hops_total;
    tot.no_path = no_path_total;
    tot.filtered = filtered_total;
    tot.batches = nb;
    tot.walk_s = t_walk;
    tot.total_s = now_s() - t0;
    return 0;
}

}  // namespace

/* ==================================================================== */
/* C ABI                                                                */
/* ==================================================================== */
extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

uint32_t orc_draw(uint64_t seed, uint32_t tag, uint64_t edge_id, uint32_t walker, uint32_t d) {
    Stream s(seed, tag, edge_id, walker);
    uint32_t r = 0;
    for (uint32_t k = 0; k <= d; ++k) r = s.next();
    return r;
}

uint32_t orc_tree_reduce(const float* cond, uint32_t W) {
    return tree_reduce(std::vector<float>(cond, cond + W));
}

orc_graph* orc_graph_new(uint32_t V, const uint64_t* row_ptr, const uint32_t* col, const float* w) {
    orc_graph* g = new orc_graph;
    g->rows.resize(V);
    for (uint32_t v = 0; v < V; ++v)
        for (uint64_t p = row_ptr[v]; p < row_ptr[v + 1]; ++p) g->rows[v].push_back({col[p], w[p]});
    return g;
}
void orc_graph_free(orc_graph* g) { delete g; }
uint32_t orc_graph_nv(const orc_graph* g) { return (uint32_t)g->rows.size(); }
uint64_t orc_graph_nslots(const orc_graph* g) {
    uint64_t n = 0;
    for (auto& r : g->rows) n += r.size();
    return n;
}
void orc_graph_export(const orc_graph* g, uint64_t* row_ptr, uint32_t* col, float* w) {
    uint64_t p = 0;
    for (size_t v = 0; v < g->rows.size(); ++v) {
        row_ptr[v] = p;
        for (auto& s : g->rows[v]) {
            col[p] = s.col;
            w[p] = s.w;
            ++p;
        }
    }
    row_ptr[g->rows.size()] = p;
}
void orc_graph_insert(orc_graph* g, uint32_t u, uint32_t v, float w) {
    g->rows[u].push_back({v, w});
    g->rows[v].push_back({u, w});
}

void orc_walk_inc_edge(const orc_graph* H, uint32_t s, uint32_t t, float w_e, float thr,
                       uint32_t W, uint32_t max_steps, uint64_t seed, uint64_t edge_id,
                       int32_t* found_at, float* cond, uint32_t* winner, uint64_t* hops) {
    inc_edge(H, s, t, w_e, thr, W, max_steps, seed, edge_id, found_at, cond, winner, hops);
}

void orc_walk_dec_edge(const orc_graph* G, uint32_t s, uint32_t t, uint32_t W,
                       uint32_t max_steps, uint64_t seed, uint64_t edge_id, int32_t* found_at,
                       float* cond, uint32_t* winner, uint64_t* hops, uint32_t* path) {
    dec_edge(G, s, t, W, max_steps, seed, edge_id, found_at, cond, winner, hops, path);
}

uint32_t orc_walk_trace(const orc_graph* g, int kind, uint32_t s, uint32_t t, float w_e,
                        float thr, uint32_t max_steps, uint64_t seed, uint64_t edge_id,
                        uint32_t walker, uint32_t* traj, int32_t* found_at, float* R) {
    std::vector<uint32_t> tr;
    uint64_t h = 0;
    float r = 0.f;
    int f;
    if (kind == 1) {
        Stream rng(seed, ORC_TAG_INSERT, edge_id, walker);
        f = walk_inc_one(g, s, t, w_e, thr, max_steps, rng, r, h, &tr);
    } else {
        Stream rng(seed, ORC_TAG_DELETE, edge_id, walker);
        f = walk_dec_one(g, s, t, max_steps, rng, r, h, &tr);
    }
    for (size_t k = 0; k < tr.size(); ++k) traj[k] = tr[k];
    *found_at = f;
    *R = r;
    return (uint32_t)tr.size();
}

void orc_walk_heuristic(const orc_graph* G, uint32_t u, uint32_t nwalk, uint64_t seed,
                        uint64_t key, uint32_t* path3) {
    heuristic_one(G, u, nwalk, seed, key, path3);
}

void orc_walk_inc_batch(const orc_graph* H, uint32_t n, const uint32_t* src, const uint32_t* dst,
                        const float* w, const uint64_t* edge_id, float thr, uint32_t W,
                        uint32_t max_steps, uint64_t seed, int threads, int32_t* found_at,
                        float* cond, uint32_t* winner, uint64_t* hops) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t e = 0; e < (int64_t)n; ++e) {
        uint64_t h = 0;
        uint32_t win = 0;
        inc_edge(H, src[e], dst[e], w[e], thr, W, max_steps, seed, edge_id[e], &found_at[e],
                 &cond[e], &win, &h);
        if (winner) winner[e] = win;
        if (hops) hops[e] = h;
    }
}

void orc_walk_dec_batch(const orc_graph* G, uint32_t n, const uint32_t* src, const uint32_t* dst,
                        const uint64_t* edge_id, uint32_t W, uint32_t max_steps, uint64_t seed,
                        int threads, int32_t* found_at, float* cond, uint32_t* winner,
                        uint64_t* hops, uint32_t* path) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int64_t e = 0; e < (int64_t)n; ++e) {
        uint64_t h = 0;
        uint32_t win = 0;
        dec_edge(G, src[e], dst[e], W, max_steps, seed, edge_id[e], &found_at[e], &cond[e], &win,
                 &h, path ? path + (uint64_t)e * max_steps : nullptr);
        if (winner) winner[e] = win;
        if (hops) hops[e] = h;
    }
}

int orc_decremental_lines(uint32_t V, const uint32_t* g_u, const uint32_t* g_v, const float* g_w,
                          uint64_t g_m, const uint32_t* h_u, const uint32_t* h_v, uint64_t h_m,
                          const uint32_t* d_u, const uint32_t* d_v, uint32_t d_m, uint64_t seed,
                          int threads, orc_run_stats* st, uint32_t* added_u, uint32_t* added_v,
                          uint64_t added_cap) {
    CSR dense;
    build_from_lines(g_u, g_v, g_w, g_m, true, dense);
    if (dense.V != V) return -4; /* V is set by the largest id, as in a file */
    const long mult = dense.multiplier;
    orc_graph* G = graph_from_csr(dense);
    std::unordered_set<long> sparse_map, added_edges;
    for (uint64_t i = 0; i < h_m; ++i) {
        long a = h_u[i], b = h_v[i];
        if (a > b) std::swap(a, b);
        sparse_map.insert(a * mult + b);
    }
    const uint64_t before = sparse_map.size();
    std::ostringstream log;
    DecTotals tot;
    int rc = dec_loop(G, dense.edge_map, mult, sparse_map, added_edges, d_u, d_v, d_m, seed,
                      threads, log, tot);
    delete G;
    if (rc != 0) return rc;
    std::vector<long> keys(added_edges.begin(), added_edges.end());
    std::sort(keys.begin(), keys.end());
    for (size_t i = 0; i < keys.size() && i < added_cap; ++i) {
        added_u[i] = (uint32_t)(keys[i] / mult);
        added_v[i] = (uint32_t)(keys[i] % mult);
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_updates = d_m;
        st->n_batches = tot.batches;
        st->hops = tot.hops;
        st->sparse_before = before;
        st->sparse_after = sparse_map.size();
        st->added_edges = added_edges.size();
        st->no_path = tot.no_path;
        st->filtered = tot.filtered;
        st->walk_seconds = tot.walk_s;
        st->total_seconds = tot.total_s;
    }
    return 0;
}

/* `ostream << float` as the reference's writer uses it
 * (random_walk_incremental/gpu_graph.cuh:434-445). */
int orc_format_float(float f, char* buf, int cap) {
    std::ostringstream os;
    os << f;
    std::string s = os.str();
    std::snprintf(buf, (size_t)cap, "%s", s.c_str());
    return (int)s.size();
}

orc_dense* orc_dense_open(const char* path) {
    CSR c;
    if (!parse_graph(path, 1, true, true, c)) return nullptr;
    orc_dense* d = new orc_dense;
    d->G = graph_from_csr(c);
    d->mult = c.multiplier;
    d->map = std::move(c.edge_map);
    return d;
}
int orc_dense_remove(orc_dense* d, uint32_t u, uint32_t v) {
    long a = u, b = v;
    if (a > b) std::swap(a, b);
    return dense_remove(d->G, d->map, d->mult, a, b);
}
orc_graph* orc_dense_graph(orc_dense* d) { return d->G; }
int64_t orc_dense_multiplier(const orc_dense* d) { return d->mult; }
void orc_dense_free(orc_dense* d) {
    if (!d) return;
    delete d->G;
    delete d;
}

uint32_t orc_batches(uint32_t m, uint32_t divisor, uint32_t* starts) {
    std::vector<uint32_t> st = batch_starts(m, divisor);
    for (size_t k = 0; k < st.size(); ++k) starts[k] = st[k];
    return (uint32_t)st.size() - 1;
}

int64_t orc_read_graph0(const char* path, uint64_t* lines, uint64_t* begin, uint32_t* adj,
                        float* weight) {
    CSR g;
    if (!parse_graph(path, 0, true, false, g)) return -1;
    if (lines) *lines = g.lines;
    if (begin) std::copy(g.begin.begin(), g.begin.end(), begin);
    if (adj) std::copy(g.adj.begin(), g.adj.end(), adj);
    if (weight) std::copy(g.weight.begin(), g.weight.end(), weight);
    return (int64_t)g.V;
}

int64_t orc_read_targets(const char* path, int weighted, uint32_t divisor, uint32_t* nodes,
                         float* weights) {
    std::vector<uint32_t> nd;
    std::vector<float> wt;
    uint64_t cnt = 0;
    if (!parse_targets(path, weighted != 0, divisor, nd, wt, cnt)) return -1;
    if (nodes) std::copy(nd.begin(), nd.end(), nodes);
    if (weights) std::copy(wt.begin(), wt.end(), weights);
    return (int64_t)cnt;
}

/* The incremental batch loop on arrays (random_walk_incremental/main.cu:297-340,
 * gpu_graph.cuh:292-336): snapshot batches of ceil(m/divisor), walks of a
 * batch on H as it stood at batch start, then the no-path edges appended in
 * batch order.  Edge i's draws are keyed by i. */
int orc_incremental_arrays(orc_graph* H, uint32_t m, const uint32_t* src, const uint32_t* dst,
                           const float* w, uint32_t divisor, float thr, uint32_t W,
                           uint32_t max_steps, uint64_t seed, int threads, int32_t* found_at,
                           orc_run_stats* st) {
    std::vector<uint32_t> starts = batch_starts(m, divisor);
    const uint32_t nb = (uint32_t)starts.size() - 1;
    std::vector<uint64_t> id(m);
    for (uint32_t i = 0; i < m; ++i) id[i] = i;
    std::vector<float> c(m);
    std::vector<uint64_t> h(m);
    uint64_t added = 0;
    double t_walk = 0, t0 = now_s();
    for (uint32_t k = 0; k < nb; ++k) {
        const uint32_t b0 = starts[k], n = starts[k + 1] - starts[k];
        double tw = now_s();
        orc_walk_inc_batch(H, n, src + b0, dst + b0, w + b0, id.data() + b0, thr, W, max_steps,
                           seed, threads, found_at + b0, c.data() + b0, nullptr, h.data() + b0);
        t_walk += now_s() - tw;
        for (uint32_t i = b0; i < b0 + n; ++i)
            if (found_at[i] == -1) {
                orc_graph_insert(H, src[i], dst[i], w[i]);
                ++added;
            }
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_updates = m;
        st->n_added = added;
        st->n_batches = nb;
        for (uint32_t i = 0; i < m; ++i) st->hops += h[i];
        st->walk_seconds = t_walk;
        st->total_seconds = now_s() - t0;
    }
    return 0;
}

/* Whole incremental run: random_walk_incremental/main.cu:240-362 with the
 * host update of gpu_graph.cuh:292-336 and the writer :427-454. */
int orc_incremental_run(const char* dir, float thr, uint32_t W, uint64_t seed, int threads,
                        const char* log_path, orc_run_stats* st) {
    std::string d(dir);
    std::string graph_file = d + "/adj_sparse.mtx";
    std::string target_file = d + "/ext.mtx";
    std::string output_file = d + "/updated_adj_sparse.mtx";
    std::ostringstream log;
    CSR csr;
    if (!parse_graph(graph_file.c_str(), 0, true, false, csr)) return -1;
    std::vector<uint32_t> nodes;
    std::vector<float> tw;
    uint64_t M = 0;
    if (!parse_targets(target_file.c_str(), true, 10, nodes, tw, M)) return -2;
    log << "after read_targets\n";
    log << "CSRGraph deleted\n";
    orc_graph* H = graph_from_csr(csr);
    log << "Graph copied to GPU\n";
    /* the 10-segment layout: batch k holds sources at nodes[2kI .. 2kI+n_k)
     * and targets right after them */
    std::vector<uint32_t> starts = batch_starts((uint32_t)M, 10);
    uint32_t nb = (uint32_t)starts.size() - 1;
    std::vector<std::tuple<uint32_t, uint32_t, float>> ext;
    uint64_t no_path = 0, hops_total = 0;
    double t_walk = 0, t0 = now_s();
    for (uint32_t k = 0; k < nb; ++k) {
        uint32_t b0 = starts[k], n = starts[k + 1] - starts[k];
        std::vector<uint32_t> s(n), t(n);
        std::vector<uint64_t> id(n);
        for (uint32_t i = 0; i < n; ++i) {
            s[i] = nodes[2 * (uint64_t)b0 + i];
            t[i] = nodes[2 * (uint64_t)b0 + n + i];
            id[i] = b0 + i;
        }
        std::vector<int32_t> f(n);
        std::vector<float> c(n);
        std::vector<uint64_t> h(n);
        double tw0 = now_s();
        orc_walk_inc_batch(H, n, s.data(), t.data(), tw.data() + b0, id.data(), thr, W, 100, seed,
                           threads, f.data(), c.data(), nullptr, h.data());
        t_walk += now_s() - tw0;
        for (uint32_t i = 0; i < n; ++i) {
            hops_total += h[i];
            if (f[i] == -1) {
                no_path++;
                ext.emplace_back(s[i], t[i], tw[b0 + i]);
                orc_graph_insert(H, s[i], t[i], tw[b0 + i]);
            }
        }
        log << "no path count: " << no_path << "\n";
    }
    /* M <= 1: the loop launches one empty batch after the first (main.cu:314-329) */
    for (uint32_t k = nb; k < 2 && M <= 1; ++k) log << "no path count: " << no_path << "\n";
    double t1 = now_s();
    log << "No path %: " << (float)no_path / (float)M * 100 << "\n";
    log << "Edges added: " << no_path << " out of " << M << " samples\n";
    log << "saving mtx to " << output_file << "\n";
    {
        std::ofstream out(output_file, std::ios::out);
        if (!out.is_open()) return -3;
        /* vertex ids pass through float (gpu_graph.cuh:174-183, 315-317) */
        for (auto& e : csr.mtx) {
            float a = (float)std::get<0>(e), b = (float)std::get<1>(e);
            out << static_cast<int>(a) + 1 << " " << static_cast<int>(b) + 1 << " "
                << std::get<2>(e) << "\n";
        }
        for (auto& e : ext) {
            float a = (float)(int)std::get<0>(e), b = (float)(int)std::get<1>(e);
            out << static_cast<int>(a) + 1 << " " << static_cast<int>(b) + 1 << " "
                << std::get<2>(e) << "\n";
        }
    }
    log << "Updated graph saved to " << output_file << "\n";
    log << "Random walk data deleted\n";
    log << "Incremental graph sparsification completed successfully!\n";
    log << "gpu_graph\ngpu_graph deleted\nTargets deleted\nCSRGraph deleted\n";
    if (log_path) {
        std::ofstream lf(log_path);
        lf << log.str();
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_updates = M;
        st->n_added = no_path;
        st->n_batches = nb;
        st->hops = hops_total;
        st->walk_seconds = t_walk;
        st->total_seconds = t1 - t0;
    }
    delete H;
    return 0;
}

/* Whole decremental run: random_walk_decremental/main.cu:318-468 with
 * removeAndUpdateDelEdges (gpu_graph.cuh:328-477), incrementalUpdateSparse
 * (:509-559), heuristicRecovery (:577-615) and save_added_edges (:633-646). */
int orc_decremental_run(const char* dir, uint64_t seed, int threads, const char* log_path,
                        orc_run_stats* st) {
    std::string d(dir);
    std::string sparse_file = d + "/updated_adj_sparse.mtx";
    std::string dense_file = d + "/updated_dense.mtx";
    std::string del_file = d + "/del.mtx";
    std::string output_file = d + "/added_edges.mtx";
    const uint32_t divisor = 10;
    std::ostringstream log;
    log << "Read graph and sparsifier\n";
    CSR dense, sparse;
    if (!parse_graph(dense_file.c_str(), 1, true, true, dense)) return -1;
    if (!parse_graph(sparse_file.c_str(), 1, true, true, sparse)) return -1;
    log << "Read targets\n";
    std::vector<uint32_t> nodes;
    std::vector<float> unused;
    uint64_t M = 0;
    if (!parse_targets(del_file.c_str(), false, divisor, nodes, unused, M)) return -2;
    if (dense.V != sparse.V || dense.multiplier != sparse.multiplier) return -4;
    const long mult = dense.multiplier;
    const uint64_t V = dense.V;
    orc_graph* G = graph_from_csr(dense);
    auto& dense_map = dense.edge_map;
    std::unordered_set<long> sparse_map;
    for (auto& kv : sparse.edge_map) sparse_map.insert(kv.first);
    std::unordered_set<long> added_edges;
    const uint64_t old_edge_num = sparse_map.size();

    /* the 10-segment layout decoded back to file order */
    std::vector<uint32_t> du(M), dv(M);
    {
        std::vector<uint32_t> st0 = batch_starts((uint32_t)M, divisor);
        for (size_t k = 0; k + 1 < st0.size(); ++k) {
            const uint32_t b0 = st0[k], n = st0[k + 1] - st0[k];
            for (uint32_t i = 0; i < n; ++i) {
                du[b0 + i] = nodes[2 * (uint64_t)b0 + i];
                dv[b0 + i] = nodes[2 * (uint64_t)b0 + n + i];
            }
        }
    }
    DecTotals tot;
    int rc = dec_loop(G, dense_map, mult, sparse_map, added_edges, du.data(), dv.data(),
                      (uint32_t)M, seed, threads, log, tot);
    if (rc != 0) {
        delete G;
        return rc;
    }
    const uint32_t nb = (uint32_t)tot.batches;
    const uint64_t hops_total = tot.hops, no_path_total = tot.no_path,
                   filtered_total = tot.filtered;
    const double t_walk = tot.walk_s;
    uint64_t new_edge_num = sparse_map.size();
    log << "Increased edges number: " << (size_t)(new_edge_num - old_edge_num) << "\n";
    log << "Density: " << (float)(size_t)(new_edge_num - (uint32_t)V) / (uint32_t)V * 100 << " %\n";
    log << "Size of added edges: " << added_edges.size() << "\n";
    {
        std::ofstream out(output_file, std::ios::out);
        if (!out.is_open()) return -3;
        for (long key : added_edges) out << key / mult + 1 << " " << key % mult + 1 << "\n";
    }
    log << "Added edges saved to: " << output_file << "\n";
    log << "Random walk data deleted\n";
    log << "Decremental graph sparsification completed successfully!\n";
    log << "CSRGraph deleted\nCSRGraph deleted\n";
    if (log_path) {
        std::ofstream lf(log_path);
        lf << log.str();
    }
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->n_updates = M;
        st->n_batches = nb;
        st->hops = hops_total;
        st->sparse_before = old_edge_num;
        st->sparse_after = new_edge_num;
        st->added_edges = added_edges.size();
        st->no_path = no_path_total;
        st->filtered = filtered_total;
        st->walk_seconds = t_walk;
        st->total_seconds = tot.total_s;
    }
    delete G;
    return 0;
}

}  // extern "C"

/* ==================================================================== */
/* Mixed insertion/deletion stream (C++ restatement of oracle/stream.py, */
/* the sequential stream oracle; same semantics, same order of edits)     */
/* ==================================================================== */
namespace {

/* Slot of b in row a, or -1 (rows are short: a scan). */
inline int row_pos(const orc_graph* g, uint32_t a, uint32_t b) {
    const std::vector<Slot>& r = g->rows[a];
    for (size_t i = 0; i < r.size(); ++i)
        if (r[i].col == b) return (int)i;
    return -1;
}

inline void row_append(orc_graph* g, uint32_t a, uint32_t b, float w) {
    g->rows[a].push_back(Slot{b, w});
    g->rows[b].push_back(Slot{a, w});
}

inline void row_add_weight(orc_graph* g, uint32_t a, uint32_t b, float dw) {
    Slot& x = g->rows[a][row_pos(g, a, b)];
    x.w = x.w + dw;
    Slot& y = g->rows[b][row_pos(g, b, a)];
    y.w = y.w + dw;
}

/* swap-with-last in row a, then in row b */
inline void row_delete(orc_graph* g, uint32_t a, uint32_t b) {
    for (int side = 0; side < 2; ++side) {
        const uint32_t x = side ? b : a, y = side ? a : b;
        std::vector<Slot>& r = g->rows[x];
        r[row_pos(g, x, y)] = r.back();
        r.pop_back();
    }
}

/* Insertion walks deciding only: an edge is found once any walker reaches
 * its target (the stream reports found / not found; the walks themselves
 * are random_walk_incremental/main.cu:79-163). */
bool inc_decide(const orc_graph* H, uint32_t s, uint32_t t, float w_e, float thr, uint32_t W,
                uint32_t max_steps, uint64_t seed, uint64_t eid, uint64_t* hops) {
    for (uint32_t j = 0; j < W; ++j) {
        Stream rng(seed, ORC_TAG_INSERT, eid, j);
        float R;
        if (walk_inc_one(H, s, t, w_e, thr, max_steps, rng, R, *hops, nullptr) != -1) return true;
    }
    return false;
}

}  // namespace

struct orc_stream {
    orc_graph G, H;
    float thr;
    uint32_t w_ins, w_del, steps_ins, steps_del, heur;
    uint64_t seed;
    int threads;
    uint64_t base = 0;
    orc_stream_stats st{};
};

namespace {

double since(double t0) { return now_s() - t0; }

void stream_recover(orc_stream* s, uint32_t x, uint32_t y) {
    if (x == y) return;
    const uint32_t a = std::min(x, y), b = std::max(x, y);
    if (row_pos(&s->H, a, b) < 0) row_append(&s->H, a, b, s->G.rows[a][row_pos(&s->G, a, b)].w);
}

void stream_insert(orc_stream* s, const uint32_t* u, const uint32_t* v, const float* w,
                   uint32_t n, int32_t* out) {
    std::vector<uint32_t> keep;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        if (a == b) {
            out[i] = 8;  // invalid
        } else if (row_pos(&s->G, a, b) >= 0) {
            row_add_weight(&s->G, a, b, w[i]);
            if (row_pos(&s->H, a, b) >= 0) row_add_weight(&s->H, a, b, w[i]);
            out[i] = 1;  // weight update
        } else {
            row_append(&s->G, a, b, w[i]);
            keep.push_back(i);
        }
    }
    const int64_t nk = (int64_t)keep.size();
    std::vector<uint8_t> found(nk);
    std::vector<uint64_t> hops(nk, 0);
    const double tw = now_s();
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->threads > 0 ? s->threads : 1)
    for (int64_t j = 0; j < nk; ++j) {
        const uint32_t i = keep[j];
        found[j] = inc_decide(&s->H, u[i], v[i], w[i], s->thr, s->w_ins, s->steps_ins, s->seed,
                              s->base + i, &hops[j]);
    }
    s->st.walk_seconds += since(tw);
    for (int64_t j = 0; j < nk; ++j) {
        const uint32_t i = keep[j];
        s->st.hops += hops[j];
        if (!found[j]) {
            row_append(&s->H, std::min(u[i], v[i]), std::max(u[i], v[i]), w[i]);
            out[i] = 3;  // added
        } else {
            out[i] = 2;  // found
        }
    }
}

void stream_delete(orc_stream* s, const uint32_t* u, const uint32_t* v, uint32_t n,
                   int32_t* out) {
    std::vector<uint32_t> keep;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        if (a == b) {
            out[i] = 8;
        } else if (row_pos(&s->G, a, b) < 0) {
            out[i] = 4;  // skip
        } else {
            row_delete(&s->G, a, b);
            if (row_pos(&s->H, a, b) >= 0) {
                row_delete(&s->H, a, b);
                keep.push_back(i);
            } else {
                out[i] = 5;  // G only
            }
        }
    }
    const int64_t nk = (int64_t)keep.size();
    const uint32_t ms = s->steps_del;
    std::vector<uint32_t> src(nk), dst(nk), path((size_t)nk * ms);
    std::vector<uint64_t> eid(nk), hops(nk);
    std::vector<int32_t> f(nk);
    std::vector<float> c(nk);
    for (int64_t j = 0; j < nk; ++j) {
        const uint32_t i = keep[j];
        src[j] = std::min(u[i], v[i]);
        dst[j] = std::max(u[i], v[i]);
        eid[j] = s->base + i;
    }
    const double tw = now_s();
    orc_walk_dec_batch(&s->G, (uint32_t)nk, src.data(), dst.data(), eid.data(), s->w_del, ms,
                       s->seed, s->threads, f.data(), c.data(), nullptr, hops.data(),
                       path.data());
    s->st.walk_seconds += since(tw);
    for (int64_t j = 0; j < nk; ++j) {
        s->st.hops += hops[j];
        if (f[j] > 0) {
            for (int q = 1; q <= f[j]; ++q)
                stream_recover(s, path[(size_t)j * ms + q - 1], path[(size_t)j * ms + q]);
            out[keep[j]] = 6;  // recovered
        }
    }
    std::vector<uint32_t> p3((size_t)s->heur * 3);
    for (int64_t j = 0; j < nk; ++j) {
        if (f[j] > 0) continue;
        const uint64_t e = s->base + keep[j];
        for (int side = 0; side < 2; ++side) {
            heuristic_one(&s->G, side ? dst[j] : src[j], s->heur, s->seed, (e << 1) | side,
                          p3.data());
            for (uint32_t q = 0; q < s->heur; ++q) {
                stream_recover(s, p3[q * 3], p3[q * 3 + 1]);
                stream_recover(s, p3[q * 3 + 1], p3[q * 3 + 2]);
            }
        }
        out[keep[j]] = 7;  // heuristic
    }
}

}  // namespace

extern "C" {

orc_stream* orc_stream_new(uint32_t V, const uint64_t* g_rp, const uint32_t* g_col,
                           const float* g_w, const uint64_t* h_rp, const uint32_t* h_col,
                           const float* h_w, float thr, uint32_t w_ins, uint32_t w_del,
                           uint32_t steps_ins, uint32_t steps_del, uint32_t heur, uint64_t seed,
                           int threads) {
    auto* s = new orc_stream;
    for (int k = 0; k < 2; ++k) {
        const uint64_t* rp = k ? h_rp : g_rp;
        const uint32_t* col = k ? h_col : g_col;
        const float* w = k ? h_w : g_w;
        orc_graph& g = k ? s->H : s->G;
        g.rows.resize(V);
        for (uint32_t x = 0; x < V; ++x)
            for (uint64_t p = rp[x]; p < rp[x + 1]; ++p) g.rows[x].push_back(Slot{col[p], w[p]});
    }
    s->thr = thr;
    s->w_ins = w_ins;
    s->w_del = w_del;
    s->steps_ins = steps_ins;
    s->steps_del = steps_del;
    s->heur = heur;
    s->seed = seed;
    s->threads = threads;
    return s;
}

int orc_stream_apply(orc_stream* s, int op, const uint32_t* u, const uint32_t* v, const float* w,
                     uint32_t n, int32_t* out) {
    const uint32_t V = (uint32_t)s->G.rows.size();
    for (uint32_t i = 0; i < n; ++i)
        if (u[i] >= V || v[i] >= V) return -1;
    const double t0 = now_s();
    if (op == 0)
        stream_insert(s, u, v, w, n, out);
    else
        stream_delete(s, u, v, n, out);
    s->base += n;
    s->st.updates += n;
    s->st.batches += 1;
    s->st.total_seconds += since(t0);
    return 0;
}

const orc_graph* orc_stream_graph(const orc_stream* s, int which) { return which ? &s->H : &s->G; }

void orc_stream_get_stats(const orc_stream* s, orc_stream_stats* st) { *st = s->st; }

void orc_stream_free(orc_stream* s) { delete s; }

}  // extern "C"
