// graph_store.hip -- device graph store with per-row slack and in-place edits.
//
// HBM layout (one graph):
//   rows[V]        uint2 (first slot, degree)                          8 B per vertex
//   adj[capacity]  uint4 (neighbour, weight bits, neighbour's first
//                         slot, neighbour's degree)                    16 B per slot
// Each row owns `cap` consecutive slots; the first `deg` are live, in the
// reference's slot order.  Inserts append at slot deg; a full row moves to
// the arena tail with its order kept (the reference's relocation,
// random_walk_incremental/gpu_graph.cuh:357-384).  Deletes move the row's
// last slot into the hole (random_walk_decremental/gpu_graph.cuh:368-399).
// Every slot also holds the row descriptor of the vertex it points to, so a
// walk hop is one load; an edit that changes a row therefore rewrites that
// copy in each neighbour's slots that point back to it ("refresh").
//
// The host keeps a mirror of rows and slots.  An edit is planned on the host
// in reference order and shipped as a delta (final slot values and row
// descriptors) that one kernel applies; the reference instead re-uploads the
// whole arena after every batch (gpu_graph.cuh:333-335, :462-465).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace dyg {

hipError_t Pinned::reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = std::max(n, (size_t)1 << 20);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e == hipSuccess) bytes = want;
    return e;
}
Pinned::~Pinned() {
    if (p) (void)hipHostFree(p);
}

hipError_t DevBuf::reserve(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = std::max(n, (size_t)1 << 20);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
}
DevBuf::~DevBuf() {
    if (p) (void)hipFree(p);
}

long key_multiplier(uint64_t V) {
    // 10^digits(V) (random_walk_decremental/functions.cpp:129-130)
    long m = 1;
    for (uint64_t n = V; n != 0; n /= 10) m *= 10;
    return m;
}

namespace {

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Grow the slot arena to hold at least `need` slots (device copy of the used
// prefix).  The host mirror grows when the host next plans an edit.  A growth
// leaves an eighth of `need` (or half the old size) on top, so that the small
// rises of a headroom reservation after each call (slots in use grow with
// every relocated row) do not each cost a reallocation and a copy.
dyg_status grow_arena(dyg_graph g, uint64_t need) {
    if (need <= g->cap_slots) return DYG_OK;
    uint64_t ncap = std::max<uint64_t>(need + need / 8, g->cap_slots + g->cap_slots / 2 + 1024);
    if (ncap > 0xFFFFFFFFull) ncap = 0xFFFFFFFFull;
    if (need > ncap) {
        set_error(g->ctx, "slot arena exceeds 2^32 slots");
        return DYG_ERR_NOMEM;
    }
    uint4* nadj = nullptr;
    if (hipMalloc(&nadj, ncap * sizeof(uint4)) != hipSuccess) {
        set_error(g->ctx, "device allocation of the slot arena failed");
        return DYG_ERR_NOMEM;
    }
    DYG_HIP(g->ctx, hipMemcpyAsync(nadj, g->d_adj, g->used_slots * sizeof(uint4),
                                   hipMemcpyDeviceToDevice, g->ctx->stream));
    DYG_HIP(g->ctx, hipStreamSynchronize(g->ctx->stream));
    DYG_HIP(g->ctx, hipFree(g->d_adj));
    g->d_adj = nadj;
    g->cap_slots = ncap;
    return DYG_OK;
}

// Rewrite, in every slot that points to one of `rows_changed`, the copy of
// that row's descriptor (host mirror), and record the slots touched.
void refresh_neighbour_copies(dyg_graph g, const std::vector<uint32_t>& rows_changed,
                              std::vector<uint32_t>& dirty) {
    for (uint32_t r : rows_changed) {
        const uint32_t sr = g->start[r], dr = g->deg[r];
        for (uint32_t q = 0; q < dr; ++q) {
            const uint32_t x = g->adj[sr + q].x;
            const uint32_t sx = g->start[x], dx = g->deg[x];
            for (uint32_t k = 0; k < dx; ++k) {
                uint4& sl = g->adj[sx + k];
                if (sl.x == r && (sl.z != sr || sl.w != dr)) {
                    sl.z = sr;
                    sl.w = dr;
                    dirty.push_back(sx + k);
                }
            }
        }
    }
}

// Ship the final values of `dirty` slots and the descriptors of `rows`.
dyg_status ship_delta(dyg_graph g, std::vector<uint32_t>& dirty, std::vector<uint32_t>& rows) {
    std::sort(dirty.begin(), dirty.end());
    dirty.erase(std::unique(dirty.begin(), dirty.end()), dirty.end());
    std::sort(rows.begin(), rows.end());
    rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
    g->delta.clear();
    g->delta.reserve(dirty.size() * 5 + rows.size() * 3);
    for (uint32_t sl : dirty) {
        const uint4 a = g->adj[sl];
        g->delta.insert(g->delta.end(), {sl, a.x, a.y, a.z, a.w});
    }
    for (uint32_t r : rows) g->delta.insert(g->delta.end(), {r, g->start[r], g->deg[r]});
    const size_t bytes = g->delta.size() * sizeof(uint32_t);
    if (bytes == 0) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    // the staging buffers are shared per context: wait for earlier users
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, g->d_delta.reserve(bytes));
    DYG_HIP(ctx, ctx->stage.reserve(bytes));
    std::memcpy(ctx->stage.p, g->delta.data(), bytes);
    DYG_HIP(ctx, hipMemcpyAsync(g->d_delta.p, ctx->stage.p, bytes, hipMemcpyHostToDevice,
                                ctx->stream));
    DYG_HIP(ctx, launch_apply_delta(g->d_rows, g->d_adj, (const uint32_t*)g->d_delta.p,
                                    (uint32_t)dirty.size(), (uint32_t)rows.size(), ctx->stream));
    return DYG_OK;
}

}  // namespace

dyg_status graph_reserve_slots(dyg_graph g, uint64_t need) { return grow_arena(g, need); }

dyg_status graph_refresh_mirror(dyg_graph g) {
    if (!g->mirror_stale) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    std::vector<uint2> rows(g->V);
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, hipMemcpy(rows.data(), g->d_rows, (size_t)g->V * sizeof(uint2),
                           hipMemcpyDeviceToHost));
    if (g->dev.cap)
        DYG_HIP(ctx, hipMemcpy(g->cap.data(), g->dev.cap, (size_t)g->V * 4, hipMemcpyDeviceToHost));
    if (g->adj.size() < g->used_slots) g->adj.resize(g->used_slots, make_uint4(0, 0, 0, 0));
    DYG_HIP(ctx, hipMemcpy(g->adj.data(), g->d_adj, g->used_slots * sizeof(uint4),
                           hipMemcpyDeviceToHost));
    for (uint32_t v = 0; v < g->V; ++v) {
        g->start[v] = rows[v].x;
        g->deg[v] = rows[v].y;
    }
    g->mirror_stale = false;
    return DYG_OK;
}

dyg_status graph_apply_inserts(dyg_graph g, const uint32_t* u, const uint32_t* v, const float* w,
                               uint32_t n) {
    if (n == 0) return DYG_OK;
    if (g->keyed && g->mirror_stale) {
        set_error(g->ctx, "insert: keyed graph edited on the device");
        return DYG_ERR_STATE;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (u[i] >= g->V || v[i] >= g->V) {
            set_error(g->ctx, "insert: vertex id out of range");
            return DYG_ERR_ARG;
        }
    {
        dyg_status rs = graph_refresh_mirror(g);
        if (rs != DYG_OK) return rs;
    }
    // appends in reference order: edge i gives (u->v) then (v->u)
    struct App {
        uint32_t row, seq, nbr;
        float w;
    };
    std::vector<App> apps(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        apps[2 * i] = {u[i], 2 * i, v[i], w[i]};
        apps[2 * i + 1] = {v[i], 2 * i + 1, u[i], w[i]};
    }
    std::stable_sort(apps.begin(), apps.end(),
                     [](const App& a, const App& b) { return a.row < b.row; });
    // arena space needed by relocations
    uint64_t extra = 0;
    for (size_t p = 0; p < apps.size();) {
        size_t q = p;
        while (q < apps.size() && apps[q].row == apps[p].row) ++q;
        const uint32_t r = apps[p].row;
        const uint64_t need = (uint64_t)g->deg[r] + (q - p);
        if (need > g->cap[r])
            extra += std::max<uint64_t>({2ull * g->cap[r], need + g->slack, 4ull});
        p = q;
    }
    dyg_status st = grow_arena(g, g->used_slots + extra);
    if (st != DYG_OK) return st;
    if (g->adj.size() < g->used_slots + extra) {
        // grow the mirror geometrically, not to the (possibly reserved) arena size
        const uint64_t want = std::max<uint64_t>(g->used_slots + extra, g->adj.size() * 3 / 2);
        g->adj.resize(std::min<uint64_t>(want, g->cap_slots), make_uint4(0, 0, 0, 0));
    }

    g->dev.cap_valid = false;  // relocations below change row capacities on the host
    std::vector<uint32_t> dirty, touched;
    std::vector<uint32_t> slot_at(2 * (size_t)n);
    dirty.reserve(4 * (size_t)n);
    for (size_t p = 0; p < apps.size();) {
        size_t q = p;
        while (q < apps.size() && apps[q].row == apps[p].row) ++q;
        const uint32_t r = apps[p].row;
        const uint32_t d0 = g->deg[r];
        const uint64_t need = (uint64_t)d0 + (q - p);
        if (need > g->cap[r]) {
            const uint64_t ncap = std::max<uint64_t>({2ull * g->cap[r], need + g->slack, 4ull});
            const uint32_t ns = (uint32_t)g->used_slots;
            g->used_slots += ncap;
            std::copy(g->adj.begin() + g->start[r], g->adj.begin() + g->start[r] + d0,
                      g->adj.begin() + ns);
            for (uint32_t k = 0; k < d0; ++k) dirty.push_back(ns + k);
            g->start[r] = ns;
            g->cap[r] = (uint32_t)ncap;
        }
        for (size_t k = p; k < q; ++k) {
            const uint32_t slot = g->start[r] + d0 + (uint32_t)(k - p);
            // the neighbour's descriptor is filled by the refresh below
            g->adj[slot] = make_uint4(apps[k].nbr, fbits(apps[k].w), 0, 0);
            dirty.push_back(slot);
            slot_at[apps[k].seq] = (uint32_t)(d0 + (k - p));
        }
        g->deg[r] = (uint32_t)need;
        touched.push_back(r);
        p = q;
    }
    refresh_neighbour_copies(g, touched, dirty);
    g->nslots += 2ull * n;
    if (g->keyed) {
        // first occurrence wins, as std::unordered_map::insert in the reference
        for (uint32_t i = 0; i < n; ++i) {
            const long a = u[i], b = v[i];
            const long key = key_of(a, b, g->mult);
            const uint32_t sa = slot_at[2 * i], sb = slot_at[2 * i + 1];
            if (a <= b) g->key_map.insert({key, {sa, sb}});
            else g->key_map.insert({key, {sb, sa}});
        }
    }
    return ship_delta(g, dirty, touched);
}

dyg_status graph_apply_deletes(dyg_graph g, const uint32_t* u, const uint32_t* v, uint32_t n,
                               int8_t* hit, bool allow_zero_degree) {
    if (!g->keyed) {
        set_error(g->ctx, "delete: graph was uploaded without its edge map");
        return DYG_ERR_ARG;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (u[i] >= g->V || v[i] >= g->V) {
            set_error(g->ctx, "delete: vertex id out of range");
            return DYG_ERR_ARG;
        }
    {
        dyg_status rs = graph_refresh_mirror(g);
        if (rs != DYG_OK) return rs;
    }
    // validation pass: no vertex may reach degree 0 (the reference asserts,
    // random_walk_decremental/gpu_graph.cuh:362-363); nothing changes on failure
    if (!allow_zero_degree) {
        std::unordered_map<long, char> gone;
        std::unordered_map<uint32_t, uint32_t> dec;
        for (uint32_t i = 0; i < n; ++i) {
            const long key = key_of(u[i], v[i], g->mult);
            if (!g->key_map.count(key) || gone.count(key)) continue;
            gone[key] = 1;
            const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
            if (++dec[a] >= g->deg[a] || ++dec[b] >= g->deg[b]) {
                set_error(g->ctx, "delete: a vertex would be left with degree 0");
                return DYG_ERR_STATE;
            }
        }
    }
    std::vector<uint32_t> dirty, touched;
    for (uint32_t i = 0; i < n; ++i) {
        long a = u[i], b = v[i];
        if (a > b) std::swap(a, b);
        const long key = a * g->mult + b;
        auto it = g->key_map.find(key);
        if (it == g->key_map.end()) {
            if (hit) hit[i] = 0;
            continue;
        }
        if (hit) hit[i] = 1;
        const uint32_t ib_a = it->second.first, ia_b = it->second.second;
        g->key_map.erase(it);
        const uint32_t da = --g->deg[a];
        const uint32_t db = --g->deg[b];
        if (da != ib_a) {
            const uint4 last = g->adj[g->start[a] + da];
            g->adj[g->start[a] + ib_a] = last;
            const long al = last.x;
            if (a < al) g->key_map.at(a * g->mult + al).first = ib_a;
            else g->key_map.at(al * g->mult + a).second = ib_a;
            dirty.push_back(g->start[a] + ib_a);
        }
        if (db != ia_b) {
            const uint4 last = g->adj[g->start[b] + db];
            g->adj[g->start[b] + ia_b] = last;
            const long bl = last.x;
            if (b < bl) g->key_map.at(b * g->mult + bl).first = ia_b;
            else g->key_map.at(bl * g->mult + b).second = ia_b;
            dirty.push_back(g->start[b] + ia_b);
        }
        touched.push_back((uint32_t)a);
        touched.push_back((uint32_t)b);
        g->nslots -= 2;
    }
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    refresh_neighbour_copies(g, touched, dirty);
    return ship_delta(g, dirty, touched);
}

dyg_status graph_add_weights(dyg_graph g, const uint32_t* u, const uint32_t* v, const float* dw,
                             uint32_t n) {
    if (!g->keyed) {
        set_error(g->ctx, "weight update: graph was uploaded without its edge map");
        return DYG_ERR_ARG;
    }
    {
        dyg_status rs = graph_refresh_mirror(g);
        if (rs != DYG_OK) return rs;
    }
    std::vector<uint32_t> dirty, rows;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        auto it = g->key_map.find(key_of(a, b, g->mult));
        if (it == g->key_map.end()) {
            set_error(g->ctx, "weight update: edge not in the graph");
            return DYG_ERR_STATE;
        }
        uint4& sa = g->adj[g->start[a] + it->second.first];
        uint4& sb = g->adj[g->start[b] + it->second.second];
        float x;
        std::memcpy(&x, &sa.y, 4);
        x = x + dw[i];
        sa.y = fbits(x);
        sb.y = fbits(x);
        dirty.push_back(g->start[a] + it->second.first);
        dirty.push_back(g->start[b] + it->second.second);
    }
    return ship_delta(g, dirty, rows);
}

}  // namespace dyg

using namespace dyg;

extern "C" dyg_status dyg_graph_upload(dyg_ctx ctx, uint32_t V, const uint64_t* row_ptr,
                                       const uint32_t* col, const float* w, uint32_t slack,
                                       int keyed, dyg_graph* out) {
    if (!ctx || !out || !row_ptr || V == 0) {
        set_error(ctx, "dyg_graph_upload: invalid argument");
        return DYG_ERR_ARG;
    }
    *out = nullptr;
    if (row_ptr[0] != 0) {
        set_error(ctx, "dyg_graph_upload: row_ptr[0] must be 0");
        return DYG_ERR_ARG;
    }
    for (uint32_t v = 0; v < V; ++v)
        if (row_ptr[v + 1] < row_ptr[v]) {
            set_error(ctx, "dyg_graph_upload: row_ptr not monotone");
            return DYG_ERR_ARG;
        }
    const uint64_t E = row_ptr[V];
    if (E && (!col || !w)) {
        set_error(ctx, "dyg_graph_upload: invalid argument");
        return DYG_ERR_ARG;
    }
    for (uint64_t p = 0; p < E; ++p)
        if (col[p] >= V) {
            set_error(ctx, "dyg_graph_upload: neighbour id out of range");
            return DYG_ERR_ARG;
        }
    auto* g = new dyg_graph_s;
    g->ctx = ctx;
    g->V = V;
    g->slack = slack;
    g->start.resize(V);
    g->deg.resize(V);
    g->cap.resize(V);
    uint64_t total = 0;
    for (uint32_t v = 0; v < V; ++v) {
        g->deg[v] = (uint32_t)(row_ptr[v + 1] - row_ptr[v]);
        g->cap[v] = g->deg[v] + slack;
        g->start[v] = (uint32_t)total;
        total += g->cap[v];
    }
    // room for relocations without an early regrow (HBM is plentiful)
    const uint64_t cap_slots = total + std::max<uint64_t>(4096, total / 2);
    if (cap_slots > 0xFFFFFFFFull) {
        delete g;
        set_error(ctx, "dyg_graph_upload: more than 2^32 slots");
        return DYG_ERR_NOMEM;
    }
    g->used_slots = total;
    g->cap_slots = cap_slots;
    g->nslots = E;
    g->adj.assign(total + total / 4 + 64, make_uint4(0, 0, 0, 0));
    for (uint32_t v = 0; v < V; ++v)
        for (uint32_t k = 0; k < g->deg[v]; ++k) {
            const uint32_t c = col[row_ptr[v] + k];
            g->adj[g->start[v] + k] = make_uint4(c, fbits(w[row_ptr[v] + k]), g->start[c], g->deg[c]);
        }
    std::vector<uint2> rows(V);
    for (uint32_t v = 0; v < V; ++v) rows[v] = make_uint2(g->start[v], g->deg[v]);
    if (hipMalloc(&g->d_rows, (size_t)V * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&g->d_adj, cap_slots * sizeof(uint4)) != hipSuccess) {
        if (g->d_rows) (void)hipFree(g->d_rows);
        if (g->d_adj) (void)hipFree(g->d_adj);
        delete g;
        set_error(ctx, "dyg_graph_upload: device allocation failed");
        return DYG_ERR_NOMEM;
    }
    hipError_t e1 = hipMemcpy(g->d_rows, rows.data(), (size_t)V * sizeof(uint2),
                              hipMemcpyHostToDevice);
    hipError_t e2 = hipMemcpy(g->d_adj, g->adj.data(), total * sizeof(uint4),
                              hipMemcpyHostToDevice);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        set_error(ctx, std::string("dyg_graph_upload: ") +
                           hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        dyg_graph_free(g);
        return DYG_ERR_HIP;
    }
    if (keyed) {
        g->keyed = true;
        g->mult = key_multiplier(V);
        g->key_map.reserve((size_t)(E / 2 * 1.5) + 16);
        // (slot of b in a's row, slot of a in b's row) for key a*mult+b, a<b;
        // the first occurrence in each row wins, as the first file line does
        for (uint32_t a = 0; a < V; ++a)
            for (uint32_t k = 0; k < g->deg[a]; ++k) {
                const uint32_t b = g->adj[g->start[a] + k].x;
                const long key = key_of(a, b, g->mult);
                auto it = g->key_map.find(key);
                if (a <= b) {
                    if (it == g->key_map.end()) g->key_map.insert({key, {k, 0xFFFFFFFFu}});
                    else if (it->second.first == 0xFFFFFFFFu) it->second.first = k;
                }
                if (a >= b) {
                    if (it == g->key_map.end()) g->key_map.insert({key, {0xFFFFFFFFu, k}});
                    else if (it->second.second == 0xFFFFFFFFu) it->second.second = k;
                }
            }
    }
    *out = g;
    return DYG_OK;
}

extern "C" dyg_status dyg_graph_info(dyg_graph g, uint32_t* V, uint64_t* nslots,
                                     uint64_t* capacity) {
    if (!g) return DYG_ERR_ARG;
    if (V) *V = g->V;
    if (nslots) *nslots = g->nslots;
    if (capacity) *capacity = g->cap_slots;
    return DYG_OK;
}

extern "C" dyg_status dyg_graph_download(dyg_graph g, uint64_t* row_ptr, uint32_t* col,
                                         float* w) {
    if (!g || !row_ptr) return DYG_ERR_ARG;
    dyg_ctx ctx = g->ctx;
    std::vector<uint2> rows(g->V);
    std::vector<uint4> adj(g->used_slots);
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, hipMemcpy(rows.data(), g->d_rows, (size_t)g->V * sizeof(uint2),
                           hipMemcpyDeviceToHost));
    DYG_HIP(ctx, hipMemcpy(adj.data(), g->d_adj, g->used_slots * sizeof(uint4),
                           hipMemcpyDeviceToHost));
    uint64_t p = 0;
    for (uint32_t v = 0; v < g->V; ++v) {
        row_ptr[v] = p;
        for (uint32_t k = 0; k < rows[v].y; ++k) {
            const uint4 sl = adj[rows[v].x + k];
            if (col) col[p] = sl.x;
            if (w) std::memcpy(&w[p], &sl.y, 4);
            ++p;
        }
    }
    row_ptr[g->V] = p;
    return DYG_OK;
}

extern "C" dyg_status dyg_graph_check(dyg_graph g, uint64_t* bad_slots) {
    // Every live slot's copy of its neighbour's row descriptor equals the
    // neighbour's row (device state); a graph that keeps no descriptors only
    // needs valid neighbour ids.
    if (!g) return DYG_ERR_ARG;
    dyg_ctx ctx = g->ctx;
    std::vector<uint2> rows(g->V);
    std::vector<uint4> adj(g->used_slots);
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, hipMemcpy(rows.data(), g->d_rows, (size_t)g->V * sizeof(uint2),
                           hipMemcpyDeviceToHost));
    DYG_HIP(ctx, hipMemcpy(adj.data(), g->d_adj, g->used_slots * sizeof(uint4),
                           hipMemcpyDeviceToHost));
    uint64_t bad = 0;
    for (uint32_t v = 0; v < g->V; ++v)
        for (uint32_t k = 0; k < rows[v].y; ++k) {
            const uint4 sl = adj[rows[v].x + k];
            if (sl.x >= g->V || (!g->nodesc && (sl.z != rows[sl.x].x || sl.w != rows[sl.x].y)))
                ++bad;
        }
    if (bad_slots) *bad_slots = bad;
    return DYG_OK;
}

extern "C" dyg_status dyg_graph_insert(dyg_graph g, const uint32_t* u, const uint32_t* v,
                                       const float* w, uint32_t n) {
    if (!g || (n && (!u || !v || !w))) return DYG_ERR_ARG;
    dyg_status st = graph_apply_inserts(g, u, v, w, n);
    if (st != DYG_OK) return st;
    DYG_HIP(g->ctx, hipStreamSynchronize(g->ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_graph_delete(dyg_graph g, const uint32_t* u, const uint32_t* v,
                                       uint32_t n, int8_t* hit) {
    if (!g || (n && (!u || !v))) return DYG_ERR_ARG;
    dyg_status st = graph_apply_deletes(g, u, v, n, hit);
    if (st != DYG_OK) return st;
    DYG_HIP(g->ctx, hipStreamSynchronize(g->ctx->stream));
    return DYG_OK;
}

extern "C" void dyg_graph_free(dyg_graph g) {
    if (!g) return;
    if (g->d_rows) (void)hipFree(g->d_rows);
    if (g->d_adj) (void)hipFree(g->d_adj);
    delete g;
}
