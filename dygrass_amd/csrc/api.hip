// api.hip -- context, walk entry points and the decremental dual graph
// (C ABI declared in include/dygrass.h).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <unordered_set>

#include "internal.h"

namespace dyg {

namespace {
thread_local std::string g_tls_error;   // last failure of this thread
thread_local std::string g_tls_report;  // what dyg_last_error returned to this thread
}

void set_error(dyg_ctx ctx, const std::string& msg) {
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->last_error = msg;
    }
    g_tls_error = msg;
}

dyg_status check_params(dyg_ctx ctx, const dyg_walk_params* p) {
    if (!p || !valid_walkers(p->walkers)) {
        set_error(ctx, "walk params: walkers must be a power of two in [1,1024]");
        return DYG_ERR_ARG;
    }
    if (p->max_steps < 1 || p->max_steps > 32767) {
        set_error(ctx, "walk params: max_steps must be in [1,32767]");
        return DYG_ERR_ARG;
    }
    return DYG_OK;
}

namespace {

// ---------------------------------------------------------------------------
// Fixed device staging of dyg_walk_incremental_noalloc: one call's inputs,
// results and counters, in an array the library holds on every GPU it runs
// on until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: every host use goes through the per-device
// address taken once by scratch_setup (hipGetSymbolAddress, with the device
// current).  All calls on one GPU share its array; calls that overlap on a
// GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 40ull << 20;
constexpr size_t kScratchStats = 256;  // counters at the array's end
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Takes the device address of `scratch` on `device`, which must be current;
// once per device.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

void* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
    if (threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Device staging area for one host-pointer walk call, carved from ctx->dscratch.
struct Carve {
    char* base;
    size_t off = 0;
    explicit Carve(void* b) : base((char*)b) {}
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base + off);
        off += n * sizeof(T);
        return p;
    }
    static size_t need(std::initializer_list<size_t> sizes) {
        size_t t = 0;
        for (size_t s : sizes) t += ((s + 255) & ~(size_t)255);
        return t + 256;
    }
};

// Host-pointer walk: stage inputs, launch, copy results back, wait.  The
// staging comes from the context's growable buffers, or, with `fixed`, from
// the fixed array (no allocation; at most DYG_NOALLOC_MAX_EDGES edges).
dyg_status walk_host(dyg_graph g, bool insert, const uint32_t* src, const uint32_t* dst,
                     const float* w, const uint64_t* edge_id, uint32_t n,
                     const dyg_walk_params* p, int32_t* found_at, float* cond, uint32_t* winner,
                     uint32_t* hops, uint32_t* path, dyg_walk_stats* st, bool fixed = false) {
    dyg_ctx ctx = g->ctx;
    dyg_status cs = check_params(ctx, p);
    if (cs != DYG_OK) return cs;
    if (n && (!src || !dst || !found_at || !cond || (insert && !w) || (!insert && !path))) {
        set_error(ctx, "walk: missing array");
        return DYG_ERR_ARG;
    }
    for (uint32_t i = 0; i < n; ++i)
        if (src[i] >= g->V || dst[i] >= g->V) {
            set_error(ctx, "walk: vertex id out of range");
            return DYG_ERR_ARG;
        }
    if (st) std::memset(st, 0, sizeof(*st));
    if (n == 0) return DYG_OK;
    const size_t pathn = insert ? 0 : (size_t)n * p->max_steps;
    const size_t bytes = Carve::need({n * 4ull, n * 4ull, n * 4ull, n * 8ull, n * 4ull, n * 4ull,
                                      n * 4ull, n * 4ull, pathn * 4, 64});
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    void* stage = nullptr;
    unsigned long long* d_stats = nullptr;
    if (fixed) {
        char* base = (char*)scratch_of(ctx->device);
        if (!base) {
            set_error(ctx, "walk (no allocation): the fixed staging array has no device address");
            return DYG_ERR_STATE;
        }
        if (bytes > kScratchBytes - kScratchStats) {
            set_error(ctx, "walk (no allocation): batch exceeds the fixed staging array");
            return DYG_ERR_ARG;
        }
        stage = base;
        d_stats = (unsigned long long*)(base + kScratchBytes - kScratchStats);
    } else {
        DYG_HIP(ctx, ctx->dscratch.reserve(bytes));
        DYG_HIP(ctx, ctx->dstats.reserve(64));
        stage = ctx->dscratch.p;
        d_stats = (unsigned long long*)ctx->dstats.p;
    }
    Carve c(stage);
    uint32_t* d_src = c.take<uint32_t>(n);
    uint32_t* d_dst = c.take<uint32_t>(n);
    float* d_w = c.take<float>(n);
    uint64_t* d_id = c.take<uint64_t>(n);
    int32_t* d_found = c.take<int32_t>(n);
    float* d_cond = c.take<float>(n);
    uint32_t* d_win = c.take<uint32_t>(n);
    uint32_t* d_hops = c.take<uint32_t>(n);
    uint32_t* d_path = c.take<uint32_t>(pathn ? pathn : 1);
    hipStream_t s = ctx->stream;
    DYG_HIP(ctx, hipMemcpyAsync(d_src, src, n * 4ull, hipMemcpyHostToDevice, s));
    DYG_HIP(ctx, hipMemcpyAsync(d_dst, dst, n * 4ull, hipMemcpyHostToDevice, s));
    if (insert) DYG_HIP(ctx, hipMemcpyAsync(d_w, w, n * 4ull, hipMemcpyHostToDevice, s));
    if (edge_id) DYG_HIP(ctx, hipMemcpyAsync(d_id, edge_id, n * 8ull, hipMemcpyHostToDevice, s));
    if (fixed) {  // a kernel, not a memset: the array keeps the last call's contents
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, d_stats, 3u);
        DYG_HIP(ctx, hipGetLastError());
    } else {
        DYG_HIP(ctx, hipMemsetAsync(d_stats, 0, 3 * sizeof(unsigned long long), s));
    }
    WalkLaunch a{};
    a.rows = g->d_rows;
    a.nodesc = g->nodesc;
    a.adj = g->d_adj;
    a.src = d_src;
    a.dst = d_dst;
    a.w = insert ? d_w : nullptr;
    a.eid = edge_id ? d_id : nullptr;
    a.eid_base = 0;
    a.n = n;
    a.walkers = p->walkers;
    a.max_steps = p->max_steps;
    a.thr = p->threshold;
    a.flags = p->flags;
    a.edges_per_block = p->block_edges;
    a.seed = p->seed;
    a.found = d_found;
    a.cond = d_cond;
    a.winner = d_win;
    a.hops = d_hops;
    a.path = insert ? nullptr : d_path;
    a.stats = d_stats;
    DYG_HIP(ctx, insert ? launch_walk_incremental(a, s, ctx->ev0, ctx->ev1)
                        : launch_walk_decremental(a, s, ctx->ev0, ctx->ev1));
    DYG_HIP(ctx, hipMemcpyAsync(found_at, d_found, n * 4ull, hipMemcpyDeviceToHost, s));
    DYG_HIP(ctx, hipMemcpyAsync(cond, d_cond, n * 4ull, hipMemcpyDeviceToHost, s));
    if (winner) DYG_HIP(ctx, hipMemcpyAsync(winner, d_win, n * 4ull, hipMemcpyDeviceToHost, s));
    if (hops) DYG_HIP(ctx, hipMemcpyAsync(hops, d_hops, n * 4ull, hipMemcpyDeviceToHost, s));
    if (!insert) DYG_HIP(ctx, hipMemcpyAsync(path, d_path, pathn * 4, hipMemcpyDeviceToHost, s));
    unsigned long long hst[3] = {0, 0, 0};
    DYG_HIP(ctx, hipMemcpyAsync(hst, d_stats, sizeof(hst), hipMemcpyDeviceToHost, s));
    DYG_HIP(ctx, hipStreamSynchronize(s));
    if (st) {
        float ms = 0.f;
        DYG_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        st->hops = hst[0];
        st->redirects = hst[1];
        st->leaf_stops = hst[2];
        st->kernel_ms = ms;
    }
    return DYG_OK;
}

}  // namespace
}  // namespace dyg

using namespace dyg;

/* ------------------------------------------------------------------ */
extern "C" const char* dyg_version(void) { return "dygrass-mi355x 0.1 (gfx950)"; }

extern "C" dyg_status dyg_ctx_create(int device, dyg_ctx* out) {
    if (!out) return DYG_ERR_ARG;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error(nullptr, "dyg_ctx_create: no HIP device");
        return DYG_ERR_HIP;
    }
    if (device < 0 || device >= n) {
        set_error(nullptr, "dyg_ctx_create: device index out of range");
        return DYG_ERR_ARG;
    }
    auto* c = new dyg_ctx_s;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        set_error(nullptr, "dyg_ctx_create: stream/event creation failed");
        delete c;
        return DYG_ERR_HIP;
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess)
        c->clock_khz = khz;
    // the fixed staging array's address on this device (device current);
    // a failure here is retried by the first call that needs it
    (void)scratch_setup(device);
    *out = c;
    return DYG_OK;
}

extern "C" void dyg_ctx_destroy(dyg_ctx ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char* dyg_last_error(dyg_ctx ctx) {
    // a copy owned by the calling thread: a server's worker may set the
    // context's message at any time
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        g_tls_report = ctx->last_error;
    } else {
        g_tls_report = g_tls_error;
    }
    return g_tls_report.c_str();
}

extern "C" void* dyg_ctx_stream(dyg_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

extern "C" dyg_status dyg_ctx_sync(dyg_ctx ctx) {
    if (!ctx) return DYG_ERR_ARG;
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_dev_alloc(dyg_ctx ctx, size_t bytes, void** out) {
    if (!ctx || !out) return DYG_ERR_ARG;
    *out = nullptr;
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        set_error(ctx, "dyg_dev_alloc: device allocation failed");
        return DYG_ERR_NOMEM;
    }
    return DYG_OK;
}

extern "C" dyg_status dyg_dev_free(dyg_ctx ctx, void* p) {
    if (!ctx) return DYG_ERR_ARG;
    if (p) {
        DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        DYG_HIP(ctx, hipFree(p));
    }
    return DYG_OK;
}

extern "C" dyg_status dyg_copy_to_device(dyg_ctx ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return DYG_ERR_ARG;
    DYG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_copy_on_device(dyg_ctx ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return DYG_ERR_ARG;
    DYG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_copy_to_host(dyg_ctx ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return DYG_ERR_ARG;
    DYG_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return DYG_OK;
}

/* ------------------------------------------------------------------ */
extern "C" dyg_status dyg_walk_incremental(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                           const float* w, const uint64_t* edge_id, uint32_t n,
                                           const dyg_walk_params* p, int32_t* found_at,
                                           float* cond, uint32_t* winner, uint32_t* hops,
                                           dyg_walk_stats* st) {
    if (!H) return DYG_ERR_ARG;
    return walk_host(H, true, src, dst, w, edge_id, n, p, found_at, cond, winner, hops, nullptr,
                     st);
}

extern "C" dyg_status dyg_walk_incremental_noalloc(dyg_graph H, const uint32_t* src,
                                                   const uint32_t* dst, const float* w,
                                                   const uint64_t* edge_id, uint32_t n,
                                                   const dyg_walk_params* p, int32_t* found_at,
                                                   float* cond, uint32_t* winner, uint32_t* hops,
                                                   dyg_walk_stats* st) {
    if (!H) return DYG_ERR_ARG;
    if (n > DYG_NOALLOC_MAX_EDGES) {
        set_error(H->ctx, "walk (no allocation): more than DYG_NOALLOC_MAX_EDGES edges");
        return DYG_ERR_ARG;
    }
    const int dev = H->ctx->device;
    if (!scratch_of(dev)) {  // set-up failed at context creation: once more, device current
        int cur = 0;
        DYG_HIP(H->ctx, hipGetDevice(&cur));
        DYG_HIP(H->ctx, hipSetDevice(dev));
        const hipError_t e = scratch_setup(dev);
        DYG_HIP(H->ctx, hipSetDevice(cur));
        DYG_HIP(H->ctx, e);
    }
    return walk_host(H, true, src, dst, w, edge_id, n, p, found_at, cond, winner, hops, nullptr,
                     st, true);
}

extern "C" dyg_status dyg_walk_decremental(dyg_graph G, const uint32_t* src, const uint32_t* dst,
                                           const uint64_t* edge_id, uint32_t n,
                                           const dyg_walk_params* p, int32_t* found_at,
                                           float* cond, uint32_t* winner, uint32_t* hops,
                                           uint32_t* path, dyg_walk_stats* st) {
    if (!G) return DYG_ERR_ARG;
    return walk_host(G, false, src, dst, nullptr, edge_id, n, p, found_at, cond, winner, hops,
                     path, st);
}

extern "C" dyg_status dyg_walk_incremental_dev(dyg_graph H, const uint32_t* src,
                                               const uint32_t* dst, const float* w,
                                               const uint64_t* edge_id, uint32_t n,
                                               const dyg_walk_params* p, int32_t* found_at,
                                               float* cond, uint32_t* winner, uint32_t* hops,
                                               unsigned long long* hops_total) {
    if (!H) return DYG_ERR_ARG;
    dyg_status cs = check_params(H->ctx, p);
    if (cs != DYG_OK) return cs;
    WalkLaunch a{};
    a.rows = H->d_rows;
    a.nodesc = H->nodesc;
    a.adj = H->d_adj;
    a.src = src;
    a.dst = dst;
    a.w = w;
    a.eid = edge_id;
    a.n = n;
    a.walkers = p->walkers;
    a.max_steps = p->max_steps;
    a.thr = p->threshold;
    a.flags = p->flags;
    a.edges_per_block = p->block_edges;
    a.seed = p->seed;
    a.found = found_at;
    a.cond = cond;
    a.winner = winner;
    a.hops = hops;
    a.stats = hops_total;
    DYG_HIP(H->ctx, launch_walk_incremental(a, H->ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_walk_decremental_dev(dyg_graph G, const uint32_t* src,
                                               const uint32_t* dst, const uint64_t* edge_id,
                                               uint32_t n, const dyg_walk_params* p,
                                               int32_t* found_at, float* cond, uint32_t* winner,
                                               uint32_t* hops, uint32_t* path,
                                               unsigned long long* hops_total) {
    if (!G) return DYG_ERR_ARG;
    dyg_status cs = check_params(G->ctx, p);
    if (cs != DYG_OK) return cs;
    WalkLaunch a{};
    a.rows = G->d_rows;
    a.nodesc = G->nodesc;
    a.adj = G->d_adj;
    a.src = src;
    a.dst = dst;
    a.eid = edge_id;
    a.n = n;
    a.walkers = p->walkers;
    a.max_steps = p->max_steps;
    a.flags = p->flags;
    a.edges_per_block = p->block_edges;
    a.seed = p->seed;
    a.found = found_at;
    a.cond = cond;
    a.winner = winner;
    a.hops = hops;
    a.path = path;
    a.stats = hops_total;
    DYG_HIP(G->ctx, launch_walk_decremental(a, G->ctx->stream));
    return DYG_OK;
}

extern "C" dyg_status dyg_walk_heuristic(dyg_graph G, const uint32_t* endpoint,
                                         const uint64_t* key, uint32_t n, uint32_t walkers,
                                         uint64_t seed, uint32_t* path3) {
    if (!G || (n && (!endpoint || !key || !path3)) || walkers == 0) return DYG_ERR_ARG;
    dyg_ctx ctx = G->ctx;
    for (uint32_t i = 0; i < n; ++i)
        if (endpoint[i] >= G->V) {
            set_error(ctx, "heuristic: vertex id out of range");
            return DYG_ERR_ARG;
        }
    if (n == 0) return DYG_OK;
    const size_t out = (size_t)n * walkers * 3;
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, ctx->dscratch.reserve(Carve::need({n * 4ull, n * 8ull, out * 4})));
    Carve c(ctx->dscratch.p);
    uint32_t* d_ep = c.take<uint32_t>(n);
    uint64_t* d_key = c.take<uint64_t>(n);
    uint32_t* d_out = c.take<uint32_t>(out);
    hipStream_t s = ctx->stream;
    DYG_HIP(ctx, hipMemcpyAsync(d_ep, endpoint, n * 4ull, hipMemcpyHostToDevice, s));
    DYG_HIP(ctx, hipMemcpyAsync(d_key, key, n * 8ull, hipMemcpyHostToDevice, s));
    DYG_HIP(ctx, launch_heuristic(G->d_rows, G->d_adj, d_ep, d_key, n, walkers, seed, d_out, s,
                                  nullptr, G->nodesc));
    DYG_HIP(ctx, hipMemcpyAsync(path3, d_out, out * 4, hipMemcpyDeviceToHost, s));
    DYG_HIP(ctx, hipStreamSynchronize(s));
    return DYG_OK;
}

/* ------------------------------------------------------------------ */
/* Dual graph (random_walk_decremental/gpu_graph.cuh:79-687)          */
/* ------------------------------------------------------------------ */
struct dyg_dual_s {
    dyg_ctx ctx = nullptr;
    dyg_graph G = nullptr;
    long mult = 1;
    std::unordered_set<long> sparse;  // H as a key set
    std::unordered_set<long> added;   // edges added by path recovery
    // deletions queued for walking (in order)
    std::vector<uint32_t> qa, qb;
    std::vector<uint64_t> qid;
    // no-path endpoints queued for the heuristic
    std::vector<uint32_t> heur;
    std::vector<uint64_t> heur_key;
};

extern "C" dyg_status dyg_dual_create(dyg_ctx ctx, uint32_t V, const uint64_t* g_row_ptr,
                                      const uint32_t* g_col, const float* g_w,
                                      const uint32_t* h_u, const uint32_t* h_v, uint64_t h_m,
                                      dyg_dual* out) {
    if (!ctx || !out || (h_m && (!h_u || !h_v))) return DYG_ERR_ARG;
    *out = nullptr;
    for (uint64_t i = 0; i < h_m; ++i)
        if (h_u[i] >= V || h_v[i] >= V) {
            set_error(ctx, "dyg_dual_create: sparsifier vertex id out of range");
            return DYG_ERR_ARG;
        }
    dyg_graph G = nullptr;
    dyg_status st = dyg_graph_upload(ctx, V, g_row_ptr, g_col, g_w, 0, 1, &G);
    if (st != DYG_OK) return st;
    auto* d = new dyg_dual_s;
    d->ctx = ctx;
    d->G = G;
    d->mult = G->mult;
    d->sparse.reserve((size_t)(h_m * 1.5) + 16);
    for (uint64_t i = 0; i < h_m; ++i) d->sparse.insert(key_of(h_u[i], h_v[i], d->mult));
    *out = d;
    return DYG_OK;
}

extern "C" int64_t dyg_dual_key_multiplier(dyg_dual d) { return d ? d->mult : 0; }

extern "C" dyg_status dyg_dual_remove(dyg_dual d, const uint32_t* u, const uint32_t* v,
                                      const uint64_t* ids, uint32_t n, int8_t* status,
                                      uint32_t* n_queued) {
    if (!d || (n && (!u || !v))) return DYG_ERR_ARG;
    std::vector<int8_t> hit(n);
    dyg_status st = graph_apply_deletes(d->G, u, v, n, hit.data());
    if (st != DYG_OK) return st;
    d->qa.clear();
    d->qb.clear();
    d->qid.clear();
    for (uint32_t i = 0; i < n; ++i) {
        int8_t s = 0;
        if (hit[i]) {
            const long key = key_of(u[i], v[i], d->mult);
            auto it = d->sparse.find(key);
            if (it != d->sparse.end()) {
                d->sparse.erase(it);
                // walks start at the smaller endpoint (gpu_graph.cuh:342, 453-454)
                d->qa.push_back(std::min(u[i], v[i]));
                d->qb.push_back(std::max(u[i], v[i]));
                d->qid.push_back(ids ? ids[i] : i);
                s = 2;
            } else {
                s = 1;
            }
        }
        if (status) status[i] = s;
    }
    if (n_queued) *n_queued = (uint32_t)d->qa.size();
    return DYG_OK;
}

extern "C" dyg_status dyg_dual_walk(dyg_dual d, const dyg_walk_params* p, uint32_t* no_path,
                                    uint64_t* added, dyg_walk_stats* st) {
    if (!d) return DYG_ERR_ARG;
    const uint32_t n = (uint32_t)d->qa.size();
    std::vector<int32_t> f(n);
    std::vector<float> c(n);
    std::vector<uint32_t> path((size_t)n * (p ? p->max_steps : 1));
    dyg_status s = dyg_walk_decremental(d->G, d->qa.data(), d->qb.data(), d->qid.data(), n, p,
                                        f.data(), c.data(), nullptr, nullptr, path.data(), st);
    if (s != DYG_OK) return s;
    d->heur.clear();
    d->heur_key.clear();
    uint64_t add_count = 0;
    const uint32_t ms = p->max_steps;
    for (uint32_t i = 0; i < n; ++i) {
        if (f[i] != -1) {
            long from = path[(size_t)i * ms];
            for (int j = 0; j < f[i]; ++j) {
                const long to = path[(size_t)i * ms + j + 1];
                const long key = key_of(from, to, d->mult);
                if (d->sparse.insert(key).second) {
                    add_count++;
                    d->added.insert(key);
                }
                from = to;
            }
        } else {
            d->heur.push_back(d->qa[i]);
            d->heur.push_back(d->qb[i]);
            d->heur_key.push_back(d->qid[i] << 1);
            d->heur_key.push_back((d->qid[i] << 1) | 1u);
        }
    }
    if (no_path) *no_path = (uint32_t)(d->heur.size() / 2);
    if (added) *added = add_count;
    return DYG_OK;
}

extern "C" dyg_status dyg_dual_heuristic(dyg_dual d, uint32_t walkers, uint64_t seed) {
    if (!d || walkers == 0) return DYG_ERR_ARG;
    const uint32_t n = (uint32_t)d->heur.size();
    if (n == 0) return DYG_OK;
    std::vector<uint32_t> p3((size_t)n * walkers * 3);
    dyg_status s = dyg_walk_heuristic(d->G, d->heur.data(), d->heur_key.data(), n, walkers, seed,
                                      p3.data());
    if (s != DYG_OK) return s;
    for (size_t e = 0; e < n; ++e)
        for (uint32_t j = 0; j < walkers; ++j) {
            const uint32_t* q = &p3[(e * walkers + j) * 3];
            d->sparse.insert(key_of(q[0], q[1], d->mult));
            d->sparse.insert(key_of(q[1], q[2], d->mult));
        }
    d->heur.clear();
    d->heur_key.clear();
    return DYG_OK;
}

extern "C" uint64_t dyg_dual_sparse_size(dyg_dual d) { return d ? d->sparse.size() : 0; }
extern "C" uint64_t dyg_dual_added_count(dyg_dual d) { return d ? d->added.size() : 0; }

extern "C" dyg_status dyg_dual_added_edges(dyg_dual d, uint32_t* u, uint32_t* v) {
    if (!d || !u || !v) return DYG_ERR_ARG;
    std::vector<long> keys(d->added.begin(), d->added.end());
    std::sort(keys.begin(), keys.end());
    for (size_t i = 0; i < keys.size(); ++i) {
        u[i] = (uint32_t)(keys[i] / d->mult);
        v[i] = (uint32_t)(keys[i] % d->mult);
    }
    return DYG_OK;
}

extern "C" dyg_graph dyg_dual_dense(dyg_dual d) { return d ? d->G : nullptr; }

extern "C" void dyg_dual_free(dyg_dual d) {
    if (!d) return;
    dyg_graph_free(d->G);
    delete d;
}
