// incremental.hip -- the incremental CLI's snapshot-batched loop
// (random_walk_incremental/main.cu:297-340 with gpu_graph.cuh:292-413).
//
// m insertions are split into batches of ceil(m/divisor); the walks of a
// batch see H as it stood at batch start; the no-path edges of the batch are
// then appended in batch order.  Edge i's draws are keyed by i.
//
// Two ways to apply a batch's appends, same resulting graph:
//   device (default)  the run's row entries are sorted once, by batch and
//                     row, before the first walk; each batch's appends then
//                     run as two kernels on the context stream right after
//                     its walk kernel, so the whole run is one uninterrupted
//                     stream of launches;
//   host              the decisions are copied back after each batch and the
//                     host plans the appends (the reference's shape).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "internal.h"

namespace dyg {
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Device buffers of one run: inputs (unless the caller passed device
// pointers), decisions, conductances, counters.  Kept on the context.
struct RunBufs {
    uint32_t* src;
    uint32_t* dst;
    float* w;
    int32_t* found;
    float* cond;
    unsigned long long* stats;
};

// The loop itself.  d_* are device pointers; h_* are host copies of the
// inputs, required only for host-planned edits.
dyg_status run_incremental(dyg_graph H, const RunBufs& b, const uint32_t* h_src,
                           const uint32_t* h_dst, const float* h_w, uint32_t m, uint32_t divisor,
                           const dyg_walk_params* p, int32_t* h_found, dyg_run_stats* out) {
    dyg_ctx ctx = H->ctx;
    hipStream_t s = ctx->stream;
    dyg_run_stats rs{};
    const double t0 = now_ms();
    const uint32_t I = m ? (m + divisor - 1) / divisor : 0;
    const uint32_t nb = I ? (m + I - 1) / I : 0;
    DYG_HIP(ctx, hipMemsetAsync(b.stats, 0, 3 * sizeof(unsigned long long), s));
    const bool device_edits = !(p->flags & DYG_RUN_HOST_EDITS) && !H->keyed;
    dyg_status ret = DYG_OK;
    ctx->reserve_events(3 * (size_t)std::max<uint32_t>(nb, 1));
    if (ctx->evpool.size() < 3 * (size_t)nb) {
        set_error(ctx, "incremental: event creation failed");
        return DYG_ERR_HIP;
    }
    auto launch = [&](uint32_t b0, uint32_t n, hipEvent_t e0, hipEvent_t e1) -> hipError_t {
        WalkLaunch a{};
        a.rows = H->d_rows;
        a.nodesc = H->nodesc;
        a.adj = H->d_adj;
        a.src = b.src + b0;
        a.dst = b.dst + b0;
        a.w = b.w + b0;
        a.eid_base = b0;
        a.n = n;
        a.walkers = p->walkers;
        a.max_steps = p->max_steps;
        a.thr = p->threshold;
        a.flags = p->flags;
        a.edges_per_block = p->block_edges;
        a.seed = p->seed;
        a.found = b.found + b0;
        a.cond = b.cond + b0;
        a.stats = b.stats;
        return launch_walk_incremental(a, s, e0, e1);  // the events time the kernel itself
    };
    if (device_edits && m > 0) {
        ret = device_edits_begin(H, I, m);
        // every batch's row entries, sorted once before the first walk: a
        // batch's appends are then two launches (take the entries of the
        // edges that found no path, refresh the touched rows) with no sort
        // of their own
        const uint32_t *orow = nullptr, *oent = nullptr;
        bool presorted = false;
        if (ret == DYG_OK)
            ret = device_edits_order_run(H, b.src, b.dst, m, I, &orow, &oent, &presorted);
        uint32_t k = 0;
        for (uint32_t b0 = 0; b0 < m && ret == DYG_OK; b0 += I, ++k) {
            const uint32_t n = std::min(I, m - b0);
            hipError_t e = launch(b0, n, ctx->evpool[3 * k], ctx->evpool[3 * k + 1]);
            if (e != hipSuccess) {
                set_error(ctx, std::string("incremental walk: ") + hipGetErrorString(e));
                ret = DYG_ERR_HIP;
                break;
            }
            ret = presorted
                      ? device_edits_sorted(H, b.src + b0, b.dst + b0, b.w + b0, orow + 2ull * b0,
                                            oent + 2ull * b0, n,
                                            EditTake{b.found + b0, -1, nullptr, nullptr, nullptr})
                      : device_edits_batch(H, b.src + b0, b.dst + b0, b.w + b0, b.found + b0, n);
            if (ret == DYG_OK && hipEventRecord(ctx->evpool[3 * k + 2], s) != hipSuccess) {
                set_error(ctx, "incremental: event record failed");
                ret = DYG_ERR_HIP;
            }
            rs.batches++;
        }
        uint64_t added = 0;
        const dyg_status end = device_edits_end(H, &added);  // waits for the stream
        if (ret == DYG_OK) ret = end;
        rs.added = added;
        for (uint32_t q = 0; ret == DYG_OK && q < k; ++q) {
            float a_ms = 0.f, e_ms = 0.f;
            (void)hipEventElapsedTime(&a_ms, ctx->evpool[3 * q], ctx->evpool[3 * q + 1]);
            (void)hipEventElapsedTime(&e_ms, ctx->evpool[3 * q + 1], ctx->evpool[3 * q + 2]);
            rs.walk_ms += a_ms;
            rs.update_ms += e_ms;
        }
        if (ret == DYG_OK && h_found)
            DYG_HIP(ctx, hipMemcpy(h_found, b.found, m * 4ull, hipMemcpyDeviceToHost));
    } else if (m > 0) {
        // host-planned edits: host copies of the inputs are needed
        std::vector<uint32_t> hs, hd;
        std::vector<float> hw;
        if (!h_src || !h_dst || !h_w) {
            hs.resize(m);
            hd.resize(m);
            hw.resize(m);
            DYG_HIP(ctx, hipMemcpyAsync(hs.data(), b.src, m * 4ull, hipMemcpyDeviceToHost, s));
            DYG_HIP(ctx, hipMemcpyAsync(hd.data(), b.dst, m * 4ull, hipMemcpyDeviceToHost, s));
            DYG_HIP(ctx, hipMemcpyAsync(hw.data(), b.w, m * 4ull, hipMemcpyDeviceToHost, s));
            DYG_HIP(ctx, hipStreamSynchronize(s));
            h_src = hs.data();
            h_dst = hd.data();
            h_w = hw.data();
        }
        std::vector<int32_t> found(m);
        std::vector<uint32_t> au, av;
        std::vector<float> aw;
        for (uint32_t b0 = 0; b0 < m && ret == DYG_OK; b0 += I) {
            const uint32_t n = std::min(I, m - b0);
            hipError_t e = launch(b0, n, ctx->ev0, ctx->ev1);
            if (e == hipSuccess)
                e = hipMemcpyAsync(found.data() + b0, b.found + b0, n * 4ull,
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
            if (e != hipSuccess) {
                set_error(ctx, std::string("incremental walk: ") + hipGetErrorString(e));
                ret = DYG_ERR_HIP;
                break;
            }
            rs.walk_ms += ms;
            const double tu = now_ms();
            au.clear();
            av.clear();
            aw.clear();
            for (uint32_t i = b0; i < b0 + n; ++i)
                if (found[i] == -1) {
                    au.push_back(h_src[i]);
                    av.push_back(h_dst[i]);
                    aw.push_back(h_w[i]);
                }
            rs.added += au.size();
            ret = graph_apply_inserts(H, au.data(), av.data(), aw.data(), (uint32_t)au.size());
            rs.update_ms += now_ms() - tu;
            rs.batches++;
        }
        if (ret == DYG_OK) {
            DYG_HIP(ctx, hipStreamSynchronize(s));
            if (h_found) std::copy(found.begin(), found.end(), h_found);
        }
    }
    if (ret != DYG_OK) return ret;
    unsigned long long hst[3] = {0, 0, 0};
    DYG_HIP(ctx, hipMemcpy(hst, b.stats, sizeof(hst), hipMemcpyDeviceToHost));
    rs.updates = m;
    rs.hops = hst[0];
    rs.redirects = hst[1];
    rs.leaf_stops = hst[2];
    rs.total_ms = now_ms() - t0;
    if (out) *out = rs;
    return DYG_OK;
}

__global__ void k_check_ids(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                            uint32_t m, uint32_t V, uint32_t* __restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && (src[i] >= V || dst[i] >= V)) atomicOr(bad, 1u);
}

// Device-pointer inputs are checked on the device before any walk reads them.
dyg_status check_ids_dev(dyg_graph H, const uint32_t* src, const uint32_t* dst, uint32_t m,
                         unsigned long long* scratch) {
    dyg_ctx ctx = H->ctx;
    if (m == 0) return DYG_OK;
    uint32_t* bad = (uint32_t*)scratch;
    hipStream_t s = ctx->stream;
    DYG_HIP(ctx, hipMemsetAsync(bad, 0, 4, s));
    hipLaunchKernelGGL(k_check_ids, dim3((m + 255) / 256), dim3(256), 0, s, src, dst, m, H->V, bad);
    DYG_HIP(ctx, hipGetLastError());
    uint32_t h = 0;
    DYG_HIP(ctx, hipMemcpyAsync(&h, bad, 4, hipMemcpyDeviceToHost, s));
    DYG_HIP(ctx, hipStreamSynchronize(s));
    if (h) {
        set_error(ctx, "incremental: vertex id out of range");
        return DYG_ERR_ARG;
    }
    return DYG_OK;
}

dyg_status check_ids(dyg_graph H, const uint32_t* src, const uint32_t* dst, uint32_t m) {
    for (uint32_t i = 0; i < m; ++i)
        if (src[i] >= H->V || dst[i] >= H->V) {
            set_error(H->ctx, "incremental: vertex id out of range");
            return DYG_ERR_ARG;
        }
    return DYG_OK;
}

}  // namespace
}  // namespace dyg

using namespace dyg;

extern "C" dyg_status dyg_graph_reserve(dyg_graph g, uint64_t max_inserts, uint32_t max_batch) {
    if (!g) return DYG_ERR_ARG;
    if (g->keyed) return DYG_OK;  // keyed graphs take host-planned edits
    const dyg_status st = device_edits_prepare(g, max_batch, max_inserts);
    if (st != DYG_OK) return st;
    // the run-wide order of a run of max_inserts updates
    return max_inserts <= 0x3FFFFFFFull ? device_edits_reserve_run(g, max_inserts) : DYG_OK;
}

extern "C" dyg_status dyg_incremental_update(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                             const float* w, uint32_t m, uint32_t divisor,
                                             const dyg_walk_params* p, int32_t* found_at,
                                             dyg_run_stats* st) {
    if (!H || divisor == 0 || (m && (!src || !dst || !w))) return DYG_ERR_ARG;
    dyg_ctx ctx = H->ctx;
    dyg_status cs = check_params(ctx, p);
    if (cs != DYG_OK) return cs;
    cs = check_ids(H, src, dst, m);
    if (cs != DYG_OK) return cs;
    const size_t mb = std::max<size_t>(m, 1);
    // one device buffer holds inputs, decisions, conductances and counters
    const size_t off_dst = mb * 4, off_w = 2 * mb * 4, off_found = 3 * mb * 4,
                 off_cond = 4 * mb * 4, off_stats = (5 * mb * 4 + 255) & ~(size_t)255;
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, ctx->dscratch.reserve(off_stats + 64));
    char* base = (char*)ctx->dscratch.p;
    RunBufs b{(uint32_t*)base, (uint32_t*)(base + off_dst), (float*)(base + off_w),
              (int32_t*)(base + off_found), (float*)(base + off_cond),
              (unsigned long long*)(base + off_stats)};
    hipStream_t s = ctx->stream;
    DYG_HIP(ctx, hipMemcpyAsync(b.src, src, m * 4ull, hipMemcpyHostToDevice, s));
    DYG_HIP(ctx, hipMemcpyAsync(b.dst, dst, m * 4ull, hipMemcpyHostToDevice, s));
    DYG_HIP(ctx, hipMemcpyAsync(b.w, w, m * 4ull, hipMemcpyHostToDevice, s));
    return run_incremental(H, b, src, dst, w, m, divisor, p, found_at, st);
}

extern "C" dyg_status dyg_incremental_update_dev(dyg_graph H, const uint32_t* src,
                                                 const uint32_t* dst, const float* w, uint32_t m,
                                                 uint32_t divisor, const dyg_walk_params* p,
                                                 int32_t* found_at, float* cond,
                                                 dyg_run_stats* st) {
    if (!H || divisor == 0 || (m && (!src || !dst || !w || !found_at || !cond)))
        return DYG_ERR_ARG;
    dyg_ctx ctx = H->ctx;
    dyg_status cs = check_params(ctx, p);
    if (cs != DYG_OK) return cs;
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    DYG_HIP(ctx, ctx->dstats.reserve(64));
    cs = check_ids_dev(H, src, dst, m, (unsigned long long*)ctx->dstats.p + 4);
    if (cs != DYG_OK) return cs;
    RunBufs b{const_cast<uint32_t*>(src), const_cast<uint32_t*>(dst), const_cast<float*>(w),
              found_at, cond, (unsigned long long*)ctx->dstats.p};
    return run_incremental(H, b, nullptr, nullptr, nullptr, m, divisor, p, nullptr, st);
}
