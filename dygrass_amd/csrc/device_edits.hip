// device_edits.hip -- snapshot-batch inserts applied on the device.
//
// After a batch's walks, the edges whose walks found no path are appended to
// H in batch order, (u,v) to u's row then (v,u) to v's row, exactly as the
// host plan does (random_walk_incremental/gpu_graph.cuh:297-329), but without
// a host round trip:
//   k_keys      append keys by row (pad keys for edges that found a path)
//               and the batch's append count
//   radix sort  stable sort of appends by row (hipCUB), so each row's
//               appends stay in batch order
//   k_heads     first append of each row records its position; counts per row
//   k_alloc     per row: relocate to the arena tail when the row is full
//               (device bump allocator, order kept) and publish (start, deg)
//   k_write     each append writes its slot at start + old degree + rank
//   k_refresh   per touched row: every slot pointing to it (found through its
//               neighbours' rows) gets the row's new (first slot, degree)
// Slot order is what the walks index, and it equals the host plan's; the
// physical position of a relocated row may differ, which no walk can see.
//
// Deletions (random_walk_decremental/gpu_graph.cuh:338-400 semantics) use the
// same keys, sort and heads; then one thread per touched row removes that
// row's entries in batch order, each by swap-with-last, and k_refresh
// publishes the new degrees.  Edge lists may carry their length in device
// memory (n_dev), so a batch whose size is decided on the device needs no
// host round trip.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "internal.h"

namespace dyg {
namespace {

// Sort keys of one batch: entry 2i / 2i+1 is the edit of src[i]'s / dst[i]'s
// row when found[i] == match (-1: no path, append), else a pad key that sorts
// after every row.  The
// values carry the entry index, so a stable sort keeps each row's appends in
// batch order.  Also counts the batch's appends and clears the other
// parity's counter for the next batch.
__global__ __launch_bounds__(256) void k_keys(
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
    const int32_t* __restrict__ found, int32_t match, uint32_t n,
    const uint32_t* __restrict__ n_dev, uint32_t* __restrict__ keys,
    uint32_t* __restrict__ vals, uint32_t* __restrict__ count,
    uint32_t* __restrict__ count_next, unsigned long long* __restrict__ added_total) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i == 0) *count_next = 0;
    const uint32_t nn = n_dev ? min(*n_dev, n) : n;
    const bool add = i < nn && found[i] == match;
    if (i < n) {
        keys[2 * i] = add ? src[i] : 0xFFFFFFFFu;
        keys[2 * i + 1] = add ? dst[i] : 0xFFFFFFFFu;
        vals[2 * i] = 2 * i;
        vals[2 * i + 1] = 2 * i + 1;
    }
    // one pair of atomics per workgroup: per-wave atomics on the same two
    // words serialise at the cache and cost more than the rest of the kernel
    const uint32_t c = (uint32_t)__syncthreads_count(add);
    if (threadIdx.x == 0 && c) {
        atomicAdd(count, 2u * c);
        atomicAdd(added_total, (unsigned long long)c);
    }
}

__device__ __forceinline__ bool takes(const EditTake& t, uint32_t i) {
    return t.keep ? (t.keep[i] && t.found[t.jmap[i]] == -1) : t.flag[i] == t.match;
}

// Presorted edits (BatchOrder row entries, all 2n of them sorted by row):
// the thread of each row's first entry does the row's whole edit -- the
// taken entries of its run, in order -- so no compaction, head or count
// kernels are needed; it marks its position in `mark` for the descriptor
// refresh (every position is written: 1 at a touched row's first entry).
// Edges edited are counted once, at their even entry, one atomic per block.
constexpr uint32_t kTakeBlock = 256;

__device__ __forceinline__ uint32_t run_end(const uint32_t* __restrict__ row, uint32_t p,
                                            uint32_t n2) {
    const uint32_t r = row[p];
    uint32_t e = p + 1;
    while (e < n2 && row[e] == r) ++e;
    return e;
}

// Appends (k_alloc + k_write of the sorted path): relocation to the arena
// tail when the row is full, order kept; the neighbour's descriptor is
// written by the refresh of that neighbour.
__global__ __launch_bounds__(kTakeBlock) void k_take_append(
    uint2* __restrict__ rows, uint4* __restrict__ adj, uint32_t* __restrict__ cap,
    const uint32_t* __restrict__ row, const uint32_t* __restrict__ ent, uint32_t n2, EditTake tk,
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
    const float* __restrict__ w, uint8_t* __restrict__ mark, unsigned long long* __restrict__ used,
    unsigned long long cap_slots, uint32_t slack, uint32_t* __restrict__ err,
    unsigned long long* __restrict__ edges_total) {
    const uint32_t p = blockIdx.x * kTakeBlock + threadIdx.x;
    bool even_taken = false, touched = false;
    if (p < n2) {
        const uint32_t x = ent[p];
        even_taken = !(x & 1u) && takes(tk, x >> 1);
        const uint32_t r = row[p];
        if (p == 0 || row[p - 1] != r) {
            const uint32_t end = run_end(row, p, n2);
            uint32_t cnt = 0;
            for (uint32_t z = p; z < end; ++z) cnt += takes(tk, ent[z] >> 1) ? 1u : 0u;
            if (cnt) {
                touched = true;
                uint2 rw = rows[r];
                const uint32_t d0 = rw.y, c = cap[r];
                bool ok = true;
                if ((unsigned long long)d0 + cnt > c) {
                    unsigned long long ncap = 2ull * c;
                    ncap = max(ncap, (unsigned long long)d0 + cnt + slack);
                    ncap = max(ncap, 4ull);
                    const unsigned long long ns = atomicAdd(used, ncap);
                    if (ns + ncap > cap_slots) {
                        atomicOr(err, 1u);
                        ok = false;
                    } else {
                        copy_slots(adj + ns, adj + rw.x, d0);
                        rw.x = (uint32_t)ns;
                        cap[r] = (uint32_t)ncap;
                    }
                }
                if (ok) {
                    uint32_t o = rw.x + d0;
                    for (uint32_t z = p; z < end; ++z) {
                        const uint32_t y = ent[z], k = y >> 1;
                        if (!takes(tk, k)) continue;
                        const uint32_t nbr = (y & 1u) ? src[k] : dst[k];
                        adj[o++] = make_uint4(nbr, __float_as_uint(w[k]), 0u, 0u);
                    }
                    rows[r] = make_uint2(rw.x, d0 + cnt);
                } else {
                    touched = false;
                }
            }
        }
        mark[p] = touched ? 1 : 0;
    }
    const uint32_t c = (uint32_t)__syncthreads_count(even_taken);
    if (threadIdx.x == 0 && c) atomicAdd(edges_total, (unsigned long long)c);
}

// Deletions (k_del_rows of the sorted path): the run's taken entries in
// order, each by swap-with-last.
__global__ __launch_bounds__(kTakeBlock) void k_take_delete(
    uint2* __restrict__ rows, uint4* __restrict__ adj, const uint32_t* __restrict__ row,
    const uint32_t* __restrict__ ent, uint32_t n2, EditTake tk, const uint32_t* __restrict__ src,
    const uint32_t* __restrict__ dst, uint8_t* __restrict__ mark, uint32_t* __restrict__ err,
    unsigned long long* __restrict__ edges_total) {
    const uint32_t p = blockIdx.x * kTakeBlock + threadIdx.x;
    bool even_taken = false, touched = false;
    if (p < n2) {
        const uint32_t x = ent[p];
        even_taken = !(x & 1u) && takes(tk, x >> 1);
        const uint32_t r = row[p];
        if (p == 0 || row[p - 1] != r) {
            const uint32_t end = run_end(row, p, n2);
            uint2 rw = rows[r];
            for (uint32_t z = p; z < end; ++z) {
                const uint32_t y = ent[z], k = y >> 1;
                if (!takes(tk, k)) continue;
                touched = true;
                const uint32_t nbr = (y & 1u) ? src[k] : dst[k];
                const int sl = slot_of(adj, rw, nbr);
                if (sl < 0) {  // not present: the caller's lists were inconsistent
                    atomicOr(err, 2u);
                    continue;
                }
                const uint32_t last = rw.y - 1u;
                if ((uint32_t)sl != last) adj[rw.x + sl] = adj[rw.x + last];
                rw.y = last;
            }
            if (touched) rows[r] = rw;
        }
        mark[p] = touched ? 1 : 0;
    }
    const uint32_t c = (uint32_t)__syncthreads_count(even_taken);
    if (threadIdx.x == 0 && c) atomicAdd(edges_total, (unsigned long long)c);
}

// The descriptor refresh of the sorted path: LPR lanes per marked entry (a
// touched row's first entry), as k_refresh.
template <uint32_t LPR>
__global__ __launch_bounds__(256) void k_refresh_marked(const uint2* __restrict__ rows,
                                                        uint4* __restrict__ adj,
                                                        const uint32_t* __restrict__ row,
                                                        const uint8_t* __restrict__ mark,
                                                        uint32_t n2) {
    const uint32_t p = (blockIdx.x * 256u + threadIdx.x) / LPR;
    const uint32_t lane = threadIdx.x & (LPR - 1u);
    if (p >= n2 || !mark[p]) return;
    const uint32_t r = row[p];
    const uint2 rr = rows[r];
    for (uint32_t q = lane; q < rr.y; q += LPR) {
        const uint32_t x = adj[rr.x + q].x;
        refresh_pointers(adj, rows[x], r, rr);
    }
}

__global__ void k_heads(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ count,
                        uint32_t* __restrict__ segstart, uint32_t* __restrict__ segcnt) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *count) return;
    const uint32_t r = keys[p];
    if (p == 0 || keys[p - 1] != r) segstart[r] = p;
    atomicAdd(&segcnt[r], 1u);
}

__global__ void k_alloc(uint2* __restrict__ rows, uint4* __restrict__ adj,
                        uint32_t* __restrict__ cap, const uint32_t* __restrict__ keys,
                        const uint32_t* __restrict__ count, const uint32_t* __restrict__ segstart,
                        const uint32_t* __restrict__ segcnt, uint32_t* __restrict__ rowbase,
                        unsigned long long* __restrict__ used, unsigned long long cap_slots,
                        uint32_t slack, uint32_t* __restrict__ err) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *count) return;
    const uint32_t r = keys[p];
    if (segstart[r] != p) return;
    const uint32_t k = segcnt[r];
    uint2 row = rows[r];
    const uint32_t d0 = row.y;
    const uint32_t c = cap[r];
    if ((unsigned long long)d0 + k > c) {
        unsigned long long ncap = 2ull * c;
        ncap = max(ncap, (unsigned long long)d0 + k + slack);
        ncap = max(ncap, 4ull);
        const unsigned long long ns = atomicAdd(used, ncap);
        if (ns + ncap > cap_slots) {
            atomicOr(err, 1u);
            rowbase[r] = 0xFFFFFFFFu;
            return;
        }
        copy_slots(adj + ns, adj + row.x, d0);
        row.x = (uint32_t)ns;
        cap[r] = (uint32_t)ncap;
    }
    rowbase[r] = row.x + d0;
    rows[r] = make_uint2(row.x, d0 + k);
}

__global__ void k_write(uint4* __restrict__ adj, const uint32_t* __restrict__ keys,
                        const uint32_t* __restrict__ vals, const uint32_t* __restrict__ count,
                        uint32_t* __restrict__ segstart, uint32_t* __restrict__ segcnt,
                        const uint32_t* __restrict__ rowbase, const uint32_t* __restrict__ src,
                        const uint32_t* __restrict__ dst, const float* __restrict__ w) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *count) return;
    const uint32_t r = keys[p];
    const uint32_t base = rowbase[r];
    const uint32_t st = segstart[r];
    if (st == p) segcnt[r] = 0;  // read by k_alloc only: ready for the next batch
    if (base != 0xFFFFFFFFu) {
        const uint32_t o = vals[p];
        const uint32_t k = o >> 1;
        const uint32_t nbr = (o & 1u) ? src[k] : dst[k];
        // the neighbour's descriptor is written by k_refresh of that neighbour
        adj[base + (p - st)] = make_uint4(nbr, __float_as_uint(w[k]), 0u, 0u);
    }
}

// One thread per touched row: remove the row's entries in batch order, each
// by swap-with-last (the last live slot moves into the hole; its copy of its
// neighbour's row descriptor moves with it).
__global__ void k_del_rows(uint2* __restrict__ rows, uint4* __restrict__ adj,
                           const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                           const uint32_t* __restrict__ count,
                           const uint32_t* __restrict__ segstart,
                           uint32_t* __restrict__ segcnt, const uint32_t* __restrict__ src,
                           const uint32_t* __restrict__ dst, uint32_t* __restrict__ err) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *count) return;
    const uint32_t r = keys[p];
    if (segstart[r] != p) return;
    const uint32_t k = segcnt[r];
    segcnt[r] = 0;  // ready for the next batch
    uint2 row = rows[r];
    for (uint32_t e = p; e < p + k; ++e) {
        const uint32_t o = vals[e];
        const uint32_t nbr = (o & 1u) ? src[o >> 1] : dst[o >> 1];
        const int s = slot_of(adj, row, nbr);
        if (s < 0) {  // not present: the caller's lists were inconsistent
            atomicOr(err, 2u);
            continue;
        }
        const uint32_t q = (uint32_t)s;
        const uint32_t last = row.y - 1u;
        if (q != last) adj[row.x + q] = adj[row.x + last];
        row.y = last;
    }
    rows[r] = row;
}

// One group of LPR lanes per sorted entry; a touched row's first entry
// rewrites r's descriptor in every slot that points to r, its lanes taking
// the row's neighbours (so a lane's chain of dependent loads is one
// neighbour's row, not all of them; LPR near the graph's mean degree keeps
// the lanes busy).  Each such slot is written only by r's group and only its
// upper half (.z,.w), so groups never race; the neighbour ids and the row
// descriptors read here are final once k_write has run.  (segcnt was
// cleared by k_write / k_del_rows.)
template <uint32_t LPR>
__global__ __launch_bounds__(256) void k_refresh(
    const uint2* __restrict__ rows, uint4* __restrict__ adj, const uint32_t* __restrict__ keys,
    const uint32_t* __restrict__ count, uint32_t* __restrict__ segstart,
    uint32_t* __restrict__ segcnt) {
    const uint32_t p = (blockIdx.x * 256u + threadIdx.x) / LPR;
    const uint32_t lane = threadIdx.x & (LPR - 1u);
    if (p >= *count) return;
    const uint32_t r = keys[p];
    if (segstart[r] != p) return;
    const uint2 rr = rows[r];
    for (uint32_t q = lane; q < rr.y; q += LPR) {
        const uint32_t x = adj[rr.x + q].x;
        refresh_pointers(adj, rows[x], r, rr);
    }
}

void launch_refresh(dyg_graph g, const uint32_t* keys, const uint32_t* count, uint32_t entries,
                    hipStream_t s) {
    if (g->nodesc) return;  // the graph keeps no descriptors
    DevEdits& d = g->dev;
    const uint32_t lpr = refresh_lanes(g);
    const uint32_t blocks = (uint32_t)(((uint64_t)entries * lpr + 255) / 256);
    if (blocks == 0) return;
#define DYG_REFRESH(L)                                                                          \
    case L:                                                                                     \
        hipLaunchKernelGGL(k_refresh<L>, dim3(blocks), dim3(256), 0, s, g->d_rows, g->d_adj, keys, \
                           count, d.segstart, d.segcnt);                                        \
        break;
    switch (lpr) {
        DYG_REFRESH(4) DYG_REFRESH(8) DYG_REFRESH(16) DYG_REFRESH(32) default: DYG_REFRESH(64)
    }
#undef DYG_REFRESH
}

void launch_refresh_marked(dyg_graph g, const uint32_t* row, uint32_t n2, hipStream_t s) {
    if (g->nodesc) return;
    const uint32_t lpr = refresh_lanes(g);
    const uint32_t blocks = (uint32_t)(((uint64_t)n2 * lpr + 255) / 256);
    if (blocks == 0) return;
#define DYG_REFRESH_M(L)                                                                          \
    case L:                                                                                       \
        hipLaunchKernelGGL(k_refresh_marked<L>, dim3(blocks), dim3(256), 0, s, g->d_rows,         \
                           g->d_adj, row, g->dev.mark, n2);                                       \
        break;
    switch (lpr) {
        DYG_REFRESH_M(4) DYG_REFRESH_M(8) DYG_REFRESH_M(16) DYG_REFRESH_M(32)
        default: DYG_REFRESH_M(64)
    }
#undef DYG_REFRESH_M
}

}  // namespace

// Device state of a graph in device-edit mode (see internal.h).
dyg_status device_edits_prepare(dyg_graph g, uint32_t max_batch, uint64_t max_inserts) {
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    // worst-case arena need: every row's lifetime allocations form a doubling
    // series, bounded by 4x its final size, plus per-relocation slack
    // (per relocated row: allocations <= 2*last <= 4*final + 2*(slack+4));
    // the slots already in use stay allocated
    const uint64_t need = g->used_slots + 4 * (g->nslots + 2 * max_inserts) +
                          4 * max_inserts * (uint64_t)(g->slack + 4);
    dyg_status st = graph_reserve_slots(g, std::min<uint64_t>(need, 0xFFFFFFFFull));
    if (st != DYG_OK) return st;
    const size_t n2 = 2 * (size_t)std::max<uint32_t>(max_batch, 1);
    if (!d.ready || d.max_batch < max_batch) {
        if (g->mirror_stale) {  // capacities live only in the buffers about to go
            st = graph_refresh_mirror(g);
            if (st != DYG_OK) return st;
        }
        d.release();
        bool ok = hipMalloc(&d.cap, (size_t)g->V * 4) == hipSuccess &&
                  hipMalloc(&d.segstart, (size_t)g->V * 4) == hipSuccess &&
                  hipMalloc(&d.segcnt, (size_t)g->V * 4) == hipSuccess &&
                  hipMalloc(&d.rowbase, (size_t)g->V * 4) == hipSuccess &&
                  hipMalloc(&d.keys_in, n2 * 4) == hipSuccess &&
                  hipMalloc(&d.keys_out, n2 * 4) == hipSuccess &&
                  hipMalloc(&d.vals_in, n2 * 4) == hipSuccess &&
                  hipMalloc(&d.vals_out, n2 * 4) == hipSuccess &&
                  hipMalloc(&d.meta, 64) == hipSuccess &&
                  hipMalloc(&d.mark, n2) == hipSuccess;
        if (!ok) {
            d.release();
            set_error(ctx, "device edits: allocation failed");
            return DYG_ERR_NOMEM;
        }
        size_t tb = 0;
        DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, d.keys_in, d.keys_out,
                                                        d.vals_in, d.vals_out, (int)n2, 0, 32,
                                                        ctx->stream));
        if (hipMalloc(&d.temp, tb) != hipSuccess) {
            d.release();
            set_error(ctx, "device edits: allocation failed");
            return DYG_ERR_NOMEM;
        }
        d.temp_bytes = tb;
        d.max_batch = max_batch;
        d.ready = true;
    }
    // row capacities: uploaded when the host's are newer than the device's
    if (!d.cap_valid) {
        DYG_HIP(ctx, hipMemcpyAsync(d.cap, g->cap.data(), (size_t)g->V * 4,
                                    hipMemcpyHostToDevice, ctx->stream));
        DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
        d.cap_valid = true;
    }
    return DYG_OK;
}

dyg_status device_edits_begin(dyg_graph g, uint32_t max_batch, uint64_t max_inserts) {
    dyg_status st = device_edits_prepare(g, max_batch, max_inserts);
    if (st != DYG_OK) return st;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    hipStream_t s = ctx->stream;
    DYG_HIP(ctx, hipMemsetAsync(d.segcnt, 0, (size_t)g->V * 4, s));
    // meta: [0] used slots (u64), [1] appended edges (u64), [2] error flags
    // (u32), [3] entries of the current batch (u32, two by batch parity),
    // [4] deleted edges (u64)
    unsigned long long meta[5] = {g->used_slots, 0, 0, 0, 0};
    DYG_HIP(ctx, ctx->stage.reserve(64));
    DYG_HIP(ctx, hipStreamSynchronize(s));
    std::memcpy(ctx->stage.p, meta, sizeof(meta));
    DYG_HIP(ctx, hipMemcpyAsync(d.meta, ctx->stage.p, sizeof(meta), hipMemcpyHostToDevice, s));
    d.seq = 0;
    d.base_nslots = g->nslots;
    g->mirror_stale = true;
    return DYG_OK;
}

namespace {
// keys, sort and heads of one batch of row edits; returns the count pointer
dyg_status edit_keys(dyg_graph g, const uint32_t* src, const uint32_t* dst, const int32_t* found,
                     int32_t match, uint32_t n, const uint32_t* n_dev, uint32_t** count_out) {
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    hipStream_t s = ctx->stream;
    if (n > d.max_batch) {
        set_error(ctx, "device edits: batch larger than reserved");
        return DYG_ERR_ARG;
    }
    unsigned long long* added = d.meta + 1;
    uint32_t* counts = (uint32_t*)(d.meta + 3);  // entries of this batch, by parity
    uint32_t* count = counts + (d.seq & 1u);
    uint32_t* count_next = counts + ((d.seq + 1u) & 1u);
    ++d.seq;
    hipLaunchKernelGGL(k_keys, dim3((n + 255) / 256), dim3(256), 0, s, src, dst, found, match, n,
                       n_dev, d.keys_in, d.vals_in, count, count_next,
                       match == -1 ? added : d.meta + 4);
    DYG_HIP(ctx, hipGetLastError());
    const int end_bit = std::min(32, 33 - __builtin_clz(std::max<uint32_t>(g->V, 1)));
    size_t tb = d.temp_bytes;
    DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(d.temp, tb, d.keys_in, d.keys_out, d.vals_in,
                                                    d.vals_out, (int)(2 * n), 0, end_bit, s));
    const uint32_t blocks = (2 * n + 255) / 256;
    hipLaunchKernelGGL(k_heads, dim3(blocks), dim3(256), 0, s, d.keys_out, count, d.segstart,
                       d.segcnt);
    DYG_HIP(ctx, hipGetLastError());
    *count_out = count;
    return DYG_OK;
}

}  // namespace

dyg_status device_deletes_batch(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                const int32_t* flag, uint32_t n, const uint32_t* n_dev) {
    if (n == 0) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    hipStream_t s = ctx->stream;
    uint32_t* count = nullptr;
    dyg_status st = edit_keys(g, src, dst, flag, 1, n, n_dev, &count);
    if (st != DYG_OK) return st;
    uint32_t* err = (uint32_t*)(d.meta + 2);
    const uint32_t blocks = (2 * n + 255) / 256;
    hipLaunchKernelGGL(k_del_rows, dim3(blocks), dim3(256), 0, s, g->d_rows, g->d_adj, d.keys_out,
                       d.vals_out, count, d.segstart, d.segcnt, src, dst, err);
    launch_refresh(g, d.keys_out, count, 2 * n, s);
    DYG_HIP(ctx, hipGetLastError());
    return DYG_OK;
}

dyg_status device_edits_batch(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                              const float* w, const int32_t* found, uint32_t n,
                              const uint32_t* n_dev) {
    if (n == 0) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    hipStream_t s = ctx->stream;
    unsigned long long* used = d.meta;
    uint32_t* err = (uint32_t*)(d.meta + 2);
    uint32_t* count = nullptr;
    dyg_status st = edit_keys(g, src, dst, found, -1, n, n_dev, &count);
    if (st != DYG_OK) return st;
    const uint32_t blocks = (2 * n + 255) / 256;
    hipLaunchKernelGGL(k_alloc, dim3(blocks), dim3(256), 0, s, g->d_rows, g->d_adj, d.cap,
                       d.keys_out, count, d.segstart, d.segcnt, d.rowbase, used,
                       (unsigned long long)g->cap_slots, g->slack, err);
    hipLaunchKernelGGL(k_write, dim3(blocks), dim3(256), 0, s, g->d_adj, d.keys_out, d.vals_out,
                       count, d.segstart, d.segcnt, d.rowbase, src, dst, w);
    launch_refresh(g, d.keys_out, count, 2 * n, s);
    DYG_HIP(ctx, hipGetLastError());
    return DYG_OK;
}

dyg_status device_edits_sorted(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                               const float* w, const uint32_t* row, const uint32_t* ent,
                               uint32_t n, const EditTake& tk) {
    if (n == 0) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    if (n > d.max_batch) {
        set_error(ctx, "device edits: batch larger than reserved");
        return DYG_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    const uint32_t n2 = 2 * n;
    hipLaunchKernelGGL(k_take_append, dim3((n2 + kTakeBlock - 1) / kTakeBlock), dim3(kTakeBlock),
                       0, s, g->d_rows, g->d_adj, d.cap, row, ent, n2, tk, src, dst, w, d.mark,
                       d.meta, (unsigned long long)g->cap_slots, g->slack,
                       (uint32_t*)(d.meta + 2), d.meta + 1);
    launch_refresh_marked(g, row, n2, s);
    DYG_HIP(ctx, hipGetLastError());
    return DYG_OK;
}

dyg_status device_deletes_sorted(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                 const uint32_t* row, const uint32_t* ent, uint32_t n,
                                 const EditTake& tk) {
    if (n == 0) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    if (n > d.max_batch) {
        set_error(ctx, "device edits: batch larger than reserved");
        return DYG_ERR_ARG;
    }
    hipStream_t s = ctx->stream;
    const uint32_t n2 = 2 * n;
    hipLaunchKernelGGL(k_take_delete, dim3((n2 + kTakeBlock - 1) / kTakeBlock), dim3(kTakeBlock),
                       0, s, g->d_rows, g->d_adj, row, ent, n2, tk, src, dst, d.mark,
                       (uint32_t*)(d.meta + 2), d.meta + 4);
    launch_refresh_marked(g, row, n2, s);
    DYG_HIP(ctx, hipGetLastError());
    return DYG_OK;
}

namespace {

// Keys of a run's row entries (device_edits_order_run): batch b = i / I in
// the bits above rb, the row below; the value is the entry's index within
// its batch.
__global__ __launch_bounds__(256) void k_run_keys(const uint32_t* __restrict__ src,
                                                  const uint32_t* __restrict__ dst, uint32_t m,
                                                  uint32_t I, int rb, uint32_t* __restrict__ key,
                                                  uint32_t* __restrict__ val) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m) return;
    const uint32_t b = i / I, l = i - b * I;
    const uint32_t hi = rb < 32 ? b << rb : 0u;
    key[2 * i] = hi | src[i];
    key[2 * i + 1] = hi | dst[i];
    val[2 * i] = 2 * l;
    val[2 * i + 1] = 2 * l + 1;
}

// The sorted keys' row bits alone (the batch bits only grouped the entries).
__global__ __launch_bounds__(256) void k_run_rows(uint32_t* __restrict__ key, uint32_t n2,
                                                  uint32_t mask) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p < n2) key[p] &= mask;
}

int bits_for(uint64_t x) {  // bits that hold every value below x (x >= 1)
    int b = 0;
    while (b < 64 && (x - 1) >> b) ++b;
    return b;
}

size_t run_array_bytes(uint64_t m) { return ((2 * m * 4 + 255) / 256) * 256; }

}  // namespace

dyg_status device_edits_reserve_run(dyg_graph g, uint64_t m) {
    DevEdits& d = g->dev;
    if (m == 0 || m <= d.run_cap) return DYG_OK;
    dyg_ctx ctx = g->ctx;
    if (2 * m > 0x7FFFFFFFull) {
        set_error(ctx, "device edits: run too large to order at once");
        return DYG_ERR_ARG;
    }
    size_t tb = 0;
    DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint32_t*)nullptr,
                                                    (uint32_t*)nullptr, (uint32_t*)nullptr,
                                                    (uint32_t*)nullptr, (int)(2 * m), 0, 32,
                                                    ctx->stream));
    DYG_HIP(ctx, d.run.reserve(4 * run_array_bytes(m) + tb));
    d.run_cap = m;
    d.run_temp_bytes = tb;
    return DYG_OK;
}

dyg_status device_edits_order_run(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                  uint32_t m, uint32_t I, const uint32_t** row,
                                  const uint32_t** ent, bool* ok) {
    *ok = false;
    // (a run too long for one sort's item count keeps the per-batch pipeline)
    if (m == 0 || I == 0 || 2ull * m > 0x7FFFFFFFull) return DYG_OK;
    const int rb = bits_for(std::max<uint32_t>(g->V, 2)), bb = bits_for((m + I - 1) / I);
    if (rb + bb > 32) return DYG_OK;
    dyg_status st = device_edits_reserve_run(g, m);
    if (st != DYG_OK) return st;
    dyg_ctx ctx = g->ctx;
    DevEdits& d = g->dev;
    hipStream_t s = ctx->stream;
    const size_t ab = run_array_bytes(m);
    char* base = (char*)d.run.p;
    auto* k_in = (uint32_t*)base;
    auto* v_in = (uint32_t*)(base + ab);
    auto* k_out = (uint32_t*)(base + 2 * ab);
    auto* v_out = (uint32_t*)(base + 3 * ab);
    void* temp = base + 4 * ab;
    const uint32_t n2 = 2 * m;
    hipLaunchKernelGGL(k_run_keys, dim3((m + 255) / 256), dim3(256), 0, s, src, dst, m, I, rb,
                       k_in, v_in);
    size_t tb = d.run_temp_bytes;
    DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(temp, tb, k_in, k_out, v_in, v_out, (int)n2,
                                                    0, rb + bb, s));
    if (bb > 0)
        hipLaunchKernelGGL(k_run_rows, dim3((n2 + 255) / 256), dim3(256), 0, s, k_out, n2,
                           rb < 32 ? (1u << rb) - 1u : 0xFFFFFFFFu);
    DYG_HIP(ctx, hipGetLastError());
    *row = k_out;
    *ent = v_out;
    *ok = true;
    return DYG_OK;
}

dyg_status device_edits_absorb(dyg_graph g, const unsigned long long* meta, uint64_t* added_out) {
    const uint32_t err = (uint32_t)meta[2];
    g->used_slots = meta[0];
    g->nslots = g->dev.base_nslots + 2 * meta[1] - 2 * meta[4];
    if (added_out) *added_out = meta[1];
    if (err & 1u) {
        set_error(g->ctx, "device edits: slot arena exhausted");
        return DYG_ERR_NOMEM;
    }
    if (err & 2u) {
        set_error(g->ctx, "device edits: deletion of an edge the graph does not hold");
        return DYG_ERR_STATE;
    }
    return DYG_OK;
}

// Waits for the stream, then brings the host's view of the arena up to date.
dyg_status device_edits_end(dyg_graph g, uint64_t* added_out) {
    dyg_ctx ctx = g->ctx;
    unsigned long long meta[5];
    DYG_HIP(ctx, hipMemcpyAsync(meta, g->dev.meta, sizeof(meta), hipMemcpyDeviceToHost,
                                ctx->stream));
    DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return device_edits_absorb(g, meta, added_out);
}

void DevEdits::release() {
    void* ps[] = {cap,  segstart, segcnt, rowbase, keys_in, keys_out,
                  vals_in, vals_out, meta, mark, temp};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    cap = segstart = segcnt = rowbase = keys_in = keys_out = vals_in = vals_out = nullptr;
    mark = nullptr;
    meta = nullptr;
    temp = nullptr;
    temp_bytes = 0;
    max_batch = 0;
    ready = false;
    cap_valid = false;
}

}  // namespace dyg
