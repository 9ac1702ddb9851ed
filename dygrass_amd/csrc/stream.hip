// stream.hip -- mixed insertion/deletion update stream over a device-resident
// dense graph G and its sparsifier H (C ABI dyg_stream_*, include/dygrass.h).
//
// Semantics follow dynamic_update/gpu_graph.cuh:361-439 (preprocessStreamEdges),
// :451-465 (weight update) and :620-700 (updateSparsiferFromResult), with the
// walks, block reduction and heuristic of the plain updaters
// (random_walk_incremental/main.cu:70-211, random_walk_decremental/main.cu:
// 77-290, random_walk_decremental/gpu_graph.cuh:509-615), per batch:
//
// insertion batch   (a,b) in G: weight += w in G and in H when there;
//                   else appended to G and kept; kept edges walk in H (state
//                   at batch start, after the batch's weight updates); those
//                   without a path are appended to H in batch order.
// deletion batch    (a,b) not in G: skipped; else removed from G and, when in
//                   H, from H and kept; kept edges walk in G after all of the
//                   batch's deletions, from min(a,b) to max(a,b); every hop of a
//                   found path that H lacks, then every heuristic edge (both
//                   endpoints, 16 walkers, two hops) that H lacks, is appended
//                   to H with G's weight, in that order.
// Duplicates inside a batch act in batch order (a second insertion of a new
// edge is a weight update of it; a second deletion is skipped).  Self-loops
// are rejected per update.  G and H must be simple graphs with H inside G.
//
// Every step runs on the device.  A batch of up to kSmallBatch updates runs
// as one-workgroup kernels around the walk (stream_small.hip).  A larger one
// takes the general pipeline: the batch's keys are sorted (stable, 64-bit
// (min,max) keys) so that one thread per distinct edge sees its occurrences
// in batch order; membership and positions come from scanning the smaller
// row; row edits use the append and swap-with-last pipelines of
// device_edits.hip; the walks use the work-queue kernels with the list
// length in device memory.  A batch's last kernel writes its outcome codes
// and counters into pinned host memory; the host waits once per batch (a
// deletion batch also reads its recovery list sizes once, after its walks).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "admission.h"
#include "internal.h"

namespace dyg {
namespace {

constexpr int32_t kOutWeight = 1, kOutFound = 2, kOutAdded = 3, kOutSkip = 4, kOutGOnly = 5,
                  kOutRecovered = 6, kOutHeuristic = 7, kOutInvalid = 8;
constexpr unsigned long long kPadKey = ~0ull;

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

__device__ __forceinline__ int row_find(const uint2* __restrict__ rows,
                                        const uint4* __restrict__ adj, uint32_t a, uint32_t b) {
    return slot_of(adj, rows[a], b);
}

// Also clears bcnt[block] (when given): the per-block counts a later
// k_compact over the same n reads.
// Range check of a batch of device inputs (larger than the one-workgroup
// path, which checks inside its pre kernel): any id >= V sets *bad.
// The device's wall clock into *w (a list's start, for its batches' spans).
__global__ void k_clock(unsigned long long* w) {
    if (threadIdx.x == 0) *w = wall_clock64();
}

__global__ void k_check_batch(const uint32_t* __restrict__ u, const uint32_t* __restrict__ v,
                              uint32_t n, uint32_t V, unsigned long long* bad) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n && (u[i] >= V || v[i] >= V)) *bad = 1ull;
}

__global__ void k_batch_keys(const uint32_t* __restrict__ u, const uint32_t* __restrict__ v,
                             uint32_t n, unsigned long long* __restrict__ key,
                             uint32_t* __restrict__ idx, uint32_t* __restrict__ bcnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (bcnt && threadIdx.x == 0) bcnt[blockIdx.x] = 0;
    if (i >= n) return;
    const uint32_t a = min(u[i], v[i]), b = max(u[i], v[i]);
    key[i] = a == b ? kPadKey : ((unsigned long long)a << 32) | b;
    idx[i] = i;
}

// k_batch_keys reading the batch from pinned host memory (u, v, w: w null for
// deletions) and staging it in device memory (su, sv, sw) for the kernels
// after it: the batch crosses PCIe once, with no separate copy.
__global__ void k_batch_keys_staged(const uint32_t* __restrict__ u, const uint32_t* __restrict__ v,
                                    const float* __restrict__ w, uint32_t n,
                                    unsigned long long* __restrict__ key,
                                    uint32_t* __restrict__ idx, uint32_t* __restrict__ bcnt,
                                    uint32_t* __restrict__ su, uint32_t* __restrict__ sv,
                                    float* __restrict__ sw) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) bcnt[blockIdx.x] = 0;
    if (i >= n) return;
    const uint32_t x = u[i], y = v[i];
    su[i] = x;
    sv[i] = y;
    if (w) sw[i] = w[i];
    const uint32_t a = min(x, y), b = max(x, y);
    key[i] = a == b ? kPadKey : ((unsigned long long)a << 32) | b;
    idx[i] = i;
}

// A batch's orders computed ahead of it (BatchOrder), before their sorts:
// its edge keys with their positions, its row entries (2i: u[i]'s row, 2i+1:
// v[i]'s) with their indices, and its cleared per-block kept counts.
__global__ void k_prep_keys(const uint32_t* __restrict__ u, const uint32_t* __restrict__ v,
                            uint32_t n, unsigned long long* __restrict__ key,
                            uint32_t* __restrict__ idx, uint32_t* __restrict__ erow,
                            uint32_t* __restrict__ eent, uint32_t* __restrict__ bcnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x == 0) bcnt[blockIdx.x] = 0;
    if (i >= n) return;
    const uint32_t x = u[i], y = v[i];
    const uint32_t a = min(x, y), b = max(x, y);
    key[i] = a == b ? kPadKey : ((unsigned long long)a << 32) | b;
    idx[i] = i;
    erow[2 * i] = x;
    erow[2 * i + 1] = y;
    eent[2 * i] = 2 * i;
    eent[2 * i + 1] = 2 * i + 1;
}

// One thread per distinct key (first position of its run in the sorted order).
// mode[p] at the head: 0 nothing more, 1 existing edge (weight updates for
// all occurrences; bit 2 set when H holds it too), 2 new edge with further
// occurrences (weight updates after the append).
__global__ void k_ins_groups(const unsigned long long* __restrict__ key,
                             const uint32_t* __restrict__ idx, uint32_t n,
                             const uint2* __restrict__ grows, const uint4* __restrict__ gadj,
                             const uint2* __restrict__ hrows, const uint4* __restrict__ hadj,
                             int32_t* __restrict__ cls, int32_t* __restrict__ flag,
                             uint8_t* __restrict__ keep, uint8_t* __restrict__ mode,
                             uint32_t* __restrict__ bcnt) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long k = key[p];
    if (p > 0 && key[p - 1] == k) return;
    uint32_t q = p + 1;
    while (q < n && key[q] == k) ++q;
    uint8_t md = 0;
    if (k == kPadKey) {
        for (uint32_t e = p; e < q; ++e) {
            cls[idx[e]] = kOutInvalid;
            flag[idx[e]] = 0;
            keep[idx[e]] = 0;
        }
    } else {
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        const bool in_g = has_edge(grows, gadj, a, b);
        for (uint32_t e = p; e < q; ++e) {
            const bool first_new = !in_g && e == p;
            cls[idx[e]] = first_new ? 0 : kOutWeight;
            flag[idx[e]] = first_new ? -1 : 0;
            keep[idx[e]] = first_new ? 1 : 0;
        }
        if (!in_g) atomicAdd(&bcnt[idx[p] >> 8], 1u);  // the kept first occurrence
        if (in_g) md = 1 | (has_edge(hrows, hadj, a, b) ? 4 : 0);
        else if (q - p > 1) md = 2;
    }
    mode[p] = md;
}

// Weight updates of one edge, in batch order, on both of its slots.
__device__ void add_weights(const uint2* __restrict__ rows, uint4* __restrict__ adj, uint32_t a,
                            uint32_t b, const uint32_t* __restrict__ idx, uint32_t e0,
                            uint32_t e1, const float* __restrict__ w) {
    const int pa = row_find(rows, adj, a, b), pb = row_find(rows, adj, b, a);
    if (pa < 0 || pb < 0) return;
    uint4* sa = &adj[rows[a].x + pa];
    uint4* sb = &adj[rows[b].x + pb];
    float x = __uint_as_float(sa->y);
    for (uint32_t e = e0; e < e1; ++e) x = x + w[idx[e]];
    sa->y = __float_as_uint(x);
    sb->y = __float_as_uint(x);
}

// An insertion batch's weight updates (after G's appends: a repeated new
// edge adds to the weight it was appended with), run by k_compact's threads
// at sorted position p (the head of each key's run).
struct InsWeights {
    const unsigned long long* key;  // null: a deletion batch, nothing to do
    const uint32_t* idx;
    const uint8_t* mode;
    const float* w;
    const uint2* grows;
    uint4* gadj;
    const uint2* hrows;
    uint4* hadj;
};
__device__ void ins_weights_at(const InsWeights& t, uint32_t p, uint32_t n) {
    const unsigned long long k = t.key[p];
    if (p > 0 && t.key[p - 1] == k) return;  // not a head: mode[p] is not written
    const uint8_t md = t.mode[p];
    if (md == 0) return;
    uint32_t q = p + 1;
    while (q < n && t.key[q] == k) ++q;
    const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
    const uint32_t e0 = (md & 1) ? p : p + 1;  // a new edge already holds its first weight
    add_weights(t.grows, t.gadj, a, b, t.idx, e0, q, t.w);
    if (md & 4) add_weights(t.hrows, t.hadj, a, b, t.idx, e0, q, t.w);
}

__global__ void k_del_groups(const unsigned long long* __restrict__ key,
                             const uint32_t* __restrict__ idx, uint32_t n,
                             const uint2* __restrict__ grows, const uint4* __restrict__ gadj,
                             const uint2* __restrict__ hrows, const uint4* __restrict__ hadj,
                             int32_t* __restrict__ cls, int32_t* __restrict__ flag_g,
                             int32_t* __restrict__ flag_h, uint8_t* __restrict__ keep,
                             uint32_t* __restrict__ bcnt) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const unsigned long long k = key[p];
    if (p > 0 && key[p - 1] == k) return;
    uint32_t q = p + 1;
    while (q < n && key[q] == k) ++q;
    bool in_g = false, in_h = false;
    if (k != kPadKey) {
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
        in_g = has_edge(grows, gadj, a, b);
        in_h = in_g && has_edge(hrows, hadj, a, b);
    }
    for (uint32_t e = p; e < q; ++e) {
        const uint32_t i = idx[e];
        const bool first = (e == p) && in_g;
        cls[i] = k == kPadKey ? kOutInvalid : !first ? kOutSkip : in_h ? 0 : kOutGOnly;
        flag_g[i] = first ? 1 : 0;
        flag_h[i] = first && in_h ? 1 : 0;
        keep[i] = first && in_h ? 1 : 0;
    }
    if (in_h) atomicAdd(&bcnt[idx[p] >> 8], 1u);  // the kept first occurrence
}

// Stable compaction of the kept updates (keep[i], batch order) in one pass:
// block b's output offset is the sum of the kept counts of blocks before it
// (bcnt, counted by the classification kernel), plus its own block scan.
// Gathers each kept update: (u, v) or, `ordered`, (min, max); its weight;
// its stream position; and (jmap set) its place in the kept list by batch
// position.  cnt[0] = kept total.  With wt.key set (an insertion batch) the
// same threads also apply the weight updates.  256 threads per block.
__global__ __launch_bounds__(256) void k_compact(
    const uint8_t* __restrict__ keep, uint32_t n, const uint32_t* __restrict__ bcnt,
    const uint32_t* __restrict__ u, const uint32_t* __restrict__ v, const float* __restrict__ w,
    uint64_t base, bool ordered, uint32_t* __restrict__ cu, uint32_t* __restrict__ cv,
    float* __restrict__ cw, uint64_t* __restrict__ ceid, uint32_t* __restrict__ sel,
    uint32_t* __restrict__ cnt, uint32_t* __restrict__ jmap, InsWeights wt) {
    using BScan = hipcub::BlockScan<uint32_t, 256>;
    using BReduce = hipcub::BlockReduce<uint32_t, 256>;
    __shared__ union {
        typename BScan::TempStorage scan;
        typename BReduce::TempStorage reduce;
    } tmp;
    __shared__ uint32_t s_prefix;
    const uint32_t b = blockIdx.x, t = threadIdx.x, i = b * 256u + t;
    uint32_t part = 0;
    for (uint32_t q = t; q < b; q += 256u) part += bcnt[q];
    const uint32_t prefix = BReduce(tmp.reduce).Sum(part);
    if (t == 0) s_prefix = prefix;
    __syncthreads();
    const bool k = i < n && keep[i];
    uint32_t pos = 0, total = 0;
    BScan(tmp.scan).ExclusiveSum(k ? 1u : 0u, pos, total);
    pos += s_prefix;
    if (k) {
        const uint32_t a = u[i], c = v[i];
        cu[pos] = ordered ? min(a, c) : a;
        cv[pos] = ordered ? max(a, c) : c;
        if (w) cw[pos] = w[i];
        ceid[pos] = base + i;
        sel[pos] = i;
        if (jmap) jmap[i] = pos;
    }
    if (b == gridDim.x - 1 && t == 0) cnt[0] = s_prefix + total;
    if (wt.key && i < n) ins_weights_at(wt, i, n);
}

// The batch's outcome codes straight into its report (k_outcome and k_report
// in one launch): a kept update's code comes from its walk (jmap: its place in
// the kept list), every other one's from the classification; block 0 also
// copies the counters.
__global__ __launch_bounds__(256) void k_outcome_report(
    Report r, const int32_t* __restrict__ cls, const uint8_t* __restrict__ keep,
    const uint32_t* __restrict__ jmap, const int32_t* __restrict__ found, uint32_t n,
    int32_t if_found, int32_t if_not, unsigned long long* __restrict__ stats) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) r.cls[i] = keep[i] ? (found[jmap[i]] == -1 ? if_not : if_found) : cls[i];
    if (blockIdx.x == 0) report_out(r, cls, 0, stats, threadIdx.x, 256);
}

__global__ void k_outcome(const uint32_t* __restrict__ sel, const uint32_t* __restrict__ nsel,
                          const int32_t* __restrict__ found, int32_t* __restrict__ cls,
                          int32_t if_found, int32_t if_not) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *nsel) return;
    cls[sel[j]] = found[j] == -1 ? if_not : if_found;
}

// Per kept deletion j: path hop count pc[j] (found walks) and the no-path
// flag; per block of 256, their sums packed as (hops << 32 | no-path) in
// bsum[block].
__global__ __launch_bounds__(256) void k_path_counts(
    const int32_t* __restrict__ found, const uint32_t* __restrict__ nsel, uint32_t n,
    uint32_t* __restrict__ pc, uint8_t* __restrict__ nopath,
    unsigned long long* __restrict__ bsum) {
    using BReduce = hipcub::BlockReduce<unsigned long long, 256>;
    __shared__ typename BReduce::TempStorage tmp;
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    const bool live = j < n && j < *nsel;
    const int32_t f = live ? found[j] : 0;
    const uint32_t c = f > 0 ? (uint32_t)f : 0u;
    const bool np = f == -1;
    if (j < n) {
        pc[j] = c;
        nopath[j] = np ? 1 : 0;
    }
    const unsigned long long sum = BReduce(tmp).Sum(((unsigned long long)c << 32) | (np ? 1u : 0u));
    if (threadIdx.x == 0) bsum[blockIdx.x] = sum;
}

// The recovery lists in one pass over the kept deletions (block prefixes
// from bsum): poff[j] = exclusive prefix of the hop counts; the no-path
// edges' endpoints ep[2k], ep[2k+1] (min, max) with their heuristic stream
// keys (position << 1 | side), k the no-path rank.  The last block writes
// cnt[1] = no-path count, cnt[2] = endpoints, cnt[3] = path hops and copies
// cnt[0..3] to `out` (pinned host memory).
__global__ __launch_bounds__(256) void k_recovery_lists(
    const uint32_t* __restrict__ pc, const uint8_t* __restrict__ nopath, uint32_t n,
    const unsigned long long* __restrict__ bsum, const uint32_t* __restrict__ ca,
    const uint32_t* __restrict__ cb, const uint64_t* __restrict__ ceid,
    uint32_t* __restrict__ poff, uint32_t* __restrict__ ep, uint64_t* __restrict__ ekey,
    uint32_t* __restrict__ cnt, uint32_t* __restrict__ out) {
    using BScan = hipcub::BlockScan<unsigned long long, 256>;
    using BReduce = hipcub::BlockReduce<unsigned long long, 256>;
    __shared__ union {
        typename BScan::TempStorage scan;
        typename BReduce::TempStorage reduce;
    } tmp;
    __shared__ unsigned long long s_prefix;
    const uint32_t b = blockIdx.x, t = threadIdx.x, j = b * 256u + t;
    unsigned long long part = 0;
    for (uint32_t q = t; q < b; q += 256u) part += bsum[q];
    const unsigned long long prefix = BReduce(tmp.reduce).Sum(part);
    if (t == 0) s_prefix = prefix;
    __syncthreads();
    const uint32_t c = j < n ? pc[j] : 0u;
    const bool np = j < n && nopath[j];
    unsigned long long off = 0, total = 0;
    BScan(tmp.scan).ExclusiveSum(((unsigned long long)c << 32) | (np ? 1u : 0u), off, total);
    off += s_prefix;
    if (j < n) poff[j] = (uint32_t)(off >> 32);
    if (np) {
        const uint32_t k = (uint32_t)off;
        ep[2 * k] = ca[j];
        ep[2 * k + 1] = cb[j];
        ekey[2 * k] = ceid[j] << 1;
        ekey[2 * k + 1] = (ceid[j] << 1) | 1ull;
    }
    if (b == gridDim.x - 1 && t == 0) {
        const unsigned long long all = s_prefix + total;
        cnt[1] = (uint32_t)all;
        cnt[2] = 2u * (uint32_t)all;
        cnt[3] = (uint32_t)(all >> 32);
        for (int q = 0; q < 4; ++q) out[q] = cnt[q];
    }
}

// Candidate list, part one: the hops of every found path, kept-edge order.
__global__ void k_cand_paths(const int32_t* __restrict__ found, const uint32_t* __restrict__ nsel,
                             const uint32_t* __restrict__ path, uint32_t max_steps,
                             const uint32_t* __restrict__ poff, uint32_t* __restrict__ xu,
                             uint32_t* __restrict__ xv) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *nsel || found[j] <= 0) return;
    const uint32_t* pj = path + (uint64_t)j * max_steps;
    const uint32_t o = poff[j];
    for (int q = 1; q <= found[j]; ++q) {
        xu[o + q - 1] = pj[q - 1];
        xv[o + q - 1] = pj[q];
    }
}

// Part two, after all path hops: per endpoint and walker, (u,n1) then (n1,n2).
__global__ void k_cand_heur(const uint32_t* __restrict__ path3, uint32_t n3, uint32_t base,
                            uint32_t* __restrict__ xu, uint32_t* __restrict__ xv) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n3) return;
    const uint32_t* t = path3 + (uint64_t)g * 3;
    xu[base + 2 * g] = t[0];
    xv[base + 2 * g] = t[1];
    xu[base + 2 * g + 1] = t[1];
    xv[base + 2 * g + 1] = t[2];
}


// Device buffers of one stream, carved from one allocation.
// Recovery candidates as row entries (entry 2c: xu[c]'s row, 2c+1: xv[c]'s),
// with their indices, for one stable sort by row.
__global__ void k_cand_entries(const uint32_t* __restrict__ xu, const uint32_t* __restrict__ xv,
                               uint32_t n, uint32_t* __restrict__ erow,
                               uint32_t* __restrict__ eent) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    erow[2 * c] = xu[c];
    erow[2 * c + 1] = xv[c];
    eent[2 * c] = 2 * c;
    eent[2 * c + 1] = 2 * c + 1;
}

// Each candidate c = (a, b), decided at its even entry (row a's run of the
// row-sorted entries, in candidate order): appended to H (flag -1, with G's
// weight) when it is the first candidate of edge {a, b} -- no earlier entry
// of the run has neighbour b, the same first occurrence a sort by edge key
// gives -- and H lacks the edge.  Every candidate's flag is written.
__global__ void k_cand_decide(const uint32_t* __restrict__ erow, const uint32_t* __restrict__ eent,
                              uint32_t n2, const uint32_t* __restrict__ xu,
                              const uint32_t* __restrict__ xv, const uint2* __restrict__ grows,
                              const uint4* __restrict__ gadj, const uint2* __restrict__ hrows,
                              const uint4* __restrict__ hadj, int32_t* __restrict__ flag,
                              float* __restrict__ xw) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n2) return;
    const uint32_t x = eent[p];
    if (x & 1u) return;
    const uint32_t c = x >> 1, a = erow[p], b = xv[c];
    int32_t f = 0;
    bool first = a != b;  // a self-loop (an isolated endpoint's heuristic) is never an edge
    for (uint32_t q = p; first && q > 0 && erow[q - 1] == a; --q) {
        const uint32_t y = eent[q - 1], k = y >> 1;
        first = ((y & 1u) ? xu[k] : xv[k]) != b;
    }
    if (first && !has_edge(hrows, hadj, a, b)) {
        const uint2 ra = grows[a];
        const int sl = slot_of(gadj, ra, b);
        if (sl >= 0) {  // always: path and heuristic hops are edges of G
            f = -1;
            xw[c] = __uint_as_float(gadj[ra.x + sl].y);
        }
    }
    flag[c] = f;
}

// Scratch for computing one batch's orders (BatchOrder) of up to n updates on
// a stream of its own: the keys' and row entries' sort inputs and the
// sorts' hipCUB scratch.
struct OrderScratch {
    void* mem = nullptr;
    unsigned long long* key = nullptr;
    uint32_t *idx = nullptr, *row = nullptr, *ent = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t alloc(uint32_t n, hipStream_t st) {
        hipError_t e = hipMalloc(&mem, 28ull * n + 1024);
        if (e != hipSuccess) {
            mem = nullptr;
            return e;
        }
        char* b = (char*)mem;
        key = (unsigned long long*)b;
        idx = (uint32_t*)(b + 8ull * n);
        row = (uint32_t*)(b + 12ull * n);
        ent = (uint32_t*)(b + 20ull * n);
        size_t t1 = 0, t2 = 0;
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, t1, key, key, idx, idx, (int)n, 0, 64, st);
        if (e == hipSuccess)
            e = hipcub::DeviceRadixSort::SortPairs(nullptr, t2, row, row, ent, ent, (int)(2 * n), 0,
                                                   32, st);
        if (e == hipSuccess) e = hipMalloc(&temp, std::max(t1, t2));
        if (e != hipSuccess) {
            temp = nullptr;
            return e;
        }
        temp_bytes = std::max(t1, t2);
        return hipSuccess;
    }
    ~OrderScratch() {
        if (mem) (void)hipFree(mem);
        if (temp) (void)hipFree(temp);
    }
};

// A host thread kept for the life of a stream, that runs one job at a time
// (a list's order enqueues): handing it a job costs a wake-up instead of a
// thread creation inside the call.
class Helper {
  public:
    explicit Helper(int device) : device_(device) {}
    ~Helper() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }
    // Starts the thread (once); false when it could not be started.
    bool start() {
        if (th_.joinable()) return true;
        try {
            th_ = std::thread([this] { loop(); });
        } catch (...) {
            return false;
        }
        return true;
    }
    // Hands `f` to the thread; the previous job has been waited for.
    void post(std::function<void()> f) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = std::move(f);
            busy_ = true;
        }
        cv_.notify_all();
    }
    // Returns once the posted job has finished.
    void wait() {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !busy_; });
    }

  private:
    void loop() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || job_ != nullptr; });
            if (job_ == nullptr) return;  // stopping, nothing left to run
            std::function<void()> f = std::move(job_);
            job_ = nullptr;
            lk.unlock();
            f();
            lk.lock();
            busy_ = false;
            cv_.notify_all();
        }
    }
    int device_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::function<void()> job_;
    bool busy_ = false, stop_ = false;
};

struct Arena {
    char* base = nullptr;
    size_t off = 0, cap = 0;
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = (T*)(base + off);
        off += std::max<size_t>(n, 1) * sizeof(T);
        return p;
    }
};

}  // namespace
}  // namespace dyg

using namespace dyg;

struct dyg_stream_s {
    dyg_ctx ctx = nullptr;
    dyg_graph G = nullptr, H = nullptr;
    dyg_stream_params sp{};
    uint64_t base = 0;  // stream position of the next update
    uint32_t cmax = 0;  // recovery candidates per deletion batch, at most
    void* mem = nullptr;
    // per-batch buffers (max_batch entries unless noted)
    uint32_t* in;  // [3 * max_batch]: a batch's u, v, w at stride n (one copy in)
    int32_t* cls;
    unsigned long long *k_in, *k_out;
    uint32_t *i_in, *i_out;
    int32_t *flag_g, *flag_h;
    uint8_t *keep, *mode;
    uint32_t* sel;
    uint32_t* cnt;  // [8] device counters: kept, no-path, endpoints, path hops
    uint32_t *cu, *cv;
    float *cw, *ccond;
    uint64_t* ceid;
    int32_t* cfound;
    uint32_t* path;  // [max_batch * max_steps]
    uint32_t *pc, *poff;
    uint8_t* nopath;
    uint32_t* bcnt;             // [blocks of 256] kept counts for k_compact
    uint32_t* jmap;             // [max_batch] kept-list place by batch position (BatchOrder)
    unsigned long long* bsum;   // [blocks of 256] path hop and no-path sums
    uint32_t* ep;     // [2 * max_batch]
    uint64_t* ekey;   // [2 * max_batch]
    uint32_t* path3;  // [2 * max_batch * heur_walkers * 3]
    uint32_t *xu, *xv;  // [cmax] recovery candidates
    float* xw;
    int32_t* xflag;
    unsigned long long *xk_in, *xk_out;
    uint32_t *xi_in, *xi_out;
    uint32_t *xr_in, *xr_out, *xe_in, *xe_out;  // [2 * cmax] candidates' row entries
    unsigned long long* stats;  // [4] walk counters
    // pinned: inputs (u, v, w), outcomes, report words, recovery sizes; the
    // device reaches it at pin_dev
    void* pin = nullptr;
    char* pin_dev = nullptr;
    hipEvent_t ev_sync = nullptr;  // polled by wait_spin
    void* temp = nullptr;
    size_t temp_bytes = 0;
    std::vector<int32_t> h_cls;
    // DYG_STREAM_GRAPH: the one-workgroup step of each operation and size
    // class, captured on first use
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        BlockGraph g{}, h{};  // G's and H's layout baked into the captured nodes
        bool failed = false;  // capture refused: the class launches directly
    };
    static constexpr int kStepClasses = 5;
    StepGraph steps[2][kStepClasses];
    StepDesc* desc = nullptr;      // pinned: the graphed batch's values
    StepDesc* desc_dev = nullptr;  // the same, as the device reaches it
    bool last_graph = false;       // the current batch ran as a step graph
    // events bracketing the current batch's walk (the context's pair, or a
    // batch's own pair while a list of batches is in flight)
    hipEvent_t wev0 = nullptr, wev1 = nullptr;
    // Orders of a list's larger batches (BatchOrder), computed on their own
    // stream while the batches before them run: the orders (grow-only), one
    // batch's sort inputs and the sorts' scratch, one event per batch.
    hipStream_t prep = nullptr;
    DevBuf order;
    OrderScratch oscr;
    std::vector<hipEvent_t> prep_ev;
    Pinned jobs_h;  // the one-workgroup batches' order jobs (SmallOrderJob), host side
    DevBuf jobs_d;  // and on the device
    // enqueues a list's larger batches' orders beside the list's own thread
    std::unique_ptr<Helper> helper;
    ~dyg_stream_s() {
        helper.reset();  // its last job has been waited for
        if (prep) {
            (void)hipStreamSynchronize(prep);
            (void)hipStreamDestroy(prep);
        }
        for (hipEvent_t e : prep_ev) (void)hipEventDestroy(e);
        for (auto& op : steps)
            for (StepGraph& g : op) {
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
                if (g.graph) (void)hipGraphDestroy(g.graph);
            }
        if (ev_sync) (void)hipEventDestroy(ev_sync);
        if (desc) (void)hipHostFree(desc);
        if (pin) (void)hipHostFree(pin);
        if (mem) (void)hipFree(mem);
        if (temp) (void)hipFree(temp);
        if (G) dyg_graph_free(G);
        if (H) dyg_graph_free(H);
    }
};

namespace {

// Pinned staging: report words at 0, recovery sizes of a deletion batch at
// 128, then one batch of n (<= max_batch) updates from 256: u, v, w at
// stride n and the outcome codes.
constexpr size_t kPinHead = 256, kPinSizes = 128;
size_t pin_bytes(uint32_t max_batch) { return kPinHead + (size_t)max_batch * 16; }
struct PinLayout {
    uint32_t *u, *v;
    float* w;
    int32_t* cls;
    unsigned long long* words;  // kReportWords
};
PinLayout pin_layout(char* base, uint32_t n) {
    PinLayout l;
    l.words = (unsigned long long*)base;
    char* b = base + kPinHead;
    l.u = (uint32_t*)b;
    l.v = (uint32_t*)(b + n * 4ull);
    l.w = (float*)(b + n * 8ull);
    l.cls = (int32_t*)(b + n * 12ull);
    return l;
}

// Largest batch of the one-workgroup path (0: every batch takes the general
// pipeline).
uint32_t small_limit(const dyg_stream_s* s) {
    if (s->sp.flags & DYG_STREAM_GENERAL_ONLY) return 0;
    const uint32_t top = (s->sp.flags & DYG_STREAM_SMALL_1024) ? kSmallBatch / 2 : kSmallBatch;
    return std::min(top, s->sp.max_batch);
}

int key_bits(uint32_t V) { return std::min(64, 32 + 33 - __builtin_clz(std::max<uint32_t>(V, 1))); }
int row_bits(uint32_t V) { return std::min(32, 33 - __builtin_clz(std::max<uint32_t>(V, 1))); }

dyg_status stream_alloc(dyg_stream s) {
    const uint32_t n = s->sp.max_batch;
    const uint32_t ms = s->sp.del.max_steps;
    const uint32_t hw = s->sp.heur_walkers;
    s->cmax = (uint32_t)std::min<uint64_t>((uint64_t)n * (ms - 1) + 4ull * n * hw, 0x7FFFFFFFull);
    const uint64_t cm = s->cmax;
    // size pass then carve pass
    Arena ar;
    for (int pass = 0; pass < 2; ++pass) {
        ar.off = 0;
        s->in = ar.take<uint32_t>(3 * (size_t)n);
        s->cls = ar.take<int32_t>(n);
        s->k_in = ar.take<unsigned long long>(n);
        s->k_out = ar.take<unsigned long long>(n);
        s->i_in = ar.take<uint32_t>(n);
        s->i_out = ar.take<uint32_t>(n);
        s->flag_g = ar.take<int32_t>(n);
        s->flag_h = ar.take<int32_t>(n);
        s->keep = ar.take<uint8_t>(n);
        s->mode = ar.take<uint8_t>(n);
        s->sel = ar.take<uint32_t>(n);
        s->cnt = ar.take<uint32_t>(8);
        s->cu = ar.take<uint32_t>(n);
        s->cv = ar.take<uint32_t>(n);
        s->cw = ar.take<float>(n);
        s->ccond = ar.take<float>(n);
        s->ceid = ar.take<uint64_t>(n);
        s->cfound = ar.take<int32_t>(n);
        s->path = ar.take<uint32_t>((size_t)n * ms);
        s->pc = ar.take<uint32_t>(n);
        s->poff = ar.take<uint32_t>(n);
        s->nopath = ar.take<uint8_t>(n);
        s->bcnt = ar.take<uint32_t>(((size_t)n + 255) / 256);
        s->jmap = ar.take<uint32_t>(n);
        s->bsum = ar.take<unsigned long long>(((size_t)n + 255) / 256);
        s->ep = ar.take<uint32_t>(2 * (size_t)n);
        s->ekey = ar.take<uint64_t>(2 * (size_t)n);
        s->path3 = ar.take<uint32_t>(2 * (size_t)n * hw * 3);
        s->xu = ar.take<uint32_t>(cm);
        s->xv = ar.take<uint32_t>(cm);
        s->xw = ar.take<float>(cm);
        s->xflag = ar.take<int32_t>(cm);
        // also the rows a small deletion batch's recovery touches (2 per candidate)
        s->xk_in = ar.take<unsigned long long>(std::max<uint64_t>(cm, kSmallCand));
        s->xk_out = ar.take<unsigned long long>(cm);
        s->xi_in = ar.take<uint32_t>(cm);
        s->xi_out = ar.take<uint32_t>(cm);
        s->xr_in = ar.take<uint32_t>(2 * cm);
        s->xr_out = ar.take<uint32_t>(2 * cm);
        s->xe_in = ar.take<uint32_t>(2 * cm);
        s->xe_out = ar.take<uint32_t>(2 * cm);
        s->stats = ar.take<unsigned long long>(4);
        if (pass == 0) {
            ar.cap = ar.off + 256;
            if (hipMalloc(&s->mem, ar.cap) != hipSuccess) {
                s->mem = nullptr;
                set_error(s->ctx, "stream: device allocation failed");
                return DYG_ERR_NOMEM;
            }
            ar.base = (char*)s->mem;
        }
    }
    // hipCUB scratch of the key sorts and of the candidates' row-entry sort
    size_t t = 0, t2 = 0;
    DYG_HIP(s->ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t, s->xk_in, s->xk_out, s->xi_in,
                                                       s->xi_out, (int)std::max<uint64_t>(cm, n),
                                                       0, 64, s->ctx->stream));
    DYG_HIP(s->ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, t2, s->xr_in, s->xr_out, s->xe_in,
                                                       s->xe_out, (int)std::min<uint64_t>(2 * cm, 0x7FFFFFFFull),
                                                       0, 32, s->ctx->stream));
    t = std::max(t, t2);
    if (hipMalloc(&s->temp, t) != hipSuccess) {
        s->temp = nullptr;
        set_error(s->ctx, "stream: device allocation failed");
        return DYG_ERR_NOMEM;
    }
    s->temp_bytes = t;
    if (hipHostMalloc(&s->pin, pin_bytes(n), hipHostMallocDefault) != hipSuccess) {
        s->pin = nullptr;
        set_error(s->ctx, "stream: pinned allocation failed");
        return DYG_ERR_NOMEM;
    }
    void* pd = nullptr;
    DYG_HIP(s->ctx, hipHostGetDevicePointer(&pd, s->pin, 0));
    s->pin_dev = (char*)pd;
    DYG_HIP(s->ctx, hipEventCreateWithFlags(&s->ev_sync, hipEventDisableTiming));
    DYG_HIP(s->ctx, hipMemsetAsync(s->stats, 0, 4 * sizeof(unsigned long long), s->ctx->stream));
    DYG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
    return DYG_OK;
}

// Waits for a recorded event.  Polls it for up to 2 ms before blocking:
// waking from a blocking wait costs tens of microseconds, about what a batch
// of a few hundred updates takes.
hipError_t spin_event(hipEvent_t ev) {
    const double t0 = now_ms();
    for (;;) {
        const hipError_t e = hipEventQuery(ev);
        if (e != hipErrorNotReady) return e;
        if (now_ms() - t0 > 2.0) break;
    }
    return hipEventSynchronize(ev);
}

// Waits for everything enqueued on `st` so far (see spin_event).
dyg_status wait_spin(dyg_ctx ctx, hipStream_t st, hipEvent_t ev) {
    DYG_HIP(ctx, hipEventRecord(ev, st));
    DYG_HIP(ctx, spin_event(ev));
    return DYG_OK;
}


dyg_status sort_keys(dyg_stream s, unsigned long long* kin, unsigned long long* kout,
                     uint32_t* iin, uint32_t* iout, uint32_t n) {
    // keys are (min << 32 | max) or all ones: bits 32 + bits(V) order them
    size_t tb = s->temp_bytes;
    DYG_HIP(s->ctx, hipcub::DeviceRadixSort::SortPairs(s->temp, tb, kin, kout, iin, iout, (int)n,
                                                       0, key_bits(s->G->V), s->ctx->stream));
    return DYG_OK;
}

WalkLaunch walk_args(dyg_stream s, dyg_graph g, const dyg_walk_params& p, uint32_t n) {
    WalkLaunch a{};
    a.rows = g->d_rows;
    a.adj = g->d_adj;
    a.src = s->cu;
    a.dst = s->cv;
    a.w = s->cw;
    a.eid = s->ceid;
    a.n = n;
    a.n_dev = s->cnt;
    a.walkers = p.walkers;
    a.max_steps = p.max_steps;
    a.thr = p.threshold;
    a.flags = p.flags;
    a.edges_per_block = p.block_edges;
    a.seed = p.seed;
    a.found = s->cfound;
    a.cond = s->ccond;
    a.stats = s->stats;
    a.nodesc = g->nodesc;
    return a;
}

#define LAUNCH_1D(kernel, n, ...)                                                              \
    do {                                                                                       \
        const uint32_t _n = (n);                                                               \
        if (_n) hipLaunchKernelGGL(kernel, dim3((_n + 255) / 256), dim3(256), 0, st, __VA_ARGS__); \
    } while (0)

BlockGraph block_graph(dyg_graph g) {
    return BlockGraph{g->d_rows, g->d_adj, g->dev.cap, g->dev.meta,
                      (unsigned long long)g->cap_slots, g->slack, refresh_lanes(g), g->nodesc};
}

// Size classes of the captured one-workgroup steps: a step graph's grids are
// sized for its class maximum, the kernels read the batch's real counts from
// device memory, and only the pre and post kernels' arguments change.
constexpr uint32_t kStepClass[dyg_stream_s::kStepClasses] = {16, 64, 256, 1024, kSmallBatch};

int step_class(uint32_t n) {
    for (int c = 0; c < dyg_stream_s::kStepClasses; ++c)
        if (n <= kStepClass[c]) return c;
    return -1;
}

// The one-workgroup insertion step for a.n updates: pre kernel (and G's
// descriptor refresh), walks over the kept list in H, post kernel (and H's
// refresh).  `timed` brackets the walk with the context's events.
dyg_status enqueue_ins_step(dyg_stream s, const SmallInsArgs& a, bool timed) {
    hipStream_t st = s->ctx->stream;
    DYG_HIP(s->ctx, launch_small_ins_pre(a, st));
    WalkLaunch w = walk_args(s, s->H, s->sp.ins, a.n);
    w.stats = a.stats;
    w.flags |= DYG_WALK_DECISION_ONLY;  // outcomes need found / not found only
    DYG_HIP(s->ctx, launch_walk_incremental(w, st, timed ? s->wev0 : nullptr,
                                            timed ? s->wev1 : nullptr));
    DYG_HIP(s->ctx, launch_small_ins_post(a, st));
    return DYG_OK;
}

// The deletion counterpart: pre kernel (and refreshes), walks over the kept
// list in G with paths, post kernel (recovery and H's refresh).
dyg_status enqueue_del_step(dyg_stream s, const SmallDelArgs& a, bool timed) {
    hipStream_t st = s->ctx->stream;
    DYG_HIP(s->ctx, launch_small_del_pre(a, st));
    WalkLaunch w = walk_args(s, s->G, s->sp.del, a.n);
    w.w = nullptr;
    w.path = s->path;
    w.stats = a.stats;
    DYG_HIP(s->ctx, launch_walk_decremental(w, st, timed ? s->wev0 : nullptr,
                                            timed ? s->wev1 : nullptr));
    DYG_HIP(s->ctx, launch_small_del_post(a, st));
    return DYG_OK;
}

dyg_status enqueue_step(dyg_stream s, const SmallInsArgs& a, bool timed) {
    return enqueue_ins_step(s, a, timed);
}
dyg_status enqueue_step(dyg_stream s, const SmallDelArgs& a, bool timed) {
    return enqueue_del_step(s, a, timed);
}

// Captures the step of one operation and size class: grids for the class
// maximum, the pre and post kernels reading the batch's values from the
// stream's step descriptor, the walk timing events included.  False when the
// runtime refuses any part of it: the class is then marked and launches
// directly, with the same results.
template <class Args>
bool capture_step(dyg_stream s, dyg_stream_s::StepGraph& g, Args a, uint32_t cmax) {
    hipStream_t st = s->ctx->stream;
    g.failed = true;
    if (!s->desc) {
        void* h = nullptr;
        void* d = nullptr;
        if (hipHostMalloc(&h, sizeof(StepDesc), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipHostFree(h);
            return false;
        }
        s->desc = (StepDesc*)h;
        s->desc_dev = (StepDesc*)d;
    }
    a.n = cmax;  // sizes the grids; the kernels take the batch's n from the descriptor
    a.desc = s->desc_dev;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const dyg_status r = enqueue_step(s, a, false);  // no event nodes in the graph
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    (void)hipGetLastError();
    if (r != DYG_OK || e != hipSuccess || !graph ||
        hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        g.exec = nullptr;
        if (graph) (void)hipGraphDestroy(graph);
        return false;
    }
    g.graph = graph;
    g.g = a.G;
    g.h = a.H;
    g.failed = false;
    return true;
}

bool same_layout(const BlockGraph& x, const BlockGraph& y) {
    return x.rows == y.rows && x.adj == y.adj && x.cap == y.cap && x.meta == y.meta &&
           x.cap_slots == y.cap_slots && x.slack == y.slack && x.lpr == y.lpr &&
           x.nodesc == y.nodesc;
}

void fill_desc(StepDesc* d, const SmallInsArgs& a) {
    d->u = a.u;
    d->v = a.v;
    d->w = a.w;
    d->rep_cls = a.rep.cls;
    d->rep_words = a.rep.words;
    d->base = a.base;
    d->n = a.n;
}
void fill_desc(StepDesc* d, const SmallDelArgs& a) {
    d->u = a.u;
    d->v = a.v;
    d->w = nullptr;
    d->rep_cls = a.rep.cls;
    d->rep_words = a.rep.words;
    d->base = a.base;
    d->n = a.n;
}

// Runs a one-workgroup step through its captured graph: the batch's values
// go into the step descriptor, then one graph launch.  Only for callers that
// wait for each batch before the next (the descriptor must not change under
// a launch still in flight).  Returns false (nothing enqueued) when the
// class has no graph.
template <class Args>
bool run_step_graph(dyg_stream s, bool insert, const Args& a, dyg_status* out) {
    const int c = step_class(a.n);
    if (c < 0) return false;
    dyg_stream_s::StepGraph& g = s->steps[insert ? 0 : 1][c];
    if (g.exec && !(same_layout(g.g, a.G) && same_layout(g.h, a.H))) {
        // the arena grew (or the edit buffers or refresh lanes changed) since
        // the capture: the nodes hold the old addresses, so capture again
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
        g.graph = nullptr;
    }
    if (!g.exec && (g.failed || !capture_step(s, g, a, kStepClass[c]))) return false;
    fill_desc(s->desc, a);
    // the walk is a node inside the graph: the timing events bracket the step
    hipError_t e = hipEventRecord(s->wev0, s->ctx->stream);
    if (e == hipSuccess) e = hipGraphLaunch(g.exec, s->ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(s->wev1, s->ctx->stream);
    if (e != hipSuccess) {
        set_error(s->ctx, std::string("stream step graph: ") + hipGetErrorString(e));
        *out = DYG_ERR_HIP;
    } else {
        *out = DYG_OK;
    }
    return true;
}

SmallInsArgs small_ins_args(dyg_stream s, const uint32_t* du, const uint32_t* dv,
                            const float* dw, uint32_t n, int32_t* dcls,
                            unsigned long long* dstats, const Report& rep) {
    SmallInsArgs a{};
    a.G = block_graph(s->G);
    a.H = block_graph(s->H);
    a.u = du;
    a.v = dv;
    a.w = dw;
    a.n = n;
    a.base = s->base;
    a.key_bits = key_bits(s->G->V);
    a.row_bits = row_bits(s->G->V);
    a.cls = dcls;
    a.sel = s->sel;
    a.cnt = s->cnt;
    a.cu = s->cu;
    a.cv = s->cv;
    a.cw = s->cw;
    a.ceid = s->ceid;
    a.found = s->cfound;
    a.touched = s->ep;  // 2 * max_batch entries, free during insertion batches
    a.stats = dstats;
    a.rep = rep;
    a.desc = nullptr;
    return a;
}

// Small insertion batch: two one-workgroup kernels around the walk (as a
// captured graph when `graph`).
dyg_status apply_insert_small(dyg_stream s, const uint32_t* du, const uint32_t* dv,
                              const float* dw, uint32_t n, int32_t* dcls,
                              unsigned long long* dstats, bool timed, const Report& rep,
                              bool graph, uint32_t vcheck, const BatchOrder* ord) {
    SmallInsArgs a = small_ins_args(s, du, dv, dw, n, dcls, dstats, rep);
    a.vcheck = vcheck;
    if (ord) a.ord = *ord;
    dyg_status r;
    if (graph && timed && run_step_graph(s, true, a, &r)) {
        s->last_graph = true;
        return r;
    }
    return enqueue_ins_step(s, a, timed);
}

// Both apply_* enqueue a whole batch (no wait except a deletion batch's one
// read-back of its recovery sizes) ending with the report `rep`.
// `ord` (batches of a list): the batch's orders, computed ahead of it on the
// prep stream; its own kernels then sort neither its keys nor its row
// entries.
dyg_status apply_insert(dyg_stream s, const uint32_t* du, const uint32_t* dv, const float* dw,
                        uint32_t n, int32_t* dcls, unsigned long long* dstats, bool timed,
                        dyg_stream_stats* out, const Report& rep, uint32_t* stage,
                        bool graph = false, uint32_t vcheck = 0,
                        const BatchOrder* ord = nullptr) {
    if (n <= small_limit(s))
        return apply_insert_small(s, du, dv, dw, n, dcls, dstats, timed, rep, graph, vcheck, ord);
    hipStream_t st = s->ctx->stream;
    dyg_graph G = s->G, H = s->H;
    const unsigned long long* key = ord ? ord->key : s->k_out;
    const uint32_t* idx = ord ? ord->idx : s->i_out;
    uint32_t* bcnt = ord ? ord->bcnt : s->bcnt;
    dyg_status r = DYG_OK;
    if (!ord) {
        if (stage) {
            LAUNCH_1D(k_batch_keys_staged, n, du, dv, dw, n, s->k_in, s->i_in, s->bcnt, stage,
                      stage + n, (float*)(stage + 2ull * n));
            du = stage;
            dv = stage + n;
            dw = (const float*)(stage + 2ull * n);
        } else {
            LAUNCH_1D(k_batch_keys, n, du, dv, n, s->k_in, s->i_in, s->bcnt);
        }
        r = sort_keys(s, s->k_in, s->k_out, s->i_in, s->i_out, n);
        if (r != DYG_OK) return r;
    }
    LAUNCH_1D(k_ins_groups, n, key, idx, n, G->d_rows, G->d_adj, H->d_rows, H->d_adj, dcls,
              s->flag_g, s->keep, s->mode, bcnt);
    r = ord ? device_edits_sorted(G, du, dv, dw, ord->row, ord->ent, n,
                                  EditTake{s->flag_g, -1, nullptr, nullptr, nullptr})
            : device_edits_batch(G, du, dv, dw, s->flag_g, n);
    if (r != DYG_OK) return r;
    LAUNCH_1D(k_compact, n, s->keep, n, bcnt, du, dv, dw, s->base, false, s->cu, s->cv, s->cw,
              s->ceid, s->sel, s->cnt, s->jmap,
              InsWeights{key, idx, s->mode, dw, G->d_rows, G->d_adj, H->d_rows, H->d_adj});
    DYG_HIP(s->ctx, hipGetLastError());
    WalkLaunch a = walk_args(s, H, s->sp.ins, n);
    a.stats = dstats;
    a.flags |= DYG_WALK_DECISION_ONLY;  // outcomes need found / not found only
    DYG_HIP(s->ctx, launch_walk_incremental(a, st, timed ? s->wev0 : nullptr,
                                            timed ? s->wev1 : nullptr));
    // the kept edges that found no path, appended to H (from the batch's own
    // arrays and presorted entries when ordered ahead)
    r = ord ? device_edits_sorted(H, du, dv, dw, ord->row, ord->ent, n,
                                  EditTake{nullptr, 0, s->keep, s->jmap, s->cfound})
            : device_edits_batch(H, s->cu, s->cv, s->cw, s->cfound, n, s->cnt);
    if (r != DYG_OK) return r;
    hipLaunchKernelGGL(k_outcome_report, dim3((n + 255) / 256), dim3(256), 0, st, rep, dcls,
                       s->keep, s->jmap, s->cfound, n, kOutFound, kOutAdded, dstats);
    (void)out;
    return DYG_OK;
}

// Recovery of a deletion batch once the walks' candidate list sizes are in
// `hc` (kept, no-path, endpoints, path hops) and the path hops are in xu/xv:
// heuristic hops, first occurrences H lacks, appended with G's weight.
// kept_map: keep[] and jmap[] describe the batch (the general pipeline's
// k_compact wrote them), so the outcome codes go out with the report.
dyg_status recover_general(dyg_stream s, uint32_t n, const uint32_t* hc, int32_t* dcls,
                           unsigned long long* dstats, dyg_stream_stats* out, const Report& rep, bool kept_map) {
    hipStream_t st = s->ctx->stream;
    dyg_graph G = s->G, H = s->H;
    const uint32_t hw = s->sp.heur_walkers;
    const uint32_t npath = hc[3], n3 = hc[2] * hw;
    const uint32_t ncand = npath + 2 * n3;
    if (ncand > s->cmax) {
        set_error(s->ctx, "stream: recovery list exceeds its reservation");
        return DYG_ERR_STATE;
    }
    LAUNCH_1D(k_cand_heur, n3, s->path3, n3, npath, s->xu, s->xv);
    // the candidates' row entries, sorted once: each candidate is decided in
    // its row's run (first occurrence of its edge, H lacking it), and the
    // appends to H take the same presorted entries
    LAUNCH_1D(k_cand_entries, ncand, s->xu, s->xv, ncand, s->xr_in, s->xe_in);
    if (ncand) {
        size_t tb = s->temp_bytes;
        DYG_HIP(s->ctx, hipcub::DeviceRadixSort::SortPairs(s->temp, tb, s->xr_in, s->xr_out, s->xe_in,
                                                           s->xe_out, (int)(2 * ncand), 0,
                                                           row_bits(G->V), st));
    }
    LAUNCH_1D(k_cand_decide, 2 * ncand, s->xr_out, s->xe_out, 2 * ncand, s->xu, s->xv, G->d_rows,
              G->d_adj, H->d_rows, H->d_adj, s->xflag, s->xw);
    dyg_status r = device_edits_sorted(H, s->xu, s->xv, s->xw, s->xr_out, s->xe_out, ncand,
                                       EditTake{s->xflag, -1, nullptr, nullptr, nullptr});
    if (r != DYG_OK) return r;
    if (kept_map) {
        hipLaunchKernelGGL(k_outcome_report, dim3((n + 255) / 256), dim3(256), 0, st, rep, dcls,
                           s->keep, s->jmap, s->cfound, n, kOutRecovered, kOutHeuristic, dstats);
    } else {
        LAUNCH_1D(k_outcome, n, s->sel, s->cnt, s->cfound, dcls, kOutRecovered, kOutHeuristic);
        DYG_HIP(s->ctx, launch_report(rep, dcls, n, dstats, st));
    }
    if (out) out->candidates = ncand;
    return DYG_OK;
}

SmallDelArgs small_del_args(dyg_stream s, const uint32_t* du, const uint32_t* dv, uint32_t n,
                            int32_t* dcls, unsigned long long* dstats, const Report& rep) {
    SmallDelArgs a{};
    a.G = block_graph(s->G);
    a.H = block_graph(s->H);
    a.u = du;
    a.v = dv;
    a.n = n;
    a.base = s->base;
    a.key_bits = key_bits(s->G->V);
    a.row_bits = row_bits(s->G->V);
    a.cls = dcls;
    a.sel = s->sel;
    a.cnt = s->cnt;
    a.touched_g = s->ep;                  // 2 * max_batch entries
    a.touched_h = (uint32_t*)s->ekey;     // 4 * max_batch entries
    a.touched_x = (uint32_t*)s->xk_in;    // 2 * cmax entries >= 2 * candidates
    a.cu = s->cu;
    a.cv = s->cv;
    a.ceid = s->ceid;
    a.found = s->cfound;
    a.path = s->path;
    a.max_steps = s->sp.del.max_steps;
    a.hw = s->sp.heur_walkers;
    a.seed = s->sp.del.seed;
    a.pc = s->pc;
    a.poff = s->poff;
    a.ep = s->ep;
    a.ekey = s->ekey;
    a.sizes = (uint32_t*)(s->pin_dev + kPinSizes);
    a.stats = dstats;
    a.rep = rep;
    a.desc = nullptr;
    return a;
}

// Captures every operation's one-workgroup step in every size class the
// stream can use (DYG_STREAM_GRAPH), at creation: a replay then costs one
// graph launch, and only a layout change (arena growth) captures again.
void capture_all_steps(dyg_stream s) {
    const Report rep{nullptr, nullptr, s->G->dev.meta, s->H->dev.meta};
    const uint32_t top = small_limit(s);
    for (int c = 0; c < dyg_stream_s::kStepClasses; ++c) {
        if (c > 0 && kStepClass[c - 1] >= top) break;  // no batch reaches this class
        const uint32_t cm = kStepClass[c];
        (void)capture_step(s, s->steps[0][c],
                           small_ins_args(s, nullptr, nullptr, nullptr, cm, s->cls, s->stats, rep),
                           cm);
        (void)capture_step(s, s->steps[1][c],
                           small_del_args(s, nullptr, nullptr, cm, s->cls, s->stats, rep), cm);
    }
}

// Small deletion batch: one-workgroup kernels around the walk (as a captured
// graph when `graph`); waits only when the recovery could outgrow the post
// kernel (then reads its flag).
dyg_status apply_delete_small(dyg_stream s, const uint32_t* du, const uint32_t* dv, uint32_t n,
                              int32_t* dcls, unsigned long long* dstats, bool timed,
                              dyg_stream_stats* out, const Report& rep, bool graph,
                              uint32_t vcheck, const BatchOrder* ord) {
    hipStream_t st = s->ctx->stream;
    const uint32_t ms = s->sp.del.max_steps, hw = s->sp.heur_walkers;
    SmallDelArgs a = small_del_args(s, du, dv, n, dcls, dstats, rep);
    a.vcheck = vcheck;
    if (ord) a.ord = *ord;
    dyg_status r = DYG_OK;
    if (graph && timed && run_step_graph(s, false, a, &r))
        s->last_graph = true;
    else
        r = enqueue_del_step(s, a, timed);
    if (r != DYG_OK) return r;
    if ((uint64_t)n * (ms - 1) + 4ull * n * hw <= small_del_candidates(n)) return DYG_OK;
    r = wait_spin(s->ctx, st, s->ev_sync);
    if (r != DYG_OK) return r;
    const uint32_t* hc = (const uint32_t*)((char*)s->pin + kPinSizes);
    // rejected by the pre kernel's id check (device inputs only; the word is
    // the stream's own, cleared before this batch): nothing to do
    if (vcheck && ((const unsigned long long*)s->pin)[kReportBad]) return DYG_OK;
    if (hc[4] == 0) return DYG_OK;  // the post kernel finished the batch
    DYG_HIP(s->ctx, launch_heuristic(s->G->d_rows, s->G->d_adj, s->ep, s->ekey, 2 * n, hw,
                                     s->sp.del.seed, s->path3, st, s->cnt + 2, s->G->nodesc));
    LAUNCH_1D(k_cand_paths, n, s->cfound, s->cnt, s->path, ms, s->poff, s->xu, s->xv);
    DYG_HIP(s->ctx, hipGetLastError());
    return recover_general(s, n, hc, dcls, dstats, out, rep, false);
}

dyg_status apply_delete(dyg_stream s, const uint32_t* du, const uint32_t* dv, uint32_t n,
                        int32_t* dcls, unsigned long long* dstats, bool timed,
                        dyg_stream_stats* out, const Report& rep, uint32_t* stage,
                        bool graph = false, uint32_t vcheck = 0,
                        const BatchOrder* ord = nullptr) {
    if (n <= small_limit(s))
        return apply_delete_small(s, du, dv, n, dcls, dstats, timed, out, rep, graph, vcheck,
                                  ord);
    hipStream_t st = s->ctx->stream;
    dyg_graph G = s->G, H = s->H;
    const uint32_t ms = s->sp.del.max_steps, hw = s->sp.heur_walkers;
    const unsigned long long* key = ord ? ord->key : s->k_out;
    const uint32_t* idx = ord ? ord->idx : s->i_out;
    uint32_t* bcnt = ord ? ord->bcnt : s->bcnt;
    dyg_status r = DYG_OK;
    if (!ord) {
        if (stage) {
            LAUNCH_1D(k_batch_keys_staged, n, du, dv, (const float*)nullptr, n, s->k_in, s->i_in,
                      s->bcnt, stage, stage + n, (float*)nullptr);
            du = stage;
            dv = stage + n;
        } else {
            LAUNCH_1D(k_batch_keys, n, du, dv, n, s->k_in, s->i_in, s->bcnt);
        }
        r = sort_keys(s, s->k_in, s->k_out, s->i_in, s->i_out, n);
        if (r != DYG_OK) return r;
    }
    LAUNCH_1D(k_del_groups, n, key, idx, n, G->d_rows, G->d_adj, H->d_rows, H->d_adj, dcls,
              s->flag_g, s->flag_h, s->keep, bcnt);
    r = ord ? device_deletes_sorted(G, du, dv, ord->row, ord->ent, n,
                                    EditTake{s->flag_g, 1, nullptr, nullptr, nullptr})
            : device_deletes_batch(G, du, dv, s->flag_g, n, nullptr);
    if (r != DYG_OK) return r;
    r = ord ? device_deletes_sorted(H, du, dv, ord->row, ord->ent, n,
                                    EditTake{s->flag_h, 1, nullptr, nullptr, nullptr})
            : device_deletes_batch(H, du, dv, s->flag_h, n, nullptr);
    if (r != DYG_OK) return r;
    LAUNCH_1D(k_compact, n, s->keep, n, bcnt, du, dv, (const float*)nullptr, s->base, true,
              s->cu, s->cv, (float*)nullptr, s->ceid, s->sel, s->cnt, s->jmap, InsWeights{});
    DYG_HIP(s->ctx, hipGetLastError());
    WalkLaunch a = walk_args(s, G, s->sp.del, n);
    a.w = nullptr;
    a.path = s->path;
    a.stats = dstats;
    DYG_HIP(s->ctx, launch_walk_decremental(a, st, timed ? s->wev0 : nullptr,
                                            timed ? s->wev1 : nullptr));
    LAUNCH_1D(k_path_counts, n, s->cfound, s->cnt, n, s->pc, s->nopath, s->bsum);
    LAUNCH_1D(k_recovery_lists, n, s->pc, s->nopath, n, s->bsum, s->cu, s->cv, s->ceid, s->poff,
              s->ep, s->ekey, s->cnt, (uint32_t*)(s->pin_dev + kPinSizes));
    // the one wait of a deletion batch is for the recovery list sizes, which
    // this kernel writes: the host waits for it (not for the heuristic walks
    // and path candidates enqueued behind it) and enqueues the recovery
    // while those run
    DYG_HIP(s->ctx, hipEventRecord(s->ev_sync, st));
    const uint32_t* hc = (const uint32_t*)((char*)s->pin + kPinSizes);
    DYG_HIP(s->ctx, launch_heuristic(G->d_rows, G->d_adj, s->ep, s->ekey, 2 * n, hw,
                                     s->sp.del.seed, s->path3, st, s->cnt + 2, G->nodesc));
    LAUNCH_1D(k_cand_paths, n, s->cfound, s->cnt, s->path, ms, s->poff, s->xu, s->xv);
    DYG_HIP(s->ctx, hipGetLastError());
    DYG_HIP(s->ctx, spin_event(s->ev_sync));
    return recover_general(s, n, hc, dcls, dstats, out, rep, true);
}

// ---------------------------------------------------------------------------
// Plain form: the reference's shape.  The host keeps both edge maps, decides
// every update, plans the row edits (shipped to the device as deltas) and
// launches one walk per batch with one workgroup per edge, waiting for each
// result before going on.  Same semantics and results as the device form.
// ---------------------------------------------------------------------------
dyg_status plain_insert(dyg_stream s, const uint32_t* u, const uint32_t* v, const float* w,
                        uint32_t n, int32_t* cls, dyg_stream_stats* rs) {
    dyg_graph G = s->G, H = s->H;
    std::vector<uint32_t> nu, nv, ku, kv, kidx, gu, gv, hu, hv;
    std::vector<float> nw, kw, gw, hw;
    std::vector<uint64_t> keid;
    std::unordered_set<long> fresh;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        if (a == b) {
            cls[i] = kOutInvalid;
            continue;
        }
        const long key = key_of(a, b, G->mult);
        if (G->key_map.count(key) || fresh.count(key)) {
            cls[i] = kOutWeight;
            gu.push_back(a); gv.push_back(b); gw.push_back(w[i]);
            if (H->key_map.count(key)) {
                hu.push_back(a); hv.push_back(b); hw.push_back(w[i]);
            }
        } else {
            fresh.insert(key);
            nu.push_back(u[i]); nv.push_back(v[i]); nw.push_back(w[i]);
            ku.push_back(u[i]); kv.push_back(v[i]); kw.push_back(w[i]);
            keid.push_back(s->base + i);
            kidx.push_back(i);
        }
    }
    dyg_status r = graph_apply_inserts(G, nu.data(), nv.data(), nw.data(), (uint32_t)nu.size());
    if (r == DYG_OK) r = graph_add_weights(G, gu.data(), gv.data(), gw.data(), (uint32_t)gu.size());
    if (r == DYG_OK) r = graph_add_weights(H, hu.data(), hv.data(), hw.data(), (uint32_t)hu.size());
    if (r != DYG_OK) return r;
    const uint32_t nk = (uint32_t)ku.size();
    std::vector<int32_t> found(nk);
    std::vector<float> cond(nk);
    dyg_walk_params p = s->sp.ins;
    p.flags |= DYG_WALK_PER_EDGE;
    dyg_walk_stats ws{};
    if (nk) {
        r = dyg_walk_incremental(H, ku.data(), kv.data(), kw.data(), keid.data(), nk, &p,
                                 found.data(), cond.data(), nullptr, nullptr, &ws);
        if (r != DYG_OK) return r;
    }
    rs->walk_ms += ws.kernel_ms;
    rs->hops += ws.hops;
    std::vector<uint32_t> au, av;
    std::vector<float> aw;
    for (uint32_t j = 0; j < nk; ++j) {
        cls[kidx[j]] = found[j] == -1 ? kOutAdded : kOutFound;
        if (found[j] == -1) {
            au.push_back(ku[j]); av.push_back(kv[j]); aw.push_back(kw[j]);
        }
    }
    return graph_apply_inserts(H, au.data(), av.data(), aw.data(), (uint32_t)au.size());
}

dyg_status plain_delete(dyg_stream s, const uint32_t* u, const uint32_t* v, uint32_t n,
                        int32_t* cls, dyg_stream_stats* rs) {
    dyg_graph G = s->G, H = s->H;
    std::vector<uint32_t> gu, gv, hu, hv, ka, kb, kidx;
    std::vector<uint64_t> keid;
    std::unordered_set<long> gone;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t a = std::min(u[i], v[i]), b = std::max(u[i], v[i]);
        if (a == b) {
            cls[i] = kOutInvalid;
            continue;
        }
        const long key = key_of(a, b, G->mult);
        if (!G->key_map.count(key) || gone.count(key)) {
            cls[i] = kOutSkip;
            continue;
        }
        gone.insert(key);
        gu.push_back(u[i]); gv.push_back(v[i]);
        if (H->key_map.count(key)) {
            hu.push_back(u[i]); hv.push_back(v[i]);
            ka.push_back(a); kb.push_back(b);
            keid.push_back(s->base + i);
            kidx.push_back(i);
        } else {
            cls[i] = kOutGOnly;
        }
    }
    dyg_status r = graph_apply_deletes(G, gu.data(), gv.data(), (uint32_t)gu.size(), nullptr, true);
    if (r == DYG_OK)
        r = graph_apply_deletes(H, hu.data(), hv.data(), (uint32_t)hu.size(), nullptr, true);
    if (r != DYG_OK) return r;
    const uint32_t nk = (uint32_t)ka.size();
    if (nk == 0) return DYG_OK;
    const uint32_t ms = s->sp.del.max_steps, hwk = s->sp.heur_walkers;
    std::vector<int32_t> found(nk);
    std::vector<float> cond(nk);
    std::vector<uint32_t> path((size_t)nk * ms);
    dyg_walk_params p = s->sp.del;
    p.flags |= DYG_WALK_PER_EDGE;
    dyg_walk_stats ws{};
    r = dyg_walk_decremental(G, ka.data(), kb.data(), keid.data(), nk, &p, found.data(),
                             cond.data(), nullptr, nullptr, path.data(), &ws);
    if (r != DYG_OK) return r;
    rs->walk_ms += ws.kernel_ms;
    rs->hops += ws.hops;
    std::vector<uint32_t> xu, xv, ep;
    std::vector<uint64_t> ek;
    for (uint32_t j = 0; j < nk; ++j) {
        cls[kidx[j]] = found[j] > 0 ? kOutRecovered : kOutHeuristic;
        if (found[j] > 0) {
            for (int q = 1; q <= found[j]; ++q) {
                xu.push_back(path[(size_t)j * ms + q - 1]);
                xv.push_back(path[(size_t)j * ms + q]);
            }
        } else {
            ep.push_back(ka[j]); ek.push_back(keid[j] << 1);
            ep.push_back(kb[j]); ek.push_back((keid[j] << 1) | 1ull);
        }
    }
    if (!ep.empty()) {
        std::vector<uint32_t> p3(ep.size() * hwk * 3);
        r = dyg_walk_heuristic(G, ep.data(), ek.data(), (uint32_t)ep.size(), hwk, s->sp.del.seed,
                               p3.data());
        if (r != DYG_OK) return r;
        for (size_t g = 0; g < ep.size() * hwk; ++g) {
            xu.push_back(p3[g * 3]); xv.push_back(p3[g * 3 + 1]);
            xu.push_back(p3[g * 3 + 1]); xv.push_back(p3[g * 3 + 2]);
        }
    }
    rs->candidates = xu.size();
    std::vector<uint32_t> au, av;
    std::vector<float> aw;
    std::unordered_set<long> added;
    for (size_t c = 0; c < xu.size(); ++c) {
        const uint32_t a = std::min(xu[c], xv[c]), b = std::max(xu[c], xv[c]);
        if (a == b) continue;
        const long key = key_of(a, b, G->mult);
        if (H->key_map.count(key) || added.count(key)) continue;
        auto it = G->key_map.find(key);
        if (it == G->key_map.end()) continue;  // unreachable: hops are edges of G
        added.insert(key);
        float x;
        std::memcpy(&x, &G->adj[G->start[a] + it->second.first].y, 4);
        au.push_back(xu[c]); av.push_back(xv[c]); aw.push_back(x);
    }
    return graph_apply_inserts(H, au.data(), av.data(), aw.data(), (uint32_t)au.size());
}

}  // namespace

namespace {
size_t order_bytes(uint64_t n);
dyg_status prep_reserve(dyg_stream s, size_t bytes, size_t events, size_t jobs);
}  // namespace

extern "C" dyg_status dyg_stream_create(dyg_ctx ctx, uint32_t V, const uint64_t* g_row_ptr,
                                        const uint32_t* g_col, const float* g_w,
                                        const uint64_t* h_row_ptr, const uint32_t* h_col,
                                        const float* h_w, const dyg_stream_params* sp,
                                        dyg_stream* out) {
    if (!ctx || !sp || !out || V == 0) return DYG_ERR_ARG;
    *out = nullptr;
    dyg_status r = check_params(ctx, &sp->ins);
    if (r == DYG_OK) r = check_params(ctx, &sp->del);
    if (r != DYG_OK) return r;
    if (sp->max_batch == 0 || sp->heur_walkers == 0 || sp->heur_walkers > 1024) {
        set_error(ctx, "stream: max_batch must be positive, heur_walkers in [1,1024]");
        return DYG_ERR_ARG;
    }
    auto* s = new dyg_stream_s;
    s->ctx = ctx;
    s->wev0 = ctx->ev0;
    s->wev1 = ctx->ev1;
    s->sp = *sp;
    const bool plain = (sp->flags & DYG_STREAM_PLAIN) != 0;
    r = dyg_graph_upload(ctx, V, g_row_ptr, g_col, g_w, sp->slack, plain, &s->G);
    if (r == DYG_OK) r = dyg_graph_upload(ctx, V, h_row_ptr, h_col, h_w, sp->slack, plain, &s->H);
    if (!plain) {
        if (r == DYG_OK && (sp->flags & DYG_STREAM_DENSE_NO_DESC)) s->G->nodesc = true;
        if (r == DYG_OK) r = stream_alloc(s);
        // arena room for the run's expected appends, and at least what each
        // call's closing reservation asks for (a batch for G, a deletion
        // batch's recovery candidates for H): calls then do not reallocate
        if (r == DYG_OK)
            r = device_edits_begin(s->G, sp->max_batch,
                                   std::max<uint64_t>(sp->max_inserts, sp->max_batch));
        if (r == DYG_OK)
            r = device_edits_begin(s->H, s->cmax, std::max<uint64_t>(sp->max_inserts, s->cmax));
        if (r == DYG_OK && (sp->flags & DYG_STREAM_GRAPH) && !(sp->flags & DYG_STREAM_GENERAL_ONLY))
            capture_all_steps(s);
        // lists compute their batches' orders on the prep stream: its
        // buffers for lists of up to max_inserts updates (at most 4M; larger
        // lists grow them), 64 larger batches and 256 one-workgroup batches,
        // so that a list call allocates nothing
        const uint64_t upd = std::min<uint64_t>(
            std::max<uint64_t>(sp->max_inserts, sp->max_batch), 4ull << 20);
        if (r == DYG_OK) r = prep_reserve(s, order_bytes(upd) + 64 * 256, 65, 256);
        if (r == DYG_OK) {
            s->helper.reset(new Helper(ctx->device));
            if (!s->helper->start()) {
                set_error(ctx, "stream: helper thread could not start");
                r = DYG_ERR_STATE;
            }
        }
    }
    if (r != DYG_OK) {
        delete s;
        return r;
    }
    *out = s;
    return DYG_OK;
}

namespace {

// Outcome codes from pinned memory to the caller (when `out`) and their counts
// by class: one bulk copy, then counts over the copy in four independent
// tables (a single table serialises on repeated codes).
void count_codes(const int32_t* codes, int32_t* out, uint32_t n, uint64_t counts[16]) {
    if (out) {
        std::memcpy(out, codes, (size_t)n * 4);
        codes = out;
    }
    uint32_t t[4][16] = {};
    uint32_t i = 0;
    for (; i + 4 <= n; i += 4) {
        ++t[0][codes[i] & 15];
        ++t[1][codes[i + 1] & 15];
        ++t[2][codes[i + 2] & 15];
        ++t[3][codes[i + 3] & 15];
    }
    for (; i < n; ++i) ++t[0][codes[i] & 15];
    for (int c = 0; c < 16; ++c) counts[c] += (uint64_t)t[0][c] + t[1][c] + t[2][c] + t[3][c];
}

// One batch through the stream; u, v, w are host arrays, or device pointers
// when `dev` (then the device checks the ids before anything changes).
dyg_status stream_apply(dyg_stream s, int op, const uint32_t* u, const uint32_t* v,
                        const float* w, uint32_t n, int32_t* outcome, dyg_stream_stats* st,
                        bool dev) {
    if (!s) return DYG_ERR_ARG;
    dyg_ctx ctx = s->ctx;
    if ((op != DYG_STREAM_INSERT && op != DYG_STREAM_DELETE) || (n && (!u || !v)) ||
        (n && op == DYG_STREAM_INSERT && !w)) {
        set_error(ctx, "stream apply: bad operation or missing array");
        return DYG_ERR_ARG;
    }
    if (n > s->sp.max_batch) {
        set_error(ctx, "stream apply: batch larger than max_batch");
        return DYG_ERR_ARG;
    }
    const bool plain = (s->sp.flags & DYG_STREAM_PLAIN) != 0;
    if (dev && plain) {
        // the plain form plans on the host: bring the batch there and go on as usual
        std::vector<uint32_t> hu(n), hv(n);
        std::vector<float> hw(op == DYG_STREAM_INSERT ? n : 0);
        if (n) {
            DYG_HIP(ctx, hipMemcpy(hu.data(), u, n * 4ull, hipMemcpyDeviceToHost));
            DYG_HIP(ctx, hipMemcpy(hv.data(), v, n * 4ull, hipMemcpyDeviceToHost));
            if (op == DYG_STREAM_INSERT)
                DYG_HIP(ctx, hipMemcpy(hw.data(), w, n * 4ull, hipMemcpyDeviceToHost));
        }
        return stream_apply(s, op, hu.data(), hv.data(), hw.empty() ? nullptr : hw.data(), n,
                            outcome, st, false);
    }
    const double t0 = now_ms();
    dyg_stream_stats rs{};
    rs.updates = n;
    if (n && !dev) {
        // range check, fused with the copy into pinned staging (device form)
        uint32_t mx = 0;
        if (plain) {
            for (uint32_t i = 0; i < n; ++i) mx = std::max(mx, std::max(u[i], v[i]));
        } else {
            const PinLayout hp = pin_layout((char*)s->pin, n);
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t a = u[i], b = v[i];
                hp.u[i] = a;
                hp.v[i] = b;
                mx = std::max(mx, std::max(a, b));
            }
        }
        if (mx >= s->G->V) {
            set_error(ctx, "stream apply: vertex id out of range");
            return DYG_ERR_ARG;
        }
    }
    uint64_t counts[16] = {};
    if (n && plain) {
        s->h_cls.assign(n, 0);
        dyg_status r = op == DYG_STREAM_INSERT ? plain_insert(s, u, v, w, n, s->h_cls.data(), &rs)
                                               : plain_delete(s, u, v, n, s->h_cls.data(), &rs);
        if (r != DYG_OK) return r;
        DYG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else if (n) {
        // host inputs: the device reads them from pinned staging (a
        // one-workgroup batch in place, a larger one through its first
        // kernel, which stages them in device memory); device inputs are
        // read where they are.  The last kernel writes the results back into
        // pinned memory, and the batch has one wait.
        hipStream_t str = ctx->stream;
        const PinLayout hp = pin_layout((char*)s->pin, n), dp = pin_layout(s->pin_dev, n);
        const bool ins = op == DYG_STREAM_INSERT;
        const bool small =
            n <= small_limit(s);
        hp.words[kReportBad] = 0;
        if (dev && !small) {
            // a larger batch of device inputs: its ids are checked before it starts
            hipLaunchKernelGGL(k_check_batch, dim3((n + 255) / 256), dim3(256), 0, str, u, v, n,
                               s->G->V, dp.words + kReportBad);
            dyg_status r = wait_spin(ctx, str, s->ev_sync);
            if (r != DYG_OK) return r;
            if (hp.words[kReportBad]) {
                set_error(ctx, "stream apply: vertex id out of range");
                return DYG_ERR_ARG;
            }
        }
        if (ins && !dev) std::memcpy(hp.w, w, n * 4ull);
        const Report rep{dp.cls, dp.words, s->G->dev.meta, s->H->dev.meta};
        // step graphs are safe here: this call waits for its batch before
        // returning (host inputs only: a graph's kernels take no id check)
        const bool graph = (s->sp.flags & DYG_STREAM_GRAPH) != 0 && !dev;
        const uint32_t vcheck = dev ? s->G->V : 0u;
        const uint32_t* bu = dev ? u : dp.u;
        const uint32_t* bv = dev ? v : dp.v;
        const float* bw = dev ? w : dp.w;
        uint32_t* stage = dev ? nullptr : s->in;
        s->last_graph = false;
        dyg_status r = ins ? apply_insert(s, bu, bv, bw, n, s->cls, s->stats, true, &rs, rep,
                                          stage, graph, vcheck)
                           : apply_delete(s, bu, bv, n, s->cls, s->stats, true, &rs, rep, stage,
                                          graph, vcheck);
        if (r != DYG_OK) return r;
        r = wait_spin(ctx, str, s->ev_sync);
        if (r != DYG_OK) return r;
        if (hp.words[kReportBad]) {  // rejected by its pre kernel: nothing changed
            set_error(ctx, "stream apply: vertex id out of range");
            return DYG_ERR_ARG;
        }
        const unsigned long long* hst = hp.words;
        r = device_edits_absorb(s->G, hp.words + 4, nullptr);
        if (r == DYG_OK) r = device_edits_absorb(s->H, hp.words + 9, nullptr);
        if (r != DYG_OK) return r;
        count_codes(hp.cls, outcome, n, counts);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, s->wev0, s->wev1) == hipSuccess) rs.walk_ms = ms;
        rs.hops = hst[0];
        rs.redirects = hst[1];
        if (hst[3]) rs.candidates = hst[3];  // counted on the device (small deletion batch)
        rs.graph = s->last_graph ? 1 : 0;
        // keep a batch's worth of arena headroom for the next batch
        r = device_edits_prepare(s->G, s->sp.max_batch, s->sp.max_batch);
        if (r == DYG_OK) r = device_edits_prepare(s->H, s->cmax, s->cmax);
        if (r != DYG_OK) return r;
    }
    if (n && plain) {
        for (int32_t c : s->h_cls) counts[c & 15]++;
        if (outcome) std::memcpy(outcome, s->h_cls.data(), n * 4ull);
    }
    rs.weight_updates = counts[kOutWeight];
    rs.found = counts[kOutFound];
    rs.added = counts[kOutAdded];
    rs.skipped = counts[kOutSkip];
    rs.g_only = counts[kOutGOnly];
    rs.recovered = counts[kOutRecovered];
    rs.heuristic = counts[kOutHeuristic];
    rs.invalid = n - (rs.weight_updates + rs.found + rs.added + rs.skipped + rs.g_only +
                      rs.recovered + rs.heuristic);
    s->base += n;
    rs.total_ms = now_ms() - t0;
    if (st) *st = rs;
    return DYG_OK;
}

// A batch of n updates takes the general pipeline (and an order computed
// ahead of it in a list) when it is larger than a one-workgroup batch.
bool general_batch(dyg_stream s, uint64_t n) {
    return n > small_limit(s);
}

// Bytes of one batch's BatchOrder arrays (each batch's region 256-aligned).
size_t order_bytes(uint64_t n) {
    const size_t b = 8 * n + 4 * n + 8 * n + 8 * n + 4 * ((n + 255) / 256);
    return (b + 255) & ~(size_t)255;
}

// The prep stream, one batch's sort inputs and scratch, `bytes` of orders,
// `events` events and room for `jobs` one-workgroup order jobs.
dyg_status prep_reserve(dyg_stream s, size_t bytes, size_t events, size_t jobs) {
    dyg_ctx ctx = s->ctx;
    if (!s->prep) DYG_HIP(ctx, hipStreamCreate(&s->prep));
    if (!s->oscr.mem && s->oscr.alloc(s->sp.max_batch, s->prep) != hipSuccess) {
        set_error(ctx, "stream: device allocation failed");
        return DYG_ERR_NOMEM;
    }
    DYG_HIP(ctx, s->order.reserve(bytes));
    if (jobs) {
        DYG_HIP(ctx, s->jobs_h.reserve(jobs * sizeof(SmallOrderJob)));
        DYG_HIP(ctx, s->jobs_d.reserve(jobs * sizeof(SmallOrderJob)));
    }
    while (s->prep_ev.size() < events) {
        hipEvent_t e = nullptr;
        DYG_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s->prep_ev.push_back(e);
    }
    return DYG_OK;
}

// Enqueues one batch's orders on stream `st` (keys, then the two stable
// sorts, with scratch `sc`) into the arrays of `region` (order_bytes(n)),
// then records `ev` (when given) after them.
dyg_status enqueue_order(dyg_stream s, OrderScratch& sc, hipStream_t st, const uint32_t* u,
                         const uint32_t* v, uint32_t n, char* region, hipEvent_t ev) {
    dyg_ctx ctx = s->ctx;
    auto* key = (unsigned long long*)region;
    auto* idx = (uint32_t*)(region + 8ull * n);
    auto* row = (uint32_t*)(region + 12ull * n);
    auto* ent = (uint32_t*)(region + 20ull * n);
    auto* bcnt = (uint32_t*)(region + 28ull * n);
    hipLaunchKernelGGL(k_prep_keys, dim3((n + 255) / 256), dim3(256), 0, st, u, v, n, sc.key,
                       sc.idx, sc.row, sc.ent, bcnt);
    size_t tb = sc.temp_bytes;
    DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(sc.temp, tb, sc.key, key, sc.idx, idx, (int)n,
                                                    0, key_bits(s->G->V), st));
    tb = sc.temp_bytes;
    DYG_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(sc.temp, tb, sc.row, row, sc.ent, ent,
                                                    (int)(2 * n), 0, row_bits(s->G->V), st));
    if (ev) DYG_HIP(ctx, hipEventRecord(ev, st));
    return DYG_OK;
}

// The arrays of a batch's orders at `region` (order_bytes(n)).
BatchOrder order_at(char* region, uint64_t n) {
    return BatchOrder{(unsigned long long*)region, (uint32_t*)(region + 8 * n),
                      (uint32_t*)(region + 12 * n), (uint32_t*)(region + 20 * n),
                      (uint32_t*)(region + 28 * n)};
}

// A list of batches already in device memory, enqueued back to back: the
// ids of the whole list are checked once before anything changes; each
// batch reports into its own region of the context's pinned report buffer
// and times its walk with its own events; the host waits only where a
// deletion batch needs its recovery list sizes (and when the arenas might
// run out of room, which a running worst-case estimate tells), then once at
// the end.  Same kernels, same order, so the same results as one
// stream_apply per batch.
constexpr size_t kListWords = 16;  // report words per batch (kReportWords used)

dyg_status stream_apply_list(dyg_stream s, uint32_t nb, const int* op, const uint64_t* off,
                             const uint32_t* u, const uint32_t* v, const float* w,
                             int32_t* outcome, dyg_stream_stats* st) {
    dyg_ctx ctx = s->ctx;
    hipStream_t str = ctx->stream;
    const uint64_t o0 = off[0], total = off[nb] - off[0];
    for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t n = off[b + 1] - off[b];
        if ((op[b] != DYG_STREAM_INSERT && op[b] != DYG_STREAM_DELETE) ||
            (n && (!u || !v)) || (n && op[b] == DYG_STREAM_INSERT && !w)) {
            set_error(ctx, "stream apply many: bad operation or missing array");
            return DYG_ERR_ARG;
        }
        if (n > s->sp.max_batch) {
            set_error(ctx, "stream apply many: batch larger than max_batch");
            return DYG_ERR_ARG;
        }
    }
    // every id of the list, before anything changes
    const PinLayout hp = pin_layout((char*)s->pin, 0), dp = pin_layout(s->pin_dev, 0);
    hp.words[kReportBad] = 0;
    for (uint64_t c = 0; c < total; c += (1u << 30)) {
        const uint32_t k = (uint32_t)std::min<uint64_t>(total - c, 1u << 30);
        hipLaunchKernelGGL(k_check_batch, dim3((k + 255) / 256), dim3(256), 0, str, u + o0 + c,
                           v + o0 + c, k, s->G->V, dp.words + kReportBad);
    }
    DYG_HIP(ctx, hipGetLastError());
    // (its result is read once the host has set up the orders below: they
    // only read the ids, never index by them, and change nothing)
    dyg_status r = DYG_OK;
    // every batch's orders, computed on the prep stream while the batches
    // before it run (after the inputs are in place on this stream): a larger
    // batch's by its own keys kernel and sorts, all one-workgroup batches'
    // by one launch per size class (a workgroup each), at the place of the
    // first of them
    std::vector<BatchOrder> ords(nb);
    std::vector<int> ord_ev(nb, -1);
    std::vector<uint32_t> gens;  // the larger batches, in list order
    size_t obytes = 0, ngen = 0, nsmall = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t n = off[b + 1] - off[b];
        if (!n) continue;
        obytes += order_bytes(n);
        if (general_batch(s, n))
            ++ngen;
        else
            ++nsmall;
    }
    const size_t nev = ngen + (nsmall ? 1 : 0);
    if (nev) {
        r = prep_reserve(s, obytes, nev + 1, nsmall);
        if (r != DYG_OK) return r;
        DYG_HIP(ctx, hipEventRecord(s->prep_ev[nev], str));
        DYG_HIP(ctx, hipStreamWaitEvent(s->prep, s->prep_ev[nev], 0));
        // regions in list order; the one-workgroup jobs, smaller class first
        std::vector<SmallOrderJob> j1, j2;
        char* region = (char*)s->order.p;
        for (uint32_t b = 0; b < nb; ++b) {
            const uint64_t o = off[b], n = off[b + 1] - o;
            if (!n) continue;
            ords[b] = order_at(region, n);
            region += order_bytes(n);
            if (general_batch(s, n)) continue;
            const BatchOrder& q = ords[b];
            const SmallOrderJob j{u + o, v + o, const_cast<unsigned long long*>(q.key),
                                  const_cast<uint32_t*>(q.idx), const_cast<uint32_t*>(q.row),
                                  const_cast<uint32_t*>(q.ent), (uint32_t)n, 0u};
            (n <= (uint64_t)kSmallBatch / 2 ? j1 : j2).push_back(j);
        }
        auto* jobs = (SmallOrderJob*)s->jobs_h.p;
        std::copy(j1.begin(), j1.end(), jobs);
        std::copy(j2.begin(), j2.end(), jobs + j1.size());
        if (nsmall) {  // first: the one launch per size class for every small batch
            const uint32_t c1 = (uint32_t)j1.size(), c2 = (uint32_t)j2.size();
            DYG_HIP(ctx, hipMemcpyAsync(s->jobs_d.p, s->jobs_h.p, nsmall * sizeof(SmallOrderJob),
                                        hipMemcpyHostToDevice, s->prep));
            auto* dj = (const SmallOrderJob*)s->jobs_d.p;
            const int kb = key_bits(s->G->V), rb = row_bits(s->G->V);
            DYG_HIP(ctx, launch_small_order(dj, c1, kSmallBatch / 2, kb, rb, s->prep));
            DYG_HIP(ctx, launch_small_order(dj + c1, c2, kSmallBatch, kb, rb, s->prep));
            DYG_HIP(ctx, hipEventRecord(s->prep_ev[ngen], s->prep));
            for (uint32_t b = 0; b < nb; ++b)
                if (off[b + 1] > off[b] && !general_batch(s, off[b + 1] - off[b]))
                    ord_ev[b] = (int)ngen;
        }
        for (uint32_t b = 0, k = 0; b < nb; ++b)
            if (off[b + 1] > off[b] && general_batch(s, off[b + 1] - off[b])) {
                gens.push_back(b);
                ord_ev[b] = (int)k++;
            }
    }
    // report regions: kListWords words per batch, then every outcome code
    const size_t words_bytes = (size_t)nb * kListWords * 8;
    DYG_HIP(ctx, ctx->report.reserve(words_bytes + total * 4 + 256));
    void* rd = nullptr;
    DYG_HIP(ctx, hipHostGetDevicePointer(&rd, ctx->report.p, 0));
    auto* rw_h = (unsigned long long*)ctx->report.p;
    auto* rw_d = (unsigned long long*)rd;
    auto* rc_h = (int32_t*)((char*)ctx->report.p + words_bytes);
    auto* rc_d = (int32_t*)((char*)rd + words_bytes);
    // events: each insertion batch's walk kernel (start, end, taken by the
    // dispatch itself)
    ctx->reserve_events(2 * (size_t)nb + 2);
    if (ctx->evpool.size() < 2 * (size_t)nb + 2) {
        set_error(ctx, "stream apply many: event creation failed");
        return DYG_ERR_HIP;
    }
    hipEvent_t* ev = ctx->evpool.data();
    std::vector<dyg_stream_stats> rs(nb);
    const uint32_t ms = s->sp.del.max_steps, hw = s->sp.heur_walkers;
    // a batch's worst-case arena growth in slots, as the served loop
    // estimates it: G (and H) 4 * 2n * (slack + 4) + 8n; a deletion batch's
    // recovery candidates, n (max_steps - 1) + 4n * heuristic walkers, 8 slots
    // each in H
    auto growth = [&](uint32_t b, uint64_t* add_g, uint64_t* add_h) {
        const uint64_t n = off[b + 1] - off[b];
        const uint64_t add = 4ull * (2ull * n) * (s->sp.slack + 4) + 8ull * n;
        const uint64_t cand = n * (ms - 1) + 4ull * n * hw;
        *add_g = add;
        *add_h = op[b] == DYG_STREAM_DELETE ? add + 8ull * cand : add;
    };
    uint64_t est_g = s->G->used_slots, est_h = s->H->used_slots;  // arena bump pointers, upper
    int last = -1;  // last batch enqueued
    // brings the host's view of both arenas up to the last enqueued batch
    // (after a wait)
    auto absorb_last = [&]() -> dyg_status {
        if (last < 0) return DYG_OK;
        const unsigned long long* wd = rw_h + (size_t)last * kListWords;
        dyg_status q = device_edits_absorb(s->G, wd + 4, nullptr);
        if (q == DYG_OK) q = device_edits_absorb(s->H, wd + 9, nullptr);
        return q;
    };
    dyg_status fail = DYG_OK;
    uint32_t done_b = 0;
    bool started = false;  // the list's start clock enqueued
    int waited = -1;       // the last prep event this stream waits for
    // The larger batches' orders are enqueued by the stream's helper thread,
    // in list order, while this thread enqueues the batches themselves: the
    // prep stream gets all its work at once (its kernels run early, between
    // and beside the first batches' kernels) and no batch waits for the host
    // to enqueue orders.  `recorded` publishes how many of their events are
    // recorded; a batch's wait on its event is enqueued only after that.  The
    // job is posted after the last allocation and waited for after the loop
    // (or before the return of a list whose ids fail the check).
    std::atomic<size_t> recorded{0};
    dyg_status order_fail = DYG_OK;
    const bool helped = !gens.empty();
    if (helped) {
        s->helper->post([&] {
            for (size_t g = 0; g < gens.size(); ++g) {
                const uint32_t b = gens[g];
                const uint64_t o = off[b];
                const dyg_status q =
                    enqueue_order(s, s->oscr, s->prep, u + o, v + o, (uint32_t)(off[b + 1] - o),
                                  (char*)ords[b].key, s->prep_ev[ord_ev[b]]);
                if (q != DYG_OK) {
                    order_fail = q;
                    recorded.store(gens.size() + 1, std::memory_order_release);
                    return;
                }
                recorded.store(g + 1, std::memory_order_release);
            }
        });
    }
    // the ids' check, read while the prep stream starts on the orders
    r = wait_spin(ctx, str, s->ev_sync);
    if (r == DYG_OK && hp.words[kReportBad]) {
        set_error(ctx, "stream apply many: vertex id out of range");
        r = DYG_ERR_ARG;
    }
    if (r != DYG_OK) {
        if (helped) s->helper->wait();
        if (s->prep) (void)hipStreamSynchronize(s->prep);
        return r;
    }
    auto order_ready = [&](int k) -> dyg_status {
        size_t c;
        while ((c = recorded.load(std::memory_order_acquire)) <= (size_t)k) std::this_thread::yield();
        return c > gens.size() ? (order_fail != DYG_OK ? order_fail : DYG_ERR_HIP) : DYG_OK;
    };
    for (uint32_t b = 0; b < nb && fail == DYG_OK; ++b) {
        const uint64_t o = off[b];
        const uint32_t n = (uint32_t)(off[b + 1] - o);
        rs[b].updates = n;
        if (n == 0) {
            done_b = b + 1;
            continue;
        }
        const bool del = op[b] == DYG_STREAM_DELETE;
        uint64_t add = 0, add_h = 0;
        growth(b, &add, &add_h);
        const uint64_t cand = (uint64_t)n * (ms - 1) + 4ull * n * hw;
        if (est_g + add > s->G->cap_slots || est_h + add_h > s->H->cap_slots) {
            // the estimate ran out: wait for the arenas' true use, and grow
            // an arena only when that still leaves too little room
            fail = wait_spin(ctx, str, s->ev_sync);
            if (fail == DYG_OK) fail = absorb_last();
            if (fail == DYG_OK && s->G->used_slots + add > s->G->cap_slots)
                fail = device_edits_prepare(s->G, s->sp.max_batch, 2ull * n + s->sp.max_batch);
            if (fail == DYG_OK && s->H->used_slots + add_h > s->H->cap_slots)
                fail = device_edits_prepare(s->H, s->cmax, 2ull * n + cand);
            if (fail != DYG_OK) break;
            est_g = s->G->used_slots;
            est_h = s->H->used_slots;
            if (est_g + add > s->G->cap_slots || est_h + add_h > s->H->cap_slots) {
                set_error(ctx, "stream apply many: arena headroom could not be reserved");
                fail = DYG_ERR_NOMEM;
                break;
            }
        }
        est_g += add;
        est_h += add_h;
        const Report rep{rc_d + (o - o0), rw_d + (size_t)b * kListWords, s->G->dev.meta,
                         s->H->dev.meta};
        // insertion walks (the stream's dominant kernel) are timed by events
        // taken with their dispatch; deletion walks are not (events on their
        // dispatch cost about 13 us per batch here)
        s->wev0 = del ? nullptr : ev[1 + 2 * b];
        s->wev1 = del ? nullptr : ev[2 + 2 * b];
        if (!started) {  // the list's start on the device clock
            hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, str, dp.words + kReportClock);
            fail = hipGetLastError() == hipSuccess ? DYG_OK : DYG_ERR_HIP;
            started = true;
        }
        const BatchOrder* ord = ord_ev[b] >= 0 ? &ords[b] : nullptr;
        if (fail == DYG_OK && ord && general_batch(s, n)) fail = order_ready(ord_ev[b]);
        if (fail == DYG_OK && ord && ord_ev[b] != waited) {
            if (hipStreamWaitEvent(str, s->prep_ev[ord_ev[b]], 0) != hipSuccess) fail = DYG_ERR_HIP;
            waited = ord_ev[b];
        }
        if (fail == DYG_OK)
            fail = del ? apply_delete(s, u + o, v + o, n, s->cls, s->stats, true, &rs[b], rep,
                                      nullptr, false, 0, ord)
                       : apply_insert(s, u + o, v + o, w + o, n, s->cls, s->stats, true, &rs[b],
                                      rep, nullptr, false, 0, ord);
        if (fail != DYG_OK) break;
        last = (int)b;
        done_b = b + 1;
        s->base += n;
    }
    s->wev0 = ctx->ev0;
    s->wev1 = ctx->ev1;
    if (helped) s->helper->wait();
    if (nev) {  // the prep stream's work is done before the call returns
        DYG_HIP(ctx, hipEventRecord(s->prep_ev[nev], s->prep));
        DYG_HIP(ctx, hipStreamWaitEvent(str, s->prep_ev[nev], 0));
    }
    // the list's one wait, then the reports of every batch enqueued
    r = wait_spin(ctx, str, s->ev_sync);
    if (r == DYG_OK) r = absorb_last();
    if (fail == DYG_OK) fail = r;
    // each batch's span on the device clock runs from the previous nonempty
    // batch's report (the list's start for the first) to its own; the spans
    // add up to the list's
    std::vector<unsigned long long> pclk(done_b);
    {
        unsigned long long c = hp.words[kReportClock];  // the list's start
        for (uint32_t b = 0; b < done_b; ++b) {
            pclk[b] = c;
            if (rs[b].updates) c = rw_h[(size_t)b * kListWords + kReportClock];
        }
    }
    auto read_reports = [&](uint32_t b0, uint32_t b1) {
        for (uint32_t b = b0; b < b1; ++b) {
            const uint64_t o = off[b] - o0;
            const uint32_t n = (uint32_t)rs[b].updates;
            if (n == 0) {
                if (st) st[b] = rs[b];
                continue;
            }
            const unsigned long long* wd = rw_h + (size_t)b * kListWords;
            uint64_t counts[16] = {};
            count_codes(rc_h + o, outcome ? outcome + o : nullptr, n, counts);
            dyg_stream_stats& x = rs[b];
            x.weight_updates = counts[kOutWeight];
            x.found = counts[kOutFound];
            x.added = counts[kOutAdded];
            x.skipped = counts[kOutSkip];
            x.g_only = counts[kOutGOnly];
            x.recovered = counts[kOutRecovered];
            x.heuristic = counts[kOutHeuristic];
            x.invalid = n - (x.weight_updates + x.found + x.added + x.skipped + x.g_only +
                             x.recovered + x.heuristic);
            x.hops = wd[0];
            x.redirects = wd[1];
            if (wd[3]) x.candidates = wd[3];
            float t = 0.f;
            if (op[b] != DYG_STREAM_DELETE &&
                hipEventElapsedTime(&t, ev[1 + 2 * b], ev[2 + 2 * b]) == hipSuccess)
                x.walk_ms = t;
            const unsigned long long clk = wd[kReportClock];
            if (ctx->clock_khz > 0 && clk >= pclk[b])
                x.total_ms = (double)(clk - pclk[b]) / ctx->clock_khz;
            if (st) st[b] = x;
        }
    };
    // a long list's reports are read by two threads (this one and the
    // helper, each about half the updates): the outcome codes come from
    // pinned memory the device wrote, so the copy runs at memory speed per
    // thread
    uint64_t done_upd = 0;
    for (uint32_t b = 0; b < done_b; ++b) done_upd += rs[b].updates;
    uint32_t mid = 0;
    for (uint64_t acc = 0; mid < done_b && 2 * acc < done_upd; ++mid) acc += rs[mid].updates;
    if (s->helper && done_upd >= (1u << 16) && mid > 0 && mid < done_b) {
        s->helper->post([&] { read_reports(0, mid); });
        read_reports(mid, done_b);
        s->helper->wait();
    } else {
        read_reports(0, done_b);
    }
    if (fail != DYG_OK) return fail;
    // a batch's worth of arena headroom for what comes next
    r = device_edits_prepare(s->G, s->sp.max_batch, s->sp.max_batch);
    if (r == DYG_OK) r = device_edits_prepare(s->H, s->cmax, s->cmax);
    return r;
}

}  // namespace

extern "C" dyg_status dyg_stream_apply(dyg_stream s, int op, const uint32_t* u, const uint32_t* v,
                                       const float* w, uint32_t n, int32_t* outcome,
                                       dyg_stream_stats* st) {
    return stream_apply(s, op, u, v, w, n, outcome, st, false);
}

extern "C" dyg_status dyg_stream_apply_dev(dyg_stream s, int op, const uint32_t* u,
                                           const uint32_t* v, const float* w, uint32_t n,
                                           int32_t* outcome, dyg_stream_stats* st) {
    return stream_apply(s, op, u, v, w, n, outcome, st, true);
}

extern "C" dyg_status dyg_stream_apply_many_dev(dyg_stream s, uint32_t nb, const int* op,
                                                const uint64_t* off, const uint32_t* u,
                                                const uint32_t* v, const float* w,
                                                int32_t* outcome, dyg_stream_stats* st) {
    if (!s) return DYG_ERR_ARG;
    if (nb && (!op || !off)) {
        set_error(s->ctx, "stream apply many: missing batch list");
        return DYG_ERR_ARG;
    }
    for (uint32_t b = 0; b < nb; ++b)  // the whole list, before anything changes
        if (off[b + 1] < off[b] || off[b + 1] - off[b] > 0xFFFFFFFFull) {
            set_error(s->ctx, "stream apply many: bad batch offsets");
            return DYG_ERR_ARG;
        }
    if (nb && !(s->sp.flags & DYG_STREAM_PLAIN))
        return stream_apply_list(s, nb, op, off, u, v, w, outcome, st);
    for (uint32_t b = 0; b < nb; ++b) {
        const uint64_t o = off[b];
        const dyg_status r =
            stream_apply(s, op[b], u + o, v + o, w ? w + o : nullptr, (uint32_t)(off[b + 1] - o),
                         outcome ? outcome + o : nullptr, st ? st + b : nullptr, true);
        if (r != DYG_OK) return r;
    }
    return DYG_OK;
}

extern "C" dyg_graph dyg_stream_graph(dyg_stream s, int which) {
    if (!s) return nullptr;
    return which == 0 ? s->G : s->H;
}

extern "C" uint64_t dyg_stream_position(dyg_stream s) { return s ? s->base : 0; }

extern "C" void dyg_stream_free(dyg_stream s) { delete s; }

// ---------------------------------------------------------------------------
// Served form: a request loop over one stream.  Requests are admitted into
// slots (pinned host + device buffers of one request) while slots and the
// memory budget allow; an admitted request's input is copied to the device
// at once on a copy stream, so it overlaps the requests ahead of it (a small
// insertion request is read in place from pinned memory instead); a worker
// thread launches requests in arrival order on the compute stream without
// waiting for the previous one to finish (deletions wait once, inside, for
// their walks); each request's last kernel writes its results into the
// slot's pinned memory, and a completion thread waits for them and
// timestamps them.  Order of application = order of admission, so every
// result equals the one the same request gets from dyg_stream_apply in the
// same sequence.
// ---------------------------------------------------------------------------
namespace {

struct ServeSlot {
    uint32_t* din = nullptr;  // u, v, w of the request at stride n
    int32_t* dcls = nullptr;
    unsigned long long* dstats = nullptr;
    void* host = nullptr;     // pinned, laid out by pin_layout; device address hdev
    char* hdev = nullptr;
    hipEvent_t e_in = nullptr, e_done = nullptr;
    hipEvent_t ew0 = nullptr, ew1 = nullptr;  // the request's walk kernel, start and end
    char* omem = nullptr;  // the request's orders (BatchOrder), computed at submit
};

struct ServeReq {
    uint64_t ticket = 0;
    int op = 0;
    uint32_t n = 0, slot = 0;
    uint64_t need = 0;  // admitted bytes
    double t_submit = 0, t_start = 0, t_done = 0;
    dyg_status status = DYG_OK;
    std::string error;
    std::vector<int32_t> outcome;
    uint64_t hops = 0, redirects = 0;
    double walk_ms = 0;
    bool copied = false;  // input copied to the slot's device buffer (else read in place)
    bool ordered = false; // its orders computed on the copy stream (e_in follows them)
    // device inputs (dyg_serve_submit_dev): the caller's arrays, read in place
    const uint32_t* du = nullptr;
    const uint32_t* dv = nullptr;
    const float* dw = nullptr;
    bool done = false;
};

}  // namespace

struct dyg_serve_s {
    dyg_stream s = nullptr;
    hipStream_t copy = nullptr;  // inputs in
    uint32_t slot_cap = 0;
    std::vector<ServeSlot> slots;
    dyg::AdmissionGate gate;  // slots and the memory budget (under mu)
    uint64_t used_est_g = 0, used_est_h = 0;  // arena bump pointers, upper estimates
    std::mutex mu;
    std::condition_variable cv_work, cv_inflight, cv_done;
    bool fail_next_copy = false;  // testing hook (dyg_serve_fail_next_copy)
    // id check of device inputs (dyg_serve_submit_dev), one at a time on the
    // copy stream: a pinned flag word and the event that orders it
    std::mutex chk_mu;
    unsigned long long* chk = nullptr;
    unsigned long long* chk_dev = nullptr;
    hipEvent_t e_chk = nullptr;
    OrderScratch oscr;  // the copy stream's scratch for requests' orders
    std::deque<ServeReq*> queue, inflight;
    std::unordered_map<uint64_t, ServeReq*> reqs;
    std::thread worker, completer;
    bool stop = false, worker_done = false;
    dyg_status failed = DYG_OK;
    std::string fail_msg;
    // the first ticket the failure stops: requests admitted before it go on
    // (a failed input copy never reached the device; a request that failed
    // on the device stops everything after it)
    uint64_t fail_ticket = ~0ull;
    uint64_t next_ticket = 0;
};

namespace {

uint64_t request_need(const dyg_serve_s* v, int op, uint32_t n) {
    // device + pinned buffers are preallocated per slot; what a request adds
    // is arena growth: 2 slots per appended edge per graph, doubled for row
    // relocation, plus the recovery appends of a deletion batch
    const uint64_t rec = op == DYG_STREAM_DELETE
                             ? (uint64_t)n * (v->s->sp.del.max_steps - 1 + 4 * v->s->sp.heur_walkers)
                             : (uint64_t)n;
    return 16ull * 4 * (2ull * n + 2 * rec) + 32ull * n;
}

// Grows the arenas (and resyncs the host's counts) when the headroom left
// could not hold the next request's worst case.
dyg_status ensure_headroom(dyg_serve_s* v, const ServeReq& r) {
    dyg_stream s = v->s;
    // the request's own worst case (a deletion request's recovery candidates
    // are bounded by its size, as in a list), not a largest batch's
    const uint64_t add = 4ull * (2ull * r.n) * (s->sp.slack + 4) + 8ull * r.n;
    const uint64_t cand =
        (uint64_t)r.n * (s->sp.del.max_steps - 1) + 4ull * r.n * s->sp.heur_walkers;
    const uint64_t add_h = r.op == DYG_STREAM_DELETE ? add + 8ull * std::min<uint64_t>(cand, s->cmax)
                                                     : add;
    if (v->used_est_g + add <= s->G->cap_slots && v->used_est_h + add_h <= s->H->cap_slots) {
        v->used_est_g += add;
        v->used_est_h += add_h;
        return DYG_OK;
    }
    dyg_status st = device_edits_end(s->G, nullptr);  // waits for the compute stream
    if (st == DYG_OK) st = device_edits_end(s->H, nullptr);
    if (st == DYG_OK) st = device_edits_prepare(s->G, s->sp.max_batch, 2ull * r.n + s->sp.max_batch);
    if (st == DYG_OK) st = device_edits_prepare(s->H, s->cmax, 2ull * r.n + s->cmax);
    if (st != DYG_OK) return st;
    v->used_est_g = s->G->used_slots + add;
    v->used_est_h = s->H->used_slots + add_h;
    size_t mfree = 0, mtotal = 0;  // arena growth is resident state from now on
    if (hipMemGetInfo(&mfree, &mtotal) == hipSuccess) {
        std::lock_guard<std::mutex> lk(v->mu);
        const uint64_t used = mtotal - mfree, adm = v->gate.admitted();
        v->gate.raise_resident(used > adm ? used - adm : 0);
    }
    if (v->used_est_g > s->G->cap_slots || v->used_est_h > s->H->cap_slots) {
        set_error(s->ctx, "serve: arena headroom could not be reserved");
        return DYG_ERR_NOMEM;
    }
    return DYG_OK;
}

void serve_worker(dyg_serve_s* v) {
    dyg_stream s = v->s;
    hipStream_t cs = s->ctx->stream;
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        ServeReq* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(v->mu);
            v->cv_work.wait(lk, [&] { return v->stop || !v->queue.empty(); });
            if (v->queue.empty()) {  // stop requested and drained
                v->worker_done = true;
                v->cv_inflight.notify_all();
                return;
            }
            r = v->queue.front();
            v->queue.pop_front();
        }
        ServeSlot& sl = v->slots[r->slot];
        r->t_start = now_ms();
        // a request whose submit failed (its input never reached the device)
        // is never applied, and neither is anything admitted after a failure
        dyg_status st = r->status;
        if (st == DYG_OK) {
            std::lock_guard<std::mutex> lk(v->mu);
            if (r->ticket >= v->fail_ticket) st = v->failed;
        }
        if (st == DYG_OK) st = ensure_headroom(v, *r);
        if (st == DYG_OK && (r->copied || r->ordered) &&
            hipStreamWaitEvent(cs, sl.e_in, 0) != hipSuccess)
            st = DYG_ERR_HIP;
        const BatchOrder ord = order_at(sl.omem, r->n);
        dyg_stream_stats rs{};
        const uint32_t n = r->n;
        const PinLayout dp = pin_layout(sl.hdev, n);
        const uint32_t *du = dp.u, *dv = dp.v;
        const float* dw = dp.w;
        if (r->copied) {
            du = sl.din;
            dv = sl.din + n;
            dw = (const float*)(sl.din + 2ull * n);
        } else if (r->du) {  // device inputs, ids checked at submit
            du = r->du;
            dv = r->dv;
            dw = r->dw;
        }
        const Report rep{dp.cls, dp.words, s->G->dev.meta, s->H->dev.meta};
        // the walk kernel takes the slot's events with its dispatch (read by
        // the completion thread, which frees the slot only after that)
        s->wev0 = sl.ew0;
        s->wev1 = sl.ew1;
        const BatchOrder* po = r->ordered ? &ord : nullptr;
        if (st == DYG_OK)
            st = r->op == DYG_STREAM_INSERT
                     ? apply_insert(s, du, dv, dw, n, sl.dcls, sl.dstats, true, &rs, rep, nullptr,
                                    false, 0, po)
                     : apply_delete(s, du, dv, n, sl.dcls, sl.dstats, true, &rs, rep, nullptr,
                                    false, 0, po);
        s->wev0 = s->ctx->ev0;
        s->wev1 = s->ctx->ev1;
        if (st == DYG_OK) {
            s->base += n;
            if (hipEventRecord(sl.e_done, cs) != hipSuccess) st = DYG_ERR_HIP;
        }
        std::lock_guard<std::mutex> lk(v->mu);
        if (st != DYG_OK) {
            if (r->status == DYG_OK) {
                r->status = st;
                r->error = v->failed != DYG_OK ? "serve: stopped after an error: " + v->fail_msg
                                               : std::string(dyg_last_error(s->ctx));
            }
            if (v->failed == DYG_OK) {
                v->failed = st;
                v->fail_msg = r->error;
                v->fail_ticket = r->ticket + 1;
            }
            v->gate.fail(v->failed);
        }
        v->inflight.push_back(r);
        v->cv_inflight.notify_one();
    }
}

void serve_completer(dyg_serve_s* v) {
    (void)hipSetDevice(v->s->ctx->device);
    for (;;) {
        ServeReq* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(v->mu);
            v->cv_inflight.wait(lk, [&] { return v->worker_done || !v->inflight.empty(); });
            if (v->inflight.empty()) return;  // the worker has finished and all is delivered
            r = v->inflight.front();
            v->inflight.pop_front();
        }
        ServeSlot& sl = v->slots[r->slot];
        if (r->status == DYG_OK) {
            if (spin_event(sl.e_done) != hipSuccess) {
                r->status = DYG_ERR_HIP;
                r->error = "serve: request failed on the device";
            } else {
                const PinLayout hp = pin_layout((char*)sl.host, r->n);
                r->outcome.assign(hp.cls, hp.cls + r->n);
                r->hops = hp.words[0];
                r->redirects = hp.words[1];
                float t = 0.f;
                if (hipEventElapsedTime(&t, sl.ew0, sl.ew1) == hipSuccess) r->walk_ms = t;
            }
        }
        r->t_done = now_ms();
        std::lock_guard<std::mutex> lk(v->mu);
        v->gate.release(r->slot, r->need);
        r->done = true;
        v->cv_done.notify_all();
    }
}

}  // namespace

extern "C" dyg_status dyg_serve_open(dyg_stream s, uint32_t slots, uint32_t max_request,
                                     uint64_t budget_bytes, dyg_serve* out) {
    if (!s || !out || slots == 0 || max_request == 0) return DYG_ERR_ARG;
    *out = nullptr;
    dyg_ctx ctx = s->ctx;
    if (s->sp.flags & DYG_STREAM_PLAIN) {
        set_error(ctx, "serve: the plain form is served by dyg_stream_apply");
        return DYG_ERR_ARG;
    }
    if (max_request > s->sp.max_batch) {
        set_error(ctx, "serve: max_request exceeds the stream's max_batch");
        return DYG_ERR_ARG;
    }
    auto* v = new dyg_serve_s;
    v->s = s;
    v->slot_cap = max_request;
    v->slots.resize(slots);
    bool ok = hipStreamCreateWithFlags(&v->copy, hipStreamNonBlocking) == hipSuccess;
    for (ServeSlot& sl : v->slots) {
        if (!ok) break;
        const size_t n = max_request;
        void* hd = nullptr;
        ok = hipMalloc(&sl.din, n * 12) == hipSuccess && hipMalloc(&sl.dcls, n * 4) == hipSuccess &&
             hipMalloc(&sl.dstats, 32) == hipSuccess &&
             hipMemsetAsync(sl.dstats, 0, 32, ctx->stream) == hipSuccess &&
             hipHostMalloc(&sl.host, pin_bytes(max_request), hipHostMallocDefault) == hipSuccess &&
             hipHostGetDevicePointer(&hd, sl.host, 0) == hipSuccess &&
             hipEventCreateWithFlags(&sl.e_in, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&sl.e_done, hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&sl.ew0) == hipSuccess && hipEventCreate(&sl.ew1) == hipSuccess &&
             hipMalloc(&sl.omem, order_bytes(max_request)) == hipSuccess;
        sl.hdev = (char*)hd;
    }
    if (ok) ok = v->oscr.alloc(max_request, v->copy) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(ctx->stream) == hipSuccess;
    size_t mfree = 0, mtotal = 0;
    if (ok) ok = hipMemGetInfo(&mfree, &mtotal) == hipSuccess;
    if (!ok) {
        set_error(ctx, "serve: slot allocation failed");
        dyg_serve_close(v);
        return DYG_ERR_NOMEM;
    }
    const uint64_t budget = budget_bytes ? std::min<uint64_t>(budget_bytes, mtotal) : mtotal;
    // graphs, stream buffers and slots already in place
    v->gate.open(slots, budget, mtotal - mfree);
    if (v->gate.resident() >= budget) {
        set_error(ctx, "serve: resident state already exceeds the memory budget");
        dyg_serve_close(v);
        return DYG_ERR_ADMISSION;
    }
    v->used_est_g = s->G->used_slots;
    v->used_est_h = s->H->used_slots;
    // served requests report into their slots; the stream's own rejection
    // word must not carry over from an earlier rejected dyg_stream_apply_dev
    ((unsigned long long*)s->pin)[kReportBad] = 0;
    void* cd = nullptr;
    if (hipHostMalloc((void**)&v->chk, 256, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer(&cd, v->chk, 0) != hipSuccess ||
        hipEventCreateWithFlags(&v->e_chk, hipEventDisableTiming) != hipSuccess) {
        set_error(ctx, "serve: pinned allocation failed");
        dyg_serve_close(v);
        return DYG_ERR_NOMEM;
    }
    v->chk_dev = (unsigned long long*)cd;
    v->worker = std::thread(serve_worker, v);
    v->completer = std::thread(serve_completer, v);
    *out = v;
    return DYG_OK;
}

namespace {

// Both submits: host arrays (dev false) are range-checked on the host and
// copied into the slot; device arrays are range-checked by a kernel on the
// copy stream before admission and read in place by the request.
dyg_status serve_submit(dyg_serve v, int op, const uint32_t* u, const uint32_t* w_v,
                        const float* w, uint32_t n, uint64_t* ticket, bool dev) {
    if (!v || !ticket || (op != DYG_STREAM_INSERT && op != DYG_STREAM_DELETE) ||
        (n && (!u || !w_v || (op == DYG_STREAM_INSERT && !w))))
        return DYG_ERR_ARG;
    dyg_ctx ctx = v->s->ctx;
    const double t0 = now_ms();
    if (n == 0 || n > v->slot_cap) {
        set_error(ctx, "serve submit: request size must be in [1, max_request]");
        return DYG_ERR_ARG;
    }
    if (dev) {
        std::lock_guard<std::mutex> lk(v->chk_mu);
        v->chk[0] = 0;
        hipLaunchKernelGGL(k_check_batch, dim3((n + 255) / 256), dim3(256), 0, v->copy, u, w_v,
                           n, v->s->G->V, v->chk_dev);
        if (hipGetLastError() != hipSuccess || hipEventRecord(v->e_chk, v->copy) != hipSuccess ||
            spin_event(v->e_chk) != hipSuccess) {
            set_error(ctx, "serve submit: id check failed on the device");
            return DYG_ERR_HIP;
        }
        if (v->chk[0]) {
            set_error(ctx, "serve submit: vertex id out of range");
            return DYG_ERR_ARG;
        }
    } else {
        uint32_t mx = 0;
        for (uint32_t i = 0; i < n; ++i) mx = std::max(mx, std::max(u[i], w_v[i]));
        if (mx >= v->s->G->V) {
            set_error(ctx, "serve submit: vertex id out of range");
            return DYG_ERR_ARG;
        }
    }
    const uint64_t need = request_need(v, op, n);
    auto* r = new ServeReq;
    r->op = op;
    r->n = n;
    r->need = need;
    r->t_submit = t0;
    uint32_t slot = 0;
    {
        std::unique_lock<std::mutex> lk(v->mu);
        // admission: a free slot and room in the budget; tickets (and the
        // order of application) follow the order of admission
        const dyg_status adm = v->gate.admit(lk, need, &slot);
        if (adm != DYG_OK) {
            set_error(ctx, adm == DYG_ERR_ADMISSION
                               ? std::string("serve submit: request cannot fit the memory budget")
                           : adm == DYG_ERR_STATE ? std::string("serve submit: server is closing")
                                                  : "serve: stopped after an error: " + v->fail_msg);
            delete r;
            return adm;
        }
        r->slot = slot;
        // the input goes to the device now, ahead of the requests before it
        // (a small insertion request is read in place from pinned memory)
        ServeSlot& sl = v->slots[slot];
        const PinLayout hp = pin_layout((char*)sl.host, n);
        const bool ins = op == DYG_STREAM_INSERT;
        bool copy_ok = true;
        if (dev) {
            r->du = u;
            r->dv = w_v;
            r->dw = ins ? w : nullptr;
            r->copied = false;
        } else {
            std::memcpy(hp.u, u, n * 4ull);
            std::memcpy(hp.v, w_v, n * 4ull);
            if (ins) std::memcpy(hp.w, w, n * 4ull);
            r->copied = !(ins && n <= small_limit(v->s));
        }
        if (r->copied) {
            if (v->fail_next_copy) {  // testing hook: report a failure, issue nothing
                v->fail_next_copy = false;
                copy_ok = false;
            } else {
                copy_ok = hipMemcpyAsync(sl.din, hp.u, n * (ins ? 12ull : 8ull),
                                         hipMemcpyHostToDevice, v->copy) == hipSuccess;
            }
        }
        if (copy_ok) {
            // the request's orders, on the copy stream behind its input: they
            // overlap the requests ahead of it on the device
            const uint32_t* ou = r->copied ? sl.din : dev ? u : pin_layout(sl.hdev, n).u;
            const uint32_t* ov = r->copied ? sl.din + n : dev ? w_v : pin_layout(sl.hdev, n).v;
            if (general_batch(v->s, n)) {
                copy_ok = enqueue_order(v->s, v->oscr, v->copy, ou, ov, n, sl.omem, nullptr) ==
                          DYG_OK;
            } else {
                const BatchOrder q = order_at(sl.omem, n);
                const SmallOrderJob j{ou, ov, const_cast<unsigned long long*>(q.key),
                                      const_cast<uint32_t*>(q.idx), const_cast<uint32_t*>(q.row),
                                      const_cast<uint32_t*>(q.ent), n, 0u};
                copy_ok = launch_small_order_one(j, key_bits(v->s->G->V), row_bits(v->s->G->V),
                                                 v->copy) == hipSuccess;
            }
            copy_ok = copy_ok && hipEventRecord(sl.e_in, v->copy) == hipSuccess;
            r->ordered = copy_ok;
        }
        if (!copy_ok) {
            // the request never reaches the worker; the device's state is no
            // longer known, so the server stops: every later request fails
            v->gate.release(slot, need);
            if (v->failed == DYG_OK) {
                v->failed = DYG_ERR_HIP;
                v->fail_msg = "serve submit: input copy failed";
                v->fail_ticket = v->next_ticket;  // this request took no ticket
            }
            v->gate.fail(v->failed);
            set_error(ctx, "serve submit: input copy failed");
            delete r;
            return DYG_ERR_HIP;
        }
        r->ticket = v->next_ticket++;
        v->reqs[r->ticket] = r;
        v->queue.push_back(r);
        *ticket = r->ticket;
    }
    v->cv_work.notify_one();
    return DYG_OK;
}

}  // namespace

extern "C" dyg_status dyg_serve_submit(dyg_serve v, int op, const uint32_t* u, const uint32_t* w_v,
                                       const float* w, uint32_t n, uint64_t* ticket) {
    return serve_submit(v, op, u, w_v, w, n, ticket, false);
}

extern "C" dyg_status dyg_serve_submit_dev(dyg_serve v, int op, const uint32_t* u,
                                           const uint32_t* w_v, const float* w, uint32_t n,
                                           uint64_t* ticket) {
    return serve_submit(v, op, u, w_v, w, n, ticket, true);
}

extern "C" dyg_status dyg_serve_fail_next_copy(dyg_serve v) {
    if (!v) return DYG_ERR_ARG;
    std::lock_guard<std::mutex> lk(v->mu);
    v->fail_next_copy = true;
    return DYG_OK;
}

extern "C" dyg_status dyg_serve_wait(dyg_serve v, uint64_t ticket, int32_t* outcome,
                                     dyg_serve_result* res) {
    if (!v) return DYG_ERR_ARG;
    ServeReq* r = nullptr;
    {
        std::unique_lock<std::mutex> lk(v->mu);
        auto it = v->reqs.find(ticket);
        if (it == v->reqs.end()) {
            set_error(v->s->ctx, "serve wait: unknown ticket");
            return DYG_ERR_ARG;
        }
        r = it->second;
        v->cv_done.wait(lk, [&] { return r->done; });
        v->reqs.erase(it);
    }
    const dyg_status st = r->status;
    if (st != DYG_OK) set_error(v->s->ctx, r->error);
    if (st == DYG_OK && outcome) std::memcpy(outcome, r->outcome.data(), r->n * 4ull);
    if (res) {
        res->n = r->n;
        res->queue_ms = r->t_start - r->t_submit;
        res->latency_ms = r->t_done - r->t_submit;
        res->hops = r->hops;
        res->redirects = r->redirects;
        res->walk_ms = r->walk_ms;
    }
    delete r;
    return st;
}

extern "C" void dyg_serve_close(dyg_serve v) {
    if (!v) return;
    {
        std::lock_guard<std::mutex> lk(v->mu);
        v->stop = true;
        v->gate.close();
    }
    v->cv_work.notify_all();
    v->cv_inflight.notify_all();
    if (v->worker.joinable()) v->worker.join();
    if (v->completer.joinable()) v->completer.join();
    if (v->s) {
        (void)hipStreamSynchronize(v->s->ctx->stream);
        (void)device_edits_end(v->s->G, nullptr);  // host counts current again
        (void)device_edits_end(v->s->H, nullptr);
    }
    if (v->copy) (void)hipStreamSynchronize(v->copy);
    for (ServeSlot& sl : v->slots) {
        void* ps[] = {sl.din, sl.dcls, sl.dstats, sl.omem};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (sl.host) (void)hipHostFree(sl.host);
        if (sl.e_in) (void)hipEventDestroy(sl.e_in);
        if (sl.e_done) (void)hipEventDestroy(sl.e_done);
        if (sl.ew0) (void)hipEventDestroy(sl.ew0);
        if (sl.ew1) (void)hipEventDestroy(sl.ew1);
    }
    for (auto& kv : v->reqs) delete kv.second;
    for (ServeReq* r : v->queue) delete r;
    for (ServeReq* r : v->inflight) delete r;
    if (v->chk) (void)hipHostFree(v->chk);
    if (v->e_chk) (void)hipEventDestroy(v->e_chk);
    if (v->copy) (void)hipStreamDestroy(v->copy);
    delete v;
}
