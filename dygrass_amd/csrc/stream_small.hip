// stream_small.hip -- batches of up to kSmallBatch updates in one-workgroup
// kernels around the walk.
//
// The general pipeline of a batch (stream.hip) is thirty to fifty launches;
// for a batch of a few hundred edges every kernel runs for a few
// microseconds and the launches themselves set the pace.  Here one
// workgroup of 1024 lanes does a phase with the batch in LDS.
//
// Insertion batch:
//   pre:  sort the (min,max) keys (stable: block radix sort, or a rank by
//         counting for a few keys), classify each distinct edge in batch
//         order (weight update / new), append the new edges to G (appends
//         sorted by row the same way, so each row's appends keep batch
//         order; relocation through the same device bump allocator as the
//         general pipeline), apply the weight updates, and compact the kept
//         edges for the walk;
//   post: append the edges that found no path to H the same way, write the
//         outcome codes and report the batch into pinned host memory.
// Deletion batch:
//   pre:  sort and classify as above (skip / G only / G and H), delete from
//         G and from H (each row's removals in batch order, swap-with-last,
//         one lane per row), compact the edges H held for the walk on G;
//   post: path hop counts and the no-path list (block scans), the heuristic
//         two-hop walks from the no-path edges' endpoints, the candidate list
//         (path hops, then heuristic hops) in LDS, the first occurrence of
//         each candidate H lacks (sorted again), appended to H with G's
//         weight, outcome codes and the report.  A candidate list longer than
//         kSmallCand is left to the general recovery: the post kernel then
//         writes the general pipeline's intermediate lists and a flag instead.
// After each phase one multi-workgroup launch refreshes the descriptors
// copied into the slots pointing at the touched rows.
// Same results as the general pipeline (tested against the oracle).
#include <hipcub/hipcub.hpp>

#include "internal.h"
#include "philox.h"

namespace dyg {
namespace {

constexpr int kThreads = 1024;
constexpr uint32_t kPadRow = 0xFFFFFFFFu;
constexpr unsigned long long kPad64 = ~0ull;
constexpr int32_t kOutWeight = 1, kOutFound = 2, kOutAdded = 3, kOutSkip = 4, kOutGOnly = 5,
                  kOutRecovered = 6, kOutHeuristic = 7, kOutInvalid = 8;

using Scan = hipcub::BlockScan<uint32_t, kThreads>;
using SortKeys64 = hipcub::BlockRadixSort<unsigned long long, kThreads, 1, uint32_t>;
using SortRows = hipcub::BlockRadixSort<uint32_t, kThreads, 2, uint32_t>;

// Up to this many items a sort ranks by counting (one wave, a few hundred
// LDS reads); above it, a block radix sort (O(items) work per pass: counting
// would be items^2 compares on one CU).
constexpr uint32_t kCountMax = 64;

// Phase-local scratch, one phase after another (a barrier between).
union TempU {
    SortKeys64::TempStorage k64;
    SortRows::TempStorage rows;
    Scan::TempStorage scan;
    unsigned long long kin[kThreads];  // counting sort of keys: keys in input order
    uint32_t rk[2 * kThreads];         // counting sort of rows: rows in entry order
};

__device__ __forceinline__ int find_in_row(const uint2* rows, const uint4* adj, uint32_t a,
                                           uint32_t b) {
    return slot_of(adj, rows[a], b);
}

__device__ __forceinline__ bool holds(const uint2* rows, const uint4* adj, uint32_t a, uint32_t b) {
    return rows[a].y <= rows[b].y ? find_in_row(rows, adj, a, b) >= 0
                                  : find_in_row(rows, adj, b, a) >= 0;
}

__device__ __forceinline__ unsigned long long edge_key(uint32_t u, uint32_t v) {
    const uint32_t lo = min(u, v), hi = max(u, v);
    return lo == hi ? kPad64 : ((unsigned long long)lo << 32) | hi;
}

__device__ __forceinline__ uint32_t row_pad(int row_bits) {
    return row_bits >= 32 ? 0xFFFFFFFFu : (1u << row_bits) - 1u;
}

// Stable sort of n keys (lane t < n holds kt, the others kPad64; bits
// [0, bits) order them): key[rank] = kt, idx[rank] = t.  Lanes past n keep
// their place.
__device__ void sort_keys64(TempU& tmp, unsigned long long* key, uint32_t* idx,
                            unsigned long long kt, uint32_t t, uint32_t n, int bits) {
    if (n <= kCountMax) {
        tmp.kin[t] = kt;
        __syncthreads();
        uint32_t rank = t;
        if (t < n) {
            rank = 0;
            for (uint32_t j = 0; j < n; ++j) {
                const unsigned long long kj = tmp.kin[j];
                rank += (kj < kt || (kj == kt && j < t)) ? 1u : 0u;
            }
        }
        key[rank] = kt;
        idx[rank] = t;
    } else {
        unsigned long long k[1] = {kt};
        uint32_t v[1] = {t};
        SortKeys64(tmp.k64).Sort(k, v, 0, bits);
        key[t] = k[0];
        idx[t] = v[0];
    }
    __syncthreads();
}

// Stable sort of the entries 2e (row r0 of lane e) and 2e+1 (r1) of the
// first `span` lanes by row, pads (row_pad) last: srow/sent[p] = row and
// entry at sorted position p < 2*span.
__device__ void sort_rows(TempU& tmp, uint32_t* srow, uint32_t* sent, uint32_t r0, uint32_t r1,
                          uint32_t t, uint32_t span, int bits) {
    if (span <= kCountMax) {
        tmp.rk[2 * t] = r0;
        tmp.rk[2 * t + 1] = r1;
        __syncthreads();
        if (t < span) {
            uint32_t a0 = 0, a1 = 0;
            for (uint32_t y = 0; y < 2 * span; ++y) {
                const uint32_t ry = tmp.rk[y];
                a0 += (ry < r0 || (ry == r0 && y < 2 * t)) ? 1u : 0u;
                a1 += (ry < r1 || (ry == r1 && y < 2 * t + 1)) ? 1u : 0u;
            }
            srow[a0] = r0;
            sent[a0] = 2 * t;
            srow[a1] = r1;
            sent[a1] = 2 * t + 1;
        }
    } else {
        uint32_t k[2] = {r0, r1};
        uint32_t v[2] = {2 * t, 2 * t + 1};
        SortRows(tmp.rows).Sort(k, v, 0, bits);
        srow[2 * t] = k[0];
        srow[2 * t + 1] = k[1];
        sent[2 * t] = v[0];
        sent[2 * t + 1] = v[1];
    }
    __syncthreads();
}

// Lists the rows r[q] flagged h[q] by this lane, in any order, and their count.
__device__ __forceinline__ void list_rows(Scan::TempStorage& scan, const bool (&h)[2],
                                          const uint32_t (&r)[2], uint32_t t,
                                          uint32_t* __restrict__ touched,
                                          uint32_t* __restrict__ ntouched) {
    uint32_t pos = 0, total = 0;
    Scan(scan).ExclusiveSum((h[0] ? 1u : 0u) + (h[1] ? 1u : 0u), pos, total);
    if (h[0]) touched[pos++] = r[0];
    if (h[1]) touched[pos] = r[1];
    if (t == 0) *ntouched = total;
}

// Row groups of a sorted entry list: sorted position p (one of this lane's
// two, 2t + j) starts a group when it holds a row and differs from p-1; the
// group runs to `end`.
__device__ __forceinline__ bool group_head(const uint32_t* srow, uint32_t p, uint32_t lim,
                                           uint32_t pad, uint32_t& end) {
    const uint32_t r = srow[p];
    if (p >= lim || r == pad || (p > 0 && srow[p - 1] == r)) return false;
    end = p + 1;
    while (end < lim && srow[end] == r) ++end;
    return true;
}

// Appends edge t (lane t < span: src[t]-dst[t] with weight w[t]) to both
// rows when `take`.  Entry 2t is the append to src[t]'s row, 2t+1 to
// dst[t]'s; each row's appends keep entry order (the general pipeline's
// order), written by the lane holding the row's first sorted entry.
// Returns after the rows are final, with the touched rows listed for
// k_refresh_rows, which rewrites the neighbours' copies of their descriptors
// across the whole GPU (in one workgroup that refresh would be deg^2
// dependent loads per lane).
__device__ void block_append(const BlockGraph& g, TempU& tmp, uint32_t* srow, uint32_t* sent,
                             const uint32_t* src, const uint32_t* dst, const float* w,
                             bool take, uint32_t t, uint32_t span, int bits,
                             uint32_t* __restrict__ touched, uint32_t* __restrict__ ntouched) {
    const uint32_t pad = row_pad(bits);
    sort_rows(tmp, srow, sent, take ? src[t] : pad, take ? dst[t] : pad, t, span, bits);
    const int appended = __syncthreads_count(take ? 1 : 0);
    bool h[2];
    uint32_t r[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t p = 2 * t + j;
        uint32_t end = 0;
        h[j] = group_head(srow, p, 2 * span, pad, end);
        r[j] = h[j] ? srow[p] : pad;
        if (!h[j]) continue;
        const uint32_t cnt = end - p;
        uint2 row = g.rows[r[j]];
        const uint32_t d0 = row.y;
        const uint32_t c = g.cap[r[j]];
        if ((unsigned long long)d0 + cnt > c) {  // relocate to the arena tail, order kept
            unsigned long long ncap = 2ull * c;
            ncap = max(ncap, (unsigned long long)d0 + cnt + g.slack);
            ncap = max(ncap, 4ull);
            const unsigned long long ns = atomicAdd(&g.meta[0], ncap);
            if (ns + ncap > g.cap_slots) {
                atomicOr((uint32_t*)(g.meta + 2), 1u);
                continue;
            }
            for (uint32_t z = 0; z < d0; ++z) g.adj[ns + z] = g.adj[row.x + z];
            row.x = (uint32_t)ns;
            g.cap[r[j]] = (uint32_t)ncap;
        }
        for (uint32_t z = p; z < end; ++z) {
            const uint32_t x = sent[z], e = x >> 1;
            const uint32_t nbr = (x & 1u) ? src[e] : dst[e];
            g.adj[row.x + d0 + (z - p)] = make_uint4(nbr, __float_as_uint(w[e]), 0u, 0u);
        }
        g.rows[r[j]] = make_uint2(row.x, d0 + cnt);
    }
    if (t == 0 && appended) atomicAdd(&g.meta[1], (unsigned long long)appended);
    list_rows(tmp.scan, h, r, t, touched, ntouched);
    __syncthreads();
}

// Deletes edge t (lane t < span: src[t]-dst[t]) from both rows when `take`;
// each row's removals in entry order (2t: src[t]'s row, 2t+1: dst[t]'s),
// each by swap-with-last (the last live slot moves into the hole with its
// copy of its neighbour's descriptor), by the lane holding the row's first
// sorted entry.  A removal the row cannot make sets error bit 2.  Lists the
// touched rows.
__device__ void block_delete(const BlockGraph& g, TempU& tmp, uint32_t* srow, uint32_t* sent,
                             const uint32_t* src, const uint32_t* dst, bool take, uint32_t t,
                             uint32_t span, int bits, uint32_t* __restrict__ touched,
                             uint32_t* __restrict__ ntouched) {
    const uint32_t pad = row_pad(bits);
    sort_rows(tmp, srow, sent, take ? src[t] : pad, take ? dst[t] : pad, t, span, bits);
    const int deleted = __syncthreads_count(take ? 1 : 0);
    bool h[2];
    uint32_t r[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const uint32_t p = 2 * t + j;
        uint32_t end = 0;
        h[j] = group_head(srow, p, 2 * span, pad, end);
        r[j] = h[j] ? srow[p] : pad;
        if (!h[j]) continue;
        uint2 row = g.rows[r[j]];
        for (uint32_t z = p; z < end; ++z) {
            const uint32_t x = sent[z], e = x >> 1;
            const uint32_t nbr = (x & 1u) ? src[e] : dst[e];
            const int s = slot_of(g.adj, row, nbr);
            if (s < 0) {  // not present: the caller's lists were inconsistent
                atomicOr((uint32_t*)(g.meta + 2), 2u);
                continue;
            }
            const uint32_t last = row.y - 1u;
            if ((uint32_t)s != last) g.adj[row.x + s] = g.adj[row.x + last];
            row.y = last;
        }
        g.rows[r[j]] = row;
    }
    if (t == 0 && deleted) atomicAdd(&g.meta[4], (unsigned long long)deleted);
    list_rows(tmp.scan, h, r, t, touched, ntouched);
    __syncthreads();
}

__device__ void add_weights_seq(const BlockGraph& g, uint32_t a, uint32_t b,
                                const uint32_t* idx, uint32_t p0, uint32_t p1,
                                const float* __restrict__ w) {
    const int pa = find_in_row(g.rows, g.adj, a, b), pb = find_in_row(g.rows, g.adj, b, a);
    if (pa < 0 || pb < 0) return;
    uint4* sa = &g.adj[g.rows[a].x + pa];
    uint4* sb = &g.adj[g.rows[b].x + pb];
    float x = __uint_as_float(sa->y);
    for (uint32_t e = p0; e < p1; ++e) x = x + w[idx[e]];
    sa->y = __float_as_uint(x);
    sb->y = __float_as_uint(x);
}

// Lane i of `lanes` copies its share of the report (see Report).
__device__ __forceinline__ void report_out(const Report& r, const int32_t* __restrict__ cls,
                                           uint32_t n, unsigned long long* __restrict__ stats,
                                           uint32_t i, uint32_t lanes) {
    for (uint32_t k = i; k < n; k += lanes) r.cls[k] = cls[k];
    if (i < 4) {
        r.words[i] = stats[i];
        stats[i] = 0;
    } else if (i < 9) {
        r.words[i] = __hip_atomic_load(r.mg + (i - 4), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (i < (uint32_t)kReportWords) {
        r.words[i] = __hip_atomic_load(r.mh + (i - 9), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------- insertions
struct InsShared {
    TempU tmp;
    unsigned long long key[kThreads];  // sorted batch keys
    uint32_t idx[kThreads];            // their batch positions
    uint32_t srow[2 * kThreads];       // appends sorted by row: row, entry
    uint32_t sent[2 * kThreads];
    uint32_t src[kThreads], dst[kThreads];
    float w[kThreads];
    uint8_t keep[kThreads];            // by batch position
    uint8_t mode[kThreads];            // by sorted position (heads)
};

__global__ __launch_bounds__(kThreads) void k_small_ins_pre(SmallInsArgs a) {
    __shared__ InsShared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t n = a.n;
    // ---- sort the batch's (min,max) keys, stable ----
    uint32_t ut = 0, vt = 0;
    float wt = 0.0f;
    unsigned long long kt = kPad64;
    if (t < n) {
        ut = a.u[t];
        vt = a.v[t];
        wt = a.w[t];
        kt = edge_key(ut, vt);
    }
    sh.keep[t] = 0;
    sh.mode[t] = 0;
    sort_keys64(sh.tmp, sh.key, sh.idx, kt, t, n, a.key_bits);
    // ---- classify each distinct edge, occurrences in batch order ----
    const unsigned long long k = sh.key[t];
    if (t < n && k != kPad64 && (t == 0 || sh.key[t - 1] != k)) {
        uint32_t q = t + 1;
        while (q < n && sh.key[q] == k) ++q;
        const uint32_t lo = (uint32_t)(k >> 32), hi = (uint32_t)k;
        const bool in_g = holds(a.G.rows, a.G.adj, lo, hi);
        for (uint32_t e = t; e < q; ++e) {
            const bool first_new = !in_g && e == t;
            a.cls[sh.idx[e]] = first_new ? 0 : kOutWeight;
            sh.keep[sh.idx[e]] = first_new ? 1 : 0;
        }
        if (in_g) sh.mode[t] = 1 | (holds(a.H.rows, a.H.adj, lo, hi) ? 4 : 0);
        else if (q - t > 1) sh.mode[t] = 2;
    } else if (t < n && k == kPad64) {
        a.cls[sh.idx[t]] = kOutInvalid;
    }
    sh.src[t] = ut;
    sh.dst[t] = vt;
    sh.w[t] = wt;
    __syncthreads();
    // ---- new edges into G, in batch order per row ----
    const bool take = t < n && sh.keep[t];
    block_append(a.G, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, sh.w, take, t, n, a.row_bits,
                 a.touched, a.cnt + 4);
    // ---- weight updates (after the appends: a repeated new edge adds to it) ----
    const uint8_t md = sh.mode[t];
    if (md) {
        uint32_t q = t + 1;
        while (q < n && sh.key[q] == k) ++q;
        const uint32_t lo = (uint32_t)(k >> 32), hi = (uint32_t)k;
        const uint32_t e0 = (md & 1) ? t : t + 1;
        add_weights_seq(a.G, lo, hi, sh.idx, e0, q, a.w);
        if (md & 4) add_weights_seq(a.H, lo, hi, sh.idx, e0, q, a.w);
    }
    // ---- compact the kept edges (batch order) for the walk ----
    uint32_t pos = 0, total = 0;
    Scan(sh.tmp.scan).ExclusiveSum(take ? 1u : 0u, pos, total);
    if (take) {
        a.cu[pos] = ut;
        a.cv[pos] = vt;
        a.cw[pos] = wt;
        a.ceid[pos] = a.base + t;
        a.sel[pos] = t;
    }
    if (t == 0) a.cnt[0] = total;
}

__global__ __launch_bounds__(kThreads) void k_small_ins_post(SmallInsArgs a) {
    __shared__ InsShared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t m = a.cnt[0];
    const bool live = t < m;
    const bool add = live && a.found[t] == -1;
    sh.src[t] = live ? a.cu[t] : 0u;
    sh.dst[t] = live ? a.cv[t] : 0u;
    sh.w[t] = live ? a.cw[t] : 0.0f;
    __syncthreads();
    block_append(a.H, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, sh.w, add, t, m, a.row_bits,
                 a.touched, a.cnt + 4);
    if (live) a.cls[a.sel[t]] = add ? kOutAdded : kOutFound;
    __syncthreads();
    report_out(a.rep, a.cls, a.n, a.stats, t, kThreads);
}

// ----------------------------------------------------------------- deletions
struct DelPreShared {
    TempU tmp;
    unsigned long long key[kThreads];
    uint32_t idx[kThreads];
    uint32_t srow[2 * kThreads];
    uint32_t sent[2 * kThreads];
    uint32_t src[kThreads], dst[kThreads];
    uint8_t fg[kThreads], fh[kThreads];  // by batch position: delete from G / from H
};

__global__ __launch_bounds__(kThreads) void k_small_del_pre(SmallDelArgs a) {
    __shared__ DelPreShared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t n = a.n;
    uint32_t ut = 0, vt = 0;
    unsigned long long kt = kPad64;
    if (t < n) {
        ut = a.u[t];
        vt = a.v[t];
        kt = edge_key(ut, vt);
    }
    sh.fg[t] = 0;
    sh.fh[t] = 0;
    sort_keys64(sh.tmp, sh.key, sh.idx, kt, t, n, a.key_bits);
    // ---- classify: the first occurrence of an edge G holds is deleted ----
    const unsigned long long k = sh.key[t];
    if (t < n && (t == 0 || sh.key[t - 1] != k)) {
        uint32_t q = t + 1;
        while (q < n && sh.key[q] == k) ++q;
        bool in_g = false, in_h = false;
        if (k != kPad64) {
            const uint32_t lo = (uint32_t)(k >> 32), hi = (uint32_t)k;
            in_g = holds(a.G.rows, a.G.adj, lo, hi);
            in_h = in_g && holds(a.H.rows, a.H.adj, lo, hi);
        }
        for (uint32_t e = t; e < q; ++e) {
            const uint32_t i = sh.idx[e];
            const bool first = e == t && in_g;
            a.cls[i] = k == kPad64 ? kOutInvalid : !first ? kOutSkip : in_h ? 0 : kOutGOnly;
            sh.fg[i] = first ? 1 : 0;
            sh.fh[i] = first && in_h ? 1 : 0;
        }
    }
    sh.src[t] = ut;
    sh.dst[t] = vt;
    __syncthreads();
    const bool dg = t < n && sh.fg[t], dh = t < n && sh.fh[t];
    block_delete(a.G, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, dg, t, n, a.row_bits,
                 a.touched_g, a.cnt + 4);
    block_delete(a.H, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, dh, t, n, a.row_bits,
                 a.touched_h, a.cnt + 5);
    // ---- compact the edges H held (batch order), as (min, max), for the walk ----
    uint32_t pos = 0, total = 0;
    Scan(sh.tmp.scan).ExclusiveSum(dh ? 1u : 0u, pos, total);
    if (dh) {
        a.cu[pos] = min(ut, vt);
        a.cv[pos] = max(ut, vt);
        a.ceid[pos] = a.base + t;
        a.sel[pos] = t;
    }
    if (t == 0) a.cnt[0] = total;
}

struct DelPostShared {
    TempU tmp;
    uint32_t np[kThreads];                    // no-path list: kept-edge index
    uint32_t cu[kSmallCand], cv[kSmallCand];  // candidates, in the general list's order
    float cw[kSmallCand];
    uint8_t add[kSmallCand];                  // by candidate: append to H
    unsigned long long key[kSmallCand];       // candidate keys, sorted
    uint32_t idx[kSmallCand];
    uint32_t srow[2 * kSmallCand];
    uint32_t sent[2 * kSmallCand];
};

__global__ __launch_bounds__(kThreads) void k_small_del_post(SmallDelArgs a) {
    __shared__ DelPostShared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t m = a.cnt[0];
    const int32_t f = t < m ? a.found[t] : 0;
    const uint32_t pc = f > 0 ? (uint32_t)f : 0u;
    const bool np = f == -1;
    uint32_t poff = 0, hops = 0, k = 0, nnp = 0;
    Scan(sh.tmp.scan).ExclusiveSum(pc, poff, hops);
    __syncthreads();
    Scan(sh.tmp.scan).ExclusiveSum(np ? 1u : 0u, k, nnp);
    const uint32_t hw = a.hw;
    const uint64_t n3 = 2ull * nnp * hw;  // heuristic walks: two endpoints per no-path edge
    const uint64_t ncand = hops + 2 * n3;
    if (ncand > kSmallCand) {
        // the general recovery takes over: its lists, as its kernels leave them
        if (t < m) {
            a.pc[t] = pc;
            a.poff[t] = poff;
        }
        if (np) {
            a.ep[2 * k] = a.cu[t];
            a.ep[2 * k + 1] = a.cv[t];
            a.ekey[2 * k] = a.ceid[t] << 1;
            a.ekey[2 * k + 1] = (a.ceid[t] << 1) | 1ull;
        }
        if (t == 0) {
            a.cnt[1] = nnp;
            a.cnt[2] = 2 * nnp;
            a.cnt[3] = hops;
            a.cnt[6] = 0;  // no rows for the refresh after this kernel
            a.sizes[0] = m;
            a.sizes[1] = nnp;
            a.sizes[2] = 2 * nnp;
            a.sizes[3] = hops;
            a.sizes[4] = 1;
        }
        return;
    }
    if (np) sh.np[k] = t;
    sh.add[t] = 0;
    // ---- candidate list: path hops (kept-edge order), then heuristic hops ----
    if (f > 0) {
        const uint32_t* pj = a.path + (uint64_t)t * a.max_steps;
        for (int q = 1; q <= f; ++q) {
            sh.cu[poff + q - 1] = pj[q - 1];
            sh.cv[poff + q - 1] = pj[q];
        }
    }
    __syncthreads();
    for (uint32_t g = t; g < n3; g += kThreads) {
        const uint32_t i = g / hw, jw = g % hw;  // endpoint i = 2*(no-path index) + side
        const uint32_t e = sh.np[i >> 1];
        const uint32_t u = (i & 1u) ? a.cv[e] : a.cu[e];
        const uint64_t key = (a.ceid[e] << 1) | (uint64_t)(i & 1u);
        const uint3 w3 = heuristic_walk(a.G.rows, a.G.adj, u, key, jw, a.seed);
        sh.cu[hops + 2 * g] = w3.x;
        sh.cv[hops + 2 * g] = w3.y;
        sh.cu[hops + 2 * g + 1] = w3.y;
        sh.cv[hops + 2 * g + 1] = w3.z;
    }
    __syncthreads();
    // ---- first occurrence of each candidate that H lacks, with G's weight ----
    const uint32_t nc = (uint32_t)ncand;
    const unsigned long long kt = t < nc ? edge_key(sh.cu[t], sh.cv[t]) : kPad64;
    sort_keys64(sh.tmp, sh.key, sh.idx, kt, t, nc, a.key_bits);
    const unsigned long long ks = sh.key[t];
    if (t < nc && ks != kPad64 && (t == 0 || sh.key[t - 1] != ks)) {
        const uint32_t lo = (uint32_t)(ks >> 32), hi = (uint32_t)ks;
        if (!holds(a.H.rows, a.H.adj, lo, hi)) {
            const int s = find_in_row(a.G.rows, a.G.adj, lo, hi);
            if (s >= 0) {  // always: path and heuristic hops are edges of G
                const uint32_t c = sh.idx[t];
                sh.add[c] = 1;
                sh.cw[c] = __uint_as_float(a.G.adj[a.G.rows[lo].x + s].y);
            }
        }
    }
    __syncthreads();
    const bool take = t < nc && sh.add[t];
    block_append(a.H, sh.tmp, sh.srow, sh.sent, sh.cu, sh.cv, sh.cw, take, t, nc, a.row_bits,
                 a.touched_x, a.cnt + 6);
    if (t < m) a.cls[a.sel[t]] = np ? kOutHeuristic : kOutRecovered;
    if (t == 0) {
        a.sizes[4] = 0;
        a.stats[3] = ncand;  // reported as the batch's candidate count
    }
    __syncthreads();
    report_out(a.rep, a.cls, a.n, a.stats, t, kThreads);
}

// lpr lanes per listed row: every slot pointing to the row gets its final
// descriptor (a lane per neighbour).  blockIdx.y selects one of two lists.
struct RefreshList {
    const uint2* rows;
    uint4* adj;
    const uint32_t* list;
    const uint32_t* count;
    uint32_t lpr;  // power of two, 4..64
};

__global__ __launch_bounds__(256) void k_refresh_rows(RefreshList l0, RefreshList l1) {
    const RefreshList l = blockIdx.y ? l1 : l0;
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t i = g / l.lpr;
    if (i >= *l.count) return;
    const uint32_t lane = g & (l.lpr - 1u);
    const uint32_t r = l.list[i];
    const uint2 rr = l.rows[r];
    for (uint32_t x = lane; x < rr.y; x += l.lpr) {
        const uint32_t nb = l.adj[rr.x + x].x;
        refresh_pointers(l.adj, l.rows[nb], r, rr);
    }
}

__global__ __launch_bounds__(256) void k_report(Report r, const int32_t* __restrict__ cls,
                                                uint32_t n, unsigned long long* __restrict__ stats) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) r.cls[i] = cls[i];
    if (blockIdx.x == 0) report_out(r, cls, 0, stats, threadIdx.x, 256);
}

// Refreshes the rows of list a (and b) in one launch; max_rows bounds both.
void refresh(const RefreshList& a, const RefreshList* b, uint32_t max_rows, hipStream_t s) {
    const uint32_t lpr = b ? std::max(a.lpr, b->lpr) : a.lpr;
    const uint32_t blocks = (uint32_t)(((uint64_t)max_rows * lpr + 255) / 256);
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_refresh_rows, dim3(blocks, b ? 2 : 1), dim3(256), 0, s, a, b ? *b : a);
}

}  // namespace

hipError_t launch_small_ins_pre(const SmallInsArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_ins_pre, dim3(1), dim3(kThreads), 0, s, a);
    refresh({a.G.rows, a.G.adj, a.touched, a.cnt + 4, a.G.lpr}, nullptr, 2 * a.n, s);
    return hipGetLastError();
}

hipError_t launch_small_ins_post(const SmallInsArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_ins_post, dim3(1), dim3(kThreads), 0, s, a);
    refresh({a.H.rows, a.H.adj, a.touched, a.cnt + 4, a.H.lpr}, nullptr, 2 * a.n, s);
    return hipGetLastError();
}

hipError_t launch_small_del_pre(const SmallDelArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_del_pre, dim3(1), dim3(kThreads), 0, s, a);
    const RefreshList h{a.H.rows, a.H.adj, a.touched_h, a.cnt + 5, a.H.lpr};
    refresh({a.G.rows, a.G.adj, a.touched_g, a.cnt + 4, a.G.lpr}, &h, 2 * a.n, s);
    return hipGetLastError();
}

hipError_t launch_small_del_post(const SmallDelArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_del_post, dim3(1), dim3(kThreads), 0, s, a);
    refresh({a.H.rows, a.H.adj, a.touched_x, a.cnt + 6, a.H.lpr}, nullptr, 2 * kSmallCand, s);
    return hipGetLastError();
}

hipError_t launch_report(const Report& r, const int32_t* cls, uint32_t n,
                         unsigned long long* stats, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, (n + 255) / 256);
    hipLaunchKernelGGL(k_report, dim3(blocks), dim3(256), 0, s, r, cls, n, stats);
    return hipGetLastError();
}

}  // namespace dyg
