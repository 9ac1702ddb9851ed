// stream_small.hip -- batches of up to kSmallBatch updates in one-workgroup
// kernels around the walk.
//
// The general pipeline of a batch (stream.hip) is thirty to fifty launches;
// for a batch of a few hundred edges every kernel runs for a few
// microseconds and the launches themselves set the pace.  Here one
// workgroup of 1024 lanes (two updates per lane) does a phase with the batch
// in LDS (up to about 130 KB of the CU's 160 KB).
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
static_assert(2 * kThreads == kSmallBatch && 2 * kThreads == kSmallCand,
              "two items per lane at most");
constexpr uint32_t kPadRow = 0xFFFFFFFFu;
constexpr unsigned long long kPad64 = ~0ull;
constexpr int32_t kOutWeight = 1, kOutFound = 2, kOutAdded = 3, kOutSkip = 4, kOutGOnly = 5,
                  kOutRecovered = 6, kOutHeuristic = 7, kOutInvalid = 8;

using Scan = hipcub::BlockScan<uint32_t, kThreads>;

// Up to this many items a sort ranks by counting (one wave, a few hundred
// LDS reads); above it, a block radix sort (O(items) work per pass: counting
// would be items^2 compares on one CU).
constexpr uint32_t kCountMax = 64;

// A phase's shape: IPT batch items (edges) per lane, 2*IPT row entries per
// lane.  One item per lane keeps the sorts and LDS small for the common small
// batches; two per lane take batches of up to 2048.
template <int IPT>
struct Shape {
    static constexpr int kIPT = IPT;
    static constexpr int kEPT = 2 * IPT;
    static constexpr uint32_t kItems = kThreads * IPT;
    using SortKeys64 = hipcub::BlockRadixSort<unsigned long long, kThreads, IPT, uint32_t>;
    using SortRows = hipcub::BlockRadixSort<uint32_t, kThreads, 2 * IPT, uint32_t>;
    // Phase-local scratch, one phase after another (a barrier between).
    union TempU {
        typename SortKeys64::TempStorage k64;
        typename SortRows::TempStorage rows;
        typename Scan::TempStorage scan;
        unsigned long long kin[kItems];  // counting sort of keys: keys in input order
        uint32_t rk[2 * kItems];         // counting sort of rows: rows in entry order
    };
};

__device__ __forceinline__ int find_in_row(const uint2* rows, const uint4* adj, uint32_t a,
                                           uint32_t b) {
    return slot_of(adj, rows[a], b);
}

__device__ __forceinline__ unsigned long long edge_key(uint32_t u, uint32_t v) {
    const uint32_t lo = min(u, v), hi = max(u, v);
    return lo == hi ? kPad64 : ((unsigned long long)lo << 32) | hi;
}

__device__ __forceinline__ uint32_t row_pad(int row_bits) {
    return row_bits >= 32 ? 0xFFFFFFFFu : (1u << row_bits) - 1u;
}

// Stable sort of n keys; lane t holds items kIPT*t + k (kPad64 past n; bits
// [0, bits) order them): key[rank] = kt, idx[rank] = item.  Items past n keep
// their place.
template <int IPT>
__device__ void sort_keys64(typename Shape<IPT>::TempU& tmp, unsigned long long* key,
                            uint32_t* idx, const unsigned long long (&kt)[IPT], uint32_t t,
                            uint32_t n, int bits) {
    constexpr int kIPT = IPT;
    if (n <= kCountMax) {
#pragma unroll
        for (int k = 0; k < kIPT; ++k) tmp.kin[kIPT * t + k] = kt[k];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            const uint32_t p = kIPT * t + k;
            uint32_t rank = p;
            if (p < n) {
                rank = 0;
                for (uint32_t j = 0; j < n; ++j) {
                    const unsigned long long kj = tmp.kin[j];
                    rank += (kj < kt[k] || (kj == kt[k] && j < p)) ? 1u : 0u;
                }
            }
            key[rank] = kt[k];
            idx[rank] = p;
        }
    } else {
        unsigned long long kk[kIPT];
        uint32_t v[kIPT];
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            kk[k] = kt[k];
            v[k] = kIPT * t + k;
        }
        typename Shape<IPT>::SortKeys64(tmp.k64).Sort(kk, v, 0, bits);
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            key[kIPT * t + k] = kk[k];
            idx[kIPT * t + k] = v[k];
        }
    }
    __syncthreads();
}

// Stable sort by row of the row entries x = kEPT*t + q (row r[q]; entries
// of the first `span` edges count, pads last): srow/sent[p] = row and entry
// at sorted position p < 2*span.
template <int IPT>
__device__ void sort_rows(typename Shape<IPT>::TempU& tmp, uint32_t* srow, uint32_t* sent,
                          const uint32_t (&r)[2 * IPT], uint32_t t, uint32_t span, int bits) {
    constexpr int kEPT = 2 * IPT;
    const uint32_t ne = 2 * span;
    if (span <= kCountMax) {
#pragma unroll
        for (int q = 0; q < kEPT; ++q) tmp.rk[kEPT * t + q] = r[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kEPT; ++q) {
            const uint32_t x = kEPT * t + q;
            if (x >= ne) continue;
            uint32_t a = 0;
            for (uint32_t y = 0; y < ne; ++y) {
                const uint32_t ry = tmp.rk[y];
                a += (ry < r[q] || (ry == r[q] && y < x)) ? 1u : 0u;
            }
            srow[a] = r[q];
            sent[a] = x;
        }
    } else {
        uint32_t k[kEPT], v[kEPT];
#pragma unroll
        for (int q = 0; q < kEPT; ++q) {
            k[q] = r[q];
            v[q] = kEPT * t + q;
        }
        typename Shape<IPT>::SortRows(tmp.rows).Sort(k, v, 0, bits);
#pragma unroll
        for (int q = 0; q < kEPT; ++q) {
            srow[kEPT * t + q] = k[q];
            sent[kEPT * t + q] = v[q];
        }
    }
    __syncthreads();
}

// Lists the rows r[q] flagged h[q] by this lane, in any order, and their count.
template <int kEPT>
__device__ __forceinline__ void list_rows(Scan::TempStorage& scan, const bool (&h)[kEPT],
                                          const uint32_t (&r)[kEPT], uint32_t t,
                                          uint32_t* __restrict__ touched,
                                          uint32_t* __restrict__ ntouched) {
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < kEPT; ++q) c += h[q] ? 1u : 0u;
    uint32_t pos = 0, total = 0;
    Scan(scan).ExclusiveSum(c, pos, total);
#pragma unroll
    for (int q = 0; q < kEPT; ++q)
        if (h[q]) touched[pos++] = r[q];
    if (t == 0) *ntouched = total;
}

// Row groups of a sorted entry list: sorted position p starts a group when
// it holds a row and differs from p-1; the group runs to `end`.
__device__ __forceinline__ bool group_head(const uint32_t* srow, uint32_t p, uint32_t lim,
                                           uint32_t pad, uint32_t& end) {
    if (p >= lim) return false;
    const uint32_t r = srow[p];
    if (r == pad || (p > 0 && srow[p - 1] == r)) return false;
    end = p + 1;
    while (end < lim && srow[end] == r) ++end;
    return true;
}

// Appends edges e = kIPT*t + k (e < span: src[e]-dst[e] with weight w[e])
// to both rows when take[k].  Entry 2e is the append to src[e]'s row, 2e+1
// to dst[e]'s; each row's appends keep entry order (the general pipeline's
// order), written by the lane holding the row's first sorted entry.  Returns
// after the rows are final, with the touched rows listed for k_refresh_rows,
// which rewrites the neighbours' copies of their descriptors across the
// whole GPU (in one workgroup that refresh would be deg^2 dependent loads
// per lane).
template <int IPT>
__device__ void block_append(const BlockGraph& g, typename Shape<IPT>::TempU& tmp,
                             uint32_t* srow, uint32_t* sent, const uint32_t* src,
                             const uint32_t* dst, const float* w, const bool (&take)[IPT],
                             uint32_t t, uint32_t span, int bits, uint32_t* __restrict__ touched,
                             uint32_t* __restrict__ ntouched) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
    const uint32_t pad = row_pad(bits);
    uint32_t re[kEPT];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t e = kIPT * t + k;
        re[2 * k] = take[k] ? src[e] : pad;
        re[2 * k + 1] = take[k] ? dst[e] : pad;
        mine += take[k] ? 1u : 0u;
    }
    sort_rows<IPT>(tmp, srow, sent, re, t, span, bits);
    bool h[kEPT];
    uint32_t r[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        uint32_t end = 0;
        h[q] = group_head(srow, p, 2 * span, pad, end);
        r[q] = h[q] ? srow[p] : pad;
        if (!h[q]) continue;
        const uint32_t cnt = end - p;
        uint2 row = g.rows[r[q]];
        const uint32_t d0 = row.y;
        const uint32_t c = g.cap[r[q]];
        if ((unsigned long long)d0 + cnt > c) {  // relocate to the arena tail, order kept
            unsigned long long ncap = 2ull * c;
            ncap = max(ncap, (unsigned long long)d0 + cnt + g.slack);
            ncap = max(ncap, 4ull);
            const unsigned long long ns = atomicAdd(&g.meta[0], ncap);
            if (ns + ncap > g.cap_slots) {
                atomicOr((uint32_t*)(g.meta + 2), 1u);
                continue;
            }
            copy_slots(g.adj + ns, g.adj + row.x, d0);
            row.x = (uint32_t)ns;
            g.cap[r[q]] = (uint32_t)ncap;
        }
        for (uint32_t z = p; z < end; ++z) {
            const uint32_t x = sent[z], e = x >> 1;
            const uint32_t nbr = (x & 1u) ? src[e] : dst[e];
            g.adj[row.x + d0 + (z - p)] = make_uint4(nbr, __float_as_uint(w[e]), 0u, 0u);
        }
        g.rows[r[q]] = make_uint2(row.x, d0 + cnt);
    }
    if (mine) atomicAdd(&g.meta[1], (unsigned long long)mine);
    list_rows(tmp.scan, h, r, t, touched, ntouched);
    __syncthreads();
}

// Deletes edges e = kIPT*t + k (e < span: src[e]-dst[e]) from both rows
// when take[k]; each row's removals in entry order (2e: src[e]'s row, 2e+1:
// dst[e]'s), each by swap-with-last (the last live slot moves into the hole
// with its copy of its neighbour's descriptor), by the lane holding the
// row's first sorted entry.  A removal the row cannot make sets error bit 2.
// Lists the touched rows.
template <int IPT>
__device__ void block_delete(const BlockGraph& g, typename Shape<IPT>::TempU& tmp,
                             uint32_t* srow, uint32_t* sent, const uint32_t* src,
                             const uint32_t* dst, const bool (&take)[IPT], uint32_t t,
                             uint32_t span, int bits, uint32_t* __restrict__ touched,
                             uint32_t* __restrict__ ntouched) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
    const uint32_t pad = row_pad(bits);
    uint32_t re[kEPT];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t e = kIPT * t + k;
        re[2 * k] = take[k] ? src[e] : pad;
        re[2 * k + 1] = take[k] ? dst[e] : pad;
        mine += take[k] ? 1u : 0u;
    }
    sort_rows<IPT>(tmp, srow, sent, re, t, span, bits);
    bool h[kEPT];
    uint32_t r[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        uint32_t end = 0;
        h[q] = group_head(srow, p, 2 * span, pad, end);
        r[q] = h[q] ? srow[p] : pad;
        if (!h[q]) continue;
        uint2 row = g.rows[r[q]];
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
        g.rows[r[q]] = row;
    }
    if (mine) atomicAdd(&g.meta[4], (unsigned long long)mine);
    list_rows(tmp.scan, h, r, t, touched, ntouched);
    __syncthreads();
}

// A batch's orders computed ahead of it (BatchOrder) into LDS: the sorted
// keys and positions (pads past n, as sort_keys64 leaves them) and the row
// entries sorted by row (p < 2n).
template <int IPT>
__device__ void load_order(const BatchOrder& o, unsigned long long* key, uint32_t* idx,
                           uint32_t* srow, uint32_t* sent, uint32_t t, uint32_t n) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        key[p] = p < n ? o.key[p] : kPad64;
        idx[p] = p < n ? o.idx[p] : p;
    }
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        if (p < 2 * n) {
            srow[p] = o.row[p];
            sent[p] = o.ent[p];
        }
    }
    __syncthreads();
}

// block_append over presorted entries: srow/sent hold every entry of the
// first `span` edges (entry 2e: src[e]'s row, 2e+1: dst[e]'s) sorted by row,
// entry order within a row; the lane at the start of a row's run appends the
// run's entries of edges with tk[e] set, in order.  The rows, their appends
// and their order are those of block_append over the taken entries alone.
template <int IPT>
__device__ void block_append_sorted(const BlockGraph& g, Scan::TempStorage& scan,
                                    const uint32_t* srow, const uint32_t* sent,
                                    const uint32_t* src, const uint32_t* dst, const float* w,
                                    const uint8_t* tk, uint32_t t, uint32_t span,
                                    uint32_t* __restrict__ touched,
                                    uint32_t* __restrict__ ntouched) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
    const uint32_t lim = 2 * span;
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t e = kIPT * t + k;
        mine += (e < span && tk[e]) ? 1u : 0u;
    }
    bool h[kEPT];
    uint32_t r[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        h[q] = false;
        r[q] = 0;
        if (p >= lim) continue;
        const uint32_t rr = srow[p];
        if (p > 0 && srow[p - 1] == rr) continue;  // not the start of its row's run
        uint32_t end = p + 1, cnt = 0;
        while (end < lim && srow[end] == rr) ++end;
        for (uint32_t z = p; z < end; ++z) cnt += tk[sent[z] >> 1] ? 1u : 0u;
        if (!cnt) continue;
        h[q] = true;
        r[q] = rr;
        uint2 row = g.rows[rr];
        const uint32_t d0 = row.y;
        const uint32_t c = g.cap[rr];
        if ((unsigned long long)d0 + cnt > c) {  // relocate to the arena tail, order kept
            unsigned long long ncap = 2ull * c;
            ncap = max(ncap, (unsigned long long)d0 + cnt + g.slack);
            ncap = max(ncap, 4ull);
            const unsigned long long ns = atomicAdd(&g.meta[0], ncap);
            if (ns + ncap > g.cap_slots) {
                atomicOr((uint32_t*)(g.meta + 2), 1u);
                continue;
            }
            copy_slots(g.adj + ns, g.adj + row.x, d0);
            row.x = (uint32_t)ns;
            g.cap[rr] = (uint32_t)ncap;
        }
        uint32_t o = row.x + d0;
        for (uint32_t z = p; z < end; ++z) {
            const uint32_t x = sent[z], e = x >> 1;
            if (!tk[e]) continue;
            const uint32_t nbr = (x & 1u) ? src[e] : dst[e];
            g.adj[o++] = make_uint4(nbr, __float_as_uint(w[e]), 0u, 0u);
        }
        g.rows[rr] = make_uint2(row.x, d0 + cnt);
    }
    if (mine) atomicAdd(&g.meta[1], (unsigned long long)mine);
    list_rows(scan, h, r, t, touched, ntouched);
    __syncthreads();
}

// block_delete over presorted entries (as block_append_sorted): each row's
// removals of edges with tk[e] set, in entry order, by swap-with-last.
template <int IPT>
__device__ void block_delete_sorted(const BlockGraph& g, Scan::TempStorage& scan,
                                    const uint32_t* srow, const uint32_t* sent,
                                    const uint32_t* src, const uint32_t* dst, const uint8_t* tk,
                                    uint32_t t, uint32_t span, uint32_t* __restrict__ touched,
                                    uint32_t* __restrict__ ntouched) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
    const uint32_t lim = 2 * span;
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t e = kIPT * t + k;
        mine += (e < span && tk[e]) ? 1u : 0u;
    }
    bool h[kEPT];
    uint32_t r[kEPT];
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        h[q] = false;
        r[q] = 0;
        if (p >= lim) continue;
        const uint32_t rr = srow[p];
        if (p > 0 && srow[p - 1] == rr) continue;
        uint32_t end = p + 1;
        bool any = false;
        while (end < lim && srow[end] == rr) ++end;
        for (uint32_t z = p; z < end; ++z) any = any || tk[sent[z] >> 1];
        if (!any) continue;
        h[q] = true;
        r[q] = rr;
        uint2 row = g.rows[rr];
        for (uint32_t z = p; z < end; ++z) {
            const uint32_t x = sent[z], e = x >> 1;
            if (!tk[e]) continue;
            const uint32_t nbr = (x & 1u) ? src[e] : dst[e];
            const int sl = slot_of(g.adj, row, nbr);
            if (sl < 0) {  // not present: the caller's lists were inconsistent
                atomicOr((uint32_t*)(g.meta + 2), 2u);
                continue;
            }
            const uint32_t last = row.y - 1u;
            if ((uint32_t)sl != last) g.adj[row.x + sl] = g.adj[row.x + last];
            row.y = last;
        }
        g.rows[rr] = row;
    }
    if (mine) atomicAdd(&g.meta[4], (unsigned long long)mine);
    list_rows(scan, h, r, t, touched, ntouched);
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

// ---------------------------------------------------------------- insertions
template <int IPT>
struct InsShared {
    static constexpr uint32_t kItems = Shape<IPT>::kItems;
    typename Shape<IPT>::TempU tmp;
    unsigned long long key[kItems];  // sorted batch keys
    uint32_t idx[kItems];            // their batch positions
    uint32_t srow[2 * kItems];       // appends sorted by row: row, entry
    uint32_t sent[2 * kItems];
    uint32_t src[kItems], dst[kItems];
    float w[kItems];
    uint8_t keep[kItems];            // by batch position
    uint8_t mode[kItems];            // by sorted position (heads)
};

// A captured step graph's batch values (StepDesc), read once at entry.  The
// host rewrites the pinned descriptor and inputs before every replay, and
// pinned memory may be cached in an XCD's L2, where the previous batch's
// lines could outlive a replay: the runtime's fences between the nodes of a
// graph need only reach device scope.  So the kernel takes a system-scope
// acquire first (L1 and L2 invalidation) and reads the descriptor with vector
// loads (volatile: not through the scalar cache).
struct DescVals {
    const uint32_t* u;
    const uint32_t* v;
    const float* w;
    int32_t* cls;
    unsigned long long* words;
    uint64_t base;
    uint32_t n;
};
__device__ inline DescVals read_desc(const StepDesc* p) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    const volatile StepDesc* d = p;
    DescVals r;
    r.u = d->u;
    r.v = d->v;
    r.w = d->w;
    r.cls = d->rep_cls;
    r.words = d->rep_words;
    r.base = d->base;
    r.n = d->n;
    return r;
}
__device__ inline void take_desc(SmallInsArgs& a) {
    if (!a.desc) return;
    const DescVals d = read_desc(a.desc);
    a.u = d.u;
    a.v = d.v;
    a.w = d.w;
    a.n = d.n;
    a.base = d.base;
    a.rep.cls = d.cls;
    a.rep.words = d.words;
}
__device__ inline void take_desc(SmallDelArgs& a) {
    if (!a.desc) return;
    const DescVals d = read_desc(a.desc);
    a.u = d.u;
    a.v = d.v;
    a.n = d.n;
    a.base = d.base;
    a.rep.cls = d.cls;
    a.rep.words = d.words;
}

// Device inputs (vcheck != 0): a batch holding an id >= vcheck is rejected
// before it changes anything: nothing kept for the walk, no touched rows for
// the refreshes, and the report's flag word set.  Uniform across the block.
template <int IPT, class Args>
__device__ bool reject_batch(const uint32_t (&ut)[IPT], const uint32_t (&vt)[IPT], uint32_t n,
                             const Args& a, uint32_t t) {
    bool bad = false;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint32_t i = IPT * t + k;
        if (i < n && (ut[k] >= a.vcheck || vt[k] >= a.vcheck)) bad = true;
    }
    if (!__syncthreads_or(bad)) return false;
    if (t == 0) {
        a.cnt[0] = 0;  // kept edges (the walk's list length)
        a.cnt[4] = 0;  // rows listed for the refreshes
        a.cnt[5] = 0;
        a.rep.words[kReportBad] = 1ull;
    }
    return true;
}

template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_ins_pre(SmallInsArgs a) {
    take_desc(a);
    constexpr int kIPT = IPT;
    __shared__ InsShared<IPT> sh;
    const uint32_t t = threadIdx.x;
    const uint32_t n = a.n;
    // ---- sort the batch's (min,max) keys, stable ----
    uint32_t ut[kIPT], vt[kIPT];
    float wt[kIPT];
    unsigned long long kt[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t i = kIPT * t + k;
        ut[k] = i < n ? a.u[i] : 0u;
        vt[k] = i < n ? a.v[i] : 0u;
        wt[k] = i < n ? a.w[i] : 0.0f;
        kt[k] = i < n ? edge_key(ut[k], vt[k]) : kPad64;
        sh.keep[i] = 0;
        sh.mode[i] = 0;
    }
    if (a.vcheck && reject_batch<IPT>(ut, vt, n, a, t)) return;
    if (a.ord.key)
        load_order<IPT>(a.ord, sh.key, sh.idx, sh.srow, sh.sent, t, n);
    else
        sort_keys64<IPT>(sh.tmp, sh.key, sh.idx, kt, t, n, a.key_bits);
    // ---- classify each distinct edge, occurrences in batch order ----
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        const unsigned long long kp = sh.key[p];
        if (p < n && kp != kPad64 && (p == 0 || sh.key[p - 1] != kp)) {
            uint32_t q = p + 1;
            while (q < n && sh.key[q] == kp) ++q;
            const uint32_t lo = (uint32_t)(kp >> 32), hi = (uint32_t)kp;
            const bool in_g = has_edge(a.G.rows, a.G.adj, lo, hi);
            for (uint32_t e = p; e < q; ++e) {
                const bool first_new = !in_g && e == p;
                a.cls[sh.idx[e]] = first_new ? 0 : kOutWeight;
                sh.keep[sh.idx[e]] = first_new ? 1 : 0;
            }
            if (in_g) sh.mode[p] = 1 | (has_edge(a.H.rows, a.H.adj, lo, hi) ? 4 : 0);
            else if (q - p > 1) sh.mode[p] = 2;
        } else if (p < n && kp == kPad64) {
            a.cls[sh.idx[p]] = kOutInvalid;
        }
        sh.src[kIPT * t + k] = ut[k];
        sh.dst[kIPT * t + k] = vt[k];
        sh.w[kIPT * t + k] = wt[k];
    }
    __syncthreads();
    // ---- new edges into G, in batch order per row ----
    bool take[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) take[k] = kIPT * t + k < n && sh.keep[kIPT * t + k];
    if (a.ord.key)
        block_append_sorted<IPT>(a.G, sh.tmp.scan, sh.srow, sh.sent, sh.src, sh.dst, sh.w,
                                 sh.keep, t, n, a.touched, a.cnt + 4);
    else
        block_append<IPT>(a.G, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, sh.w, take, t, n,
                          a.row_bits, a.touched, a.cnt + 4);
    // ---- weight updates (after the appends: a repeated new edge adds to it) ----
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        const uint8_t md = sh.mode[p];
        if (!md) continue;
        const unsigned long long kp = sh.key[p];
        uint32_t q = p + 1;
        while (q < n && sh.key[q] == kp) ++q;
        const uint32_t lo = (uint32_t)(kp >> 32), hi = (uint32_t)kp;
        const uint32_t e0 = (md & 1) ? p : p + 1;
        add_weights_seq(a.G, lo, hi, sh.idx, e0, q, a.w);
        if (md & 4) add_weights_seq(a.H, lo, hi, sh.idx, e0, q, a.w);
    }
    // ---- compact the kept edges (batch order) for the walk ----
    uint32_t in[kIPT], pos[kIPT], total = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) in[k] = take[k] ? 1u : 0u;
    Scan(sh.tmp.scan).ExclusiveSum(in, pos, total);
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        if (!take[k]) continue;
        const uint32_t i = kIPT * t + k;
        a.cu[pos[k]] = ut[k];
        a.cv[pos[k]] = vt[k];
        a.cw[pos[k]] = wt[k];
        a.ceid[pos[k]] = a.base + i;
        a.sel[pos[k]] = i;
    }
    if (t == 0) a.cnt[0] = total;
}

template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_ins_post(SmallInsArgs a) {
    take_desc(a);
    constexpr int kIPT = IPT;
    __shared__ InsShared<IPT> sh;
    const uint32_t t = threadIdx.x;
    const uint32_t m = a.cnt[0];
    bool live[kIPT], add[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t j = kIPT * t + k;
        live[k] = j < m;
        add[k] = live[k] && a.found[j] == -1;
    }
    if (a.ord.key) {
        // presorted entries are by batch position: the kept edges go back
        // to their batch positions, flagged when they found no path
        const uint32_t n = a.n;
#pragma unroll
        for (int k = 0; k < kIPT; ++k) sh.keep[kIPT * t + k] = 0;
#pragma unroll
        for (int q = 0; q < 2 * kIPT; ++q) {
            const uint32_t p = 2 * kIPT * t + q;
            if (p < 2 * n) {
                sh.srow[p] = a.ord.row[p];
                sh.sent[p] = a.ord.ent[p];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            const uint32_t j = kIPT * t + k;
            if (!live[k]) continue;
            const uint32_t e = a.sel[j];
            sh.src[e] = a.cu[j];
            sh.dst[e] = a.cv[j];
            sh.w[e] = a.cw[j];
            sh.keep[e] = add[k] ? 1 : 0;
        }
        __syncthreads();
        block_append_sorted<IPT>(a.H, sh.tmp.scan, sh.srow, sh.sent, sh.src, sh.dst, sh.w,
                                 sh.keep, t, n, a.touched, a.cnt + 4);
    } else {
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            const uint32_t j = kIPT * t + k;
            sh.src[j] = live[k] ? a.cu[j] : 0u;
            sh.dst[j] = live[k] ? a.cv[j] : 0u;
            sh.w[j] = live[k] ? a.cw[j] : 0.0f;
        }
        __syncthreads();
        block_append<IPT>(a.H, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, sh.w, add, t, m,
                          a.row_bits, a.touched, a.cnt + 4);
    }
#pragma unroll
    for (int k = 0; k < kIPT; ++k)
        if (live[k]) a.cls[a.sel[kIPT * t + k]] = add[k] ? kOutAdded : kOutFound;
    __syncthreads();
    report_out(a.rep, a.cls, a.n, a.stats, t, kThreads);
}

// ----------------------------------------------------------------- deletions
template <int IPT>
struct DelPreShared {
    static constexpr uint32_t kItems = Shape<IPT>::kItems;
    typename Shape<IPT>::TempU tmp;
    unsigned long long key[kItems];
    uint32_t idx[kItems];
    uint32_t srow[2 * kItems];
    uint32_t sent[2 * kItems];
    uint32_t src[kItems], dst[kItems];
    uint8_t fg[kItems], fh[kItems];  // by batch position: delete from G / from H
};

template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_del_pre(SmallDelArgs a) {
    take_desc(a);
    constexpr int kIPT = IPT;
    __shared__ DelPreShared<IPT> sh;
    const uint32_t t = threadIdx.x;
    const uint32_t n = a.n;
    uint32_t ut[kIPT], vt[kIPT];
    unsigned long long kt[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t i = kIPT * t + k;
        ut[k] = i < n ? a.u[i] : 0u;
        vt[k] = i < n ? a.v[i] : 0u;
        kt[k] = i < n ? edge_key(ut[k], vt[k]) : kPad64;
        sh.fg[i] = 0;
        sh.fh[i] = 0;
    }
    if (a.vcheck && reject_batch<IPT>(ut, vt, n, a, t)) return;
    if (a.ord.key)
        load_order<IPT>(a.ord, sh.key, sh.idx, sh.srow, sh.sent, t, n);
    else
        sort_keys64<IPT>(sh.tmp, sh.key, sh.idx, kt, t, n, a.key_bits);
    // ---- classify: the first occurrence of an edge G holds is deleted ----
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        const unsigned long long kp = sh.key[p];
        if (p < n && (p == 0 || sh.key[p - 1] != kp)) {
            uint32_t q = p + 1;
            while (q < n && sh.key[q] == kp) ++q;
            bool in_g = false, in_h = false;
            if (kp != kPad64) {
                const uint32_t lo = (uint32_t)(kp >> 32), hi = (uint32_t)kp;
                in_g = has_edge(a.G.rows, a.G.adj, lo, hi);
                in_h = in_g && has_edge(a.H.rows, a.H.adj, lo, hi);
            }
            for (uint32_t e = p; e < q; ++e) {
                const uint32_t i = sh.idx[e];
                const bool first = e == p && in_g;
                a.cls[i] = kp == kPad64 ? kOutInvalid : !first ? kOutSkip : in_h ? 0 : kOutGOnly;
                sh.fg[i] = first ? 1 : 0;
                sh.fh[i] = first && in_h ? 1 : 0;
            }
        }
        sh.src[kIPT * t + k] = ut[k];
        sh.dst[kIPT * t + k] = vt[k];
    }
    __syncthreads();
    bool dg[kIPT], dh[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t i = kIPT * t + k;
        dg[k] = i < n && sh.fg[i];
        dh[k] = i < n && sh.fh[i];
    }
    if (a.ord.key) {
        block_delete_sorted<IPT>(a.G, sh.tmp.scan, sh.srow, sh.sent, sh.src, sh.dst, sh.fg, t, n,
                                 a.touched_g, a.cnt + 4);
        block_delete_sorted<IPT>(a.H, sh.tmp.scan, sh.srow, sh.sent, sh.src, sh.dst, sh.fh, t, n,
                                 a.touched_h, a.cnt + 5);
    } else {
        block_delete<IPT>(a.G, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, dg, t, n, a.row_bits,
                          a.touched_g, a.cnt + 4);
        block_delete<IPT>(a.H, sh.tmp, sh.srow, sh.sent, sh.src, sh.dst, dh, t, n, a.row_bits,
                          a.touched_h, a.cnt + 5);
    }
    // ---- compact the edges H held (batch order), as (min, max), for the walk ----
    uint32_t in[kIPT], pos[kIPT], total = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) in[k] = dh[k] ? 1u : 0u;
    Scan(sh.tmp.scan).ExclusiveSum(in, pos, total);
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        if (!dh[k]) continue;
        const uint32_t i = kIPT * t + k;
        a.cu[pos[k]] = min(ut[k], vt[k]);
        a.cv[pos[k]] = max(ut[k], vt[k]);
        a.ceid[pos[k]] = a.base + i;
        a.sel[pos[k]] = i;
    }
    if (t == 0) a.cnt[0] = total;
}

template <int IPT>
struct DelPostShared {
    static constexpr uint32_t kItems = Shape<IPT>::kItems;
    typename Shape<IPT>::TempU tmp;
    uint32_t np[kItems];                  // no-path list: kept-edge index
    uint32_t cu[kItems], cv[kItems];      // candidates, in the general list's order
    float cw[kItems];
    uint8_t add[kItems];                  // by candidate: append to H
    unsigned long long key[kItems];       // candidate keys, sorted
    uint32_t idx[kItems];
    uint32_t srow[2 * kItems];
    uint32_t sent[2 * kItems];
};

template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_del_post(SmallDelArgs a) {
    take_desc(a);
    constexpr int kIPT = IPT;
    __shared__ DelPostShared<IPT> sh;
    const uint32_t t = threadIdx.x;
    const uint32_t m = a.cnt[0];
    int32_t f[kIPT];
    uint32_t pc[kIPT], npf[kIPT], poff[kIPT], kr[kIPT], hops = 0, nnp = 0;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t j = kIPT * t + k;
        f[k] = j < m ? a.found[j] : 0;
        pc[k] = f[k] > 0 ? (uint32_t)f[k] : 0u;
        npf[k] = f[k] == -1 ? 1u : 0u;
    }
    Scan(sh.tmp.scan).ExclusiveSum(pc, poff, hops);
    __syncthreads();
    Scan(sh.tmp.scan).ExclusiveSum(npf, kr, nnp);
    const uint32_t hw = a.hw;
    const uint64_t n3 = 2ull * nnp * hw;  // heuristic walks: two endpoints per no-path edge
    const uint64_t ncand = hops + 2 * n3;
    if (ncand > Shape<IPT>::kItems) {
        // the general recovery takes over: its lists, as its kernels leave them
#pragma unroll
        for (int k = 0; k < kIPT; ++k) {
            const uint32_t j = kIPT * t + k;
            if (j < m) {
                a.pc[j] = pc[k];
                a.poff[j] = poff[k];
            }
            if (npf[k]) {
                const uint32_t q = kr[k];
                a.ep[2 * q] = a.cu[j];
                a.ep[2 * q + 1] = a.cv[j];
                a.ekey[2 * q] = a.ceid[j] << 1;
                a.ekey[2 * q + 1] = (a.ceid[j] << 1) | 1ull;
            }
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
    // ---- candidate list: path hops (kept-edge order), then heuristic hops ----
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t j = kIPT * t + k;
        if (npf[k]) sh.np[kr[k]] = j;
        sh.add[j] = 0;
        if (f[k] > 0) {
            const uint32_t* pj = a.path + (uint64_t)j * a.max_steps;
            for (int q = 1; q <= f[k]; ++q) {
                sh.cu[poff[k] + q - 1] = pj[q - 1];
                sh.cv[poff[k] + q - 1] = pj[q];
            }
        }
    }
    __syncthreads();
    for (uint32_t g = t; g < n3; g += kThreads) {
        const uint32_t i = g / hw, jw = g % hw;  // endpoint i = 2*(no-path index) + side
        const uint32_t e = sh.np[i >> 1];
        const uint32_t u = (i & 1u) ? a.cv[e] : a.cu[e];
        const uint64_t key = (a.ceid[e] << 1) | (uint64_t)(i & 1u);
        const uint3 w3 = heuristic_walk(a.G.rows, a.G.adj, u, key, jw, a.seed, a.G.nodesc);
        sh.cu[hops + 2 * g] = w3.x;
        sh.cv[hops + 2 * g] = w3.y;
        sh.cu[hops + 2 * g + 1] = w3.y;
        sh.cv[hops + 2 * g + 1] = w3.z;
    }
    __syncthreads();
    // ---- first occurrence of each candidate that H lacks, with G's weight ----
    const uint32_t nc = (uint32_t)ncand;
    unsigned long long kt[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t c = kIPT * t + k;
        kt[k] = c < nc ? edge_key(sh.cu[c], sh.cv[c]) : kPad64;
    }
    sort_keys64<IPT>(sh.tmp, sh.key, sh.idx, kt, t, nc, a.key_bits);
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        const unsigned long long ks = sh.key[p];
        if (p >= nc || ks == kPad64 || (p > 0 && sh.key[p - 1] == ks)) continue;
        const uint32_t lo = (uint32_t)(ks >> 32), hi = (uint32_t)ks;
        if (has_edge(a.H.rows, a.H.adj, lo, hi)) continue;
        const int s = find_in_row(a.G.rows, a.G.adj, lo, hi);
        if (s < 0) continue;  // unreachable: path and heuristic hops are edges of G
        const uint32_t c = sh.idx[p];
        sh.add[c] = 1;
        sh.cw[c] = __uint_as_float(a.G.adj[a.G.rows[lo].x + s].y);
    }
    __syncthreads();
    bool take[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) take[k] = kIPT * t + k < nc && sh.add[kIPT * t + k];
    block_append<IPT>(a.H, sh.tmp, sh.srow, sh.sent, sh.cu, sh.cv, sh.cw, take, t, nc, a.row_bits,
                 a.touched_x, a.cnt + 6);
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t j = kIPT * t + k;
        if (j < m) a.cls[a.sel[j]] = npf[k] ? kOutHeuristic : kOutRecovered;
    }
    if (t == 0) {
        a.sizes[4] = 0;
        a.stats[3] = ncand;  // reported as the batch's candidate count
    }
    __syncthreads();
    report_out(a.rep, a.cls, a.n, a.stats, t, kThreads);
}

// ------------------------------------------------------------ batch orders
template <int IPT>
struct OrderShared {
    static constexpr uint32_t kItems = Shape<IPT>::kItems;
    typename Shape<IPT>::TempU tmp;
    unsigned long long key[kItems];
    uint32_t idx[kItems];
    uint32_t srow[2 * kItems];
    uint32_t sent[2 * kItems];
};

// A job's orders: the batch's keys and its row entries, sorted by the pre
// kernel's own sorts (so the orders are the ones it would compute).
template <int IPT>
__device__ void small_order(OrderShared<IPT>& sh, const SmallOrderJob& j, int key_bits,
                            int row_bits) {
    constexpr int kIPT = IPT, kEPT = 2 * IPT;
    const uint32_t t = threadIdx.x, n = j.n;
    const uint32_t pad = row_pad(row_bits);
    unsigned long long kt[kIPT];
    uint32_t re[kEPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t i = kIPT * t + k;
        const uint32_t x = i < n ? j.u[i] : 0u, y = i < n ? j.v[i] : 0u;
        kt[k] = i < n ? edge_key(x, y) : kPad64;
        re[2 * k] = i < n ? x : pad;
        re[2 * k + 1] = i < n ? y : pad;
    }
    sort_keys64<IPT>(sh.tmp, sh.key, sh.idx, kt, t, n, key_bits);
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t p = kIPT * t + k;
        if (p < n) {
            j.key[p] = sh.key[p];
            j.idx[p] = sh.idx[p];
        }
    }
    sort_rows<IPT>(sh.tmp, sh.srow, sh.sent, re, t, n, row_bits);
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
        const uint32_t p = kEPT * t + q;
        if (p < 2 * n) {
            j.row[p] = sh.srow[p];
            j.ent[p] = sh.sent[p];
        }
    }
}

// One workgroup per job of a device array (a list's one-workgroup batches).
template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_order(const SmallOrderJob* __restrict__ jobs,
                                                          int key_bits, int row_bits) {
    __shared__ OrderShared<IPT> sh;
    const SmallOrderJob j = jobs[blockIdx.x];
    small_order<IPT>(sh, j, key_bits, row_bits);
}

// One job given by value (a served request).
template <int IPT>
__global__ __launch_bounds__(kThreads) void k_small_order_one(SmallOrderJob j, int key_bits,
                                                              int row_bits) {
    __shared__ OrderShared<IPT> sh;
    small_order<IPT>(sh, j, key_bits, row_bits);
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
    if (a.n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_ins_pre<1>, dim3(1), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_small_ins_pre<2>, dim3(1), dim3(kThreads), 0, s, a);
    if (!a.G.nodesc)
        refresh({a.G.rows, a.G.adj, a.touched, a.cnt + 4, a.G.lpr}, nullptr, 2 * a.n, s);
    return hipGetLastError();
}

hipError_t launch_small_ins_post(const SmallInsArgs& a, hipStream_t s) {
    if (a.n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_ins_post<1>, dim3(1), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_small_ins_post<2>, dim3(1), dim3(kThreads), 0, s, a);
    refresh({a.H.rows, a.H.adj, a.touched, a.cnt + 4, a.H.lpr}, nullptr, 2 * a.n, s);
    return hipGetLastError();
}

hipError_t launch_small_del_pre(const SmallDelArgs& a, hipStream_t s) {
    if (a.n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_del_pre<1>, dim3(1), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_small_del_pre<2>, dim3(1), dim3(kThreads), 0, s, a);
    const RefreshList h{a.H.rows, a.H.adj, a.touched_h, a.cnt + 5, a.H.lpr};
    if (a.G.nodesc)
        refresh(h, nullptr, 2 * a.n, s);
    else
        refresh({a.G.rows, a.G.adj, a.touched_g, a.cnt + 4, a.G.lpr}, &h, 2 * a.n, s);
    return hipGetLastError();
}

uint32_t small_del_candidates(uint32_t n) {
    return n <= (uint32_t)kThreads ? (uint32_t)kThreads : 2u * kThreads;
}

hipError_t launch_small_del_post(const SmallDelArgs& a, hipStream_t s) {
    if (a.n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_del_post<1>, dim3(1), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_small_del_post<2>, dim3(1), dim3(kThreads), 0, s, a);
    refresh({a.H.rows, a.H.adj, a.touched_x, a.cnt + 6, a.H.lpr}, nullptr,
            2 * small_del_candidates(a.n), s);
    return hipGetLastError();
}

hipError_t launch_small_order(const SmallOrderJob* jobs, uint32_t count, uint32_t max_n,
                              int key_bits, int row_bits, hipStream_t s) {
    if (!count) return hipSuccess;
    if (max_n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_order<1>, dim3(count), dim3(kThreads), 0, s, jobs, key_bits,
                           row_bits);
    else
        hipLaunchKernelGGL(k_small_order<2>, dim3(count), dim3(kThreads), 0, s, jobs, key_bits,
                           row_bits);
    return hipGetLastError();
}

hipError_t launch_small_order_one(const SmallOrderJob& j, int key_bits, int row_bits,
                                  hipStream_t s) {
    if (j.n <= (uint32_t)kThreads)
        hipLaunchKernelGGL(k_small_order_one<1>, dim3(1), dim3(kThreads), 0, s, j, key_bits,
                           row_bits);
    else
        hipLaunchKernelGGL(k_small_order_one<2>, dim3(1), dim3(kThreads), 0, s, j, key_bits,
                           row_bits);
    return hipGetLastError();
}

hipError_t launch_report(const Report& r, const int32_t* cls, uint32_t n,
                         unsigned long long* stats, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, (n + 255) / 256);
    hipLaunchKernelGGL(k_report, dim3(blocks), dim3(256), 0, s, r, cls, n, stats);
    return hipGetLastError();
}

}  // namespace dyg
