// stream_small.hip -- insertion batches of up to kSmallBatch edges in two
// one-workgroup kernels around the walk.
//
// The general pipeline of an insertion batch (stream.hip) is about forty
// launches; for a batch of a few hundred edges every kernel runs for a few
// microseconds and the launches themselves set the pace.  Here one
// workgroup of 1024 lanes does, with the batch in LDS:
//   pre:  sort the (min,max) keys (stable rank by counting: a batch of n
//         keys costs n LDS reads per lane), classify each distinct edge in
//         batch order (weight update / new), append the new edges to G (each
//         row's appends in batch order, counted the same way; relocation
//         through the same device bump allocator as the general pipeline),
//         apply the weight updates, and compact the kept edges for the walk;
//   post: append the edges that found no path to H the same way, write the
//         outcome codes and report the batch into pinned host memory.
// After each, one multi-workgroup launch refreshes the descriptors copied
// into the slots pointing at the touched rows.
// Same results as the general pipeline (tested against the oracle).
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace dyg {
namespace {

constexpr int kThreads = 1024;
constexpr uint32_t kPadRow = 0xFFFFFFFFu;
constexpr unsigned long long kPad64 = ~0ull;
constexpr int32_t kOutWeight = 1, kOutFound = 2, kOutAdded = 3, kOutInvalid = 8;

using Scan = hipcub::BlockScan<uint32_t, kThreads>;

struct Shared {
    typename Scan::TempStorage scan;
    unsigned long long kin[kThreads];  // batch keys, batch order
    unsigned long long key[kThreads];  // sorted batch keys
    uint32_t idx[kThreads];            // their batch positions
    uint32_t rk[2 * kThreads];         // row of entry 2t+j (edge t, side j), or kPadRow
    uint32_t rb[2 * kThreads];         // head entries: first slot of the row's appends
    uint32_t src[kThreads], dst[kThreads];
    float w[kThreads];
    uint8_t keep[kThreads];            // by batch position
    uint8_t mode[kThreads];            // by sorted position (heads)
};

// Slot of b in row a, or -1; four independent loads in flight per round.
__device__ __forceinline__ int find_in_row(const uint2* rows, const uint4* adj, uint32_t a,
                                           uint32_t b) {
    const uint2 r = rows[a];
    const uint4* p = adj + r.x;
    uint32_t q = 0;
    for (; q + 4 <= r.y; q += 4) {
        const uint32_t c0 = p[q].x, c1 = p[q + 1].x, c2 = p[q + 2].x, c3 = p[q + 3].x;
        if (c0 == b) return (int)q;
        if (c1 == b) return (int)q + 1;
        if (c2 == b) return (int)q + 2;
        if (c3 == b) return (int)q + 3;
    }
    for (; q < r.y; ++q)
        if (p[q].x == b) return (int)q;
    return -1;
}

__device__ __forceinline__ bool holds(const uint2* rows, const uint4* adj, uint32_t a, uint32_t b) {
    return rows[a].y <= rows[b].y ? find_in_row(rows, adj, a, b) >= 0
                                  : find_in_row(rows, adj, b, a) >= 0;
}

// Appends edge t (lane t < span: src[t]-dst[t], weight w[t]) to both rows
// when `take`; each row's appends keep batch order.  Returns after the rows
// are final.  The touched rows are listed in `touched` (count in *ntouched)
// for k_refresh_rows, which rewrites the neighbours' copies of their
// descriptors across the whole GPU (in one workgroup that refresh would be
// deg^2 dependent loads per lane).
__device__ void block_append(const BlockGraph& g, Shared& sh, bool take, uint32_t t,
                             uint32_t span, uint32_t* __restrict__ touched,
                             uint32_t* __restrict__ ntouched) {
    const uint32_t r0 = take ? sh.src[t] : kPadRow, r1 = take ? sh.dst[t] : kPadRow;
    sh.rk[2 * t] = r0;
    sh.rk[2 * t + 1] = r1;
    const int appended = __syncthreads_count(take ? 1 : 0);
    // per entry: earlier entries of its row, all entries of its row, first one
    uint32_t before0 = 0, total0 = 0, head0 = 2 * t;
    uint32_t before1 = 0, total1 = 0, head1 = 2 * t + 1;
    if (take) {
        for (uint32_t y = 0; y < 2 * span; ++y) {
            const uint32_t ry = sh.rk[y];
            if (ry == r0) {
                if (total0 == 0) head0 = y;
                before0 += y < 2 * t ? 1u : 0u;
                ++total0;
            }
            if (ry == r1) {
                if (total1 == 0) head1 = y;
                before1 += y < 2 * t + 1 ? 1u : 0u;
                ++total1;
            }
        }
    }
    // the first entry of each row reserves the row's room
    for (int j = 0; j < 2; ++j) {
        const uint32_t before = j ? before1 : before0;
        if (!take || before != 0) continue;
        const uint32_t r = j ? r1 : r0;
        const uint32_t cnt = j ? total1 : total0;
        uint2 row = g.rows[r];
        const uint32_t d0 = row.y;
        const uint32_t c = g.cap[r];
        uint32_t base = row.x + d0;
        if ((unsigned long long)d0 + cnt > c) {  // relocate to the arena tail, order kept
            unsigned long long ncap = 2ull * c;
            ncap = max(ncap, (unsigned long long)d0 + cnt + g.slack);
            ncap = max(ncap, 4ull);
            const unsigned long long ns = atomicAdd(&g.meta[0], ncap);
            if (ns + ncap > g.cap_slots) {
                atomicOr((uint32_t*)(g.meta + 2), 1u);
                sh.rb[2 * t + j] = kPadRow;
                continue;
            }
            for (uint32_t x = 0; x < d0; ++x) g.adj[ns + x] = g.adj[row.x + x];
            row.x = (uint32_t)ns;
            g.cap[r] = (uint32_t)ncap;
            base = row.x + d0;
        }
        g.rows[r] = make_uint2(row.x, d0 + cnt);
        sh.rb[2 * t + j] = base;
    }
    __syncthreads();
    if (take) {
        const uint32_t b0 = sh.rb[head0], b1 = sh.rb[head1];
        const uint32_t wb = __float_as_uint(sh.w[t]);
        if (b0 != kPadRow) g.adj[b0 + before0] = make_uint4(r1, wb, 0u, 0u);
        if (b1 != kPadRow) g.adj[b1 + before1] = make_uint4(r0, wb, 0u, 0u);
    }
    if (t == 0 && appended) atomicAdd(&g.meta[1], (unsigned long long)appended);
    // list the touched rows (one per row: its first entry), in any order
    const bool h0 = take && before0 == 0, h1 = take && before1 == 0;
    uint32_t pos = 0, total = 0;
    Scan(sh.scan).ExclusiveSum((h0 ? 1u : 0u) + (h1 ? 1u : 0u), pos, total);
    if (h0) touched[pos++] = r0;
    if (h1) touched[pos] = r1;
    if (t == 0) *ntouched = total;
    __syncthreads();
}

__device__ void add_weights_seq(const BlockGraph& g, uint32_t a, uint32_t b, const Shared& sh,
                                uint32_t p0, uint32_t p1, const float* __restrict__ w) {
    const int pa = find_in_row(g.rows, g.adj, a, b), pb = find_in_row(g.rows, g.adj, b, a);
    if (pa < 0 || pb < 0) return;
    uint4* sa = &g.adj[g.rows[a].x + pa];
    uint4* sb = &g.adj[g.rows[b].x + pb];
    float x = __uint_as_float(sa->y);
    for (uint32_t e = p0; e < p1; ++e) x = x + w[sh.idx[e]];
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

__global__ __launch_bounds__(kThreads) void k_small_ins_pre(SmallInsArgs a) {
    __shared__ Shared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t n = a.n;
    // ---- sort the batch's (min,max) keys, stable: rank by counting ----
    uint32_t ut = 0, vt = 0;
    float wt = 0.0f;
    unsigned long long kt = kPad64;
    if (t < n) {
        ut = a.u[t];
        vt = a.v[t];
        wt = a.w[t];
        const uint32_t lo = min(ut, vt), hi = max(ut, vt);
        kt = lo == hi ? kPad64 : ((unsigned long long)lo << 32) | hi;
    }
    sh.kin[t] = kt;
    sh.keep[t] = 0;
    sh.mode[t] = 0;
    __syncthreads();
    uint32_t rank = t;  // lanes past the batch keep their place
    if (t < n) {
        rank = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const unsigned long long kj = sh.kin[j];
            rank += (kj < kt || (kj == kt && j < t)) ? 1u : 0u;
        }
    }
    sh.key[rank] = kt;
    sh.idx[rank] = t;
    __syncthreads();
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
    block_append(a.G, sh, take, t, n, a.touched, a.cnt + 4);
    // ---- weight updates (after the appends: a repeated new edge adds to it) ----
    const uint8_t md = sh.mode[t];
    if (md) {
        uint32_t q = t + 1;
        while (q < n && sh.key[q] == k) ++q;
        const uint32_t lo = (uint32_t)(k >> 32), hi = (uint32_t)k;
        const uint32_t e0 = (md & 1) ? t : t + 1;
        add_weights_seq(a.G, lo, hi, sh, e0, q, a.w);
        if (md & 4) add_weights_seq(a.H, lo, hi, sh, e0, q, a.w);
    }
    // ---- compact the kept edges (batch order) for the walk ----
    uint32_t pos = 0, total = 0;
    Scan(sh.scan).ExclusiveSum(take ? 1u : 0u, pos, total);
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
    __shared__ Shared sh;
    const uint32_t t = threadIdx.x;
    const uint32_t m = a.cnt[0];
    const bool live = t < m;
    const bool add = live && a.found[t] == -1;
    sh.src[t] = live ? a.cu[t] : 0u;
    sh.dst[t] = live ? a.cv[t] : 0u;
    sh.w[t] = live ? a.cw[t] : 0.0f;
    __syncthreads();
    block_append(a.H, sh, add, t, m, a.touched, a.cnt + 4);
    if (live) a.cls[a.sel[t]] = add ? kOutAdded : kOutFound;
    __syncthreads();
    report_out(a.rep, a.cls, a.n, a.stats, t, kThreads);
}

// One wave per listed row: every slot pointing to the row gets its final
// descriptor (a lane per neighbour).
__global__ __launch_bounds__(256) void k_refresh_rows(const uint2* __restrict__ rows,
                                                      uint4* __restrict__ adj,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ count) {
    const uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (i >= *count) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = list[i];
    const uint2 rr = rows[r];
    for (uint32_t x = lane; x < rr.y; x += 64u) {
        const uint32_t nb = adj[rr.x + x].x;
        const uint2 rn = rows[nb];
        for (uint32_t y = 0; y < rn.y; ++y) {
            uint4* sl = &adj[rn.x + y];
            if (sl->x == r) *reinterpret_cast<uint2*>(&sl->z) = rr;
        }
    }
}

__global__ __launch_bounds__(256) void k_report(Report r, const int32_t* __restrict__ cls,
                                                uint32_t n, unsigned long long* __restrict__ stats) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) r.cls[i] = cls[i];
    if (blockIdx.x == 0) report_out(r, cls, 0, stats, threadIdx.x, 256);
}

}  // namespace

hipError_t launch_small_ins_pre(const SmallInsArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_ins_pre, dim3(1), dim3(kThreads), 0, s, a);
    hipLaunchKernelGGL(k_refresh_rows, dim3(2 * kSmallBatch / 4), dim3(256), 0, s, a.G.rows,
                       a.G.adj, a.touched, a.cnt + 4);
    return hipGetLastError();
}

hipError_t launch_small_ins_post(const SmallInsArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_small_ins_post, dim3(1), dim3(kThreads), 0, s, a);
    hipLaunchKernelGGL(k_refresh_rows, dim3(2 * kSmallBatch / 4), dim3(256), 0, s, a.H.rows,
                       a.H.adj, a.touched, a.cnt + 4);
    return hipGetLastError();
}

hipError_t launch_report(const Report& r, const int32_t* cls, uint32_t n,
                         unsigned long long* stats, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1u, (n + 255) / 256);
    hipLaunchKernelGGL(k_report, dim3(blocks), dim3(256), 0, s, r, cls, n, stats);
    return hipGetLastError();
}

}  // namespace dyg
