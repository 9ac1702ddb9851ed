// walk_kernels.hip -- non-backtracking random-walk kernels for CDNA4 (gfx950).
//
// Layout read by every kernel (device graph store):
//   rows[v] = (first slot, degree)                              8 B, source only
//   adj[p]  = (neighbour, weight bits, neighbour's first slot,
//              neighbour's degree)                              16 B per draw
// A walk hop is one dependent 16 B load: the drawn slot already carries the
// row descriptor of the vertex it leads to.  The slot that the
// non-backtracking rule may redirect to (the row's last slot) is loaded in the
// same step as the drawn slot, so a redirect costs no extra latency.
//
// One workgroup per updated edge, one wave64 lane per walker (W/64 waves).
// The winner of an edge is the walker with the largest conductance 1/R; ties
// go to the walker whose index, bit-reversed over log2(W) bits, is smallest.
// That is the closed form of the reference's tree reduction
// (random_walk_incremental/main.cu:179-201).  Here it is one 64-bit max per
// lane: key = (bits(1/R)+1) << 32 | (0xffff - bitrev(j)) << 16 | found_at,
// reduced by wave64 butterflies and then across waves through LDS.
//
// Draws come from Philox4x32-10 keyed by (seed, tag, edge id, walker): each
// 10-round block yields the next four draws of a walker, so results do not
// depend on how walkers are mapped to lanes, workgroups or launches.
#include <hip/hip_ext.h>

#include <atomic>

#include "internal.h"
#include "philox.h"

namespace dyg {
namespace {

struct Draws {
    uint32_t k0, k1, c1, c2, c3, blk;
    uint32_t b0, b1, b2, b3, avail;

    __device__ __forceinline__ Draws(uint64_t seed, uint32_t tag, uint64_t id, uint32_t walker)
        : k0((uint32_t)seed),
          k1((uint32_t)(seed >> 32)),
          c1(walker),
          c2((uint32_t)id),
          c3(stream_c3(tag, id)),
          blk(0),
          b0(0), b1(0), b2(0), b3(0),
          avail(0) {}

    __device__ __forceinline__ void refill() {
        const uint4 x = philox_block(blk, c1, c2, c3, k0, k1);
        b0 = x.x; b1 = x.y; b2 = x.z; b3 = x.w;
        avail = 4;
        ++blk;
    }

    __device__ __forceinline__ uint32_t next() {
        if (avail == 0) refill();
        const uint32_t r = b0;
        b0 = b1; b1 = b2; b2 = b3;
        --avail;
        return r;
    }
};

struct WalkOut {
    int found;
    float R;
    uint32_t hops, redir, leaf;
};

__device__ __forceinline__ uint2 ld(const uint2* __restrict__ p, uint32_t i) { return p[i]; }
__device__ __forceinline__ uint4 ld(const uint4* __restrict__ p, uint32_t i) { return p[i]; }

// One walker.  INSERT: distortion test R*w_e >= thr before the target test
// (random_walk_incremental/main.cu:97-163).  Otherwise the deletion walk
// (random_walk_decremental/main.cu:104-170).  RECORD writes path[0..].
// PRUNE: stop once fl(1/R) is strictly below the best conductance already
// found by a walker of the same edge; R never decreases, so such a walker can
// no longer win and the edge's winner is unchanged.
template <bool INSERT, bool RECORD, bool PRUNE, bool DESC = true>
__device__ __forceinline__ void walk_lane(const uint2* __restrict__ rows,
                                          const uint4* __restrict__ adj, uint32_t s, uint32_t t,
                                          float we, float thr, int max_steps, Draws& rng,
                                          uint32_t* best, uint32_t* __restrict__ path,
                                          WalkOut& o) {
    o.found = -1;
    o.R = 0.0f;
    o.hops = 0;
    o.redir = 0;
    o.leaf = 0;
    if (RECORD) path[0] = s;
    uint2 row = ld(rows, s);
    if (row.y == 0) return;  // isolated source: nothing to find
    uint32_t i = rng.next() % row.y;
    o.hops = 1;
    uint4 sl = ld(adj, row.x + i);
    float R = 1.0f / __uint_as_float(sl.y);
    if (RECORD) path[1] = sl.x;
    if (sl.x == t) {
        o.found = 1;
        o.R = R;
        if (PRUNE) atomicMax(best, __float_as_uint(1.0f / R) + 1u);
        return;
    }
    uint32_t prev = s, cur = sl.x;
    row = DESC ? make_uint2(sl.z, sl.w) : ld(rows, sl.x);
    int step = 2;
    while (step < max_steps) {
        if (row.y <= 1) {  // leaf (degree 0 cannot occur: deletions keep degree >= 1)
            o.leaf = 1;
            break;
        }
        i = rng.next() % (row.y - 1);
        ++o.hops;
        const uint4 a = ld(adj, row.x + i);
        const uint4 b = ld(adj, row.x + row.y - 1);
        const bool back = (a.x == prev);
        o.redir += back ? 1u : 0u;
        sl = back ? b : a;
        R += 1.0f / __uint_as_float(sl.y);
        if (RECORD) path[step] = sl.x;
        if (INSERT && R * we >= thr) break;
        if (sl.x == t) {
            o.found = step;
            if (PRUNE) atomicMax(best, __float_as_uint(1.0f / R) + 1u);
            break;
        }
        if (PRUNE) {
            const uint32_t bb = __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (bb != 0u && __float_as_uint(1.0f / R) + 1u < bb) break;
        }
        prev = cur;
        cur = sl.x;
        row = DESC ? make_uint2(sl.z, sl.w) : ld(rows, sl.x);
        ++step;
    }
    o.R = R;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int LOGW>
__device__ __forceinline__ uint32_t tie_field(uint32_t j) {
    if constexpr (LOGW == 0) {
        return 0xFFFFu;
    } else {
        return 0xFFFFu - (__brev(j) >> (32 - LOGW));
    }
}

template <int LOGW>
__device__ __forceinline__ uint32_t tie_to_walker(uint32_t tie) {
    if constexpr (LOGW == 0) {
        return 0;
    } else {
        return __brev((0xFFFFu - tie) << (32 - LOGW));
    }
}

// Block of max(W,64) threads; threads j >= W hold no walker.
template <int LOGW, bool INSERT, bool PRUNE>
__global__ __launch_bounds__((1 << LOGW) < 64 ? 64 : (1 << LOGW)) void k_walk(WalkLaunch a) {
    constexpr int W = 1 << LOGW;
    constexpr int BS = W < 64 ? 64 : W;
    constexpr int NWAVE = BS / 64;
    __shared__ unsigned long long s_key[NWAVE];
    __shared__ uint32_t s_cnt[NWAVE][3];
    __shared__ uint32_t s_best;

    const uint32_t e = blockIdx.x;
    if (a.n_dev && e >= *a.n_dev) return;  // list shorter than the grid (uniform per block)
    const uint32_t j = threadIdx.x;
    const uint32_t lane = j & 63u, wave = j >> 6;
    if (j == 0) s_best = 0u;
    const uint32_t s = a.src[e], t = a.dst[e];
    const float we = INSERT ? a.w[e] : 0.0f;
    const uint64_t eid = a.eid ? a.eid[e] : a.eid_base + e;
    __syncthreads();

    WalkOut o;
    o.found = -1; o.R = 0.0f; o.hops = 0; o.redir = 0; o.leaf = 0;
    if (j < (uint32_t)W) {
        Draws rng(a.seed, INSERT ? kTagInsert : kTagDelete, eid, j);
        walk_lane<INSERT, false, PRUNE>(a.rows, a.adj, s, t, we, a.thr, (int)a.max_steps, rng,
                                        &s_best, nullptr, o);
    }
    const bool found = (o.found != -1) && (j < (uint32_t)W);
    const uint32_t hi = found ? __float_as_uint(1.0f / o.R) + 1u : 0u;
    const uint32_t tie = (j < (uint32_t)W) ? tie_field<LOGW>(j) : 0u;
    unsigned long long key =
        ((unsigned long long)hi << 32) | ((unsigned long long)tie << 16) |
        (unsigned long long)((uint32_t)o.found & 0xFFFFu);
    key = wave_max_u64(key);
    const uint32_t h = wave_sum_u32(o.hops);
    const uint32_t r = wave_sum_u32(o.redir);
    const uint32_t l = wave_sum_u32(o.leaf);
    if (lane == 0) {
        s_key[wave] = key;
        s_cnt[wave][0] = h;
        s_cnt[wave][1] = r;
        s_cnt[wave][2] = l;
    }
    __syncthreads();
    if (j == 0) {
        unsigned long long best = s_key[0];
        uint32_t th = s_cnt[0][0], tr = s_cnt[0][1], tl = s_cnt[0][2];
#pragma unroll
        for (int k = 1; k < NWAVE; ++k) {
            best = s_key[k] > best ? s_key[k] : best;
            th += s_cnt[k][0];
            tr += s_cnt[k][1];
            tl += s_cnt[k][2];
        }
        const int f = (int)(int16_t)(uint16_t)(best & 0xFFFFu);
        const uint32_t win = tie_to_walker<LOGW>((uint32_t)(best >> 16) & 0xFFFFu);
        a.found[e] = f;
        a.cond[e] = (f != -1) ? __uint_as_float((uint32_t)(best >> 32) - 1u) : -1.0f;
        if (a.winner) a.winner[e] = win;
        if (a.hops) a.hops[e] = th;
        if (a.stats) {
            atomicAdd(&a.stats[0], (unsigned long long)th);
            atomicAdd(&a.stats[1], (unsigned long long)tr);
            atomicAdd(&a.stats[2], (unsigned long long)tl);
        }
        if (!INSERT && f > 0) {
            // The winner's path, replayed: same draws, same graph snapshot,
            // so the same trajectory (the winner is never pruned).
            Draws rng(a.seed, kTagDelete, eid, win);
            WalkOut ro;
            walk_lane<false, true, false>(a.rows, a.adj, s, t, 0.0f, 0.0f, (int)a.max_steps, rng,
                                          nullptr, a.path + (uint64_t)e * a.max_steps, ro);
        }
    }
}

// ---------------------------------------------------------------------------
// Work-queue form.  A workgroup of 256 lanes owns `E` consecutive edges and
// their E*W (edge, walker) tasks, taken from an LDS counter; a lane whose
// walker ends takes the next task at once, so a wave is never held by its
// longest walker.  Each wave keeps a ring of up to 64 claimed tasks with
// their first draw block; a claim fills the ring up, by the whole wave (one
// LDS atomic and one Philox block per lane), so refills cost no divergent
// Philox and the workgroup's pool can be large: many tasks per lane keep the
// tail after the pool drains short.  Per-edge winner keys and hop counts live in LDS (64-bit LDS max);
// pruning reads the edge's key.  Results equal the one-workgroup-per-edge
// kernel's bit for bit: a walker's trajectory depends only on (edge, walker)
// and the graph.
//
// The kernel is bound by instruction issue and by the latency of each hop's
// dependent load, so the loop keeps every lane on one code path and keeps a
// load in flight while it computes:
//  * with neighbour descriptors in the slots, the loop is software-pipelined:
//    a lane's next slot loads are issued as soon as the slot they depend on
//    has arrived (draw, modulo, address), and the hop's arithmetic and tests
//    (1/w, distortion, target, pruning), the task refills and the Philox
//    rounds run while they are in flight; a walk that ends leaves its last
//    loads unused, and a leaf ends the walk on arrival;
//  * first and later hops share the draw, the load phase and the tests;
//  * each edge's source row is staged in LDS (a refill reads no global memory);
//  * draws come from an 8-entry ring of two Philox blocks (a column of LDS
//    in the pipelined loop, registers otherwise).  Block 0 of a task comes
//    with it from the wave's task ring; later blocks are computed every
//    fourth turn by all active lanes at once, into the half of the ring
//    already consumed (a walker draws exactly once per turn, so the half not
//    holding the next draw is always spent then).  The Philox rounds thus
//    run once per four turns for the whole wave, instead of whenever any one
//    lane runs dry.
// ---------------------------------------------------------------------------
constexpr int kPoolThreads = 256;
constexpr int kPoolWaves = kPoolThreads / 64;
constexpr float kPruneSlack = 1.0f + 0x1p-20f;
constexpr int kPoolMaxEdges = 128;  // edges owned by one workgroup, at most
constexpr uint32_t kRing = 64;      // a wave's claimed tasks (a claim fills it up)

struct Lane {
    uint32_t edge;  // local edge index
    uint32_t walker;
    uint32_t cur, prev, t;
    uint2 row;      // descriptor of the source row (the first hop draws from it)
    float R, we;
    int step;
};

// Draw ring in LDS (the pipelined loop): a lane's column of an 8-row array
// holds its draws, draw d at row d & 7, so block b fills rows 4(b & 1) ..
// 4(b & 1) + 3.  A draw is one LDS read (issued ahead of the slot wait),
// instead of a chain of register selects.
struct LdsRing {
    uint32_t* col;  // the lane's column; rows kPoolThreads words apart
    uint32_t used;  // draws taken so far (the next one sits at row used & 7)
    uint32_t blk;   // next block to compute
    uint32_t c2, c3;

    __device__ __forceinline__ void start(uint4 b0, uint64_t id, uint32_t tag) {
        col[0] = b0.x;
        col[kPoolThreads] = b0.y;
        col[2 * kPoolThreads] = b0.z;
        col[3 * kPoolThreads] = b0.w;
        used = 0;
        blk = 1;
        c2 = (uint32_t)id;
        c3 = stream_c3(tag, id);
    }
    __device__ __forceinline__ void append(uint32_t walker, uint32_t k0, uint32_t k1) {
        const uint4 x = philox_block(blk, walker, c2, c3, k0, k1);
        uint32_t* const p = col + (blk & 1u) * (4u * kPoolThreads);
        p[0] = x.x;
        p[kPoolThreads] = x.y;
        p[2 * kPoolThreads] = x.z;
        p[3 * kPoolThreads] = x.w;
        ++blk;
    }
    __device__ __forceinline__ uint32_t next() {
        const uint32_t v = col[(used & 7u) * kPoolThreads];
        ++used;
        return v;
    }
    // the half the next block goes to has been read (at most 4 draws left)
    __device__ __forceinline__ bool room() const { return 4u * blk - used <= 4u; }
};

// Draw ring: r[d & 7] holds draw d of the walker; block b sits in half b & 1.
struct Ring {
    uint32_t r[8];
    uint32_t q;    // index of the next draw (mod 8)
    uint32_t blk;  // next block to compute
    uint32_t c2, c3;

    __device__ __forceinline__ void start(uint4 b0, uint64_t id, uint32_t tag) {
        r[0] = b0.x; r[1] = b0.y; r[2] = b0.z; r[3] = b0.w;
        q = 0;
        blk = 1;
        c2 = (uint32_t)id;
        c3 = stream_c3(tag, id);
    }
    __device__ __forceinline__ void append(uint32_t walker, uint32_t k0, uint32_t k1) {
        const uint4 x = philox_block(blk, walker, c2, c3, k0, k1);
        const bool hi = (blk & 1u) != 0u;
        r[0] = hi ? r[0] : x.x; r[1] = hi ? r[1] : x.y;
        r[2] = hi ? r[2] : x.z; r[3] = hi ? r[3] : x.w;
        r[4] = hi ? x.x : r[4]; r[5] = hi ? x.y : r[5];
        r[6] = hi ? x.z : r[6]; r[7] = hi ? x.w : r[7];
        ++blk;
    }
    __device__ __forceinline__ uint32_t next() {
        const bool b0 = (q & 1u) != 0u, b1 = (q & 2u) != 0u, b2 = (q & 4u) != 0u;
        const uint32_t a0 = b0 ? r[1] : r[0], a1 = b0 ? r[3] : r[2];
        const uint32_t a2 = b0 ? r[5] : r[4], a3 = b0 ? r[7] : r[6];
        const uint32_t l = b1 ? a1 : a0, h = b1 ? a3 : a2;
        q = (q + 1u) & 7u;
        return b2 ? h : l;
    }
};

// DESC false: the graph's slots carry no neighbour descriptors, so a hop
// reads the next row from rows[] (a second dependent load).
template <int LOGW, bool INSERT, bool PRUNE, bool DESC = true>
__global__ __launch_bounds__(kPoolThreads) void k_walk_pool(WalkLaunch a, uint32_t E) {
    constexpr uint32_t W = 1u << LOGW;
    constexpr uint32_t kE = kPoolMaxEdges;
    constexpr uint32_t kTag = INSERT ? kTagInsert : kTagDelete;
    __shared__ unsigned long long s_key[kE];
    __shared__ uint32_t s_hops[kE];
    __shared__ uint32_t s_src[kE], s_dst[kE];
    __shared__ float s_w[kE];
    __shared__ unsigned long long s_eid[kE];
    __shared__ uint2 s_row[kE];  // source rows: every walker of an edge starts there
    __shared__ uint4 s_rblk[kPoolWaves][kRing];     // per wave: claimed tasks' block 0
    __shared__ uint32_t s_rtask[kPoolWaves][kRing];  // and their task numbers
    __shared__ uint32_t s_next;
    __shared__ uint32_t s_cnt[3];
    __shared__ uint32_t s_draw[DESC ? 8 : 1][kPoolThreads];  // LdsRing columns

    const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;  // list length (device or host)
    const bool decide = INSERT && (a.flags & DYG_WALK_DECISION_ONLY) != 0u;
    const uint32_t e0 = blockIdx.x * E;
    if (e0 >= n) return;  // uniform per workgroup
    const uint32_t ne = min(E, n - e0);
    const uint32_t tid = threadIdx.x;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    if (tid < ne) {
        const uint32_t e = e0 + tid;
        s_src[tid] = a.src[e];
        s_row[tid] = ld(a.rows, a.src[e]);
        s_dst[tid] = a.dst[e];
        s_w[tid] = INSERT ? a.w[e] : 0.0f;
        s_eid[tid] = a.eid ? a.eid[e] : a.eid_base + e;
        // walker 0 not found: the key every found walker beats
        s_key[tid] = (0xFFFFull << 16) | 0xFFFFull;
        s_hops[tid] = 0;
    }
    if (tid == 0) {
        s_next = 0;
        s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
    }
    __syncthreads();
    const uint32_t total = ne * W;

    const uint32_t lane = __lane_id();
    uint4* const rblk = s_rblk[tid >> 6];
    uint32_t* const rtask = s_rtask[tid >> 6];
    uint32_t rhead = 0, rtail = 0;  // the wave's task ring (same in every lane)
    bool pool_empty = false;        // the workgroup's counter has run out (same in every lane)
    bool active = false;
    bool drained = false;
    Lane L{};
    Ring rng{};
    uint4 x{}, y{};  // the slots of the lane's next hop: the drawn one and the row's last
    uint32_t hops = 0, redir = 0, leaf = 0;  // this lane's running counts
    uint32_t my_hops = 0;                    // current walker's draws

    if constexpr (DESC) {
        // slots carry the next row: the next hop's loads go out as soon as
        // a slot arrives (software-pipelined loop)
        LdsRing drw{};
        drw.col = &s_draw[0][tid];
        // Deletion walks (in the dense graph, too large to stay in cache)
        // load the row's last slot only when the drawn slot turns out to be
        // the previous vertex: that hop then completes one turn later, with
        // the draw kept, while the wave's other lanes go on.  Insertion walks
        // (in the sparsifier, cache-resident) load both slots at once.
        constexpr bool kLazy = !INSERT;
        bool pend = false;   // the slot in flight is the redirect of the current hop
        uint32_t r_keep = 0; // the next hop's draw, kept while the redirect is in flight
        uint2 crow{};        // the row the current hop drew from
        for (uint32_t turn = 0;; ++turn) {
            // ---- refill idle lanes from the wave's task ring ----
            const unsigned long long need = __ballot(!active && !drained);
            if (need) {
                const uint32_t cnt = __popcll(need);
                if (rtail - rhead < cnt && !pool_empty) {
                    // fill the ring (one LDS atomic), one block 0 per lane; it
                    // holds fewer than cnt <= 64 tasks, so there is room for one
                    const uint32_t room = kRing - (rtail - rhead);
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&s_next, room);
                    base = __shfl(base, 0, 64);
                    const uint32_t task = base + lane;
                    if (lane < room && task < total) {
                        const uint64_t id = s_eid[task >> LOGW];
                        const uint32_t q = (rtail + lane) & (kRing - 1u);
                        rblk[q] = philox_block(0u, task & (W - 1u), (uint32_t)id, stream_c3(kTag, id),
                                               k0, k1);
                        rtask[q] = task;
                    }
                    const uint32_t got = base >= total ? 0u : min(room, total - base);
                    rtail += got;
                    pool_empty = got < room;
                    // the ring entries this wave just wrote, before its lanes read them
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                const uint32_t avail = rtail - rhead;
                if (!active && !drained) {
                    const uint32_t rank = __popcll(need & ((1ull << lane) - 1ull));
                    if (rank < avail) {
                        const uint32_t q = (rhead + rank) & (kRing - 1u);
                        const uint32_t task = rtask[q];
                        L.edge = task >> LOGW;
                        L.walker = task & (W - 1u);
                        L.cur = s_src[L.edge];
                        L.row = s_row[L.edge];
                        L.t = s_dst[L.edge];
                        L.we = s_w[L.edge];
                        L.prev = 0;
                        L.R = 0.0f;
                        L.step = 0;  // fresh: its first slot load is issued this turn
                        drw.start(rblk[q], s_eid[L.edge], kTag);
                        pend = false;
                        my_hops = 0;
                        active = true;
                    } else {
                        drained = true;  // the counter is out: nothing left to claim
                    }
                }
                rhead += min(cnt, avail);
            }
            if (!__any(active)) break;
            // ---- every fourth turn: the next draw block of every active lane
            //      (while the lanes' slot loads are in flight) ----
            // (a lane that kept draws back for redirects takes a block only
            // once it has read the half the block goes to)
            if ((turn & 3u) == 0u && active && (!kLazy || drw.room()))
                drw.append(L.walker, k0, k1);
            if (!active) continue;

            // ---- one hop: take the slot loaded for it in the last turn, issue the
            //      next hop's loads at once, then decide whether the walk goes on
            //      (a walk that ends here leaves those loads unused) ----
            bool done = false;
            int found = -1;
            const bool fresh = (L.step == 0);  // task started this turn: nothing loaded yet
            const bool first = (L.step == 1);  // the loaded slot is the first hop's
            bool back = false, defer = false;
            uint4 sl = make_uint4(0u, 0u, 0u, 0u);
            uint2 nrow = L.row;  // the row the next hop draws from
            // read before the wait for the slots
            const uint32_t r = (kLazy && pend) ? r_keep : drw.next();
            if (!fresh) {
                // non-backtracking: a draw of the previous vertex is redirected to
                // the row's last slot (the first hop draws over the whole row)
                if (kLazy && pend) {
                    sl = x;  // the last slot, loaded in the last turn
                    back = true;
                } else {
                    back = !first && (x.x == L.prev);
                    sl = (back && !kLazy) ? y : x;
                    defer = kLazy && back;
                }
                nrow = make_uint2(sl.z, sl.w);
            }
            {
                // one load site for every active lane, with no branch: a lane
                // whose walk cannot go on (isolated source, leaf) reads the
                // arena's slot 0 instead (every arena holds at least 4096
                // slots) and never uses it
                // the first hop draws over the whole row, later ones over all
                // but the last slot: d = 0 for a fresh task, else 1
                const uint32_t d = (uint32_t)min(L.step, 1);
                const bool go = nrow.y > d;
                const uint32_t span = go ? nrow.y - d : 1u;
                const uint32_t i = r % span;
                if constexpr (kLazy) {
                    // the next hop's drawn slot, or the redirect of this hop
                    x = a.adj[defer ? crow.x + crow.y - 1u : (go ? nrow.x + i : 0u)];
                    if (!defer) crow = nrow;
                } else {
                    x = a.adj[go ? nrow.x + i : 0u];
                    y = a.adj[go ? nrow.x + nrow.y - 1u : 0u];
                }
            }
            if constexpr (kLazy) {
                pend = defer;
                if (defer) {  // this hop completes next turn, with this turn's draw
                    r_keep = r;
                    continue;
                }
            }
            if (fresh) {
                if (nrow.y == 0u) done = true;  // isolated source: nothing to find
                else L.step = 1;
            } else {
                ++my_hops;
                redir += back ? 1u : 0u;
                L.R += 1.0f / __uint_as_float(sl.y);  // L.R starts at 0: first hop gives 1/w
                if (INSERT && !first && L.R * L.we >= a.thr) {
                    done = true;
                } else if (sl.x == L.t) {
                    found = L.step;
                    done = true;
                } else {
                    if (PRUNE && !first) {
                        // conservative: prune only when fl(1/R) < c is certain;
                        // rcp is within 1 ulp of 1/R, the margin covers 2^-20.
                        // Decision only: any walker having found the target ends
                        // the edge's walk.
                        const uint32_t bb = (uint32_t)(s_key[L.edge] >> 32);
                        if (bb != 0u &&
                            (decide ||
                             __builtin_amdgcn_rcpf(L.R) * kPruneSlack < __uint_as_float(bb - 1u)))
                            done = true;
                    }
                    L.prev = L.cur;
                    L.cur = sl.x;
                    ++L.step;
                    if (!done) {
                        if (L.step >= (int)a.max_steps) {
                            done = true;
                        } else if (nrow.y <= 1u) {
                            // a leaf: the walk would stop at its next hop without a draw
                            leaf += 1u;
                            done = true;
                        }
                    }
                }
            }
            if (done) {
                if (found != -1) {
                    const unsigned long long key =
                        ((unsigned long long)(__float_as_uint(1.0f / L.R) + 1u) << 32) |
                        ((unsigned long long)tie_field<LOGW>(L.walker) << 16) |
                        (unsigned long long)((uint32_t)found & 0xFFFFu);
                    atomicMax(&s_key[L.edge], key);
                }
                atomicAdd(&s_hops[L.edge], my_hops);
                hops += my_hops;
                active = false;
            }
        }
    } else {
        // the next row comes from rows[] (a second dependent load): a walk
        // that ends would wait for it too, so each hop completes in its turn
        for (uint32_t turn = 0;; ++turn) {
            // ---- refill idle lanes from the wave's task ring ----
            const unsigned long long need = __ballot(!active && !drained);
            if (need) {
                const uint32_t cnt = __popcll(need);
                if (rtail - rhead < cnt && !pool_empty) {
                    // fill the ring (one LDS atomic), one block 0 per lane; it
                    // holds fewer than cnt <= 64 tasks, so there is room for one
                    const uint32_t room = kRing - (rtail - rhead);
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&s_next, room);
                    base = __shfl(base, 0, 64);
                    const uint32_t task = base + lane;
                    if (lane < room && task < total) {
                        const uint64_t id = s_eid[task >> LOGW];
                        const uint32_t q = (rtail + lane) & (kRing - 1u);
                        rblk[q] = philox_block(0u, task & (W - 1u), (uint32_t)id, stream_c3(kTag, id),
                                               k0, k1);
                        rtask[q] = task;
                    }
                    const uint32_t got = base >= total ? 0u : min(room, total - base);
                    rtail += got;
                    pool_empty = got < room;
                    // the ring entries this wave just wrote, before its lanes read them
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
                const uint32_t avail = rtail - rhead;
                if (!active && !drained) {
                    const uint32_t rank = __popcll(need & ((1ull << lane) - 1ull));
                    if (rank < avail) {
                        const uint32_t q = (rhead + rank) & (kRing - 1u);
                        const uint32_t task = rtask[q];
                        L.edge = task >> LOGW;
                        L.walker = task & (W - 1u);
                        L.cur = s_src[L.edge];
                        L.row = s_row[L.edge];
                        L.t = s_dst[L.edge];
                        L.we = s_w[L.edge];
                        L.prev = 0;
                        L.R = 0.0f;
                        L.step = 1;
                        rng.start(rblk[q], s_eid[L.edge], kTag);
                        my_hops = 0;
                        active = true;
                    } else {
                        drained = true;  // the counter is out: nothing left to claim
                    }
                }
                rhead += min(cnt, avail);
            }
            if (!__any(active)) break;
            // ---- every fourth turn: the next draw block of every active lane ----
            if ((turn & 3u) == 0u && active) rng.append(L.walker, k0, k1);
            if (!active) continue;

            // ---- one hop: first and later hops share one path ----
            bool done = false;
            int found = -1;
            const uint2 row = L.row;
            const bool first = (L.step == 1);
            // first hop: any slot of the source row; later hops: the first deg-1
            // slots, with the previous vertex redirected to the last slot
            const uint32_t span = first ? row.y : row.y - 1u;
            if (row.y == 0u || span == 0u) {
                leaf += first ? 0u : 1u;  // isolated source, or a leaf
                done = true;
            } else {
                const uint32_t i = rng.next() % span;
                ++my_hops;
                const uint4 x = ld(a.adj, row.x + i);
                const uint4 y = ld(a.adj, row.x + row.y - 1u);
                const bool back = !first && (x.x == L.prev);
                redir += back ? 1u : 0u;
                const uint4 sl = back ? y : x;
                L.R += 1.0f / __uint_as_float(sl.y);  // L.R starts at 0: first hop gives 1/w
                if (INSERT && !first && L.R * L.we >= a.thr) {
                    done = true;
                } else if (sl.x == L.t) {
                    found = L.step;
                    done = true;
                } else {
                    if (PRUNE && !first) {
                        // conservative: prune only when fl(1/R) < c is certain;
                        // rcp is within 1 ulp of 1/R, the margin covers 2^-20.
                        // Decision only: any walker having found the target ends
                        // the edge's walk.
                        const uint32_t bb = (uint32_t)(s_key[L.edge] >> 32);
                        if (bb != 0u &&
                            (decide ||
                             __builtin_amdgcn_rcpf(L.R) * kPruneSlack < __uint_as_float(bb - 1u)))
                            done = true;
                    }
                    L.prev = L.cur;
                    L.cur = sl.x;
                    L.row = DESC ? make_uint2(sl.z, sl.w) : ld(a.rows, sl.x);
                    ++L.step;
                    if (L.step >= (int)a.max_steps) done = true;
                }
            }
            if (done) {
                if (found != -1) {
                    const unsigned long long key =
                        ((unsigned long long)(__float_as_uint(1.0f / L.R) + 1u) << 32) |
                        ((unsigned long long)tie_field<LOGW>(L.walker) << 16) |
                        (unsigned long long)((uint32_t)found & 0xFFFFu);
                    atomicMax(&s_key[L.edge], key);
                }
                atomicAdd(&s_hops[L.edge], my_hops);
                hops += my_hops;
                active = false;
            }
        }
    }

    // block totals
    const uint32_t h = wave_sum_u32(hops), r = wave_sum_u32(redir), l = wave_sum_u32(leaf);
    if (lane == 0) {
        atomicAdd(&s_cnt[0], h);
        atomicAdd(&s_cnt[1], r);
        atomicAdd(&s_cnt[2], l);
    }
    __syncthreads();
    if (tid < ne) {
        const uint32_t e = e0 + tid;
        const unsigned long long best = s_key[tid];
        const int f = (int)(int16_t)(uint16_t)(best & 0xFFFFu);
        const uint32_t win = tie_to_walker<LOGW>((uint32_t)(best >> 16) & 0xFFFFu);
        a.found[e] = f;
        a.cond[e] = (f != -1) ? __uint_as_float((uint32_t)(best >> 32) - 1u) : -1.0f;
        if (a.winner) a.winner[e] = win;
        if (a.hops) a.hops[e] = s_hops[tid];
        if (!INSERT && f > 0) {
            Draws rr(a.seed, kTagDelete, s_eid[tid], win);
            WalkOut ro;
            walk_lane<false, true, false, DESC>(a.rows, a.adj, s_src[tid], s_dst[tid], 0.0f, 0.0f,
                                          (int)a.max_steps, rr, nullptr,
                                          a.path + (uint64_t)e * a.max_steps, ro);
        }
    }
    if (tid == 0 && a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)s_cnt[0]);
        atomicAdd(&a.stats[1], (unsigned long long)s_cnt[1]);
        atomicAdd(&a.stats[2], (unsigned long long)s_cnt[2]);
    }
}

__global__ __launch_bounds__(256) void k_heuristic(const uint2* __restrict__ rows,
                                                   const uint4* __restrict__ adj,
                                                   const uint32_t* __restrict__ endpoint,
                                                   const uint64_t* __restrict__ key, uint32_t n,
                                                   uint32_t walkers, uint64_t seed,
                                                   uint32_t* __restrict__ path3,
                                                   const uint32_t* __restrict__ n_dev,
                                                   bool nodesc) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev && *n_dev < n) n = *n_dev;
    if (gid >= (uint64_t)n * walkers) return;
    const uint32_t i = (uint32_t)(gid / walkers), j = (uint32_t)(gid % walkers);
    const uint3 t = heuristic_walk(rows, adj, endpoint[i], key[i], j, seed, nodesc);
    path3[gid * 3 + 0] = t.x;
    path3[gid * 3 + 1] = t.y;
    path3[gid * 3 + 2] = t.z;
}

// Graph edits: final slot values and row descriptors planned on the host.
// Every written slot and row appears once, so one launch applies them all.
__global__ __launch_bounds__(256) void k_apply_delta(uint2* __restrict__ rows,
                                                     uint4* __restrict__ adj,
                                                     const uint32_t* __restrict__ d,
                                                     uint32_t n_write, uint32_t n_row) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t* rw = d + (uint64_t)n_write * 5;
    if (g < n_write) {
        const uint32_t* q = d + (uint64_t)g * 5;
        adj[q[0]] = make_uint4(q[1], q[2], q[3], q[4]);
    } else if (g - n_write < n_row) {
        const uint32_t k = g - n_write;
        rows[rw[k * 3 + 0]] = make_uint2(rw[k * 3 + 1], rw[k * 3 + 2]);
    }
}

template <bool INSERT, bool PRUNE>
hipError_t launch_walk_t(const WalkLaunch& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const dim3 grid(a.n);
    switch (a.walkers) {
#define DYG_CASE(LW)                                                                       \
    case (1u << LW):                                                                       \
        hipExtLaunchKernelGGL((k_walk<LW, INSERT, PRUNE>), grid,                           \
                              dim3((1 << LW) < 64 ? 64 : (1 << LW)), 0, s, e0, e1, 0, a);  \
        break;
        DYG_CASE(0) DYG_CASE(1) DYG_CASE(2) DYG_CASE(3) DYG_CASE(4) DYG_CASE(5)
        DYG_CASE(6) DYG_CASE(7) DYG_CASE(8) DYG_CASE(9) DYG_CASE(10)
#undef DYG_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Workgroups of the work-queue kernel that the device holds at once (CUs x
// resident workgroups per CU), computed once per device.
template <int LOGW, bool INSERT, bool PRUNE, bool DESC>
uint32_t pool_slots() {
    static std::atomic<uint32_t> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2048;
    uint32_t v = cached[dev].load(std::memory_order_relaxed);
    if (v) return v;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(&k_walk_pool<LOGW, INSERT, PRUNE, DESC>),
            kPoolThreads, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu <= 0 || cus <= 0)
        return 2048;
    v = (uint32_t)(per_cu * cus);
    cached[dev].store(v, std::memory_order_relaxed);
    return v;
}

// Edges per workgroup: by default the fewest that keep the launch within one
// round of resident workgroups (so no workgroup starts late and adds its own
// tail), and at least one task per lane.
inline uint32_t pool_edges(uint32_t W, uint32_t requested, uint32_t n, uint32_t slots) {
    const uint32_t cap = kPoolMaxEdges;
    uint32_t E = requested;
    if (!E) {
        E = (uint32_t)(((uint64_t)n + slots - 1) / slots);
        const uint32_t min_e = (kPoolThreads + W - 1) / W;
        if (E < min_e) E = min_e;
    }
    if (E < 1) E = 1;
    if (E > cap) E = cap;
    return E;
}

template <bool INSERT, bool PRUNE, bool DESC>
hipError_t launch_pool_t(const WalkLaunch& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    uint32_t slots = 2048;
    switch (a.walkers) {
#define DYG_CASE(LW) \
    case (1u << LW): slots = pool_slots<LW, INSERT, PRUNE, DESC>(); break;
        DYG_CASE(0) DYG_CASE(1) DYG_CASE(2) DYG_CASE(3) DYG_CASE(4) DYG_CASE(5)
        DYG_CASE(6) DYG_CASE(7) DYG_CASE(8) DYG_CASE(9) DYG_CASE(10)
#undef DYG_CASE
        default:
            return hipErrorInvalidValue;
    }
    const uint32_t E = pool_edges(a.walkers, a.edges_per_block, a.n, slots);
    const dim3 grid((a.n + E - 1) / E);
    switch (a.walkers) {
#define DYG_CASE(LW)                                                                        \
    case (1u << LW):                                                                        \
        hipExtLaunchKernelGGL((k_walk_pool<LW, INSERT, PRUNE, DESC>), grid, dim3(kPoolThreads), \
                              0, s, e0, e1, 0, a, E);                                       \
        break;
        DYG_CASE(0) DYG_CASE(1) DYG_CASE(2) DYG_CASE(3) DYG_CASE(4) DYG_CASE(5)
        DYG_CASE(6) DYG_CASE(7) DYG_CASE(8) DYG_CASE(9) DYG_CASE(10)
#undef DYG_CASE
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool INSERT>
hipError_t launch_any(const WalkLaunch& a, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    const bool prune = !(a.flags & DYG_WALK_NO_PRUNE);
    if constexpr (INSERT) {
        if (a.nodesc) return hipErrorNotSupported;  // insertion walks run on sparsifiers
    }
    if (!valid_walkers(a.walkers)) return hipErrorInvalidValue;
    if (a.n == 0) {  // nothing to launch: the events still mark the place
        hipError_t e = e0 ? hipEventRecord(e0, s) : hipSuccess;
        if (e == hipSuccess && e1) e = hipEventRecord(e1, s);
        return e;
    }
    if constexpr (!INSERT) {
        // a graph without descriptors (the stream's dense graph) is walked by
        // the work-queue kernel reading rows[] per hop; same results
        if (a.nodesc)
            return prune ? launch_pool_t<false, true, false>(a, s, e0, e1)
                         : launch_pool_t<false, false, false>(a, s, e0, e1);
    }
    if (a.flags & DYG_WALK_PER_EDGE)
        return prune ? launch_walk_t<INSERT, true>(a, s, e0, e1)
                     : launch_walk_t<INSERT, false>(a, s, e0, e1);
    return prune ? launch_pool_t<INSERT, true, true>(a, s, e0, e1)
                 : launch_pool_t<INSERT, false, true>(a, s, e0, e1);
}

}  // namespace

bool valid_walkers(uint32_t w) { return w >= 1 && w <= 1024 && (w & (w - 1)) == 0; }

hipError_t launch_walk_incremental(const WalkLaunch& a, hipStream_t s, hipEvent_t ev_begin,
                                   hipEvent_t ev_end) {
    return launch_any<true>(a, s, ev_begin, ev_end);
}

hipError_t launch_walk_decremental(const WalkLaunch& a, hipStream_t s, hipEvent_t ev_begin,
                                   hipEvent_t ev_end) {
    return launch_any<false>(a, s, ev_begin, ev_end);
}

hipError_t launch_heuristic(const uint2* rows, const uint4* adj, const uint32_t* endpoint,
                            const uint64_t* key, uint32_t n, uint32_t walkers, uint64_t seed,
                            uint32_t* path3, hipStream_t s, const uint32_t* n_dev,
                            bool nodesc) {
    const uint64_t total = (uint64_t)n * walkers;
    if (total == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)((total + 255) / 256);
    hipLaunchKernelGGL(k_heuristic, dim3(blocks), dim3(256), 0, s, rows, adj, endpoint, key, n,
                       walkers, seed, path3, n_dev, nodesc);
    return hipGetLastError();
}

hipError_t launch_apply_delta(uint2* rows, uint4* adj, const uint32_t* delta, uint32_t n_write,
                              uint32_t n_row, hipStream_t s) {
    const uint32_t blocks = (uint32_t)(((uint64_t)n_write + n_row + 255) / 256);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_apply_delta, dim3(blocks), dim3(256), 0, s, rows, adj, delta, n_write,
                       n_row);
    return hipGetLastError();
}

}  // namespace dyg
