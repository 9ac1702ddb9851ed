// internal.h -- shared internals of libdygrass (context, graph store, helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/dygrass.h"

namespace dyg {

// Random stream tags: top byte of Philox counter word 3.
constexpr uint32_t kTagInsert = 1u;
constexpr uint32_t kTagDelete = 2u;
constexpr uint32_t kTagHeuristic = 3u;

void set_error(dyg_ctx ctx, const std::string& msg);

// Pinned host staging buffer that grows on demand.
struct Pinned {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n);
    ~Pinned();
};

// Device buffer that grows on demand (contents not preserved on growth).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t n);
    ~DevBuf();
};

}  // namespace dyg

struct dyg_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double clock_khz = 0.0;  // the device's wall clock rate (wall_clock64 ticks per ms)
    // written by the caller's thread and, while a server is open, by its
    // worker threads: always under err_mu
    std::mutex err_mu;
    std::string last_error;
    dyg::Pinned stage;         // H2D / D2H staging
    dyg::Pinned report;        // per-batch reports of a list of stream batches in flight
    dyg::DevBuf dscratch;      // per-call device staging of inputs/outputs
    dyg::DevBuf dstats;        // device counters (hops, redirects, leaf stops)
    std::vector<hipEvent_t> evpool;  // timing events for multi-launch runs
    void reserve_events(size_t n) {
        while (evpool.size() < n) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) break;
            evpool.push_back(e);
        }
    }
    ~dyg_ctx_s() {
        for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    }
};

namespace dyg {
// Device-side state of the insert pipeline (device_edits.hip).
struct DevEdits {
    bool ready = false;
    bool cap_valid = false;  // `cap` holds the current capacities
    uint32_t max_batch = 0;
    uint32_t *cap = nullptr, *segstart = nullptr, *segcnt = nullptr, *rowbase = nullptr;
    uint32_t seq = 0;  // batches since device_edits_begin (selects the count slot)
    uint64_t base_nslots = 0;  // live slots at device_edits_begin
    uint32_t *keys_in = nullptr, *keys_out = nullptr, *vals_in = nullptr, *vals_out = nullptr;
    unsigned long long* meta = nullptr;  // used slots, added edges, error, counts
    uint8_t* mark = nullptr;   // [2 * max_batch] touched rows of a presorted edit (by position)
    void* temp = nullptr;
    size_t temp_bytes = 0;
    // a whole run's row entries, every batch's sorted ahead of its walks
    // (device_edits_order_run): keys and values in and out, then sort scratch
    DevBuf run;
    uint64_t run_cap = 0;  // updates the buffer holds
    size_t run_temp_bytes = 0;
    void release();
    ~DevEdits() { release(); }
};
}  // namespace dyg

// One row descriptor: first slot and degree, 8 bytes.
// One slot: (neighbour id, weight bits, neighbour's first slot, neighbour's
// degree), 16 bytes read in one load -- a walk hop needs no row lookup.
// Edits keep every slot's copy of its neighbour's row descriptor current.
struct dyg_graph_s {
    dyg_ctx ctx = nullptr;
    uint32_t V = 0;
    uint2* d_rows = nullptr;  // [V] (start, deg)
    uint4* d_adj = nullptr;   // [cap_slots] (col, float bits of w, start(col), deg(col))
    uint64_t cap_slots = 0;   // slots allocated on the device
    uint64_t used_slots = 0;  // bump pointer of the slot arena
    uint32_t slack = 0;       // spare slots given to a relocated row
    // host mirror of the device state (same slot positions)
    std::vector<uint32_t> start, deg, cap;
    std::vector<uint4> adj;
    uint64_t nslots = 0;  // sum of degrees
    // slots do not keep their neighbour's descriptor current (the stream's
    // dense graph: no refresh after its edits; walks on it read rows[] per hop)
    bool nodesc = false;
    // reference edge map (random_walk_decremental/functions.cpp:243-258)
    bool keyed = false;
    long mult = 1;
    std::unordered_map<long, std::pair<uint32_t, uint32_t>> key_map;
    // delta staging
    std::vector<uint32_t> delta;
    dyg::DevBuf d_delta;
    // device-edit mode: rows/cap/slots were changed on the device and the
    // host mirror must be re-read before the host plans an edit
    dyg::DevEdits dev;
    bool mirror_stale = false;
};

namespace dyg {

inline long key_of(long a, long b, long mult) { return a < b ? a * mult + b : b * mult + a; }
long key_multiplier(uint64_t V);

// Graph edit plans (graph_store.hip).  Both build a delta in g->delta and
// apply it on the device; the host mirror is updated in place.
dyg_status graph_apply_inserts(dyg_graph g, const uint32_t* u, const uint32_t* v, const float* w,
                               uint32_t n);
dyg_status graph_apply_deletes(dyg_graph g, const uint32_t* u, const uint32_t* v, uint32_t n,
                               int8_t* hit, bool allow_zero_degree = false);
// weight(u[i],v[i]) += dw[i] in order, both slots (keyed graphs).
dyg_status graph_add_weights(dyg_graph g, const uint32_t* u, const uint32_t* v, const float* dw,
                             uint32_t n);
// Grow the slot arena to at least `need` slots.
dyg_status graph_reserve_slots(dyg_graph g, uint64_t need);
// Re-read rows, capacities and slots from the device after device edits.
dyg_status graph_refresh_mirror(dyg_graph g);

// Device-side insert pipeline (device_edits.hip).  begin() reserves arena
// space for up to `max_inserts` appends in batches of at most `max_batch`
// edges; batch() enqueues one batch's edits (device pointers; found[i] == -1
// marks an edge to add); end() waits and reports the edges added.
dyg_status device_edits_begin(dyg_graph g, uint32_t max_batch, uint64_t max_inserts);
// The allocation part of begin(): arena growth and pipeline buffers (no-op
// when already large enough).  Lets a caller move all allocation out of a
// timed or latency-critical run.
dyg_status device_edits_prepare(dyg_graph g, uint32_t max_batch, uint64_t max_inserts);
dyg_status device_edits_batch(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                              const float* w, const int32_t* found, uint32_t n,
                              const uint32_t* n_dev = nullptr);
// The row entries of a run of m updates in batches of I (batch b = i / I),
// computed in one sort ahead of the run's walks: entry 2i / 2i+1 of update i
// is src[i]'s / dst[i]'s row, valued by its entry index within its batch;
// batch b's 2n entries end up at [2bI, 2bI + 2n) of *row / *ent, by row and
// then by entry, as BatchOrder::row / ent are for one batch.  *ok is false
// (and nothing enqueued) when batch and row bits exceed 32 or 2m entries
// exceed one sort's item count.
dyg_status device_edits_reserve_run(dyg_graph g, uint64_t m);
dyg_status device_edits_order_run(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                  uint32_t m, uint32_t I, const uint32_t** row,
                                  const uint32_t** ent, bool* ok);
// Swap-with-last removal of (src[i], dst[i]) for every i < n (or < *n_dev)
// with flag[i] == 1, in batch order per row.  Each edge must be present.
dyg_status device_deletes_batch(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                const int32_t* flag, uint32_t n, const uint32_t* n_dev);
// The orders of one batch of n updates computed ahead of it (on another
// stream, so they overlap the batches before it): its edge keys (min << 32 |
// max, or all ones for a self-loop) stably sorted with their batch
// positions, and its 2n row entries (entry 2i is u[i]'s row, 2i+1 is v[i]'s)
// stably sorted by row with their entry index; bcnt is the batch's cleared
// array of per-256 kept counts.
struct BatchOrder {
    const unsigned long long* key;  // [n]
    const uint32_t* idx;            // [n]
    const uint32_t* row;            // [2n]
    const uint32_t* ent;            // [2n]
    uint32_t* bcnt;                 // [(n + 255) / 256]
};
// Which presorted entries an edit takes, by batch position i = entry >> 1:
// flag[i] == match, or, when keep is set, keep[i] && found[jmap[i]] == -1
// (an edge kept for the walk, jmap[i] its place in the kept list, that found
// no path).
struct EditTake {
    const int32_t* flag;
    int32_t match;
    const uint8_t* keep;
    const uint32_t* jmap;
    const int32_t* found;
};
// device_edits_batch / device_deletes_batch over a batch whose entries are
// presorted (BatchOrder::row / ent): each row's run of entries holds its
// taken entries in the order the sort of only those entries gives, so one
// thread per row edits the row as the sorted pipeline does, with no sort in
// the batch's own pipeline.  src, dst, w are the batch's arrays (indexed by
// batch position).
dyg_status device_edits_sorted(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                               const float* w, const uint32_t* row, const uint32_t* ent,
                               uint32_t n, const EditTake& tk);
dyg_status device_deletes_sorted(dyg_graph g, const uint32_t* src, const uint32_t* dst,
                                 const uint32_t* row, const uint32_t* ent, uint32_t n,
                                 const EditTake& tk);
// Waits for the stream; updates the host's slot counts; reports the edges
// appended since begin().  The pipeline stays usable afterwards.
dyg_status device_edits_end(dyg_graph g, uint64_t* added);
// The host-side half of end() for a copy of the 5 meta words already taken
// (after the caller's own wait on the stream).
dyg_status device_edits_absorb(dyg_graph g, const unsigned long long* meta, uint64_t* added);

// Walk launchers (walk_kernels.hip).  All pointers are device pointers.
struct WalkLaunch {
    const uint2* rows;
    const uint4* adj;
    const uint32_t* src;
    const uint32_t* dst;
    const float* w;          // insertion weights (nullptr for deletions)
    const uint64_t* eid;     // nullptr -> eid_base + index
    uint64_t eid_base;
    uint32_t n;
    uint32_t walkers;
    uint32_t max_steps;
    float thr;
    uint32_t flags;
    uint32_t edges_per_block;  // work-queue kernel: edges per workgroup (0 = automatic)
    uint64_t seed;
    int32_t* found;
    float* cond;
    uint32_t* winner;        // optional
    uint32_t* hops;          // optional, per edge
    uint32_t* path;          // deletions: [n*max_steps]
    unsigned long long* stats;  // optional [3]: hops, redirects, leaf stops
    const uint32_t* n_dev;   // optional: list length in device memory (<= n; grid sized by n)
    bool nodesc;             // the graph's slots carry no neighbour descriptors
};
// ev_begin / ev_end (optional) take the walk kernel's own start and end
// times (recorded with the dispatch, not as marker packets around it); an
// empty launch records both where it would have run.
hipError_t launch_walk_incremental(const WalkLaunch& a, hipStream_t s,
                                   hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
hipError_t launch_walk_decremental(const WalkLaunch& a, hipStream_t s,
                                   hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);
hipError_t launch_heuristic(const uint2* rows, const uint4* adj, const uint32_t* endpoint,
                            const uint64_t* key, uint32_t n, uint32_t walkers, uint64_t seed,
                            uint32_t* path3, hipStream_t s, const uint32_t* n_dev = nullptr,
                            bool nodesc = false);
// Delta layout (uint32 words): writes (slot, col, wbits, nstart, ndeg), then
// rows (v, start, deg).
hipError_t launch_apply_delta(uint2* rows, uint4* adj, const uint32_t* delta, uint32_t n_write,
                              uint32_t n_row, hipStream_t s);

// Slot of column b in a row r = (first slot, degree), or -1: the first match
// in slot order.  The columns are read sixteen at a time; a round past the
// row's end rereads its last slot (a later duplicate never wins over it), so
// the loads need no masks, and the round's matches are gathered into a bit
// mask before the first one is taken, so no compare sits between two loads:
// a row of up to sixteen slots costs one round trip to memory.
__device__ __forceinline__ int slot_of(const uint4* __restrict__ adj, uint2 r, uint32_t b) {
    const uint4* p = adj + r.x;
    for (uint32_t q = 0; q < r.y; q += 16) {
        uint32_t c[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) c[j] = p[min(q + j, r.y - 1u)].x;
        uint32_t m = 0;
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) m |= (c[j] == b ? 1u : 0u) << j;
        if (m) return (int)(q + (uint32_t)__builtin_ctz(m));
    }
    return -1;
}

// Whether the graph holds edge a-b: both row descriptors are loaded whole
// and together (selecting between the two loaded words keeps the compiler
// from loading the chosen row's start after the degree compare), then the
// shorter row is scanned.
__device__ __forceinline__ bool has_edge(const uint2* __restrict__ rows,
                                         const uint4* __restrict__ adj, uint32_t a, uint32_t b) {
    const unsigned long long xa = reinterpret_cast<const unsigned long long*>(rows)[a];
    const unsigned long long xb = reinterpret_cast<const unsigned long long*>(rows)[b];
    const bool sa = (xa >> 32) <= (xb >> 32);
    const unsigned long long x = sa ? xa : xb;
    return slot_of(adj, make_uint2((uint32_t)x, (uint32_t)(x >> 32)), sa ? b : a) >= 0;
}

// Moves a row's n slots to a fresh range (relocation to the arena tail; the
// ranges do not overlap), eight loads in flight before their stores.
__device__ __forceinline__ void copy_slots(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                           uint32_t n) {
    for (uint32_t z = 0; z < n; z += 8) {
        uint4 t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (z + j < n) t[j] = src[z + j];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (z + j < n) dst[z + j] = t[j];
    }
}

// In a row (first slot, degree) rx, every slot pointing to r gets r's
// descriptor rr.  Only the slots' upper halves are written; the columns are
// read sixteen at a time as independent loads (a round past the row's end
// rereads its last slot and writes nothing there), so a neighbour's row is
// one short chain of loads, not one load per slot.
__device__ __forceinline__ void refresh_pointers(uint4* __restrict__ adj, uint2 rx, uint32_t r,
                                                 uint2 rr) {
    uint4* p = adj + rx.x;
    for (uint32_t k = 0; k < rx.y; k += 16) {
        uint32_t c[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) c[j] = p[min(k + j, rx.y - 1u)].x;
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j)
            if (k + j < rx.y && c[j] == r) *reinterpret_cast<uint2*>(&p[k + j].z) = rr;
    }
}

// Lanes per touched row for the descriptor refresh: the power of two at or
// above the graph's mean degree, within [4, 64].
inline uint32_t refresh_lanes(const dyg_graph_s* g) {
    const uint64_t mean = g->V ? (g->nslots + g->V - 1) / g->V : 1;
    uint32_t l = 4;
    while (l < 64 && l < mean) l <<= 1;
    return l;
}

// One graph's device state for edits made inside a single workgroup
// (stream_small.hip): same arena, capacities and counters as the edit pipeline.
struct BlockGraph {
    uint2* rows;
    uint4* adj;
    uint32_t* cap;
    unsigned long long* meta;  // DevEdits::meta of the graph
    unsigned long long cap_slots;
    uint32_t slack;
    uint32_t lpr;  // lanes per row of its descriptor refresh (refresh_lanes)
    bool nodesc;   // no descriptors kept: no refresh (dyg_graph_s::nodesc)
};
// A batch's results, written by its last kernel straight into pinned host
// memory (device-visible addresses): the outcome codes and 14 counter words,
// [0..3] the walk counters (zeroed once read, ready for the next batch),
// [4..8] G's and [9..13] H's DevEdits::meta.
constexpr int kReportWords = 14;
// One more pinned word after them, cleared by the host before a batch: set
// when a batch of device inputs held a vertex id out of range (the batch then
// changed nothing).
constexpr int kReportBad = 14;
// And after it the device's wall clock (wall_clock64, `dyg_ctx_s::clock_khz`
// ticks per ms) when the batch's last kernel reported it: consecutive
// batches' clocks give their spans without events between their kernels.
constexpr int kReportClock = 15;
struct Report {
    int32_t* cls;                  // [n] outcome codes
    unsigned long long* words;     // [kReportWords]
    const unsigned long long* mg;  // device meta words of G (5)
    const unsigned long long* mh;  // device meta words of H (5)
};
// Copies cls[0..n) and the counters out as described above.
hipError_t launch_report(const Report& r, const int32_t* cls, uint32_t n,
                         unsigned long long* stats, hipStream_t s);
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
    } else if (i == (uint32_t)kReportClock) {
        r.words[i] = wall_clock64();
    }
}


// The per-batch values of a one-workgroup step run as a captured graph,
// whose kernel arguments are fixed at capture: the host writes them into
// pinned memory before each launch, and the pre and post kernels read them
// at entry in place of the same fields of their arguments.
struct StepDesc {
    const uint32_t* u;
    const uint32_t* v;
    const float* w;               // insertions only
    int32_t* rep_cls;             // Report::cls
    unsigned long long* rep_words;  // Report::words
    uint64_t base;
    uint32_t n;
    uint32_t pad;
};

// An insertion batch of at most kSmallBatch edges in two one-workgroup
// kernels around the walk (device pointers; u, v and w may point into
// pinned host memory).
constexpr uint32_t kSmallBatch = 2048;
struct SmallInsArgs {
    BlockGraph G, H;
    const uint32_t* u;
    const uint32_t* v;
    const float* w;
    uint32_t n;
    uint64_t base;   // stream position of the batch's first update
    int key_bits;    // significant bits of the (min << 32 | max) keys
    int row_bits;    // significant bits of a vertex id (and of the pad above them)
    int32_t* cls;    // outcome per update
    uint32_t* sel;   // batch position of each kept edge
    uint32_t* cnt;   // [0]: kept edges, [4]: touched rows
    uint32_t* touched;  // [2 * kSmallBatch] rows touched by the last append
    uint32_t* cu;
    uint32_t* cv;
    float* cw;
    uint64_t* ceid;
    const int32_t* found;  // walk result per kept edge (post kernel)
    unsigned long long* stats;  // walk counters, reported and zeroed by the post kernel
    Report rep;
    const StepDesc* desc;  // when set: u, v, w, n, base and rep's outputs come from here
    uint32_t vcheck;       // nonzero (device inputs): ids must be below it, checked by the
                           // pre kernel before any change; a failing batch changes nothing
    BatchOrder ord;        // key set: the batch's orders, computed ahead (no sorts of the
                           // batch's keys or row entries in the pre and post kernels)
};
hipError_t launch_small_ins_pre(const SmallInsArgs& a, hipStream_t s);
hipError_t launch_small_ins_post(const SmallInsArgs& a, hipStream_t s);

// A deletion batch of at most kSmallBatch edges: one-workgroup kernels
// before and after the walk on G; a recovery of more candidate edges than
// the post kernel holds (small_del_candidates, at most kSmallCand) is left
// to the general kernels (flag sizes[4]).
constexpr uint32_t kSmallCand = 2048;
struct SmallDelArgs {
    BlockGraph G, H;
    const uint32_t* u;
    const uint32_t* v;
    uint32_t n;
    uint64_t base;
    int key_bits;
    int row_bits;
    int32_t* cls;
    uint32_t* sel;
    uint32_t* cnt;  // [0] kept, [1] no-path, [2] endpoints, [3] path hops,
                    // [4], [5], [6] rows listed in touched_g, touched_h, touched_x
    uint32_t* touched_g;  // [2n] rows of G touched by the deletions
    uint32_t* touched_h;  // [2n] rows of H touched by the deletions
    uint32_t* touched_x;  // [2 * candidates] rows of H touched by the recovery
    uint32_t* cu;
    uint32_t* cv;
    uint64_t* ceid;
    const int32_t* found;  // walk results (post kernel)
    const uint32_t* path;  // [kept * max_steps]
    uint32_t max_steps;
    uint32_t hw;           // heuristic walkers per endpoint
    uint64_t seed;         // of the heuristic walks
    uint32_t* pc;          // general recovery's lists (written when it takes over)
    uint32_t* poff;
    uint32_t* ep;
    uint64_t* ekey;
    uint32_t* sizes;       // pinned: kept, no-path, endpoints, path hops, taken-over flag
    unsigned long long* stats;
    Report rep;
    const StepDesc* desc;  // when set: u, v, n, base and rep's outputs come from here
    uint32_t vcheck;       // as SmallInsArgs::vcheck
    BatchOrder ord;        // as SmallInsArgs::ord (the pre kernel's sorts)
};
// A list's one-workgroup batches get their orders (BatchOrder: keys and row
// entries, no kept counts) from one launch, a workgroup per batch, sorting
// exactly as the pre kernel would.
struct SmallOrderJob {
    const uint32_t* u;
    const uint32_t* v;
    unsigned long long* key;
    uint32_t* idx;
    uint32_t* row;
    uint32_t* ent;
    uint32_t n;
    uint32_t pad;
};
// jobs: device array of `count` jobs, all of at most `max_n` updates.
hipError_t launch_small_order(const SmallOrderJob* jobs, uint32_t count, uint32_t max_n,
                              int key_bits, int row_bits, hipStream_t s);
// One job (j.n <= kSmallBatch) passed by value.
hipError_t launch_small_order_one(const SmallOrderJob& j, int key_bits, int row_bits,
                                  hipStream_t s);
// Candidate edges the post kernel of an n-update deletion batch can hold.
uint32_t small_del_candidates(uint32_t n);
hipError_t launch_small_del_pre(const SmallDelArgs& a, hipStream_t s);
hipError_t launch_small_del_post(const SmallDelArgs& a, hipStream_t s);


bool valid_walkers(uint32_t w);
// Validates walkers (power of two, 1..1024) and max_steps (1..32767).
dyg_status check_params(dyg_ctx ctx, const dyg_walk_params* p);

}  // namespace dyg

#define DYG_HIP(ctx, expr)                                                              \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            dyg::set_error((ctx), std::string(#expr) + ": " + hipGetErrorString(_e));   \
            return DYG_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)
