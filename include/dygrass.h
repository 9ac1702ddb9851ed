/*
 * dygrass.h -- C ABI of the MI355X-native dyGRASS random-walk updaters.
 *
 * The reference has no library interface: its boundary is the two CLIs
 * (random_walk_incremental/main.cu:240-362, random_walk_decremental/main.cu:318-468)
 * and, inside them, kernel launches on a bitwise copy of a host object.
 * Each entry point below replaces one reference unit; the unit is cited on
 * the declaration.  Conventions:
 *   - every call returns dyg_status (0 = success, < 0 = error) and never
 *     exits the process; dyg_last_error() describes the last failure;
 *   - vertex ids are 0-based uint32, weights are IEEE float32;
 *   - host-pointer entry points copy inputs in and results out and return
 *     after the device work has finished; *_dev entry points take device
 *     pointers, enqueue on the context's stream and return at once;
 *   - the caller owns every array it passes; handles are opaque and owned
 *     by the library until the matching *_free / *_destroy;
 *   - one context = one HIP stream = one host thread.  Distinct contexts
 *     may be used from distinct threads concurrently.
 */
#ifndef DYGRASS_H
#define DYGRASS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dyg_status;
#define DYG_OK 0
#define DYG_ERR_ARG (-1)       /* invalid argument (sizes, ids, walker count) */
#define DYG_ERR_HIP (-2)       /* HIP runtime failure */
#define DYG_ERR_NOMEM (-3)     /* device or host allocation failed */
#define DYG_ERR_IO (-4)        /* file could not be opened / written */
#define DYG_ERR_FORMAT (-5)    /* malformed .mtx input */
#define DYG_ERR_STATE (-6)     /* graph invariant violated (e.g. degree would reach 0) */
#define DYG_ERR_ADMISSION (-7) /* request rejected by the memory budget */

typedef struct dyg_ctx_s* dyg_ctx;
typedef struct dyg_graph_s* dyg_graph;
typedef struct dyg_dual_s* dyg_dual;
typedef struct dyg_stream_s* dyg_stream;
typedef struct dyg_serve_s* dyg_serve;

/* ------------------------------------------------------------------ */
/* Context                                                            */
/* ------------------------------------------------------------------ */
/* Replaces the implicit CUDA context / default stream the reference uses
 * (cudaMalloc at random_walk_incremental/main.cu:276). */
dyg_status dyg_ctx_create(int device, dyg_ctx* out);
void dyg_ctx_destroy(dyg_ctx ctx);
/* Message of the last failed call on ctx (ctx may be NULL for calls that
 * failed before a context existed).  Never NULL. */
const char* dyg_last_error(dyg_ctx ctx);
const char* dyg_version(void);
/* Stream of the context (a hipStream_t), for callers that enqueue their own
 * work around the *_dev entry points. */
void* dyg_ctx_stream(dyg_ctx ctx);
/* Forces all work enqueued on ctx to finish. */
dyg_status dyg_ctx_sync(dyg_ctx ctx);
/* Device buffers for the *_dev entry points (inputs kept resident in HBM).
 * Copies are ordered on the context stream and complete before returning.
 * No counterpart in the reference (its buffers live inside its classes). */
dyg_status dyg_dev_alloc(dyg_ctx ctx, size_t bytes, void** out);
dyg_status dyg_dev_free(dyg_ctx ctx, void* p);
dyg_status dyg_copy_to_device(dyg_ctx ctx, void* dst, const void* src, size_t bytes);
dyg_status dyg_copy_to_host(dyg_ctx ctx, void* dst, const void* src, size_t bytes);
dyg_status dyg_copy_on_device(dyg_ctx ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------ */
/* Device graph store                                                 */
/* ------------------------------------------------------------------ */
/* Upload an undirected graph given as per-vertex slot lists (CSR).  Slot
 * order is preserved and is what the walks index.  Each row gets
 * `slack` spare slots so inserts append in place.  keyed != 0 also builds
 * the reference's edge map (key a*10^digits(V)+b -> slot pair) needed by
 * dyg_graph_delete.
 * Replaces gpu_graph::gpu_graph + createNeighborArray + copyToDevice
 * (random_walk_incremental/gpu_graph.cuh:123-162,205-275) and
 * GPU_Dual_Graph's dense layout (random_walk_decremental/gpu_graph.cuh:179-268). */
dyg_status dyg_graph_upload(dyg_ctx ctx, uint32_t V, const uint64_t* row_ptr,
                            const uint32_t* col, const float* w, uint32_t slack, int keyed,
                            dyg_graph* out);
/* V, directed slots in use (sum of degrees), slots allocated on the device. */
dyg_status dyg_graph_info(dyg_graph g, uint32_t* V, uint64_t* nslots, uint64_t* capacity);
/* Copy the current state back as compact CSR in slot order (row_ptr[V+1],
 * col/w[nslots]).  Reads the device copy, not the host mirror. */
dyg_status dyg_graph_download(dyg_graph g, uint64_t* row_ptr, uint32_t* col, float* w);
/* Append edges in order: (u,v,w) goes to slot deg(u) of u, then (v,u,w) to
 * slot deg(v) of v; a full row is relocated with its order kept.
 * Replaces updateGraphFromResult/manageExtraNeighborsData
 * (random_walk_incremental/gpu_graph.cuh:292-384) without the full re-upload. */
dyg_status dyg_graph_insert(dyg_graph g, const uint32_t* u, const uint32_t* v, const float* w,
                            uint32_t n);
/* Delete edges in order by swapping the row's last slot into the hole (keyed
 * graphs only).  hit[i] (optional) = 1 when edge i existed and was removed.
 * Fails with DYG_ERR_STATE, leaving the graph unchanged, when a deletion
 * would leave a vertex with degree 0.
 * Replaces the dense half of removeAndUpdateDelEdges
 * (random_walk_decremental/gpu_graph.cuh:338-400). */
dyg_status dyg_graph_delete(dyg_graph g, const uint32_t* u, const uint32_t* v, uint32_t n,
                            int8_t* hit);
void dyg_graph_free(dyg_graph g);
/* Consistency check of the device state: counts live slots whose copy of the
 * neighbour's row descriptor differs from that row (0 when consistent). */
dyg_status dyg_graph_check(dyg_graph g, uint64_t* bad_slots);

/* ------------------------------------------------------------------ */
/* Walks                                                              */
/* ------------------------------------------------------------------ */
#define DYG_WALK_NO_PRUNE 1u /* run every walker to its own end (no early stop) */
/* Launch shape.  Default: work-queue kernel (each workgroup owns a run of
 * edges and hands their (edge, walker) tasks to lanes as lanes free up).
 * DYG_WALK_PER_EDGE: one workgroup per edge, one lane per walker (the
 * reference's <<<edges, walkers>>> shape).  Results are identical. */
#define DYG_WALK_PER_EDGE 2u
/* dyg_incremental_update: plan each batch's appends on the host and ship
 * them as a delta (the reference's host-ordered update between batches).
 * Default: the appends are compacted, ordered and applied on the device, so
 * batches follow each other on the stream with no host round trip.  Same
 * resulting graph either way (slot order per row). */
#define DYG_RUN_HOST_EDITS 4u
/* Insertion walks, work-queue kernel: once one walker of an edge has found
 * its target, the edge's other walkers stop.  The decision (found or not,
 * which is all random_walk_incremental/gpu_graph.cuh:302 reads) is exactly
 * the reference's; found_at and cond then describe a walker that found the
 * target, not necessarily the reference's winner.  The stream forms always
 * use it (they report outcome codes only). */
#define DYG_WALK_DECISION_ONLY 8u

typedef struct {
    uint32_t walkers;   /* walkers per edge, power of two in [64,1024] (argv[3]; 512 in decremental) */
    uint32_t max_steps; /* walk bound, reference: 100 (random_walk_incremental/main.cu:44) */
    float threshold;    /* insertion distortion threshold (argv[2]) */
    uint32_t flags;     /* DYG_WALK_* */
    uint64_t seed;      /* Philox key of the draw streams */
    uint32_t block_edges; /* work-queue kernel: edges owned by one workgroup (0 = automatic,
                             max 128); a launch-shape knob, results do not depend on it */
    uint32_t reserved;
} dyg_walk_params;

typedef struct {
    uint64_t hops;       /* neighbour draws executed over all walkers */
    uint64_t redirects;  /* non-backtracking redirects to the row's last slot */
    uint64_t leaf_stops; /* walks ended by a degree-1 row */
    double kernel_ms;    /* device time of the walk kernel(s), HIP events */
} dyg_walk_stats;

/* Insertion walks, one edge per workgroup, walkers as wave64 lanes.
 * Per edge: found_at = winner's hop index or -1, cond = winner's 1/R or -1,
 * winner (optional) = winning walker, hops (optional) = draws of the edge.
 * edge_id (optional, else 0..n-1) keys the draw streams.
 * Replaces random_walk (random_walk_incremental/main.cu:70-211) + get_result
 * (random_walk_incremental/gpu_graph.cuh:551-554). */
dyg_status dyg_walk_incremental(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                const float* w, const uint64_t* edge_id, uint32_t n,
                                const dyg_walk_params* p, int32_t* found_at, float* cond,
                                uint32_t* winner, uint32_t* hops, dyg_walk_stats* st);
/* dyg_walk_incremental without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc).  Inputs, results and
 * counters are staged through one fixed 40 MiB device array that the library
 * holds on each GPU it runs on until the process ends, so n is limited to
 * DYG_NOALLOC_MAX_EDGES (DYG_ERR_ARG beyond).  Same results as
 * dyg_walk_incremental.  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (from several threads or contexts) race. */
#define DYG_NOALLOC_MAX_EDGES 1048576u
dyg_status dyg_walk_incremental_noalloc(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                        const float* w, const uint64_t* edge_id, uint32_t n,
                                        const dyg_walk_params* p, int32_t* found_at, float* cond,
                                        uint32_t* winner, uint32_t* hops, dyg_walk_stats* st);
/* Deletion walks; path[i*max_steps + 0..found_at[i]] receives the winner's
 * vertex path when found.  Replaces random_walk_decremental
 * (random_walk_decremental/main.cu:77-220) + get_result (gpu_graph.cuh:831-841). */
dyg_status dyg_walk_decremental(dyg_graph G, const uint32_t* src, const uint32_t* dst,
                                const uint64_t* edge_id, uint32_t n, const dyg_walk_params* p,
                                int32_t* found_at, float* cond, uint32_t* winner, uint32_t* hops,
                                uint32_t* path, dyg_walk_stats* st);
/* Two-hop recovery walks: `walkers` per endpoint, path3[(i*walkers+j)*3+0..2]
 * = (u, n1, n2); key[i] keys endpoint i's draws.  Replaces random_walk_small
 * (random_walk_decremental/main.cu:243-290) with its hop-two row fixed. */
dyg_status dyg_walk_heuristic(dyg_graph G, const uint32_t* endpoint, const uint64_t* key,
                              uint32_t n, uint32_t walkers, uint64_t seed, uint32_t* path3);

/* Device-pointer forms of the two walks: all arrays are device pointers,
 * work is enqueued on the context stream; `hops_total` (device, optional)
 * accumulates draws.  Same results as the host forms. */
dyg_status dyg_walk_incremental_dev(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                    const float* w, const uint64_t* edge_id, uint32_t n,
                                    const dyg_walk_params* p, int32_t* found_at, float* cond,
                                    uint32_t* winner, uint32_t* hops,
                                    unsigned long long* hops_total);
dyg_status dyg_walk_decremental_dev(dyg_graph G, const uint32_t* src, const uint32_t* dst,
                                    const uint64_t* edge_id, uint32_t n,
                                    const dyg_walk_params* p, int32_t* found_at, float* cond,
                                    uint32_t* winner, uint32_t* hops, uint32_t* path,
                                    unsigned long long* hops_total);

/* ------------------------------------------------------------------ */
/* Whole incremental update (the incremental CLI's batch loop)        */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t updates;      /* edges processed */
    uint64_t added;        /* edges added to the sparsifier (no path found) */
    uint64_t batches;      /* snapshot batches */
    uint64_t hops;         /* neighbour draws over all walkers */
    uint64_t redirects;
    uint64_t leaf_stops;
    double walk_ms;        /* device time in walk kernels (HIP events) */
    double update_ms;      /* host planning + device apply of graph edits */
    double total_ms;       /* wall time of the whole loop */
} dyg_run_stats;

/* Split m insertions into ceil(m/divisor)-sized snapshot batches; for each
 * batch run the walks on H as it stood at batch start, then append the
 * no-path edges in batch order.  found_at[i] (optional) = outcome of edge i;
 * edge i's draws are keyed by i.  Replaces the loop of
 * random_walk_incremental/main.cu:297-340 with gpu_graph.cuh:292-413. */
dyg_status dyg_incremental_update(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                  const float* w, uint32_t m, uint32_t divisor,
                                  const dyg_walk_params* p, int32_t* found_at,
                                  dyg_run_stats* st);
/* Same loop with inputs already in device memory: src/dst/w and the outputs
 * found_at/cond[m] are device pointers.  Returns when the run has finished. */
dyg_status dyg_incremental_update_dev(dyg_graph H, const uint32_t* src, const uint32_t* dst,
                                      const float* w, uint32_t m, uint32_t divisor,
                                      const dyg_walk_params* p, int32_t* found_at, float* cond,
                                      dyg_run_stats* st);
/* Pre-allocate for device-side edits: slot-arena room for `max_inserts`
 * appended edges, the ordering buffers for batches of up to `max_batch`
 * edges and for a run of `max_inserts` updates sorted at once, so that a
 * following dyg_incremental_update* allocates nothing.
 * No counterpart in the reference (its arena is fixed at 2x, unchecked:
 * random_walk_incremental/gpu_graph.cuh:210,357-384). */
dyg_status dyg_graph_reserve(dyg_graph g, uint64_t max_inserts, uint32_t max_batch);

/* ------------------------------------------------------------------ */
/* Dual graph for deletions (GPU_Dual_Graph)                          */
/* ------------------------------------------------------------------ */
/* G (dense, on the device, keyed) plus H as a host key set, as in
 * GPU_Dual_Graph (random_walk_decremental/gpu_graph.cuh:79-268).  G is given
 * as CSR in slot order; H as an edge list (duplicates collapse). */
dyg_status dyg_dual_create(dyg_ctx ctx, uint32_t V, const uint64_t* g_row_ptr,
                           const uint32_t* g_col, const float* g_w, const uint32_t* h_u,
                           const uint32_t* h_v, uint64_t h_m, dyg_dual* out);
/* Key multiplier 10^digits(V) of the edge keys a*mult+b (a<b). */
int64_t dyg_dual_key_multiplier(dyg_dual d);
/* Process deletions in order (removeAndUpdateDelEdges,
 * random_walk_decremental/gpu_graph.cuh:328-477): status[i] = 0 not in G,
 * 1 removed from G only, 2 removed from G and H (queued for a walk).  ids[i]
 * keys the walk of deletion i.  *n_queued = deletions queued for the walk. */
dyg_status dyg_dual_remove(dyg_dual d, const uint32_t* u, const uint32_t* v, const uint64_t* ids,
                           uint32_t n, int8_t* status, uint32_t* n_queued);
/* Walk the queued deletions on G, then add the winners' path edges to H and
 * to the added set in queue order (random_walk_decremental/main.cu:392-397,
 * gpu_graph.cuh:509-559).  *no_path = queued edges without a path (they are
 * queued for dyg_dual_heuristic), *added = edges newly added to H. */
dyg_status dyg_dual_walk(dyg_dual d, const dyg_walk_params* p, uint32_t* no_path,
                         uint64_t* added, dyg_walk_stats* st);
/* Two-hop recovery for the no-path endpoints (main.cu:400-408,
 * gpu_graph.cuh:577-615): inserts (u,n1),(n1,n2) into H only. */
dyg_status dyg_dual_heuristic(dyg_dual d, uint32_t walkers, uint64_t seed);
uint64_t dyg_dual_sparse_size(dyg_dual d);
uint64_t dyg_dual_added_count(dyg_dual d);
/* The added set as (u,v) pairs, u < v, sorted by key. */
dyg_status dyg_dual_added_edges(dyg_dual d, uint32_t* u, uint32_t* v);
dyg_graph dyg_dual_dense(dyg_dual d);
void dyg_dual_free(dyg_dual d);

/* ------------------------------------------------------------------ */
/* Mixed insertion/deletion update stream (device-resident G and H)   */
/* ------------------------------------------------------------------ */
/* The stream form of dynamic_update (dynamic_update/main.cu:92-124 batch
 * loop; preprocessStreamEdges dynamic_update/gpu_graph.cuh:361-439, weight
 * update :451-465 with the weight, not the neighbour id, updated;
 * updateSparsiferFromResult :620-700), with the plain updaters' walks,
 * reduction and heuristic.  One batch is one operation kind and one snapshot:
 *   insertion: an edge already in G gets weight += w in G (and in H when
 *   there); a new edge is appended to G and walks in H; without a path it is
 *   appended to H.
 *   deletion: an edge not in G is skipped; otherwise it leaves G and, when in
 *   H, leaves H and walks in G (after all of the batch's deletions) from
 *   min(a,b) to max(a,b); found paths' hops, then the heuristic edges of
 *   no-path deletions, join H (when absent) with G's weight.
 * Duplicates in a batch act in batch order.  G and H must be simple graphs,
 * H inside G.  Edge ids of the draw streams: the update's position in the
 * whole stream (dyg_stream_position before the batch + index). */
#define DYG_STREAM_INSERT 0
#define DYG_STREAM_DELETE 1
/* per-update outcome codes */
#define DYG_OUT_WEIGHT 1     /* insertion of an existing edge: weight update */
#define DYG_OUT_FOUND 2      /* insertion: path found, H unchanged */
#define DYG_OUT_ADDED 3      /* insertion: no path, appended to H */
#define DYG_OUT_SKIP 4       /* deletion of an edge not in G (or repeated) */
#define DYG_OUT_G_ONLY 5     /* deletion from G only */
#define DYG_OUT_RECOVERED 6  /* deletion from G and H, path hops joined H */
#define DYG_OUT_HEURISTIC 7  /* deletion from G and H, no path: heuristic edges joined H */
#define DYG_OUT_INVALID 8    /* self-loop */
typedef struct {
    dyg_walk_params ins;    /* insertion walks (threshold used) */
    dyg_walk_params del;    /* deletion walks (threshold ignored) */
    uint32_t heur_walkers;  /* heuristic walkers per endpoint (the reference: 16) */
    uint32_t max_batch;     /* largest batch apply() accepts */
    uint64_t max_inserts;   /* arena reservation hint: appends expected per graph */
    uint32_t slack;         /* spare slots given to a relocated row */
    uint32_t flags;         /* DYG_STREAM_* bits, or 0 */
} dyg_stream_params;
/* The reference's shape: host edge maps and host-planned edits, one
 * one-workgroup-per-edge walk launch per batch, waiting for every result
 * (same results as the default device-resident form). */
#define DYG_STREAM_PLAIN 1u
/* Device form only: always use the general per-batch pipeline (by default
 * batches of at most 2048 updates take the one-workgroup path). */
#define DYG_STREAM_GENERAL_ONLY 2u
/* Device form only: the dense graph G keeps no neighbour descriptors in its
 * slots, so its edits need no descriptor refresh and deletion walks read the
 * row table once per hop (a second dependent load).  Same results; faster
 * when refreshing G after insertions costs more than the longer deletion
 * hops (high-degree G, insertion-heavy streams). */
#define DYG_STREAM_DENSE_NO_DESC 4u
/* Device form, dyg_stream_apply only: a batch of at most 2048 updates runs as
 * a hipGraph of its one-workgroup step (pre kernel, descriptor refreshes,
 * walk with its timing events, post kernel, refresh), captured once per
 * operation and size class (16, 64, 256, 1024, 2048 updates) with grids for
 * the class maximum, and replayed with the batch's arguments set on its pre
 * and post kernel nodes.  A class whose capture the runtime refuses launches
 * directly.  Same results.  (No reference counterpart: the reference launches
 * each kernel on the default stream.) */
#define DYG_STREAM_GRAPH 8u
/* Device form only: the one-workgroup path takes batches of at most 1024
 * updates (one per lane) instead of 2048; larger ones take the general
 * pipeline, whose kernels spread a batch's row scans over the whole GPU.
 * Same results; a shape knob of the autotuner.  (No reference counterpart.) */
#define DYG_STREAM_SMALL_1024 16u
typedef struct {
    uint64_t updates, weight_updates, found, added, skipped, g_only, recovered, heuristic,
        invalid;
    uint64_t candidates;  /* deletion batches: recovery edges examined */
    uint64_t hops;        /* neighbour draws of the batch's walks */
    double walk_ms;       /* walk kernel (HIP events) */
    double total_ms;      /* whole apply() call, host clock */
    uint64_t graph;       /* 1 when the batch ran as a captured step graph (its walk_ms
                             then times the whole step: the walk is a node inside it) */
    uint64_t redirects;   /* non-backtracking redirects to a row's last slot (each one
                             more slot read) in the batch's walks */
} dyg_stream_stats;
/* G and H in CSR (slot order as given), uploaded and owned by the stream.
 * The device form reserves both slot arenas at creation for max_inserts
 * appends and for what each call keeps in hand for the next (a largest batch
 * in G, a deletion batch's recovery candidates in H), and starts one host
 * helper thread (a list call hands it the orders of its larger batches; it
 * ends with dyg_stream_free). */
dyg_status dyg_stream_create(dyg_ctx ctx, uint32_t V, const uint64_t* g_row_ptr,
                             const uint32_t* g_col, const float* g_w, const uint64_t* h_row_ptr,
                             const uint32_t* h_col, const float* h_w,
                             const dyg_stream_params* params, dyg_stream* out);
/* Apply one batch (host arrays; w ignored for deletions).  outcome[i]
 * (optional) receives DYG_OUT_* of update i.  Returns when the batch is done. */
dyg_status dyg_stream_apply(dyg_stream s, int op, const uint32_t* u, const uint32_t* v,
                            const float* w, uint32_t n, int32_t* outcome, dyg_stream_stats* st);
/* Same, with the batch already in device memory: u, v, w are device pointers
 * (w ignored for deletions); outcome stays a host array.  The device checks
 * the ids (a batch of at most 2048 updates inside its first kernel, a larger
 * one before it starts); a batch holding an id >= V fails with DYG_ERR_ARG
 * and changes nothing.  No reference counterpart: the reference reads its
 * batches from files into host memory (random_walk_incremental/functions.cpp:
 * 261-370) and copies them per batch (gpu_graph.cuh:395-413). */
dyg_status dyg_stream_apply_dev(dyg_stream s, int op, const uint32_t* u, const uint32_t* v,
                                const float* w, uint32_t n, int32_t* outcome,
                                dyg_stream_stats* st);
/* A list of nb batches in one call, all in device memory: batch b is op[b]
 * over updates [off[b], off[b+1]) of u, v and w (w entries of deletion
 * batches are ignored; w may be NULL when every batch deletes); the results
 * of nb calls of dyg_stream_apply_dev in order.  The batches are enqueued
 * back to back: the host waits only where a deletion batch's recovery list
 * sizes must reach it, and once at the end; each batch reports into its own
 * region, and an insertion batch times its walk with events taken with the
 * walk's dispatch (a deletion batch's walk_ms is 0 here: events on its
 * dispatch would cost more than its walk's share of the time).  Each
 * batch's orders (its sorted keys and row entries) are computed ahead on a
 * second stream.  outcome (optional, host) receives every update's code at
 * its position, st (optional) nb per-batch stats (total_ms: the batch's
 * span on the device clock, from the previous batch's end).  Offsets, and
 * then every vertex id of the list (on the device), are checked first
 * (DYG_ERR_ARG, nothing applied); a later failure stops the call with the
 * batches before it applied. */
dyg_status dyg_stream_apply_many_dev(dyg_stream s, uint32_t nb, const int* op,
                                     const uint64_t* off, const uint32_t* u, const uint32_t* v,
                                     const float* w, int32_t* outcome, dyg_stream_stats* st);
/* 0: G, 1: H (owned by the stream; use with dyg_graph_info/download/check). */
dyg_graph dyg_stream_graph(dyg_stream s, int which);
/* Updates applied so far (the edge id of the next update). */
uint64_t dyg_stream_position(dyg_stream s);
void dyg_stream_free(dyg_stream s);

/* ------------------------------------------------------------------ */
/* Served form: request loop over a stream (no reference counterpart;  */
/* the reference runs one batch at a time, dynamic_update/main.cu:92)  */
/* ------------------------------------------------------------------ */
/* `slots` requests of at most `max_request` updates may be admitted at
 * once; budget_bytes (0 = device memory size) bounds resident state plus the
 * worst-case growth of the admitted requests.  Requests are applied in
 * admission order with the stream's per-batch semantics, so each result
 * equals what dyg_stream_apply gives the same sequence.  While a server is
 * open, do not call dyg_stream_apply on its stream. */
typedef struct {
    uint32_t n;          /* updates in the request */
    double queue_ms;     /* submit -> launch */
    double latency_ms;   /* submit -> results on the host */
    uint64_t hops;       /* neighbour draws of its walks */
    uint64_t redirects;  /* non-backtracking redirects of its walks */
    double walk_ms;      /* its walk kernel's own duration on the device */
} dyg_serve_result;
dyg_status dyg_serve_open(dyg_stream s, uint32_t slots, uint32_t max_request,
                          uint64_t budget_bytes, dyg_serve* out);
/* Blocks until the request is admitted (a slot and memory free); returns
 * DYG_ERR_ADMISSION when the request cannot fit the budget (at once, or
 * after the server's resident state grew while it waited), DYG_ERR_STATE
 * when the server is closing, and the server's failure status once a
 * request has failed (the server then stops: no later request is applied).
 * The arrays are copied before the call returns.  A failed input copy fails
 * the call with DYG_ERR_HIP, applies nothing and stops the server. */
dyg_status dyg_serve_submit(dyg_serve v, int op, const uint32_t* u, const uint32_t* v_,
                            const float* w, uint32_t n, uint64_t* ticket);
/* dyg_serve_submit with the request already in device memory: u, v_ and w
 * are device pointers (w ignored for deletions), ready when the call is made
 * and left unchanged until dyg_serve_wait returns for the ticket (the
 * request reads them in place; nothing is copied).  Its ids are checked on
 * the device (on the server's copy stream, beside the requests in flight)
 * before admission: an id >= V fails the call with DYG_ERR_ARG and the server
 * goes on.  No reference counterpart (the reference reads batches from files
 * into host memory, random_walk_incremental/functions.cpp:261-370). */
dyg_status dyg_serve_submit_dev(dyg_serve v, int op, const uint32_t* u, const uint32_t* v_,
                                const float* w, uint32_t n, uint64_t* ticket);
/* Testing hook: the next admitted request that copies its input reports the
 * copy as failed without issuing it (the failure path of dyg_serve_submit,
 * forced from the host).  No reference counterpart. */
dyg_status dyg_serve_fail_next_copy(dyg_serve v);
/* Waits for one request; outcome[i] (optional) receives DYG_OUT_* codes. */
dyg_status dyg_serve_wait(dyg_serve v, uint64_t ticket, int32_t* outcome,
                          dyg_serve_result* res);
/* Drains the queue and releases the slots (the stream stays usable). */
void dyg_serve_close(dyg_serve v);

/* ------------------------------------------------------------------ */
/* .mtx I/O with the reference's parsing rules                        */
/* ------------------------------------------------------------------ */
typedef struct dyg_mtx_s* dyg_mtx;
/* base 0: random_walk_incremental/functions.cpp:58-239 (read_graph);
 * base 1: random_walk_decremental/functions.cpp:69-281 (V = v_max+1).
 * Slots are laid out in file order; each line (a,b,w) appends b to a's
 * row then a to b's row. */
dyg_status dyg_mtx_read_graph(const char* path, int base, int weighted, dyg_mtx* out);
/* Edge list in file order: 1-based ids made 0-based
 * (random_walk_incremental/functions.cpp:261-370). */
dyg_status dyg_mtx_read_edges(const char* path, int weighted, dyg_mtx* out);
/* Tolerant reader (the loader of dynamic_update/functions.cpp:35-291,
 * restated; the reference-exact readers above stay the default): leading
 * '%' header lines (a MatrixMarket banner's "pattern" / "symmetric" are
 * honoured), an optional "n n k" dimensions line (taken as one when the data
 * that follows fits n vertices), 0- or 1-based ids (from the smallest id),
 * weighted or pattern lines, Laplacian input (any negative weight: diagonal
 * dropped, edge weight = -value), and full matrices (the first off-diagonal
 * entry also stored transposed: only entries above the diagonal kept).
 * Self-loops are dropped.  Gives 0-based edges and a CSR in file order. */
typedef struct {
    int has_header, has_dimensions, weighted, laplacian, triangular, base;
    uint64_t entries;  /* data lines read (after the dimensions line) */
} dyg_mtx_info;
dyg_status dyg_mtx_read_auto(const char* path, dyg_mtx_info* info, dyg_mtx* out);
uint32_t dyg_mtx_nv(dyg_mtx m);
uint64_t dyg_mtx_nlines(dyg_mtx m);
const uint64_t* dyg_mtx_row_ptr(dyg_mtx m); /* [V+1], graphs only */
const uint32_t* dyg_mtx_col(dyg_mtx m);     /* [2*lines], graphs only */
const float* dyg_mtx_w(dyg_mtx m);          /* [2*lines], graphs only */
const uint32_t* dyg_mtx_line_u(dyg_mtx m);  /* [lines] first id of each line (0-based) */
const uint32_t* dyg_mtx_line_v(dyg_mtx m);  /* [lines] */
const float* dyg_mtx_line_w(dyg_mtx m);     /* [lines] (1.0 when unweighted) */
void dyg_mtx_free(dyg_mtx m);
/* Write lines "u+1 v+1 w" with w printed as C++ `ostream << float` does
 * (6 significant digits) -- random_walk_incremental/gpu_graph.cuh:427-454. */
dyg_status dyg_mtx_write_weighted(const char* path, const uint32_t* u, const uint32_t* v,
                                  const float* w, uint64_t n, int append);

#ifdef __cplusplus
}
#endif
#endif /* DYGRASS_H */
