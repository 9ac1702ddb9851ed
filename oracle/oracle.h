/*
 * oracle.h -- CPU restatement of the dyGRASS incremental / decremental
 * non-backtracking random-walk updaters.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker the HIP product is compared
 * against: only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark may load it.  The product library never links or calls it.
 *
 * Every function restates one reference unit; the citation is given beside
 * each declaration (paths relative to the reference repository root).
 * cuRAND (closed NVIDIA library) is replaced by Philox4x32-10 draws keyed by
 * (seed, stream tag, edge id, walker, draw index); the product uses the same
 * keying, so walks are comparable bit for bit.
 */
#ifndef DYG_ORACLE_H
#define DYG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random stream tags (counter word 3, top byte). */
#define ORC_TAG_INSERT 1u
#define ORC_TAG_DELETE 2u
#define ORC_TAG_HEUR 3u

/* Philox4x32-10 block function (Salmon et al., SC'11).  Replaces
 * curand_init/curand at random_walk_incremental/main.cu:82-83,101,126. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Draw number `d` of walker `walker` of update `edge_id` in stream `tag`. */
uint32_t orc_draw(uint64_t seed, uint32_t tag, uint64_t edge_id, uint32_t walker, uint32_t d);

/* Literal restatement of the block tree reduction
 * random_walk_incremental/main.cu:179-201 (same at
 * random_walk_decremental/main.cu:185-205).  Returns the winning thread. */
uint32_t orc_tree_reduce(const float* cond, uint32_t W);

/* ---- graph (per-row slot lists; slot order is what the walks see) ---- */
typedef struct orc_graph orc_graph;
orc_graph* orc_graph_new(uint32_t V, const uint64_t* row_ptr, const uint32_t* col, const float* w);
void orc_graph_free(orc_graph* g);
uint32_t orc_graph_nv(const orc_graph* g);
uint64_t orc_graph_nslots(const orc_graph* g);
void orc_graph_export(const orc_graph* g, uint64_t* row_ptr, uint32_t* col, float* w);
/* Append (u,v,w) then (v,u,w): random_walk_incremental/gpu_graph.cuh:319-325,357-384. */
void orc_graph_insert(orc_graph* g, uint32_t u, uint32_t v, float w);

/* ---- single-edge walks, all W walkers, block reduction ----
 * found_at: winner's hop index (-1 none), cond: winner's 1/R (-1 none),
 * winner: winning walker, hops: sum over walkers of neighbour draws. */
/* random_walk_incremental/main.cu:70-211 */
void orc_walk_inc_edge(const orc_graph* H, uint32_t s, uint32_t t, float w_e, float thr,
                       uint32_t W, uint32_t max_steps, uint64_t seed, uint64_t edge_id,
                       int32_t* found_at, float* cond, uint32_t* winner, uint64_t* hops);
/* random_walk_decremental/main.cu:77-220; path gets path[0..found_at]. */
void orc_walk_dec_edge(const orc_graph* G, uint32_t s, uint32_t t, uint32_t W,
                       uint32_t max_steps, uint64_t seed, uint64_t edge_id, int32_t* found_at,
                       float* cond, uint32_t* winner, uint64_t* hops, uint32_t* path);
/* One walker's full trajectory (vertices visited, starting with s).  Returns
 * the number of vertices written to `traj` (at most max_steps+1); *found_at
 * and *R receive the walker's own outcome.  kind: 1 insertion, 2 deletion. */
uint32_t orc_walk_trace(const orc_graph* g, int kind, uint32_t s, uint32_t t, float w_e,
                        float thr, uint32_t max_steps, uint64_t seed, uint64_t edge_id,
                        uint32_t walker, uint32_t* traj, int32_t* found_at, float* R);
/* random_walk_decremental/main.cu:243-290 with defect D3 resolved (hop two
 * reads n1's own row).  Writes 3 vertices per walker into path3. */
void orc_walk_heuristic(const orc_graph* G, uint32_t u, uint32_t nwalk, uint64_t seed,
                        uint64_t key, uint32_t* path3);

/* ---- batched walks over many edges (OpenMP over edges: the CPU baseline) ---- */
void orc_walk_inc_batch(const orc_graph* H, uint32_t n, const uint32_t* src, const uint32_t* dst,
                        const float* w, const uint64_t* edge_id, float thr, uint32_t W,
                        uint32_t max_steps, uint64_t seed, int threads, int32_t* found_at,
                        float* cond, uint32_t* winner, uint64_t* hops);
void orc_walk_dec_batch(const orc_graph* G, uint32_t n, const uint32_t* src, const uint32_t* dst,
                        const uint64_t* edge_id, uint32_t W, uint32_t max_steps, uint64_t seed,
                        int threads, int32_t* found_at, float* cond, uint32_t* winner,
                        uint64_t* hops, uint32_t* path);

/* `ostream << float` text of f (the reference writer's weight format). */
int orc_format_float(float f, char* buf, int cap);

/* ---- dense graph with the reference's edge map ----
 * Parsed as random_walk_decremental/functions.cpp:69-281 (1-based, weighted;
 * V = v_max+1); removal as gpu_graph.cuh:340-400.  orc_dense_remove returns
 * -1 (edge absent), -2 (a degree would reach 0; nothing changed), 1 (removed). */
typedef struct orc_dense orc_dense;
orc_dense* orc_dense_open(const char* path);
int orc_dense_remove(orc_dense* d, uint32_t u, uint32_t v);
orc_graph* orc_dense_graph(orc_dense* d);
int64_t orc_dense_multiplier(const orc_dense* d);
void orc_dense_free(orc_dense* d);

/* ---- batching (random_walk_incremental/functions.cpp:310-361,
 *      gpu_graph.cuh:135,395-413; random_walk_decremental/gpu_graph.cuh:333-336) ----
 * Writes the start of each snapshot batch into starts[0..n] (starts[n] = m)
 * and returns n. */
uint32_t orc_batches(uint32_t m, uint32_t divisor, uint32_t* starts);

/* ---- whole runs mirroring the two CLIs ----
 * dir holds the reference's files.  Outputs are written into dir; `log_path`
 * (may be NULL) receives the reference's stdout lines without the timing line. */
typedef struct {
    uint64_t n_updates;     /* targets read */
    uint64_t n_added;       /* incremental: edges added to H */
    uint64_t n_batches;
    uint64_t hops;          /* total neighbour draws over all walkers */
    uint64_t sparse_before; /* decremental: |H keys| before */
    uint64_t sparse_after;  /* decremental: |H keys| after */
    uint64_t added_edges;   /* decremental: size of added_edges */
    uint64_t no_path;       /* decremental: deletions with no path */
    uint64_t filtered;      /* decremental: deletions that hit H and were walked */
    double walk_seconds;    /* wall time inside the batch loop spent in walks */
    double total_seconds;   /* wall time of the batch loop */
} orc_run_stats;

/* The incremental batch loop on arrays (random_walk_incremental/main.cu:297-340):
 * H is modified in place; found_at[m] receives every decision. */
int orc_incremental_arrays(orc_graph* H, uint32_t m, const uint32_t* src, const uint32_t* dst,
                           const float* w, uint32_t divisor, float thr, uint32_t W,
                           uint32_t max_steps, uint64_t seed, int threads, int32_t* found_at,
                           orc_run_stats* st);

/* random_walk_incremental/main.cu:240-362 */
int orc_incremental_run(const char* dir, float thr, uint32_t W, uint64_t seed, int threads,
                        const char* log_path, orc_run_stats* st);
/* The decremental loop (random_walk_decremental/main.cu:363-444) on arrays:
 * G as lines in file order (0-based ids; V = largest id + 1), H as an edge
 * list, deletions in file order.  added_u/v receive the added set sorted by
 * key (up to added_cap pairs; st->added_edges gives the full count). */
int orc_decremental_lines(uint32_t V, const uint32_t* g_u, const uint32_t* g_v, const float* g_w,
                          uint64_t g_m, const uint32_t* h_u, const uint32_t* h_v, uint64_t h_m,
                          const uint32_t* d_u, const uint32_t* d_v, uint32_t d_m, uint64_t seed,
                          int threads, orc_run_stats* st, uint32_t* added_u, uint32_t* added_v,
                          uint64_t added_cap);

/* random_walk_decremental/main.cu:318-468 */
int orc_decremental_run(const char* dir, uint64_t seed, int threads, const char* log_path,
                        orc_run_stats* st);

/* ---- mixed insertion/deletion stream ----
 * The C++ restatement of oracle/stream.py (same semantics and order of edits:
 * dynamic_update/gpu_graph.cuh:361-439 preprocessing, :451-465 weight update,
 * :620-700 sparsifier update, with the plain updaters' walks and heuristic).
 * Insertion walks decide only (an edge is found once any walker reaches the
 * target); deletion walks run every walker and recover the winner's path.
 * Walks are OpenMP-parallel over a batch's edges; the graph edits are
 * sequential, in batch order.  Outcome codes as oracle/stream.py. */
typedef struct orc_stream orc_stream;
typedef struct {
    uint64_t updates, batches, hops;
    double walk_seconds;  /* wall time in the walks */
    double total_seconds; /* wall time of orc_stream_apply calls */
} orc_stream_stats;
orc_stream* orc_stream_new(uint32_t V, const uint64_t* g_rp, const uint32_t* g_col,
                           const float* g_w, const uint64_t* h_rp, const uint32_t* h_col,
                           const float* h_w, float thr, uint32_t w_ins, uint32_t w_del,
                           uint32_t steps_ins, uint32_t steps_del, uint32_t heur, uint64_t seed,
                           int threads);
/* op 0 insertion, 1 deletion; out[n] receives outcome codes.  -1 on a bad id. */
int orc_stream_apply(orc_stream* s, int op, const uint32_t* u, const uint32_t* v, const float* w,
                     uint32_t n, int32_t* out);
const orc_graph* orc_stream_graph(const orc_stream* s, int which); /* 0 G, 1 H */
void orc_stream_get_stats(const orc_stream* s, orc_stream_stats* st);
void orc_stream_free(orc_stream* s);

/* ---- parse restatements, exposed for pinning against the reference's own parser ---- */
/* random_walk_incremental/functions.cpp:58-239 (0-based, undirected).  Returns V;
 * fills begin[V+1], adj/weight[2L] when the pointers are non-NULL.  *lines gets L. */
int64_t orc_read_graph0(const char* path, uint64_t* lines, uint64_t* begin, uint32_t* adj,
                        float* weight);
/* random_walk_incremental/functions.cpp:261-370 (1-based -> 0-based, 10-segment layout). */
int64_t orc_read_targets(const char* path, int weighted, uint32_t divisor, uint32_t* nodes,
                         float* weights);

#ifdef __cplusplus
}
#endif
#endif
