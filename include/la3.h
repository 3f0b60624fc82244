/*
 * la3.h — C ABI of the MI355X (gfx950) engine for LA3's vertex-program path.
 *
 * The reference boundary is the header-only C++ template pair
 *   Graph<W>::load_directed / load_undirected      (src/matrix/graph.h:38-47)
 *   VertexProgram<W,M,A,S>::initialize / execute   (src/vprogram/vertex_program.h:33-220)
 * with the graph-analytics programs of src/apps/graph_analytics/ (degree, pr, bfs,
 * sssp, cc, tc). Device code
 * cannot call host virtuals, so each program is selected by its entry point
 * here; the C++ drop-in wrappers in include/la3.hpp keep the reference's class
 * and method names on top of these calls.
 *
 * Conventions
 *  - Every call returns la3_status; la3_last_error(ctx) gives a message for the
 *    last failure on that context. No C++ exception crosses this boundary.
 *  - The library owns all device memory; the caller owns every host buffer.
 *  - Per-vertex outputs are indexed by ORIGINAL vertex id and have
 *    la3_graph_info.nslots entries (ids 0..n, the one-rank tile of
 *    src/matrix/matrix2d.hpp:6-8 with n/1+1 slots).
 *  - One host thread per context at a time (calls on a context serialize on its
 *    stream); results do not depend on batch composition.
 */
#ifndef LA3_H
#define LA3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LA3_OK = 0,
  LA3_EINVAL = 1,   /* bad argument: null pointer, id out of range, wrong graph mode */
  LA3_ENOMEM = 2,   /* device or host allocation failed, or admission rejected */
  LA3_EIO = 3,      /* file could not be opened / read / stat'ed (graph.hpp:69-98) */
  LA3_EDEVICE = 4,  /* a HIP runtime call failed */
  LA3_ENOTSUP = 5,  /* size beyond what this build supports (e.g. nnz >= 2^32) */
  LA3_EAGAIN = 6    /* serving: no completed query yet */
} la3_status;

/* Load modes: the flag combinations of Graph<W>::load_directed(reverse_edges,
 * remove_cycles) and load_undirected (src/matrix/graph.hpp:224-246). */
typedef enum {
  LA3_DIRECTED = 0,          /* load_directed(reverse=false): stores A^T, messages flow a->b */
  LA3_REVERSED = 1,          /* load_directed(reverse=true):  messages flow b->a */
  LA3_UNDIRECTED = 2,        /* load_undirected: both orientations */
  LA3_ACYCLIC = 3,           /* load_directed(false, remove_cycles=true): low id -> high id */
  LA3_REVERSED_ACYCLIC = 4   /* load_directed(true, true): high id -> low id */
} la3_mode;

/* Vertex-id hashing (src/matrix/graph.h:10-15). */
typedef enum { LA3_HASH_NONE = 0, LA3_HASH_BUCKET = 1 } la3_hashing;

typedef struct la3_ctx la3_ctx;
typedef struct la3_graph la3_graph;

typedef struct {
  uint32_t n;          /* nvertices as passed (or read from the header) */
  uint32_t nslots;     /* n + 1 */
  uint32_t nregular;   /* vertex classes (src/matrix/processed_matrix2d.hpp:182-229) */
  uint32_t nsink;
  uint32_t nsource;
  uint32_t nisolated;
  uint32_t ncols;      /* non-empty columns = nregular + nsource */
  uint32_t directed;
  uint32_t weighted;
  uint32_t mode;
  uint64_t nnz;        /* deduplicated stored entries */
  uint64_t nnz_regular;/* entries whose row is a regular vertex */
  uint64_t nnz_sink;   /* entries whose row is a sink vertex */
  uint64_t nrecords;   /* records read from the input (before self-loop/dedup) */
  uint64_t device_bytes;
} la3_graph_info;

/* ------------------------------------------------------------ context --- */
/* Creates a context on HIP device `device`: one stream, device scratch.
 * la3_ctx_destroy releases the caller's reference; the context is freed when
 * its last graph is freed.
 * device == -1 creates a host-only context: graphs are ingested and their
 * layout can be exported (la3_graph_info_get / slots / csc), but programs
 * return LA3_EDEVICE. */
la3_status la3_ctx_create(int device, la3_ctx** out);
void la3_ctx_destroy(la3_ctx* ctx);
const char* la3_last_error(const la3_ctx* ctx);
/* Waits for all work queued on the context's stream. */
la3_status la3_ctx_synchronize(la3_ctx* ctx);
/* Frees the device scratch the context caches between calls (per-query
 * state, work lists, outputs left in device memory, captured step graphs):
 * later calls allocate again. LA3_EAGAIN while a session runs on it. */
la3_status la3_ctx_release_scratch(la3_ctx* ctx);
/* Where la3_graph_load_* builds a graph of a device context: on the device
 * (default: records uploaded, classes / slots / sorted DCSC and row-major
 * companion built with device scans and radix sorts) or on the host
 * (multi-threaded builder). Both give the same arrays. */
la3_status la3_ctx_set_ingest(la3_ctx* ctx, int on_device);

/* -------------------------------------------------------------- graphs --- */
/* Graph<W>::load_binary (src/matrix/graph.hpp:38-214) at np=1: reads the LA3
 * binary file (u32 pairs, or u32 triples when weighted != 0), drops
 * self-loops, orients per `mode`, hashes ids, deduplicates with the first
 * record winning, classifies vertices and builds the device-resident DCSC and
 * its row-major companion. nvertices == 0 means the file starts with the
 * 16-byte header {u32 n-1, u32 m-1, u64 nnz}. Ids must lie in [0, n]. */
la3_status la3_graph_load_file(la3_ctx* ctx, const char* path, uint32_t nvertices, int weighted,
                               la3_mode mode, la3_hashing hashing, la3_graph** out);
/* Same, from records already in host memory (nrecords records of 2 or 3 u32). */
la3_status la3_graph_load_records(la3_ctx* ctx, const uint32_t* records, uint64_t nrecords,
                                  uint32_t nvertices, int weighted, la3_mode mode,
                                  la3_hashing hashing, la3_graph** out);
/* Graph<W>::free (src/matrix/graph.hpp:28-34). */
void la3_graph_free(la3_graph* g);
la3_status la3_graph_info_get(const la3_graph* g, la3_graph_info* out);
/* A graph from the two CSCs of the reference's Graph<W> at np=1 (the
 * arrays of CSCMatrix2D's tile, src/utils/csc.h:61-73, as la3_graph_csc
 * exports them; [0] = regular rows (`csc`), [1] = sink rows (`sink_csc`)):
 * for a caller whose graph is already built on the host. ncols = regular +
 * source vertices; per CSC colptrs[ncols + 1], colidxs[ncols] (hashed column
 * ids; unused for a column with no entries in that CSC) and per entry
 * global_idx (position among the regular or the sink rows; may be NULL),
 * idx (hashed row id) and weight (NULL when unweighted). `mode` is the load
 * the arrays came from (directedness; TC needs LA3_ACYCLIC). The arrays are
 * taken as the stored matrix: the graph is rebuilt from their entries with
 * the same hashing, and its CSCs must equal the given arrays entry for
 * entry, else LA3_EINVAL names the first difference (arrays that are not a
 * reference layout are rejected, not reinterpreted). */
typedef struct {
  uint32_t nvertices;          /* n: vertex ids [0, n] */
  int weighted;
  la3_mode mode;
  la3_hashing hashing;
  uint32_t ncols;
  const uint32_t* colptrs[2];
  const uint32_t* colidxs[2];
  const uint32_t* global_idx[2];
  const uint32_t* idx[2];
  const uint32_t* weight[2];
} la3_csc_arrays;
la3_status la3_graph_from_csc(la3_ctx* ctx, const la3_csc_arrays* arrays, la3_graph** out);
/* Layout export for parity checks: hid_of[nslots] maps a vertex-state slot
 * (order [regular | sink | source | isolated], each by increasing hashed id,
 * src/utils/locator.h:84-120) to its hashed id. */
la3_status la3_graph_slots(const la3_graph* g, uint32_t* hid_of);
/* Export of one of the tile's two DCSCs (which = 0 regular rows, 1 sink rows),
 * in the reference's CSC<W> layout (src/utils/csc.h:61-73): colptrs[ncols+1],
 * colidxs[ncols] (hashed column ids), per entry global_idx (position among the
 * regular or the sink rows), idx (hashed row id) and weight. */
la3_status la3_graph_csc(const la3_graph* g, int which, uint32_t* colptrs, uint32_t* colidxs,
                         uint32_t* global_idx, uint32_t* idx, uint32_t* weight);
/* Any array of the graph's device layout, as stored (the layout the kernels
 * read: DCSC with a regular/sink split per column, row-major companion in
 * message order, id maps by slot, long-row and PPR-segment lists). With
 * out == NULL only *count is set; cap = capacity of out in elements. For
 * parity checks of the ingest paths and for debugging; no counterpart in the
 * reference. */
typedef enum {
  LA3_ARR_COL_PTR = 0, LA3_ARR_COL_SPLIT, LA3_ARR_COL_ROWS, LA3_ARR_COL_W,
  LA3_ARR_ROW_PTR, LA3_ARR_ROW_SPLIT, LA3_ARR_ROW_COLS, LA3_ARR_ROW_W,
  LA3_ARR_ID_OF, LA3_ARR_SLOT_OF_ID, LA3_ARR_LONG_ROWS,
  LA3_ARR_PLONG, LA3_ARR_PSEG_FIRST, LA3_ARR_PSEG_ROW, LA3_ARR_PSEG_BEG
} la3_array;
la3_status la3_graph_array(const la3_graph* g, la3_array which, uint32_t* out, uint64_t cap,
                           uint64_t* count);
/* Device view of a graph's layout, for device code compiled outside the
 * library (the VertexProgram template of include/la3_vertex_program.hpp):
 * the arrays the library's kernels read, by vertex slot (order [regular |
 * sink | source | isolated], each by increasing hashed id,
 * src/utils/locator.h:84-120). Columns c = [regular | source]; column c's
 * sender slot is c (regular) or c + nsink (source). Rows are slots
 * [0, nreg + nsink). col_ptr/col_split/col_rows: the DCSC, entries of a
 * column by row slot, [col_ptr, col_split) the regular rows (the reference's
 * `csc`), [col_split, col_ptr + 1) the sink rows (`sink_csc`,
 * src/utils/csc.h:61-130). row_ptr/row_split/row_cols: the row-major
 * companion, a row's senders in the reference's message order (source
 * columns, then regular columns, each ascending,
 * src/vprogram/vertex_program_execute.hpp:464-483); row_split marks the
 * first regular sender. id_of / hid_of: original / hashed id of a slot;
 * slot_of_id: slot of an original id. long_rows: the rows with more than 64
 * entries. The pointers stay valid until la3_graph_free. No counterpart in
 * the reference (its CSC<W> arrays are host memory). */
typedef struct {
  int device;
  uint32_t nslots, nreg, nsink, nsrc, ncols, nrows;
  uint64_t nnz;
  uint32_t directed, weighted;
  const uint32_t* col_ptr;
  const uint32_t* col_split;
  const uint32_t* col_rows;
  const uint32_t* col_w;      /* NULL when unweighted */
  const uint32_t* row_ptr;
  const uint32_t* row_split;
  const uint32_t* row_cols;
  const uint32_t* row_w;      /* NULL when unweighted */
  const uint32_t* id_of;
  const uint32_t* hid_of;
  const uint32_t* slot_of_id;
  const uint32_t* long_rows;
  uint32_t nlong;
} la3_device_view;
la3_status la3_graph_device_view(const la3_graph* g, la3_device_view* out);
/* The context's HIP stream (a hipStream_t), on which every call on the
 * context runs; device code launched by the caller on it is ordered with
 * those calls. NULL for a host-only context. */
void* la3_ctx_stream(const la3_ctx* ctx);

/* hasher->hash / unhash (src/matrix/hashers.h:55-67). */
uint32_t la3_graph_hash(const la3_graph* g, uint32_t v);
uint32_t la3_graph_unhash(const la3_graph* g, uint32_t v);

/* ------------------------------------------------- single-query programs --- */
/* DegVertex (src/apps/graph_analytics/degree.h:25-35) with execute(1): the
 * in-degree of the stored graph per vertex (out-degree on a REVERSED load). */
la3_status la3_degree(la3_graph* g, uint32_t* degree);

/* PrVertex (pr.h:28-46, pr.cpp:11-55) on a DIRECTED load: out-degrees as the
 * reversed-graph DegVertex computes them, then execute(iters) (0 = until no
 * rank moves by more than 1e-5). Ranks are unnormalized (0.15 + 0.85*sum);
 * sources and isolated vertices keep 0.15; sinks are computed once after the
 * loop from the final regular ranks (vertex_program_execute.hpp:131-148).
 * FP64, deterministic summation. degree and iters_run may be NULL. */
la3_status la3_pagerank(la3_graph* g, uint32_t iters, double* rank, uint32_t* degree,
                        uint32_t* iters_run);

/* Every traversal takes the reference's max_iters
 * (VertexProgram::execute, src/vprogram/vertex_program_execute.hpp:11-52):
 * 0 runs until an iteration activates no regular vertex; 1 is
 * execute_single (one step for regular and sink rows from the initial
 * messages); m > 1 runs exactly m iterations (and iters_run reports m even if
 * the traversal settled earlier, as the reference's loop does). */

/* BfsVertex (bfs.h:27-49), execute(max_iters), on any load mode
 * (bfs.cpp:14 uses UNDIRECTED). hops: u8 level, 255 = unreached (levels
 * above max_iters stay 255). parent (may be NULL): the hashed id of the
 * in-neighbour whose message the reference combines last (last writer wins,
 * bfs.h:40), 0 for the root and unreached. */
la3_status la3_bfs(la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops, uint32_t* parent,
                   uint32_t* iters_run);

/* SpVertex (sssp.h:26-46) on a weighted DIRECTED load, execute(max_iters).
 * dist: u32, 2^31-1 for unreachable. With max_iters = m > 1 the regular
 * vertices hold the shortest distances over paths of at most m edges and the
 * sinks one edge more (the post-loop sink step); with max_iters = 1 every
 * row is one edge from the root. Bounded runs relax in the reference's
 * synchronous order (each iteration reads the distances the previous one
 * left); unbounded runs relax asynchronously (same distances, fewer device
 * rounds), and iters_run then counts device rounds, not the reference's
 * iterations. */
la3_status la3_sssp(la3_graph* g, uint32_t root, uint32_t max_iters, uint32_t* dist, uint32_t* iters_run);

/* The same two programs on another context's stream and scratch (same
 * device as the graph's): concurrent callers, such as the serving loop's
 * workers, use one context each. No counterpart in the reference. */
la3_status la3_bfs_ctx(la3_ctx* ctx, la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops,
                       uint32_t* parent, uint32_t* iters_run);
la3_status la3_sssp_ctx(la3_ctx* ctx, la3_graph* g, uint32_t root, uint32_t max_iters, uint32_t* dist,
                        uint32_t* iters_run);

/* la3_bfs_ctx that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * traversal's work buffers (levels, marks, two frontier queues: about
 * 21 bytes per vertex slot plus nnz/16 bytes) live in one fixed device array
 * of 64 MiB per device, held by the library from its load to the end of the
 * process; larger graphs return LA3_ENOTSUP. hops_dev (nslots bytes) and
 * parent_dev (nslots u32, may be NULL) are DEVICE pointers the caller
 * allocated; results are by original id, as la3_bfs_ctx returns them. Every
 * call on a device shares that device's array: calls that overlap on one
 * device (from any context) race. The array's device address is looked up
 * when a context is created on the device. */
la3_status la3_bfs_noalloc(la3_ctx* ctx, la3_graph* g, uint32_t root, uint32_t max_iters, uint8_t* hops_dev,
                           uint32_t* parent_dev, uint32_t* iters_run);

/* CcVertex (cc.h:22-38), execute(max_iters), run by the vertex-program
 * engine of include/la3_vertex_program.hpp with the reference's operations:
 * each row combines its messages in the reference's order with cc.h:31's
 * rule (an accumulator of 0 counts as empty), so labels and iters_run equal
 * the reference's on any load, including graphs where vertex 0 has edges
 * (its label 0 then survives only where it is the last message a row
 * combines). On an UNDIRECTED load with vertex 0 isolated, the labels are
 * the smallest original id of each connected component. */
la3_status la3_cc(la3_graph* g, uint32_t max_iters, uint32_t* label, uint32_t* iters_run);

/* Connected components by union-find (no counterpart in the reference):
 * label = smallest original id in the component, the fixed point CcVertex
 * reaches when vertex 0 is isolated; UNDIRECTED loads only. Faster than
 * la3_cc (no iterations); no iteration count. */
la3_status la3_components(la3_graph* g, uint32_t* label);

/* GnVertex + CtVertex (tc.h:47-99, tc.cpp:11-52) on an ACYCLIC load: total
 * triangle count (u64) and, if per_vertex != NULL, the count each vertex holds
 * in the reference (triangles in which it is the middle id). */
la3_status la3_tc(la3_graph* g, uint64_t* total, uint32_t* per_vertex);

/* Personalized PageRank (not in the reference; PR with teleport alpha at
 * `seed` only, initial rank = teleport, same loop and sink step as
 * la3_pagerank). FP64 single query. */
la3_status la3_ppr(la3_graph* g, uint32_t seed, uint32_t iters, double alpha, double* rank);

/* ------------------------------------------------------- batched queries --- */
/* k BFS roots in one traversal with k-bit frontiers (k <= 1024). hops is
 * k x nslots, query-major: hops[q * nslots + id]. Identical per query to
 * la3_bfs levels with the same max_iters. */
la3_status la3_bfs_batch(la3_graph* g, const uint32_t* roots, uint32_t k, uint32_t max_iters, uint8_t* hops,
                         uint32_t* iters_run);
/* k SSSP roots with k-wide u32 distances (k <= 1024); dist[q * nslots + id].
 * Identical per query to la3_sssp with the same max_iters; a bounded batch
 * (max_iters > 0) runs its queries one after another in the synchronous
 * order. */
la3_status la3_sssp_batch(la3_graph* g, const uint32_t* roots, uint32_t k, uint32_t max_iters, uint32_t* dist,
                          uint32_t* iters_run);
/* k PPR seeds (k <= 1024), `iters` iterations, FP32 with compensation: each
 * vertex and query carries its rank (f32) and the rank's compensation term
 * (f32); messages are f32; each row sums its messages with Kahan summation in
 * the reference's per-row message order. rank[q * nslots + id] is the f32
 * rank, rank_comp (may be NULL) its compensation (rank + rank_comp is the
 * carried value), residual (may be NULL) per query the FP64 L1 change of the
 * last iteration. */
la3_status la3_ppr_batch(la3_graph* g, const uint32_t* seeds, uint32_t k, uint32_t iters,
                         double alpha, float* rank, float* rank_comp, double* residual);

/* The batched programs on an explicit context: the batch runs on `ctx`'s
 * stream with `ctx`'s scratch instead of the graph's own context, so several
 * host threads (each with its own context on the same device) can run
 * batches on one resident graph concurrently. */
la3_status la3_bfs_batch_ctx(la3_ctx* ctx, la3_graph* g, const uint32_t* roots, uint32_t k,
                             uint32_t max_iters, uint8_t* hops, uint32_t* iters_run);
la3_status la3_sssp_batch_ctx(la3_ctx* ctx, la3_graph* g, const uint32_t* roots, uint32_t k,
                              uint32_t max_iters, uint32_t* dist, uint32_t* iters_run);
la3_status la3_ppr_batch_ctx(la3_ctx* ctx, la3_graph* g, const uint32_t* seeds, uint32_t k,
                             uint32_t iters, double alpha, float* rank, float* rank_comp,
                             double* residual);

typedef enum { LA3_Q_BFS = 0, LA3_Q_SSSP = 1, LA3_Q_PPR = 2 } la3_query_kind;

/* ------------------------------------------------- fused batch runs --- */
/* Several batches — BFS, SSSP and PPR, on one or several resident graphs of
 * the context's device — run as ONE traversal: each step is one fused launch
 * over the work of every batch (blocks split among the batches in
 * proportion to their work), one launch for the work that depends on it, and
 * a one-block control launch that advances each batch's state on the device
 * (BFS direction, SSSP threshold, buffer flips, termination). The steps are
 * captured once per context in a hipGraph and replayed; the host reads one
 * word per replay of 4 steps. Batches of more than 256 queries are split into
 * tasks of 256 (at most 64 tasks per run). Every query's result is identical
 * to the single-family batch call (la3_bfs_batch / la3_sssp_batch /
 * la3_ppr_batch) and so to the query alone. No counterpart in the reference;
 * it replaces several VertexProgram::execute calls (vertex_program_execute.hpp:
 * 11-52) issued back to back. */
typedef struct {
  la3_graph* graph;
  la3_query_kind kind;     /* LA3_Q_BFS, LA3_Q_SSSP or LA3_Q_PPR */
  uint32_t k;              /* queries, 1..1024 */
  const uint32_t* roots;   /* k root ids (PPR: seed ids) */
  uint32_t iters;          /* PPR iteration count (> 0); ignored otherwise */
  void* out;               /* k x nslots by id: u8 hops / u32 distances / f32 ranks;
                              NULL leaves the result in device memory */
  float* out_comp;         /* PPR: rank compensation terms, or NULL */
  double* residual;        /* PPR: FP64 L1 change of the last iteration per query, or NULL */
  uint32_t iters_run;      /* out: as the single-family batch call reports it */
} la3_batch;
la3_status la3_batch_run(la3_ctx* ctx, la3_batch* batches, uint32_t nbatches, double ppr_alpha);

/* Sessions: continuous batching at step granularity. A session is a fused
 * run that batches join and leave while it runs: la3_session_add sets a batch
 * up (it joins the next step), la3_session_step runs one step of every
 * running batch (one replay of a captured graph) and finishes the batches whose
 * traversals ended (results copied to their `out`, which must stay valid
 * until then), and la3_session_poll returns the ids of finished batches.
 * la3_batch_run is one session run to completion. A context holds at most one
 * session at a time (LA3_EAGAIN otherwise); a session has 64 task slots of
 * 256 queries (la3_session_add returns LA3_EAGAIN when the batch does not
 * fit). Each batch's results are those of la3_batch_run for it alone. A
 * task slot keeps its device buffers for the next task up to 256 MB (the
 * environment variable LA3_SESSION_KEEP_BYTES, read at session creation,
 * changes the limit) and frees larger ones when its task finishes
 * (la3_batch_run keeps them all until la3_ctx_release_scratch). A finished
 * batch whose `out` was NULL leaves its rows as the context's last result
 * only while its slot's buffers are kept: once they are freed,
 * la3_result_fetch and la3_result_topk return LA3_EINVAL. */
typedef struct la3_session la3_session;
la3_status la3_session_create(la3_ctx* ctx, double ppr_alpha, la3_session** out);
la3_status la3_session_add(la3_session* s, const la3_batch* batch, uint64_t* id);
/* *running = batches not finished after the step. */
la3_status la3_session_step(la3_session* s, uint32_t* running);
/* Pops up to cap finished batches: ids, iters_run and status (may be NULL). */
la3_status la3_session_poll(la3_session* s, uint64_t* ids, uint32_t* iters_run, int32_t* status,
                            uint32_t cap, uint32_t* n);
uint32_t la3_session_free_slots(const la3_session* s);
void la3_session_destroy(la3_session* s);

/* ------------------------------------------------------------- tuning --- */
/* SSSP relaxes in near/far rounds: only (vertex, query) pairs whose distance
 * is below a threshold T are relaxed; the others stay in the frontier, and
 * T rises to (smallest frontier distance + delta) when nothing below it
 * remains. delta = 0 relaxes everything every round (Bellman-Ford). The
 * distances are exact for every delta; delta only changes the work. */
/* Fused runs (la3_batch_run, sessions) split a batch of this graph into
 * tasks of LA3_TUNE_{BFS,SSSP,PPR}_TASK_K queries (64, 128 or 256: the k-wide
 * rows each task streams), and give their A / pull launches
 * LA3_TUNE_FUSED_BLOCKS blocks per CU (8, 16 or 32; a run uses the largest
 * value among its graphs). */
typedef enum {
  LA3_TUNE_SSSP_DELTA = 0,
  LA3_TUNE_BATCH_SSSP_DELTA = 1,
  LA3_TUNE_BFS_TASK_K = 2,
  LA3_TUNE_SSSP_TASK_K = 3,
  LA3_TUNE_PPR_TASK_K = 4,
  LA3_TUNE_FUSED_BLOCKS = 5
} la3_tunable;
la3_status la3_graph_set_tuning(la3_graph* g, la3_tunable what, uint32_t value);
la3_status la3_graph_get_tuning(const la3_graph* g, la3_tunable what, uint32_t* value);
/* Autotuner: times single-query and batched SSSP from `roots` (a batch of
 * min(nroots, 1024) and up to 4 single queries) for each candidate delta and
 * keeps the fastest of each on the graph. With at least 64 roots it then
 * times fused runs of min(nroots, 256) queries: per program (BFS, SSSP, PPR
 * of 5 iterations) each task width 64 / 128 / 256, then a BFS + SSSP run at
 * 8 / 16 / 32 blocks per CU, and keeps the fastest of each. */
typedef struct {
  uint32_t ncandidates;
  uint32_t candidates[8];
  double single_ms[8];   /* mean single-query time per candidate */
  double batch_ms[8];    /* batch time per candidate */
  uint32_t sssp_delta, batch_sssp_delta, batch_k;
  uint32_t task_k[3];        /* chosen task width: BFS, SSSP, PPR (0: not tuned) */
  double task_ms[3][3];      /* fused-run time per program at widths 64 / 128 / 256 */
  uint32_t fused_blocks;     /* chosen blocks per CU (0: not tuned) */
  double fused_ms[3];        /* BFS + SSSP fused run at 8 / 16 / 32 blocks per CU */
} la3_autotune_report;
la3_status la3_autotune(la3_graph* g, const uint32_t* roots, uint32_t nroots,
                        la3_autotune_report* report);

/* ------------------------------------------------------------ serving --- */
/* A request loop over resident graphs. Queries (BFS, SSSP, PPR) are
 * submitted one by one; `nworkers` host threads, each with its own context
 * (stream, scratch) on the device, repeatedly form a batch from the pending
 * queries — a pending query whose key (graph, program and, for PPR,
 * iteration count) has no batch running fixes the key (the oldest such
 * query, or by max_wait_s the one of least expected work), and up to
 * max_batch pending queries with that key join it in arrival order; queries
 * of a key with a batch running wait for that batch to finish and form the
 * key's next one — as long as the per-query state bytes of
 * the admitted batches stay within state_budget, run it as one batched
 * launch sequence, and record each query's completion. Queries that arrive
 * while batches are on the GPU are admitted as workers and budget free up;
 * batches of different workers overlap on the device. Each query's result
 * is the one the query gets alone (la3_*_batch with k = 1), whatever batch
 * it lands in. No counterpart in the reference. */
typedef struct la3_server la3_server;
typedef struct {
  uint32_t nworkers;      /* concurrent batches (host threads, one context each) */
  uint32_t max_batch;     /* queries per batch, 1..1024 */
  uint64_t state_budget;  /* bytes of per-query state admitted at once */
  int keep_results;       /* copy each result to host and keep it for la3_server_result */
  double ppr_alpha;       /* teleport probability of PPR queries */
  int fused;              /* 0: a batch is one (kind, graph, iters) group run by
                             la3_*_batch_ctx; 1: each worker keeps a session
                             (la3_session_*): between steps it admits
                             pending queries as new (kind, graph, iters) groups of up
                             to max_batch while task slots and budget allow, and
                             retires finished groups, so queries of every program and
                             graph share each fused launch */
  double max_wait_s;      /* batch order (fused = 0): 0 runs the group of the oldest
                             pending query first; > 0 runs the group of least
                             expected work first (the graph's entries times 1 for
                             BFS, 4 for SSSP, the iterations for PPR), unless the
                             oldest pending query has waited longer than this */
} la3_server_config;
typedef struct {
  uint64_t qid;
  uint32_t kind, graph, root, batch_size;
  int32_t status;                   /* la3_status of the batch that ran it */
  double submit_s, start_s, done_s; /* server clock, seconds */
} la3_query_info;
la3_status la3_server_create(int device, const la3_server_config* cfg, la3_server** out);
/* Registers a resident graph (loaded on any context of the same device). */
la3_status la3_server_add_graph(la3_server* s, la3_graph* g, uint32_t* graph_id);
/* Warms every worker before a stream: each runs, on its own context, one
 * query and a batch of max_batch queries of each program on each registered
 * graph, so that its scratch has reached its working size and no scratch
 * growth (whose hipFree waits for the whole device and stalls the other
 * workers) falls into the stream. Blocks until every worker is done; call
 * it with no queries pending. A no-op in fused mode. No counterpart in the
 * reference. */
la3_status la3_server_warm(la3_server* s);
/* Bytes of per-query state a query of `kind` on the graph is admitted with. */
uint64_t la3_server_query_bytes(const la3_server* s, la3_query_kind kind, uint32_t graph_id);
la3_status la3_server_submit(la3_server* s, la3_query_kind kind, uint32_t graph_id, uint32_t root,
                             uint32_t iters, uint64_t* qid);
/* Blocks until every submitted query has completed. */
la3_status la3_server_drain(la3_server* s);
/* Pops up to `cap` completion records (oldest first). */
la3_status la3_server_completed(la3_server* s, la3_query_info* out, uint32_t cap, uint32_t* n);
/* Copies a kept result (BFS u8 hops, SSSP u32 dist, PPR f32 rank; nslots
 * entries by original id) and releases it. */
la3_status la3_server_result(la3_server* s, uint64_t qid, void* dst, uint64_t bytes);
double la3_server_clock(const la3_server* s);
void la3_server_destroy(la3_server* s);

/* ------------------------------------------------------------- timing --- */
/* Device time (ms) of the kernels of the last la3_* program call on this
 * context, measured with HIP events on the context's stream, and the name of
 * its dominant kernel with its launch count and summed duration. */
typedef struct {
  double total_ms;
  double dominant_ms;
  uint64_t dominant_launches;
  uint64_t edges_scanned;   /* entries read by the traversal kernels */
  char dominant_kernel[64];
} la3_timing;
la3_status la3_last_timing(const la3_ctx* ctx, la3_timing* out);
/* Enables per-launch HIP-event timing (adds one event pair per launch). */
la3_status la3_set_profiling(la3_ctx* ctx, int enabled);

/* Cumulative per-kernel statistics since the last reset, gathered while
 * profiling is enabled: launches, summed HIP-event duration on the context's
 * stream, and the algorithmic bytes those launches moved (the compulsory
 * traffic of the vertex-program step they implement: index and weight
 * entries scanned, column pointers and messages of active columns, state of
 * activated rows; per-kernel formulas in the library's driver code). */
typedef struct {
  char name[48];
  uint64_t launches;
  double ms;
  double alg_bytes;
  uint64_t entries;     /* index entries read */
} la3_kernel_stat;
la3_status la3_kernel_stats(const la3_ctx* ctx, la3_kernel_stat* out, uint32_t cap,
                            uint32_t* count);
la3_status la3_kernel_stats_reset(la3_ctx* ctx);

/* Every program's main per-vertex output (hops, dist, rank, label, ...) may be
 * NULL: the result then stays in device memory, and la3_result_fetch copies
 * the last program's main output (as laid out for the host buffer) into
 * `dst`. The result lives in the context's scratch: it is gone (LA3_EINVAL)
 * after la3_ctx_release_scratch or when a later call frees or regrows that
 * buffer. Device-resident results let a caller time or chain programs without
 * the PCIe transfer. */
la3_status la3_result_fetch(la3_ctx* ctx, void* dst, uint64_t bytes);
/* The first `rows` rows (nslots elements of elem_bytes = 1, 2, 4 or 8 each, by
 * original id) of the last device-resident result, reordered on the device
 * to the reference's master-slot order (the vertex-state positions
 * [regular | sink | source | isolated], each class by hashed id,
 * src/utils/locator.h:84-120; the order VertexProgram::reduce and topk walk,
 * src/vprogram/vertex_program.hpp:558-680) and copied into `dst`. */
la3_status la3_result_fetch_slots(la3_ctx* ctx, const la3_graph* g, uint32_t elem_bytes, uint32_t rows, void* dst);

/* Top-k of each row of the last device-resident result (VertexProgram::topk,
 * vertex_program.hpp:607-680, one row per query for batched results): the
 * result is read as rows x n values of `dtype` by original id (the layout of
 * every program's output; rows = 1 for single queries); per row the k
 * largest (largest != 0) or smallest values, ordered by value and then by
 * smaller id (the reference's std::partial_sort leaves the order of equal
 * values unspecified). ids[r * k + j] and values[r * k + j] (values may be
 * NULL); rows with fewer than k values pad with id 0xffffffff. Selection
 * runs on the device (radix select and an ordered compaction); only k
 * entries per row cross PCIe. A fused run leaves its last task's rows as the
 * last result. */
typedef enum { LA3_DT_U8 = 0, LA3_DT_U32 = 1, LA3_DT_F32 = 2, LA3_DT_F64 = 3 } la3_dtype;
la3_status la3_result_topk(la3_ctx* ctx, la3_dtype dtype, uint32_t rows, uint64_t n, uint32_t k, int largest,
                           uint32_t* ids, void* values);

/* Per-row reductions of the last device-resident result, for the checksums
 * the reference's drivers print with VertexProgram::reduce
 * (vertex_program.hpp:558-604; bfs.cpp:28-31, sssp.cpp:28-36, cc.cpp:27-30,
 * pr.cpp:51-54) without copying the result to the host: the result is read
 * as rows x n values of `dtype` (by original id, as every program writes
 * it). LA3_REDUCE_SUM: the sum of each row, into uint64_t out[rows] for
 * LA3_DT_U8 / LA3_DT_U32 (exact) and double out[rows] for LA3_DT_F32 /
 * LA3_DT_F64 (FP64, in a fixed order: the same on every run, possibly not in
 * the last bits the reference's sequential sum). LA3_REDUCE_COUNT_NE: the
 * number of values other than `value` (converted to `dtype`), into
 * uint64_t out[rows] (e.g. reachable vertices: hops != 255, distances !=
 * 2^31 - 1). */
typedef enum { LA3_REDUCE_SUM = 0, LA3_REDUCE_COUNT_NE = 1 } la3_reduce_op;
la3_status la3_result_reduce(la3_ctx* ctx, la3_dtype dtype, uint32_t rows, uint64_t n, la3_reduce_op op,
                             double value, void* out);

/* Page-locked host memory (hipHostMalloc) for result buffers: device-to-host
 * copies into it run at full link speed. Any la3_* output pointer may point
 * into it. */
la3_status la3_host_alloc(uint64_t bytes, void** out);
void la3_host_free(void* p);

/* Device copy bandwidth: copies a `bytes` buffer `reps` times on the
 * context's stream with a plain vectorized copy kernel (la3_copy) and
 * returns (read + written bytes) / time in GB/s. A measured counterpart to
 * the 8 TB/s datasheet peak, and the known-size kernel that calibrates HBM
 * counter readings. */
la3_status la3_copy_bandwidth(la3_ctx* ctx, uint64_t bytes, uint32_t reps, double* gbs);

/* -------------------------------------------------------------- tools --- */
/* Seeded R-MAT generator (a, b, c, d = 1-a-b-c), 2^scale vertices with ids
 * 1..2^scale, edgefactor*2^scale samples, self-loops and duplicates removed,
 * records sorted by (src, dst); with weighted != 0 each record carries a u32
 * weight uniform in [1, 128]. Call with out == NULL to get the record count
 * in *nrecords, then with a buffer of nrecords*(2 or 3) u32. Deterministic
 * for a seed, independent of thread count. */
la3_status la3_rmat_generate(uint32_t scale, uint32_t edgefactor, double a, double b, double c,
                             uint64_t seed, int weighted, uint32_t* out, uint64_t* nrecords);

#ifdef __cplusplus
}
#endif

#endif /* LA3_H */
