/* knng.h -- C-ABI of the MI355X kNN-graph builder (libknng.so).
 *
 * The reference engine (src/main.cpp of Aftab5550/SISAP-2026-task1) has no
 * library API: its boundary is the process plus the HDF5 file contract
 * (`main --task <name> --dataset <file.h5>`, src/main.cpp:114-141, and
 * results/<task>/apex_simd.h5, src/main.cpp:459-480).  knng_main reproduces
 * that boundary on top of the entry points below; each entry point names
 * the part of src/main.cpp it replaces.
 *
 * Conventions: every function returns 0 on success and a negative code on
 * error (knng_last_error() then describes it); no exception crosses the ABI.
 * Host buffers are caller-owned, device buffers context-owned unless said
 * otherwise.  One context drives one GPU on one stream and is not
 * thread-safe.  Calls are stream-ordered; functions that return host data
 * synchronise.
 */
#ifndef KNNG_H
#define KNNG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KNNG_OK 0
#define KNNG_EINVAL -1   /* bad argument */
#define KNNG_EHIP -2     /* HIP runtime error */
#define KNNG_ENOMEM -3   /* device or pinned allocation failed */
#define KNNG_ESTATE -4   /* call out of order (e.g. join before init) */
#define KNNG_ERANGE -5   /* a size exceeds a kernel limit */

/* Build parameters.  Defaults (knng_default_params) are the reference's
 * compile-time constants, src/main.cpp:19-29. */
typedef struct knng_params {
  int32_t dim;             /* DIM = 1024 (any multiple of 64, <= 1024) */
  int32_t k_eval;          /* K_EVAL = 15 */
  int32_t k_build;         /* K_BUILD = 50 (<= 64) */
  int32_t num_trees;       /* NUM_TREES = 16 (<= 64) */
  int32_t rerank_mult;     /* RERANK_MULTIPLIER = 8: sketch prefilter keeps k_build*8 */
  int32_t leaf_size;       /* LEAF_SIZE = 64 (2..2048; leaves above 64 rows keep each
                              member's 64 best mates) */
  int32_t max_iters;       /* MAX_ITERATIONS = 30 */
  float early_stop_rate;   /* EARLY_STOP_RATE = 0.04 */
  float rate_diff_tol;     /* RATE_DIFF_TOLERANCE = 0.01 */
  uint64_t seed;           /* pivot-hash seed (1337, the reference's tree seed base) */
  int32_t keep_fp32;       /* 1: keep the float32 rows in HBM for the FP32 rerank */
  int32_t rerank_width;    /* list entries re-scored in FP32 (k_eval..k_build); 0 = k_build */
  int32_t stop_rule;       /* KNNG_STOP_*: which rate the stop rule of src/main.cpp:432-439 watches */
  float insert_stop_rate;  /* KNNG_STOP_INSERTED: stop once fewer than this share of slots got a new id */
  int32_t incremental;     /* 1: a 2-hop pair is a candidate only if one of its edges is new
                              since the previous iteration (NN-Descent's new/old rule) */
  int32_t cells;           /* target cell count of the k-means partition that seeds every
                              tree (0: plain RP trees over all rows, as the reference;
                              -1: n / 1536) */
  int32_t kmeans_iters;    /* Lloyd iterations on the row sample */
  int32_t cell_partitions; /* independently seeded partitions; tree t uses t % partitions */
  int32_t cell_oversample; /* seeds = cells * cell_oversample before merging */
  int32_t cell_sample;     /* Lloyd and merging run on every cell_sample-th row */
  float cell_merge;        /* merge mutually nearest cells whose centroid cosine exceeds this */
  int32_t cell_full_iters; /* Lloyd iterations over all rows after merging */
  float cell_prune;        /* after merging, drop cells holding fewer sample rows than
                              cell_prune * the median cell (their rows go to the others) */
  int32_t reverse_width;   /* R (0..8, k_build <= 50): NN-Descent's reverse candidates in the
                              join.  Row v's reverse list holds the R best rows (score desc, id
                              asc) whose lists hold v as an entry the iteration treats as new;
                              row i also takes its reverse entries, its neighbours' reverse
                              entries and the lists of its reverse entries.  Single builds only
                              (knng_build_batch needs 0).  0 (default): pull-only 2-hop join as
                              src/main.cpp:335-419 */
} knng_params;

#define KNNG_STOP_REFERENCE 0  /* slots whose id changed, shifts included (src/main.cpp:421-427) */
#define KNNG_STOP_INSERTED 1   /* ids that newly entered a list (default) */
#define KNNG_STOP_NONE 2       /* always run max_iters */

/* Per-build statistics.  Times are HIP-event milliseconds on the context's
 * stream; bytes are algorithmic (counted from the work done, no cache credit). */
typedef struct knng_stats {
  int64_t n;
  int32_t dim;
  int32_t iterations;          /* join iterations run */
  int32_t forest_levels;
  int32_t pad0;
  int64_t num_leaves;
  int64_t updates[32];         /* per-iteration id changes (src/main.cpp:421-428 rule) */
  float rates[32];             /* updates / (n * k_build) */
  float ms_ingest;             /* K1 quantise + sketch (device time) */
  float ms_forest;
  float ms_init;
  float ms_join;               /* all iterations */
  float ms_join_iter[32];
  float ms_export;
  float ms_build;              /* forest + init + join, the reference's buildtime span */
  float ms_kmeans;             /* cell k-means (part of ms_forest) */
  int32_t cells;               /* cells over all partitions after merging (0: none) */
  double join_bytes[32];       /* algorithmic HBM bytes per join iteration */
  int64_t join_candidates[32]; /* unique 2-hop candidates scored by sketch */
  int64_t join_rescored[32];   /* candidates re-scored with int8 dots */
  int64_t join_inserted[32];   /* ids that newly entered a list */
  float insert_rates[32];      /* join_inserted / (n * k_build) */
  int32_t join_variant;        /* k_join shape the autotuner chose (0..2) */
  float ms_tune;               /* time spent timing k_join shapes (first build of a shape) */
  int64_t join_passed_on[32];  /* rows the join passed to its block-per-row kernel because they
                                  hold too many candidates for the wave-per-row phases (or, with
                                  KNNG_JOIN_FUSED=1, that the small-row kernel passed on); -1:
                                  the block-per-row kernel ran every row */
  int64_t join_leafmates[32];  /* 2-hop paths i -> n1 -> c the leaf-mate rule dropped (both edges
                                  share a leaf of one tree, so c shares it with i: scored exactly at
                                  init, it cannot enter i's list) */
  double init_tile_ops;        /* int8 ops (2 per multiply-add) of the init's leaf tiles: every
                                  member of each leaf above 64 rows against every member, m * m * 2 *
                                  dim per leaf (padding not counted) */
  double assign_ops;           /* int8 ops of the cell k-means assignments, rows * cells * 2 * dim
                                  per assignment (sample Lloyd steps, merge rounds, full rows) */
  int32_t leaf_variant;        /* leaf-tile shape the autotuner chose for leaves of <= 64 rows:
                                  0 block per leaf (LDS-staged quadrants), 1 wave per leaf */
  float ms_tune_leaf;          /* time spent timing leaf-tile shapes (first init of a shape) */
} knng_stats;

typedef struct knng_ctx knng_ctx;

/* Fill p with the reference's constants (src/main.cpp:19-29). */
int knng_default_params(knng_params* p);

/* Allocate a context for n rows on HIP device `device`.  Replaces the
 * buffer set-up of src/main.cpp:143-165.  n >= 2. */
int knng_create(knng_ctx** out, int32_t device, const knng_params* p, int64_t n);
int knng_destroy(knng_ctx* ctx);
const char* knng_last_error(const knng_ctx* ctx);

/* Ingest rows [row0, row0+nrows) from HOST memory: pinned double-buffered
 * H2D copy + fused quantise/sketch kernel.  Replaces the per-chunk loop body
 * of src/main.cpp:168-192. */
int knng_ingest_chunk(knng_ctx* ctx, const float* rows, int64_t row0, int64_t nrows);
/* Page-locked host buffer for zero-copy ingest: knng_ingest_chunk DMAs
 * straight from it instead of staging through its own pinned buffers.  Lets a
 * reader thread fill one buffer while the previous one is copied (the
 * reference reads and quantises serially, src/main.cpp:168-193). */
int knng_host_alloc(knng_ctx* ctx, void** ptr, int64_t bytes);
int knng_host_free(knng_ctx* ctx, void* ptr);
/* Same, rows already in device memory (not retained). */
int knng_ingest_device(knng_ctx* ctx, const float* dev_rows, int64_t row0, int64_t nrows);
/* Borrow a resident float32 copy of all n rows (device pointer) for the FP32
 * rerank instead of keeping an owned copy.  The caller keeps it alive. */
int knng_attach_fp32_device(knng_ctx* ctx, const float* dev_x);

/* Whole build: forest + init + join loop with the reference's stop rule.
 * Replaces src/main.cpp:203-442. */
int knng_build(knng_ctx* ctx);
/* Stages, for tests and tuning. */
int knng_build_forest(knng_ctx* ctx);                      /* src/main.cpp:205-232 */
int knng_build_init(knng_ctx* ctx);                        /* src/main.cpp:234-314 */
int knng_join_iteration(knng_ctx* ctx, int64_t* updates);  /* src/main.cpp:321-430, one pass */

/* Export the best k per row: knns = id+1 (int32), dists as float32, each row
 * ascending with the best neighbour last.  rerank=0: dists = float(int8 dot),
 * exactly src/main.cpp:449-457.  rerank=1: the first rerank_width list
 * entries are re-scored with FP32 dot products of the original rows and the
 * best k of those are emitted with their FP32 scores.  Host buffers n*k. */
int knng_export(knng_ctx* ctx, int32_t* knns, float* dists, int32_t k, int32_t rerank);
/* Same into device buffers (no host sync). */
int knng_export_device(knng_ctx* ctx, int32_t* dev_knns, float* dev_dists, int32_t k,
                       int32_t rerank);

int knng_get_stats(knng_ctx* ctx, knng_stats* st);
/* Block until all work queued on the context's stream is done. */
int knng_synchronize(knng_ctx* ctx);

/* Introspection for parity tests (host buffers, synchronising). */
int knng_copy_quantized(knng_ctx* ctx, int8_t* q, uint64_t* sketch);   /* n*dim, n*dim/64 */
/* perm: num_trees*n leaf-ordered point ids; leaves: (tree,start,size) triples
 * sorted by (tree,start); *nleaf receives the count (<= cap). */
int knng_copy_forest(knng_ctx* ctx, int32_t* perm, int64_t* leaves, int64_t cap, int64_t* nleaf);
int knng_copy_graph(knng_ctx* ctx, int32_t* ids, int32_t* scores);      /* n*k_build each */
/* Cells of partition `part` after knng_build_forest: *count cells, assign[n]
 * (cell of every row), z[count*dim] int8 centroid directions, w[count] = 1/|z|.
 * Any output may be NULL (count alone sizes the buffers); *count = 0 when the
 * build used no cells. */
int knng_copy_cells(knng_ctx* ctx, int32_t part, int32_t* count, int32_t* assign, int8_t* z,
                    float* w);
int knng_load_graph(knng_ctx* ctx, const int32_t* ids, const int32_t* scores);

/* Stand-alone primitives on device buffers (synchronous, device 0 stream):
 * K1 quantise + sketch of n rows of dimension d (d % 64 == 0), src/main.cpp:179-192. */
int knng_quantize_device(const float* dev_x, int64_t n, int32_t d, int8_t* dev_q,
                         uint64_t* dev_sketch);
/* Gathered pair scores: dot[j] = <q[a[j]], q[b[j]]> (int32, src/main.cpp:403-409) and
 * sim[j] = d - popcount(sk[a[j]] ^ sk[b[j]]) (src/main.cpp:354-370). */
int knng_score_pairs_device(const int8_t* dev_q, const uint64_t* dev_sketch, int32_t d,
                            const int32_t* dev_a, const int32_t* dev_b, int64_t npairs,
                            int32_t* dev_dot, int32_t* dev_sim);

/* One join iteration (the full join: own list, 2-hop candidates, sketch
 * prefilter keeping k_build * rerank_mult, int8 re-scoring, rank merge) over
 * caller-owned DEVICE buffers, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving or refinement loops, harnesses that
 * forbid hipMalloc.  It computes exactly the lists knng_join_iteration
 * computes from the same graph without the new/old rule.
 *   q [n][d] int8, sketch [n][d/64]: the quantised rows (knng_quantize_device);
 *   ids_in / scores_in [n][k_build]: the current lists, (score desc, id asc);
 *   ids_out / scores_out [n][k_build]: the next lists (must not alias the inputs);
 *   order: optional processing order of the rows (n int32, a permutation; NULL
 *          = index order; the result does not depend on it);
 *   counts [6] uint64 (device): changed slots (src/main.cpp:421-427 rule), unique
 *          candidates, re-scored candidates, ids new to their list, candidates
 *          whose sketches were gathered, 0 (no leaf-mate rule without a forest);
 *   stream: a hipStream_t (NULL: the null stream).
 * Stream-ordered, no host synchronisation.  Its only scratch is a static
 * device array of 51,200 bytes per GPU (join counters and row queues), held
 * until the process ends and shared by every call on that GPU: calls that
 * overlap on one GPU race.  Limits: d a multiple of 64 up to 1024, k_build
 * 1..64, rerank_mult >= 1, 2 <= n < 2^31 - 1 (KNNG_ERANGE otherwise). */
int knng_join_iteration_noalloc(const int8_t* dev_q, const uint64_t* dev_sketch, int32_t d, int64_t n,
                                int32_t k_build, int32_t rerank_mult, const int32_t* dev_ids_in,
                                const int32_t* dev_scores_in, int32_t* dev_ids_out,
                                int32_t* dev_scores_out, const int32_t* dev_order,
                                uint64_t* dev_counts, void* stream);

/* ------------------------------------------------------------------------
 * Batched builds of many independent graphs of uneven size (no reference
 * counterpart: the reference builds one graph of DIM = 1024).  Each job is a
 * whole build (ingest, forest, init, join, export); `workers` host threads
 * each drive their own context and HIP stream and take jobs largest-first
 * from a shared counter, so small builds overlap on the device.
 * ------------------------------------------------------------------------ */
typedef struct knng_job {
  const float* rows;       /* n x params.dim float32 rows */
  int32_t rows_on_device;  /* 1: rows is a device pointer on `device` */
  int32_t rerank;          /* export with the FP32 rerank */
  int64_t n;
  knng_params params;      /* per-job dim, k_build, k_eval, trees, ... */
  int32_t* knns;           /* host out, n x k_eval (1-based, ascending, best last) */
  float* dists;            /* host out, n x k_eval */
  int32_t status;          /* out: 0 ok, else a KNNG_E* code */
  int32_t iterations;      /* out: join iterations run */
  double build_ms;         /* out: wall time of this job's build + export */
} knng_job;

int knng_build_batch(knng_job* jobs, int32_t njobs, int32_t device, int32_t workers);

/* ------------------------------------------------------------------------
 * Serving: kNN queries and inserts against a built, device-resident graph.
 * The reference has no query path (it writes querytime = 0,
 * src/main.cpp:479); these entry points are new.  Search = descent of
 * `entry_trees` RP trees to their leaves for entry points, then beam search
 * over the K-list graph with int8 scores, then (rerank=1) FP32 re-scoring of
 * the final beam.  Result ids are 0-based, best first.
 * ------------------------------------------------------------------------ */
typedef struct knng_server knng_server;

typedef struct knng_server_params {
  int32_t beam;            /* beam width (<= 64) */
  int32_t k;               /* neighbours returned per query (<= beam) */
  int32_t entry_trees;     /* RP trees descended for entry points (<= num_trees) */
  int32_t rerank;          /* 1: FP32 re-score of the final beam */
  int32_t max_steps;       /* beam-search steps (4 expansions each) */
  int32_t max_batch;       /* queries per launch (per slot) */
  int32_t max_slots;       /* concurrent launches, one HIP stream each */
  int32_t pad0;
  int64_t insert_capacity; /* rows that later inserts may add */
  int64_t reserve_bytes;   /* HBM left free when admitting a new slot */
  int32_t express_rows;    /* query requests of at most this many rows take the express
                              lane (0: no express lane) */
  int32_t express_slots;   /* express slots, created with the server */
  int32_t express_cus;     /* CUs reserved for the express slots' streams through
                              hipExtStreamCreateWithCUMask (0, the default: none; where the
                              runtime refuses CU masks the server runs without them) */
  int32_t strict_order;    /* 0 (default): inserts run beside searches on their own thread and
                              stream; 1: an insert waits for the queries submitted before it and
                              a query for the inserts submitted before it */
  int32_t bulk_yield;      /* while express requests wait or run, start no bulk launch once this
                              many bulk slots are busy (0, the default: no limit); launches in
                              flight are never held back, so bulk work always progresses */
} knng_server_params;

int knng_server_default_params(knng_server_params* sp);

/* Copy the built graph, int8 rows and float32 rows of `ctx` into a
 * device-resident index with room for insert_capacity more rows, and start
 * the scheduler thread.  ctx must stay alive (its forest is used for entry
 * points). */
int knng_server_create(knng_server** out, knng_ctx* ctx, const knng_server_params* sp);
int knng_server_destroy(knng_server* srv);
const char* knng_server_last_error(const knng_server* srv);

/* Queue a request (rows are copied): kind 0 = query batch of nrows vectors,
 * kind 1 = insert of nrows new vectors (ids n, n+1, ... in submission order).
 * Returns at once with a ticket.  Queries are split into max_batch pieces and
 * admitted as slots and HBM free up; pieces of different requests run
 * concurrently on different streams.  Query requests of at most express_rows
 * rows bypass the bulk queue: they run on the express slots, whose streams
 * (with express_cus > 0) own a reserved set of CUs that bulk launches never
 * occupy, so their latency does not grow with the bulk backlog.  Inserts run
 * on their own thread and stream beside the searches (pending ones together):
 * a query sees every insert finished before it was dispatched, and may see
 * one that overlaps it.  With strict_order = 1 an insert instead waits until
 * the queries submitted before it have finished, and queries submitted after
 * it (either lane) wait for it. */
int knng_server_submit(knng_server* srv, int32_t kind, const float* rows, int32_t nrows,
                       int64_t* ticket);
/* state: 0 queued, 1 running, 2 done. */
int knng_server_poll(knng_server* srv, int64_t ticket, int32_t* state);
/* Block until done; copy results (queries: nrows*k ids and scores; inserts:
 * the new ids in ids[0..nrows)) and the submit-to-done latency; release the
 * ticket.  ids/scores may be NULL. */
int knng_server_wait(knng_server* srv, int64_t ticket, int32_t* ids, float* scores,
                     double* latency_ms);
/* knng_server_wait plus *index_rows: for a query, the index size when it was
 * dispatched (it saw at least those rows); for an insert, the size just after
 * it.  May be NULL. */
int knng_server_wait_info(knng_server* srv, int64_t ticket, int32_t* ids, float* scores,
                          double* latency_ms, int64_t* index_rows);
/* Lanes in force: *cu_masked = 1 when the express streams own reserved CUs,
 * and the bulk / express slot counts.  Any output may be NULL. */
int knng_server_info(const knng_server* srv, int32_t* cu_masked, int32_t* bulk_slots,
                     int32_t* express_slots);
/* Why the runtime refused the express CU masks ("" when they are in force or
 * were not asked for). */
const char* knng_server_mask_error(const knng_server* srv);
/* Rows currently in the index. */
int64_t knng_server_size(const knng_server* srv);

/* One synchronous batch search on the server's index (no scheduler): host
 * queries in, host results out.  Same kernel as the server's slots. */
int knng_server_search(knng_server* srv, const float* queries, int32_t nq, int32_t* ids,
                       float* scores);

#ifdef __cplusplus
}
#endif
#endif /* KNNG_H */
