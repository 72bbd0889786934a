/*
 * evoc_mi355x — C ABI of the MI355X (gfx950) kNN-graph library.
 *
 * Drop-in for the reference's kNN-graph path (eugenehp/evoc-rs 0.0.1):
 *   - knn_graph(data, options, rng)               src/knn.rs:345-359
 *   - RlxBackend::knn_graph(data, opts, rng, sp)  src/rlx_backend/mod.rs:135-144
 * A Rust `Mi355xBackend` implements that trait method by calling
 * evoc_mi355x_knn_graph_seeded() after drawing the same NumPy-RNG values the
 * reference draws (src/knn.rs:416 and :124-129).
 *
 * Plain pointers and sizes only.  All functions are thread-safe per context
 * (calls on one context are serialised by an internal mutex).
 */
#ifndef EVOC_MI355X_H
#define EVOC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 6: adds evoc_mi355x_exact_plan, evoc_mi355x_make_forest, evoc_mi355x_knn_graph_noalloc,
 *    evoc_mi355x_server_set_batching / _batch_stats and evoc_knn_output.slabs; rayon_threads = 0 follows the host's
 *    thread count
 * 5: adds evoc_npy_read / evoc_npz_read / evoc_npy_write / evoc_npy_free (struct layouts of 4 unchanged) */
#define EVOC_MI355X_ABI_VERSION 6

/* EmbeddingData::{Float32, Int8, UInt8}, src/knn.rs:16-20.  UInt8 rows are
 * packed bits compared with bitwise Jaccard; d counts bytes. */
enum evoc_dtype { EVOC_DTYPE_F32 = 0, EVOC_DTYPE_I8 = 1, EVOC_DTYPE_U8 = 2 };

/* where a caller buffer lives */
enum evoc_location { EVOC_MEM_HOST = 0, EVOC_MEM_DEVICE = 1 };

/* search method */
enum evoc_mode {
  EVOC_MODE_EXACT = 0,     /* tiled exact all-pairs + top-k (default) */
  EVOC_MODE_NNDESCENT = 1  /* NN-descent local join on the device */
};

/* f32 scoring precision (ignored for int8 / binary, which are exact) */
enum evoc_precision {
  EVOC_PREC_FP32 = 0,         /* FP32 MFMA scoring */
  EVOC_PREC_BF16_RESCORE = 1  /* BF16 MFMA candidates, FP32 re-score of k + margin */
};

/* status codes (KnnError, src/knn.rs:33-41; BackendError, src/rlx_backend/mod.rs:191-202) */
enum evoc_status {
  EVOC_OK = 0,
  EVOC_ERR_EMPTY_DATA = 1,        /* KnnError::EmptyData */
  EVOC_ERR_INVALID_NEIGHBORS = 2, /* KnnError::InvalidNeighbors */
  EVOC_ERR_INVALID_ARGUMENT = 3,
  EVOC_ERR_UNSUPPORTED = 4,
  EVOC_ERR_OUT_OF_BUDGET = 5,     /* request does not fit the context's HBM budget */
  EVOC_ERR_HIP = 6,               /* HIP runtime error (message in last_error) */
  EVOC_ERR_CANCELLED = 7,
  EVOC_ERR_NOT_BITWISE = 8,       /* strict_precision requested: results are not bitwise the CPU path's */
  EVOC_ERR_NO_DEVICE = 9,
  EVOC_ERR_PENDING = 10           /* serving: the request has not finished yet */
};

/* KnnGraphOptions, src/knn.rs:43-76, plus the mi355x extensions.
 * Fields <= 0 marked "default" are derived exactly as src/knn.rs:402-414. */
typedef struct evoc_knn_options {
  int64_t n_neighbors;       /* default 30 */
  int64_t n_trees;           /* <= 0: clamp(rayon_threads, 4, 8) */
  int64_t leaf_size;         /* <= 0: max(10, k) */
  int64_t max_candidates;    /* <= 0: min(60, floor(1.5 k)) */
  int64_t max_rptree_depth;  /* default 200 */
  int64_t n_iters;           /* <= 0: max(5, round(log2 n)) */
  float delta;               /* default 0.001 */
  float delta_improv;        /* default 0.001 */
  int32_t use_sorted_updates;/* default 1 */
  int32_t deterministic;     /* default 0 */
  int64_t brute_force_max_n; /* default 0 */
  /* mi355x extensions */
  int32_t mode;              /* enum evoc_mode, default EXACT */
  int32_t precision;         /* enum evoc_precision, default BF16_RESCORE */
  int32_t margin;            /* candidates kept beyond k, default 16: the BF16 form re-scores them
                                in FP32; NN-descent keeps k + margin list entries per vertex */
  int32_t strict_precision;  /* RlxBackend strict_precision; nonzero -> EVOC_ERR_NOT_BITWISE */
  int32_t rayon_threads;     /* caller's rayon::current_num_threads() for the n_trees default; 0: the
                                count rayon's global pool has in this process (RAYON_NUM_THREADS when
                                set, else the CPUs the process may run on) */
  int32_t reserved0;
} evoc_knn_options;

typedef struct evoc_knn_input {
  int32_t dtype;      /* enum evoc_dtype */
  int32_t location;   /* enum evoc_location */
  const void* data;   /* row-major; row r starts at data + r * ld elements */
  int64_t n;          /* rows */
  int64_t d;          /* columns (u8: bytes per row) */
  int64_t ld;         /* leading dimension in elements, >= d */
} evoc_knn_input;

typedef struct evoc_knn_output {
  int32_t location;   /* enum evoc_location of indices/distances */
  int32_t reserved0;
  int32_t* indices;   /* n x k_out, row-major (Array2<i32>) */
  float* distances;   /* n x k_out, transformed as src/knn.rs:156-187 */
  int64_t capacity_k; /* columns the caller allocated; must be >= evoc_knn_effective_k() */
  int64_t k_out;      /* set: min(n_neighbors, n - 1).max(1) */
  /* statistics of the call */
  double seconds;          /* wall time of the call */
  double device_seconds;   /* device time between the first and last kernel */
  double score_seconds;    /* device time of the all-pairs scoring kernel (HIP events) */
  int64_t iterations;      /* NN-descent iterations run (0 for exact) */
  int64_t rescore_rounds;  /* exact: extra scoring rounds for overflowing buffers */
  int64_t offers_dropped;  /* NN-descent: offers that found a better offer in both of their
                              hashed buffer slots on arrival (a diagnostic that may vary between
                              calls; the lists do not depend on arrival order).  Counted only
                              when the environment sets EVOC_MI355X_COUNT_DROPPED=1 (counting
                              makes every offer wait for its atomics), else -1 */
  double queue_seconds;    /* serving: time from submit until a slot started the request */
  int64_t score_pairs;     /* exact: row pairs the scoring kernel scored (one per unordered pair
                              in the symmetric pass, one per query and candidate otherwise) */
  int64_t slabs;           /* exact: query slabs the search ran in (1 = the whole matrix at once) */
  int64_t join_bytes;      /* NN-descent: algorithmic bytes the local joins gathered (each work
                              item's candidate rows once, rows padded to 128 bytes) */
  double join_seconds;     /* NN-descent: device time of the local-join launches (HIP events) */
  int64_t join_launches;   /* NN-descent: local-join launches */
} evoc_knn_output;

typedef struct evoc_mi355x_ctx evoc_mi355x_ctx;
typedef struct evoc_numpy_rng evoc_numpy_rng;

/* ---- library / context ---- */
const char* evoc_mi355x_version(void);
int evoc_mi355x_device_count(int32_t* count);
/* hbm_budget_bytes = 0: 90% of the device's free memory at creation */
int evoc_mi355x_create(int32_t device, uint64_t hbm_budget_bytes, evoc_mi355x_ctx** out);
void evoc_mi355x_destroy(evoc_mi355x_ctx* ctx);
const char* evoc_mi355x_last_error(const evoc_mi355x_ctx* ctx);
/* A context keeps the device scratch of its last call for the next one (calls
 * of the same shape then allocate nothing); this returns it to the device.
 * No counterpart in the reference (its CPU path has no device memory). */
void evoc_mi355x_trim(evoc_mi355x_ctx* ctx);

/* ---- options and contract helpers ---- */
void evoc_knn_options_default(evoc_knn_options* opts);           /* KnnGraphOptions::default() */
int64_t evoc_knn_effective_k(int64_t n_neighbors, int64_t n);   /* src/knn.rs:402 */
/* number of randint(INT32_MIN+1, INT32_MAX-1) draws knn_graph takes from the
 * caller's RNG: 3 (randint3_for_tau) + 3 * n_trees, or 3 on the brute-force
 * branch (src/knn.rs:416-431) */
int64_t evoc_knn_rng_draws(int32_t dtype, int64_t n, const evoc_knn_options* opts);

/* ---- kNN graph ----
 * knn_graph (src/knn.rs:345-454) with the caller's NumPy RandomState: advances
 * `rng` exactly as the reference does, then builds the graph on the device. */
int evoc_mi355x_knn_graph(evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                          const evoc_knn_options* opts, evoc_numpy_rng* rng,
                          evoc_knn_output* out);

/* Same, with the draws already made by the caller (the Rust binding draws
 * from its own NumpyRandomState): tau_state = randint3_for_tau(), tree_states =
 * the n_tree_rows x 3 forest seeds of make_forest (may be NULL). */
int evoc_mi355x_knn_graph_seeded(evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                                 const evoc_knn_options* opts, const int64_t tau_state[3],
                                 const int64_t* tree_states, int64_t n_tree_rows,
                                 evoc_knn_output* out);

/* Allocation-free knn_graph.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It computes exactly what
 * evoc_mi355x_knn_graph_seeded computes, with every device buffer of the call
 * carved from one fixed scratch array of EVOC_MI355X_NOALLOC_SCRATCH_BYTES per
 * GPU (a file-scope device array the GPU holds until the process ends), and it
 * calls no hipMalloc, hipFree or allocating library routine.  The caller passes
 * inputs and outputs as for the seeded call (device or host; host ones are
 * staged through the scratch).  A call whose minimum working set
 * (evoc_mi355x_request_bytes) exceeds the scratch returns
 * EVOC_ERR_OUT_OF_BUDGET before any device work; the exact search runs in query
 * slabs that fit it.  Calls on one GPU share its scratch: within a process they
 * are serialised by a per-device lock (they would race otherwise), and the
 * scratch keeps the last call's contents (every buffer a call reads is written
 * or cleared by that call first). */
#define EVOC_MI355X_NOALLOC_SCRATCH_BYTES (256ull << 20)
int evoc_mi355x_knn_graph_noalloc(evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                                  const evoc_knn_options* opts, const int64_t tau_state[3],
                                  const int64_t* tree_states, int64_t n_tree_rows,
                                  evoc_knn_output* out);

/* Device bytes a call with these arguments holds at its peak (`full`) and the
 * least it can run in (`minimum`: exact search in small query slabs).  No
 * counterpart in the reference; used for admission by HBM budget. */
int evoc_mi355x_request_bytes(const evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                              const evoc_knn_options* opts, int32_t out_location,
                              uint64_t* full, uint64_t* minimum);

/* The plan an exact search (mode EXACT) of `in` and `opts` follows on a device
 * with n_cu compute units, given budget for the whole matrix in one slab.  Host
 * only: no device is touched and `in->data` is not read.  No counterpart in the
 * reference; a diagnostic of the scoring and selection passes. */
typedef struct evoc_exact_plan {
  int32_t symmetric;   /* 1: each unordered pair of row tiles is scored once */
  int32_t n_splits;    /* candidate splits of the scoring pass (one buffer each per query) */
  int32_t caps;        /* key slots per split and query */
  int32_t ccap;        /* symmetric pass: candidate-side key slots per row */
  int32_t stride;      /* threshold pass: every stride-th row is sampled (0: no threshold pass) */
  int32_t n_sample;    /* threshold pass: rows sampled */
  int32_t segments;    /* key lists the selection merges per query */
  int32_t tile_shape;  /* MFMA tile shape (-1: binary rows) */
  int32_t candidates;  /* K: keys kept per query (k, or k + margin for the BF16 form) */
  int32_t bound_rank;  /* threshold pass: scoring tests the bound_rank-th largest sample group maximum
                          (< candidates: an estimated bound; queries it misses are scored again) */
  int64_t work_items;  /* workgroup work items of the scoring pass */
  uint64_t scratch_bytes; /* device scratch of the search (outside the rows and results) */
} evoc_exact_plan;
/* serving != 0: the plan of a serving slot (per-split pass over time-bounded query slabs) */
int evoc_mi355x_exact_plan(const evoc_knn_input* in, const evoc_knn_options* opts, int32_t n_cu,
                           int32_t serving, evoc_exact_plan* out);

/* make_forest (src/knn.rs:115-154) on the device: the random-projection forest
 * NN-descent mode starts from, n_trees trees seeded by tree_states (n_trees x 3,
 * the make_forest draws), leaf size and depth from opts as knn_graph derives them
 * (leaves of at most 64 rows on the device), rows prepared as knn_graph prepares
 * them.  Returned in the reference's leaf-array layout: one row per leaf, each
 * tree's leaves in the order of their first row in the tree's permutation,
 * padded with -1 to *cols = max(leaf_size, widest leaf) (float.rs:148-151);
 * *rows = leaves in all trees.  tree_leaves[t] (may be NULL) = leaves of tree t,
 * depths[t] (may be NULL) = levels split in tree t (at most max_rptree_depth).
 * With leaf_array NULL only the sizes are returned; with capacity_rows < *rows,
 * EVOC_ERR_INVALID_ARGUMENT.  The split rule is the reference's (hyperplane
 * between two random members, int8 pivots normalised, binary margins by
 * popcount); the random choices come from the device's counter-based hash, not
 * the reference's tau RNG, so trees differ from the reference's on the same seeds. */
int evoc_mi355x_make_forest(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* opts,
                            const int64_t* tree_states, int64_t n_trees, int32_t* leaf_array,
                            int64_t capacity_rows, int64_t* rows, int64_t* cols, int64_t* tree_leaves,
                            int32_t* depths);

/* Tile-shape autotuner.  No counterpart in the reference.  Times the exact
 * search on the first min(n, max_rows) rows of `in` (max_rows <= 0: all) once
 * per MFMA tile shape (0: 128x128, 1: 128x256, 2: 256x256 in 8 waves,
 * 3: 256x256 in 4 waves) and makes the fastest the shape of every later exact
 * search of the same element kind, row count (power-of-two bucket) and row
 * width in this process.  ms[4] receives
 * the device milliseconds per shape (may be NULL).  Binary rows have a single
 * engine: *best_shape = -1.  Shapes change speed only, never results. */
int evoc_mi355x_autotune(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* opts,
                         int64_t max_rows, int32_t* best_shape, double ms[4]);

/* ---- serving: continuous batching of knn_graph requests ----
 * No counterpart in the reference (its knn_graph is one blocking call,
 * src/knn.rs:345-359).  A server owns `slots` device executors, each with its
 * own HIP stream, on one GPU.  Requests queue in arrival order and are admitted
 * whenever a slot is free and their device memory fits what the running
 * requests leave of the HBM budget; a request that does not fit may be passed
 * by smaller ones a bounded number of times.  Running requests overlap on the
 * device.  Each request computes exactly what evoc_mi355x_knn_graph_seeded
 * computes for it alone.  Input and output buffers stay owned by the caller and
 * must stay valid until the request is waited for. */
typedef struct evoc_mi355x_server evoc_mi355x_server;
/* slots <= 0: 4; hbm_budget_bytes = 0: 72% of free device memory (the slots'
 * memory pools keep up to a quarter of the budget cached between requests, so
 * 1.25 budgets must fit the device) */
int evoc_mi355x_server_create(int32_t device, int32_t slots, uint64_t hbm_budget_bytes,
                              evoc_mi355x_server** out);
/* waits for running requests; queued ones finish with EVOC_ERR_CANCELLED */
void evoc_mi355x_server_destroy(evoc_mi355x_server* srv);
/* copies *in, *opts, tau_state and tree_states; `out` must stay valid */
int evoc_mi355x_submit(evoc_mi355x_server* srv, const evoc_knn_input* in, const evoc_knn_options* opts,
                       const int64_t tau_state[3], const int64_t* tree_states, int64_t n_tree_rows,
                       evoc_knn_output* out, uint64_t* ticket);
/* EVOC_ERR_PENDING while the request runs or waits, else its final status
 * (the ticket stays valid until evoc_mi355x_wait returns it) */
int evoc_mi355x_poll(evoc_mi355x_server* srv, uint64_t ticket);
/* blocks up to timeout_seconds (< 0: forever); returns EVOC_ERR_PENDING on
 * timeout, else the request's status, copies its error message into err
 * (err_len bytes, may be NULL) and releases the ticket */
int evoc_mi355x_wait(evoc_mi355x_server* srv, uint64_t ticket, double timeout_seconds, char* err,
                     int64_t err_len);
/* the (full, minimum) device bytes the server's admission reserves for such a
 * request: evoc_mi355x_request_bytes on a serving slot plus the server's headroom */
int evoc_mi355x_server_request_bytes(const evoc_mi355x_server* srv, const evoc_knn_input* in,
                                     const evoc_knn_options* opts, int32_t out_location,
                                     uint64_t* full_bytes, uint64_t* min_bytes);
/* Continuous batching of small requests (no counterpart in the reference):
 * requests of at most max_rows rows in exact mode over f32 or int8 rows with
 * k + margin <= 64 are queued for a batch lane, which runs every queued one of
 * an element kind together, each pass one persistent launch over all of them
 * (replayed from a hipGraph captured once per kind), on its own stream beside
 * the slots.  Each result equals the request's result alone, bit for bit.
 * Applies to requests submitted afterwards; max_rows = 0 turns it off.  On by
 * default with max_rows = EVOC_MI355X_BATCH_ROWS or 16384 when the budget
 * leaves the lane a workspace (min(4 GiB, budget / 8), at least 256 MiB,
 * reserved from the budget at creation); EVOC_ERR_UNSUPPORTED without one. */
int evoc_mi355x_server_set_batching(evoc_mi355x_server* srv, int64_t max_rows);
/* batches run, requests they held, the largest batch, and batches run by replaying
 * a captured hipGraph (the others launched pass by pass) */
void evoc_mi355x_server_batch_stats(const evoc_mi355x_server* srv, int64_t* batches, int64_t* requests,
                                    int64_t* largest, int64_t* graph_replays);
/* requests queued / running and bytes granted to running requests */
void evoc_mi355x_server_stats(const evoc_mi355x_server* srv, int64_t* queued, int64_t* running,
                              uint64_t* granted_bytes, uint64_t* budget_bytes);

/* ---- the steps after knn_graph that consume its output ----
 * Fuzzy simplicial set: neighbor_graph_matrix_with_coo(n_neighbors,
 * knn_indices, knn_dists, symmetrize) (src/graph_construction.rs:175-226,
 * without its golden-value override): sigma / rho by smooth_knn_dist (:15-81),
 * membership strengths (:83-115) and, with symmetrize, the fuzzy union
 * A + A^T - A o A^T (src/csr_matmul.rs:62-147), returned as CSR with each row's
 * columns ascending and exact zeros dropped (:228-240).  Evoc::fit_predict calls
 * it on knn_graph's output with n_neighbors = neighbor_scale * n_neighbors
 * (src/clustering.rs:427-433).  -1 entries are skipped; a row must not list a
 * column twice; other indices outside [0, n) are EVOC_ERR_INVALID_ARGUMENT. */
typedef struct evoc_fuzzy_input {
  int32_t location;            /* enum evoc_location of knn_indices / knn_dists */
  int32_t symmetrize;          /* nonzero: the fuzzy union (the reference's default) */
  const int32_t* knn_indices;  /* n x k, row-major */
  const float* knn_dists;      /* n x k, as knn_graph returns them */
  int64_t n;
  int64_t k;
  float n_neighbors;           /* k of log2(k) in smooth_knn_dist */
  int32_t reserved0;
} evoc_fuzzy_input;

typedef struct evoc_fuzzy_output {
  int32_t location;   /* enum evoc_location of every array below */
  int32_t reserved0;
  int64_t* indptr;    /* n + 1 */
  int32_t* indices;   /* capacity */
  float* data;        /* capacity */
  int64_t capacity;   /* entries allocated in indices / data; 2 n k always suffices */
  float* sigmas;      /* n, may be NULL */
  float* rhos;        /* n, may be NULL */
  int64_t nnz;        /* set: entries of the result (also when capacity is too small) */
  double seconds;     /* wall time of the call */
  double device_seconds;
} evoc_fuzzy_output;

int evoc_mi355x_fuzzy_graph(evoc_mi355x_ctx* ctx, const evoc_fuzzy_input* in, evoc_fuzzy_output* out);

/* find_duplicates(knn_inds, knn_dists) (src/cluster_util.rs:430-449): the pairs
 * (i, j), i < j, that the kNN graph lists at a distance <= the largest distance
 * of its first column, as ascending keys (i << 32 | j) (the reference returns
 * them as a HashSet).  -1 entries are skipped.  *n_pairs is set also when
 * `capacity` is too small (EVOC_ERR_INVALID_ARGUMENT); n k always suffices. */
int evoc_mi355x_find_duplicates(evoc_mi355x_ctx* ctx, int32_t location, const int32_t* knn_indices,
                                const float* knn_dists, int64_t n, int64_t k, int32_t out_location,
                                uint64_t* pairs, int64_t capacity, int64_t* n_pairs);

/* ---- .npy / .npz arrays: the on-disk format of the reference's fixtures, goldens
 * and bench inputs (ndarray-npy ReadNpyExt / NpzReader::by_name, used by
 * tests/parity.rs:22-30, benches/evoc_bench.rs:4-15, src/knn.rs:460-472 and
 * src/graph_construction.rs:117-140).  Host-only; C order,
 * little-endian scalars; .npz members stored uncompressed (numpy.savez), zip64
 * included.  Readers allocate `data` (release with evoc_npy_free); errors return
 * EVOC_ERR_INVALID_ARGUMENT (missing file, malformed or truncated data, unknown
 * member) or EVOC_ERR_UNSUPPORTED (Fortran order, big-endian or other dtypes,
 * compressed members, format versions beyond 3.0) or EVOC_ERR_OUT_OF_BUDGET
 * (the host cannot allocate the array), with a message in err (err_len bytes,
 * may be NULL). ---- */
#define EVOC_NPY_MAX_DIMS 8
typedef enum {
  EVOC_NPY_BOOL = 0,
  EVOC_NPY_I8 = 1,
  EVOC_NPY_U8 = 2,
  EVOC_NPY_I32 = 3,
  EVOC_NPY_U32 = 4,
  EVOC_NPY_I64 = 5,
  EVOC_NPY_U64 = 6,
  EVOC_NPY_F32 = 7,
  EVOC_NPY_F64 = 8
} evoc_npy_dtype;

typedef struct {
  int32_t dtype; /* evoc_npy_dtype */
  int32_t ndim;  /* 0 .. EVOC_NPY_MAX_DIMS */
  int64_t shape[EVOC_NPY_MAX_DIMS];
  void* data;      /* C order */
  uint64_t nbytes; /* product(shape) x element size */
} evoc_npy_array;

int evoc_npy_read(const char* path, evoc_npy_array* out, char* err, int64_t err_len);
/* member: the array's name in the archive, with or without ".npy" */
int evoc_npz_read(const char* path, const char* member, evoc_npy_array* out, char* err, int64_t err_len);
/* writes format 1.0 (2.0 for headers beyond 64 KiB), header padded to 64 bytes */
int evoc_npy_write(const char* path, const evoc_npy_array* a, char* err, int64_t err_len);
void evoc_npy_free(evoc_npy_array* a);

/* ---- NumPy RandomState (src/numpy_rng.rs:12-226), for C / Python callers ---- */
evoc_numpy_rng* evoc_numpy_rng_new(uint32_t seed);        /* check_random_state(Some(seed)) */
void evoc_numpy_rng_free(evoc_numpy_rng* rng);
int64_t evoc_numpy_rng_randint(evoc_numpy_rng* rng, int64_t low, int64_t high);
double evoc_numpy_rng_normal(evoc_numpy_rng* rng);
double evoc_numpy_rng_random(evoc_numpy_rng* rng);
void evoc_numpy_rng_get_state(const evoc_numpy_rng* rng, uint32_t key[624], int32_t* pos,
                              int32_t* has_gauss, double* gauss);
void evoc_numpy_rng_set_state(evoc_numpy_rng* rng, const uint32_t key[624], int32_t pos,
                              int32_t has_gauss, double gauss);

#ifdef __cplusplus
}
#endif
#endif /* EVOC_MI355X_H */
