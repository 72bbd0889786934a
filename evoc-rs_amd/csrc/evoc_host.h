// Internal host-side types of the MI355X kNN-graph library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "evoc_mi355x.h"

struct evoc_mi355x_ctx {
  int device = -1;
  int n_cu = 0;
  uint64_t budget = 0;   // HBM bytes this context may hold at once
  uint64_t in_use = 0;   // bytes currently held
  hipStream_t stream = nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  hipEvent_t ev_score0 = nullptr, ev_score1 = nullptr;  // around the exact search's scoring launch
  // event pairs around NN-descent's local-join launches (created with the
  // context; their times are collected after each host synchronisation)
  static constexpr int JOIN_EVENTS = 32;
  hipEvent_t ev_join[2 * JOIN_EVENTS] = {};
  // Arena mode (evoc_mi355x_knn_graph_noalloc): while `arena` is set, every
  // DevBuf is carved from [arena, arena + arena_size) by a stack (a released
  // block returns its bytes when it is the top one), budget is the arena size
  // and in_use its top; nothing is allocated from or freed to the pool.
  uint8_t* arena = nullptr;
  size_t arena_size = 0, arena_top = 0;
  // stream-ordered device memory of this context: frees do not synchronise the
  // device, so contexts on other streams keep running (serving slots)
  hipMemPool_t pool = nullptr;
  // serving slot: searches work in bounded query slabs (host syncs between
  // them are where other slots' kernels get the device), smaller while
  // `contention` (requests queued or running beside this one) is non-zero
  bool serving = false;
  const std::atomic<int>* contention = nullptr;
  // serving slot: pinned staging ring for host inputs and outputs, so their
  // copies are asynchronous DMA on this stream (pageable copies go through
  // the runtime's own staging and can wait behind other streams' work)
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;
  // Non-serving contexts keep the scratch blocks of their last call for the
  // next one: a released block goes to `cache` instead of back to the pool,
  // and an allocation of about its size takes it again.  Every use is on this
  // context's stream, so stream order makes the reuse safe.  (Releasing tens
  // of GiB to the pool and taking them back costs hundreds of milliseconds
  // per call.)  A call frees the cached blocks it did not reuse when it
  // returns (CallScope); evoc_mi355x_trim frees all of them.
  struct Cached {
    void* ptr;
    size_t bytes;
    uint64_t gen;  // call that released it
  };
  std::vector<Cached> cache;
  uint64_t cached = 0;  // bytes in `cache` (not in in_use)
  uint64_t gen = 0;     // calls begun
  void free_cached(size_t i);
  void free_stale();    // free cached blocks released before the current call
  std::mutex mu;
  std::string last_error;

  int fail(int code, const std::string& msg);
  int hip_fail(hipError_t e, const char* what);
  ~evoc_mi355x_ctx();
};
typedef evoc_mi355x_ctx Ctx;

// One top-level call on a context: blocks cached by earlier calls and not
// reused by this one are freed when it ends.  Construct it before the call's
// first DevBuf, so that it ends after their release.
struct CallScope {
  Ctx* c;
  explicit CallScope(Ctx* ctx) : c(ctx) { c->gen++; }
  ~CallScope() { c->free_stale(); }
};

// Device allocation accounted against the context's HBM budget, from the
// context's memory pool in the order of its stream.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  Ctx* owner = nullptr;
  bool in_arena = false;
  int alloc(Ctx* c, size_t nbytes, const char* what);
  void release();
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
};

// The allocation-free scratch (noalloc.hip): a file-scope device array of
// EVOC_MI355X_NOALLOC_SCRATCH_BYTES per GPU.  noalloc_scratch_init looks up
// this GPU's address of it once per device (with the device current);
// noalloc_scratch returns that address (nullptr if the lookup failed) and the
// lock that serialises its users on that device within this process.
int noalloc_scratch_init(int device);
void* noalloc_scratch(int device, std::mutex** lock);

// Host <-> device copies through the context's pinned staging ring (serving
// slots); both return when the copy is complete.  staged_download writes
// `rows` rows of `row_bytes` from a packed device array into host rows
// `dst_pitch` bytes apart.
hipError_t staged_upload(Ctx* c, void* dst, const void* src, size_t bytes);
hipError_t staged_download(Ctx* c, void* dst, size_t dst_pitch, const void* src, size_t row_bytes, size_t rows);
int enable_staging(Ctx* c, size_t bytes);

struct ExactStats {
  int64_t extra_rounds = 0;  // re-scoring rounds caused by buffer overflow
  int64_t slabs = 0;
  double pass2_ms = 0.0;     // device time of the scoring pass (all slabs)
  int64_t score_pairs = 0;   // row pairs scored by the scoring pass (all slabs, first round)
};

// Host -> device copy on the context stream that returns only when the copy is
// done, so the source may be a short-lived host object.
inline hipError_t upload_sync(void* dst, const void* src, size_t bytes, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e;
}

struct NndParams {
  int n_iters = 10;
  int max_candidates = 22;
  int n_trees = 4;
  int64_t leaf_size = 15;
  int64_t max_depth = 200;
  float delta = 0.001f;
  float delta_improv = 0.001f;
  bool use_delta_improv = true;
  int k = 15;                       // neighbours the caller asked for (stopping rule)
  uint64_t seed = 0;                // candidate priorities / random fill
  std::vector<uint64_t> tree_seeds; // one per tree
};

struct NndStats {
  int64_t iterations = 0;
  int64_t leaves = 0;
  int64_t dropped = -1;  // lost offers, -1 when not counted
  std::vector<uint32_t> changed;
  uint64_t join_bytes = 0;   // algorithmic bytes the local joins gathered
  double join_ms = 0.0;      // device time of the local-join launches
  int64_t join_launches = 0;
};

// Device random-projection forest (nnd_driver.cpp over nndescent.hip's level
// kernels).  ForestScratch holds the working set of a forest; forest_build
// builds one tree per seed, all of them level by level together, and leaves
// their row permutations and (start, length) leaves in that scratch.
struct ForestScratch {
  DevBuf perm, perm2, seg_of, side, segs, nsegs, leaves, fcnt, hyper, seeds, depth;
  int max_seg = 0;
  int n_trees = 0;
  int64_t hyper_ld = 0;
};
// A forest built by forest_build: tree t's rows in leaf order at positions
// [t n, t n + n) of perm; leaves are (flat start, length) into perm, in no
// particular order, tree start / n.
struct ForestSet {
  const int32_t* perm = nullptr;  // device, n_trees x n
  const int2* leaves = nullptr;   // device, n_leaves
  int n_leaves = 0;
  std::vector<int32_t> depth;     // levels split per tree (the root is level 0)
};
namespace evoc {
struct ForestArgs;
}
// scratch for up to n_trees trees of n rows built together (n_trees * n < 2^31)
int forest_alloc(Ctx* ctx, ForestScratch& fs, int kind, int64_t n, int dim, int64_t row_bytes, int leaf_size,
                 int n_trees);
evoc::ForestArgs forest_args(const ForestScratch& fs, int kind, int64_t n, int dim, int leaf_size,
                             const uint8_t* rows, int64_t row_bytes, const float* frows, int64_t frow_ld);
// builds one tree per seed (at most fs.n_trees), every level of all trees in one set of launches
int forest_build(Ctx* ctx, ForestScratch& fs, evoc::ForestArgs fa, int64_t max_depth,
                 const std::vector<uint64_t>& seeds, ForestSet* out);
// device bytes of forest_alloc
uint64_t forest_bytes(int kind, int64_t n, int dim, int64_t row_bytes, int leaf_size, int n_trees);

// NN-descent over all rows (nndescent.hip); out_keys: n x Kg exact-format keys.
int nndescent_search(Ctx* ctx, int kind, const uint8_t* rows, int64_t row_bytes, const uint32_t* pop,
                     const float* frows, int64_t frow_ld, int dim, int64_t n, int Kg,
                     const NndParams& P, uint64_t* out_keys, NndStats* st);

// Device bytes of the NN-descent working set (nnd_driver.cpp).
uint64_t nndescent_bytes(int kind, int64_t n, int Kg, int max_candidates, int64_t leaf_size, int64_t row_bytes,
                         int dim);

// Tile shape for a search over n rows of row_bytes (tuner.cpp): the
// autotuner's choice for the size bucket, else a size heuristic.
int choose_tile_shape(int kind, int64_t n, int64_t row_bytes);

// Exact search planning (exact.cpp): the query slab that fits `avail` bytes of
// scratch (called with the search's tile shape in force), and the scratch
// bytes of the whole search in one slab or in the smallest slab.
int64_t exact_slab(int kind, int64_t n, int K, int n_cu, uint64_t avail);
int64_t time_slab(int kind, int64_t n, int64_t row_bytes, double ms);
uint64_t exact_scratch_bytes(int kind, int64_t n, int64_t row_bytes, int K, int n_cu, bool minimum,
                             bool serving);

// The plan exact_search would use for these arguments with enough budget for
// the symmetric pass (a host-only diagnostic; evoc_mi355x_exact_plan).
struct ExactPlanInfo {
  int symmetric = 0, n_splits = 0, caps = 0, ccap = 0, stride = 0, n_sample = 0, segments = 0, tile_shape = -1;
  int bound_rank = 0;  // J of the plan (the K-th group maximum when J == K)
  int64_t work_items = 0;
  uint64_t scratch_bytes = 0;
};
int exact_plan_info(int kind, int64_t n, int64_t row_bytes, int K, int n_cu, bool serving, ExactPlanInfo* out);

// The plan of one per-split slab of all n queries (threshold sample, splits,
// buffer slots) under the calling thread's tile shape: what exact_search uses
// for a slab of the whole matrix outside the symmetric pass.
struct SlabPlan {
  int n_splits = 1, tiles_per_split = 1, caps = 0;
  bool pass1 = false;
  int stride = 1, n_sample = 0, ng = 0, p1_tiles_per_split = 1, p1_splits = 1;
};
SlabPlan exact_slab_plan(int kind, int64_t n, int K, int n_cu);

// Batched exact search of small requests (batch.cpp): many requests of one
// element kind per launch of each pass, from one fixed device workspace; the
// launch sequence is captured once per kind as a hipGraph and replayed for
// every batch.
struct BatchJob {
  evoc_knn_input in{};
  evoc_knn_options opts{};
  evoc_knn_output* out = nullptr;
  int status = EVOC_ERR_PENDING;
  bool overflow = false;  // a query's buffers overflowed: run the request alone instead
  std::string error;
  size_t bytes = 0;       // workspace bytes (batch_fit)
  int kind = -1;
};
struct BatchEngine;
// workspace_bytes: the device workspace (held by the engine's context from the
// first batch on); use_graph: replay captured launch sequences (else launch them)
BatchEngine* batch_engine_create(int device, size_t workspace_bytes, int max_requests, bool use_graph);
void batch_engine_destroy(BatchEngine* e);
// Whether a request can run batched (exact mode, f32 or int8 rows, n <= max_rows,
// k + margin <= 64) and fits the workspace; sets job->kind and job->bytes.
bool batch_fit(const BatchEngine* e, BatchJob* job, int64_t max_rows);
// Runs the jobs (one kind, workspace bytes within the workspace) as one batch
// on the stream of `run` (the engine's workspace, the caller's stream; one
// batch at a time per engine).  Each job's status is set; jobs with overflow
// set need the per-request path.
int batch_run(BatchEngine* e, Ctx* run, const std::vector<BatchJob*>& jobs, double* device_ms);
size_t batch_workspace_bytes(const BatchEngine* e);
int batch_max_requests(const BatchEngine* e);
int64_t batch_graph_replays(const BatchEngine* e);  // batches run from a captured graph

// Exact K best candidates per query over all rows (see knn_exact.hip).
// rows: n x row_bytes scored rows; pop: per-row popcounts (binary only).
// keys: n x K output keys, descending.
int exact_search(Ctx* ctx, int kind, const uint8_t* rows, int64_t row_bytes, const uint32_t* pop,
                 int64_t n, int K, bool seed_self, uint64_t* keys, ExactStats* stats);
