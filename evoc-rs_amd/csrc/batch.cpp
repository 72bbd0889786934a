// Batched exact search for small requests: the serving loop's continuous
// batching of knn_graph calls below a size limit.
//
// A batch runs on the stream of the context the caller passes (a serving
// slot's), one batch at a time per engine.  A batch is a set of queued
// requests of one element kind (BF16 + FP32
// re-score, FP32 or int8), each of at most max_rows rows and k + margin <= 64.
// Their rows, thresholds, key buffers and outputs are laid out in one device
// workspace; a table of per-request arguments (BatchReq) and the work-item
// lists of every pass are copied in front of it.  One persistent launch per
// pass then covers all requests (knn_exact.hip, prep.hip), each item doing
// exactly the work of the single-request kernel with its request's arguments,
// so a request's graph equals the one it gets alone, bit for bit.  The grids
// and every kernel argument are fixed (the table and lists sit at fixed
// addresses), so the launch sequence of a kind is captured once into a hipGraph
// and replayed for each batch.  A request whose candidate buffers overflow (the
// single-request search would then re-score those query blocks) is handed back
// to run alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "evoc_host.h"
#include "evoc_kernels.h"
#include "evoc_mi355x.h"

using namespace evoc;

namespace {

constexpr int MAX_ITEMS = 1 << 20;  // work items of one batch (all lists)
constexpr int KINDS = 3;            // BF16, FP32, int8

size_t r256(size_t b) { return (b + 255) & ~(size_t)255; }
int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
int64_t round_up(int64_t v, int64_t m) { return cdiv(v, m) * m; }

// Everything a batched request needs to know about itself.
struct Shape {
  int kind, dtype;
  bool bf16, host_in, dev_out, seed_self;
  int64_t n, d, k, K, row_bytes, f32_ld, elem;
  SlabPlan p;
};

}  // namespace

struct BatchEngine {
  int device = 0;
  int n_cu = 0;
  Ctx* ctx = nullptr;
  DevBuf ws;
  size_t ws_bytes = 0, head_bytes = 0;
  int max_reqs = 0;
  bool use_graph = true;
  void* pinned = nullptr;  // header + table + items, staged for the upload
  hipGraph_t graph_def[KINDS] = {nullptr, nullptr, nullptr};  // kept to find the kernel nodes
  hipGraphExec_t graph[KINDS] = {nullptr, nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::atomic<int64_t> replays{0};  // batches run by replaying a captured graph (read by other threads)
};

static int kind_index(int kind) { return kind == EVOC_ELEM_BF16 ? 0 : kind == EVOC_ELEM_F32 ? 1 : 2; }

// (the tile shape of batches: 128 x 128, whatever the thread's shape is)
struct Shape0 {
  Shape0() { set_tile_shape(0); }
  ~Shape0() { set_tile_shape(-1); }
};

static bool shape_of(const BatchEngine* e, const evoc_knn_input* in, const evoc_knn_options* o, int32_t out_location,
                     int64_t capacity_k, Shape* s) {
  if (in->dtype != EVOC_DTYPE_F32 && in->dtype != EVOC_DTYPE_I8) return false;
  if (o->mode != EVOC_MODE_EXACT || o->strict_precision || o->n_neighbors <= 0 || in->n < 2 || in->d <= 0 ||
      in->ld < in->d || !in->data)
    return false;
  s->dtype = in->dtype;
  s->n = in->n;
  s->d = in->d;
  s->k = evoc_knn_effective_k(o->n_neighbors, in->n);
  s->bf16 = in->dtype == EVOC_DTYPE_F32 && o->precision == EVOC_PREC_BF16_RESCORE;
  s->K = s->bf16 ? std::min<int64_t>(s->k + std::max(0, o->margin), s->n) : s->k;
  if (s->K > 64) return false;
  // brute_force_knn (src/knn.rs:418-422) skips self; the NN-descent layout seeds it
  s->seed_self = !(s->n <= o->brute_force_max_n && in->dtype != EVOC_DTYPE_F32);
  s->kind = in->dtype == EVOC_DTYPE_I8 ? EVOC_ELEM_I8 : s->bf16 ? EVOC_ELEM_BF16 : EVOC_ELEM_F32;
  s->elem = in->dtype == EVOC_DTYPE_F32 ? 4 : 1;
  s->row_bytes = round_up(s->d * (s->bf16 ? 2 : s->elem), 128);
  s->f32_ld = round_up(s->d, 32);
  s->host_in = in->location != EVOC_MEM_DEVICE;
  s->dev_out = out_location == EVOC_MEM_DEVICE && capacity_k == s->k;
  Shape0 shape0;
  // the batch kernels run 128 x 128 tiles: the plan must be made for them (a
  // process-wide tile pin overrides the thread's shape)
  if (tile_rows_c(s->kind) != 128 || tile_rows_q(s->kind) != 128 || groups_per_tile(s->kind) != 8) return false;
  s->p = exact_slab_plan(s->kind, s->n, (int)s->K, e->n_cu);
  return true;
}

// workspace bytes of one request, in the order batch_run lays them out
static size_t request_bytes(const Shape& s, const evoc_knn_input* in) {
  const int64_t n = s.n;
  size_t b = 0;
  if (s.host_in) b += r256((size_t)((n - 1) * in->ld + s.d) * (size_t)s.elem);
  if (s.dtype == EVOC_DTYPE_F32) b += r256((size_t)(n * s.f32_ld * 4)) + r256((size_t)n * 4) + 256;
  if (s.dtype != EVOC_DTYPE_F32 || s.bf16) b += r256((size_t)(n * s.row_bytes));
  if (s.p.pass1) b += r256((size_t)n * s.p.ng * 4) + r256((size_t)n * 4);
  b += r256((size_t)s.p.n_splits * n * s.p.caps * 8) + r256((size_t)s.p.n_splits * n * 4);
  b += r256((size_t)(n * s.K * 8)) + r256((size_t)n * 4) + r256((size_t)n * 8);
  if (s.bf16) b += r256((size_t)(n * s.k * 8));
  if (!s.dev_out) b += 2 * r256((size_t)(n * s.k * 4));
  return b;
}

static int64_t request_items(const Shape& s) {
  const int64_t nqb = cdiv(s.n, 128);
  return cdiv(s.n, 64) + (s.p.pass1 ? nqb * s.p.p1_splits : 0) + nqb * s.p.n_splits + cdiv(s.n, 4);
}

BatchEngine* batch_engine_create(int device, size_t workspace_bytes, int max_requests, bool use_graph) {
  auto* e = new (std::nothrow) BatchEngine;
  if (!e) return nullptr;
  e->device = device;
  e->max_reqs = std::max(1, max_requests);
  e->use_graph = use_graph;
  e->head_bytes = r256(sizeof(BatchHeader)) + r256(sizeof(BatchReq) * (size_t)e->max_reqs) +
                  r256(sizeof(int4) * (size_t)MAX_ITEMS) + r256(sizeof(int32_t) * (size_t)e->max_reqs);
  e->ws_bytes = workspace_bytes;
  if (evoc_mi355x_create(device, workspace_bytes + (64u << 20), &e->ctx) != EVOC_OK) {
    delete e;
    return nullptr;
  }
  e->n_cu = e->ctx->n_cu;
  if (hipHostMalloc(&e->pinned, e->head_bytes, hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&e->e0) != hipSuccess || hipEventCreate(&e->e1) != hipSuccess) {
    (void)hipGetLastError();
    batch_engine_destroy(e);
    return nullptr;
  }
  return e;
}

void batch_engine_destroy(BatchEngine* e) {
  if (!e) return;
  if (e->ctx) (void)hipSetDevice(e->device);
  for (auto& g : e->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : e->graph_def)
    if (g) (void)hipGraphDestroy(g);
  if (e->e0) (void)hipEventDestroy(e->e0);
  if (e->e1) (void)hipEventDestroy(e->e1);
  if (e->pinned) (void)hipHostFree(e->pinned);
  if (e->ctx) {
    e->ws.release();
    evoc_mi355x_destroy(e->ctx);
  }
  delete e;
}

size_t batch_workspace_bytes(const BatchEngine* e) { return e->ws_bytes; }
int batch_max_requests(const BatchEngine* e) { return e->max_reqs; }
int64_t batch_graph_replays(const BatchEngine* e) { return e->replays.load(); }

bool batch_fit(const BatchEngine* e, BatchJob* job, int64_t max_rows) {
  if (job->in.n > max_rows) return false;
  Shape s;
  if (!shape_of(e, &job->in, &job->opts, job->out->location, job->out->capacity_k, &s)) return false;
  job->kind = s.kind;
  job->bytes = request_bytes(s, &job->in);
  return job->bytes + e->head_bytes <= e->ws_bytes && request_items(s) <= MAX_ITEMS;
}

int batch_run(BatchEngine* e, Ctx* run, const std::vector<BatchJob*>& jobs, double* device_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  Ctx* ctx = e->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  std::lock_guard<std::mutex> run_lock(run->mu);
  hipStream_t st = run->stream;
  auto fail_all = [&](int code, const std::string& msg) {
    for (BatchJob* j : jobs) {
      j->status = code;
      j->error = msg;
    }
    return code;
  };
  if (jobs.empty()) return EVOC_OK;
  if ((int)jobs.size() > e->max_reqs) return fail_all(EVOC_ERR_INVALID_ARGUMENT, "batch larger than the table");
  if (hipSetDevice(e->device) != hipSuccess) return fail_all(EVOC_ERR_HIP, "hipSetDevice failed");
  if (!e->ws.ptr) {
    // allocated in the engine context's stream order; that stream is drained
    // before the workspace is used on the caller's stream
    const int rc = e->ws.alloc(ctx, e->ws_bytes, "batch workspace");
    if (rc) return fail_all(rc, ctx->last_error);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail_all(EVOC_ERR_HIP, "batch workspace");
  }
  const int kind = jobs[0]->kind;
  // ---- layout: header | table | items | overflow flags | per-request areas ----
  uint8_t* dev = (uint8_t*)e->ws.ptr;
  uint8_t* host = (uint8_t*)e->pinned;
  BatchHeader* hh = (BatchHeader*)host;
  BatchReq* ht = (BatchReq*)(host + r256(sizeof(BatchHeader)));
  int4* hi = (int4*)(host + r256(sizeof(BatchHeader)) + r256(sizeof(BatchReq) * (size_t)e->max_reqs));
  BatchHeader* d_hdr = (BatchHeader*)dev;
  BatchReq* d_tab = (BatchReq*)(dev + r256(sizeof(BatchHeader)));
  const int4* d_items = (const int4*)(dev + r256(sizeof(BatchHeader)) + r256(sizeof(BatchReq) * (size_t)e->max_reqs));
  int32_t* d_ovf = (int32_t*)(dev + e->head_bytes - r256(sizeof(int32_t) * (size_t)e->max_reqs));
  int32_t* h_ovf = (int32_t*)(host + e->head_bytes - r256(sizeof(int32_t) * (size_t)e->max_reqs));
  size_t off = e->head_bytes;
  auto carve = [&](size_t b) {
    uint8_t* p = dev + off;
    off += r256(b);
    return p;
  };
  std::vector<Shape> shapes(jobs.size());
  std::vector<int4> rows_it, p1_it, p2_it, q_it;
  std::vector<std::pair<void*, const void*>> uploads;  // (device, host) input copies
  std::vector<size_t> upload_bytes;
  for (size_t r = 0; r < jobs.size(); r++) {
    BatchJob* j = jobs[r];
    Shape& s = shapes[r];
    if (!shape_of(e, &j->in, &j->opts, j->out->location, j->out->capacity_k, &s) || s.kind != kind)
      return fail_all(EVOC_ERR_INVALID_ARGUMENT, "request does not fit the batch");
    if (off + request_bytes(s, &j->in) > e->ws_bytes) return fail_all(EVOC_ERR_OUT_OF_BUDGET, "batch workspace full");
    const int64_t n = s.n;
    BatchReq q;
    std::memset(&q, 0, sizeof(q));
    const uint8_t* x = (const uint8_t*)j->in.data;
    if (s.host_in) {
      const size_t span = (size_t)((n - 1) * j->in.ld + s.d) * (size_t)s.elem;
      uint8_t* d = carve(span);
      uploads.push_back({d, j->in.data});
      upload_bytes.push_back(span);
      x = d;
    }
    q.x_in = x;
    q.ld_in = j->in.ld;
    if (s.dtype == EVOC_DTYPE_F32) {
      q.xf32 = (float*)carve((size_t)(n * s.f32_ld * 4));
      q.norms = (float*)carve((size_t)n * 4);
      q.flag = (int32_t*)carve(4);
    }
    q.f32_ld = s.f32_ld;
    q.rows = (s.dtype == EVOC_DTYPE_F32 && !s.bf16) ? (uint8_t*)q.xf32 : carve((size_t)(n * s.row_bytes));
    q.row_bytes = s.row_bytes;
    if (s.p.pass1) {
      q.gmax = (float*)carve((size_t)n * s.p.ng * 4);
      q.tau = (float*)carve((size_t)n * 4);
    }
    uint64_t* bufs = (uint64_t*)carve((size_t)s.p.n_splits * n * s.p.caps * 8);
    int32_t* counts = (int32_t*)carve((size_t)s.p.n_splits * n * 4);
    uint64_t* cand = (uint64_t*)carve((size_t)(n * s.K * 8));
    int32_t* qovf = (int32_t*)carve((size_t)n * 4);
    uint64_t* tkey = (uint64_t*)carve((size_t)n * 8);
    if (s.bf16) q.rkeys = (uint64_t*)carve((size_t)(n * s.k * 8));
    if (s.dev_out) {
      q.out_idx = j->out->indices;
      q.out_dist = j->out->distances;
    } else {
      q.out_idx = (int32_t*)carve((size_t)(n * s.k * 4));
      q.out_dist = (float*)carve((size_t)(n * s.k * 4));
    }
    q.ovf = d_ovf + r;
    q.n = (int32_t)n;
    q.d = (int32_t)s.d;
    q.K = (int32_t)s.K;
    q.k = (int32_t)s.k;
    q.ng = s.p.ng;
    q.has_p1 = s.p.pass1 ? 1 : 0;
    q.dtype = s.dtype;
    q.bf16 = s.bf16 ? 1 : 0;
    // the search's tile arguments, as exact_search sets them for one slab of all n queries
    TileArgs a = {};
    a.q_rows = q.rows;
    a.c_rows = q.rows;
    a.row_bytes = s.row_bytes;
    a.n_q = (int32_t)n;
    a.q_offset = 0;
    a.skip_self = s.seed_self ? 0 : 1;
    a.group_q = 1;
    const int64_t nqb = cdiv(n, 128);
    a.n_qblocks = (int32_t)nqb;
    if (s.p.pass1) {
      q.p1 = a;
      q.p1.n_c = s.p.n_sample;
      q.p1.c_stride = s.p.stride;
      q.p1.n_ctiles = (int32_t)cdiv(s.p.n_sample, 128);
      q.p1.tiles_per_split = s.p.p1_tiles_per_split;
      q.p1.n_splits = s.p.p1_splits;
      q.p1.gmax = q.gmax;
      q.p1.ng = s.p.ng;
    }
    q.p2 = a;
    q.p2.n_c = (int32_t)n;
    q.p2.c_stride = 1;
    q.p2.n_ctiles = (int32_t)cdiv(n, 128);
    q.p2.tiles_per_split = s.p.tiles_per_split;
    q.p2.n_splits = s.p.n_splits;
    q.p2.tau = s.p.pass1 ? q.tau : nullptr;
    q.p2.bufs = bufs;
    q.p2.counts = counts;
    q.p2.caps = s.p.caps;
    q.sel.bufs = bufs;
    q.sel.counts = counts;
    q.sel.n_q = (int32_t)n;
    q.sel.n_splits = s.p.n_splits;
    q.sel.caps = s.p.caps;
    q.sel.K = (int32_t)s.K;
    q.sel.q_offset = 0;
    q.sel.out_keys = cand;
    q.sel.overflow = qovf;
    q.sel.tau_key_out = tkey;
    ht[r] = q;
    const int ri = (int)r;
    for (int64_t r0 = 0; r0 < n; r0 += 64) rows_it.push_back(make_int4(ri, (int)r0, 0, 0));
    for (int64_t b = 0; b < nqb; b++) {
      if (s.p.pass1)
        for (int sp = 0; sp < s.p.p1_splits; sp++) p1_it.push_back(make_int4(ri, (int)b, sp, 0));
      for (int sp = 0; sp < s.p.n_splits; sp++) p2_it.push_back(make_int4(ri, (int)b, sp, 0));
    }
    for (int64_t q0 = 0; q0 < n; q0 += 4) q_it.push_back(make_int4(ri, (int)q0, 0, 0));
  }
  const size_t n_items = rows_it.size() + p1_it.size() + p2_it.size() + q_it.size();
  if (n_items > (size_t)MAX_ITEMS) return fail_all(EVOC_ERR_OUT_OF_BUDGET, "batch has too many work items");
  std::memset(hh, 0, sizeof(*hh));
  hh->n_reqs = (int32_t)jobs.size();
  size_t pos = 0;
  auto put = [&](const std::vector<int4>& v, int32_t& cnt, int32_t& at) {
    cnt = (int32_t)v.size();
    at = (int32_t)pos;
    if (!v.empty()) std::memcpy(hi + pos, v.data(), v.size() * sizeof(int4));
    pos += v.size();
  };
  put(rows_it, hh->n_rows, hh->off_rows);
  put(p1_it, hh->n_p1, hh->off_p1);
  put(p2_it, hh->n_p2, hh->off_p2);
  put(q_it, hh->n_q, hh->off_q);
  // ---- copies in, the launch sequence (captured once per kind), copies out ----
  hipError_t err = hipSuccess;
  auto ok = [&](hipError_t x) {
    if (err == hipSuccess) err = x;
    return err == hipSuccess;
  };
  const size_t up = (size_t)((uint8_t*)(hi + pos) - host);
  ok(hipMemcpyAsync(dev, host, up, hipMemcpyHostToDevice, st));
  // host inputs through the context's pinned staging ring when it has one
  // (pageable copies go through the runtime's own staging, far slower)
  for (size_t u = 0; u < uploads.size(); u++)
    ok(run->stage_bytes ? staged_upload(run, uploads[u].first, uploads[u].second, upload_bytes[u])
                        : hipMemcpyAsync(uploads[u].first, uploads[u].second, upload_bytes[u], hipMemcpyHostToDevice,
                                         st));
  const int ki = kind_index(kind);
  bool replay = false;
  if (err == hipSuccess && e->use_graph) {
    if (!e->graph[ki]) {
      hipGraph_t g = nullptr;
      ok(batch_prepare_attributes(kind));
      if (ok(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal))) {
        const hipError_t le = launch_batch(kind, e->n_cu, d_tab, d_items, d_hdr, *hh, st);
        const hipError_t ce = hipStreamEndCapture(st, &g);
        ok(le);
        ok(ce);
      }
      if (err == hipSuccess && ok(hipGraphInstantiate(&e->graph[ki], g, nullptr, nullptr, 0))) {
        e->graph_def[ki] = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    // fit the captured grids to this batch; if the graph cannot be updated the
    // passes are launched directly instead (same work, grids per launch)
    replay = err == hipSuccess && e->graph[ki] &&
             batch_set_grids(e->graph_def[ki], e->graph[ki], kind, e->n_cu, *hh) == hipSuccess;
    if (!replay) (void)hipGetLastError();
  }
  ok(hipEventRecord(e->e0, st));
  if (err == hipSuccess) {
    if (replay) {
      ok(hipGraphLaunch(e->graph[ki], st));
      e->replays++;
    } else {
      ok(batch_prepare_attributes(kind));
      ok(launch_batch(kind, e->n_cu, d_tab, d_items, d_hdr, *hh, st));
    }
  }
  ok(hipEventRecord(e->e1, st));
  ok(hipMemcpyAsync(h_ovf, d_ovf, sizeof(int32_t) * jobs.size(), hipMemcpyDeviceToHost, st));
  for (size_t r = 0; r < jobs.size(); r++) {
    const Shape& s = shapes[r];
    evoc_knn_output* o = jobs[r]->out;
    if (s.dev_out) continue;
    int32_t* oi = ht[r].out_idx;
    float* od = ht[r].out_dist;
    if (o->location != EVOC_MEM_DEVICE && run->stage_bytes) {
      ok(staged_download(run, o->indices, (size_t)o->capacity_k * 4, oi, (size_t)s.k * 4, (size_t)s.n));
      ok(staged_download(run, o->distances, (size_t)o->capacity_k * 4, od, (size_t)s.k * 4, (size_t)s.n));
      continue;
    }
    const hipMemcpyKind kd = o->location == EVOC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    ok(hipMemcpy2DAsync(o->indices, (size_t)o->capacity_k * 4, oi, (size_t)s.k * 4, (size_t)s.k * 4, (size_t)s.n, kd,
                        st));
    ok(hipMemcpy2DAsync(o->distances, (size_t)o->capacity_k * 4, od, (size_t)s.k * 4, (size_t)s.k * 4, (size_t)s.n,
                        kd, st));
  }
  ok(hipStreamSynchronize(st));
  std::vector<int32_t> ovf(h_ovf, h_ovf + jobs.size());
  if (err != hipSuccess) {
    (void)hipGetLastError();
    return fail_all(EVOC_ERR_HIP, std::string("batched search: ") + hipGetErrorString(err));
  }
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e->e0, e->e1);
  if (device_ms) *device_ms = ms;
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (size_t r = 0; r < jobs.size(); r++) {
    BatchJob* j = jobs[r];
    evoc_knn_output* o = j->out;
    j->overflow = ovf[r] != 0;
    j->status = j->overflow ? EVOC_ERR_PENDING : EVOC_OK;
    o->k_out = shapes[r].k;
    o->seconds = wall;
    o->device_seconds = ms * 1e-3;
    o->score_seconds = 0.0;
    o->iterations = 0;
    o->rescore_rounds = 0;
    o->offers_dropped = 0;
    o->score_pairs = shapes[r].n * shapes[r].n;
    o->slabs = 1;
  }
  return EVOC_OK;
}
