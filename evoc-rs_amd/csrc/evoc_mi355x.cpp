// Host side of the MI355X kNN-graph library: the C ABI declared in
// include/evoc_mi355x.h.
//
// knn_graph_impl (src/knn.rs:377-454) is restated here over device kernels:
// argument validation, the derived defaults (src/knn.rs:402-414), the caller's
// RNG advance (src/knn.rs:416, :124-129), prepare_float (src/knn.rs:95-113),
// the search itself, and the distance transform (src/knn.rs:156-187).
#include "evoc_mi355x.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <sched.h>

#include "evoc_host.h"
#include "evoc_kernels.h"

using namespace evoc;

// ---------------------------------------------------------------------------
// NumPy RandomState (MT19937 + legacy gauss): src/numpy_rng.rs:12-226.
// ---------------------------------------------------------------------------
struct evoc_numpy_rng {
  uint32_t key[624];
  int32_t pos;
  int32_t has_gauss;
  double gauss;
};

static void mt_seed(evoc_numpy_rng* r, uint32_t seed) {  // numpy_rng.rs:35-45
  r->key[0] = seed;
  for (uint32_t i = 1; i < 624; i++) {
    seed = 1812433253u * (seed ^ (seed >> 30)) + i;
    r->key[i] = seed;
  }
  r->pos = 624;
  r->has_gauss = 0;
  r->gauss = 0.0;
}

static void mt_refill(evoc_numpy_rng* r) {  // numpy_rng.rs:47-59
  static const uint32_t A = 0x9908b0dfu, UP = 0x80000000u, LO = 0x7fffffffu;
  uint32_t* k = r->key;
  for (int i = 0; i < 624; i++) {
    const uint32_t y = (k[i] & UP) | (k[(i + 1) % 624] & LO);
    k[i] = k[(i + 397) % 624] ^ (y >> 1) ^ ((0u - (y & 1u)) & A);
  }
  r->pos = 0;
}

static uint32_t mt_next(evoc_numpy_rng* r) {  // numpy_rng.rs:61-72
  if (r->pos >= 624) mt_refill(r);
  uint32_t y = r->key[r->pos++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

static double mt_double(evoc_numpy_rng* r) {  // numpy_rng.rs:74-78
  const int32_t a = (int32_t)(mt_next(r) >> 5), b = (int32_t)(mt_next(r) >> 6);
  return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

extern "C" {

evoc_numpy_rng* evoc_numpy_rng_new(uint32_t seed) {
  evoc_numpy_rng* r = new (std::nothrow) evoc_numpy_rng;
  if (r) mt_seed(r, seed);
  return r;
}

void evoc_numpy_rng_free(evoc_numpy_rng* r) { delete r; }

int64_t evoc_numpy_rng_randint(evoc_numpy_rng* r, int64_t low, int64_t high) {  // :81-112
  const uint64_t range = (uint64_t)(high - low);
  if (range == 0) return low;
  const uint64_t max = range - 1;
  if (max == 0) return low;
  uint64_t mask = max;
  mask |= mask >> 1;
  mask |= mask >> 2;
  mask |= mask >> 4;
  mask |= mask >> 8;
  mask |= mask >> 16;
  mask |= mask >> 32;
  for (;;) {
    const uint64_t v = (uint64_t)mt_next(r) & mask;
    if (v <= max) return low + (int64_t)v;
  }
}

double evoc_numpy_rng_random(evoc_numpy_rng* r) { return mt_double(r); }

double evoc_numpy_rng_normal(evoc_numpy_rng* r) {  // numpy_rng.rs:115-133
  if (r->has_gauss) {
    r->has_gauss = 0;
    const double t = r->gauss;
    r->gauss = 0.0;
    return t;
  }
  for (;;) {
    const double x1 = 2.0 * mt_double(r) - 1.0, x2 = 2.0 * mt_double(r) - 1.0;
    const double r2 = x1 * x1 + x2 * x2;
    if (r2 < 1.0 && r2 != 0.0) {
      const double f = std::sqrt(-2.0 * std::log(r2) / r2);
      r->gauss = f * x1;
      r->has_gauss = 1;
      return f * x2;
    }
  }
}

void evoc_numpy_rng_get_state(const evoc_numpy_rng* r, uint32_t key[624], int32_t* pos,
                              int32_t* has_gauss, double* gauss) {
  std::memcpy(key, r->key, sizeof(r->key));
  if (pos) *pos = r->pos;
  if (has_gauss) *has_gauss = r->has_gauss;
  if (gauss) *gauss = r->gauss;
}

void evoc_numpy_rng_set_state(evoc_numpy_rng* r, const uint32_t key[624], int32_t pos,
                              int32_t has_gauss, double gauss) {
  std::memcpy(r->key, key, sizeof(r->key));
  r->pos = pos;
  r->has_gauss = has_gauss;
  r->gauss = gauss;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Context, errors, device memory under an HBM budget
// ---------------------------------------------------------------------------
evoc_mi355x_ctx::~evoc_mi355x_ctx() {
  if (device >= 0) {
    (void)hipSetDevice(device);
    if (ev_begin) (void)hipEventDestroy(ev_begin);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (ev_score0) (void)hipEventDestroy(ev_score0);
    if (ev_score1) (void)hipEventDestroy(ev_score1);
    for (hipEvent_t ev : ev_join)
      if (ev) (void)hipEventDestroy(ev);
    while (!cache.empty()) free_cached(cache.size() - 1);
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
    if (pool) (void)hipMemPoolDestroy(pool);
    for (int i = 0; i < 2; i++) {
      if (stage[i]) (void)hipHostFree(stage[i]);
      if (stage_ev[i]) (void)hipEventDestroy(stage_ev[i]);
    }
  }
}

int Ctx::fail(int code, const std::string& msg) {
  last_error = msg;
  return code;
}

int Ctx::hip_fail(hipError_t e, const char* what) {
  last_error = std::string(what) + ": " + hipGetErrorString(e);
  return EVOC_ERR_HIP;
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release() {
  if (ptr && in_arena) {
    // a stack: the top block gives its bytes back; others wait for the call's end
    if ((uint8_t*)ptr + bytes == owner->arena + owner->arena_top) {
      owner->arena_top = (size_t)((uint8_t*)ptr - owner->arena);
      owner->in_use = owner->arena_top;
    }
    ptr = nullptr;
    bytes = 0;
    in_arena = false;
    return;
  }
  if (ptr) {
    if (owner->serving) {
      (void)hipFreeAsync(ptr, owner->stream);
    } else {
      owner->cache.push_back({ptr, bytes, owner->gen});
      owner->cached += bytes;
    }
    owner->in_use -= bytes;
    ptr = nullptr;
    bytes = 0;
  }
}

void Ctx::free_cached(size_t i) {
  (void)hipFreeAsync(cache[i].ptr, stream);
  cached -= cache[i].bytes;
  cache[i] = cache.back();
  cache.pop_back();
}

void Ctx::free_stale() {
  for (size_t i = cache.size(); i-- > 0;)
    if (cache[i].gen < gen) free_cached(i);
}

int DevBuf::alloc(Ctx* c, size_t nbytes, const char* what) {
  release();
  if (nbytes == 0) nbytes = 256;
  nbytes = (nbytes + 255) & ~(size_t)255;
  if (c->arena) {
    if (c->arena_top + nbytes > c->arena_size) {
      char buf[256];
      snprintf(buf, sizeof(buf),
               "allocation-free scratch exhausted allocating %s: %zu bytes requested, %zu of %zu in use",
               what, nbytes, c->arena_top, c->arena_size);
      return c->fail(EVOC_ERR_OUT_OF_BUDGET, buf);
    }
    ptr = c->arena + c->arena_top;
    bytes = nbytes;
    owner = c;
    in_arena = true;
    c->arena_top += nbytes;
    c->in_use = c->arena_top;
    return EVOC_OK;
  }
  // a cached block of about this size (at most 1/8 + 1 MiB larger)
  size_t best = SIZE_MAX;
  for (size_t i = 0; i < c->cache.size(); i++) {
    const size_t b = c->cache[i].bytes;
    if (b >= nbytes && b <= nbytes + nbytes / 8 + (1u << 20) && c->in_use + b <= c->budget &&
        (best == SIZE_MAX || b < c->cache[best].bytes))
      best = i;
  }
  if (best != SIZE_MAX) {
    ptr = c->cache[best].ptr;
    bytes = c->cache[best].bytes;
    owner = c;
    c->cached -= bytes;
    c->cache[best] = c->cache.back();
    c->cache.pop_back();
    c->in_use += bytes;
    return EVOC_OK;
  }
  // room for a new block: cached blocks go first, largest first
  while (!c->cache.empty() && c->in_use + c->cached + nbytes > c->budget) {
    size_t big = 0;
    for (size_t i = 1; i < c->cache.size(); i++)
      if (c->cache[i].bytes > c->cache[big].bytes) big = i;
    c->free_cached(big);
  }
  if (c->in_use + nbytes > c->budget) {
    char buf[256];
    snprintf(buf, sizeof(buf), "HBM budget exceeded allocating %s: %zu bytes requested, %zu of %zu in use",
             what, nbytes, (size_t)c->in_use, (size_t)c->budget);
    return c->fail(EVOC_ERR_OUT_OF_BUDGET, buf);
  }
  hipError_t e = hipMallocFromPoolAsync(&ptr, nbytes, c->pool, c->stream);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return c->hip_fail(e, what);
  }
  owner = c;
  bytes = nbytes;
  c->in_use += nbytes;
  return EVOC_OK;
}

int enable_staging(Ctx* c, size_t bytes) {
  if (c->stage_bytes) return EVOC_OK;
  if (hipSetDevice(c->device) != hipSuccess) return EVOC_ERR_HIP;
  for (int i = 0; i < 2; i++) {
    if (hipHostMalloc(&c->stage[i], bytes, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      return EVOC_ERR_HIP;
    }
  }
  c->stage_bytes = bytes;
  return EVOC_OK;
}

hipError_t staged_upload(Ctx* c, void* dst, const void* src, size_t bytes) {
  const size_t CH = c->stage_bytes;
  hipError_t e = hipSuccess;
  size_t i = 0;
  for (size_t off = 0; off < bytes && e == hipSuccess; off += CH, i++) {
    const int b = (int)(i & 1);
    const size_t len = std::min(CH, bytes - off);
    if (i >= 2) e = hipEventSynchronize(c->stage_ev[b]);  // the DMA that last read this buffer
    if (e != hipSuccess) break;
    std::memcpy(c->stage[b], (const uint8_t*)src + off, len);
    e = hipMemcpyAsync((uint8_t*)dst + off, c->stage[b], len, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->stage_ev[b], c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  return e;
}

hipError_t staged_download(Ctx* c, void* dst, size_t dst_pitch, const void* src, size_t row_bytes, size_t rows) {
  const size_t per = std::max<size_t>(1, c->stage_bytes / row_bytes);
  hipError_t e = hipSuccess;
  size_t r0 = 0, pending_r0 = 0, pending_n = 0;
  int b = 0, pending_b = 0;
  auto drain = [&]() {
    e = hipEventSynchronize(c->stage_ev[pending_b]);
    if (e != hipSuccess) return;
    for (size_t r = 0; r < pending_n; r++)
      std::memcpy((uint8_t*)dst + (pending_r0 + r) * dst_pitch, (const uint8_t*)c->stage[pending_b] + r * row_bytes,
                  row_bytes);
    pending_n = 0;
  };
  while (r0 < rows && e == hipSuccess) {
    const size_t nr = std::min(per, rows - r0);
    e = hipMemcpyAsync(c->stage[b], (const uint8_t*)src + r0 * row_bytes, nr * row_bytes, hipMemcpyDeviceToHost,
                       c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->stage_ev[b], c->stream);
    if (e != hipSuccess) break;
    if (pending_n) drain();  // the previous chunk, while this one is in flight
    pending_r0 = r0;
    pending_n = nr;
    pending_b = b;
    b ^= 1;
    r0 += nr;
  }
  if (e == hipSuccess && pending_n) drain();
  return e;
}

extern "C" {

#define EVOC_STR2(x) #x
#define EVOC_STR(x) EVOC_STR2(x)
const char* evoc_mi355x_version(void) {
  return "evoc_mi355x 0.1.0 (gfx950, abi " EVOC_STR(EVOC_MI355X_ABI_VERSION) ")";
}

int evoc_mi355x_device_count(int32_t* count) {
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return EVOC_ERR_NO_DEVICE;
  }
  *count = c;
  return EVOC_OK;
}

int evoc_mi355x_create(int32_t device, uint64_t hbm_budget_bytes, evoc_mi355x_ctx** out) {
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return EVOC_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= count) return EVOC_ERR_INVALID_ARGUMENT;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return EVOC_ERR_HIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return EVOC_ERR_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return EVOC_ERR_HIP;
  Ctx* c = new (std::nothrow) Ctx;
  if (!c) return EVOC_ERR_OUT_OF_BUDGET;
  c->device = device;
  c->n_cu = prop.multiProcessorCount;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = total_b = 0;
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)((double)free_b * 0.9);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
      hipEventCreate(&c->ev_score0) != hipSuccess || hipEventCreate(&c->ev_score1) != hipSuccess) {
    delete c;
    return EVOC_ERR_HIP;
  }
  for (hipEvent_t& ev : c->ev_join)
    if (hipEventCreate(&ev) != hipSuccess) {
      delete c;
      return EVOC_ERR_HIP;
    }
  // the allocation-free entry point's scratch address on this device (looked up
  // once per device; a failure is reported by that entry point, not here)
  (void)noalloc_scratch_init(device);
  // private pool; freed blocks stay cached for the next call until destroy
  hipMemPoolProps props = {};
  props.allocType = hipMemAllocationTypePinned;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  uint64_t keep = UINT64_MAX;
  if (hipMemPoolCreate(&c->pool, &props) != hipSuccess ||
      hipMemPoolSetAttribute(c->pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return EVOC_ERR_HIP;
  }
  *out = c;
  return EVOC_OK;
}

void evoc_mi355x_destroy(evoc_mi355x_ctx* ctx) { delete ctx; }

void evoc_mi355x_trim(evoc_mi355x_ctx* ctx) {
  if (!ctx) return;
  std::lock_guard<std::mutex> lock(ctx->mu);
  (void)hipSetDevice(ctx->device);
  while (!ctx->cache.empty()) ctx->free_cached(ctx->cache.size() - 1);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipMemPoolTrimTo(ctx->pool, 0);
}

const char* evoc_mi355x_last_error(const evoc_mi355x_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

void evoc_knn_options_default(evoc_knn_options* o) {  // src/knn.rs:60-76
  std::memset(o, 0, sizeof(*o));
  o->n_neighbors = 30;
  o->max_rptree_depth = 200;
  o->delta = 0.001f;
  o->delta_improv = 0.001f;
  o->use_sorted_updates = 1;
  o->brute_force_max_n = 0;
  o->deterministic = 0;
  o->mode = EVOC_MODE_EXACT;
  o->precision = EVOC_PREC_BF16_RESCORE;
  o->margin = 16;
  o->strict_precision = 0;
  o->rayon_threads = 0;
}

int64_t evoc_knn_effective_k(int64_t n_neighbors, int64_t n) {  // src/knn.rs:402
  int64_t k = std::min(n_neighbors, n > 0 ? n - 1 : 0);
  return k < 1 ? 1 : k;
}

// rayon::current_num_threads() of a caller that runs in rayon's global pool:
// RAYON_NUM_THREADS when set to a positive count, else the CPUs this process
// may run on (std::thread::available_parallelism, which rayon uses).
static int64_t host_rayon_threads() {
  if (const char* e = std::getenv("RAYON_NUM_THREADS")) {
    const long long v = std::atoll(e);
    if (v > 0) return (int64_t)v;
  }
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    const int c = CPU_COUNT(&set);
    if (c > 0) return c;
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h > 0 ? (int64_t)h : 1;
}

static int64_t default_n_trees(const evoc_knn_options* o) {  // src/knn.rs:404-407
  if (o->n_trees > 0) return o->n_trees;
  const int64_t t = o->deterministic ? 1 : (o->rayon_threads > 0 ? o->rayon_threads : host_rayon_threads());
  return std::min<int64_t>(std::max<int64_t>(t, 4), 8);
}

static bool brute_force_branch(int32_t dtype, int64_t n, const evoc_knn_options* o) {
  return n <= o->brute_force_max_n && dtype != EVOC_DTYPE_F32;  // src/knn.rs:418-419
}

int64_t evoc_knn_rng_draws(int32_t dtype, int64_t n, const evoc_knn_options* o) {
  if (brute_force_branch(dtype, n, o)) return 3;
  return 3 + 3 * default_n_trees(o);
}

int evoc_mi355x_knn_graph(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* o,
                          evoc_numpy_rng* rng, evoc_knn_output* out) {
  if (!ctx || !in || !o || !rng || !out) return EVOC_ERR_INVALID_ARGUMENT;
  // Validation happens before any draw, as in the reference (src/knn.rs:382-393).
  if (o->n_neighbors <= 0) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "n_neighbors must be positive");
  if (in->n <= 0) return ctx->fail(EVOC_ERR_EMPTY_DATA, "empty embedding matrix");
  const int64_t LO = (int64_t)INT32_MIN + 1, HI = (int64_t)INT32_MAX - 1;  // src/knn.rs:11-12
  int64_t tau[3];
  for (int i = 0; i < 3; i++) tau[i] = evoc_numpy_rng_randint(rng, LO, HI);
  std::vector<int64_t> trees;
  if (!brute_force_branch(in->dtype, in->n, o)) {
    const int64_t nt = default_n_trees(o);
    trees.resize((size_t)(3 * nt));
    for (auto& v : trees) v = evoc_numpy_rng_randint(rng, LO, HI);
  }
  return evoc_mi355x_knn_graph_seeded(ctx, in, o, tau, trees.empty() ? nullptr : trees.data(),
                                      (int64_t)(trees.size() / 3), out);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// knn_graph on the device
// ---------------------------------------------------------------------------
static size_t elem_size(int dtype) { return dtype == EVOC_DTYPE_F32 ? 4 : 1; }

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }
static uint64_t r256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

// Everything about a call that its validated arguments fix: search method,
// scored element kind, candidate count K, row layouts.  Shared by the entry
// point and by the device-memory estimate of the serving loop.
struct CallShape {
  int dtype = 0, kind = 0;
  bool bf16 = false, seed_self = true, nnd = false, dev_out = false, host_in = false;
  int64_t n = 0, d = 0, k = 0, K = 0;
  int64_t row_bytes = 0, f32_ld = 0, in_span = 0;
};

static CallShape call_shape(const evoc_knn_input* in, const evoc_knn_options* o, int32_t out_location,
                            int64_t out_capacity_k) {
  CallShape c;
  c.dtype = in->dtype;
  c.n = in->n;
  c.d = in->d;
  c.k = evoc_knn_effective_k(o->n_neighbors, c.n);
  c.bf16 = c.dtype == EVOC_DTYPE_F32 && o->precision == EVOC_PREC_BF16_RESCORE;
  // brute_force_knn (src/knn.rs:189-281) skips self; the NN-descent layout seeds it.
  c.seed_self = !brute_force_branch(c.dtype, c.n, o);
  // NN-descent mode runs the device NN-descent; the brute-force branch
  // (src/knn.rs:418-422) and the exact mode run the tiled exact search.
  c.nnd = o->mode == EVOC_MODE_NNDESCENT && c.seed_self;
  // k + margin candidates: the BF16 form re-scores them in FP32; NN-descent
  // keeps that many list entries per vertex in every precision (wider lists
  // explore more per iteration; the best k are emitted)
  c.K = (c.bf16 || c.nnd) ? std::min<int64_t>(c.k + std::max(0, o->margin), c.n) : c.k;
  if (c.nnd) c.K = std::min<int64_t>(c.K, 64);
  int64_t elem_bytes = 1;
  if (c.dtype == EVOC_DTYPE_F32) {
    c.kind = c.bf16 ? EVOC_ELEM_BF16 : EVOC_ELEM_F32;
    elem_bytes = c.bf16 ? 2 : 4;
    c.f32_ld = round_up(c.d, 32);
  } else {
    c.kind = c.dtype == EVOC_DTYPE_I8 ? EVOC_ELEM_I8 : EVOC_ELEM_U8BITS;
  }
  c.row_bytes = round_up(c.d * elem_bytes, 128);
  c.dev_out = out_location == EVOC_MEM_DEVICE && out_capacity_k == c.k;
  c.host_in = in->location != EVOC_MEM_DEVICE;
  c.in_span = (c.n - 1) * in->ld + c.d;
  return c;
}

// The FP32 re-score takes its queries in a cluster-coherent order unless
// EVOC_MI355X_RESCORE_ORDER=0 (the order changes speed only).
static bool rescore_order_enabled() {
  const char* e = std::getenv("EVOC_MI355X_RESCORE_ORDER");
  return !(e && e[0] == '0');
}

// Device bytes of everything the call holds besides the search's own scratch.
static uint64_t fixed_bytes(const CallShape& c) {
  uint64_t b = 0;
  if (c.host_in) b += r256((uint64_t)c.in_span * elem_size(c.dtype));
  b += r256((uint64_t)(c.n * c.row_bytes));
  if (c.dtype == EVOC_DTYPE_F32) {
    if (c.bf16) b += r256((uint64_t)(c.n * c.f32_ld * 4));
    b += r256((uint64_t)c.n * 4) + 256;
  }
  if (c.dtype == EVOC_DTYPE_U8) b += r256((uint64_t)c.n * 4);
  b += r256((uint64_t)(c.n * c.K * 8));
  if (c.bf16 || c.K != c.k) b += r256((uint64_t)(c.n * c.k * 8));
  if ((c.bf16 || c.K != c.k) && c.dtype == EVOC_DTYPE_F32 && rescore_order_enabled() && c.n >= 4096) {
    size_t tb = 0;  // the re-score's query order: four n-entry arrays and the sort's scratch
    (void)locality_order(nullptr, tb, nullptr, 0, c.n, (int)c.d, nullptr, nullptr, nullptr, nullptr, nullptr);
    b += 4 * r256((uint64_t)c.n * 4) + r256(tb > 0 ? tb : 1);
  }
  if (!c.dev_out) b += 2 * r256((uint64_t)(c.n * c.k * 4));
  return b;
}

static int64_t nnd_max_candidates(const evoc_knn_options* o, int64_t k) {
  return o->max_candidates > 0 ? o->max_candidates : std::min<int64_t>(60, (int64_t)((float)k * 1.5f));
}
static int64_t nnd_leaf_size(const evoc_knn_options* o, int64_t k) {
  return std::max<int64_t>(o->leaf_size > 0 ? o->leaf_size : std::max<int64_t>(10, k), std::max<int64_t>(10, k));
}

extern "C" int evoc_mi355x_request_bytes(const evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                                         const evoc_knn_options* o, int32_t out_location, uint64_t* full,
                                         uint64_t* minimum) {
  if (!ctx || !in || !o || in->n <= 0 || in->d <= 0 || o->n_neighbors <= 0) return EVOC_ERR_INVALID_ARGUMENT;
  const CallShape c = call_shape(in, o, out_location, evoc_knn_effective_k(o->n_neighbors, in->n));
  const uint64_t f = fixed_bytes(c);
  uint64_t hi, lo;
  if (c.nnd) {
    hi = lo = f + nndescent_bytes(c.kind, c.n, (int)c.K, (int)nnd_max_candidates(o, c.k), nnd_leaf_size(o, c.k),
                                  c.row_bytes, (int)c.d);
  } else {
    const int K = (int)std::min<int64_t>(c.K, 1024);
    hi = f + exact_scratch_bytes(c.kind, c.n, c.row_bytes, K, ctx->n_cu, false, ctx->serving);
    lo = f + exact_scratch_bytes(c.kind, c.n, c.row_bytes, K, ctx->n_cu, true, ctx->serving);
  }
  if (full) *full = hi;
  if (minimum) *minimum = lo;
  return EVOC_OK;
}

extern "C" int evoc_mi355x_exact_plan(const evoc_knn_input* in, const evoc_knn_options* o, int32_t n_cu,
                                      int32_t serving, evoc_exact_plan* out) {
  if (!in || !o || !out || in->n <= 0 || in->d <= 0 || o->n_neighbors <= 0 || n_cu <= 0 ||
      in->dtype < EVOC_DTYPE_F32 || in->dtype > EVOC_DTYPE_U8)
    return EVOC_ERR_INVALID_ARGUMENT;
  const CallShape c = call_shape(in, o, EVOC_MEM_DEVICE, evoc_knn_effective_k(o->n_neighbors, in->n));
  if (c.K > 1024) return EVOC_ERR_UNSUPPORTED;
  ExactPlanInfo p;
  exact_plan_info(c.kind, c.n, c.row_bytes, (int)c.K, n_cu, serving != 0, &p);
  std::memset(out, 0, sizeof(*out));
  out->symmetric = p.symmetric;
  out->n_splits = p.n_splits;
  out->caps = p.caps;
  out->ccap = p.ccap;
  out->stride = p.stride;
  out->n_sample = p.n_sample;
  out->segments = p.segments;
  out->tile_shape = p.tile_shape;
  out->candidates = (int32_t)c.K;
  out->bound_rank = p.bound_rank;
  out->work_items = p.work_items;
  out->scratch_bytes = p.scratch_bytes;
  return EVOC_OK;
}

// The whole of knn_graph_seeded once the context is locked: validation, input
// preparation, search, re-score, transform, output.  Every device buffer is a
// DevBuf of the context (its pool, or its arena in allocation-free calls).
static int knn_graph_run(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* o,
                         const int64_t tau_state[3], const int64_t* tree_states, int64_t n_tree_rows,
                         evoc_knn_output* out) {
  const auto t_start = std::chrono::steady_clock::now();
  ctx->last_error.clear();
  if (o->n_neighbors <= 0) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "n_neighbors must be positive");
  if (in->n <= 0) return ctx->fail(EVOC_ERR_EMPTY_DATA, "empty embedding matrix");
  if (in->dtype < EVOC_DTYPE_F32 || in->dtype > EVOC_DTYPE_U8)
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "unknown dtype");
  if (in->d <= 0 || in->ld < in->d || !in->data)
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "bad matrix shape (need d > 0, ld >= d, data != NULL)");
  if (in->n > (int64_t)INT32_MAX - 1)
    return ctx->fail(EVOC_ERR_UNSUPPORTED, "more than 2^31-2 rows");
  if (o->strict_precision)
    return ctx->fail(EVOC_ERR_NOT_BITWISE,
                     "strict_precision=true asks for results bitwise identical to the CPU strict "
                     "backend; the mi355x backend reproduces its neighbour sets with a different "
                     "f32 summation order (distances agree within 1e-6). Set strict_precision=false.");
  const int64_t n = in->n, d = in->d;
  const int64_t k = evoc_knn_effective_k(o->n_neighbors, n);
  if (out->capacity_k < k) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "output capacity_k < effective k");
  if (!out->indices || !out->distances) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "null output arrays");
  out->k_out = k;
  out->offers_dropped = 0;
  out->slabs = 0;
  out->score_pairs = 0;
  out->join_bytes = 0;
  out->join_seconds = 0.0;
  out->join_launches = 0;
  out->score_seconds = 0.0;
  out->iterations = 0;
  out->rescore_rounds = 0;
  if (o->mode != EVOC_MODE_EXACT && o->mode != EVOC_MODE_NNDESCENT)
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "unknown mode");
  const CallShape cs = call_shape(in, o, out->location, out->capacity_k);
  const bool bf16 = cs.bf16, nnd = cs.nnd;
  const int64_t K = cs.K;
  if (bf16 && K > 256)
    return ctx->fail(EVOC_ERR_UNSUPPORTED, "BF16 re-score keeps at most 256 candidates: need k + margin <= 256");
  if (K > 1024) return ctx->fail(EVOC_ERR_UNSUPPORTED, "exact search supports k <= 1024");
  if (nnd && k > 64) return ctx->fail(EVOC_ERR_UNSUPPORTED, "NN-descent mode supports k <= 64");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(EVOC_ERR_HIP, "hipSetDevice failed");

  hipStream_t st = ctx->stream;
  const int dtype = in->dtype;
  const bool seed_self = cs.seed_self;

  // ---- stage the input on the device ----
  const int64_t in_row_bytes = d * (int64_t)elem_size(dtype);
  DevBuf src;
  const uint8_t* d_src;
  if (!cs.host_in) {
    d_src = (const uint8_t*)in->data;
  } else {
    // the caller's matrix spans (n - 1) * ld + d elements
    const size_t span = (size_t)cs.in_span * elem_size(dtype);
    int rc = src.alloc(ctx, span, "input copy");
    if (rc) return rc;
    hipError_t e = ctx->stage_bytes ? staged_upload(ctx, src.ptr, in->data, span)
                                    : hipMemcpyAsync(src.ptr, in->data, span, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "input upload");
    d_src = (const uint8_t*)src.ptr;
  }
  (void)hipEventRecord(ctx->ev_begin, st);

  // scored rows: padded to a multiple of 128 bytes of the scored element type
  const int kind = cs.kind;
  const int64_t row_bytes = cs.row_bytes, f32_ld = cs.f32_ld;
  DevBuf rows, xf32, norms, flag, pop;
  int rc = rows.alloc(ctx, (size_t)(n * row_bytes), "scored rows");
  if (rc) return rc;
  hipError_t e = hipSuccess;
  if (dtype == EVOC_DTYPE_F32) {
    // prepared f32 rows (kept for the FP32 re-score) + the scored copy
    if (bf16) {
      if ((rc = xf32.alloc(ctx, (size_t)(n * f32_ld * 4), "prepared f32 rows"))) return rc;
    }
    if ((rc = norms.alloc(ctx, (size_t)n * 4, "row norms"))) return rc;
    if ((rc = flag.alloc(ctx, 4, "renorm flag"))) return rc;
    float* fout = bf16 ? (float*)xf32.ptr : (float*)rows.ptr;
    e = launch_prepare_float((const float*)d_src, in->ld, n, (int)d, fout, f32_ld,
                             bf16 ? (uint16_t*)rows.ptr : nullptr, bf16 ? row_bytes / 2 : 0,
                             (float*)norms.ptr, (int32_t*)flag.ptr, st);
  } else {
    e = launch_pad_rows(d_src, in->ld, n, (int)in_row_bytes, (uint8_t*)rows.ptr, row_bytes, st);
    if (e == hipSuccess && dtype == EVOC_DTYPE_U8) {
      if ((rc = pop.alloc(ctx, (size_t)n * 4, "row popcounts"))) return rc;
      e = launch_row_popcount((const uint8_t*)rows.ptr, row_bytes, n, (int)row_bytes,
                              (uint32_t*)pop.ptr, st);
    }
  }
  if (e != hipSuccess) return ctx->hip_fail(e, "input preparation");

  // ---- result buffers first, so the search can plan with what is left ----
  DevBuf keys, fkeys, oidx, odist;
  if ((rc = keys.alloc(ctx, (size_t)(n * K * 8), "candidate keys"))) return rc;
  if (bf16 || K != k) {
    if ((rc = fkeys.alloc(ctx, (size_t)(n * k * 8), "rescored keys"))) return rc;
  }
  int32_t* d_idx;
  float* d_dist;
  const bool dev_out = cs.dev_out;
  if (dev_out) {
    d_idx = out->indices;
    d_dist = out->distances;
  } else {
    if ((rc = oidx.alloc(ctx, (size_t)(n * k * 4), "output indices"))) return rc;
    if ((rc = odist.alloc(ctx, (size_t)(n * k * 4), "output distances"))) return rc;
    d_idx = (int32_t*)oidx.ptr;
    d_dist = (float*)odist.ptr;
  }

  // ---- search ----
  if (nnd) {
    NndParams P;
    P.k = (int)k;
    P.n_iters = o->n_iters > 0 ? (int)o->n_iters
                               : (int)std::max<int64_t>(5, (int64_t)std::llround(std::log2((double)n)));
    P.max_candidates = (int)nnd_max_candidates(o, k);
    P.n_trees = (int)default_n_trees(o);
    P.leaf_size = nnd_leaf_size(o, k);
    P.max_depth = o->max_rptree_depth;
    P.delta = o->delta;
    P.delta_improv = o->delta_improv;
    // the relative-improvement stop (float.rs:853-870) can only fire for a
    // positive threshold (the relative change is >= 0, and a NaN threshold
    // compares false), so its per-iteration distance sum is skipped otherwise
    P.use_delta_improv = o->delta_improv > 0.0f;
    P.seed = ((uint64_t)tau_state[0] * 0x9e3779b97f4a7c15ull) ^ ((uint64_t)tau_state[1] << 21) ^
             ((uint64_t)tau_state[2] << 42);
    for (int t = 0; t < P.n_trees; t++) {
      uint64_t ts = P.seed + 0x632be59bd9b4e019ull * (uint64_t)(t + 1);
      if (tree_states && t < n_tree_rows)
        ts = ((uint64_t)tree_states[3 * t] << 40) ^ ((uint64_t)tree_states[3 * t + 1] << 20) ^
             (uint64_t)tree_states[3 * t + 2];
      P.tree_seeds.push_back(ts);
    }
    const float* frows = dtype == EVOC_DTYPE_F32 ? (const float*)(bf16 ? xf32.ptr : rows.ptr) : nullptr;
    NndStats ns;
    rc = nndescent_search(ctx, kind, (const uint8_t*)rows.ptr, row_bytes, (const uint32_t*)pop.ptr, frows,
                          f32_ld, (int)d, n, (int)K, P, (uint64_t*)keys.ptr, &ns);
    if (rc) return rc;
    out->iterations = ns.iterations;
    out->offers_dropped = ns.dropped;
    out->join_bytes = (int64_t)ns.join_bytes;
    out->join_seconds = ns.join_ms * 1e-3;
    out->join_launches = ns.join_launches;
  } else {
    ExactStats xs;
    rc = exact_search(ctx, kind, (const uint8_t*)rows.ptr, row_bytes, (const uint32_t*)pop.ptr, n, (int)K,
                      seed_self, (uint64_t*)keys.ptr, &xs);
    if (rc) return rc;
    out->rescore_rounds = xs.extra_rounds;
    out->score_seconds = xs.pass2_ms * 1e-3;
    out->score_pairs = xs.score_pairs;
    out->slabs = xs.slabs;
  }

  const uint64_t* final_keys = (const uint64_t*)keys.ptr;
  if (bf16 || K != k) {
    if (dtype == EVOC_DTYPE_F32) {
      const float* xr = (const float*)(bf16 ? xf32.ptr : rows.ptr);
      // queries in a cluster-coherent order, so the candidate rows they gather
      // are shared by the queries in flight (cache hits instead of HBM reads)
      DevBuf ok, ok2, oin, oord, otemp;
      const int32_t* order = nullptr;
      if (rescore_order_enabled() && n >= 4096) {
        size_t tb = 0;
        e = locality_order(nullptr, tb, xr, f32_ld, n, (int)d, nullptr, nullptr, nullptr, nullptr, st);
        if (e != hipSuccess) return ctx->hip_fail(e, "rescore order");
        if ((rc = ok.alloc(ctx, (size_t)n * 4, "order keys")) || (rc = ok2.alloc(ctx, (size_t)n * 4, "order keys")) ||
            (rc = oin.alloc(ctx, (size_t)n * 4, "order")) || (rc = oord.alloc(ctx, (size_t)n * 4, "order")) ||
            (rc = otemp.alloc(ctx, tb > 0 ? tb : 1, "order sort scratch")))
          return rc;
        e = locality_order(otemp.ptr, tb, xr, f32_ld, n, (int)d, (uint32_t*)ok.ptr, (uint32_t*)ok2.ptr,
                           (int32_t*)oin.ptr, (int32_t*)oord.ptr, st);
        if (e != hipSuccess) return ctx->hip_fail(e, "rescore order");
        order = (const int32_t*)oord.ptr;
      }
      e = launch_rescore(xr, f32_ld, (int)d, (int)n, 0, (const uint64_t*)keys.ptr, (int)K, (int)k,
                         (uint64_t*)fkeys.ptr, st, order);
    } else {
      e = hipMemcpy2DAsync(fkeys.ptr, (size_t)k * 8, keys.ptr, (size_t)K * 8, (size_t)k * 8, (size_t)n,
                           hipMemcpyDeviceToDevice, st);
    }
    if (e != hipSuccess) return ctx->hip_fail(e, "rescore");
    final_keys = (const uint64_t*)fkeys.ptr;
  }

  // ---- transform + write out ----
  e = launch_finalize(final_keys, (int)n, 0, (int)k, dtype, d_idx, d_dist, st);
  if (e != hipSuccess) return ctx->hip_fail(e, "finalize");
  (void)hipEventRecord(ctx->ev_end, st);
  if (!dev_out && out->location != EVOC_MEM_DEVICE && ctx->stage_bytes) {
    e = staged_download(ctx, out->indices, (size_t)out->capacity_k * 4, d_idx, (size_t)k * 4, (size_t)n);
    if (e == hipSuccess)
      e = staged_download(ctx, out->distances, (size_t)out->capacity_k * 4, d_dist, (size_t)k * 4, (size_t)n);
    if (e != hipSuccess) return ctx->hip_fail(e, "output copy");
  } else if (!dev_out) {
    const hipMemcpyKind kd = out->location == EVOC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    e = hipMemcpy2DAsync(out->indices, (size_t)out->capacity_k * 4, d_idx, (size_t)k * 4, (size_t)k * 4,
                         (size_t)n, kd, st);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(out->distances, (size_t)out->capacity_k * 4, d_dist, (size_t)k * 4,
                           (size_t)k * 4, (size_t)n, kd, st);
    if (e != hipSuccess) return ctx->hip_fail(e, "output copy");
  }
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ctx->hip_fail(e, "knn_graph");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
  out->device_seconds = ms * 1e-3;
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  return EVOC_OK;
}

extern "C" int evoc_mi355x_knn_graph_seeded(evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                                            const evoc_knn_options* o, const int64_t tau_state[3],
                                            const int64_t* tree_states, int64_t n_tree_rows,
                                            evoc_knn_output* out) {
  if (!ctx || !in || !o || !out || !tau_state) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  CallScope call(ctx);
  return knn_graph_run(ctx, in, o, tau_state, tree_states, n_tree_rows, out);
}

namespace {
// The context in arena mode for one call: budget = the scratch, nothing taken
// from or returned to the pool, and the stream drained before the scratch is
// handed to the next call.
struct ArenaScope {
  Ctx* c;
  uint64_t budget, in_use;
  ArenaScope(Ctx* ctx, void* base, size_t bytes) : c(ctx), budget(ctx->budget), in_use(ctx->in_use) {
    c->arena = (uint8_t*)base;
    c->arena_size = bytes;
    c->arena_top = 0;
    c->budget = bytes;
    c->in_use = 0;
  }
  ~ArenaScope() {
    (void)hipStreamSynchronize(c->stream);
    c->arena = nullptr;
    c->arena_size = c->arena_top = 0;
    c->budget = budget;
    c->in_use = in_use;
  }
};
}  // namespace

extern "C" int evoc_mi355x_knn_graph_noalloc(evoc_mi355x_ctx* ctx, const evoc_knn_input* in,
                                             const evoc_knn_options* o, const int64_t tau_state[3],
                                             const int64_t* tree_states, int64_t n_tree_rows,
                                             evoc_knn_output* out) {
  if (!ctx || !in || !o || !out || !tau_state) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  ctx->last_error.clear();
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(EVOC_ERR_HIP, "hipSetDevice failed");
  std::mutex* dev_lock = nullptr;
  void* base = noalloc_scratch(ctx->device, &dev_lock);
  if (!base && noalloc_scratch_init(ctx->device) == EVOC_OK) base = noalloc_scratch(ctx->device, &dev_lock);
  if (!base || !dev_lock)
    return ctx->fail(EVOC_ERR_HIP, "the allocation-free scratch has no address on this device");
  // refuse before any device work a call whose smallest working set cannot fit
  if (o->n_neighbors > 0 && in->n > 0 && in->d > 0 && in->dtype >= EVOC_DTYPE_F32 && in->dtype <= EVOC_DTYPE_U8) {
    uint64_t full = 0, minimum = 0;
    if (evoc_mi355x_request_bytes(ctx, in, o, out->location, &full, &minimum) == EVOC_OK &&
        minimum > EVOC_MI355X_NOALLOC_SCRATCH_BYTES) {
      char buf[256];
      snprintf(buf, sizeof(buf),
               "the allocation-free entry point has %llu bytes of scratch; this call needs at least %llu "
               "(use evoc_mi355x_knn_graph_seeded)",
               (unsigned long long)EVOC_MI355X_NOALLOC_SCRATCH_BYTES, (unsigned long long)minimum);
      return ctx->fail(EVOC_ERR_OUT_OF_BUDGET, buf);
    }
  }
  std::lock_guard<std::mutex> dev_guard(*dev_lock);
  ArenaScope arena(ctx, base, (size_t)EVOC_MI355X_NOALLOC_SCRATCH_BYTES);
  return knn_graph_run(ctx, in, o, tau_state, tree_states, n_tree_rows, out);
}
