// NN-descent driver (host side of nndescent.hip): forest initialisation,
// random fill, and the candidate / local-join / merge iterations with the
// reference's stopping rules (src/nndescent/float.rs:775-879).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <utility>
#include <vector>

#include "evoc_host.h"
#include "evoc_kernels.h"
#include "evoc_mi355x.h"
#include "nndescent.h"

using namespace evoc;

namespace {

constexpr int UCAP = NND_OFFER_CAP;  // offers buffered per vertex between merges

uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Check {
  Ctx* ctx;
  hipError_t e = hipSuccess;
  const char* what = "";
  bool operator()(hipError_t r, const char* w) {
    if (r != hipSuccess && e == hipSuccess) {
      e = r;
      what = w;
    }
    return e == hipSuccess;
  }
  int fail() { return ctx->hip_fail(e, what); }
};

// Device time of the local-join launches: each launch between a pair of the
// context's events; the times are read after a host synchronisation (or, when
// every pair is in use, after waiting for the last one).
struct JoinTimer {
  Ctx* ctx;
  int used = 0;
  double ms = 0.0;
  int64_t launches = 0;
  void collect() {
    for (int i = 0; i < used; i++) {
      float t = 0.0f;
      if (hipEventElapsedTime(&t, ctx->ev_join[2 * i], ctx->ev_join[2 * i + 1]) == hipSuccess) ms += t;
    }
    launches += used;
    used = 0;
  }
  void begin(hipStream_t s) {
    if (used == Ctx::JOIN_EVENTS) {
      (void)hipEventSynchronize(ctx->ev_join[2 * used - 1]);
      collect();
    }
    (void)hipEventRecord(ctx->ev_join[2 * used], s);
  }
  void end(hipStream_t s) { (void)hipEventRecord(ctx->ev_join[2 * used++ + 1], s); }
};

}  // namespace

static int64_t forest_hyper_ld(int kind, int64_t row_bytes, int dim) {
  // binary: positive and negative masks; else floats, padded to a multiple of 4
  return kind == EVOC_ELEM_U8BITS ? (row_bytes * 2 + 3) / 4 : ((int64_t)dim + 3) / 4 * 4;
}

uint64_t forest_bytes(int kind, int64_t n, int dim, int64_t row_bytes, int leaf_size, int n_trees) {
  auto r = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  const int64_t T = std::max(1, n_trees);
  const int64_t max_seg = T * (n / (leaf_size + 1) + 2);
  return 3 * r(T * n * 4) + r(T * n) + 2 * r(max_seg * 16) + r(T * (2 * n + 2) * 8) + 256 +
         r(max_seg * forest_hyper_ld(kind, row_bytes, dim) * 4) + r(T * 8) + r(T * 4);
}

int forest_alloc(Ctx* ctx, ForestScratch& fs, int kind, int64_t n, int dim, int64_t row_bytes, int leaf_size,
                 int n_trees) {
  fs.n_trees = std::max(1, n_trees);
  const int64_t T = fs.n_trees;
  if (T * n > (int64_t)INT32_MAX - 1) return ctx->fail(EVOC_ERR_UNSUPPORTED, "forest: n_trees x rows >= 2^31");
  fs.max_seg = (int)(T * (n / (leaf_size + 1) + 2));
  fs.hyper_ld = forest_hyper_ld(kind, row_bytes, dim);
  int rc;
  if ((rc = fs.perm.alloc(ctx, (size_t)(T * n) * 4, "forest perm")) ||
      (rc = fs.perm2.alloc(ctx, (size_t)(T * n) * 4, "forest perm")) ||
      (rc = fs.seg_of.alloc(ctx, (size_t)(T * n) * 4, "forest segments")) ||
      (rc = fs.side.alloc(ctx, (size_t)(T * n), "forest sides")) ||
      (rc = fs.segs.alloc(ctx, (size_t)fs.max_seg * 16, "forest segments")) ||
      (rc = fs.nsegs.alloc(ctx, (size_t)fs.max_seg * 16, "forest segments")) ||
      (rc = fs.leaves.alloc(ctx, (size_t)(T * (2 * n + 2)) * 8, "forest leaves")) ||
      (rc = fs.fcnt.alloc(ctx, 16, "forest counters")) ||
      (rc = fs.hyper.alloc(ctx, (size_t)fs.max_seg * fs.hyper_ld * 4, "hyperplanes")) ||
      (rc = fs.seeds.alloc(ctx, (size_t)T * 8, "tree seeds")) || (rc = fs.depth.alloc(ctx, (size_t)T * 4, "tree depths")))
    return rc;
  return EVOC_OK;
}

ForestArgs forest_args(const ForestScratch& fs, int kind, int64_t n, int dim, int leaf_size, const uint8_t* rows,
                       int64_t row_bytes, const float* frows, int64_t frow_ld) {
  ForestArgs fa = {};
  fa.kind = kind;
  fa.n = (int32_t)n;
  fa.dim = dim;
  fa.leaf_size = leaf_size;
  fa.rows = rows;
  fa.row_bytes = row_bytes;
  fa.frows = frows;
  fa.frow_ld = frow_ld;
  fa.hyper = (float*)fs.hyper.ptr;
  fa.hyper_ld = fs.hyper_ld;
  fa.seeds = (const uint64_t*)fs.seeds.ptr;
  fa.depth = (int32_t*)fs.depth.ptr;
  fa.n_trees = fs.n_trees;
  return fa;
}

// All trees level by level together (float.rs:97-126 per tree): every segment
// longer than leaf_size is split until max_depth levels have run; the leaves
// are (flat start, length) ranges of the trees' row permutations.  One set of
// launches and one host synchronisation (the next level's segment count) per
// level for the whole forest.
int forest_build(Ctx* ctx, ForestScratch& fs, ForestArgs fa, int64_t max_depth, const std::vector<uint64_t>& seeds,
                 ForestSet* out) {
  hipStream_t s = ctx->stream;
  Check ck{ctx};
  const int n = fa.n;
  const int T = (int)seeds.size();
  if (T < 1 || T > fs.n_trees) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "forest: tree count");
  fa.n_trees = T;
  int32_t* d_perm = (int32_t*)fs.perm.ptr;
  int32_t* d_perm2 = (int32_t*)fs.perm2.ptr;
  int4* d_segs = (int4*)fs.segs.ptr;
  int4* d_next = (int4*)fs.nsegs.ptr;
  int32_t* d_fc = (int32_t*)fs.fcnt.ptr;  // [0] next segments, [1] leaves
  ck(upload_sync(fs.seeds.ptr, seeds.data(), (size_t)T * 8, s), "tree seeds");
  ck(launch_fill_i32((int32_t*)fs.depth.ptr, T, 0, s), "tree depths");
  ck(launch_fill_i32(d_fc, 4, 0, s), "forest counters");
  ck(rp_launch_init(d_perm, fa, s), "forest perm");
  int nseg;
  if (n > fa.leaf_size && max_depth > 0) {
    std::vector<int4> roots((size_t)T);
    for (int t = 0; t < T; t++) roots[(size_t)t] = make_int4(t * n, n, t, 0);
    ck(upload_sync(d_segs, roots.data(), roots.size() * sizeof(int4), s), "forest roots");
    nseg = T;
  } else {
    std::vector<int2> lv((size_t)T);
    for (int t = 0; t < T; t++) lv[(size_t)t] = make_int2(t * n, n);
    const int32_t cnt = T;
    ck(upload_sync(fs.leaves.ptr, lv.data(), lv.size() * sizeof(int2), s), "forest leaves");
    ck(upload_sync(d_fc + 1, &cnt, 4, s), "forest leaves");
    nseg = 0;
  }
  for (int level = 0; nseg > 0 && ck.e == hipSuccess; level++) {
    const int last = level + 1 >= max_depth ? 1 : 0;
    ck(launch_fill_i32(d_fc, 1, 0, s), "forest counters");
    ck(rp_launch_level(d_segs, nseg, d_perm, d_perm2, (int32_t*)fs.seg_of.ptr, (int8_t*)fs.side.ptr, fa, level,
                       d_next, d_fc, (int2*)fs.leaves.ptr, d_fc + 1, last, s),
       "forest level");
    std::swap(d_perm, d_perm2);
    std::swap(d_segs, d_next);
    int32_t nn = 0;
    ck(hipMemcpyAsync(&nn, d_fc, 4, hipMemcpyDeviceToHost, s), "forest counters");
    ck(hipStreamSynchronize(s), "forest level");
    nseg = nn;
  }
  int32_t nl = 0;
  out->depth.assign((size_t)T, 0);
  ck(hipMemcpyAsync(&nl, d_fc + 1, 4, hipMemcpyDeviceToHost, s), "forest leaves");
  ck(hipMemcpyAsync(out->depth.data(), fs.depth.ptr, (size_t)T * 4, hipMemcpyDeviceToHost, s), "tree depths");
  ck(hipStreamSynchronize(s), "forest");
  if (ck.e != hipSuccess) return ck.fail();
  out->perm = d_perm;
  out->leaves = (const int2*)fs.leaves.ptr;
  out->n_leaves = nl;
  return EVOC_OK;
}

uint64_t nndescent_bytes(int kind, int64_t n, int Kg, int max_candidates, int64_t leaf_size, int64_t row_bytes,
                         int dim) {
  auto r = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  const int64_t mc = std::min(max_candidates, 64);
  const int64_t ls = std::min<int64_t>(std::max<int64_t>(leaf_size, 2), 64);
  const int64_t nb = (n + 1023) / 1024;
  uint64_t b = r(n * Kg * 8) + 2 * r(n * 4) + r(n * UCAP * 8) + 2 * r(n * mc * 4) + 2 * r(n * 4) +
               r((n + 1) * 4) + r(n * Kg * 4) + r((nb + 2) * 4) + r(NND_COUNTER_BYTES) + 256;
  b += forest_bytes(kind, n, dim, row_bytes, (int)ls, NND_FOREST_TREES);
  return b;
}

int nndescent_search(Ctx* ctx, int kind, const uint8_t* rows, int64_t row_bytes, const uint32_t* pop,
                     const float* frows, int64_t frow_ld, int dim, int64_t n64, int Kg,
                     const NndParams& P, uint64_t* out_keys, NndStats* st) {
  hipStream_t s = ctx->stream;
  const int n = (int)n64;
  if (Kg < 1 || Kg > 64) return ctx->fail(EVOC_ERR_UNSUPPORTED, "NN-descent keeps at most 64 neighbours per vertex");
  const int mc = std::min(P.max_candidates, 64);
  if (mc < 1) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "max_candidates must be positive");
  const int leaf_size = std::min<int64_t>(std::max<int64_t>(P.leaf_size, 2), 64);
  const int nb = (n + 1023) / 1024;

  DevBuf keys, thr, umark, ubuf, newc, oldc, revcnt, revoff, revcur, rev, sums, counters, sumd;
  int rc;
  if ((rc = keys.alloc(ctx, (size_t)n * Kg * 8, "neighbour lists")) || (rc = thr.alloc(ctx, (size_t)n * 4, "list bounds")) ||
      (rc = umark.alloc(ctx, (size_t)n * 4, "offer marks")) ||
      (rc = ubuf.alloc(ctx, (size_t)n * UCAP * 8, "offer buffers")) ||
      (rc = newc.alloc(ctx, (size_t)n * mc * 4, "new candidates")) ||
      (rc = oldc.alloc(ctx, (size_t)n * mc * 4, "old candidates")) ||
      (rc = revcnt.alloc(ctx, (size_t)n * 4, "reverse counts")) ||
      (rc = revoff.alloc(ctx, (size_t)(n + 1) * 4, "reverse offsets")) ||
      (rc = revcur.alloc(ctx, (size_t)n * 4, "reverse cursors")) ||
      (rc = rev.alloc(ctx, (size_t)n * Kg * 4, "reverse edges")) ||
      (rc = sums.alloc(ctx, (size_t)(nb + 2) * 4, "scan sums")) ||
      (rc = counters.alloc(ctx, NND_COUNTER_BYTES, "counters")) || (rc = sumd.alloc(ctx, 8, "distance sum")))
    return rc;
  // inserted-entry slots, byte slots, lost-offer count (NND_COUNTER_BYTES)
  uint32_t* d_changed = (uint32_t*)counters.ptr;
  unsigned long long* d_jbytes = (unsigned long long*)((uint8_t*)counters.ptr + NND_COUNTER_SLOTS * 128);
  uint32_t* d_dropped = (uint32_t*)((uint8_t*)counters.ptr + NND_COUNTER_SLOTS * 128 * 2);
  // host copies of the slots (on the stack: the allocation-free entry point
  // runs this search too)
  uint32_t h_changed[NND_COUNTER_SLOTS * NND_SLOT_U32];
  unsigned long long h_jbytes[NND_COUNTER_SLOTS * NND_SLOT_U64];
  Check ck{ctx};
  JoinTimer jt{ctx};
  auto join = [&](const JoinArgs& args, const char* what) {
    jt.begin(s);
    ck(nnd_launch_join(kind, args, s), what);
    jt.end(s);
  };
  ck(launch_fill_i32((int32_t*)counters.ptr, (int64_t)(NND_COUNTER_BYTES / 4), 0, s), "counters");
  ck(nnd_launch_init((uint64_t*)keys.ptr, (float*)thr.ptr, (int32_t*)umark.ptr, n, Kg, s), "list init");
  ck(launch_fill_i32((int32_t*)ubuf.ptr, (int64_t)n * UCAP * 2, 0, s), "offer slots");  // EVOC_KEY_NONE

  JoinArgs ja = {};
  ja.rows = rows;
  ja.row_bytes = row_bytes;
  ja.row_pop = pop;
  ja.thr = (const float*)thr.ptr;
  ja.umark = (int32_t*)umark.ptr;
  ja.ubuf = (uint64_t*)ubuf.ptr;
  ja.ucap = UCAP;
  ja.n = n;
  // lost offers are counted only on request (a diagnostic: counting makes each
  // offer wait for its atomics' round trip, ~15 % of a 100k x 768 BF16 call)
  const char* cd = std::getenv("EVOC_MI355X_COUNT_DROPPED");
  const bool count_dropped = cd && cd[0] == '1';
  ja.dropped = count_dropped ? d_dropped : nullptr;
  ja.newc = (const int32_t*)newc.ptr;
  ja.oldc = (const int32_t*)oldc.ptr;
  ja.mc = mc;
  ja.bytes = d_jbytes;

  auto merge_all = [&]() {
    return ck(nnd_launch_merge((uint64_t*)keys.ptr, (float*)thr.ptr, (int32_t*)umark.ptr,
                               (uint64_t*)ubuf.ptr, UCAP, n, Kg, 0, n, d_changed, s),
              "merge");
  };

  // ---- random-projection forest: leaves seed the lists (float.rs:243-348) ----
  // All trees are split together (one set of launches per level); their leaves
  // are joined two trees at a time, with a merge after each pair so the later
  // trees' offers meet the lists the earlier ones left.
  if (P.n_trees > 0 && n > 1 && ck.e == hipSuccess) {
    const int T = std::min<int>(P.n_trees, NND_FOREST_TREES);
    for (int t0 = 0; t0 < P.n_trees && ck.e == hipSuccess; t0 += T) {
      const int nt = std::min(T, P.n_trees - t0);
      ForestScratch fs;
      if ((rc = forest_alloc(ctx, fs, kind, n, dim, row_bytes, leaf_size, nt))) return rc;
      ForestArgs fa = forest_args(fs, kind, n, dim, leaf_size, rows, row_bytes, frows, frow_ld);
      std::vector<uint64_t> seeds;
      for (int t = 0; t < nt; t++) seeds.push_back(mix(P.tree_seeds[(size_t)(t0 + t)]));
      ForestSet fo;
      if ((rc = forest_build(ctx, fs, fa, P.max_depth, seeds, &fo))) return rc;
      for (int g = 0; g < nt && fo.n_leaves > 0; g += 2) {
        JoinArgs jl = ja;
        jl.mode = 1;
        jl.n_items = fo.n_leaves;
        jl.leaves = fo.leaves;
        jl.perm = fo.perm;
        jl.tree_n = n;
        jl.tree_lo = g;
        jl.tree_hi = std::min(nt, g + 2);
        jl.salt = 1 + (uint32_t)(t0 + g);
        join(jl, "leaf join");
        merge_all();
      }
      if (st) st->leaves += fo.n_leaves;
    }
  }

  // ---- random fill of rows with empty slots (float.rs:350-390) ----
  ck(nnd_launch_random_lists((const uint64_t*)keys.ptr, n, Kg, mc, mix(P.seed ^ 0x5851f42d4c957f2dull),
                             (int32_t*)newc.ptr, (int32_t*)oldc.ptr, s),
     "random lists");
  {
    JoinArgs jr = ja;
    jr.mode = 0;
    jr.v0 = 0;
    jr.n_items = n;
    jr.salt = 2;
    join(jr, "random join");
    merge_all();
  }

  // ---- iterations (float.rs:807-874) ----
  double prev_sum = 0.0;
  bool have_prev = false;
  int64_t block = std::max<int64_t>(std::min<int64_t>(n, 16384), n / 8);
  if (const char* e = std::getenv("EVOC_MI355X_NND_BLOCK")) block = std::max<int64_t>(1024, std::atoll(e));
  int it = 0;
  for (; it < P.n_iters && ck.e == hipSuccess; it++) {
    ck(launch_fill_i32((int32_t*)d_changed, (int64_t)NND_COUNTER_SLOTS * NND_SLOT_U32, 0, s), "counters");
    ck(nnd_launch_candidates((uint64_t*)keys.ptr, n, Kg, (int32_t*)revcnt.ptr, (int32_t*)revoff.ptr,
                             (int32_t*)revcur.ptr, (int32_t*)rev.ptr, (int32_t*)sums.ptr, mc, P.seed, it,
                             (int32_t*)newc.ptr, (int32_t*)oldc.ptr, s),
       "candidates");
    for (int64_t v0 = 0; v0 < n && ck.e == hipSuccess; v0 += block) {
      JoinArgs jj = ja;
      jj.mode = 0;
      jj.v0 = (int32_t)v0;
      jj.n_items = (int32_t)std::min<int64_t>(block, n - v0);
      jj.salt = 3 + (uint32_t)it * 4096u + (uint32_t)(v0 / 1024);
      join(jj, "local join");
      merge_all();
    }
    block = std::min<int64_t>(block, n) * 2;  // float.rs:872-873
    ck(hipMemcpyAsync(h_changed, d_changed, sizeof(h_changed), hipMemcpyDeviceToHost, s), "counters");
    if (P.use_delta_improv) {
      const double zero = 0.0;
      ck(upload_sync(sumd.ptr, &zero, 8, s), "distance sum");
      ck(nnd_launch_sum_dist((const uint64_t*)keys.ptr, (int64_t)n * Kg, (double*)sumd.ptr, s), "distance sum");
    }
    double sum = 0.0;
    if (P.use_delta_improv) ck(hipMemcpyAsync(&sum, sumd.ptr, 8, hipMemcpyDeviceToHost, s), "distance sum");
    ck(hipStreamSynchronize(s), "iteration");
    uint32_t c = 0;
    for (int i = 0; i < NND_COUNTER_SLOTS; i++) c += h_changed[(size_t)i * NND_SLOT_U32];
    if (st) st->changed.push_back(c);
    // float.rs:847: c <= (delta * k * n) as u32
    const float lim = P.delta * (float)P.k * (float)n;
    if ((double)c <= (lim <= 0.0f ? 0.0 : (double)(uint32_t)lim)) {
      it++;
      break;
    }
    if (P.use_delta_improv) {  // float.rs:853-870
      if (have_prev && std::abs(sum - prev_sum) / std::abs(prev_sum) < (double)P.delta_improv) {
        it++;
        break;
      }
      prev_sum = sum;
      have_prev = true;
    }
  }
  if (st) st->iterations = it;
  ck(nnd_launch_export((const uint64_t*)keys.ptr, n, Kg, out_keys, s), "export");
  uint32_t dropped = 0;
  ck(hipMemcpyAsync(&dropped, d_dropped, 4, hipMemcpyDeviceToHost, s), "counters");
  ck(hipMemcpyAsync(h_jbytes, d_jbytes, sizeof(h_jbytes), hipMemcpyDeviceToHost, s), "counters");
  ck(hipStreamSynchronize(s), "NN-descent");
  unsigned long long jbytes = 0;
  for (int i = 0; i < NND_COUNTER_SLOTS; i++) jbytes += h_jbytes[(size_t)i * NND_SLOT_U64];
  jt.collect();
  if (st) {
    st->dropped = count_dropped ? (int64_t)dropped : -1;
    st->join_bytes = jbytes;
    st->join_ms = jt.ms;
    st->join_launches = jt.launches;
  }
  if (ck.e != hipSuccess) return ck.fail();
  return EVOC_OK;
}

// make_forest (src/knn.rs:115-154) on the device: the forest NN-descent mode
// starts from, returned in the reference's leaf-array layout.
extern "C" int evoc_mi355x_make_forest(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* o,
                                       const int64_t* tree_states, int64_t n_trees, int32_t* leaf_array,
                                       int64_t capacity_rows, int64_t* rows_out, int64_t* cols_out,
                                       int64_t* tree_leaves, int32_t* depths) {
  if (!ctx || !in || !o || !rows_out || !cols_out) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  CallScope call(ctx);
  ctx->last_error.clear();
  *rows_out = *cols_out = 0;
  if (n_trees < 0 || (n_trees > 0 && !tree_states)) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "tree states");
  if (in->n <= 0) return ctx->fail(EVOC_ERR_EMPTY_DATA, "empty embedding matrix");
  if (in->n > (int64_t)INT32_MAX - 1) return ctx->fail(EVOC_ERR_UNSUPPORTED, "more than 2^31-2 rows");
  if (in->dtype < EVOC_DTYPE_F32 || in->dtype > EVOC_DTYPE_U8 || in->d <= 0 || in->ld < in->d || !in->data)
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "bad matrix (dtype, d > 0, ld >= d, data != NULL)");
  if (o->n_neighbors <= 0) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "n_neighbors must be positive");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(EVOC_ERR_HIP, "hipSetDevice failed");
  const int64_t n = in->n, d = in->d;
  const int64_t k = std::max<int64_t>(1, std::min<int64_t>(o->n_neighbors, n - 1));
  // leaf_size as knn_graph derives it (src/knn.rs:408-411, float.rs:148-151), at most 64 on the device
  const int64_t ls_ref = std::max<int64_t>(o->leaf_size > 0 ? o->leaf_size : std::max<int64_t>(10, k),
                                           std::max<int64_t>(10, k));
  const int leaf_size = (int)std::min<int64_t>(std::max<int64_t>(ls_ref, 2), 64);
  if (ls_ref > 64) return ctx->fail(EVOC_ERR_UNSUPPORTED, "the device forest keeps leaves of at most 64 rows");
  const int64_t max_depth = o->max_rptree_depth;
  hipStream_t s = ctx->stream;
  const bool f32 = in->dtype == EVOC_DTYPE_F32;
  const int kind = f32 ? EVOC_ELEM_F32 : in->dtype == EVOC_DTYPE_I8 ? EVOC_ELEM_I8 : EVOC_ELEM_U8BITS;
  const int64_t elem = f32 ? 4 : 1;
  const int64_t row_bytes = ((d * elem + 127) / 128) * 128;
  const int64_t f32_ld = ((d + 31) / 32) * 32;
  // the input on the device, then the rows the forest reads (normalised f32 as
  // prepare_float leaves them, src/knn.rs:95-113; padded i8 / binary rows)
  DevBuf src, rows, norms, flag;
  const uint8_t* d_src = (const uint8_t*)in->data;
  int rc;
  if (in->location != EVOC_MEM_DEVICE) {
    const size_t span = (size_t)((n - 1) * in->ld + d) * (size_t)elem;
    if ((rc = src.alloc(ctx, span, "input copy"))) return rc;
    if (hipMemcpyAsync(src.ptr, in->data, span, hipMemcpyHostToDevice, s) != hipSuccess)
      return ctx->fail(EVOC_ERR_HIP, "input upload");
    d_src = (const uint8_t*)src.ptr;
  }
  hipError_t e;
  if (f32) {
    if ((rc = rows.alloc(ctx, (size_t)(n * f32_ld * 4), "prepared f32 rows")) ||
        (rc = norms.alloc(ctx, (size_t)n * 4, "row norms")) || (rc = flag.alloc(ctx, 4, "renorm flag")))
      return rc;
    e = launch_prepare_float((const float*)d_src, in->ld, n, (int)d, (float*)rows.ptr, f32_ld, nullptr, 0,
                             (float*)norms.ptr, (int32_t*)flag.ptr, s);
  } else {
    if ((rc = rows.alloc(ctx, (size_t)(n * row_bytes), "scored rows"))) return rc;
    e = launch_pad_rows(d_src, in->ld, n, (int)d, (uint8_t*)rows.ptr, row_bytes, s);
  }
  if (e != hipSuccess) return ctx->hip_fail(e, "input preparation");
  std::vector<std::vector<int32_t>> tree_rows((size_t)n_trees);  // concatenated leaf members
  std::vector<std::vector<int2>> tree_lv((size_t)n_trees);
  int64_t total = 0, width = leaf_size;
  for (int64_t t0 = 0; t0 < n_trees; t0 += NND_FOREST_TREES) {
    const int nt = (int)std::min<int64_t>(NND_FOREST_TREES, n_trees - t0);
    ForestScratch fs;
    if ((rc = forest_alloc(ctx, fs, kind, n, (int)d, row_bytes, leaf_size, nt))) return rc;
    ForestArgs fa = forest_args(fs, kind, n, (int)d, leaf_size, (const uint8_t*)rows.ptr, row_bytes,
                                f32 ? (const float*)rows.ptr : nullptr, f32_ld);
    // each tree's seed from its three draws, as NN-descent mode takes it
    std::vector<uint64_t> seeds;
    for (int t = 0; t < nt; t++) {
      const int64_t tt = t0 + t;
      seeds.push_back(mix(((uint64_t)tree_states[3 * tt] << 40) ^ ((uint64_t)tree_states[3 * tt + 1] << 20) ^
                          (uint64_t)tree_states[3 * tt + 2]));
    }
    ForestSet fo;
    if ((rc = forest_build(ctx, fs, fa, max_depth, seeds, &fo))) return rc;
    std::vector<int2> all((size_t)fo.n_leaves);
    std::vector<int32_t> perm((size_t)(nt * n));
    e = hipMemcpyAsync(all.data(), fo.leaves, all.size() * sizeof(int2), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(perm.data(), fo.perm, perm.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ctx->hip_fail(e, "forest download");
    for (int t = 0; t < nt; t++) {
      const int64_t tt = t0 + t;
      tree_rows[(size_t)tt].assign(perm.begin() + (int64_t)t * n, perm.begin() + (int64_t)(t + 1) * n);
    }
    // leaves of tree t, starts made tree-local
    for (const int2& l : all) {
      const int t = (int)(l.x / n);
      tree_lv[(size_t)(t0 + t)].push_back(make_int2((int)(l.x - (int64_t)t * n), l.y));
    }
    for (int t = 0; t < nt; t++) {
      const int64_t tt = t0 + t;
      std::vector<int2>& lv = tree_lv[(size_t)tt];
      // each tree: its leaves in the order of their first row in the tree's permutation
      std::sort(lv.begin(), lv.end(), [](const int2& a, const int2& b) { return a.x < b.x; });
      for (const int2& l : lv) width = std::max<int64_t>(width, l.y);
      total += (int64_t)lv.size();
      if (tree_leaves) tree_leaves[tt] = (int64_t)lv.size();
      if (depths) depths[tt] = fo.depth[(size_t)t];
    }
  }
  *rows_out = total;
  *cols_out = width;
  if (!leaf_array || capacity_rows < total)
    return leaf_array ? ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "leaf_array holds fewer rows than the forest has leaves")
                      : EVOC_OK;
  int64_t r = 0;
  for (int64_t t = 0; t < n_trees; t++)
    for (const int2& l : tree_lv[(size_t)t]) {
      int32_t* dst = leaf_array + r * width;
      for (int64_t j = 0; j < width; j++) dst[j] = j < l.y ? tree_rows[(size_t)t][(size_t)(l.x + j)] : -1;
      r++;
    }
  return EVOC_OK;
}
