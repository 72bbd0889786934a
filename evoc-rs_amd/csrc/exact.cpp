// Exact search orchestration: plans the three passes of knn_exact.hip for a
// matrix, processes queries in slabs that fit the HBM budget, and re-scores
// query blocks whose candidate buffers overflowed with a tighter bound until
// every query's K best are exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "evoc_host.h"
#include "evoc_kernels.h"
#include "evoc_mi355x.h"

using namespace evoc;

namespace {

constexpr int MAX_SORT = 2048;  // largest per-query candidate list one wave sorts
constexpr int MAX_ROUNDS = 8;
constexpr int SELECT_MAX_SEGMENTS = 64;  // segments select_topk_kernel merges per query (one per lane)

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
int64_t pow2ceil(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Tiles per candidate split: the split count whose workgroups fill the last
// dispatch round best (busy fraction of the resident slots), fewer splits on ties.
// At least 4 splits when there are enough tiles: the workgroups resident on an
// XCD then share query tiles as well as candidate tiles (see tile_coords).
int best_split(int64_t nqb, int64_t n_ctiles, int64_t slots) {
  int64_t best_tps = n_ctiles;
  double best_eff = -1.0;
  const int64_t ns_min = n_ctiles >= 64 ? 4 : 1;
  for (int64_t ns = ns_min; ns <= std::min<int64_t>(16, n_ctiles); ns++) {
    const int64_t tps = cdiv(n_ctiles, ns);
    const int64_t real_ns = cdiv(n_ctiles, tps);
    const double rounds = (double)cdiv(nqb * real_ns, slots);
    const double eff = (double)(nqb * n_ctiles) / (rounds * (double)tps * (double)slots);
    if (eff > best_eff + 1e-3) {
      best_eff = eff;
      best_tps = tps;
    }
  }
  return (int)best_tps;
}

struct Plan {
  int n_splits = 1, tiles_per_split = 1, caps = 0;
  bool pass1 = false;
  int stride = 1, n_sample = 0, ng = 0;
  int J = 0;  // pass 2 tests the J-th largest group maximum (J < K: an estimate, see bound_rank)
  int p1_tiles_per_split = 1, p1_splits = 1;
  int group_q = 1, p1_group_q = 1;  // query blocks per XCD group (tile_coords)
  int K = 0;
  // device bytes exact_search allocates for a slab of nq queries (each buffer
  // rounded to 256 bytes by DevBuf)
  size_t scratch_bytes(int64_t nq, int64_t nqb) const {
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t b = r((size_t)n_splits * nq * caps * 8) + r((size_t)n_splits * nq * 4) + r((size_t)nq * 4) +
               r((size_t)nq * 8) + r((size_t)(nqb + nq) * 4);
    if (pass1) b += r((size_t)nq * ng * 4) + r((size_t)nq * 4) + (J < K ? r((size_t)nq * 4) : 0);
    return b;
  }
};

// Candidate rows the threshold pass samples (every stride-th row): a larger
// sample costs pass 1 more and tightens the bound, so fewer keys are appended
// and selected.  The symmetric pass takes 8192 (the threshold pass costs
// n x sample, the appends it saves about n^2 K / sample; on 768-d rows 6k-8k
// rows balance them at 20k-200k rows), the per-split pass of query slabs and
// serving requests 16384 (with 8192 a serving stream at 60 requests/s fell
// behind: p50 8.7 ms -> 2.8 s).  The stride cap in make_plan decides the
// sample beyond ~130k / ~260k rows.  EVOC_MI355X_SAMPLE_ROWS overrides both
// (a tuning knob; results do not depend on it).
int64_t sample_rows(bool symmetric) {
  static const int64_t env = [] {
    const char* e = std::getenv("EVOC_MI355X_SAMPLE_ROWS");
    const long long r = e ? std::atoll(e) : 0;
    return r > 0 ? (int64_t)r : (int64_t)0;
  }();
  return env ? env : symmetric ? 8192 : 16384;
}

// Per-query probability that a query's estimated bound misses (see bound_rank):
// 1e-6 by default; EVOC_MI355X_BOUND_EPS overrides it (a tuning knob, read at
// every plan: results do not depend on it, a missed bound only costs a re-score;
// 1 makes the bound the guaranteed one).
double bound_eps() {
  const char* e = std::getenv("EVOC_MI355X_BOUND_EPS");
  const double v = e ? std::atof(e) : 0.0;
  return v > 0.0 ? v : 1e-6;
}

// The rank J of the group maximum that bounds pass 2.  The K-th largest of a
// query's group maxima is a guaranteed lower bound on its K-th best score, but
// a loose one: the strided sample holds about K / stride of the query's K best,
// so that bound sits near rank K * stride of all candidates, and every
// candidate above it is appended.  The J-th largest group maximum exceeds the
// K-th best score only if J sampled candidates rank among the query's K - 1
// best; with every stride-th row sampled that count is about Binomial(K - 1,
// 1 / stride).  J is the smallest rank whose tail probability is at most
// bound_eps(): appended keys fall from about K * stride to J * stride per query
// (at 1,000,000 rows and K = 46: rank 46 -> 11), and a query the estimate
// misses ends pass 2 with fewer than K keys, is flagged by selection and
// scored again against the guaranteed bound.  The result does not depend on J.
int bound_rank(int K, int stride) {
  if (stride <= 1 || K <= 1) return K;
  const int m = K - 1;
  const double p = 1.0 / stride, eps = bound_eps();
  if (eps >= 1.0) return K;
  // tail(j) = P(X >= j), X ~ Binomial(m, p), summed from the top down
  std::vector<double> pmf((size_t)m + 1);
  for (int i = 0; i <= m; i++)
    pmf[(size_t)i] = std::exp(std::lgamma(m + 1.0) - std::lgamma(i + 1.0) - std::lgamma(m - i + 1.0) +
                              i * std::log(p) + (m - i) * std::log1p(-p));
  double tail = 0.0;
  int J = K;  // P(X >= K) = 0: the guaranteed bound
  for (int j = m; j >= 1; j--) {
    tail += pmf[(size_t)j];
    if (tail > eps) break;
    J = j;
  }
  return J;
}

Plan make_plan(int kind, int64_t n, int64_t nq, int K, int n_cu, bool symmetric = false,
               bool guaranteed = false) {
  Plan p;
  p.K = K;
  p.J = K;
  const int TC = tile_rows_c(kind), TQ = tile_rows_q(kind);
  const int64_t n_ctiles = cdiv(n, TC);
  const int64_t nqb = cdiv(nq, TQ);
  // workgroups resident per CU: LDS-limited (64 KiB tiles fit twice in 160 KiB)
  const int64_t per_cu = tile_lds_bytes(kind, 2) <= 80 * 1024 ? 2 : 1;
  const int64_t slots = (int64_t)n_cu * per_cu;
  p.tiles_per_split = best_split(nqb, n_ctiles, slots);
  p.n_splits = (int)cdiv(n_ctiles, p.tiles_per_split);
  const int64_t per_xcd = std::max<int64_t>(slots / 8, 1);  // workgroups resident per XCD
  p.group_q = (int)std::max<int64_t>(1, per_xcd / p.n_splits);
  if (n <= MAX_SORT) {
    // every candidate fits one sorted list: no threshold pass
    p.pass1 = false;
    p.caps = p.tiles_per_split * TC;
    return p;
  }
  p.pass1 = true;
  const int gpt = groups_per_tile(kind);
  int64_t stride = std::min<int64_t>(std::max<int64_t>(cdiv(n, sample_rows(symmetric)), 1), 16);
  for (;;) {
    const int64_t nsamp = cdiv(n, stride);
    const int64_t ng = cdiv(nsamp, TC) * gpt;
    if (ng <= MAX_SORT || stride >= n) {
      p.stride = (int)stride;
      p.n_sample = (int)nsamp;
      p.ng = (int)ng;
      break;
    }
    stride++;
  }
  p.J = guaranteed ? K : bound_rank(K, p.stride);
  // a query's bound ranks near its J-th .. 2J-th sample group, i.e. about
  // stride * 2J candidates of the full set pass it; keep 4x headroom
  const int64_t expect = (int64_t)p.stride * 2 * p.J;
  int64_t caps = std::max<int64_t>(pow2ceil(cdiv(4 * expect, p.n_splits)), 128);
  // the kept subset must hold 2K keys so an overflow round tightens the bound
  caps = std::max<int64_t>(caps, pow2ceil(cdiv(2LL * K, p.n_splits)));
  p.caps = (int)std::min<int64_t>(caps, MAX_SORT);
  const int64_t s_tiles = cdiv(p.n_sample, TC);
  p.p1_tiles_per_split = best_split(nqb, s_tiles, slots);
  p.p1_splits = (int)cdiv(s_tiles, p.p1_tiles_per_split);
  p.p1_group_q = (int)std::max<int64_t>(1, per_xcd / p.p1_splits);
  return p;
}

// Symmetric plan (query rows = candidate rows, the whole matrix one slab).
// Candidate tiles are cut into splits of tiles_per_split as for pass 2, and
// query block Q takes split s's tiles from max(Q, s * tiles_per_split) on, so
// each pair of tiles Q <= t is scored once; a row's keys from query blocks
// before its own arrive on the candidate side (ccap slots per row).  Splits
// are chosen so that about 16 work items run per resident workgroup slot (the
// last dispatch round then costs little), at least 4 when there are enough
// tiles.  Items run in groups of group_q query blocks, split-major inside a
// group, as tile_coords orders pass 2: the items resident on an XCD at once
// share candidate tiles (same split) and query tiles (same block).
struct SymPlan {
  Plan p;
  int ccap = 0;
  std::vector<int4> work;
  size_t scratch_bytes(int64_t n) const {
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return p.scratch_bytes(n, nqb) + r((size_t)n * ccap * 8) + r((size_t)n * 4) + r(work.size() * sizeof(int4));
  }
  int64_t nqb = 0;
};

SymPlan make_sym_plan(int kind, int64_t n, int K, int n_cu) {
  SymPlan sp;
  const int T = tile_rows_c(kind);
  const int64_t nt = cdiv(n, T);
  sp.nqb = nt;
  sp.p = make_plan(kind, n, n, K, n_cu, true);
  Plan& p = sp.p;
  if (!p.pass1 || tile_rows_q(kind) != T) return sp;
  const int64_t per_cu = tile_lds_bytes(kind, 3) <= 80 * 1024 ? 2 : 1;
  const int64_t slots = (int64_t)n_cu * per_cu;
  const int64_t pairs = nt * (nt + 1) / 2;
  int64_t tpi = std::min<int64_t>(nt, std::max<int64_t>(1, pairs / (16 * slots)));
  if (nt >= 64 && cdiv(nt, tpi) < 4) tpi = cdiv(nt, 4);
  // selection merges a row's split buffers and its candidate-side list in
  // registers only for at most SELECT_MAX_SEGMENTS segments (more go through
  // the slower streaming kernel), so at most that many minus one splits
  tpi = std::max<int64_t>(tpi, cdiv(nt, SELECT_MAX_SEGMENTS - 1));
  p.tiles_per_split = (int)tpi;
  p.n_splits = (int)cdiv(nt, tpi);
  // the items resident on one XCD at a time form a block of about group_q query
  // blocks x 8 splits: each candidate tile of a step is then fetched once for
  // group_q items, each query tile once for its whole row of splits
  p.group_q = (int)std::max<int64_t>(1, std::max<int64_t>(slots / 8, 1) / std::min<int64_t>(p.n_splits, 8));
  if (const char* e = std::getenv("EVOC_MI355X_SYM_GROUP")) p.group_q = std::max(1, std::atoi(e));
  const int64_t expect = (int64_t)p.stride * 2 * p.J;
  int64_t caps = std::max<int64_t>(pow2ceil(cdiv(4 * expect, p.n_splits)), 128);
  caps = std::max<int64_t>(caps, pow2ceil(cdiv(2LL * K, p.n_splits)));
  p.caps = (int)std::min<int64_t>(caps, MAX_SORT);
  // the candidate side of the last rows holds about all of their keys: 2x headroom
  sp.ccap = (int)std::max<int64_t>(std::max<int64_t>(pow2ceil(2 * expect), pow2ceil(2LL * K)), 256);
  const int64_t G = p.group_q;
  for (int64_t g0 = 0; g0 < nt; g0 += G)
    for (int64_t s = 0; s < p.n_splits; s++)
      for (int64_t Q = g0; Q < std::min(nt, g0 + G); Q++) {
        const int64_t t0 = std::max(Q, s * tpi), t1 = std::min(nt, (s + 1) * tpi);
        if (t0 < t1) sp.work.push_back(make_int4((int)Q, (int)t0, (int)t1, (int)s));
      }
  return sp;
}

// The symmetric pass is on unless EVOC_MI355X_SYMMETRIC=0 (both passes give
// the same graph; the switch lets tests run either).
bool symmetric_enabled() {
  const char* e = std::getenv("EVOC_MI355X_SYMMETRIC");
  return !(e && e[0] == '0');
}

}  // namespace

namespace {
// the tile shape of one search, on this thread, for the duration of the search
struct ShapeScope {
  explicit ShapeScope(int shape) { set_tile_shape(shape); }
  ~ShapeScope() { set_tile_shape(-1); }
};
}  // namespace

// Query rows whose all-pairs scoring takes about `ms` milliseconds at the
// sustained rate of the element kind's scorer (measured on the MI355X, rounded
// down), in whole query tiles, at least four of them.
int64_t time_slab(int kind, int64_t n, int64_t row_bytes, double ms) {
  double rate, elem;  // pair-element operations per second, bytes per element
  switch (kind) {
    case EVOC_ELEM_BF16: rate = 7e14; elem = 2; break;
    case EVOC_ELEM_F32: rate = 1.1e14; elem = 4; break;
    case EVOC_ELEM_I8: rate = 1.2e15; elem = 1; break;
    default: rate = 2e13; elem = 1.0 / 8; break;  // binary: bits, AND + popcount per 32-bit word
  }
  const int64_t TQ = tile_rows_q(kind);
  const double per_row = 2.0 * (double)n * ((double)row_bytes / elem);
  const int64_t rows = (int64_t)(ms * 1e-3 * rate / per_row);
  return std::max<int64_t>(4 * TQ, rows / TQ * TQ);
}

int64_t exact_slab(int kind, int64_t n, int K, int n_cu, uint64_t avail) {
  const int TQ = tile_rows_q(kind);
  int64_t slab = n;
  for (;;) {
    const Plan p = make_plan(kind, n, slab, K, n_cu);
    if (p.scratch_bytes(slab, cdiv(slab, TQ)) <= avail || slab <= 4 * TQ) return slab;
    slab = std::max<int64_t>(4 * TQ, cdiv(slab / 2, TQ) * TQ);
  }
}

static uint64_t slab_scratch(int kind, int64_t n, int64_t nq, int K, int n_cu) {
  const int TQ = tile_rows_q(kind);
  nq = std::min(nq, n);
  return make_plan(kind, n, nq, K, n_cu).scratch_bytes(nq, cdiv(nq, TQ));
}

uint64_t exact_scratch_bytes(int kind, int64_t n, int64_t row_bytes, int K, int n_cu, bool minimum,
                             bool serving) {
  ShapeScope shape_scope(choose_tile_shape(kind, n, row_bytes));
  const int TQ = tile_rows_q(kind);
  if (minimum) return slab_scratch(kind, n, 4 * (int64_t)TQ, K, n_cu);
  // the plan's bytes per query are not monotone in the slab size (more splits
  // for smaller slabs): take the largest over the slabs the search may use
  uint64_t b = slab_scratch(kind, n, n, K, n_cu);
  if (!serving && symmetric_enabled()) {
    const SymPlan sp = make_sym_plan(kind, n, K, n_cu);
    if (!sp.work.empty()) b = std::max<uint64_t>(b, sp.scratch_bytes(n));
  }
  if (serving)
    for (double ms : {8.0, 2.0}) b = std::max(b, slab_scratch(kind, n, time_slab(kind, n, row_bytes, ms), K, n_cu));
  return b;
}

int exact_search(Ctx* ctx, int kind, const uint8_t* rows, int64_t row_bytes, const uint32_t* pop,
                 int64_t n, int K, bool seed_self, uint64_t* keys, ExactStats* stats) {
  ShapeScope shape_scope(choose_tile_shape(kind, n, row_bytes));
  hipStream_t s = ctx->stream;
  const int TC = tile_rows_c(kind), TQ = tile_rows_q(kind);
  if (K < 1 || K > MAX_SORT / 2 || K > n)
    return ctx->fail(EVOC_ERR_UNSUPPORTED, "exact search needs 1 <= K <= min(n, 1024)");
  // largest slab of queries whose scratch fits the rest of the budget (the
  // caller has allocated its outputs already)
  const uint64_t avail = ctx->budget > ctx->in_use ? ctx->budget - ctx->in_use : 0;
  const int64_t mem_slab = exact_slab(kind, n, K, ctx->n_cu, avail);
  const int64_t n_ctiles = cdiv(n, TC);
  // the symmetric pass when the whole matrix fits one slab with its buffers
  SymPlan sp;
  bool sym = false;
  if (!ctx->serving && symmetric_enabled()) {
    sp = make_sym_plan(kind, n, K, ctx->n_cu);
    sym = !sp.work.empty() && sp.scratch_bytes(n) <= avail;
  }
  hipEvent_t e0 = ctx->ev_score0, e1 = ctx->ev_score1;
  int rc = EVOC_OK;
  int64_t nq = 0;
  for (int64_t q0 = 0; q0 < n && rc == EVOC_OK; q0 += nq) {
    int64_t slab = sym ? n : mem_slab;
    if (ctx->serving) {
      // about 8 ms of scoring per slab, 2 ms while other requests wait or run
      const bool busy = ctx->contention && ctx->contention->load(std::memory_order_relaxed) > 0;
      slab = std::min(slab, time_slab(kind, n, row_bytes, busy ? 2.0 : 8.0));
    }
    // a slab whose plan needs more than what is left is cut further
    const uint64_t left = ctx->budget > ctx->in_use ? ctx->budget - ctx->in_use : 0;
    while (!sym && slab > 4 * TQ && slab_scratch(kind, n, std::min(slab, n - q0), K, ctx->n_cu) > left)
      slab = std::max<int64_t>(4 * TQ, cdiv(slab / 2, TQ) * TQ);
    nq = std::min(slab, n - q0);
    const int64_t nqb = cdiv(nq, TQ);
    const Plan p = sym ? sp.p : make_plan(kind, n, nq, K, ctx->n_cu);
    DevBuf gmax, tau, tau_safe, bufs, counts, ovf, tkey, list, cbufs, ccounts, work;
    const bool estimate = p.pass1 && p.J < K;  // pass 2 tests an estimated bound
    if (p.pass1) {
      if ((rc = gmax.alloc(ctx, (size_t)nq * p.ng * 4, "group maxima"))) break;
      if ((rc = tau.alloc(ctx, (size_t)nq * 4, "score bounds"))) break;
      if (estimate && (rc = tau_safe.alloc(ctx, (size_t)nq * 4, "guaranteed score bounds"))) break;
    }
    if ((rc = bufs.alloc(ctx, (size_t)p.n_splits * nq * p.caps * 8, "candidate buffers"))) break;
    if ((rc = counts.alloc(ctx, (size_t)p.n_splits * nq * 4, "candidate counts"))) break;
    if ((rc = ovf.alloc(ctx, (size_t)nq * 4, "overflow flags"))) break;
    if ((rc = tkey.alloc(ctx, (size_t)nq * 8, "key bounds"))) break;
    if ((rc = list.alloc(ctx, (size_t)(nqb + nq) * 4, "re-score lists"))) break;
    if (sym) {
      if ((rc = cbufs.alloc(ctx, (size_t)n * sp.ccap * 8, "candidate-side buffers"))) break;
      if ((rc = ccounts.alloc(ctx, (size_t)n * 4, "candidate-side counts"))) break;
      if ((rc = work.alloc(ctx, sp.work.size() * sizeof(int4), "work items"))) break;
    }

    TileArgs a = {};
    a.q_rows = rows + q0 * row_bytes;
    a.c_rows = rows;
    a.row_bytes = row_bytes;
    a.n_q = (int32_t)nq;
    a.q_offset = (int32_t)q0;
    a.row_pop = pop;
    a.skip_self = seed_self ? 0 : 1;
    hipError_t e = hipSuccess;
    if (p.pass1) {
      TileArgs a1 = a;
      a1.n_c = p.n_sample;
      a1.c_stride = p.stride;
      a1.n_ctiles = (int32_t)cdiv(p.n_sample, TC);
      a1.tiles_per_split = p.p1_tiles_per_split;
      a1.gmax = (float*)gmax.ptr;
      a1.ng = p.ng;
      a1.group_q = p.p1_group_q;
      e = launch_pass1(kind, a1, (int)nqb, p.p1_splits, s);
      if (e == hipSuccess) e = launch_tau_select((const float*)gmax.ptr, (int)nq, p.ng, p.J, (float*)tau.ptr, s);
      if (e != hipSuccess) {
        rc = ctx->hip_fail(e, "threshold pass");
        break;
      }
    }
    TileArgs a2 = a;
    a2.n_c = (int32_t)n;
    a2.c_stride = 1;
    a2.n_ctiles = (int32_t)n_ctiles;
    a2.tiles_per_split = p.tiles_per_split;
    a2.tau = p.pass1 ? (const float*)tau.ptr : nullptr;
    a2.bufs = (uint64_t*)bufs.ptr;
    a2.counts = (int32_t*)counts.ptr;
    a2.caps = p.caps;
    a2.group_q = p.group_q;
    SelectArgs sa = {};
    sa.bufs = (const uint64_t*)bufs.ptr;
    sa.counts = (const int32_t*)counts.ptr;
    sa.n_q = (int32_t)nq;
    sa.n_splits = p.n_splits;
    sa.caps = p.caps;
    sa.K = K;
    sa.q_offset = (int32_t)q0;
    sa.out_keys = keys + q0 * K;
    sa.overflow = (int32_t*)ovf.ptr;
    sa.tau_key_out = (uint64_t*)tkey.ptr;
    // with an estimated bound, a query left with fewer than K keys (or every
    // candidate it has, if fewer) is scored again against the guaranteed one
    if (estimate) sa.min_count = (int32_t)std::min<int64_t>(K, n - (seed_self ? 0 : 1));

    if (sym) {
      // query blocks skip the splits before their own tile: those counts stay 0
      e = upload_sync(work.ptr, sp.work.data(), sp.work.size() * sizeof(int4), s);
      if (e == hipSuccess) e = hipMemsetAsync(counts.ptr, 0, (size_t)p.n_splits * nq * 4, s);
      if (e == hipSuccess) e = hipMemsetAsync(ccounts.ptr, 0, (size_t)n * 4, s);
      if (e != hipSuccess) {
        rc = ctx->hip_fail(e, "scoring pass");
        break;
      }
      a2.work = (const int4*)work.ptr;
      a2.n_work = (int32_t)sp.work.size();
      // keys expected per 32 x 32 block (both sides): about stride * J per row
      // spread over n candidates, 2 x 1024 tests per block; the early-out test
      // pays where fewer than about one block in four holds a key (e^-1.5)
      a2.early_out = p.pass1 && 2048.0 * p.stride * p.J / (double)n < 1.5 ? 1 : 0;
      // EVOC_MI355X_EARLY_OUT=0/1 forces it (results do not depend on it)
      if (const char* e = std::getenv("EVOC_MI355X_EARLY_OUT")) a2.early_out = e[0] == '1' ? 1 : 0;
      a2.cbufs = (uint64_t*)cbufs.ptr;
      a2.ccounts = (int32_t*)ccounts.ptr;
      a2.ccap = sp.ccap;
      sa.cbufs = (const uint64_t*)cbufs.ptr;
      sa.ccounts = (const int32_t*)ccounts.ptr;
      sa.ccap = sp.ccap;
    }
    (void)hipEventRecord(e0, s);
    e = sym ? launch_pass2_sym(kind, a2, s) : launch_pass2(kind, a2, (int)nqb, p.n_splits, s);
    (void)hipEventRecord(e1, s);
    if (e == hipSuccess) e = launch_select(sa, s);
    if (e != hipSuccess) {
      rc = ctx->hip_fail(e, "scoring pass");
      break;
    }
    std::vector<int32_t> flags((size_t)nq);
    for (int round = 0;; round++) {
      e = hipMemcpyAsync(flags.data(), ovf.ptr, (size_t)nq * 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        rc = ctx->hip_fail(e, "scoring pass");
        break;
      }
      if (round == 0) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (stats) {
          stats->pass2_ms += ms;
          stats->score_pairs += sym ? n * (n + 1) / 2 : nq * n;
        }
      }
      std::vector<int32_t> blocks;
      for (int64_t b = 0; b < nqb; b++) {
        bool any = false;
        for (int64_t q = b * TQ; q < std::min<int64_t>(nq, (b + 1) * TQ); q++) any |= flags[(size_t)q] != 0;
        if (any) blocks.push_back((int32_t)b);
      }
      if (blocks.empty()) break;
      if (round >= MAX_ROUNDS) {
        rc = ctx->fail(EVOC_ERR_UNSUPPORTED,
                       "exact search: candidate lists still overflow after 8 rounds (massive exact ties)");
        break;
      }
      if (stats) stats->extra_rounds++;
      std::vector<int32_t> qlist;
      for (int32_t b : blocks)
        for (int64_t q = (int64_t)b * TQ; q < std::min<int64_t>(nq, ((int64_t)b + 1) * TQ); q++)
          qlist.push_back((int32_t)q);
      int32_t* d_blocks = (int32_t*)list.ptr;
      int32_t* d_q = d_blocks + nqb;
      e = upload_sync(d_blocks, blocks.data(), blocks.size() * 4, s);
      if (e == hipSuccess) e = upload_sync(d_q, qlist.data(), qlist.size() * 4, s);
      // the re-scored queries' candidate-side lists are covered by pass 2 now
      if (e == hipSuccess && sym) e = launch_zero_list((int32_t*)ccounts.ptr, d_q, (int)qlist.size(), s);
      // their guaranteed bounds (the K-th group maximum), for those whose
      // estimated bound missed; every round's bounds are guaranteed ones
      if (e == hipSuccess && estimate)
        e = launch_tau_select((const float*)gmax.ptr, (int)qlist.size(), p.ng, K, (float*)tau_safe.ptr, s, d_q);
      // the round's bounds are guaranteed ones: the K-th kept key of an
      // overflowed query (the K-th best of a subset), the K-th group maximum
      // (a query whose estimated bound missed has no K-th kept key)
      TileArgs ar = a2;
      ar.work = nullptr;
      ar.block_list = d_blocks;
      ar.tau = estimate ? (const float*)tau_safe.ptr : nullptr;
      ar.tau_key = (const uint64_t*)tkey.ptr;
      if (e == hipSuccess) e = launch_pass2(kind, ar, (int)blocks.size(), p.n_splits, s);
      SelectArgs sr = sa;
      sr.min_count = 0;
      sr.query_list = d_q;
      sr.n_list = (int32_t)qlist.size();
      if (e == hipSuccess) e = launch_select(sr, s);
      if (e != hipSuccess) {
        rc = ctx->hip_fail(e, "re-scoring pass");
        break;
      }
      // queries outside the re-scored blocks keep flags == 0 from the last full pass
    }
    if (stats) stats->slabs++;
  }
  return rc;
}

int exact_plan_info(int kind, int64_t n, int64_t row_bytes, int K, int n_cu, bool serving, ExactPlanInfo* out) {
  ShapeScope shape_scope(choose_tile_shape(kind, n, row_bytes));
  const int TQ = tile_rows_q(kind);
  *out = ExactPlanInfo{};
  out->tile_shape = tile_shape(kind);
  if (!serving && symmetric_enabled()) {
    const SymPlan sp = make_sym_plan(kind, n, K, n_cu);
    if (!sp.work.empty()) {
      out->symmetric = 1;
      out->n_splits = sp.p.n_splits;
      out->caps = sp.p.caps;
      out->ccap = sp.ccap;
      out->stride = sp.p.pass1 ? sp.p.stride : 0;
      out->n_sample = sp.p.pass1 ? sp.p.n_sample : 0;
      out->bound_rank = sp.p.pass1 ? sp.p.J : 0;
      out->segments = sp.p.n_splits + 1;
      out->work_items = (int64_t)sp.work.size();
      out->scratch_bytes = sp.scratch_bytes(n);
      return 0;
    }
  }
  const int64_t nq = serving ? std::min<int64_t>(n, time_slab(kind, n, row_bytes, 8.0)) : n;
  const Plan p = make_plan(kind, n, nq, K, n_cu);
  out->n_splits = p.n_splits;
  out->caps = p.caps;
  out->stride = p.pass1 ? p.stride : 0;
  out->n_sample = p.pass1 ? p.n_sample : 0;
  out->bound_rank = p.pass1 ? p.J : 0;
  out->segments = p.n_splits;
  out->work_items = cdiv(nq, TQ) * p.n_splits;
  out->scratch_bytes = p.scratch_bytes(nq, cdiv(nq, TQ));
  return 0;
}

SlabPlan exact_slab_plan(int kind, int64_t n, int K, int n_cu) {
  // the batched search tests the guaranteed bound (the K-th group maximum):
  // its buffers are sized for that bound's survivors
  const Plan p = make_plan(kind, n, n, K, n_cu, false, true);
  SlabPlan o;
  o.n_splits = p.n_splits;
  o.tiles_per_split = p.tiles_per_split;
  o.caps = p.caps;
  o.pass1 = p.pass1;
  o.stride = p.stride;
  o.n_sample = p.n_sample;
  o.ng = p.ng;
  o.p1_tiles_per_split = p.p1_tiles_per_split;
  o.p1_splits = p.p1_splits;
  return o;
}
