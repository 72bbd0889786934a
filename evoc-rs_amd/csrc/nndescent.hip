// NN-descent on the device (gfx950).
//
// The reference's loop (src/nndescent/float.rs:775-879, int8.rs:575-678,
// uint8.rs:563-666) re-expressed for one GPU:
//   lists      per-vertex neighbour lists of Kg 64-bit keys kept sorted best
//              first (the reference's bounded max-heaps, src/heap.rs:27-200)
//   candidates forward edges + a CSR of reverse edges; each vertex keeps the
//              mc smallest priorities per new/old class (src/heap.rs:203-343)
//   local join per vertex, MFMA Gram of its new list against new + old lists
//              (float.rs:517-669); a pair is offered to a vertex when it beats
//              that vertex's current worst entry
//   apply      offers land in bounded per-vertex buffers; one wave per vertex
//              merges buffer and list (sort, de-duplicate, keep Kg)
//              (src/heap.rs:425-502)
//   forest     random-projection trees split level by level (float.rs:13-177)
//              whose leaves seed the lists (float.rs:179-348)
// Priorities and pivots come from a counter-based hash (splitmix64 finalizer,
// Steele, Lea & Flood 2014) instead of the sequential tau RNG, so every vertex
// and every split can run in parallel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evoc_device.h"
#include "evoc_kernels.h"
#include "mfma_traits.h"
#include "nndescent.h"

namespace evoc {

// ---- graph keys: [ord(score):32][0x3fffffff - idx:30][from_list:1][new:1] ----
__device__ __forceinline__ uint64_t gkey(float score, int32_t idx, uint32_t from_list, uint32_t flag) {
  return ((uint64_t)ord_f32(score) << 32) | ((uint64_t)(0x3fffffffu - (uint32_t)idx) << 2) |
         ((uint64_t)from_list << 1) | (uint64_t)flag;
}
__device__ __forceinline__ int32_t gkey_idx(uint64_t k) {
  return (int32_t)(0x3fffffffu - (uint32_t)((k >> 2) & 0x3fffffffu));
}
__device__ __forceinline__ float gkey_score(uint64_t k) { return unord_f32((uint32_t)(k >> 32)); }

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// List initialisation: self seeded with the reference's -1.0 (score 1.0,
// float.rs:210-214), the rest empty.
// ---------------------------------------------------------------------------
__global__ void nnd_init_lists_kernel(uint64_t* __restrict__ keys, float* __restrict__ thr,
                                      int32_t* __restrict__ umark, int n, int Kg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * Kg) return;
  const int v = (int)(i / Kg), s = (int)(i % Kg);
  keys[i] = s == 0 ? gkey(1.0f, v, 0, 1) : EVOC_KEY_NONE;
  if (s == 0) {
    thr[v] = Kg == 1 ? 1.0f : -INFINITY;
    umark[v] = 0;
  }
}

// ---------------------------------------------------------------------------
// Local join.  One wave per work item: a "new" list N (<= 64 vertices) and an
// "old" list O; C = N followed by O (<= 128).  Scores of N x C on MFMA from
// rows gathered straight from HBM, then every pair (a < c inside N, all of
// N x O) is offered to both endpoints.
// ---------------------------------------------------------------------------
struct JoinShared {
  int32_t nidx[64];
  int32_t cidx[128];
  float cthr[128];
};

// An offer of cand to target: the 64-bit maximum into two hashed slots of the
// target's buffer.  A slot's final content is the best key hashed to it, so the
// merged lists do not depend on the order in which waves arrive; an offer is
// lost only when a better one holds both of its slots.  The atomics return
// nothing unless a.dropped asks for the count of lost offers: a returning
// atomic makes the wave wait for its round trip before the next offer.
__device__ __forceinline__ void offer(const JoinArgs& a, int32_t target, int32_t cand, float s) {
  const unsigned long long key = gkey(s, cand, 0, 1);
  const uint64_t h = mix64((((uint64_t)(uint32_t)target << 32) | (uint32_t)cand) ^
                           ((uint64_t)a.salt * 0xd1b54a32d192ed03ull));
  const uint32_t s1 = (uint32_t)(((h & 0xffffffffull) * (uint64_t)a.ucap) >> 32);
  const uint32_t s2 = (uint32_t)(((h >> 32) * (uint64_t)a.ucap) >> 32);
  unsigned long long* b = (unsigned long long*)(a.ubuf + (int64_t)target * a.ucap);
  if (a.dropped) {
    const unsigned long long o1 = atomicMax(b + s1, key);
    const unsigned long long o2 = s2 != s1 ? atomicMax(b + s2, key) : o1;
    if (o1 > key && o2 > key) atomicAdd(a.dropped, 1u);
  } else {
    (void)atomicMax(b + s1, key);
    if (s2 != s1) (void)atomicMax(b + s2, key);
  }
  a.umark[target] = 1;
}

// gather the item's lists into LDS; returns (nN, nC) through refs
__device__ __forceinline__ void load_item(const JoinArgs& a, int item, JoinShared& sh, int& nN,
                                          int& nC) {
  const int lane = threadIdx.x & 63;
  int32_t nv = -1, ov = -1;
  if (a.mode == 0) {
    // iteration: new/old candidate lists of vertex a.v0 + item (mc <= 64 each)
    const int v = a.v0 + item;
    if (lane < a.mc) {
      nv = a.newc[(int64_t)v * a.mc + lane];
      ov = a.oldc[(int64_t)v * a.mc + lane];
    }
  } else {
    // forest leaf: members perm[start .. start + len), len <= 64 (only the
    // leaves of trees [tree_lo, tree_hi) when a tree filter is given)
    const int2 lf = a.leaves[item];
    const int tr = a.tree_n > 0 ? lf.x / a.tree_n : 0;
    if (lane < lf.y && (a.tree_n == 0 || (tr >= a.tree_lo && tr < a.tree_hi))) nv = a.perm[lf.x + lane];
  }
  const uint64_t mn = __ballot(nv >= 0), mo = __ballot(ov >= 0);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  nN = __popcll(mn);
  nC = nN + __popcll(mo);
  if (nv >= 0) {
    const int p = __popcll(mn & below);
    sh.nidx[p] = nv;
    sh.cidx[p] = nv;
    sh.cthr[p] = a.thr[nv];
  }
  if (ov >= 0) {
    const int p = nN + __popcll(mo & below);
    sh.cidx[p] = ov;
    sh.cthr[p] = a.thr[ov];
  }
}

template <int KIND>
__global__ __launch_bounds__(256, 2) void nnd_join_mfma_kernel(JoinArgs a) {
  using M = Mfma<KIND>;
  __shared__ JoinShared shm[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int item = blockIdx.x * 4 + w;
  if (item >= a.n_items) return;
  JoinShared& sh = shm[w];
  int nN, nC;
  load_item(a, item, sh, nN, nC);
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  if (nN == 0 || nC < 2) return;
  // the item's algorithmic traffic: its nC candidate rows (the new rows are
  // the first nN of them), each read once
  if (a.bytes && lane == 0)
    atomicAdd(a.bytes + NND_SLOT_U64 * (item % NND_COUNTER_SLOTS),
              (unsigned long long)nC * (unsigned long long)a.row_bytes);
  const int nbm = (nN + 31) >> 5, nbn = (nC + 31) >> 5;
  // Fragment row pointers of C only: C starts with N (cidx[r] = nidx[r] for
  // r < nN), so the A operand's block bm is C's block bm.  Rows of that block
  // at or past nN are C's old rows (or row 0 past nC) where N would have had
  // row 0; every such row is masked below, so the scores that are kept are the
  // same products in the same K order.  Invalid rows read a valid row.
  const uint8_t* brow[4];
#pragma unroll
  for (int bn = 0; bn < 4; bn++) {
    const int c = bn * 32 + (lane & 31);
    brow[bn] = a.rows + (int64_t)sh.cidx[c < nC ? c : 0] * a.row_bytes + (lane >> 5) * M::LANE_BYTES;
  }
  typename M::acc_t acc[2][4];
#pragma unroll
  for (int bm = 0; bm < 2; bm++)
#pragma unroll
    for (int bn = 0; bn < 4; bn++) acc[bm][bn] = M::zero();
  // Fragments straight from HBM through a ring of D = 4 register sets: the
  // loads of the D - 1 steps after the current one are in flight while its
  // MFMAs run.  Every load is unconditional (the blocks past nbn read row
  // cidx[0], the steps past the end re-read the last step into sets nothing
  // reads), so
  // the compiler's vmcnt waits leave the D - 1 newer steps in flight; a load
  // under a branch would make it drain them all at every step.  Only the
  // nbm x nbn blocks that hold pairs are multiplied (wave-uniform tests; the
  // blocks past them are never read below).  The step count is a multiple of
  // D (rows are padded to 128 bytes), so the sets rotate with static indices.
  if constexpr (KIND == EVOC_ELEM_F32) {
    // FP32 rows: 16 bytes per lane cover four K steps of 32x32x2 (32 bytes of
    // a row: lane half h holds elements 8S + 4h .. 8S + 4h + 3 of super-step
    // S), so a gather is one 16-byte load per lane instead of four 4-byte ones.
    // K step j of S pairs elements 8S + j and 8S + 4 + j, the same in both
    // operands: every product of the row pair is summed once (in another
    // order than element pairs 2i, 2i + 1).  Ring of 4 super-steps, loads as
    // above (unconditional, clamped, in step order).
    constexpr int DS = 4;
    static_assert(128 % (DS * 32) == 0, "128-byte rows hold whole rounds of the ring");
    const int nsup = (int)(a.row_bytes / 32);
    const uint8_t* crow[4];
#pragma unroll
    for (int bn = 0; bn < 4; bn++) crow[bn] = brow[bn] - (lane >> 5) * M::LANE_BYTES + (lane >> 5) * 16;
    float4 fq[DS][4];
    auto load_sup = [&](float4 (&rb)[4], int S) __attribute__((always_inline)) {
      const int off = min(S, nsup - 1) * 32;
#pragma unroll
      for (int bn = 0; bn < 4; bn++) rb[bn] = *(const float4*)(crow[bn] + off);
    };
    auto comp = [](const float4& v, int j) __attribute__((always_inline)) {
      return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
    };
    auto mma_sup = [&](const float4 (&rb)[4]) __attribute__((always_inline)) {
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int bm = 0; bm < 2; bm++)
#pragma unroll
          for (int bn = 0; bn < 4; bn++)
            if (bm < nbm && bn < nbn) M::mma(acc[bm][bn], comp(rb[bm], j), comp(rb[bn], j));
    };
    static_for<DS - 1>([&](auto sc) __attribute__((always_inline)) {
      load_sup(fq[decltype(sc)::value], decltype(sc)::value);
      __builtin_amdgcn_sched_barrier(0);
    });
    for (int S = 0; S < nsup; S += DS) {
      static_for<DS>([&](auto ic) __attribute__((always_inline)) {
        constexpr int i = decltype(ic)::value;
        load_sup(fq[(i + DS - 1) % DS], S + i + DS - 1);
        __builtin_amdgcn_sched_barrier(0);
        mma_sup(fq[i]);
      });
    }
  } else {
    constexpr int D = 4;
    static_assert(128 % (D * M::KSTEP_BYTES) == 0, "128-byte rows hold whole rounds of the ring");
    const int ksteps = (int)(a.row_bytes / M::KSTEP_BYTES);
    typename M::frag_t fb[D][4];
    auto load_step = [&](typename M::frag_t (&rb)[4], int ks) __attribute__((always_inline)) {
      const int off = min(ks, ksteps - 1) * M::KSTEP_BYTES;
#pragma unroll
      for (int bn = 0; bn < 4; bn++) rb[bn] = M::load(brow[bn] + off);
    };
    auto mma_step = [&](const typename M::frag_t (&rb)[4]) __attribute__((always_inline)) {
#pragma unroll
      for (int bm = 0; bm < 2; bm++)
#pragma unroll
        for (int bn = 0; bn < 4; bn++)
          if (bm < nbm && bn < nbn) M::mma(acc[bm][bn], rb[bm], rb[bn]);
    };
    // (the scheduling barriers keep the steps' loads in step order, in the
    // first D - 1 steps as in the loop: the vmcnt each step waits for is then
    // the same on entry as around the loop)
    static_for<D - 1>([&](auto sc) __attribute__((always_inline)) {
      load_step(fb[decltype(sc)::value], decltype(sc)::value);
      __builtin_amdgcn_sched_barrier(0);
    });
    for (int ks = 0; ks < ksteps; ks += D) {
      static_for<D>([&](auto ic) __attribute__((always_inline)) {
        constexpr int i = decltype(ic)::value;
        load_step(fb[(i + D - 1) % D], ks + i + D - 1);
        __builtin_amdgcn_sched_barrier(0);
        mma_step(fb[i]);
      });
    }
  }
  // Offers, one 32 x 32 block at a time (rolled: the block's values are
  // copied out under a uniform branch, so no accumulator is indexed at run time)
#pragma unroll 1
  for (int b = 0; b < 8; b++) {
    const int bm = b >> 2, bn = b & 3;
    if (bm >= nbm || bn >= nbn) continue;
    float v[16];
    static_for<8>([&](auto bc) __attribute__((always_inline)) {
      constexpr int cc = decltype(bc)::value;
      if (b == cc) {
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = M::value(acc[cc >> 2][cc & 3], j);
      }
    });
    const int c = bn * 32 + (lane & 31);
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int r = bm * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
      if (r >= nN || c >= nC) continue;
      if (c < nN && r >= c) continue;  // each new-new pair once
      const int32_t p = sh.nidx[r], q = sh.cidx[c];
      if (p == q) continue;
      const float s = v[j];
      if (s > sh.cthr[r]) offer(a, p, q, s);
      if (s > sh.cthr[c]) offer(a, q, p, s);
    }
  }
}

// Binary rows: Jaccard on the VALU, one wave per item, pairs spread over lanes.
__global__ __launch_bounds__(256) void nnd_join_jaccard_kernel(JoinArgs a) {
  __shared__ JoinShared shm[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int item = blockIdx.x * 4 + w;
  if (item >= a.n_items) return;
  JoinShared& sh = shm[w];
  int nN, nC;
  load_item(a, item, sh, nN, nC);
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  if (a.bytes && lane == 0 && nN > 0 && nC >= 2)
    atomicAdd(a.bytes + NND_SLOT_U64 * (item % NND_COUNTER_SLOTS),
              (unsigned long long)nC * (unsigned long long)a.row_bytes);
  const int words = (int)(a.row_bytes / 4);
  for (int pi = lane; pi < nN * nC; pi += 64) {
    const int r = pi / nC, c = pi % nC;
    if (c < nN && r >= c) continue;
    const int32_t p = sh.nidx[r], q = sh.cidx[c];
    if (p == q) continue;
    const uint32_t* xp = (const uint32_t*)(a.rows + (int64_t)p * a.row_bytes);
    const uint32_t* xq = (const uint32_t*)(a.rows + (int64_t)q * a.row_bytes);
    uint32_t inter = 0;
    for (int t = 0; t < words; t++) inter += __popc(xp[t] & xq[t]);
    const uint32_t uni = a.row_pop[p] + a.row_pop[q] - inter;
    const float s = uni > 0 ? __fdiv_rn((float)inter, (float)uni) : 0.0f;
    if (s > a.thr[p]) offer(a, p, q, s);
    if (s > a.thr[q]) offer(a, q, p, s);
  }
}

// ---------------------------------------------------------------------------
// Merge offers into the lists: one wave per marked vertex, list (<= 64) +
// offer slots (NND_OFFER_CAP) sorted best first, duplicates of one neighbour
// collapse onto the copy already in the list (flagged_heap_push rejects a
// present index, src/heap.rs:156-160), the first Kg survive.  New entries
// carry flag 1.  The slots are emptied for the next round of offers.
// The list and the non-empty slots are first compacted through the wave's LDS
// stash, so the wave sort runs over as many registers as they fill (one
// register per 64 keys), not over every slot.
// ---------------------------------------------------------------------------
template <int E>
__device__ __forceinline__ void merge_sorted(const uint64_t* st, int cnt, uint64_t* __restrict__ keys,
                                             float* __restrict__ thr, int v, int Kg, int& inserted) {
  const int lane = threadIdx.x & 63;
  uint64_t x[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * 64 + lane;
    x[e] = i < cnt ? st[i] : EVOC_KEY_NONE;
  }
  wave_sort_desc<E>(x);
  // keep[i]: non-empty and not a repeat of element i - 1 (same score and index)
  bool keep[E];
  uint64_t last_prev = 0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint64_t up = (uint64_t)__shfl_up((long long)x[e], 1, 64);
    const uint64_t prev = lane == 0 ? last_prev : up;
    keep[e] = x[e] != EVOC_KEY_NONE && (e == 0 && lane == 0 ? true : (x[e] >> 2) != (prev >> 2));
    last_prev = (uint64_t)__shfl((long long)x[e], 63, 64);
  }
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int base = 0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint64_t m = __ballot(keep[e]);
    const int pos = base + __popcll(m & below);
    if (keep[e] && pos < Kg) {
      keys[(int64_t)v * Kg + pos] = x[e] & ~2ull;
      if (!(x[e] & 2ull)) inserted++;
      if (pos == Kg - 1) thr[v] = gkey_score(x[e]);
    }
    base += __popcll(m);
  }
  // slots beyond the kept entries stay empty
  for (int pos = base + lane; pos < Kg; pos += 64) keys[(int64_t)v * Kg + pos] = EVOC_KEY_NONE;
  if (base < Kg && lane == 0) thr[v] = -INFINITY;
}

__global__ __launch_bounds__(256) void nnd_merge_kernel(uint64_t* __restrict__ keys,
                                                        float* __restrict__ thr,
                                                        int32_t* __restrict__ umark,
                                                        uint64_t* __restrict__ ubuf, int ucap,
                                                        int n, int Kg, int v0, int nv,
                                                        uint32_t* __restrict__ changed) {
  constexpr int SLOTS = NND_OFFER_CAP / 64;  // registers of offer slots
  static_assert(NND_OFFER_CAP % 64 == 0, "offer buffers fill whole registers of the wave sort");
  __shared__ uint64_t stash[4][64 * (1 + SLOTS)];
  const int lane = threadIdx.x & 63;
  const int vi = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (vi >= nv) return;
  const int v = v0 + vi;
  if (umark[v] == 0) return;
  uint64_t* st = stash[threadIdx.x >> 6];
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the list (from_list bit set, so a listed neighbour sorts before its offers)
  const uint64_t lk = lane < Kg ? keys[(int64_t)v * Kg + lane] : EVOC_KEY_NONE;
  const uint64_t l = lk | (lk ? 2ull : 0ull);
  const uint64_t ml = __ballot(l != EVOC_KEY_NONE);
  if (l != EVOC_KEY_NONE) st[__popcll(ml & below)] = l;
  int cnt = __popcll(ml);
  // then the non-empty offer slots, each emptied for the next round
#pragma unroll
  for (int e = 0; e < SLOTS; e++) {
    const int u = e * 64 + lane;
    const uint64_t o = u < ucap ? ubuf[(int64_t)v * ucap + u] : EVOC_KEY_NONE;
    const uint64_t mo = __ballot(o != EVOC_KEY_NONE);
    if (o != EVOC_KEY_NONE) {
      st[cnt + __popcll(mo & below)] = o;
      ubuf[(int64_t)v * ucap + u] = EVOC_KEY_NONE;
    }
    cnt += __popcll(mo);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int inserted = 0;
  if (cnt <= 64) merge_sorted<1>(st, cnt, keys, thr, v, Kg, inserted);
  else if (cnt <= 128) merge_sorted<2>(st, cnt, keys, thr, v, Kg, inserted);
  else if (cnt <= 256) merge_sorted<4>(st, cnt, keys, thr, v, Kg, inserted);
  else merge_sorted<1 + SLOTS>(st, cnt, keys, thr, v, Kg, inserted);
  inserted = wave_sum_i32(inserted);
  if (lane == 0) {
    // (slots NND_SLOT_U32 apart, see NND_COUNTER_SLOTS)
    if (inserted) atomicAdd(changed + NND_SLOT_U32 * (vi % NND_COUNTER_SLOTS), (uint32_t)inserted);
    umark[v] = 0;
  }
}

// ---------------------------------------------------------------------------
// Candidate sampling (src/heap.rs:203-343).
// ---------------------------------------------------------------------------
__global__ void nnd_rev_count_kernel(const uint64_t* __restrict__ keys, int n, int Kg,
                                     int32_t* __restrict__ revcnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * Kg) return;
  const uint64_t k = keys[i];
  if (k == EVOC_KEY_NONE) return;
  const int v = (int)(i / Kg), t = gkey_idx(k);
  if (t != v) atomicAdd(&revcnt[t], 1);
}

__global__ void nnd_rev_fill_kernel(const uint64_t* __restrict__ keys, int n, int Kg,
                                    const int32_t* __restrict__ revoff, int32_t* __restrict__ revcur,
                                    int32_t* __restrict__ rev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * Kg) return;
  const uint64_t k = keys[i];
  if (k == EVOC_KEY_NONE) return;
  const int v = (int)(i / Kg), t = gkey_idx(k);
  if (t == v) return;
  const int pos = revoff[t] + atomicAdd(&revcur[t], 1);
  rev[pos] = v | (int32_t)((uint32_t)(k & 1ull) << 31);
}

// priority of edge {a, b} in iteration `it`: the same for both directions
__device__ __forceinline__ uint32_t edge_priority(int a, int b, uint64_t seed, int it) {
  const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
  return (uint32_t)(mix64(((uint64_t)lo << 32 | hi) ^ (seed + 0x632be59bd9b4e019ull * (uint64_t)(it + 1))) >> 32);
}

// sort key: larger = smaller priority, then smaller candidate
__device__ __forceinline__ uint64_t cand_key(uint32_t prio, int32_t c) {
  return ((uint64_t)(~prio) << 32) | (uint64_t)(0xffffffffu - (uint32_t)c);
}
__device__ __forceinline__ int32_t cand_of(uint64_t k) { return (int32_t)(0xffffffffu - (uint32_t)k); }

// keep the mc best distinct keys of (best ∪ chunk), one key per lane each.
// The kept keys (non-empty, first of a run of equal keys) are compacted in
// sorted order through the wave's 64-key LDS stash: kept key i goes to slot i,
// and lane L then reads slot L.
__device__ __forceinline__ uint64_t keep_best(uint64_t best, uint64_t chunk, int mc, uint64_t* stash) {
  uint64_t x[2] = {best, chunk};
  wave_sort_desc<2>(x);
  const int lane = threadIdx.x & 63;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t last_prev = 0;
  int base = 0;
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const uint64_t up = (uint64_t)__shfl_up((long long)x[e], 1, 64);
    const uint64_t prev = lane == 0 ? last_prev : up;
    const bool keep = x[e] != EVOC_KEY_NONE && ((e == 0 && lane == 0) || x[e] != prev);
    last_prev = (uint64_t)__shfl((long long)x[e], 63, 64);
    const uint64_t m = __ballot(keep);
    const int pos = base + __popcll(m & below);
    if (keep && pos < mc) stash[pos] = x[e];
    base += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint64_t out = lane < min(base, mc) ? stash[lane] : EVOC_KEY_NONE;
  // the next call's writes must not pass this read
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return out;
}

__global__ __launch_bounds__(256) void nnd_build_candidates_kernel(
    uint64_t* __restrict__ keys, int n, int Kg, const int32_t* __restrict__ revoff,
    const int32_t* __restrict__ rev, int mc, uint64_t seed, int it, int32_t* __restrict__ newc,
    int32_t* __restrict__ oldc) {
  __shared__ uint64_t stash_all[4][64];
  const int lane = threadIdx.x & 63;
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= n) return;
  uint64_t* stash = stash_all[threadIdx.x >> 6];
  const uint64_t fk = lane < Kg ? keys[(int64_t)v * Kg + lane] : EVOC_KEY_NONE;
  const int32_t fc = fk != EVOC_KEY_NONE ? gkey_idx(fk) : -1;
  const int fflag = (int)(fk & 1ull);
  const int r0 = revoff[v], r1 = revoff[v + 1];
  uint64_t best[2];
  uint64_t fwd_new = EVOC_KEY_NONE;  // this lane's forward entry as a new-class candidate key
#pragma unroll
  for (int cls = 0; cls < 2; cls++) {  // cls 1 = new, 0 = old
    // the forward entries (distinct neighbours, one per lane) go straight into
    // the first selection with the first chunk of reverse edges: the mc best
    // of (the mc best of F) ∪ C are the mc best of F ∪ C, so no selection of
    // F alone is needed; one selection runs even when there is no reverse edge
    uint64_t b = (fc >= 0 && fc != v && fflag == cls) ? cand_key(edge_priority(v, fc, seed, it), fc)
                                                      : EVOC_KEY_NONE;
    if (cls == 1) fwd_new = b;
    int base = r0;
    do {
      const int i = base + lane;
      uint64_t rk = EVOC_KEY_NONE;
      if (i < r1) {
        const int32_t e = rev[i];
        const int c = e & 0x7fffffff, f = (int)((uint32_t)e >> 31);
        if (f == cls) rk = cand_key(edge_priority(v, c, seed, it), c);
      }
      b = keep_best(b, rk, mc, stash);
      base += 64;
    } while (base < r1);
    best[cls] = b;
  }
  if (lane < mc) {
    newc[(int64_t)v * mc + lane] = best[1] != EVOC_KEY_NONE ? cand_of(best[1]) : -1;
    oldc[(int64_t)v * mc + lane] = best[0] != EVOC_KEY_NONE ? cand_of(best[0]) : -1;
  }
  // clear the flag of forward entries sampled into the new list (heap.rs:319-340).
  // The new list holds the mc largest distinct keys, descending in lanes
  // 0..mc-1 (a candidate's key is the same from either direction), so a
  // forward entry was sampled when its key is at least the list's last one
  // (EVOC_KEY_NONE, below every key, while the list is not full).
  const uint64_t last = (uint64_t)__shfl((long long)best[1], mc - 1, 64);
  const bool sampled = fwd_new != EVOC_KEY_NONE && fwd_new >= last;
  if (lane < Kg && fc >= 0 && fflag && sampled) keys[(int64_t)v * Kg + lane] = fk & ~1ull;
}

// ---------------------------------------------------------------------------
// Random fill (float.rs:350-390): a vertex whose list still has empty slots
// gets the item N = {v}, O = random vertices; the join kernel scores them with
// the same MFMA path as every other pair, so repeated offers de-duplicate.
// ---------------------------------------------------------------------------
__global__ void nnd_random_lists_kernel(const uint64_t* __restrict__ keys, int n, int Kg, int mc,
                                        uint64_t seed, int32_t* __restrict__ newc,
                                        int32_t* __restrict__ oldc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int filled = 0;
  for (int s = 0; s < Kg; s++) filled += keys[(int64_t)v * Kg + s] != EVOC_KEY_NONE;
  int32_t* nc = newc + (int64_t)v * mc;
  int32_t* oc = oldc + (int64_t)v * mc;
  for (int s = 0; s < mc; s++) {
    nc[s] = -1;
    oc[s] = -1;
  }
  if (filled >= Kg || n < 2) return;
  nc[0] = v;
  int j = 0;
  for (int t = 0; t < mc && j < Kg - filled + 1; t++) {
    const int c = (int)(mix64(seed ^ ((uint64_t)(uint32_t)v << 24) ^ (uint64_t)t) % (uint64_t)n);
    if (c != v) oc[j++] = c;
  }
}

// ---------------------------------------------------------------------------
// Level-synchronous random-projection forest (float.rs:13-177 semantics:
// hyperplane between two random members, margin sign, random side when
// |margin| < 1e-8, random halves when one side is empty), all trees of the
// forest split together: one set of launches per level instead of one per
// tree and level.  Tree t's rows occupy positions [t n, t n + n) of the flat
// permutation; random choices hash the tree's seed with the level and the
// position within the tree, so a tree comes out the same alone or with others.
// ---------------------------------------------------------------------------
__global__ void rp_init_perm_kernel(int32_t* __restrict__ perm, int n, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) perm[i] = (int32_t)(i % n);
}

__global__ void rp_seg_fill_kernel(const int4* __restrict__ segs, int nseg, int32_t* __restrict__ seg_of) {
  const int s = blockIdx.x;
  if (s >= nseg) return;
  const int4 g = segs[s];
  for (int i = threadIdx.x; i < g.y; i += blockDim.x) seg_of[g.x + i] = s;
}

// one wave per active segment: hyperplane (float) or masks (binary) into hyper[s]
__global__ __launch_bounds__(256) void rp_hyperplane_kernel(const int4* __restrict__ segs, int nseg,
                                                            const int32_t* __restrict__ perm,
                                                            ForestArgs f, int level) {
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= nseg) return;
  const int4 g = segs[s];
  const int local = g.x - g.z * f.n;  // start within the tree
  const uint64_t h = mix64(f.seeds[g.z] ^ ((uint64_t)level << 40) ^ (uint64_t)(uint32_t)local);
  const int li = (int)(h % (uint64_t)g.y);
  int ri = (int)((h >> 32) % (uint64_t)g.y);
  ri += (li == ri);
  ri %= g.y;
  const int L = perm[g.x + li], R = perm[g.x + ri];
  float* hv = f.hyper + (int64_t)s * f.hyper_ld;
  if (f.kind == EVOC_ELEM_U8BITS) {
    const uint8_t* xl = f.rows + (int64_t)L * f.row_bytes;
    const uint8_t* xr = f.rows + (int64_t)R * f.row_bytes;
    uint8_t* pm = (uint8_t*)hv;  // positive mask then negative mask
    for (int i = lane; i < f.row_bytes; i += 64) {
      const uint8_t xo = xl[i] ^ xr[i];
      pm[i] = xo & xl[i];
      pm[f.row_bytes + i] = xo & xr[i];
    }
    return;
  }
  // per-row L2 norms so the int8 split normalises both pivots (int8.rs:43-76);
  // f32 rows are unit already and their norms are ~1
  float nl = 0.0f, nr = 0.0f;
  for (int i = lane; i < f.dim; i += 64) {
    const float a = f.kind == EVOC_ELEM_I8 ? (float)((const int8_t*)(f.rows + (int64_t)L * f.row_bytes))[i]
                                           : ((const float*)(f.frows + (int64_t)L * f.frow_ld))[i];
    const float b = f.kind == EVOC_ELEM_I8 ? (float)((const int8_t*)(f.rows + (int64_t)R * f.row_bytes))[i]
                                           : ((const float*)(f.frows + (int64_t)R * f.frow_ld))[i];
    nl = fmaf(a, a, nl);
    nr = fmaf(b, b, nr);
  }
  nl = sqrtf(wave_sum_f32(nl));
  nr = sqrtf(wave_sum_f32(nr));
  float hn = 0.0f;
  for (int i = lane; i < f.dim; i += 64) {
    const float a = f.kind == EVOC_ELEM_I8 ? (float)((const int8_t*)(f.rows + (int64_t)L * f.row_bytes))[i]
                                           : ((const float*)(f.frows + (int64_t)L * f.frow_ld))[i];
    const float b = f.kind == EVOC_ELEM_I8 ? (float)((const int8_t*)(f.rows + (int64_t)R * f.row_bytes))[i]
                                           : ((const float*)(f.frows + (int64_t)R * f.frow_ld))[i];
    const float d = (nl > 0.0f ? a / nl : 0.0f) - (nr > 0.0f ? b / nr : 0.0f);
    hv[i] = d;
    hn = fmaf(d, d, hn);
  }
  hn = sqrtf(wave_sum_f32(hn));
  if (hn < 1e-8f) hn = 1.0f;
  for (int i = lane; i < f.dim; i += 64) hv[i] = hv[i] / hn;
}

// one wave per position of an active segment (over all trees): side (0 left, 1 right)
__global__ __launch_bounds__(256) void rp_side_kernel(const int32_t* __restrict__ perm,
                                                      const int32_t* __restrict__ seg_of,
                                                      const int4* __restrict__ segs, int64_t n_pos,
                                                      ForestArgs f, int level, int8_t* __restrict__ side) {
  const int lane = threadIdx.x & 63;
  const int64_t pos = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pos >= n_pos) return;
  const int s = seg_of[pos];
  if (s < 0) return;
  const int p = perm[pos];
  const float* hv = f.hyper + (int64_t)s * f.hyper_ld;
  float m = 0.0f;
  if (f.kind == EVOC_ELEM_U8BITS) {
    const uint8_t* pm = (const uint8_t*)hv;
    const uint8_t* x = f.rows + (int64_t)p * f.row_bytes;
    int c = 0;
    for (int i = lane; i < f.row_bytes; i += 64) c += __popc(pm[i] & x[i]) - __popc(pm[f.row_bytes + i] & x[i]);
    m = (float)wave_sum_i32(c);
  } else if (f.kind == EVOC_ELEM_I8) {
    const int8_t* x = (const int8_t*)(f.rows + (int64_t)p * f.row_bytes);
    for (int i = lane; i < f.dim; i += 64) m = fmaf(hv[i], (float)x[i], m);
    m = wave_sum_f32(m);
  } else if (f.kind == EVOC_ELEM_BF16) {
    // the scored BF16 copy of the row (half the bytes of the FP32 rows), two
    // elements per lane and load; the rows are zero padded to row_bytes
    const uint32_t* x = (const uint32_t*)(f.rows + (int64_t)p * f.row_bytes);
    const float2* h2 = (const float2*)hv;
    for (int i = lane; 2 * i < f.dim; i += 64) {
      const uint32_t u = x[i];
      const float2 hh = h2[i];
      m = fmaf(hh.x, __uint_as_float(u << 16), m);
      if (2 * i + 1 < f.dim) m = fmaf(hh.y, __uint_as_float(u & 0xffff0000u), m);
    }
    m = wave_sum_f32(m);
  } else {
    const float* x = (const float*)(f.frows + (int64_t)p * f.frow_ld);
    for (int i = lane; i < f.dim; i += 64) m = fmaf(hv[i], x[i], m);
    m = wave_sum_f32(m);
  }
  if (lane == 0) {
    int8_t sd;
    if (fabsf(m) < 1e-8f) {
      const int t = (int)(pos / f.n);
      sd = (int8_t)(mix64(f.seeds[t] ^ ((uint64_t)level << 40) ^ ((uint64_t)(uint32_t)p << 1) ^ 0x5bd1e995ull) &
                    1ull);
    } else {
      sd = m > 0.0f ? 0 : 1;
    }
    side[pos] = sd;
  }
}

// BF16 and FP32 rows: 16 lanes per position, four positions per wave; each
// lane takes 16 bytes of the row per step (8 BF16 or 4 FP32 elements) and the
// matching hyperplane floats (entries at or past dim are masked: the rows are
// zero padded, the hyperplane's padding is not written).  A quarter of the
// waves of rp_side_kernel, and 16-byte loads instead of 4-byte ones.
template <int KIND>
__global__ __launch_bounds__(256) void rp_side_wide_kernel(const int32_t* __restrict__ perm,
                                                           const int32_t* __restrict__ seg_of, int64_t n_pos,
                                                           ForestArgs f, int level, int8_t* __restrict__ side) {
  constexpr int EPS = KIND == EVOC_ELEM_BF16 ? 8 : 4;  // elements per 16-byte step
  const int sub = threadIdx.x & 15;
  const int64_t pos = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
  const int s = pos < n_pos ? seg_of[pos] : -1;
  int p = 0;
  float m = 0.0f;
  if (s >= 0) {
    p = perm[pos];
    const uint4* x = KIND == EVOC_ELEM_BF16 ? (const uint4*)(f.rows + (int64_t)p * f.row_bytes)
                                            : (const uint4*)(f.frows + (int64_t)p * f.frow_ld);
    const float* hv = f.hyper + (int64_t)s * f.hyper_ld;
    for (int i = sub; EPS * i < f.dim; i += 16) {
      const int e0 = EPS * i;
      const uint4 u = x[i];
      // e0 and hyper_ld are multiples of 4, so e0 < dim keeps e0 + 3 (and
      // e0 + 4 < dim keeps e0 + 7) inside this segment's hyperplane
      const float4 h0 = *(const float4*)(hv + e0);
      if constexpr (KIND == EVOC_ELEM_BF16) {
        const float4 h1 = e0 + 4 < f.dim ? *(const float4*)(hv + e0 + 4) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float hh[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const float xe = __uint_as_float((q & 1) ? (w[q >> 1] & 0xffff0000u) : (w[q >> 1] << 16));
          if (e0 + q < f.dim) m = fmaf(hh[q], xe, m);
        }
      } else {
        const float hh[4] = {h0.x, h0.y, h0.z, h0.w};
        const float xx[4] = {__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w)};
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (e0 + q < f.dim) m = fmaf(hh[q], xx[q], m);
      }
    }
  }
  // sum over the position's 16 lanes (every lane of the wave takes part)
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) m += __shfl_xor(m, o, 16);
  if (sub == 0 && s >= 0) {
    int8_t sd;
    if (fabsf(m) < 1e-8f) {
      const int t = (int)(pos / f.n);
      sd = (int8_t)(mix64(f.seeds[t] ^ ((uint64_t)level << 40) ^ ((uint64_t)(uint32_t)p << 1) ^ 0x5bd1e995ull) &
                    1ull);
    } else {
      sd = m > 0.0f ? 0 : 1;
    }
    side[pos] = sd;
  }
}

// one workgroup per active segment: stable partition (left first) into perm2,
// random halves when a side is empty, children appended to the next level or
// to the leaf list
__global__ __launch_bounds__(256) void rp_partition_kernel(const int4* __restrict__ segs, int nseg,
                                                           const int32_t* __restrict__ perm,
                                                           int32_t* __restrict__ perm2,
                                                           int8_t* __restrict__ side, ForestArgs f,
                                                           int level, int4* __restrict__ next_segs,
                                                           int32_t* __restrict__ n_next,
                                                           int2* __restrict__ leaves,
                                                           int32_t* __restrict__ n_leaves,
                                                           int last_level) {
  __shared__ int wsum[4];
  __shared__ int total_left;
  const int s = blockIdx.x;
  if (s >= nseg) return;
  const int4 g = segs[s];
  const int local = g.x - g.z * f.n;  // start within the tree
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int attempt = 0; attempt < 2; attempt++) {
    // count the left side
    int c = 0;
    for (int i = t; i < g.y; i += 256) c += side[g.x + i] == 0;
    c = wave_sum_i32(c);
    if (lane == 0) wsum[w] = c;
    __syncthreads();
    if (t == 0) total_left = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const int nl = total_left;
    if (nl > 0 && nl < g.y) break;
    // one side empty: random halves (float.rs:77-90)
    for (int i = t; i < g.y; i += 256)
      side[g.x + i] = (int8_t)(mix64(f.seeds[g.z] ^ ((uint64_t)level << 40) ^ ((uint64_t)(uint32_t)(local + i) << 1) ^
                                     0x2545f4914f6cdd1dull) & 1ull);
    __syncthreads();
  }
  const int nl = total_left;
  // stable scatter: chunks of 256 positions with a block prefix count
  int left_base = 0, right_base = nl;
  for (int c0 = 0; c0 < g.y; c0 += 256) {
    const int i = c0 + t;
    const int sd = i < g.y ? side[g.x + i] : 2;
    const uint64_t ml = __ballot(sd == 0), mr = __ballot(sd == 1);
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __shared__ int wl[4], wr[4];
    if (lane == 0) {
      wl[w] = __popcll(ml);
      wr[w] = __popcll(mr);
    }
    __syncthreads();
    int ol = 0, orr = 0;
    for (int k = 0; k < w; k++) {
      ol += wl[k];
      orr += wr[k];
    }
    if (sd == 0) perm2[g.x + left_base + ol + __popcll(ml & below)] = perm[g.x + i];
    if (sd == 1) perm2[g.x + right_base + orr + __popcll(mr & below)] = perm[g.x + i];
    left_base += wl[0] + wl[1] + wl[2] + wl[3];
    right_base += wr[0] + wr[1] + wr[2] + wr[3];
    __syncthreads();
  }
  if (t == 0) {
    atomicMax(&f.depth[g.z], level + 1);
    const int2 kids[2] = {make_int2(g.x, nl), make_int2(g.x + nl, g.y - nl)};
    for (int k = 0; k < 2; k++) {
      if (kids[k].y > f.leaf_size && !last_level) {
        const int slot = atomicAdd(n_next, 1);
        next_segs[slot] = make_int4(kids[k].x, kids[k].y, g.z, 0);
      } else {
        const int slot = atomicAdd(n_leaves, 1);
        leaves[slot] = kids[k];
      }
    }
  }
}

// positions outside active segments keep their order
__global__ void rp_copy_kernel(const int32_t* __restrict__ perm, int32_t* __restrict__ perm2,
                               const int32_t* __restrict__ seg_of, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && seg_of[i] < 0) perm2[i] = perm[i];
}

// graph keys (sorted best first) -> exact-format keys for the re-score/finalise kernels
__global__ void nnd_export_keys_kernel(const uint64_t* __restrict__ keys, int n, int Kg,
                                       uint64_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * Kg) return;
  const uint64_t k = keys[i];
  const int v = (int)(i / Kg);
  if (k == EVOC_KEY_NONE) {
    out[i] = EVOC_KEY_NONE;
    return;
  }
  const int c = gkey_idx(k);
  out[i] = make_key(gkey_score(k), c, c == v);
}

// sum over all filled slots of the raw distance -score, in f64 (float.rs:854-859)
__global__ void nnd_sum_dist_kernel(const uint64_t* __restrict__ keys, int64_t m,
                                    double* __restrict__ out) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    if (k != EVOC_KEY_NONE) s += -(double)gkey_score(k);
  }
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, s);
}

// ---------------------------------------------------------------------------
// scans (exclusive, int32) for the reverse-edge CSR
// ---------------------------------------------------------------------------
__global__ void scan_block_kernel(const int32_t* __restrict__ in, int n, int32_t* __restrict__ out,
                                  int32_t* __restrict__ block_sums) {
  __shared__ int ws[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i = blockIdx.x * 1024 + t;
  const int x = i < n ? in[i] : 0;
  int incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) ws[w] = incl;
  __syncthreads();
  int off = 0;
  for (int k = 0; k < w; k++) off += ws[k];
  if (i < n) out[i] = off + incl - x;
  if (t == 1023) block_sums[blockIdx.x] = off + incl;
}

__global__ void scan_sums_kernel(int32_t* __restrict__ sums, int nb, int32_t* __restrict__ total) {
  if (threadIdx.x != 0) return;
  int run = 0;
  for (int b = 0; b < nb; b++) {
    const int v = sums[b];
    sums[b] = run;
    run += v;
  }
  *total = run;
}

__global__ void scan_add_kernel(int32_t* __restrict__ out, int n, const int32_t* __restrict__ sums,
                                const int32_t* __restrict__ total) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += sums[blockIdx.x];
  if (i == 0) out[n] = *total;
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

hipError_t nnd_launch_init(uint64_t* keys, float* thr, int32_t* umark, int n, int Kg, hipStream_t s) {
  hipLaunchKernelGGL(nnd_init_lists_kernel, dim3(blocks_for((int64_t)n * Kg, 256)), dim3(256), 0, s, keys,
                     thr, umark, n, Kg);
  return hipGetLastError();
}

hipError_t nnd_launch_join(int kind, const JoinArgs& a, hipStream_t s) {
  const dim3 g(blocks_for(a.n_items, 4)), b(256);
  switch (kind) {
    case EVOC_ELEM_BF16: hipLaunchKernelGGL(nnd_join_mfma_kernel<EVOC_ELEM_BF16>, g, b, 0, s, a); break;
    case EVOC_ELEM_F32: hipLaunchKernelGGL(nnd_join_mfma_kernel<EVOC_ELEM_F32>, g, b, 0, s, a); break;
    case EVOC_ELEM_I8: hipLaunchKernelGGL(nnd_join_mfma_kernel<EVOC_ELEM_I8>, g, b, 0, s, a); break;
    case EVOC_ELEM_U8BITS: hipLaunchKernelGGL(nnd_join_jaccard_kernel, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t nnd_launch_random_lists(const uint64_t* keys, int n, int Kg, int mc, uint64_t seed, int32_t* newc,
                                   int32_t* oldc, hipStream_t s) {
  hipLaunchKernelGGL(nnd_random_lists_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, Kg, mc, seed,
                     newc, oldc);
  return hipGetLastError();
}

hipError_t nnd_launch_merge(uint64_t* keys, float* thr, int32_t* umark, uint64_t* ubuf, int ucap, int n,
                            int Kg, int v0, int nv, uint32_t* changed, hipStream_t s) {
  if (ucap < 1 || ucap > NND_OFFER_CAP || Kg < 1 || Kg > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(nnd_merge_kernel, dim3(blocks_for(nv, 4)), dim3(256), 0, s, keys, thr, umark, ubuf, ucap,
                     n, Kg, v0, nv, changed);
  return hipGetLastError();
}

hipError_t nnd_launch_candidates(uint64_t* keys, int n, int Kg, int32_t* revcnt, int32_t* revoff,
                                 int32_t* revcur, int32_t* rev, int32_t* scan_sums, int mc, uint64_t seed,
                                 int it, int32_t* newc, int32_t* oldc, hipStream_t s) {
  const int64_t nk = (int64_t)n * Kg;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, revcnt, (int64_t)n, 0);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, revcur, (int64_t)n, 0);
  hipLaunchKernelGGL(nnd_rev_count_kernel, dim3(blocks_for(nk, 256)), dim3(256), 0, s, keys, n, Kg, revcnt);
  const int nb = (int)blocks_for(n, 1024);
  hipLaunchKernelGGL(scan_block_kernel, dim3(nb), dim3(1024), 0, s, revcnt, n, revoff, scan_sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(64), 0, s, scan_sums, nb, scan_sums + nb);
  hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(1024), 0, s, revoff, n, scan_sums, scan_sums + nb);
  hipLaunchKernelGGL(nnd_rev_fill_kernel, dim3(blocks_for(nk, 256)), dim3(256), 0, s, keys, n, Kg, revoff,
                     revcur, rev);
  hipLaunchKernelGGL(nnd_build_candidates_kernel, dim3(blocks_for(n, 4)), dim3(256), 0, s, keys, n, Kg, revoff,
                     rev, mc, seed, it, newc, oldc);
  return hipGetLastError();
}

hipError_t nnd_launch_export(const uint64_t* keys, int n, int Kg, uint64_t* out, hipStream_t s) {
  hipLaunchKernelGGL(nnd_export_keys_kernel, dim3(blocks_for((int64_t)n * Kg, 256)), dim3(256), 0, s, keys, n,
                     Kg, out);
  return hipGetLastError();
}

hipError_t nnd_launch_sum_dist(const uint64_t* keys, int64_t m, double* out, hipStream_t s) {
  hipLaunchKernelGGL(nnd_sum_dist_kernel, dim3(1024), dim3(256), 0, s, keys, m, out);
  return hipGetLastError();
}

hipError_t rp_launch_init(int32_t* perm, const ForestArgs& f, hipStream_t s) {
  const int64_t total = (int64_t)f.n * f.n_trees;
  hipLaunchKernelGGL(rp_init_perm_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, perm, f.n, total);
  return hipGetLastError();
}

hipError_t rp_launch_level(const int4* segs, int nseg, const int32_t* perm, int32_t* perm2, int32_t* seg_of,
                           int8_t* side, const ForestArgs& f, int level, int4* next_segs, int32_t* n_next,
                           int2* leaves, int32_t* n_leaves, int last_level, hipStream_t s) {
  const int64_t total = (int64_t)f.n * f.n_trees;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, seg_of, total, -1);
  hipLaunchKernelGGL(rp_seg_fill_kernel, dim3(nseg), dim3(256), 0, s, segs, nseg, seg_of);
  hipLaunchKernelGGL(rp_hyperplane_kernel, dim3(blocks_for(nseg, 4)), dim3(256), 0, s, segs, nseg, perm, f, level);
  if (f.kind == EVOC_ELEM_BF16)
    hipLaunchKernelGGL(rp_side_wide_kernel<EVOC_ELEM_BF16>, dim3(blocks_for(total, 16)), dim3(256), 0, s, perm,
                       seg_of, total, f, level, side);
  else if (f.kind == EVOC_ELEM_F32)
    hipLaunchKernelGGL(rp_side_wide_kernel<EVOC_ELEM_F32>, dim3(blocks_for(total, 16)), dim3(256), 0, s, perm,
                       seg_of, total, f, level, side);
  else
    hipLaunchKernelGGL(rp_side_kernel, dim3(blocks_for(total, 4)), dim3(256), 0, s, perm, seg_of, segs, total, f,
                       level, side);
  hipLaunchKernelGGL(rp_copy_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, perm, perm2, seg_of, total);
  hipLaunchKernelGGL(rp_partition_kernel, dim3(nseg), dim3(256), 0, s, segs, nseg, perm, perm2, side, f, level,
                     next_segs, n_next, leaves, n_leaves, last_level);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL(fill_i32_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace evoc
