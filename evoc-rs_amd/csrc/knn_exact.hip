// Tiled exact all-pairs scoring with fused candidate filtering (gfx950).
//
// This is the GPU form of the reference's exact search (brute_force_knn,
// src/knn.rs:189-281) and of the answer its NN-descent converges to
// (src/nndescent/float.rs:775-879): for every query, the K best candidates
// under the ranking in evoc_device.h.
//
// Three passes, all over row tiles of 128 queries x 128 candidates:
//   pass 1 (threshold): queries x a strided sample of candidates, MFMA scores
//          max-pooled over groups of 16 candidates per lane.  The K-th largest
//          group maximum of a query is a lower bound tau_q on its true K-th
//          best score (K groups hold at least K candidates >= tau_q).
//   pass 2 (score):     queries x all candidates; a score >= tau_q is appended
//          as a 64-bit key to the query's bounded buffer (LDS counter, rare).
//   pass 3 (select):    per query, one wave sorts its buffered keys and keeps
//          the K best.  A buffer that overflowed yields a tighter bound from
//          its kept entries and its query block is scored again.
// The f32 path can score in BF16 (MFMA 32x32x16, FP32 accumulate) and then
// re-score the K = k + margin survivors in FP32 (rescore_kernel); the int8 path
// is exact i8 MFMA into i32; binary Jaccard uses VALU popcounts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include <type_traits>
#include <vector>

#include "evoc_device.h"
#include "evoc_kernels.h"
#include "mfma_traits.h"

namespace evoc {

// ---------------------------------------------------------------------------
// MFMA tile engine.  WM x WN waves, each owning BM x BN MFMA 32x32 blocks, so a
// workgroup tile is (32 BM WM) candidates x (32 BN WN) queries.  Shapes (see
// tile_shape()), LDS = stage ring + pass-2 per-query arrays:
//   0: <2,2> x (2,2)  128 x 128, 64 accumulators per lane, 64 KiB: two workgroups per CU
//   1: <2,4> x (2,2)  128 x 256, 128 accumulators, a 3-stage ring of 72 KiB: two per CU
//   2: <2,4> x (4,2)  256 x 256, 8 waves of 64 x 128 (two per SIMD), 128-byte stages in
//                     a 2-stage ring of 128 KiB (the default: half the stage barriers
//                     and load groups of 64-byte stages, 4-5 % faster)
//   3: <4,4> x (2,2)  256 x 256, 4 waves of 128 x 128, 256 accumulators in AGPRs,
//                     128 KiB: a third less LDS operand traffic per MFMA than 64 x 128
//   4: as 2 with 64-byte stages in a 4-stage ring (EVOC_MI355X_TILE=4 only)
// MFMA orientation: M (rows) = candidates (A operand), N (cols) = queries
// (B operand), so lane l of a 32x32 block holds query column l % 32 and the 16
// candidate rows 8*(j/4) + 4*(l/32) + j%4, j = 0..15.
// ---------------------------------------------------------------------------
constexpr int STAGE_BYTES = 128;     // row bytes per stage of the Jaccard tiles
constexpr int MFMA_SB = 64;          // row bytes per stage of the MFMA tiles (shape 2: 128)
// Stages in flight (LDS ring buffers): four 64-byte stages, three for the
// 128 x 256 tile (its ring then fits two workgroups in a CU's 160 KiB), two
// 128-byte stages for the default 256 x 256 shape (2).
constexpr int mfma_ring(int stage_rows, int sb = MFMA_SB) { return sb == 128 ? 2 : stage_rows == 384 ? 3 : 4; }

// MFMA stage loads go straight from global memory into LDS
// (global_load_lds_dwordx4: lane i of a wave writes 16 bytes at M0 + 16 * i),
// 16 rows of 64 bytes per instruction.  Logical 16-byte chunk c of local row r
// sits in physical slot c ^ key(r), key(r) = ((r >> 2) ^ (r >> 1)) & 3.  A b128
// fragment read takes the same chunk of 16 consecutive rows per LDS cycle over
// 64 banks of 4 bytes; four 64-byte rows share a 256-byte bank line, and key()
// gives rows r, r + 4, r + 8, r + 12 four different slots, so the 16 reads cover
// all 64 banks.  Over 32 banks (8 rows per cycle, two rows per line) it is
// conflict-free as well.  Rows past the end read the last valid row; the
// epilogues mask them.
__device__ __forceinline__ int lds_chunk_key(int r) { return ((r >> 2) ^ (r >> 1)) & 3; }
// 128-byte rows (eight 16-byte chunks, two rows per 256-byte bank line): chunk
// c of row r at slot c ^ ((r >> 1) & 7), so the 16 rows one ds_read_b128 pass
// takes, eight even and eight odd, land on 16 distinct groups of 4 banks
template <int SB>
__device__ __forceinline__ int stage_chunk_key(int r) {
  if constexpr (SB == 128) return (r >> 1) & 7;
  else return lds_chunk_key(r);
}

template <int ROWS, int WAVES, int SB = MFMA_SB>
__device__ __forceinline__ void dma_stage(uint8_t* lds_operand, const uint8_t* __restrict__ base,
                                          int64_t row_bytes, int row0, int nrows, int rstride, int kbyte) {
  constexpr int LPR = SB / 16;    // lanes per row (16 bytes each)
  constexpr int RPI = 64 / LPR;   // rows per load instruction (1 KiB)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the same wave index as a scalar, for the LDS destination (M0) of each
  // load: scalar arithmetic rather than a vector value read back per load
  const int wave_s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int rl = lane / LPR;                // row within the instruction's rows
  static_assert(ROWS % (RPI * WAVES) == 0, "stage rows must split evenly over the waves");
  // uniform 64-bit base of the tile's first row; per-lane 32-bit offsets
  // (a tile spans ROWS * rstride rows, well under 4 GiB)
  const uint32_t pitch = (uint32_t)(rstride * row_bytes);
  const uint8_t* tile = base + (int64_t)row0 * rstride * row_bytes;
  const int last = nrows - 1 - row0;
#pragma unroll
  for (int i = 0; i < ROWS / RPI / WAVES; i++) {
    const int grp = wave + WAVES * i;
    // logical chunk stored at physical slot lane % LPR of local row grp * RPI + rl
    const int chunk = (lane % LPR) ^ stage_chunk_key<SB>(grp * RPI + rl);
    const int r = min(grp * RPI + rl, last);
    const uint32_t off = (uint32_t)r * pitch + (uint32_t)(kbyte + chunk * 16);
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(tile + off),
                                     (__attribute__((address_space(3))) void*)(lds_operand + (wave_s + WAVES * i) * 1024), 16, 0,
                                     0);
  }
}

// Workgroup barrier without the release/acquire fences of __syncthreads(),
// which would drain every outstanding load (vmcnt(0)) and so the stage ring.
// Sufficient for the ring: each wave waits for its own DMA writes (vmcnt)
// before arriving, and every LDS read it made of the retiring stage has been
// consumed by its MFMAs.  The empty asm statements keep the compiler from
// moving memory accesses across it.
__device__ __forceinline__ void ring_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// s_waitcnt with only the vector-memory count bounded (gfx9 encoding: vmcnt in
// bits 3:0 and 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// A wave-uniform pointer the compiler cannot see through (volatile asm, not
// hoisted out of loops), returned in SGPRs (readfirstlane).
__device__ __forceinline__ const uint8_t* opaque_uniform_ptr(const uint8_t* p) {
  uint32_t lo = (uint32_t)(uint64_t)p, hi = (uint32_t)((uint64_t)p >> 32);
  asm volatile("" : "+v"(lo), "+v"(hi));
  lo = __builtin_amdgcn_readfirstlane(lo);
  hi = __builtin_amdgcn_readfirstlane(hi);
  return (const uint8_t*)(((uint64_t)hi << 32) | lo);
}

template <int KIND, int BM, int BN, int WM = 2, int WN = 2, int SB = MFMA_SB>
struct Engine {
  using M = Mfma<KIND>;
  typedef typename M::acc_t acc_t;
  static constexpr int WAVES = WM * WN;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int TC = 32 * BM * WM;               // candidate rows per tile
  static constexpr int TQ = 32 * BN * WN;               // query rows per tile
  static constexpr int STAGE_LDS = (TC + TQ) * SB;
  static constexpr int NS = mfma_ring(TC + TQ, SB);     // ring depth
  static constexpr int LDS = NS * STAGE_LDS;            // ring of NS stages
  static constexpr int GROUPS = WM * BM * 2;            // pass-1 groups per tile
  // DMA instructions each wave issues per stage
  static constexpr int DMA_PER_STAGE = (TC + TQ) / (1024 / SB) / WAVES;
  __device__ static int wave_m() { return (threadIdx.x >> 6) / WN; }
  __device__ static int wave_n() { return (threadIdx.x >> 6) % WN; }

  // stage `st` of the tile (c0, q0) into ring buffer `buf`
  __device__ __forceinline__ static void issue_stage(uint8_t* buf, const uint8_t* c_rows, const uint8_t* q_rows,
                                                     const TileArgs& a, int c0, int q0, int st) {
    dma_stage<TC, WAVES, SB>(buf, c_rows, a.row_bytes, c0, a.n_c, a.c_stride, st * SB);
    dma_stage<TQ, WAVES, SB>(buf + TC * SB, q_rows, a.row_bytes, q0, a.n_q, 1, st * SB);
  }

  // Whether a workgroup's next tile can be primed by this one: its first
  // NS stages are loaded during this tile's last NS stages, which
  // needs at least NS stages per tile.
  __device__ static bool can_prime(const TileArgs& a) { return a.row_bytes / SB >= NS; }

  // Accumulate the TC x TQ score tile (candidates c0.., queries q0..) over all
  // stages.  Stage st of the tile lives in ring buffer (ph + st) % NS;
  // `ph` carries the ring phase from tile to tile.  The stage boundary is
  // crossed inside the previous stage: once a wave has issued (and waited
  // for) its LDS reads of the last K step of stage s, it waits for its own
  // loads of stage s + 1, and one barrier then publishes stage s + 1 and
  // retires buffer (ph + s) % NS, which the loads of stage s + NS
  // refill (loads run NS stages ahead).  The first reads of stage s + 1
  // then overlap the last MFMAs of stage s, so no stage starts on an LDS
  // bubble.  Within a K step, the reads of the next step alternate with the
  // MFMAs.
  // With `has_next` the loads of the last NS crossings are the first
  // NS stages of the next tile (nc0, nq0): they land during this tile's
  // epilogue, and the next call, `primed`, starts on them.  Memory operations
  // issued between two tiles (the epilogue's) only make the vmcnt waits
  // stricter: everything older than a stage's loads has landed with them.
  // `published`: the primed stages have landed in every wave and a workgroup
  // barrier has passed since (the caller's), so the first stage needs no wait.
  __device__ __forceinline__ static void score_tile(acc_t (&acc)[BM][BN], uint8_t* lds, const TileArgs& a, int c0,
                                                    int q0, int& ph, bool primed, bool has_next, int nc0,
                                                    int nq0, bool published = false) {
    const int lane = threadIdx.x & 63;
    const int wm = wave_m(), wn = wave_n();
    // (the accumulators start at zero through the first K step's MFMAs,
    // which add to an inline zero instead of the previous tile's values)
    const int nst = (int)(a.row_bytes / SB);
    // fragment rows of this lane; every block starts at a multiple of 32 rows,
    // so the swizzle key of the lane's row is the same in every block
    const int arow = (wm * 32 * BM + (lane & 31)) * SB;
    const int brow = TC * SB + (wn * 32 * BN + (lane & 31)) * SB;
    const int sw = stage_chunk_key<SB>(lane & 31);
    const uint8_t* c_rows = opaque_uniform_ptr(a.c_rows);
    const uint8_t* q_rows = opaque_uniform_ptr(a.q_rows);
    constexpr int KS = SB / M::KSTEP_BYTES;
    static_assert(KS % 2 == 0, "fragment double buffer parity must carry across stages");
    typename M::frag_t fa[2][BM], fb[2][BN];
    // (ph + st >= 0: unsigned, so a ring of 2 or 4 indexes with one AND)
    auto ring = [&](int st) __attribute__((always_inline)) {
      return lds + ((unsigned)(ph + st) % (unsigned)NS) * STAGE_LDS;
    };
    auto read_frags = [&](const uint8_t* buf, int ks, int par) __attribute__((always_inline)) {
      const int kb = ks * M::KSTEP_BYTES + (lane >> 5) * M::LANE_BYTES;
      const int off = (((kb >> 4) ^ sw) << 4) + (kb & 15);
#pragma unroll
      for (int b = 0; b < BM; b++) fa[par][b] = M::load(buf + arow + off + b * 32 * SB);
#pragma unroll
      for (int b = 0; b < BN; b++) fb[par][b] = M::load(buf + brow + off + b * 32 * SB);
    };
    // wait until this wave's loads of `stage` have landed; `newer` later stages may stay in flight
    auto wait_stage = [&](int newer) __attribute__((always_inline)) {
      if (newer >= 3) wait_vmcnt<3 * DMA_PER_STAGE>();
      else if (newer == 2) wait_vmcnt<2 * DMA_PER_STAGE>();
      else if (newer == 1) wait_vmcnt<DMA_PER_STAGE>();
      else wait_vmcnt<0>();
    };
    // stage st of this tile, or stage st - nst of the next one
    auto issue = [&](int st) __attribute__((always_inline)) {
      if (st < nst) issue_stage(ring(st), c_rows, q_rows, a, c0, q0, st);
      else issue_stage(ring(st), c_rows, q_rows, a, nc0, nq0, st - nst);
    };
    if (!primed) {
      // the ring may still be read by waves finishing the previous tile
      __syncthreads();
      // (kept rolled: unrolled, the compiler precomputes each stage's 64-bit
      // addresses, runs out of registers and spills them)
#pragma unroll 1
      for (int st = 0; st < NS; st++)
        if (st < nst) issue(st);
    }
    if (!published) {
      wait_stage(min(NS - 1, nst - 1));
      ring_barrier();
    }
    read_frags(ring(0), 0, 0);
    // MFMAs of one K step from fragment buffer `par`: one MFMA, then all the
    // fragment reads of the next step issued just before (the first MFMA's
    // operand wait covers the previous step's reads only), then the remaining
    // MFMAs, which give the reads the whole step to land
    // (`first`: the tile's first K step, whose MFMAs start from zero)
    auto mma_step = [&](int par, auto first) __attribute__((always_inline)) {
#pragma unroll
      for (int bm = 0; bm < BM; bm++)
#pragma unroll
        for (int bn = 0; bn < BN; bn++) {
          if constexpr (decltype(first)::value) M::mma_zero(acc[bm][bn], fa[par][bm], fb[par][bn]);
          else M::mma(acc[bm][bn], fa[par][bm], fb[par][bn]);
        }
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);        // MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, BM + BN, 0);  // DS reads
      __builtin_amdgcn_sched_group_barrier(0x008, BM * BN - 1, 0);
      __builtin_amdgcn_sched_barrier(0);
    };
    using no_t = std::integral_constant<bool, false>;
    using yes_t = std::integral_constant<bool, true>;
    // one stage; the first is peeled so that its first K step starts from zero
    auto stage = [&](int s, auto first_stage) __attribute__((always_inline)) {
      const uint8_t* cur = ring(s);
#pragma unroll
      for (int ks = 0; ks + 1 < KS; ks++) {
        const int p = ks & 1;
        read_frags(cur, ks + 1, p ^ 1);
        if constexpr (decltype(first_stage)::value) {
          if (ks == 0) mma_step(p, yes_t{});
          else mma_step(p, no_t{});
        } else {
          mma_step(p, no_t{});
        }
      }
      constexpr int PL = (KS - 1) & 1;  // buffer of the last K step
      // this wave's reads of stage s are complete (lgkmcnt)
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0), vmcnt / expcnt untouched
      if (s + 1 < nst) {
        // cross into stage s + 1: its loads have landed (vmcnt; two later
        // stages stay in flight while the stream continues)
        wait_stage(has_next ? NS - 2 : min(NS - 2, nst - 2 - s));
        ring_barrier();
        if (s + NS < nst || has_next) issue(s + NS);
        read_frags(ring(s + 1), 0, PL ^ 1);
      } else {
        // last stage: every wave is done reading the ring
        ring_barrier();
        if (has_next) issue(s + NS);
      }
      mma_step(PL, no_t{});
    };
    stage(0, yes_t{});
    for (int s = 1; s < nst; s++) stage(s, no_t{});
    // without a prefetch every stage load has landed (the last crossing
    // waited for all); saying so lets the epilogue's LDS accesses go without
    // further vmcnt waits
    if (!has_next) wait_vmcnt<0>();
    ph = (int)((unsigned)(ph + nst) % (unsigned)NS);
  }
};

// Workgroup -> (query block, candidate split), XCD-aware.  The hardware deals
// workgroups of a launch round-robin to the 8 XCDs (linear id mod 8), each with
// its own L2.  The grid is 1-D, padded to a multiple of 8, and XCD x takes the
// x-th contiguous range of logical ids.  Logical ids run over groups of
// group_q query blocks; inside a group, split-major.  The workgroups resident
// on one XCD at a time then cover about group_q query blocks x n_splits
// candidate splits: each query tile is shared by n_splits workgroups and each
// candidate tile by group_q, so the L2 of that XCD fetches group_q + n_splits
// tiles per tile step instead of one per workgroup.
__device__ __forceinline__ bool tile_coords(const TileArgs& a, int& qb, int& split) {
  const int per_xcd = (int)gridDim.x >> 3;
  const int logical = ((int)blockIdx.x & 7) * per_xcd + ((int)blockIdx.x >> 3);
  if (logical >= a.n_qblocks * a.n_splits) return false;
  const int span = a.group_q * a.n_splits;
  const int g = logical / span;
  const int r = logical - g * span;
  const int gq = min(a.group_q, a.n_qblocks - g * a.group_q);
  split = r / gq;
  const int slot = g * a.group_q + (r - split * gq);
  qb = a.block_list ? a.block_list[slot] : slot;
  return true;
}

// A lower bound on the score test that never rejects what the exact test
// accepts: f32 scores are dot > 0 ? dot : -FLT_MIN, so for a bound above
// -FLT_MIN the raw dot must reach it; below that everything may pass.
template <int KIND>
__device__ __forceinline__ float fast_bound(float tau) {
  if (KIND == EVOC_ELEM_I8) return tau;
  return tau > -1.17549435e-38f ? tau : -INFINITY;
}

// Whether the cheap bound test of a row is its exact test for every value
// outside the self pair: no key bound (first round), and integer scores (the
// score is the accumulator as float) or a positive bound (an f32 score is
// dot > 0 ? dot : -FLT_MIN, so for tau > 0 the raw dot reaches tau exactly
// when the score does).  Rows past the end carry an infinite bound: no value
// passes on their side.
template <int KIND>
__device__ __forceinline__ bool bound_is_exact(float tau, bool key_bound) {
  return !key_bound && (KIND == EVOC_ELEM_I8 || tau > 0.0f);
}

// Bits j of a lane's 16 accumulator rows (candidates cb + 8 (j / 4) +
// 4 (lane / 32) + j % 4) that are real candidates (< n_c).
__device__ __forceinline__ uint32_t valid_rows(int cb, int lane, int n_c) {
  if (cb + 32 <= n_c) return 0xffffu;
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) v |= (cb + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3) < n_c ? 1u : 0u) << j;
  return v;
}

// ---------------------------------------------------------------------------
// Pass 1: group maxima over a strided candidate sample (GROUPS per tile:
// wave row half x BM blocks x lane half, 16 candidates each).
// ---------------------------------------------------------------------------
// The work of one (query block, candidate split): group maxima of its tiles.
template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__device__ __forceinline__ void pass1_work(const TileArgs& a, int qb, int split, uint8_t* lds) {
  using En = Engine<KIND, BM, BN, WM, WN, SB>;
  using M = Mfma<KIND>;
  typename En::acc_t acc[BM][BN];
  const int lane = threadIdx.x & 63;
  const int wm = En::wave_m(), wn = En::wave_n();
  const int q0 = qb * En::TQ;
  // candidate tiles of this split; group maxima are per tile
  const int t_begin = split * a.tiles_per_split;
  const int t_end = min(a.n_ctiles, t_begin + a.tiles_per_split);
  const bool chain = En::can_prime(a);
  int ph = 0;
  for (int t = t_begin; t < t_end; t++) {
    const int c0 = t * En::TC;
    En::score_tile(acc, lds, a, c0, q0, ph, chain && t > t_begin, chain && t + 1 < t_end, c0 + En::TC, q0);
    // (see pass 2: indices opaque here so derived values are not held through the MFMA loop)
    int lane_e = lane, wm_e = wm, wn_e = wn;
    asm volatile("" : "+v"(lane_e), "+v"(wm_e), "+v"(wn_e));
    static_for<BM>([&](auto bmc) __attribute__((always_inline)) {
      constexpr int bm = decltype(bmc)::value;
      static_for<BN>([&](auto bnc) __attribute__((always_inline)) {
        constexpr int bn = decltype(bnc)::value;
        __builtin_amdgcn_sched_barrier(0);  // keep this block's accumulator reads here
        const int q = q0 + wn_e * 32 * BN + bn * 32 + (lane_e & 31);
        const int cb = c0 + wm_e * 32 * BM + bm * 32;  // first candidate of the block
        // the block's 32 candidates (data rows cb * c_stride ..) and the wave's 32
        // queries (data rows q - lane % 32 + q_offset ..) may share a row only where
        // the two ranges meet; elsewhere, with every candidate in range, the group
        // maximum is the score of the largest raw value (scores are monotone in it)
        const int64_t qd0 = (int64_t)(q - (lane_e & 31)) + a.q_offset;
        const bool plain = cb + 32 <= a.n_c && ((int64_t)(cb + 31) * a.c_stride < qd0 ||
                                                (int64_t)cb * a.c_stride > qd0 + 31);
        float m = -INFINITY;
        if (plain) {
          float r = (float)acc[bm][bn][0];
#pragma unroll
          for (int j = 1; j < 16; j++) r = fmaxf(r, (float)acc[bm][bn][j]);
          m = KIND == EVOC_ELEM_I8 ? r : f32_score(r);
        } else {
#pragma unroll
          for (int j = 0; j < 16; j++) {
            const int c = cb + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
            const bool self = (int64_t)c * a.c_stride == (int64_t)q + a.q_offset;
            if (c < a.n_c && !(self && a.skip_self)) {
              const float s = self ? 1.0f : M::value(acc[bm][bn], j);
              m = fmaxf(m, s);
            }
          }
        }
        const int g = ((t * WM + wm_e) * BM + bm) * 2 + (lane_e >> 5);
        if (q < a.n_q) a.gmax[(int64_t)q * a.ng + g] = m;
      });
    });
  }
}

template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__global__ __launch_bounds__(64 * WM * WN, BM * BN == 16 ? 1 : 2) void pass1_groupmax_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  int qb, split;
  if (!tile_coords(a, qb, split)) return;
  pass1_work<KIND, BM, BN, WM, WN, SB>(a, qb, split, lds);
}

// ---------------------------------------------------------------------------
// Pass 2: threshold filter + append, one candidate split per workgroup.  Each
// accumulator value is tested against a cheap bound first; the exact test and
// the append run only in lanes whose value passes (exec-masked, inline).
// ---------------------------------------------------------------------------
// The work of one (query block, candidate split): appends of its tiles and the
// split's per-query counts.
template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__device__ __forceinline__ void pass2_work(const TileArgs& a, int qb, int split, uint8_t* lds) {
  using En = Engine<KIND, BM, BN, WM, WN, SB>;
  using M = Mfma<KIND>;
  // LDS after the stage ring: per-query append counters and bounds (the bounds
  // live here rather than in registers, which the MFMA tile needs)
  int* cnt = (int*)(lds + En::LDS);
  float* tau_l = (float*)(cnt + En::TQ);
  float* tfast_l = tau_l + En::TQ;
  uint64_t* tkey_l = (uint64_t*)(tfast_l + En::TQ);
  typename En::acc_t acc[BM][BN];
  const int lane = threadIdx.x & 63;
  const int wm = En::wave_m(), wn = En::wave_n();
  const int q0 = qb * En::TQ;
  for (int i = threadIdx.x; i < En::TQ; i += En::THREADS) {
    const int q = q0 + i;
    const float tau = q < a.n_q ? (a.tau ? a.tau[q] : -INFINITY) : INFINITY;
    const uint64_t tkey = (q < a.n_q && a.tau_key) ? a.tau_key[q] : 0ull;
    float tb = tau;
    if (a.tau_key && tkey != EVOC_KEY_NONE) tb = fmaxf(tb, key_score(tkey));
    cnt[i] = 0;
    tau_l[i] = tau;
    tfast_l[i] = fast_bound<KIND>(tb);
    tkey_l[i] = tkey;
  }
  // (score_tile begins with a barrier, which publishes these)
  const int t_begin = split * a.tiles_per_split;
  const int t_end = min(a.n_ctiles, t_begin + a.tiles_per_split);
  const bool chain = En::can_prime(a);
  int ph = 0;
  for (int t = t_begin; t < t_end; t++) {
    const int c0 = t * En::TC;
    En::score_tile(acc, lds, a, c0, q0, ph, chain && t > t_begin, chain && t + 1 < t_end, c0 + En::TC, q0);
    // the next tile's first stages may still be landing in LDS: waiting for
    // them once here spares each block's LDS atomic a wait (vmcnt) that would
    // also cover the previous block's global stores
    wait_vmcnt<0>();
    // lane / wave indices made opaque here: what the epilogue derives from
    // them is recomputed per tile instead of living in registers through the
    // MFMA loop (which needs them all)
    int lane_e = lane, wm_e = wm, wn_e = wn;
    asm volatile("" : "+v"(lane_e), "+v"(wm_e), "+v"(wn_e));
    static_for<BN>([&](auto bnc) __attribute__((always_inline)) {
      constexpr int bn = decltype(bnc)::value;
      const int ql = wn_e * 32 * BN + bn * 32 + (lane_e & 31);
      const int q = q0 + ql;
      const int qg = q + a.q_offset;
      const float tfast = tfast_l[ql];
      static_for<BM>([&](auto bmc) __attribute__((always_inline)) {
        constexpr int bm = decltype(bmc)::value;
        __builtin_amdgcn_sched_barrier(0);  // keep this block's accumulator reads here
        const int cb = c0 + wm_e * 32 * BM + bm * 32;      // first candidate of the block
        const bool diag = cb <= qg + 31 - (lane_e & 31) && qg - (lane_e & 31) <= cb + 31;
        // values of this lane's 16 candidates that pass the cheap bound
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) m |= ((float)acc[bm][bn][j] >= tfast ? 1u : 0u) << j;
        if (diag) m = 0xffffu;  // the self pair takes the exact path
        if (__ballot(m != 0) == 0) return;
        const float tau = tau_l[ql];
        uint32_t take;
        if (!diag && bound_is_exact<KIND>(tau, a.tau_key != nullptr)) {
          take = m & valid_rows(cb, lane_e, a.n_c);
        } else {
          // the exact test on those values
          const uint64_t tkey = tkey_l[ql];
          take = 0;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            if (m & (1u << j)) {
              const int c = cb + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
              const bool self = (c == qg);
              const float sc = self ? 1.0f : M::value(acc[bm][bn], j);
              bool t = (c < a.n_c) && (q < a.n_q) && (sc >= tau) && !(self && a.skip_self);
              if (t && a.tau_key) t = make_key(sc, c, self) >= tkey;
              take |= (t ? 1u : 0u) << j;
            }
          }
        }
        // one slot reservation per lane and block, then the appends
        if (take) {
          int pos = atomicAdd(&cnt[ql], __popc(take));
          uint64_t* dst = a.bufs + ((int64_t)split * a.n_q + q) * a.caps;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            if (take & (1u << j)) {
              const int c = cb + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
              const bool self = (c == qg);
              if (pos < a.caps) dst[pos] = make_key(self ? 1.0f : M::value(acc[bm][bn], j), c, self);
              pos++;
            }
          }
        }
      });
    });
  }
  __syncthreads();
  for (int i = threadIdx.x; i < En::TQ; i += En::THREADS) {
    const int q = q0 + i;
    if (q < a.n_q) a.counts[(int64_t)split * a.n_q + q] = cnt[i];
  }
}

template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__global__ __launch_bounds__(64 * WM * WN, BM * BN == 16 ? 1 : 2) void pass2_append_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  int qb, split;
  if (!tile_coords(a, qb, split)) return;
  pass2_work<KIND, BM, BN, WM, WN, SB>(a, qb, split, lds);
}

// ---------------------------------------------------------------------------
// Symmetric pass 2.  Every score here is symmetric (s(q, c) = s(c, q): the
// same products in the same K order), so each unordered pair of tiles is
// scored once: the workgroup of work item (Q, [t0, t1), split) scores query
// block Q against candidate tiles t >= Q and tests every value twice, against
// the query row's bound (appended to the workgroup's own split buffer, LDS
// counters, as pass 2) and, off the diagonal tile, against the candidate row's
// bound (the key of the pair seen from the candidate row, with the query as
// neighbour).  Candidate-side keys are staged per candidate row in LDS and
// flushed once per tile with one global atomicAdd per row that has any; a row
// with more than SYM_SLOTS keys in a tile appends the rest directly.
// ---------------------------------------------------------------------------
constexpr int SYM_SLOTS = 8;

// An integer b with acc >= b whenever (float)acc >= t, for an i32
// accumulator: floor(t), less 256 where floats of that size are more than one
// apart (|(float)acc - acc| <= 128 below 2^31); INT_MIN / INT_MAX at the ends
// (a NaN bound lets everything through to the exact test).
__device__ __forceinline__ int early_int_bound(float t) {
  if (!(t > -2147483520.0f)) return INT_MIN;
  if (t >= 2147483520.0f) return INT_MAX;
  int b = (int)floorf(t);
  if (fabsf(t) >= 16777216.0f) b = b > INT_MIN + 256 ? b - 256 : INT_MIN;
  return b;
}

template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__global__ __launch_bounds__(64 * WM * WN, BM * BN == 16 ? 1 : 2) void pass2_sym_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  using En = Engine<KIND, BM, BN, WM, WN, SB>;
  using M = Mfma<KIND>;
  static_assert(En::TC == En::TQ, "the symmetric pass needs square tiles");
  constexpr int T = En::TC;
  static_assert(En::THREADS >= T && T % 64 == 0, "one candidate row per thread of the first T / 64 waves");
  // LDS after the ring: query rows (counters, bounds, fast bounds, key bounds),
  // candidate rows (bounds, fast bounds, key bounds, staged counts), staged
  // keys, and one flag per wave of the candidate-row loaders (all of its rows'
  // bounds exact)
  int* cnt = (int*)(lds + En::LDS);
  float* tau_l = (float*)(cnt + T);
  float* tfast_l = tau_l + T;
  uint64_t* tkey_l = (uint64_t*)(tfast_l + T);
  float* ctau_l = (float*)(tkey_l + T);
  float* cfast_l = ctau_l + T;
  uint64_t* ckey_l = (uint64_t*)(cfast_l + T);
  int* scnt = (int*)(ckey_l + T);
  uint64_t* stage = (uint64_t*)(scnt + T);
  int* cexact = (int*)(stage + T * SYM_SLOTS);
  // int8: the fast bounds of query and candidate rows as integer lower bounds
  // of the accumulators (early_int_bound), for the blocks' early-out test
  int* tfast_i = cexact + 4;
  int* cfast_i = tfast_i + T;
  typename En::acc_t acc[BM][BN];
  const int lane = threadIdx.x & 63;
  const int wm = En::wave_m(), wn = En::wave_n();
  const int per_xcd = (int)gridDim.x >> 3;
  const int logical = ((int)blockIdx.x & 7) * per_xcd + ((int)blockIdx.x >> 3);
  if (logical >= a.n_work) return;
  const int4 w = a.work[logical];
  const int q0 = w.x * T, split = w.w;
  const bool key_bound = a.tau_key != nullptr;
  // bounds of row r (queries and candidates are the same rows); past the end
  // the bound is infinite, so nothing passes on that row's side
  auto load_bound = [&](int r, float& tau, uint64_t& key) __attribute__((always_inline)) {
    tau = r < a.n_q ? (a.tau ? a.tau[r] : -INFINITY) : INFINITY;
    key = (r < a.n_q && key_bound) ? a.tau_key[r] : 0ull;
  };
  auto fast_of = [&](float tau, uint64_t key) __attribute__((always_inline)) {
    float tb = tau;
    if (key_bound && key != EVOC_KEY_NONE) tb = fmaxf(tb, key_score(key));
    return fast_bound<KIND>(tb);
  };
  for (int i = threadIdx.x; i < T; i += En::THREADS) {
    float tau;
    uint64_t key;
    load_bound(q0 + i, tau, key);
    cnt[i] = 0;
    tau_l[i] = tau;
    tfast_l[i] = fast_of(tau, key);
    tfast_i[i] = early_int_bound(tfast_l[i]);
    tkey_l[i] = key;
  }
  // candidate-row bounds of a tile (row threadIdx.x of it), into LDS with the
  // tile's exactness flags
  auto store_cbounds = [&](int c0, float tau, uint64_t key) __attribute__((always_inline)) {
    const int i = threadIdx.x;
    scnt[i] = 0;
    ctau_l[i] = tau;
    cfast_l[i] = fast_of(tau, key);
    cfast_i[i] = early_int_bound(cfast_l[i]);
    ckey_l[i] = key;
    const bool ex = c0 + i >= a.n_q || bound_is_exact<KIND>(tau, key_bound);
    const uint64_t all = __ballot(ex);
    if (lane == 0) cexact[i >> 6] = all == ~0ull ? 1 : 0;
  };
  // the first tile's bounds now (the diagonal tile has no candidate side);
  // each later tile's are loaded before the previous tile's MFMA loop and
  // stored after its epilogue, so no tile waits on a load of them
  if (w.y != w.x && threadIdx.x < T) {
    float tau;
    uint64_t key;
    load_bound(w.y * T + (int)threadIdx.x, tau, key);
    store_cbounds(w.y * T, tau, key);
  }
  float pf_tau = INFINITY;
  uint64_t pf_key = 0;
  const bool chain = En::can_prime(a);
  int ph = 0;
  for (int t = w.y; t < w.z; t++) {
    const int c0 = t * T;
    const bool diag = t == w.x;  // the query block's own tile: query side only
    // The previous tile ended with a workgroup barrier after every wave's
    // epilogue (and after each wave's loads of this tile's first stages had
    // landed: the barrier waits for them), then the flush by the threads that
    // now store this tile's candidate bounds (thread i: row i, both times).
    // So no barrier here, and this tile's first stage is already published:
    // the waves that do not flush start on its MFMAs at once.  The bounds
    // written here reach the other waves through the stage barriers below,
    // every one of which follows an LDS wait (lgkmcnt(0)).
    if (t != w.y && threadIdx.x < T) store_cbounds(c0, pf_tau, pf_key);  // (t > w.y >= w.x: never diagonal)
    if (t + 1 < w.z && threadIdx.x < T) load_bound(c0 + T + (int)threadIdx.x, pf_tau, pf_key);
    const bool primed = chain && t > w.y;
    En::score_tile(acc, lds, a, c0, q0, ph, primed, chain && t + 1 < w.z, c0 + T, q0, primed);
    int lane_e = lane, wm_e = wm, wn_e = wn;
    asm volatile("" : "+v"(lane_e), "+v"(wm_e), "+v"(wn_e));
    // every candidate row of the tile has an exact cheap test
    bool c_fast = !diag;
#pragma unroll
    for (int g = 0; g < T / 64; g++) c_fast = c_fast && cexact[g] != 0;
    // Early-out, unrolled over the blocks: a block takes the full epilogue
    // below only if some value of it reaches the cheap bound of its query row
    // or of its candidate row (ballots of compares straight on the
    // accumulators, ORed in scalar registers), or if it lies on the diagonal
    // tile.  With the estimated bounds most blocks of a large matrix hold no
    // key, and this test costs a fifth of the masks it saves; the host turns
    // it off (early_out = 0) where most blocks hold keys.
    uint32_t need = 0;
    if (diag || !a.early_out) {
      need = (1u << (BM * BN)) - 1u;
    } else {
      static_for<BM>([&](auto bmc) __attribute__((always_inline)) {
        constexpr int bm = decltype(bmc)::value;
        using B = typename std::conditional<KIND == EVOC_ELEM_I8, int, float>::type;
        // bounds of the lane's 16 candidate rows 8 (j / 4) + 4 (lane / 32) + j % 4
        const int rl = wm_e * 32 * BM + bm * 32 + 4 * (lane_e >> 5);
        B cb[16];
#pragma unroll
        for (int g = 0; g < 4; g++) {
          if constexpr (KIND == EVOC_ELEM_I8) {
            const int4 f = *(const int4*)(cfast_i + rl + 8 * g);
            cb[4 * g] = f.x; cb[4 * g + 1] = f.y; cb[4 * g + 2] = f.z; cb[4 * g + 3] = f.w;
          } else {
            const float4 f = *(const float4*)(cfast_l + rl + 8 * g);
            cb[4 * g] = f.x; cb[4 * g + 1] = f.y; cb[4 * g + 2] = f.z; cb[4 * g + 3] = f.w;
          }
        }
        static_for<BN>([&](auto bnc) __attribute__((always_inline)) {
          constexpr int bn = decltype(bnc)::value;
          const int ql = wn_e * 32 * BN + bn * 32 + (lane_e & 31);
          B tb;
          if constexpr (KIND == EVOC_ELEM_I8) tb = tfast_i[ql];
          else tb = tfast_l[ql];
          // One ballot per block: the lane's largest value against its query
          // row's bound, and its largest margin x - cb over its candidate
          // rows' bounds against zero (max trees, one scalar OR per block
          // rather than two per value).  f32: x >= cb gives x - cb >= +0, and
          // a difference that rounds to -0 only lets the block through to the
          // exact test; i32: the saturating difference keeps the sign of
          // x - cb.
          B mq = acc[bm][bn][0], mc;
          if constexpr (KIND == EVOC_ELEM_I8) mc = __builtin_elementwise_sub_sat(acc[bm][bn][0], cb[0]);
          else mc = acc[bm][bn][0] - cb[0];
#pragma unroll
          for (int j = 1; j < 16; j++) {
            if constexpr (KIND == EVOC_ELEM_I8) {
              mq = max(mq, acc[bm][bn][j]);
              mc = max(mc, __builtin_elementwise_sub_sat(acc[bm][bn][j], cb[j]));
            } else {
              mq = fmaxf(mq, acc[bm][bn][j]);
              mc = fmaxf(mc, acc[bm][bn][j] - cb[j]);
            }
          }
          if (__ballot(mq >= tb || mc >= (B)0)) need |= 1u << (bm * BN + bn);
        });
      });
    }
    // The next tile's first stages may still be landing in LDS.  The compiler
    // makes every LDS write or atomic wait (vmcnt) for such loads, which also
    // waits for every global store issued before it: one wait here, before the
    // appends, lets the LDS atomics below go without waiting on the stores.
    if (need) wait_vmcnt<0>();
    // The blocks that need it, one at a time, rolled: the block's 16 values are
    // copied out under a uniform branch, so the epilogue's code is one block's
    // worth (unrolled over all BM x BN blocks it outgrows the instruction cache
    // and is fetched again on every tile).
#pragma unroll 1
    while (need) {
      const int sb = __builtin_ctz(need);
      need &= need - 1u;
      const int bm = sb / BN, bn = sb % BN;
      float v[16];  // raw accumulator values of this block
      static_for<BM * BN>([&](auto sbc) __attribute__((always_inline)) {
        constexpr int c = decltype(sbc)::value;
        if (sb == c) {
#pragma unroll
          for (int j = 0; j < 16; j++) v[j] = (float)acc[c / BN][c % BN][j];
        }
      });
      auto val = [&](int j) __attribute__((always_inline)) {
        return KIND == EVOC_ELEM_I8 ? v[j] : f32_score(v[j]);
      };
      // value j of the block for a per-lane j: a tree of bitwise selects on
      // all-ones / zero masks, no branches.  The masks are opaque so that the
      // tree is not folded back into an indexed load (which would put the
      // block in scratch memory).
      auto pick_bits = [&](int j) __attribute__((always_inline)) {
        uint32_t m0 = 0u - (uint32_t)(j & 1), m1 = 0u - (uint32_t)((j >> 1) & 1);
        uint32_t m2 = 0u - (uint32_t)((j >> 2) & 1), m3 = 0u - (uint32_t)((j >> 3) & 1);
        asm volatile("" : "+v"(m0), "+v"(m1), "+v"(m2), "+v"(m3));
        auto sel = [](uint32_t m, uint32_t hi, uint32_t lo) __attribute__((always_inline)) {
          return (m & hi) | (~m & lo);
        };
        uint32_t t8[8], t4[4], t2[2];
#pragma unroll
        for (int i = 0; i < 8; i++) t8[i] = sel(m0, __float_as_uint(v[2 * i + 1]), __float_as_uint(v[2 * i]));
#pragma unroll
        for (int i = 0; i < 4; i++) t4[i] = sel(m1, t8[2 * i + 1], t8[2 * i]);
#pragma unroll
        for (int i = 0; i < 2; i++) t2[i] = sel(m2, t4[2 * i + 1], t4[2 * i]);
        return sel(m3, t2[1], t2[0]);
      };
      auto pick = [&](int j) __attribute__((always_inline)) {
        const float raw = __uint_as_float(pick_bits(j));
        return KIND == EVOC_ELEM_I8 ? raw : f32_score(raw);
      };
      const int cb = c0 + wm_e * 32 * BM + bm * 32;  // first candidate of the block row
      // candidate-side fast bounds of this lane's 16 rows 8*(j/4) + 4*(lane/32) + j%4
      float cf[16];
      const int rl = cb - c0 + 4 * (lane_e >> 5);
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const float4 f = diag ? make_float4(INFINITY, INFINITY, INFINITY, INFINITY)
                              : *(const float4*)(cfast_l + rl + 8 * g);
        cf[4 * g] = f.x;
        cf[4 * g + 1] = f.y;
        cf[4 * g + 2] = f.z;
        cf[4 * g + 3] = f.w;
      }
      const int ql = wn_e * 32 * BN + bn * 32 + (lane_e & 31);
      const int q = q0 + ql;
      const float tfast = tfast_l[ql];
      const bool dblk = diag && cb <= q + 31 - (lane_e & 31) && q - (lane_e & 31) <= cb + 31;
      uint32_t m = 0, mc = 0;
#pragma unroll
      for (int j = 0; j < 16; j++) {
        m |= (v[j] >= tfast ? 1u : 0u) << j;
        mc |= (v[j] >= cf[j] ? 1u : 0u) << j;
      }
      if (dblk) m = 0xffffu;  // the self pair takes the exact path
      if (q >= a.n_q) mc = 0;  // a query row past the end is nobody's neighbour
      if (__ballot((m | mc) != 0) == 0) continue;
      const float tau = tau_l[ql];
      // Common case, decided per wave: off the diagonal, every row's bound
      // exact (first round, positive for f32 scores) on both sides.  Then a
      // value that passes is the score itself (positive for f32: its order
      // key is its bits with the sign bit set) and never a self pair, and one
      // loop over the union of both sides' bits appends each value once.
      if (c_fast && __ballot(!bound_is_exact<KIND>(tau, key_bound)) == 0) {
        const uint32_t take = m & valid_rows(cb, lane_e, a.n_c);
        int pos = 0;
        if (take) pos = atomicAdd(&cnt[ql], __popc(take));
        uint64_t* dst = a.bufs + ((int64_t)split * a.n_q + q) * a.caps;
        uint32_t all = take | mc;
        while (all) {
          const int j = __builtin_ctz(all);
          all &= all - 1;
          const uint32_t bits = pick_bits(j);
          const uint32_t hi = KIND == EVOC_ELEM_I8 ? ord_f32(__uint_as_float(bits)) : (bits | 0x80000000u);
          const int roff = 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
          if ((take >> j) & 1u) {  // make_key(score, cb + roff, false)
            if (pos < a.caps) dst[pos] = ((uint64_t)hi << 32) | (uint64_t)(0xfffffffeu - (uint32_t)(cb + roff));
            pos++;
          }
          if ((mc >> j) & 1u) {  // make_key(score, q, false) for row cb + roff
            const int cr = cb - c0 + roff;
            const uint64_t key = ((uint64_t)hi << 32) | (uint64_t)(0xfffffffeu - (uint32_t)q);
            const int slot = atomicAdd(&scnt[cr], 1);
            if (slot < SYM_SLOTS) {
              stage[cr * SYM_SLOTS + slot] = key;
            } else {
              const int c = c0 + cr;
              const int cp = atomicAdd(&a.ccounts[c], 1);
              if (cp < a.ccap) a.cbufs[(int64_t)c * a.ccap + cp] = key;
            }
          }
        }
        continue;
      }
      if (m) {  // query side, as pass 2
        uint32_t take;
        if (!dblk && bound_is_exact<KIND>(tau, key_bound)) {
          take = m & valid_rows(cb, lane_e, a.n_c);
        } else {
          const uint64_t tkey = tkey_l[ql];
          take = 0;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            if (m & (1u << j)) {
              const int c = cb + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
              const bool self = (c == q);
              const float sc = self ? 1.0f : val(j);
              bool tk = (c < a.n_c) && (q < a.n_q) && (sc >= tau) && !(self && a.skip_self);
              if (tk && key_bound) tk = make_key(sc, c, self) >= tkey;
              take |= (tk ? 1u : 0u) << j;
            }
          }
        }
        if (take) {
          // one append per set bit, lowest first: the wave loops as often as
          // its fullest lane has bits, not once per value
          int pos = atomicAdd(&cnt[ql], __popc(take));
          uint64_t* dst = a.bufs + ((int64_t)split * a.n_q + q) * a.caps;
          while (take) {
            const int j = __builtin_ctz(take);
            take &= take - 1;
            const int c = cb + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
            const bool self = (c == q);
            const float sc = pick(j);
            if (pos < a.caps) dst[pos] = make_key(self ? 1.0f : sc, c, self);
            pos++;
          }
        }
      }
      if (mc) {  // candidate side: the pair seen from row c (never self here)
        uint32_t takec;
        if (c_fast) {
          takec = mc;  // rows past the end have infinite bounds
        } else {
          takec = 0;
#pragma unroll
          for (int j = 0; j < 16; j++) {
            if (mc & (1u << j)) {
              const int cr = cb - c0 + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
              const float sc = val(j);
              bool tk = (c0 + cr < a.n_c) && (sc >= ctau_l[cr]);
              if (tk && key_bound) tk = make_key(sc, q, false) >= ckey_l[cr];
              takec |= (tk ? 1u : 0u) << j;
            }
          }
        }
        while (takec) {
          const int j = __builtin_ctz(takec);
          takec &= takec - 1;
          const int cr = cb - c0 + 8 * (j >> 2) + 4 * (lane_e >> 5) + (j & 3);
          const uint64_t key = make_key(pick(j), q, false);
          const int slot = atomicAdd(&scnt[cr], 1);
          if (slot < SYM_SLOTS) {
            stage[cr * SYM_SLOTS + slot] = key;
          } else {
            const int c = c0 + cr;
            const int pos = atomicAdd(&a.ccounts[c], 1);
            if (pos < a.ccap) a.cbufs[(int64_t)c * a.ccap + pos] = key;
          }
        }
      }
    }
    // every wave's staged keys are in, every wave is done reading this tile's
    // bounds, and every wave's loads of the next tile's first stages have
    // landed (on the diagonal tile too, which has no flush: the next tile
    // relies on this barrier)
    __syncthreads();
    if (!diag) {
      for (int i = threadIdx.x; i < T; i += En::THREADS) {
        const int k = min(scnt[i], SYM_SLOTS);
        const int c = c0 + i;
        if (k > 0 && c < a.n_c) {
          const int pos = atomicAdd(&a.ccounts[c], k);
          for (int j = 0; j < k; j++)
            if (pos + j < a.ccap) a.cbufs[(int64_t)c * a.ccap + pos + j] = stage[i * SYM_SLOTS + j];
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < T; i += En::THREADS) {
    const int q = q0 + i;
    if (q < a.n_q) a.counts[(int64_t)split * a.n_q + q] = cnt[i];
  }
}

__global__ void zero_list_kernel(int32_t* __restrict__ p, const int32_t* __restrict__ list, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[list[i]] = 0;
}

// ---------------------------------------------------------------------------
// Binary Jaccard tiles (uint8.rs:21-36) on the VALU: 16 x 16 threads, each
// owning 8 candidates x 8 queries; popcount of AND per 32-bit word, union from
// the per-row popcounts.  Same two epilogues as the MFMA kernels.
// ---------------------------------------------------------------------------
constexpr int TILE = 128;                   // Jaccard tile rows (both operands)
constexpr int JAC_PITCH = STAGE_BYTES + 4;  // LDS row pitch in bytes (word-padded)

__device__ __forceinline__ void jac_tile(uint32_t (&inter)[8][8], uint8_t* lds, const TileArgs& a,
                                         int c0, int q0) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) inter[i][j] = 0;
  const int nst = (int)(a.row_bytes / STAGE_BYTES);
  uint8_t* A = lds;
  uint8_t* B = lds + TILE * JAC_PITCH;
  for (int s = 0; s < nst; s++) {
    // 128 rows x 128 B per operand: 8 x 16 B per row, 4 chunks per thread
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int id = threadIdx.x + 256 * i;
      int row = id >> 3, c16 = id & 7;
      int cg = c0 + row, qg = q0 + row;
      i32x4 va = i32x4{0, 0, 0, 0}, vb = i32x4{0, 0, 0, 0};
      if (cg < a.n_c)
        va = *(const i32x4*)(a.c_rows + (int64_t)cg * a.c_stride * a.row_bytes + s * STAGE_BYTES + c16 * 16);
      if (qg < a.n_q)
        vb = *(const i32x4*)(a.q_rows + (int64_t)qg * a.row_bytes + s * STAGE_BYTES + c16 * 16);
      uint32_t* pa = (uint32_t*)(A + row * JAC_PITCH + c16 * 16);
      uint32_t* pb = (uint32_t*)(B + row * JAC_PITCH + c16 * 16);
      pa[0] = va[0]; pa[1] = va[1]; pa[2] = va[2]; pa[3] = va[3];
      pb[0] = vb[0]; pb[1] = vb[1]; pb[2] = vb[2]; pb[3] = vb[3];
    }
    __syncthreads();
#pragma unroll 4
    for (int w = 0; w < STAGE_BYTES / 4; w++) {
      uint32_t av[8], bv[8];
#pragma unroll
      for (int i = 0; i < 8; i++) av[i] = *(const uint32_t*)(A + (ty + 16 * i) * JAC_PITCH + 4 * w);
#pragma unroll
      for (int j = 0; j < 8; j++) bv[j] = *(const uint32_t*)(B + (tx + 16 * j) * JAC_PITCH + 4 * w);
#pragma unroll
      for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) inter[i][j] += __popc(av[i] & bv[j]);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ float jac_score(uint32_t inter, uint32_t pa, uint32_t pb) {
  uint32_t uni = pa + pb - inter;
  return uni > 0 ? __fdiv_rn((float)inter, (float)uni) : 0.0f;
}

// Whether the Jaccard score inter / uni (IEEE quotient, uni counts below 2^24)
// may reach the bound tau, without the division: false only when
// inter < fl(fl(tau uni) (1 - 2^-20)), i.e. inter < tau uni (1 - 2^-21), where
// the rounded quotient is below tau.  Bounds <= 0, -inf and NaN products never
// rule a pair out; +inf rules every pair out.  The exact test follows for the
// pairs it lets through.
__device__ __forceinline__ bool jac_may_reach(uint32_t inter, uint32_t uni, float tau) {
  return !((float)inter < tau * (float)uni * (1.0f - 0x1p-20f));
}

// An integer count below which no pair with one row of `pop` bits reaches the
// bound tau, for a test of one compare per pair ahead of jac_may_reach: the
// union of that row with any other is at least pop, so for tau > 0 the
// rounded products keep t = fl(fl(tau pop) (1 - 2^-20)) at or below
// jac_may_reach's fl(fl(tau uni) (1 - 2^-20)), and inter < ceil(t) gives
// inter < t: jac_may_reach is false.  Bounds <= 0 or NaN (and pop 0 with an
// infinite bound) give 0, which rules out no pair; t beyond the int range
// (tau = +inf) rules out every count.
__device__ __forceinline__ int jac_quick_bound(float tau, uint32_t pop) {
  const float t = tau * (float)pop * (1.0f - 0x1p-20f);
  if (!(t > 0.0f)) return 0;
  if (t >= 2147483520.0f) return INT_MAX;
  return (int)ceilf(t);
}

__global__ __launch_bounds__(256) void jac_pass1_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  uint32_t inter[8][8];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int qb, split;
  if (!tile_coords(a, qb, split)) return;
  const int q0 = qb * TILE;
  const int t_begin = split * a.tiles_per_split;
  const int t_end = min(a.n_ctiles, t_begin + a.tiles_per_split);
  for (int t = t_begin; t < t_end; t++) {
    const int c0 = t * TILE;
    jac_tile(inter, lds, a, c0, q0);
    // the popcounts of the thread's 8 candidate rows, once per tile
    uint32_t pc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int c = c0 + ty + 16 * i;
      pc[i] = c < a.n_c ? a.row_pop[(int64_t)c * a.c_stride] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int q = q0 + tx + 16 * j;
      if (q >= a.n_q) continue;
      const uint32_t pq = a.row_pop[q + a.q_offset];
      // The group maximum of the IEEE quotients inter / uni is the quotient of
      // the largest fraction (a correctly rounded division is monotone in its
      // exact value), so the fractions are compared by cross-multiplication
      // and one division is left per group; uni = 0 scores 0, as jac_score.
      uint32_t bn = 0, bd = 1;
      bool any = false, self_in = false;
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int c = c0 + ty + 16 * i;
        const bool self = (int64_t)c * a.c_stride == (int64_t)q + a.q_offset;
        if (c < a.n_c && !(self && a.skip_self)) {
          if (self) {
            self_in = true;
          } else {
            const uint32_t uni = pc[i] + pq - inter[i][j];
            const uint32_t num = uni > 0 ? inter[i][j] : 0u, den = uni > 0 ? uni : 1u;
            if (!any || (uint64_t)num * bd > (uint64_t)bn * den) {
              bn = num;
              bd = den;
            }
            any = true;
          }
        }
      }
      const float m = self_in ? 1.0f : !any ? -INFINITY : bn == 0 ? 0.0f : __fdiv_rn((float)bn, (float)bd);
      // 16 groups (one per ty) of 8 candidates per tile
      a.gmax[(int64_t)q * a.ng + t * 16 + ty] = m;
    }
  }
}

__global__ __launch_bounds__(256) void jac_pass2_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  int* cnt = (int*)(lds + 2 * TILE * JAC_PITCH);
  uint32_t inter[8][8];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int qb, split;
  if (!tile_coords(a, qb, split)) return;
  const int q0 = qb * TILE;
  if (threadIdx.x < TILE) cnt[threadIdx.x] = 0;
  float tau[8];
  uint64_t tkey[8];
  uint32_t pq[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int q = q0 + tx + 16 * j;
    tau[j] = (q < a.n_q && a.tau) ? a.tau[q] : -INFINITY;
    tkey[j] = (q < a.n_q && a.tau_key) ? a.tau_key[q] : 0ull;
    pq[j] = q < a.n_q ? a.row_pop[q + a.q_offset] : 0u;
  }
  const int t_begin = split * a.tiles_per_split;
  const int t_end = min(a.n_ctiles, t_begin + a.tiles_per_split);
  for (int t = t_begin; t < t_end; t++) {
    const int c0 = t * TILE;
    jac_tile(inter, lds, a, c0, q0);
    // (per tile, so that they are not held in registers through the popcounts)
    int tq[8];
#pragma unroll
    for (int j = 0; j < 8; j++) tq[j] = jac_quick_bound(tau[j], pq[j]);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int c = c0 + ty + 16 * i;
      if (c >= a.n_c) continue;
      const uint32_t pc = a.row_pop[c];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int ql = tx + 16 * j;
        const int q = q0 + ql;
        if (q >= a.n_q) continue;
        const bool self = (c == q + a.q_offset);
        // one integer compare rules most pairs out (jac_quick_bound), then
        // jac_may_reach most of the rest, before the division
        if (!self && (int)inter[i][j] < tq[j]) continue;
        if (!self && !jac_may_reach(inter[i][j], pc + pq[j] - inter[i][j], tau[j])) continue;
        const float s = self ? 1.0f : jac_score(inter[i][j], pc, pq[j]);
        bool take = s >= tau[j] && !(self && a.skip_self);
        if (take && a.tau_key) take = make_key(s, c, self) >= tkey[j];
        if (take) {
          int pos = atomicAdd(&cnt[ql], 1);
          if (pos < a.caps) a.bufs[((int64_t)split * a.n_q + q) * a.caps + pos] = make_key(s, c, self);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < TILE) {
    const int q = q0 + threadIdx.x;
    if (q < a.n_q) a.counts[(int64_t)split * a.n_q + q] = cnt[threadIdx.x];
  }
}

// Symmetric Jaccard pass 2 (inter / union is symmetric: the same counts and
// the same IEEE division for (q, c) and (c, q)): work items as pass2_sym_kernel,
// each value tested against the query row's bound (its split buffer, as
// jac_pass2_kernel) and, off the diagonal tile, against the candidate row's
// bound (staged per candidate row in LDS, flushed once per tile).
__global__ __launch_bounds__(256) void jac_pass2_sym_kernel(TileArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  int* cnt = (int*)(lds + 2 * TILE * JAC_PITCH);
  float* ctau_l = (float*)(cnt + TILE);
  int* scnt = (int*)(ctau_l + TILE);
  uint64_t* ckey_l = (uint64_t*)(scnt + TILE);
  uint64_t* stage = ckey_l + TILE;
  uint32_t inter[8][8];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int per_xcd = (int)gridDim.x >> 3;
  const int logical = ((int)blockIdx.x & 7) * per_xcd + ((int)blockIdx.x >> 3);
  if (logical >= a.n_work) return;
  const int4 w = a.work[logical];
  const int q0 = w.x * TILE, split = w.w;
  const bool key_bound = a.tau_key != nullptr;
  if (threadIdx.x < TILE) cnt[threadIdx.x] = 0;
  float tau[8];
  uint64_t tkey[8];
  uint32_t pq[8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int q = q0 + tx + 16 * j;
    tau[j] = (q < a.n_q && a.tau) ? a.tau[q] : -INFINITY;
    tkey[j] = (q < a.n_q && key_bound) ? a.tau_key[q] : 0ull;
    pq[j] = q < a.n_q ? a.row_pop[q] : 0u;
  }
  for (int t = w.y; t < w.z; t++) {
    const int c0 = t * TILE;
    const bool diag = t == w.x;  // the query block's own tile: query side only
    __syncthreads();             // the last tile's flush is done with the staging area
    if (!diag && threadIdx.x < TILE) {
      const int c = c0 + threadIdx.x;
      scnt[threadIdx.x] = 0;
      ctau_l[threadIdx.x] = c < a.n_c ? (a.tau ? a.tau[c] : -INFINITY) : INFINITY;
      ckey_l[threadIdx.x] = (c < a.n_c && key_bound) ? a.tau_key[c] : 0ull;
    }
    jac_tile(inter, lds, a, c0, q0);  // (its barriers publish the bounds)
    // (per tile, so that they are not held in registers through the popcounts)
    int tq[8];
#pragma unroll
    for (int j = 0; j < 8; j++) tq[j] = jac_quick_bound(tau[j], pq[j]);
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int cl = ty + 16 * i;
      const int c = c0 + cl;
      if (c >= a.n_c) continue;
      const uint32_t pc = a.row_pop[c];
      const float ctau = diag ? INFINITY : ctau_l[cl];
      const uint64_t ckey = diag ? 0ull : ckey_l[cl];
      const int tc = jac_quick_bound(ctau, pc);
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int ql = tx + 16 * j;
        const int q = q0 + ql;
        if (q >= a.n_q) continue;
        const bool self = (c == q);
        // most pairs reach neither row's bound: one integer compare per side
        // rules most of them out, jac_may_reach most of the rest, and only
        // what is left takes the division
        if (!self && (int)inter[i][j] < tq[j] && (int)inter[i][j] < tc) continue;
        const uint32_t uni = pc + pq[j] - inter[i][j];
        if (!self && !jac_may_reach(inter[i][j], uni, tau[j]) && !jac_may_reach(inter[i][j], uni, ctau)) continue;
        const float s = self ? 1.0f : jac_score(inter[i][j], pc, pq[j]);
        bool take = s >= tau[j] && !(self && a.skip_self);
        if (take && key_bound) take = make_key(s, c, self) >= tkey[j];
        if (take) {
          int pos = atomicAdd(&cnt[ql], 1);
          if (pos < a.caps) a.bufs[((int64_t)split * a.n_q + q) * a.caps + pos] = make_key(s, c, self);
        }
        // the pair seen from row c (never self off the diagonal tile)
        bool takec = s >= ctau;
        if (takec && key_bound) takec = make_key(s, q, false) >= ckey;
        if (takec) {
          const uint64_t key = make_key(s, q, false);
          const int slot = atomicAdd(&scnt[cl], 1);
          if (slot < SYM_SLOTS) {
            stage[cl * SYM_SLOTS + slot] = key;
          } else {
            const int pos = atomicAdd(&a.ccounts[c], 1);
            if (pos < a.ccap) a.cbufs[(int64_t)c * a.ccap + pos] = key;
          }
        }
      }
    }
    if (!diag) {
      __syncthreads();  // every thread's staged keys are in
      if (threadIdx.x < TILE) {
        const int i = threadIdx.x;
        const int k = min(scnt[i], SYM_SLOTS);
        const int c = c0 + i;
        if (k > 0 && c < a.n_c) {
          const int pos = atomicAdd(&a.ccounts[c], k);
          for (int j = 0; j < k; j++)
            if (pos + j < a.ccap) a.cbufs[(int64_t)c * a.ccap + pos + j] = stage[i * SYM_SLOTS + j];
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < TILE) {
    const int q = q0 + threadIdx.x;
    if (q < a.n_q) a.counts[(int64_t)split * a.n_q + q] = cnt[threadIdx.x];
  }
}

// ---------------------------------------------------------------------------
// Thresholds from group maxima: out_q = the rank-th largest of gmax[q, 0..ng),
// one wave per query, ng <= 64 * E, found by bisection (no sort).
//   rank K: a guaranteed lower bound on the query's K-th best score (K groups
//           each hold a candidate at least that good);
//   rank J < K (the scoring pass's estimate): below the K-th best score unless
//           J of the sampled candidates rank among the query's K - 1 best,
//           which the sample rate makes rare (the host picks J from that
//           binomial tail).  A query it misses ends the scoring pass with
//           fewer than K keys; selection flags it, its guaranteed bound is
//           taken (rank K, for the flagged queries only) and it is scored again.
// ---------------------------------------------------------------------------
template <int E>
__device__ __forceinline__ void tau_select_one(const float* __restrict__ gmax, int ng, int rank,
                                               float* __restrict__ out, int q) {
  const int lane = threadIdx.x & 63;
  uint32_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int g = e * 64 + lane;
    const float m = g < ng ? gmax[(int64_t)q * ng + g] : -INFINITY;
    v[e] = ord_f32(m);
  }
  // the -inf padding past ng decides no count that fixes the rank-th when rank <= ng
  const uint32_t kth = wave_kth_largest<E>(v, rank, ng);
  if (lane == 0) out[q] = (rank <= ng) ? unord_f32(kth) : -INFINITY;
}

// queries 0 .. n_q - 1, or list[0 .. n_q) when a list is given
template <int E>
__global__ __launch_bounds__(256) void tau_select_kernel(const float* __restrict__ gmax, int n_q, int ng, int rank,
                                                         float* __restrict__ out, const int32_t* __restrict__ list) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n_q) return;
  tau_select_one<E>(gmax, ng, rank, out, list ? list[i] : i);
}

// ---------------------------------------------------------------------------
// Pass 3: per query, stream the buffered keys of all splits through a wave
// sort, keep K.  Overflow (a split's buffer dropped keys) marks the query and
// stores the K-th kept key as a tighter bound for the next round:
// the K-th best of any subset is <= the true K-th best.
// ---------------------------------------------------------------------------
// Keys of query q from all splits, streamed through a 64*E-key wave sort: each
// round loads as many unread keys as fit beside the K best so far, sorts, and
// keeps the K best.  Writes the K best.
template <int E>
__device__ __forceinline__ void select_body(const SelectArgs& a, int q, bool overflow, int total) {
  const int lane = threadIdx.x & 63;
  constexpr int CAP = 64 * E;
  uint64_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) v[e] = EVOC_KEY_NONE;
  int keep = 0;  // best keys so far, at positions [0, keep)
  int g0 = 0;    // index (over all splits, in split order) of the next unread key
  do {
    const int take = min(CAP - keep, total - g0);
    // keys g0 .. g0 + take - 1 go to positions keep .. keep + take - 1
    int base = 0;  // index of the current segment's first key
    auto segment = [&](const uint64_t* src, int c) __attribute__((always_inline)) {
      const int lo = max(base, g0), hi = min(base + c, g0 + take);
      if (lo < hi) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int gi = g0 + (e * 64 + lane - keep);
          if (e * 64 + lane >= keep && gi >= lo && gi < hi) v[e] = src[gi - base];
        }
      }
      base += c;
    };
    for (int s = 0; s < a.n_splits; s++)
      segment(a.bufs + ((int64_t)s * a.n_q + q) * a.caps, min(a.counts[(int64_t)s * a.n_q + q], a.caps));
    if (a.cbufs) segment(a.cbufs + (int64_t)q * a.ccap, min(a.ccounts[q], a.ccap));
#pragma unroll
    for (int e = 0; e < E; e++)
      if (e * 64 + lane >= keep + take) v[e] = EVOC_KEY_NONE;
    wave_sort_desc<E>(v);
    keep = min(a.K, keep + take);
    g0 += take;
  } while (g0 < total);
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * 64 + lane;
    if (i < a.K) a.out_keys[(int64_t)q * a.K + i] = v[e];
  }
  if (lane == 0) a.overflow[q] = overflow ? 1 : 0;
  // the K-th kept key bounds the true K-th key from below (kept keys are a
  // subset of all candidates); it drives the next round for this query block
  const uint64_t kth = wave_elem<E>(v, a.K - 1);
  if (lane == 0) a.tau_key_out[q] = kth;
}

// Keys of query q from all its segments (the split buffers, then the
// candidate-side buffer), total <= 64 * E, held register-major.  Lane s holds
// segment s's kept count `kept`, its first key's position `base` and its
// address `src`.  The K-th largest key is found by bisection, the keys above it
// are compacted into the wave's LDS slice (with the K-th key repeated up to K:
// keys are distinct, so it repeats only as the empty key when fewer than K
// exist) and only those K are sorted.  Same output as sorting all keys and
// keeping K.
template <int E, int EK>
__device__ __forceinline__ void select_bisect(const SelectArgs& a, int q, bool overflow, int total, int nseg,
                                              int kept, int base, const uint64_t* src, uint64_t* sel) {
  const int lane = threadIdx.x & 63;
  // the segments pick each kept key's address; the loads then go out together
  // (a load per segment into the same register would wait for the previous one)
  uint64_t addr[E];
#pragma unroll
  for (int e = 0; e < E; e++) addr[e] = 0;
  for (int sg = 0; sg < nseg; sg++) {
    const int c = __builtin_amdgcn_readlane(kept, sg);
    if (c == 0) continue;
    const int b0 = __builtin_amdgcn_readlane(base, sg);
    // (both halves zero-extended: readlane returns int)
    const uint64_t p = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)src, sg) |
                       ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)src >> 32), sg) << 32);
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int i = e * 64 + lane;
      if (i >= b0 && i < b0 + c) addr[e] = p + (uint64_t)(i - b0) * sizeof(uint64_t);
    }
  }
  uint64_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) v[e] = addr[e] ? *(const uint64_t*)addr[e] : EVOC_KEY_NONE;
  const uint64_t kth = wave_kth_largest<E>(v, a.K, total);
  // keys above the K-th, in lane order, then the K-th up to K
  int above = 0;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const bool up = v[e] > kth;
    const uint64_t m = __ballot(up);
    if (up) sel[above + __popcll(m & ((1ull << lane) - 1ull))] = v[e];
    above += __popcll(m);
  }
  for (int i = above + lane; i < a.K; i += 64) sel[i] = kth;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint64_t w[EK];
#pragma unroll
  for (int e = 0; e < EK; e++) {
    const int i = e * 64 + lane;
    w[e] = i < a.K ? sel[i] : EVOC_KEY_NONE;
  }
  wave_sort_desc<EK>(w);
#pragma unroll
  for (int e = 0; e < EK; e++) {
    const int i = e * 64 + lane;
    if (i < a.K) a.out_keys[(int64_t)q * a.K + i] = w[e];
  }
  if (lane == 0) {
    a.overflow[q] = overflow ? 1 : 0;
    // the K-th kept key bounds the true K-th key from below (kept keys are a
    // subset of all candidates); it drives the next round for this query block
    a.tau_key_out[q] = kth;
  }
}

// Segment counts of query q in one vector load: lane s holds segment s (split
// buffer s, then the candidate-side buffer), with its kept count, position of
// its first key among all kept keys, and address.  nseg <= 64.
struct SegLane {
  int kept, base, total;
  bool overflow;
  const uint64_t* src;
};
__device__ __forceinline__ SegLane segment_lanes(const SelectArgs& a, int q) {
  const int lane = threadIdx.x & 63;
  int c = 0, cap = 0;
  const uint64_t* src = nullptr;
  if (lane < a.n_splits) {
    c = a.counts[(int64_t)lane * a.n_q + q];
    cap = a.caps;
    src = a.bufs + ((int64_t)lane * a.n_q + q) * a.caps;
  } else if (lane == a.n_splits && a.cbufs) {
    c = a.ccounts[q];
    cap = a.ccap;
    src = a.cbufs + (int64_t)q * a.ccap;
  }
  SegLane r;
  r.overflow = __ballot(c > cap) != 0;
  if (a.min_count > 0) {
    // every appended key counts, kept or not: fewer than min_count means the
    // query's estimated bound was above its true K-th score
    int all = c;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) all += __shfl_xor(all, o, 64);
    r.overflow = r.overflow || all < a.min_count;
  }
  r.kept = min(c, cap);
  int incl = r.kept;  // inclusive prefix over the lanes (segment order)
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  r.base = incl - r.kept;
  r.total = __shfl(incl, 63, 64);
  r.src = src;
  return r;
}

constexpr int SELECT_REG_KEYS = 2048;  // keys a query may have for the register path

__device__ __forceinline__ int select_query(const SelectArgs& a, int qi) {
  if (a.query_list) return qi < a.n_list ? a.query_list[qi] : -1;
  return qi < a.n_q ? qi : -1;
}

// One wave per query, K <= 64 * EK: queries with at most 64 segments and
// SELECT_REG_KEYS kept keys (the usual case), their keys held in registers
// (width from the query's own key count, a wave-uniform branch) and selected
// by bisection.  select_stream_kernel takes the others.
template <int EK>
__device__ __forceinline__ void select_topk_one(const SelectArgs& a, int q, uint64_t* sel) {
  const int nseg = a.n_splits + (a.cbufs ? 1 : 0);
  if (nseg > 64) return;
  const SegLane g = segment_lanes(a, q);
  const int total = g.total;
  if (total <= 64) select_bisect<1, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
  else if (total <= 128) select_bisect<2, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
  else if (total <= 256) select_bisect<4, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
  else if (total <= 512) select_bisect<8, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
  else if (total <= 1024) select_bisect<16, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
  else if (total <= SELECT_REG_KEYS) select_bisect<32, EK>(a, q, g.overflow, total, nseg, g.kept, g.base, g.src, sel);
}

template <int EK>
__global__ __launch_bounds__(256) void select_topk_kernel(SelectArgs a) {
  extern __shared__ uint64_t sel_lds[];
  const int q = select_query(a, blockIdx.x * 4 + (threadIdx.x >> 6));
  if (q < 0) return;
  select_topk_one<EK>(a, q, sel_lds + (threadIdx.x >> 6) * 64 * EK);
}

// The queries select_topk_kernel leaves: their keys streamed through a
// 2048-key wave sort.
__device__ __forceinline__ void select_stream_one(const SelectArgs& a, int q) {
  const int nseg = a.n_splits + (a.cbufs ? 1 : 0);
  if (nseg <= 64) {
    const SegLane g = segment_lanes(a, q);
    if (g.total > SELECT_REG_KEYS) select_body<32>(a, q, g.overflow, g.total);
    return;
  }
  int total = 0, all = 0;
  bool overflow = false;
  for (int s = 0; s < a.n_splits; s++) {
    const int c = a.counts[(int64_t)s * a.n_q + q];
    overflow |= c > a.caps;
    total += min(c, a.caps);
    all += c;
  }
  if (a.cbufs) {
    const int c = a.ccounts[q];
    overflow |= c > a.ccap;
    total += min(c, a.ccap);
    all += c;
  }
  if (a.min_count > 0 && all < a.min_count) overflow = true;
  select_body<32>(a, q, overflow, total);
}

__global__ __launch_bounds__(256) void select_stream_kernel(SelectArgs a) {
  const int q = select_query(a, blockIdx.x * 4 + (threadIdx.x >> 6));
  if (q < 0) return;
  select_stream_one(a, q);
}

// ---------------------------------------------------------------------------
// FP32 re-score of the BF16 survivors with a fixed reduction order: lane l
// accumulates elements l, l+64, ... with fmaf, then a butterfly sum.  The k
// best by exact FP32 score are kept (Kc <= 64*E candidates per query).
// ---------------------------------------------------------------------------
template <int E>
__device__ __forceinline__ void rescore_one(const float* __restrict__ X, int64_t ld, int d, int q_offset,
                                            const uint64_t* __restrict__ cand_keys, int Kc, int k,
                                            uint64_t* __restrict__ out_keys, const int32_t* __restrict__ order,
                                            int qi) {
  // four candidates at a time, one 16-lane group each; lane t of a group sums
  // the float4 chunks t, t + 16, ... of the row with fmaf (fixed order), then a
  // fixed xor-butterfly over the 16 lanes.  Queries are taken in `order` when
  // given (a permutation of the slab's queries: only the order of the work
  // changes, each query's result is its own).  A lane issues all its chunk
  // loads of a candidate (panels of CH chunks, one panel up to 1024 floats)
  // before the fmaf chain, holds the query's chunks in registers when the row
  // is one panel, and takes the query's candidate keys with one load per lane.
  constexpr int CH = 16;
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 4, t = lane & 15;
  const int ql = order ? order[qi] : qi;
  const int q = ql + q_offset;
  const float* xq = X + (int64_t)q * ld;
  const float4* xq4 = (const float4*)xq;
  const int nv4 = d >> 2;  // full float4 chunks; ld is a multiple of 32 and rows are zero padded
  const int tail = d & 3;
  const bool one_panel = nv4 <= 16 * CH;
  float4 qa[CH];
#pragma unroll
  for (int u = 0; u < CH; u++) {
    const int j = t + 16 * u;
    qa[u] = (one_panel && j < nv4) ? xq4[j] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  uint64_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) v[e] = EVOC_KEY_NONE;
  // candidate keys: key i in register i / 64 of lane i % 64
  uint64_t ck[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * 64 + lane;
    ck[e] = i < Kc ? cand_keys[(int64_t)ql * Kc + i] : EVOC_KEY_NONE;
  }
  for (int i0 = 0; i0 < Kc; i0 += 4) {
    const int i = i0 + grp;
    // (the broadcast runs in every lane: a shuffle from a lane switched off
    // by a branch would read nothing defined)
    const uint64_t kb = wave_elem<E>(ck, min(i, Kc - 1));
    const uint64_t key = i < Kc ? kb : EVOC_KEY_NONE;
    const int c = key != EVOC_KEY_NONE ? key_index(key, q) : q;
    const float* xc = X + (int64_t)c * ld;
    const float4* xc4 = (const float4*)xc;
    float part = 0.0f;
    for (int p0 = 0; p0 < nv4; p0 += 16 * CH) {
      float4 b[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int j = p0 + t + 16 * u;
        b[u] = j < nv4 ? xc4[j] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int j = p0 + t + 16 * u;
        if (j < nv4) {
          const float4 a4 = one_panel ? qa[u] : xq4[j];
          part = fmaf(a4.x, b[u].x, part);
          part = fmaf(a4.y, b[u].y, part);
          part = fmaf(a4.z, b[u].z, part);
          part = fmaf(a4.w, b[u].w, part);
        }
      }
    }
    if (tail && t == 0)
      for (int r = 4 * nv4; r < d; r++) part = fmaf(xq[r], xc[r], part);
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) part += __shfl_xor(part, m, 64);
    const bool self = (c == q);
    const uint64_t nk = key == EVOC_KEY_NONE ? EVOC_KEY_NONE : make_key(self ? 1.0f : f32_score(part), c, self);
    // candidate i0 + g lives in lanes 16g..16g+15; element i of the sort array
    // sits in register i / 64 of lane i % 64
#pragma unroll
    for (int g2 = 0; g2 < 4; g2++) {
      const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)nk, 16 * g2, 64);
      const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(nk >> 32), 16 * g2, 64);
      const int ii = i0 + g2;
#pragma unroll
      for (int e = 0; e < E; e++)
        if (e * 64 + lane == ii && ii < Kc) v[e] = ((uint64_t)hi << 32) | lo;
    }
  }
  wave_sort_desc<E>(v);
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * 64 + lane;
    if (i < k) out_keys[(int64_t)ql * k + i] = v[e];
  }
}

template <int E>
__global__ __launch_bounds__(256) void rescore_kernel(const float* __restrict__ X, int64_t ld, int d,
                                                      int n_q, int q_offset,
                                                      const uint64_t* __restrict__ cand_keys, int Kc,
                                                      int k, uint64_t* __restrict__ out_keys,
                                                      const int32_t* __restrict__ order) {
  const int qi = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (qi >= n_q) return;
  rescore_one<E>(X, ld, d, q_offset, cand_keys, Kc, k, out_keys, order, qi);
}

// ---------------------------------------------------------------------------
// Keys -> (indices, transformed distances) (src/knn.rs:156-187).
// ---------------------------------------------------------------------------
// key i of query q (global index) -> out_idx[i], out_dist[i]
__device__ __forceinline__ void finalize_elem(const uint64_t* __restrict__ keys, int64_t i, int q, int dtype,
                                              int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  const uint64_t key = keys[i];
  if (key == EVOC_KEY_NONE) {
    out_idx[i] = -1;
    out_dist[i] = 0.0f;  // an empty slot holds raw INF (f32::MAX), which maps to 0
    return;
  }
  const int c = key_index(key, q);
  const float score = key_score(key);
  const float raw = -score;  // f32 self: -1.0; FLT_MIN rule: -(-FLT_MIN)
  out_idx[i] = c;
  float d;
  if (dtype == EVOC_DTYPE_F32) {
    if (raw >= 0.0f) {
      d = 0.0f;
    } else {
      const double sim = (double)(-raw);
      if (sim >= 1.0) {
        d = 0.0f;
      } else {
        const double t = -log2(sim);
        d = (float)(t > 0.0 ? t : 0.0);
      }
    }
  } else if (dtype == EVOC_DTYPE_I8) {
    d = raw >= 0.0f ? 0.0f : __fdiv_rn(1.0f, -raw);
  } else {
    d = raw >= 0.0f ? 0.0f : (float)(-log2((double)(-raw)));
  }
  out_dist[i] = d;
}

__global__ void finalize_kernel(const uint64_t* __restrict__ keys, int n_q, int q_offset, int k,
                                int dtype, int32_t* __restrict__ out_idx,
                                float* __restrict__ out_dist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_q * k) return;
  finalize_elem(keys, i, (int)(i / k) + q_offset, dtype, out_idx, out_dist);
}

// ---------------------------------------------------------------------------
// Batched exact search: the passes above over many requests per launch.  Each
// kernel is persistent: a fixed grid walks its work-item list (device memory)
// with a stride of the grid, so the same launches serve any batch.  Per item
// the work is exactly the single-request kernel's, with that request's
// arguments from the table, so every request's keys are the ones it gets alone.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void batch_begin_kernel(BatchReq* __restrict__ reqs,
                                                          const BatchHeader* __restrict__ hdr) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < hdr->n_reqs; r += gridDim.x * blockDim.x) {
    *reqs[r].ovf = 0;
    if (reqs[r].flag) *reqs[r].flag = 0;
  }
}

template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__global__ __launch_bounds__(64 * WM * WN, BM * BN == 16 ? 1 : 2) void pass1_batch_kernel(const BatchReq* __restrict__ reqs,
                                                                   const int4* __restrict__ items,
                                                                   const BatchHeader* __restrict__ hdr) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int n = hdr->n_p1;
  const int4* it = items + hdr->off_p1;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    __syncthreads();  // the previous item's waves are done with the LDS
    pass1_work<KIND, BM, BN, WM, WN>(reqs[w.x].p1, w.y, w.z, lds);
  }
}

template <int KIND, int BM, int BN, int WM, int WN, int SB = MFMA_SB>
__global__ __launch_bounds__(64 * WM * WN, BM * BN == 16 ? 1 : 2) void pass2_batch_kernel(const BatchReq* __restrict__ reqs,
                                                                   const int4* __restrict__ items,
                                                                   const BatchHeader* __restrict__ hdr) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int n = hdr->n_p2;
  const int4* it = items + hdr->off_p2;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    __syncthreads();  // the previous item's counters and bounds are written out
    pass2_work<KIND, BM, BN, WM, WN>(reqs[w.x].p2, w.y, w.z, lds);
  }
}

// per-query stages: one wave per query, four queries per item
__global__ __launch_bounds__(256) void tau_batch_kernel(const BatchReq* __restrict__ reqs,
                                                        const int4* __restrict__ items,
                                                        const BatchHeader* __restrict__ hdr) {
  const int n = hdr->n_q;
  const int4* it = items + hdr->off_q;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    const int q = w.y + (int)(threadIdx.x >> 6);
    if (r.has_p1 && q < r.n) tau_select_one<32>(r.gmax, r.ng, r.K, r.tau, q);
  }
}

__global__ __launch_bounds__(256) void select_batch_kernel(const BatchReq* __restrict__ reqs,
                                                           const int4* __restrict__ items,
                                                           const BatchHeader* __restrict__ hdr) {
  __shared__ uint64_t sel[4 * 64];  // K <= 64: one 64-key slice per wave
  const int n = hdr->n_q;
  const int4* it = items + hdr->off_q;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    const int q = w.y + (int)(threadIdx.x >> 6);
    if (q >= r.n) continue;
    // the register path takes a query with at most 64 segments and
    // SELECT_REG_KEYS keys, the streaming path every other one
    select_topk_one<1>(r.sel, q, sel + (threadIdx.x >> 6) * 64);
    select_stream_one(r.sel, q);
    if ((threadIdx.x & 63) == 0 && r.sel.overflow[q]) atomicOr(r.ovf, 1);
  }
}

__global__ __launch_bounds__(256) void rescore_batch_kernel(const BatchReq* __restrict__ reqs,
                                                            const int4* __restrict__ items,
                                                            const BatchHeader* __restrict__ hdr) {
  const int n = hdr->n_q;
  const int4* it = items + hdr->off_q;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    const int q = w.y + (int)(threadIdx.x >> 6);
    if (r.bf16 && q < r.n) rescore_one<1>(r.xf32, r.f32_ld, r.d, 0, r.sel.out_keys, r.K, r.k, r.rkeys, nullptr, q);
  }
}

__global__ __launch_bounds__(256) void finalize_batch_kernel(const BatchReq* __restrict__ reqs,
                                                             const int4* __restrict__ items,
                                                             const BatchHeader* __restrict__ hdr) {
  const int n = hdr->n_q;
  const int4* it = items + hdr->off_q;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    const int q = w.y + (int)(threadIdx.x >> 6);
    if (q >= r.n) continue;
    // the BF16 form outputs its re-scored keys (k per query), the others the
    // selected ones (K = k per query)
    const uint64_t* keys = r.bf16 ? r.rkeys : r.sel.out_keys;
    for (int j = (int)(threadIdx.x & 63); j < r.k; j += 64)
      finalize_elem(keys, (int64_t)q * r.k + j, q, r.dtype, r.out_idx, r.out_dist);
  }
}

}  // namespace evoc

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
namespace evoc {

// MFMA tile shapes (see Engine): 0 = 4 waves of 2x2 blocks (128 x 128 tile),
// 1 = 4 waves of 2x4 (128 x 256), 2 = 8 waves of 2x4 (256 x 256), 3 = 4 waves
// of 4x4 (256 x 256, accumulators in AGPRs, least LDS traffic per MFMA).  The shape
// in force is the calling thread's (set per search by the host, which takes it
// from the autotuner), unless EVOC_MI355X_TILE pins one for the process.  The
// shape never changes results: every shape accumulates each pair's products in
// the same MFMA K order.
static int g_env_shape = -2;
static thread_local int t_shape = -1;

int tile_shape(int kind) {
  if (kind == EVOC_ELEM_U8BITS) return -1;
  if (g_env_shape == -2) {
    const char* e = getenv("EVOC_MI355X_TILE");
    g_env_shape = e ? atoi(e) : -1;
  }
  if (g_env_shape >= 0 && g_env_shape <= 4) return g_env_shape;
  if (t_shape >= 0 && t_shape <= 4) return t_shape;
  return 2;
}

void set_tile_shape(int shape) { t_shape = shape; }

static void shape_dims(int kind, int& tc, int& tq, int& gpt) {
  if (kind == EVOC_ELEM_U8BITS) {
    tc = tq = TILE;
    gpt = 16;
    return;
  }
  switch (tile_shape(kind)) {
    case 1: tc = 128; tq = 256; gpt = 8; return;
    case 2: tc = 256; tq = 256; gpt = 16; return;
    case 3: case 4: tc = 256; tq = 256; gpt = 16; return;
    default: tc = 128; tq = 128; gpt = 8; return;
  }
}

int tile_rows_c(int kind) { int tc, tq, g; shape_dims(kind, tc, tq, g); return tc; }
int tile_rows_q(int kind) { int tc, tq, g; shape_dims(kind, tc, tq, g); return tq; }
int groups_per_tile(int kind) { int tc, tq, g; shape_dims(kind, tc, tq, g); return g; }

// LDS bytes of the MFMA tile kernels of `shape` for pass 1 / 2 / 3
size_t tile_lds_bytes_shape(int shape, int pass) {
  int tc, tq;
  switch (shape) {
    case 1: tc = 128; tq = 256; break;
    case 2: case 3: case 4: tc = 256; tq = 256; break;
    default: tc = 128; tq = 128; break;
  }
  const int sb = shape == 2 ? 128 : MFMA_SB;
  const size_t ring = (size_t)mfma_ring(tc + tq, sb) * (tc + tq) * sb;
  if (pass == 3) return ring + (size_t)tq * 40 + (size_t)tc * SYM_SLOTS * 8 + 16 + (size_t)(tq + tc) * 4;
  return ring + (pass == 2 ? (size_t)tq * (4 + 4 + 4 + 8) : 0);
}

size_t tile_lds_bytes(int kind, int pass) {
  if (kind == EVOC_ELEM_U8BITS) {
    // pass 2: per-query counters; symmetric pass (3): also the candidate rows'
    // bounds, staged counts, key bounds and SYM_SLOTS staged keys per row
    if (pass == 3) return 2 * TILE * JAC_PITCH + TILE * (4 + 4 + 4 + 8) + TILE * SYM_SLOTS * 8;
    return 2 * TILE * JAC_PITCH + (pass == 2 ? TILE * sizeof(int) : 0);
  }
  // pass 2 adds per-query counters, score bounds, fast bounds and key bounds;
  // the symmetric pass (3) adds the candidate rows' bounds, staged counts,
  // SYM_SLOTS staged keys per candidate row and four exactness flags
  return tile_lds_bytes_shape(tile_shape(kind), pass);
}

template <class Kern>
static hipError_t launch_tile(Kern kern, size_t lds, dim3 grid, hipStream_t s, const TileArgs& a,
                              int threads = 256) {
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, dim3(threads), lds, s, a);
  return hipGetLastError();
}

int tile_threads(int kind) { return tile_shape(kind) == 2 || tile_shape(kind) == 4 ? 512 : 256; }  // shape 3: 4 waves

template <int KIND>
static hipError_t launch_pass1_kind(const TileArgs& a, dim3 g, size_t lds, hipStream_t s) {
  switch (tile_shape(KIND)) {
    case 1: return launch_tile(pass1_groupmax_kernel<KIND, 2, 4, 2, 2>, lds, g, s, a);
    case 2: return launch_tile(pass1_groupmax_kernel<KIND, 2, 4, 4, 2, 128>, lds, g, s, a, 512);
    case 3: return launch_tile(pass1_groupmax_kernel<KIND, 4, 4, 2, 2>, lds, g, s, a);
    case 4: return launch_tile(pass1_groupmax_kernel<KIND, 2, 4, 4, 2, 64>, lds, g, s, a, 512);
    default: return launch_tile(pass1_groupmax_kernel<KIND, 2, 2, 2, 2>, lds, g, s, a);
  }
}

template <int KIND>
static hipError_t launch_pass2_kind(const TileArgs& a, dim3 g, size_t lds, hipStream_t s) {
  switch (tile_shape(KIND)) {
    case 1: return launch_tile(pass2_append_kernel<KIND, 2, 4, 2, 2>, lds, g, s, a);
    case 2: return launch_tile(pass2_append_kernel<KIND, 2, 4, 4, 2, 128>, lds, g, s, a, 512);
    case 3: return launch_tile(pass2_append_kernel<KIND, 4, 4, 2, 2>, lds, g, s, a);
    case 4: return launch_tile(pass2_append_kernel<KIND, 2, 4, 4, 2, 64>, lds, g, s, a, 512);
    default: return launch_tile(pass2_append_kernel<KIND, 2, 2, 2, 2>, lds, g, s, a);
  }
}

// 1-D grid of n_blocks x n_splits workgroups rounded up to whole XCD rounds
// (tile_coords turns the padding into empty workgroups).
static dim3 tile_grid(TileArgs& a, int n_blocks, int n_splits) {
  a.n_qblocks = n_blocks;
  a.n_splits = n_splits;
  if (a.group_q < 1) a.group_q = 1;
  const int64_t w = (int64_t)n_blocks * n_splits;
  return dim3((unsigned)((w + 7) / 8 * 8));
}

hipError_t launch_pass1(int kind, const TileArgs& a0, int n_blocks, int n_splits, hipStream_t s) {
  const size_t lds = tile_lds_bytes(kind, 1);
  TileArgs a = a0;
  const dim3 g = tile_grid(a, n_blocks, n_splits);
  switch (kind) {
    case EVOC_ELEM_BF16: return launch_pass1_kind<EVOC_ELEM_BF16>(a, g, lds, s);
    case EVOC_ELEM_F32: return launch_pass1_kind<EVOC_ELEM_F32>(a, g, lds, s);
    case EVOC_ELEM_I8: return launch_pass1_kind<EVOC_ELEM_I8>(a, g, lds, s);
    case EVOC_ELEM_U8BITS: return launch_tile(jac_pass1_kernel, lds, g, s, a);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_pass2(int kind, const TileArgs& a0, int n_blocks, int n_splits, hipStream_t s) {
  const size_t lds = tile_lds_bytes(kind, 2);
  TileArgs a = a0;
  const dim3 g = tile_grid(a, n_blocks, n_splits);
  switch (kind) {
    case EVOC_ELEM_BF16: return launch_pass2_kind<EVOC_ELEM_BF16>(a, g, lds, s);
    case EVOC_ELEM_F32: return launch_pass2_kind<EVOC_ELEM_F32>(a, g, lds, s);
    case EVOC_ELEM_I8: return launch_pass2_kind<EVOC_ELEM_I8>(a, g, lds, s);
    case EVOC_ELEM_U8BITS: return launch_tile(jac_pass2_kernel, lds, g, s, a);
  }
  return hipErrorInvalidValue;
}

template <int KIND>
static hipError_t launch_pass2_sym_kind(const TileArgs& a, dim3 g, size_t lds, hipStream_t s) {
  switch (tile_shape(KIND)) {
    case 2: return launch_tile(pass2_sym_kernel<KIND, 2, 4, 4, 2, 128>, lds, g, s, a, 512);
    case 3: return launch_tile(pass2_sym_kernel<KIND, 4, 4, 2, 2>, lds, g, s, a);
    case 4: return launch_tile(pass2_sym_kernel<KIND, 2, 4, 4, 2, 64>, lds, g, s, a, 512);
    case 0: return launch_tile(pass2_sym_kernel<KIND, 2, 2, 2, 2>, lds, g, s, a);
    default: return hipErrorInvalidValue;  // 128 x 256 tiles are not square
  }
}

hipError_t launch_pass2_sym(int kind, const TileArgs& a, hipStream_t s) {
  int tc, tq, g;
  shape_dims(kind, tc, tq, g);
  if (tc != tq) return hipErrorInvalidValue;
  const size_t lds = tile_lds_bytes(kind, 3);
  const dim3 grid((unsigned)(((int64_t)a.n_work + 7) / 8 * 8));
  switch (kind) {
    case EVOC_ELEM_U8BITS: return launch_tile(jac_pass2_sym_kernel, lds, grid, s, a);
    case EVOC_ELEM_BF16: return launch_pass2_sym_kind<EVOC_ELEM_BF16>(a, grid, lds, s);
    case EVOC_ELEM_F32: return launch_pass2_sym_kind<EVOC_ELEM_F32>(a, grid, lds, s);
    case EVOC_ELEM_I8: return launch_pass2_sym_kind<EVOC_ELEM_I8>(a, grid, lds, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_zero_list(int32_t* p, const int32_t* list, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_list_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, list, n);
  return hipGetLastError();
}

static int pow2_per_lane(int entries) {
  int e = 1;
  while (e * 64 < entries) e <<= 1;
  return e;
}

hipError_t launch_tau_select(const float* gmax, int n_q, int ng, int rank, float* out, hipStream_t s,
                             const int32_t* list) {
  if (n_q <= 0) return hipSuccess;
  const dim3 g((unsigned)((n_q + 3) / 4)), b(256);
  switch (pow2_per_lane(ng)) {
#define EVOC_TAU_CASE(E) \
  case E: hipLaunchKernelGGL(tau_select_kernel<E>, g, b, 0, s, gmax, n_q, ng, rank, out, list); break;
    EVOC_TAU_CASE(1)
    EVOC_TAU_CASE(2)
    EVOC_TAU_CASE(4)
    EVOC_TAU_CASE(8)
    EVOC_TAU_CASE(16)
    EVOC_TAU_CASE(32)
#undef EVOC_TAU_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_select(const SelectArgs& a, hipStream_t s) {
  const int nq = a.query_list ? a.n_list : a.n_q;
  const dim3 g((unsigned)((nq + 3) / 4)), b(256);
  const int ek = pow2_per_lane(a.K);
  const size_t lds = (size_t)4 * 64 * ek * sizeof(uint64_t);  // one K-key slice per wave
  switch (ek) {
    case 1: hipLaunchKernelGGL(select_topk_kernel<1>, g, b, lds, s, a); break;
    case 2: hipLaunchKernelGGL(select_topk_kernel<2>, g, b, lds, s, a); break;
    case 4: hipLaunchKernelGGL(select_topk_kernel<4>, g, b, lds, s, a); break;
    case 8: hipLaunchKernelGGL(select_topk_kernel<8>, g, b, lds, s, a); break;
    case 16: hipLaunchKernelGGL(select_topk_kernel<16>, g, b, lds, s, a); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(select_stream_kernel, g, b, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_rescore(const float* X, int64_t ld, int d, int n_q, int q_offset,
                          const uint64_t* cand, int Kc, int k, uint64_t* out, hipStream_t s,
                          const int32_t* order) {
  const dim3 g((unsigned)((n_q + 3) / 4)), b(256);
  switch (pow2_per_lane(Kc)) {
    case 1: hipLaunchKernelGGL(rescore_kernel<1>, g, b, 0, s, X, ld, d, n_q, q_offset, cand, Kc, k, out, order); break;
    case 2: hipLaunchKernelGGL(rescore_kernel<2>, g, b, 0, s, X, ld, d, n_q, q_offset, cand, Kc, k, out, order); break;
    case 4: hipLaunchKernelGGL(rescore_kernel<4>, g, b, 0, s, X, ld, d, n_q, q_offset, cand, Kc, k, out, order); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The batch kernels run 128 x 128 tiles (shape 0: 4 waves, two workgroups per CU).
template <int KIND>
static hipError_t batch_attributes_kind() {
  const int l1 = (int)tile_lds_bytes_shape(0, 1), l2 = (int)tile_lds_bytes_shape(0, 2);
  hipError_t e = hipFuncSetAttribute((const void*)pass1_batch_kernel<KIND, 2, 2, 2, 2>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, l1);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)pass2_batch_kernel<KIND, 2, 2, 2, 2>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, l2);
  return e;
}

hipError_t batch_prepare_attributes(int kind) {
  switch (kind) {
    case EVOC_ELEM_BF16: return batch_attributes_kind<EVOC_ELEM_BF16>();
    case EVOC_ELEM_F32: return batch_attributes_kind<EVOC_ELEM_F32>();
    case EVOC_ELEM_I8: return batch_attributes_kind<EVOC_ELEM_I8>();
  }
  return hipErrorInvalidValue;
}

// Grid of each batch pass: one workgroup per item up to the persistent size
// (two tile workgroups or four query workgroups per CU).  A grid never changes
// what is computed, only how many workgroups wait for a CU: a small batch
// beside a large running search would otherwise queue hundreds of empty ones.
static unsigned item_grid(int items, int cap) { return (unsigned)max(1, min(items, cap)); }

template <int KIND>
static const void* batch_tile_kernel(bool pass1) {
  return pass1 ? (const void*)pass1_batch_kernel<KIND, 2, 2, 2, 2> : (const void*)pass2_batch_kernel<KIND, 2, 2, 2, 2>;
}

static const void* batch_tile_kernel_of(int kind, bool pass1) {
  switch (kind) {
    case EVOC_ELEM_BF16: return batch_tile_kernel<EVOC_ELEM_BF16>(pass1);
    case EVOC_ELEM_F32: return batch_tile_kernel<EVOC_ELEM_F32>(pass1);
    default: return batch_tile_kernel<EVOC_ELEM_I8>(pass1);
  }
}

// the grid this batch gives kernel `f` of the launch sequence (0: not one of them)
static unsigned batch_grid_of(const void* f, int kind, int n_cu, const BatchHeader& h) {
  if (f == (const void*)batch_begin_kernel) return 1;
  if (f == batch_norm_kernel()) return item_grid(h.n_rows, 4 * n_cu);
  if (f == batch_write_kernel()) return item_grid(h.n_rows, 2 * n_cu);
  if (f == batch_tile_kernel_of(kind, true)) return item_grid(h.n_p1, 2 * n_cu);
  if (f == batch_tile_kernel_of(kind, false)) return item_grid(h.n_p2, 2 * n_cu);
  if (f == (const void*)tau_batch_kernel || f == (const void*)select_batch_kernel ||
      f == (const void*)rescore_batch_kernel || f == (const void*)finalize_batch_kernel)
    return item_grid(h.n_q, 4 * n_cu);
  return 0;
}

hipError_t launch_batch(int kind, int n_cu, BatchReq* reqs, const int4* items, const BatchHeader* hdr,
                        const BatchHeader& h, hipStream_t s) {
  if (kind != EVOC_ELEM_BF16 && kind != EVOC_ELEM_F32 && kind != EVOC_ELEM_I8) return hipErrorInvalidValue;
  const dim3 b(256), gq(item_grid(h.n_q, 4 * n_cu));
  hipLaunchKernelGGL(batch_begin_kernel, dim3(1), b, 0, s, reqs, hdr);
  hipError_t e = launch_batch_prepare(item_grid(h.n_rows, 4 * n_cu), item_grid(h.n_rows, 2 * n_cu), reqs, items,
                                      hdr, s);
  if (e != hipSuccess) return e;
  const void* args[3] = {&reqs, &items, &hdr};
  e = hipLaunchKernel(batch_tile_kernel_of(kind, true), dim3(item_grid(h.n_p1, 2 * n_cu)), b, (void**)args,
                      tile_lds_bytes_shape(0, 1), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(tau_batch_kernel, gq, b, 0, s, reqs, items, hdr);
  e = hipLaunchKernel(batch_tile_kernel_of(kind, false), dim3(item_grid(h.n_p2, 2 * n_cu)), b, (void**)args,
                      tile_lds_bytes_shape(0, 2), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(select_batch_kernel, gq, b, 0, s, reqs, items, hdr);
  if (kind == EVOC_ELEM_BF16) hipLaunchKernelGGL(rescore_batch_kernel, gq, b, 0, s, reqs, items, hdr);
  hipLaunchKernelGGL(finalize_batch_kernel, gq, b, 0, s, reqs, items, hdr);
  return hipGetLastError();
}

hipError_t batch_set_grids(hipGraph_t graph, hipGraphExec_t exec, int kind, int n_cu, const BatchHeader& h) {
  size_t count = 0;
  hipError_t e = hipGraphGetNodes(graph, nullptr, &count);
  if (e != hipSuccess) return e;
  std::vector<hipGraphNode_t> nodes(count);
  e = hipGraphGetNodes(graph, nodes.data(), &count);
  for (size_t i = 0; i < count && e == hipSuccess; i++) {
    hipGraphNodeType type;
    e = hipGraphNodeGetType(nodes[i], &type);
    if (e != hipSuccess || type != hipGraphNodeTypeKernel) continue;
    hipKernelNodeParams p;
    e = hipGraphKernelNodeGetParams(nodes[i], &p);
    if (e != hipSuccess) break;
    const unsigned g = batch_grid_of(p.func, kind, n_cu, h);
    if (g == 0) return hipErrorInvalidValue;  // a node this sequence does not launch
    if (p.gridDim.x == g) continue;
    p.gridDim = dim3(g);
    e = hipGraphExecKernelNodeSetParams(exec, nodes[i], &p);
  }
  return e;
}

hipError_t launch_finalize(const uint64_t* keys, int n_q, int q_offset, int k, int dtype,
                           int32_t* idx, float* dist, hipStream_t s) {
  const int64_t tot = (int64_t)n_q * k;
  hipLaunchKernelGGL(finalize_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, keys, n_q,
                     q_offset, k, dtype, idx, dist);
  return hipGetLastError();
}

}  // namespace evoc
