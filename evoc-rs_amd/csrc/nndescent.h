// NN-descent kernel arguments and launchers (nndescent.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace evoc {

// Offer slots per vertex between merges (a multiple of 64: the merge sorts the
// list and the slots in one wave sort of 64 + NND_OFFER_CAP keys).  An offer
// takes the 64-bit maximum into two slots picked by a hash of (target,
// candidate, salt), so each slot ends as the best offer hashed to it whatever
// the arrival order, and repeated offers of one pair land on the same slots.
constexpr int NND_OFFER_CAP = 448;

// The iteration's inserted-entry count and the joins' gathered bytes are
// summed in NND_COUNTER_SLOTS slots one 128-byte line apart (slot chosen by
// vertex or item): one address taking an atomic from every wave of a launch
// serialises them (about 1 ms per 100k waves); the host adds the slots up.
constexpr int NND_COUNTER_SLOTS = 64;
constexpr int NND_SLOT_U32 = 32;  // uint32 stride of the inserted-entry slots
constexpr int NND_SLOT_U64 = 16;  // uint64 stride of the byte slots
// counters buffer: inserted-entry slots, byte slots, lost-offer count
constexpr size_t NND_COUNTER_BYTES = (size_t)NND_COUNTER_SLOTS * 128 * 2 + 128;

// Trees of the random-projection forest split together (one set of launches
// per level for all of them); a larger forest is built in groups of this many.
constexpr int NND_FOREST_TREES = 8;

struct JoinArgs {
  int32_t mode;     // 0: candidate lists of vertices v0.., 1: forest leaves
  int32_t n_items;  // work items (one wave each)
  int32_t v0;       // first vertex (mode 0)
  int32_t mc;       // list width of newc / oldc (mode 0, <= 64)
  const int32_t* newc;  // [n][mc], -1 = empty
  const int32_t* oldc;  // [n][mc]
  const int2* leaves;   // mode 1: (start, len) into perm, len <= 64
  const int32_t* perm;
  int32_t tree_n;       // mode 1: rows per tree (leaf of tree start / tree_n); 0: no filter
  int32_t tree_lo, tree_hi;  // mode 1: only leaves of trees [tree_lo, tree_hi)
  int32_t pad1;
  const uint8_t* rows;  // scored rows, row_bytes apart (multiple of 128)
  int64_t row_bytes;
  const uint32_t* row_pop;  // binary rows
  const float* thr;         // per-vertex score of the current Kg-th entry
  int32_t* umark;           // per-vertex flag: offer slots written since the last merge
  uint64_t* ubuf;           // [n][ucap] offer slots, EVOC_KEY_NONE when empty
  int32_t ucap;
  int32_t n;
  uint32_t salt;            // varies the slot hash between launches
  uint32_t* dropped;        // offers that met a better offer on both slots when they arrived
                            // (null: not counted, and the offers' atomics return nothing)
  unsigned long long* bytes;  // algorithmic bytes gathered: each item's candidate rows once, summed in
                              // NND_COUNTER_SLOTS slots NND_SLOT_U64 apart (may be null)
};

// The random-projection forest of n_trees trees, built level by level for all
// trees at once: tree t's rows occupy positions [t n, t n + n) of the flat
// permutation, a segment is (flat start, length, tree, 0) and a leaf (flat
// start, length).
struct ForestArgs {
  int32_t kind;
  int32_t n;          // rows (per tree)
  int32_t dim;        // elements per row (f32 / i8)
  int32_t leaf_size;
  const uint8_t* rows;   // i8 / binary / BF16 rows, row_bytes apart
  int64_t row_bytes;
  const float* frows;    // prepared f32 rows (f32 data)
  int64_t frow_ld;
  float* hyper;          // [segments][hyper_ld]
  int64_t hyper_ld;
  const uint64_t* seeds;  // [n_trees] per-tree seeds (device)
  int32_t* depth;         // [n_trees] levels split per tree (device)
  int32_t n_trees;
  int32_t pad0;
};

hipError_t nnd_launch_init(uint64_t* keys, float* thr, int32_t* umark, int n, int Kg, hipStream_t s);
hipError_t nnd_launch_join(int kind, const JoinArgs& a, hipStream_t s);
hipError_t nnd_launch_merge(uint64_t* keys, float* thr, int32_t* umark, uint64_t* ubuf, int ucap,
                            int n, int Kg, int v0, int nv, uint32_t* changed, hipStream_t s);
hipError_t nnd_launch_candidates(uint64_t* keys, int n, int Kg, int32_t* revcnt, int32_t* revoff,
                                 int32_t* revcur, int32_t* rev, int32_t* scan_sums, int mc,
                                 uint64_t seed, int it, int32_t* newc, int32_t* oldc, hipStream_t s);
hipError_t nnd_launch_export(const uint64_t* keys, int n, int Kg, uint64_t* out, hipStream_t s);
hipError_t nnd_launch_sum_dist(const uint64_t* keys, int64_t m, double* out, hipStream_t s);
// perm[t n + i] = i for every tree t
hipError_t rp_launch_init(int32_t* perm, const ForestArgs& f, hipStream_t s);
hipError_t rp_launch_level(const int4* segs, int nseg, const int32_t* perm, int32_t* perm2,
                           int32_t* seg_of, int8_t* side, const ForestArgs& f, int level,
                           int4* next_segs, int32_t* n_next, int2* leaves, int32_t* n_leaves,
                           int last_level, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
hipError_t nnd_launch_random_lists(const uint64_t* keys, int n, int Kg, int mc, uint64_t seed,
                                   int32_t* newc, int32_t* oldc, hipStream_t s);

}  // namespace evoc
