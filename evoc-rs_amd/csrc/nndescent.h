// NN-descent kernel arguments and launchers (nndescent.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace evoc {

struct JoinArgs {
  int32_t mode;     // 0: candidate lists of vertices v0.., 1: forest leaves
  int32_t n_items;  // work items (one wave each)
  int32_t v0;       // first vertex (mode 0)
  int32_t mc;       // list width of newc / oldc (mode 0, <= 64)
  const int32_t* newc;  // [n][mc], -1 = empty
  const int32_t* oldc;  // [n][mc]
  const int2* leaves;   // mode 1: (start, len) into perm, len <= 64
  const int32_t* perm;
  const uint8_t* rows;  // scored rows, row_bytes apart (multiple of 128)
  int64_t row_bytes;
  const uint32_t* row_pop;  // binary rows
  const float* thr;         // per-vertex score of the current Kg-th entry
  int32_t* ucnt;            // per-vertex offer counters
  uint64_t* ubuf;           // [n][ucap] offers
  int32_t ucap;
  int32_t n;
  uint32_t* dropped;        // offers beyond ucap
};

struct ForestArgs {
  int32_t kind;
  int32_t n;
  int32_t dim;        // elements per row (f32 / i8)
  int32_t leaf_size;
  const uint8_t* rows;   // i8 / binary rows
  int64_t row_bytes;
  const float* frows;    // prepared f32 rows (f32 data)
  int64_t frow_ld;
  float* hyper;          // [segments][hyper_ld]
  int64_t hyper_ld;
  uint64_t seed;
};

hipError_t nnd_launch_init(uint64_t* keys, float* thr, int32_t* ucnt, int n, int Kg, hipStream_t s);
hipError_t nnd_launch_join(int kind, const JoinArgs& a, hipStream_t s);
hipError_t nnd_launch_merge(uint64_t* keys, float* thr, int32_t* ucnt, const uint64_t* ubuf, int ucap,
                            int n, int Kg, int v0, int nv, uint32_t* changed, hipStream_t s);
hipError_t nnd_launch_candidates(uint64_t* keys, int n, int Kg, int32_t* revcnt, int32_t* revoff,
                                 int32_t* revcur, int32_t* rev, int32_t* scan_sums, int mc,
                                 uint64_t seed, int it, int32_t* newc, int32_t* oldc, hipStream_t s);
hipError_t nnd_launch_export(const uint64_t* keys, int n, int Kg, uint64_t* out, hipStream_t s);
hipError_t nnd_launch_sum_dist(const uint64_t* keys, int64_t m, double* out, hipStream_t s);
hipError_t rp_launch_level(const int4* segs, int nseg, const int32_t* perm, int32_t* perm2,
                           int32_t* seg_of, int8_t* side, const ForestArgs& f, int level,
                           int4* next_segs, int32_t* n_next, int2* leaves, int32_t* n_leaves,
                           int last_level, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
hipError_t nnd_launch_random_lists(const uint64_t* keys, int n, int Kg, int mc, uint64_t seed,
                                   int32_t* newc, int32_t* oldc, hipStream_t s);

}  // namespace evoc
