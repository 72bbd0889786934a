// Device-side helpers shared by the gfx950 kernels.
//
// Ordering contract (the exact-mode answer): a candidate c of query q is ranked
// by (score descending, self first, index ascending), where `score` is the
// similarity form of the reference's raw heap distance (raw = -score):
//   f32 : dot if dot > 0 else -FLT_MIN        (src/fast_cosine.rs:33-45)
//   i8  : (float)dot_i32                      (src/nndescent/int8.rs:15-21)
//   u8  : inter / union, 0 if union == 0      (src/nndescent/uint8.rs:21-36)
// and self carries the reference's NN-descent seed value raw = -1.0, i.e.
// score 1.0 (src/nndescent/float.rs:210-214).  A (score, index) pair is packed
// into one 64-bit key whose unsigned order is that ranking, so selection is a
// plain descending sort of keys.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define EVOC_WAVE 64

#include <utility>

namespace evoc {

// Compile-time loop: f(std::integral_constant<int, i>) for i = 0 .. N-1, with
// no loop for the unroller to decline (accumulator blocks must be indexed by
// constants or they leave registers).
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// float -> uint32 with the same order; -0.0 is folded onto +0.0 so the two
// compare equal, as they do as IEEE floats in the reference.
__device__ __forceinline__ uint32_t ord_f32(float s) {
  uint32_t u = __float_as_uint(s);
  if ((u & 0x7fffffffu) == 0u) u = 0u;
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __uint_as_float(u);
}

// rank slot: self -> 0, candidate c -> c + 1; smaller slot wins ties.
__device__ __forceinline__ uint64_t make_key(float score, int32_t cand, bool self) {
  uint32_t slot = self ? 0u : (uint32_t)cand + 1u;
  return ((uint64_t)ord_f32(score) << 32) | (uint64_t)(0xffffffffu - slot);
}
__device__ __forceinline__ int32_t key_index(uint64_t key, int32_t query) {
  uint32_t slot = 0xffffffffu - (uint32_t)(key & 0xffffffffu);
  return slot == 0u ? query : (int32_t)(slot - 1u);
}
__device__ __forceinline__ float key_score(uint64_t key) { return unord_f32((uint32_t)(key >> 32)); }

// Empty slot: below every real key.
#define EVOC_KEY_NONE 0ull

// f32 similarity with the reference's FLT_MIN rule (src/fast_cosine.rs:33-45).
__device__ __forceinline__ float f32_score(float dot) { return dot > 0.0f ? dot : -1.17549435e-38f; }

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = (uint32_t)__shfl_xor((int)lo, m, EVOC_WAVE);
  hi = (uint32_t)__shfl_xor((int)hi, m, EVOC_WAVE);
  return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort of 64*E keys held E per lane, register-major: element
// i = e*64 + lane lives in register e of lane `lane`.  Result is DESCENDING
// (element 0 = largest).  Distances j < 64 exchange across lanes, j >= 64
// swap registers within a lane.  Fully unrolled for compile-time E.
__device__ __forceinline__ uint32_t shfl_xor_key(uint32_t v, int m) {
  return (uint32_t)__shfl_xor((int)v, m, EVOC_WAVE);
}
__device__ __forceinline__ uint64_t shfl_xor_key(uint64_t v, int m) { return shfl_xor_u64(v, m); }

template <int E, typename T>
__device__ __forceinline__ void wave_sort_desc(T (&v)[E]) {
  const int lane = threadIdx.x & (EVOC_WAVE - 1);
  constexpr int N = EVOC_WAVE * E;
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= EVOC_WAVE) {
        const int em = j / EVOC_WAVE;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int p = e ^ em;
          if (p > e) {
            const int i = e * EVOC_WAVE + lane;  // the lower element of the pair
            // segments with (i & k) == 0 sort descending, the others ascending
            const bool desc = (i & k) == 0;
            const T a = v[e], b = v[p];
            const bool sw = desc ? (a < b) : (a > b);
            v[e] = sw ? b : a;
            v[p] = sw ? a : b;
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int i = e * EVOC_WAVE + lane;
          const T o = shfl_xor_key(v[e], j);
          const bool lower = (i & j) == 0;
          const bool desc = (i & k) == 0;
          const bool keep_max = (lower == desc);
          v[e] = keep_max ? (v[e] > o ? v[e] : o) : (v[e] < o ? v[e] : o);
        }
      }
    }
  }
}

// K-th largest of the 64*E values of a wave (register-major, as above; K >= 1),
// counting repeats: the largest t with #(values >= t) >= K, built bit by bit
// from the top.  Each count is one ballot per register (no data movement).
// Registers e with 64 e >= n_used are taken to hold only values below every
// candidate t (0), and are skipped.  Returns 0 when fewer than K values are >= 1.
// The values at positions below n_used (element i in register i / 64 of lane
// i % 64) sort above the padding after them.  When at least K of them exist
// their K-th largest is the result and lies between their smallest and the
// largest value, so it has the leading bits those two share: the bisection
// starts below them.  All lanes of the wave must be active.
template <typename T>
__device__ __forceinline__ T wave_shfl_xor_bits(T x, int m) {
  if constexpr (sizeof(T) == 8) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), m, 64);
    return ((T)hi << 32) | (T)lo;
  } else {
    return (T)(uint32_t)__shfl_xor((int)(uint32_t)x, m, 64);
  }
}

template <int E, typename T>
__device__ __forceinline__ T wave_kth_largest(const T (&v)[E], int K, int n_used) {
  constexpr int W = (int)sizeof(T) * 8;
  const int lane = threadIdx.x & 63;
  T hi = 0, lo = ~(T)0;
#pragma unroll
  for (int e = 0; e < E; e++)
    if (e * 64 + lane < n_used) {
      hi = v[e] > hi ? v[e] : hi;
      lo = v[e] < lo ? v[e] : lo;
    }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const T h = wave_shfl_xor_bits(hi, m), l = wave_shfl_xor_bits(lo, m);
    hi = h > hi ? h : hi;
    lo = l < lo ? l : lo;
  }
  T t = 0;
  int top = W - 1;
  if (n_used >= K) {
    const T diff = hi ^ lo;
    if (diff == 0) return hi;
    if constexpr (sizeof(T) == 8) top = 63 - __builtin_clzll((unsigned long long)diff);
    else top = 31 - __builtin_clz((uint32_t)diff);
    t = top + 1 >= W ? (T)0 : (T)(hi & (~(T)0 << (top + 1)));
  }
#pragma unroll 1
  for (int b = top; b >= 0; b--) {
    const T cand = t | ((T)1 << b);
    int cnt = 0;
#pragma unroll
    for (int e = 0; e < E; e++)
      if (e * 64 < n_used) cnt += __popcll(__ballot(v[e] >= cand));
    if (cnt >= K) t = cand;
  }
  return t;
}

// Element i of a register-major wave array, broadcast to every lane.
template <int E>
__device__ __forceinline__ uint64_t wave_elem(const uint64_t (&v)[E], int i) {
  uint64_t mine = 0;
#pragma unroll
  for (int e = 0; e < E; e++)
    if (e == i / EVOC_WAVE) mine = v[e];
  const int src = i % EVOC_WAVE;
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)mine, src, EVOC_WAVE);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(mine >> 32), src, EVOC_WAVE);
  return ((uint64_t)hi << 32) | lo;
}
template <int E>
__device__ __forceinline__ uint32_t wave_elem(const uint32_t (&v)[E], int i) {
  uint32_t mine = 0;
#pragma unroll
  for (int e = 0; e < E; e++)
    if (e == i / EVOC_WAVE) mine = v[e];
  return (uint32_t)__shfl((int)mine, i % EVOC_WAVE, EVOC_WAVE);
}

// Sum over the wave with a fixed butterfly order (deterministic).
__device__ __forceinline__ float wave_sum_f32(float x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, EVOC_WAVE);
  return x;
}
__device__ __forceinline__ int wave_sum_i32(int x) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, EVOC_WAVE);
  return x;
}

}  // namespace evoc
