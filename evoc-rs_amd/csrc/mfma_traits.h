// MFMA building blocks shared by the gfx950 kernels: the 32x32 f32/i32
// accumulator blocks for BF16, FP32 and INT8 operands.
//
// Operand layout used everywhere (v_mfma_*_32x32x*): lane l supplies row
// l % 32 of A (and column l % 32 of B) and the K slice starting at element
// (l / 32) * LANE_ELEMS of each K step; both operands use the same K mapping,
// so each product sums the right pairs.  The 32x32 accumulator block puts
// column l % 32 and rows 8*(j/4) + 4*(l/32) + j%4 (j = 0..15) in lane l.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evoc_device.h"
#include "evoc_kernels.h"

namespace evoc {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;

template <int KIND> struct Mfma;

template <> struct Mfma<EVOC_ELEM_BF16> {
  static constexpr int KSTEP_BYTES = 32;  // 16 bf16 per MFMA K step
  static constexpr int LANE_BYTES = 16;   // 8 bf16 per lane
  typedef f32x16 acc_t;
  typedef i32x4 frag_t;
  __device__ static frag_t load(const uint8_t* p) { return *(const i32x4*)p; }
  __device__ static void mma(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                0, 0, 0);
  }
  // c = a b, accumulating into zero (the block's first K step)
  __device__ static void mma_zero(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                acc_t{}, 0, 0, 0);
  }
  __device__ static acc_t zero() { return acc_t{}; }
  __device__ static void step(acc_t& c, const uint8_t* a, const uint8_t* b) {
    bf16x8 av = __builtin_bit_cast(bf16x8, *(const i32x4*)a);
    bf16x8 bv = __builtin_bit_cast(bf16x8, *(const i32x4*)b);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
  }
  __device__ static float value(const acc_t& c, int j) { return f32_score(c[j]); }
};

template <> struct Mfma<EVOC_ELEM_F32> {
  static constexpr int KSTEP_BYTES = 8;  // 2 f32 per MFMA K step
  static constexpr int LANE_BYTES = 4;
  typedef f32x16 acc_t;
  typedef float frag_t;
  __device__ static frag_t load(const uint8_t* p) { return *(const float*)p; }
  __device__ static void mma(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  __device__ static void mma_zero(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc_t{}, 0, 0, 0);
  }
  __device__ static acc_t zero() { return acc_t{}; }
  __device__ static void step(acc_t& c, const uint8_t* a, const uint8_t* b) {
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(*(const float*)a, *(const float*)b, c, 0, 0, 0);
  }
  __device__ static float value(const acc_t& c, int j) { return f32_score(c[j]); }
};

template <> struct Mfma<EVOC_ELEM_I8> {
  static constexpr int KSTEP_BYTES = 32;  // 32 i8 per MFMA K step
  static constexpr int LANE_BYTES = 16;
  typedef i32x16 acc_t;
  typedef i32x4 frag_t;
  __device__ static frag_t load(const uint8_t* p) { return *(const i32x4*)p; }
  __device__ static void mma(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  }
  __device__ static void mma_zero(acc_t& c, frag_t a, frag_t b) {
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc_t{}, 0, 0, 0);
  }
  __device__ static acc_t zero() { return acc_t{}; }
  __device__ static void step(acc_t& c, const uint8_t* a, const uint8_t* b) {
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(*(const i32x4*)a, *(const i32x4*)b, c, 0, 0, 0);
  }
  // int8.rs:15-21: -(dot as f32); the score is dot as f32 (exact below 2^24)
  __device__ static float value(const acc_t& c, int j) { return (float)c[j]; }
};

}  // namespace evoc
