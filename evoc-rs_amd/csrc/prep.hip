// Input preparation on the device.
//
// prepare_float (src/knn.rs:78-113): the reference sums v*v serially in f32 per
// row, tests |sqrt(sum) - 1| > 1e-5 on every row, and if any row fails divides
// every row by its norm (a zero norm becomes 1).  The row norm is computed here
// by one thread per row in the same serial order with non-contracted IEEE
// operations, so the renormalisation decision and the divided values are the
// reference's bit for bit (rows staged through LDS so the loads coalesce).
// The element-wise pass is coalesced and also writes
// the zero-padded BF16 copy used by the MFMA scorer.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evoc_kernels.h"

namespace evoc {

// One 64-thread block per 64 rows.  Columns are staged 64 at a time through
// LDS with row-contiguous (coalesced) loads; thread t then adds row t's staged
// values in column order, so each row is still one serial f32 sum.
constexpr int NORM_ROWS = 64;
__global__ __launch_bounds__(NORM_ROWS) void row_norm_serial_kernel(const float* __restrict__ X, int64_t ld,
                                                                    int64_t n, int d,
                                                                    float* __restrict__ norms,
                                                                    int32_t* __restrict__ flag) {
  __shared__ float tile[NORM_ROWS][NORM_ROWS + 1];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * NORM_ROWS;
  const int rows = (int)(n - r0 < NORM_ROWS ? n - r0 : NORM_ROWS);
  float s = 0.0f;
  for (int c0 = 0; c0 < d; c0 += NORM_ROWS) {
    const int cols = min(NORM_ROWS, d - c0);
    __syncthreads();  // the last chunk's sums are done with the tile
    if (t < cols)
      for (int i = 0; i < rows; i++) tile[i][t] = X[(r0 + i) * ld + c0 + t];
    __syncthreads();
    if (t < rows)
      for (int j = 0; j < cols; j++) s = __fadd_rn(s, __fmul_rn(tile[t][j], tile[t][j]));
  }
  if (t >= rows) return;
  const float nr = __fsqrt_rn(s);
  if (fabsf(__fsub_rn(nr, 1.0f)) > 1e-5f) atomicOr(flag, 1);
  norms[r0 + t] = nr == 0.0f ? 1.0f : nr;
}

__global__ void zero_i32_kernel(int32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__device__ __forceinline__ uint16_t bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0u;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// grid: x over padded columns (256 per block), y over rows
__global__ void normalize_rows_kernel(const float* __restrict__ X, int64_t ld, int64_t n, int d,
                                      const float* __restrict__ norms,
                                      const int32_t* __restrict__ flag, float* __restrict__ out,
                                      int64_t out_ld, uint16_t* __restrict__ out_bf16,
                                      int64_t bf_ld) {
  const int64_t r = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const bool renorm = *flag != 0;
  float v = c < d ? X[r * ld + c] : 0.0f;
  if (renorm && c < d) v = __fdiv_rn(v, norms[r]);
  if (out && c < out_ld) out[r * out_ld + c] = v;
  if (out_bf16 && c < bf_ld) out_bf16[r * bf_ld + c] = bf16_rne(v);
}

__global__ void pad_rows_kernel(const uint8_t* __restrict__ X, int64_t ld, int64_t n, int row_bytes,
                                uint8_t* __restrict__ out, int64_t out_ld) {
  const int64_t r = blockIdx.y;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || c >= out_ld) return;
  out[r * out_ld + c] = c < row_bytes ? X[r * ld + c] : (uint8_t)0;
}

// rows padded to a multiple of 128 bytes (16-byte aligned, zero padded)
__global__ void row_popcount_kernel(const uint8_t* __restrict__ X, int64_t ld, int64_t n,
                                    int row_bytes, uint32_t* __restrict__ pop) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const uint4* x = (const uint4*)(X + r * ld);
  uint32_t s = 0;
  for (int t = 0; t < row_bytes / 16; t++) {
    const uint4 w = x[t];
    s += __popc(w.x) + __popc(w.y) + __popc(w.z) + __popc(w.w);
  }
  pop[r] = s;
}

static inline dim3 grid_rows(int64_t n, int64_t cols, int bs) {
  return dim3((unsigned)((cols + bs - 1) / bs), (unsigned)n);
}

// Batched preparation: one persistent pass per stage over (request, 64-row)
// items, each item the work of the single-request kernels above on its rows.
// Norms: 64 threads per item, as row_norm_serial_kernel (same serial order).
__global__ __launch_bounds__(NORM_ROWS) void norm_batch_kernel(const BatchReq* __restrict__ reqs,
                                                               const int4* __restrict__ items,
                                                               const BatchHeader* __restrict__ hdr) {
  __shared__ float tile[NORM_ROWS][NORM_ROWS + 1];
  const int t = threadIdx.x;
  const int n_items = hdr->n_rows;
  const int4* it = items + hdr->off_rows;
  for (int i = blockIdx.x; i < n_items; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    if (r.dtype != EVOC_DTYPE_F32) continue;  // (uniform: the whole block skips)
    const float* X = (const float*)r.x_in;
    const int64_t r0 = w.y;
    const int rows = (int)(r.n - r0 < NORM_ROWS ? r.n - r0 : NORM_ROWS);
    float acc = 0.0f;
    for (int c0 = 0; c0 < r.d; c0 += NORM_ROWS) {
      const int cols = min(NORM_ROWS, r.d - c0);
      __syncthreads();  // the last chunk's (or item's) sums are done with the tile
      if (t < cols)
        for (int j = 0; j < rows; j++) tile[j][t] = X[(r0 + j) * r.ld_in + c0 + t];
      __syncthreads();
      if (t < rows)
        for (int j = 0; j < cols; j++) acc = __fadd_rn(acc, __fmul_rn(tile[t][j], tile[t][j]));
    }
    if (t < rows) {
      const float nr = __fsqrt_rn(acc);
      if (fabsf(__fsub_rn(nr, 1.0f)) > 1e-5f) atomicOr(r.flag, 1);
      r.norms[r0 + t] = nr == 0.0f ? 1.0f : nr;
    }
  }
}

// Rows: f32 -> prepared FP32 (f32_ld apart) and, for the BF16 form, the BF16
// scored rows, as normalize_rows_kernel; int8 -> padded scored rows, as
// pad_rows_kernel.  256 threads per item, columns across the threads.
__global__ __launch_bounds__(256) void write_batch_kernel(const BatchReq* __restrict__ reqs,
                                                          const int4* __restrict__ items,
                                                          const BatchHeader* __restrict__ hdr) {
  const int n_items = hdr->n_rows;
  const int4* it = items + hdr->off_rows;
  for (int i = blockIdx.x; i < n_items; i += gridDim.x) {
    const int4 w = it[i];
    const BatchReq& r = reqs[w.x];
    const int64_t r0 = w.y;
    const int rows = (int)(r.n - r0 < 64 ? r.n - r0 : 64);
    if (r.dtype == EVOC_DTYPE_F32) {
      const float* X = (const float*)r.x_in;
      const bool renorm = *r.flag != 0;
      const int64_t bf_ld = r.bf16 ? r.row_bytes / 2 : 0;
      const int64_t cols = r.f32_ld > bf_ld ? r.f32_ld : bf_ld;
      for (int j = 0; j < rows; j++) {
        const int64_t row = r0 + j;
        for (int64_t c = threadIdx.x; c < cols; c += blockDim.x) {
          float v = c < r.d ? X[row * r.ld_in + c] : 0.0f;
          if (renorm && c < r.d) v = __fdiv_rn(v, r.norms[row]);
          if (c < r.f32_ld) r.xf32[row * r.f32_ld + c] = v;
          if (r.bf16 && c < bf_ld) ((uint16_t*)r.rows)[row * bf_ld + c] = bf16_rne(v);
        }
      }
    } else {
      for (int j = 0; j < rows; j++) {
        const int64_t row = r0 + j;
        for (int64_t c = threadIdx.x; c < r.row_bytes; c += blockDim.x)
          r.rows[row * r.row_bytes + c] = c < r.d ? r.x_in[row * r.ld_in + c] : (uint8_t)0;
      }
    }
  }
}

hipError_t launch_batch_prepare(unsigned norm_grid, unsigned write_grid, BatchReq* reqs, const int4* items,
                                const BatchHeader* hdr, hipStream_t s) {
  hipLaunchKernelGGL(norm_batch_kernel, dim3(norm_grid), dim3(NORM_ROWS), 0, s, reqs, items, hdr);
  hipLaunchKernelGGL(write_batch_kernel, dim3(write_grid), dim3(256), 0, s, reqs, items, hdr);
  return hipGetLastError();
}

const void* batch_norm_kernel() { return (const void*)norm_batch_kernel; }
const void* batch_write_kernel() { return (const void*)write_batch_kernel; }

hipError_t launch_prepare_float(const float* X, int64_t ld, int64_t n, int d, float* out,
                                int64_t out_ld, uint16_t* out_bf16, int64_t bf_ld, float* norms,
                                int32_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, flag, (int64_t)1);
  hipLaunchKernelGGL(row_norm_serial_kernel, dim3((unsigned)((n + NORM_ROWS - 1) / NORM_ROWS)), dim3(NORM_ROWS),
                     0, s, X, ld, n, d, norms, flag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t cols = out_ld > bf_ld ? out_ld : bf_ld;
  if (n > 65535) {
    // y-dimension limit: launch in row slabs
    for (int64_t r0 = 0; r0 < n; r0 += 65535) {
      const int64_t nr = (n - r0) < 65535 ? (n - r0) : 65535;
      hipLaunchKernelGGL(normalize_rows_kernel, grid_rows(nr, cols, 256), dim3(256), 0, s,
                         X + r0 * ld, ld, nr, d, norms + r0, flag, out ? out + r0 * out_ld : nullptr,
                         out_ld, out_bf16 ? out_bf16 + r0 * bf_ld : nullptr, bf_ld);
    }
  } else {
    hipLaunchKernelGGL(normalize_rows_kernel, grid_rows(n, cols, 256), dim3(256), 0, s, X, ld, n, d,
                       norms, flag, out, out_ld, out_bf16, bf_ld);
  }
  return hipGetLastError();
}

hipError_t launch_pad_rows(const uint8_t* X, int64_t ld, int64_t n, int row_bytes, uint8_t* out,
                           int64_t out_ld, hipStream_t s) {
  for (int64_t r0 = 0; r0 < n; r0 += 65535) {
    const int64_t nr = (n - r0) < 65535 ? (n - r0) : 65535;
    hipLaunchKernelGGL(pad_rows_kernel, grid_rows(nr, out_ld, 256), dim3(256), 0, s, X + r0 * ld, ld,
                       nr, row_bytes, out + r0 * out_ld, out_ld);
  }
  return hipGetLastError();
}

hipError_t launch_row_popcount(const uint8_t* X, int64_t ld, int64_t n, int row_bytes, uint32_t* pop,
                               hipStream_t s) {
  hipLaunchKernelGGL(row_popcount_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, ld,
                     n, row_bytes, pop);
  return hipGetLastError();
}

}  // namespace evoc
