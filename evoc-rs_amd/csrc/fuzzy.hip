// The steps after knn_graph that consume its output directly, on the device:
//
//   neighbor_graph_matrix_with_coo (src/graph_construction.rs:175-226): per-row
//     sigma / rho (smooth_knn_dist, :15-81), membership strengths
//     (compute_membership_strengths, :83-115) and the fuzzy union
//     A + A^T - A o A^T (src/csr_matmul.rs:62-147) as CSR, rows ascending by
//     column, exact zeros dropped (canonicalize_csr_rows, :228-240);
//   find_duplicates (src/cluster_util.rs:430-449).
//
// Every kNN edge i -> j writes its entries into two fixed slots (no atomics):
// (i, j) with its union value, and (j, i) when j does not list i.  One radix
// sort of the keys row * n + col (over the bits n^2 needs) then orders the
// whole matrix; a row's start is a binary search over the sorted keys.
// Arithmetic follows the reference operation by operation: f64 for sigma /
// rho, f32 for strengths, every product and sum rounded on its own (no
// contraction).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "evoc_device.h"
#include "evoc_kernels.h"

namespace evoc {

constexpr uint64_t FZ_NONE = ~0ull;       // empty slot: sorts after every entry (all sorted bits set)
constexpr int FZ_SUM_BLOCKS = 1024;       // partial sums of the distance mean

// Deterministic f64 sum of m floats: fixed per-thread stride, fixed tree.
__global__ __launch_bounds__(256) void fz_sum_kernel(const float* __restrict__ x, int64_t m,
                                                     double* __restrict__ partial) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (int64_t)gridDim.x * 256)
    s += (double)x[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// partial[nb] = mean of all distances, rounded to f32 as ndarray's f32 mean
// returns it (src/graph_construction.rs:19)
__global__ __launch_bounds__(256) void fz_mean_kernel(double* __restrict__ partial, int nb, int64_t m) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[nb] = m > 0 ? (double)(float)(sh[0] / (double)m) : 0.0;
}

// smooth_knn_dist (src/graph_construction.rs:15-81), one thread per row, in
// f64 and in row order: rho = first positive distance; sigma from at most 64
// bisection steps on sum_{j >= 1} exp(-(d_j - rho) / sigma) = log2(k), floored
// at 1e-3 x the row mean (rho > 0) or 1e-3 x the mean of all distances.
__global__ __launch_bounds__(256) void fz_sigma_kernel(const float* __restrict__ dist, int64_t n, int k,
                                                       double target, const double* __restrict__ mean_all,
                                                       float* __restrict__ sigma, float* __restrict__ rho) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float* row = dist + i * k;
  double rho_i = 0.0;
  for (int j = 0; j < k; j++) {
    const double d = (double)row[j];
    if (d > 0.0) {
      rho_i = d;
      break;
    }
  }
  const double tol = (double)1e-5f;  // SMOOTH_K_TOLERANCE, src/graph_construction.rs:12
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; it++) {
    double psum = 0.0;
    for (int j = 1; j < k; j++) {
      const double d = __dsub_rn((double)row[j], rho_i);
      psum = __dadd_rn(psum, d > 0.0 ? exp(-__ddiv_rn(d, mid)) : 1.0);
    }
    if (fabs(__dsub_rn(psum, target)) < tol) break;
    if (psum > target) {
      hi = mid;
      mid = __ddiv_rn(__dadd_rn(lo, hi), 2.0);
    } else {
      lo = mid;
      mid = isinf(hi) ? __dmul_rn(mid, 2.0) : __ddiv_rn(__dadd_rn(lo, hi), 2.0);
    }
  }
  double s = mid;
  const double scale = (double)1e-3f;  // MIN_K_DIST_SCALE, src/graph_construction.rs:13
  double fl;
  if (rho_i > 0.0) {
    double m = 0.0;
    for (int j = 0; j < k; j++) m = __dadd_rn(m, (double)row[j]);
    fl = __dmul_rn(scale, __ddiv_rn(m, (double)k));
  } else {
    fl = __dmul_rn(scale, *mean_all);
  }
  if (s < fl) s = fl;
  sigma[i] = (float)s;
  rho[i] = (float)rho_i;
}

// compute_membership_strengths (src/graph_construction.rs:83-115), f32
__device__ __forceinline__ float fz_membership(int64_t i, int32_t j, float d, float sg, float rh) {
  if ((int64_t)j == i) return 0.0f;
  const float x = __fsub_rn(d, rh);
  if (x <= 0.0f || sg == 0.0f) return 1.0f;
  return expf(-__fdiv_rn(x, sg));
}

// One thread per kNN entry e = i * k + t (neighbour j): slots 2e and 2e + 1.
// With symmetrize the union value of (i, j) is (a + b) - a b where b is j's
// strength for i (both stored), or a where only i lists j; in the latter case
// (j, i) holds a as well (src/csr_matmul.rs:62-74 with a missing operand 0).
// Indices outside [-1, n) set *bad.
__global__ __launch_bounds__(256) void fz_emit_kernel(const int32_t* __restrict__ inds,
                                                      const float* __restrict__ dist, int64_t n, int k,
                                                      const float* __restrict__ sigma,
                                                      const float* __restrict__ rho, int symmetrize,
                                                      uint64_t* __restrict__ keys, float* __restrict__ vals,
                                                      int32_t* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int32_t j = inds[e];
  uint64_t k0 = FZ_NONE, k1 = FZ_NONE;
  float v0 = 0.0f, v1 = 0.0f;
  if (j < -1 || (int64_t)j >= n) {
    atomicOr(bad, 1);
  } else if (j >= 0) {
    const float a = fz_membership(i, j, dist[e], sigma[i], rho[i]);
    if (a != 0.0f) {
      k0 = (uint64_t)i * (uint64_t)n + (uint64_t)j;
      v0 = a;
      if (symmetrize) {
        float b = 0.0f;
        const int64_t rj = (int64_t)j * k;
        for (int t = 0; t < k; t++)
          if ((int64_t)inds[rj + t] == i) {
            b = fz_membership(j, (int32_t)i, dist[rj + t], sigma[j], rho[j]);
            break;
          }
        if (b != 0.0f) {
          v0 = __fsub_rn(__fadd_rn(a, b), __fmul_rn(a, b));
        } else {
          k1 = (uint64_t)j * (uint64_t)n + (uint64_t)i;
          v1 = a;
        }
      }
    }
  }
  keys[2 * e] = k0;
  keys[2 * e + 1] = k1;
  vals[2 * e] = v0;
  vals[2 * e + 1] = v1;
}

// indptr[r] = first sorted position whose row is >= r (r = n: the entry count)
__global__ __launch_bounds__(256) void fz_indptr_kernel(const uint64_t* __restrict__ keys, int64_t m, int64_t n,
                                                        int64_t* __restrict__ indptr) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r > n) return;
  const uint64_t target = (uint64_t)r * (uint64_t)n;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  indptr[r] = lo;
}

__global__ __launch_bounds__(256) void fz_cols_kernel(const uint64_t* __restrict__ keys, int64_t nnz, int64_t n,
                                                      int32_t* __restrict__ cols) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < nnz) cols[p] = (int32_t)(keys[p] % (uint64_t)n);
}

// find_duplicates (src/cluster_util.rs:430-449): the threshold is the largest
// first-column distance (NaN ignored, as f32::max does), kept as an
// order-preserving integer (0: no value).
__global__ __launch_bounds__(256) void fz_colmax_kernel(const float* __restrict__ dist, int64_t n, int k,
                                                        uint32_t* __restrict__ dmax) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float d = dist[i * k];
  if (!isnan(d)) atomicMax(dmax, ord_f32(d));
}

// One key per kNN entry: min(i, j) * n + max(i, j) when i != j and the
// distance is <= the threshold, else empty.
__global__ __launch_bounds__(256) void fz_dup_emit_kernel(const int32_t* __restrict__ inds,
                                                          const float* __restrict__ dist, int64_t n, int k,
                                                          const uint32_t* __restrict__ dmax,
                                                          uint64_t* __restrict__ keys, int32_t* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int32_t j = inds[e];
  uint64_t key = FZ_NONE;
  if (j < -1 || (int64_t)j >= n) {
    atomicOr(bad, 1);
  } else if (j >= 0 && (int64_t)j != i && *dmax != 0u && dist[e] <= unord_f32(*dmax)) {
    const uint64_t a = (uint64_t)(i < j ? i : j), b = (uint64_t)(i < j ? j : i);
    key = a * (uint64_t)n + b;
  }
  keys[e] = key;
}

// distinct pair keys a * n + b -> the ABI's (a << 32 | b)
__global__ __launch_bounds__(256) void fz_pairs_kernel(const uint64_t* __restrict__ keys, int64_t m, int64_t n,
                                                       uint64_t* __restrict__ pairs) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p < m) pairs[p] = ((keys[p] / (uint64_t)n) << 32) | (keys[p] % (uint64_t)n);
}

static inline unsigned blocks_for(int64_t m) { return (unsigned)((m + 255) / 256); }

// Projection of each row on the vector whose element i is +1 or -1 by a hash
// bit of i, as an order-preserving integer; one wave per row.
__global__ __launch_bounds__(256) void proj_key_kernel(const float* __restrict__ X, int64_t ld, int64_t n, int d,
                                                       uint32_t* __restrict__ keys, int32_t* __restrict__ iota) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  float acc = 0.0f;
  for (int i = lane; i < d; i += 64) {
    const uint32_t h = (uint32_t)i * 0x9E3779B1u;
    acc += (h & 0x80000000u) ? -X[r * ld + i] : X[r * ld + i];
  }
  acc = wave_sum_f32(acc);
  if (lane == 0) {
    keys[r] = ord_f32(acc);
    iota[r] = (int32_t)r;
  }
}

hipError_t locality_order(void* temp, size_t& temp_bytes, const float* X, int64_t ld, int64_t n, int d,
                          uint32_t* keys, uint32_t* keys_out, int32_t* order_in, int32_t* order, hipStream_t s) {
  if (!temp)
    return rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, keys_out, order_in, order, (size_t)n, 0, 32, s);
  hipLaunchKernelGGL(proj_key_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, X, ld, n, d, keys, order_in);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_out, order_in, order, (size_t)n, 0, 32, s);
}

hipError_t launch_dup_pairs(const uint64_t* keys, int64_t m, int64_t n, uint64_t* pairs, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(fz_pairs_kernel, dim3(blocks_for(m)), dim3(256), 0, s, keys, m, n, pairs);
  return hipGetLastError();
}

hipError_t launch_fuzzy_sigma(const float* dist, int64_t n, int k, float n_neighbors, double* partial,
                              float* sigma, float* rho, hipStream_t s) {
  const int64_t m = n * k;
  hipLaunchKernelGGL(fz_sum_kernel, dim3(FZ_SUM_BLOCKS), dim3(256), 0, s, dist, m, partial);
  hipLaunchKernelGGL(fz_mean_kernel, dim3(1), dim3(256), 0, s, partial, FZ_SUM_BLOCKS, m);
  hipLaunchKernelGGL(fz_sigma_kernel, dim3(blocks_for(n)), dim3(256), 0, s, dist, n, k,
                     log2((double)n_neighbors), (const double*)(partial + FZ_SUM_BLOCKS), sigma, rho);
  return hipGetLastError();
}

size_t fuzzy_partial_doubles() { return FZ_SUM_BLOCKS + 1; }

hipError_t launch_fuzzy_emit(const int32_t* inds, const float* dist, int64_t n, int k, const float* sigma,
                             const float* rho, int symmetrize, uint64_t* keys, float* vals, int32_t* bad,
                             hipStream_t s) {
  hipLaunchKernelGGL(fz_emit_kernel, dim3(blocks_for(n * k)), dim3(256), 0, s, inds, dist, n, k, sigma, rho,
                     symmetrize, keys, vals, bad);
  return hipGetLastError();
}

// Bits of the keys row * n + col: the empty key's low bits (all set) then
// still exceed every real key (<= n^2 - 1 < 2^bits).
unsigned fuzzy_key_bits(int64_t n) {
  const uint64_t top = (uint64_t)n * (uint64_t)n;
  unsigned b = 1;
  while (b < 64 && (top >> b) != 0) b++;
  return b;
}

// rocPRIM radix sort of (key, value) pairs over the low `bits` key bits;
// temp == nullptr queries temp_bytes
hipError_t fuzzy_sort_pairs(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const float* vals_in, float* vals_out, int64_t m, unsigned bits, hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)m, 0, bits, s);
}

hipError_t launch_fuzzy_csr(const uint64_t* keys, int64_t m, int64_t n, int64_t* indptr, hipStream_t s) {
  hipLaunchKernelGGL(fz_indptr_kernel, dim3(blocks_for(n + 1)), dim3(256), 0, s, keys, m, n, indptr);
  return hipGetLastError();
}

hipError_t launch_fuzzy_cols(const uint64_t* keys, int64_t nnz, int64_t n, int32_t* cols, hipStream_t s) {
  if (nnz <= 0) return hipSuccess;
  hipLaunchKernelGGL(fz_cols_kernel, dim3(blocks_for(nnz)), dim3(256), 0, s, keys, nnz, n, cols);
  return hipGetLastError();
}

hipError_t launch_dup_emit(const int32_t* inds, const float* dist, int64_t n, int k, uint32_t* dmax,
                           uint64_t* keys, int32_t* bad, hipStream_t s) {
  hipLaunchKernelGGL(fz_colmax_kernel, dim3(blocks_for(n)), dim3(256), 0, s, dist, n, k, dmax);
  hipLaunchKernelGGL(fz_dup_emit_kernel, dim3(blocks_for(n * k)), dim3(256), 0, s, inds, dist, n, k,
                     (const uint32_t*)dmax, keys, bad);
  return hipGetLastError();
}

// sort of keys (low `bits`), then unique; temp == nullptr queries temp_bytes
// (the larger of the two steps' needs)
hipError_t dup_sort_unique(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_sorted,
                           uint64_t* keys_unique, int64_t* n_unique, int64_t m, unsigned bits, hipStream_t s) {
  if (!temp) {
    size_t a = 0, b = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, a, keys_in, keys_sorted, (size_t)m, 0, bits, s);
    if (e != hipSuccess) return e;
    e = rocprim::unique(nullptr, b, keys_sorted, keys_unique, n_unique, (size_t)m, rocprim::equal_to<uint64_t>(), s);
    temp_bytes = a > b ? a : b;
    return e;
  }
  size_t a = temp_bytes;
  hipError_t e = rocprim::radix_sort_keys(temp, a, keys_in, keys_sorted, (size_t)m, 0, bits, s);
  if (e != hipSuccess) return e;
  size_t b = temp_bytes;
  return rocprim::unique(temp, b, keys_sorted, keys_unique, n_unique, (size_t)m, rocprim::equal_to<uint64_t>(), s);
}

}  // namespace evoc
