// C ABI of the steps after knn_graph that consume its output (fuzzy.hip):
// the fuzzy graph (src/graph_construction.rs:175-226) and find_duplicates
// (src/cluster_util.rs:430-449).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "evoc_host.h"
#include "evoc_kernels.h"
#include "evoc_mi355x.h"

using namespace evoc;

namespace {

// The caller's n x k kNN arrays on the device: as given, or copied into `own`.
int stage_knn(Ctx* ctx, int32_t location, const int32_t* inds, const float* dist, int64_t n, int64_t k,
              DevBuf& own_i, DevBuf& own_d, const int32_t** d_inds, const float** d_dist) {
  if (location == EVOC_MEM_DEVICE) {
    *d_inds = inds;
    *d_dist = dist;
    return EVOC_OK;
  }
  const size_t bytes = (size_t)(n * k) * 4;
  int rc = own_i.alloc(ctx, bytes, "kNN indices");
  if (!rc) rc = own_d.alloc(ctx, bytes, "kNN distances");
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(own_i.ptr, inds, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(own_d.ptr, dist, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return ctx->hip_fail(e, "kNN upload");
  *d_inds = (const int32_t*)own_i.ptr;
  *d_dist = (const float*)own_d.ptr;
  return EVOC_OK;
}

hipMemcpyKind out_kind(int32_t location) {
  return location == EVOC_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

}  // namespace

extern "C" int evoc_mi355x_fuzzy_graph(evoc_mi355x_ctx* ctx, const evoc_fuzzy_input* in, evoc_fuzzy_output* out) {
  if (!ctx || !in || !out) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  CallScope call(ctx);
  const auto t_start = std::chrono::steady_clock::now();
  ctx->last_error.clear();
  out->nnz = 0;
  const int64_t n = in->n, k = in->k;
  if (n <= 0) return ctx->fail(EVOC_ERR_EMPTY_DATA, "empty kNN graph");
  if (k <= 0) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "k must be positive");
  if (n > (int64_t)INT32_MAX - 1 || k > (int64_t)INT32_MAX / 2)
    return ctx->fail(EVOC_ERR_UNSUPPORTED, "more than 2^31-2 rows");
  if (!in->knn_indices || !in->knn_dists || !out->indptr || (out->capacity > 0 && (!out->indices || !out->data)))
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "null array");
  if (!(in->n_neighbors > 0.0f)) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "n_neighbors must be positive");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(EVOC_ERR_HIP, "hipSetDevice failed");
  hipStream_t st = ctx->stream;

  DevBuf own_i, own_d;
  const int32_t* d_inds = nullptr;
  const float* d_dist = nullptr;
  int rc = stage_knn(ctx, in->location, in->knn_indices, in->knn_dists, n, k, own_i, own_d, &d_inds, &d_dist);
  if (rc) return rc;
  (void)hipEventRecord(ctx->ev_begin, st);

  const int64_t m = 2 * n * k;  // entry slots
  DevBuf partial, sig, rho, keys_a, keys_b, vals_a, vals_b, indptr, bad, temp;
  if ((rc = partial.alloc(ctx, fuzzy_partial_doubles() * sizeof(double), "mean partial sums"))) return rc;
  if ((rc = sig.alloc(ctx, (size_t)n * 4, "sigmas"))) return rc;
  if ((rc = rho.alloc(ctx, (size_t)n * 4, "rhos"))) return rc;
  if ((rc = keys_a.alloc(ctx, (size_t)m * 8, "entry keys"))) return rc;
  if ((rc = keys_b.alloc(ctx, (size_t)m * 8, "sorted keys"))) return rc;
  if ((rc = vals_a.alloc(ctx, (size_t)m * 4, "entry values"))) return rc;
  if ((rc = vals_b.alloc(ctx, (size_t)m * 4, "sorted values"))) return rc;
  if ((rc = indptr.alloc(ctx, (size_t)(n + 1) * 8, "row starts"))) return rc;
  if ((rc = bad.alloc(ctx, 4, "index check"))) return rc;
  size_t temp_bytes = 0;
  const unsigned bits = fuzzy_key_bits(n);
  hipError_t e = fuzzy_sort_pairs(nullptr, temp_bytes, (const uint64_t*)keys_a.ptr, (uint64_t*)keys_b.ptr,
                                  (const float*)vals_a.ptr, (float*)vals_b.ptr, m, bits, st);
  if (e != hipSuccess) return ctx->hip_fail(e, "sort sizing");
  if ((rc = temp.alloc(ctx, temp_bytes > 0 ? temp_bytes : 1, "sort scratch"))) return rc;

  e = hipMemsetAsync(bad.ptr, 0, 4, st);
  if (e == hipSuccess)
    e = launch_fuzzy_sigma(d_dist, n, (int)k, in->n_neighbors, (double*)partial.ptr, (float*)sig.ptr,
                           (float*)rho.ptr, st);
  if (e == hipSuccess)
    e = launch_fuzzy_emit(d_inds, d_dist, n, (int)k, (const float*)sig.ptr, (const float*)rho.ptr,
                          in->symmetrize ? 1 : 0, (uint64_t*)keys_a.ptr, (float*)vals_a.ptr, (int32_t*)bad.ptr, st);
  if (e == hipSuccess)
    e = fuzzy_sort_pairs(temp.ptr, temp_bytes, (const uint64_t*)keys_a.ptr, (uint64_t*)keys_b.ptr,
                         (const float*)vals_a.ptr, (float*)vals_b.ptr, m, bits, st);
  if (e == hipSuccess) e = launch_fuzzy_csr((const uint64_t*)keys_b.ptr, m, n, (int64_t*)indptr.ptr, st);
  int64_t nnz = 0;
  int32_t bad_h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nnz, (int64_t*)indptr.ptr + n, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&bad_h, bad.ptr, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ctx->hip_fail(e, "fuzzy graph");
  if (bad_h) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "kNN index outside [-1, n)");
  out->nnz = nnz;
  if (nnz > out->capacity)
    return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "output capacity < nnz (nnz is set; 2 n k always suffices)");

  // ---- CSR out: columns from the sorted keys, values, row starts ----
  const hipMemcpyKind kd = out_kind(out->location);
  if (out->location == EVOC_MEM_DEVICE) {
    e = launch_fuzzy_cols((const uint64_t*)keys_b.ptr, nnz, n, out->indices, st);
  } else {
    // the columns go through the (now unused) entry-key buffer
    e = launch_fuzzy_cols((const uint64_t*)keys_b.ptr, nnz, n, (int32_t*)keys_a.ptr, st);
    if (e == hipSuccess && nnz > 0) e = hipMemcpyAsync(out->indices, keys_a.ptr, (size_t)nnz * 4, kd, st);
  }
  if (e == hipSuccess && nnz > 0) e = hipMemcpyAsync(out->data, vals_b.ptr, (size_t)nnz * 4, kd, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out->indptr, indptr.ptr, (size_t)(n + 1) * 8, kd, st);
  if (e == hipSuccess && out->sigmas) e = hipMemcpyAsync(out->sigmas, sig.ptr, (size_t)n * 4, kd, st);
  if (e == hipSuccess && out->rhos) e = hipMemcpyAsync(out->rhos, rho.ptr, (size_t)n * 4, kd, st);
  (void)hipEventRecord(ctx->ev_end, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ctx->hip_fail(e, "fuzzy graph output");
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
  out->device_seconds = ms * 1e-3;
  out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  return EVOC_OK;
}

extern "C" int evoc_mi355x_find_duplicates(evoc_mi355x_ctx* ctx, int32_t location, const int32_t* knn_indices,
                                           const float* knn_dists, int64_t n, int64_t k, int32_t out_location,
                                           uint64_t* pairs, int64_t capacity, int64_t* n_pairs) {
  if (!ctx || !n_pairs) return EVOC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(ctx->mu);
  CallScope call(ctx);
  ctx->last_error.clear();
  *n_pairs = 0;
  if (n <= 0) return ctx->fail(EVOC_ERR_EMPTY_DATA, "empty kNN graph");
  if (k <= 0) return ctx->fail(EVOC_ERR_INVALID_NEIGHBORS, "k must be positive");
  if (n > (int64_t)INT32_MAX - 1 || k > (int64_t)INT32_MAX / 2)
    return ctx->fail(EVOC_ERR_UNSUPPORTED, "more than 2^31-2 rows");
  if (!knn_indices || !knn_dists || (capacity > 0 && !pairs)) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "null array");
  if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(EVOC_ERR_HIP, "hipSetDevice failed");
  hipStream_t st = ctx->stream;
  DevBuf own_i, own_d;
  const int32_t* d_inds = nullptr;
  const float* d_dist = nullptr;
  int rc = stage_knn(ctx, location, knn_indices, knn_dists, n, k, own_i, own_d, &d_inds, &d_dist);
  if (rc) return rc;
  const int64_t m = n * k;
  DevBuf keys, sorted, uniq, cnt, dmax, bad, temp;
  if ((rc = keys.alloc(ctx, (size_t)m * 8, "pair keys"))) return rc;
  if ((rc = sorted.alloc(ctx, (size_t)m * 8, "sorted pair keys"))) return rc;
  if ((rc = uniq.alloc(ctx, (size_t)m * 8, "distinct pair keys"))) return rc;
  if ((rc = cnt.alloc(ctx, 8, "distinct count"))) return rc;
  if ((rc = dmax.alloc(ctx, 4, "duplicate distance"))) return rc;
  if ((rc = bad.alloc(ctx, 4, "index check"))) return rc;
  size_t temp_bytes = 0;
  const unsigned bits = fuzzy_key_bits(n);
  hipError_t e = dup_sort_unique(nullptr, temp_bytes, (const uint64_t*)keys.ptr, (uint64_t*)sorted.ptr,
                                 (uint64_t*)uniq.ptr, (int64_t*)cnt.ptr, m, bits, st);
  if (e != hipSuccess) return ctx->hip_fail(e, "sort sizing");
  if ((rc = temp.alloc(ctx, temp_bytes > 0 ? temp_bytes : 1, "sort scratch"))) return rc;
  e = hipMemsetAsync(dmax.ptr, 0, 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(bad.ptr, 0, 4, st);
  if (e == hipSuccess)
    e = launch_dup_emit(d_inds, d_dist, n, (int)k, (uint32_t*)dmax.ptr, (uint64_t*)keys.ptr, (int32_t*)bad.ptr, st);
  if (e == hipSuccess)
    e = dup_sort_unique(temp.ptr, temp_bytes, (const uint64_t*)keys.ptr, (uint64_t*)sorted.ptr, (uint64_t*)uniq.ptr,
                        (int64_t*)cnt.ptr, m, bits, st);
  int64_t nu = 0;
  int32_t bad_h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nu, cnt.ptr, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&bad_h, bad.ptr, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return ctx->hip_fail(e, "find_duplicates");
  if (bad_h) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "kNN index outside [-1, n)");
  // the distinct keys end with the empty key when any slot was empty
  uint64_t last = 0;
  if (nu > 0) {
    e = hipMemcpy(&last, (uint64_t*)uniq.ptr + (nu - 1), 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return ctx->hip_fail(e, "find_duplicates");
    if (last == ~0ull) nu--;
  }
  *n_pairs = nu;
  if (nu > capacity) return ctx->fail(EVOC_ERR_INVALID_ARGUMENT, "pairs capacity < n_pairs (n_pairs is set)");
  if (nu > 0) {
    e = launch_dup_pairs((const uint64_t*)uniq.ptr, nu, n, (uint64_t*)sorted.ptr, st);
    if (e == hipSuccess) e = hipMemcpyAsync(pairs, sorted.ptr, (size_t)nu * 8, out_kind(out_location), st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ctx->hip_fail(e, "find_duplicates output");
  }
  return EVOC_OK;
}
