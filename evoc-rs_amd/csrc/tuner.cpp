// Tile-shape autotuner.
//
// The MFMA scorer has four tile shapes (knn_exact.hip).  Which is fastest
// depends on the element kind and on the problem size: small matrices need
// small tiles to fill 256 CUs, large ones want the 256 x 256 tile with the
// least operand traffic per MFMA.  evoc_mi355x_autotune times a knn_graph call
// on (a slice of) the caller's rows once per shape and records the fastest for
// that (kind, row-count bucket, row-bytes bucket); exact_search then runs every
// call in that bucket with it.  Shapes never change results, only time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "evoc_host.h"
#include "evoc_kernels.h"
#include "evoc_mi355x.h"

using namespace evoc;

namespace {

std::mutex g_mu;
std::map<std::tuple<int, int, int>, int> g_table;  // (kind, log2 n, row_bytes / 128) -> shape
thread_local int t_force = -1;                     // shape under test (autotuner)

std::tuple<int, int, int> bucket(int kind, int64_t n, int64_t row_bytes) {
  const int nb = n > 1 ? (int)std::floor(std::log2((double)n)) : 0;
  return std::make_tuple(kind, nb, (int)((row_bytes + 127) / 128));
}

}  // namespace

int choose_tile_shape(int kind, int64_t n, int64_t row_bytes) {
  if (kind == EVOC_ELEM_U8BITS) return -1;
  if (t_force >= 0) return t_force;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(bucket(kind, n, row_bytes));
    if (it != g_table.end()) return it->second;
  }
  // untuned: the 256 x 256 tile once there are enough tiles to fill the chip
  return n >= 16384 ? 2 : 0;
}

extern "C" int evoc_mi355x_autotune(evoc_mi355x_ctx* ctx, const evoc_knn_input* in, const evoc_knn_options* o,
                                    int64_t max_rows, int32_t* best_shape, double ms[4]) {
  if (!ctx || !in || !o || in->n <= 1) return EVOC_ERR_INVALID_ARGUMENT;
  if (in->dtype == EVOC_DTYPE_U8) {  // one (VALU) engine: nothing to choose
    if (best_shape) *best_shape = -1;
    return EVOC_OK;
  }
  evoc_knn_input slice = *in;
  slice.n = max_rows > 0 ? std::min<int64_t>(in->n, max_rows) : in->n;
  evoc_knn_options opts = *o;
  opts.mode = EVOC_MODE_EXACT;
  opts.strict_precision = 0;
  const int64_t k = evoc_knn_effective_k(opts.n_neighbors, slice.n);
  const int64_t n = slice.n;
  int32_t* d_idx = nullptr;
  float* d_dist = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(&d_idx, (size_t)(n * k * 4)) != hipSuccess ||
      hipMalloc(&d_dist, (size_t)(n * k * 4)) != hipSuccess) {
    (void)hipFree(d_idx);
    (void)hipGetLastError();
    return ctx->fail(EVOC_ERR_HIP, "autotune: output allocation failed");
  }
  const int64_t tau[3] = {1, 2, 3};
  double best_ms = 1e30;
  int best = -1, rc = EVOC_OK;
  // every shape must reproduce the first shape's graph bit for bit
  std::vector<int32_t> ref_idx, got_idx;
  std::vector<float> ref_dist, got_dist;
  for (int shape = 0; shape < 4 && rc == EVOC_OK; shape++) {
    t_force = shape;
    double t = 1e30;
    for (int rep = 0; rep < 3 && rc == EVOC_OK; rep++) {  // first run warms up
      evoc_knn_output out = {};
      out.location = EVOC_MEM_DEVICE;
      out.indices = d_idx;
      out.distances = d_dist;
      out.capacity_k = k;
      rc = evoc_mi355x_knn_graph_seeded(ctx, &slice, &opts, tau, nullptr, 0, &out);
      if (rep > 0) t = std::min(t, out.device_seconds * 1e3);
    }
    if (rc != EVOC_OK) break;
    std::vector<int32_t>& hi = shape == 0 ? ref_idx : got_idx;
    std::vector<float>& hd = shape == 0 ? ref_dist : got_dist;
    hi.resize((size_t)(n * k));
    hd.resize((size_t)(n * k));
    if (hipMemcpy(hi.data(), d_idx, hi.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hd.data(), d_dist, hd.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = ctx->fail(EVOC_ERR_HIP, "autotune: result copy failed");
      break;
    }
    if (shape > 0 && (hi != ref_idx || std::memcmp(hd.data(), ref_dist.data(), hd.size() * 4) != 0)) {
      rc = ctx->fail(EVOC_ERR_UNSUPPORTED, "autotune: tile shapes produced different graphs");
      break;
    }
    if (ms) ms[shape] = t;
    if (t < best_ms) {
      best_ms = t;
      best = shape;
    }
  }
  t_force = -1;
  (void)hipFree(d_idx);
  (void)hipFree(d_dist);
  if (rc != EVOC_OK) return rc;
  const int kind = in->dtype == EVOC_DTYPE_I8 ? EVOC_ELEM_I8
                   : opts.precision == EVOC_PREC_BF16_RESCORE ? EVOC_ELEM_BF16 : EVOC_ELEM_F32;
  const int64_t elem = kind == EVOC_ELEM_BF16 ? 2 : kind == EVOC_ELEM_F32 ? 4 : 1;
  const int64_t row_bytes = (in->d * elem + 127) / 128 * 128;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table[bucket(kind, in->n, row_bytes)] = best;
  }
  if (best_shape) *best_shape = best;
  return EVOC_OK;
}
