// Kernel argument blocks and host-side launchers of the gfx950 kernels.
// Shared by the .hip translation units and the C++ host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "evoc_mi355x.h"  // enum evoc_dtype

// element kinds of the MFMA tile engine
#define EVOC_ELEM_BF16 0
#define EVOC_ELEM_F32 1
#define EVOC_ELEM_I8 2
#define EVOC_ELEM_U8BITS 3  // binary rows, VALU popcount engine

namespace evoc {

// One query block against a range of candidate tiles (tile_rows() rows each).
struct TileArgs {
  const uint8_t* q_rows;   // query rows (row_bytes apart, zero padded)
  const uint8_t* c_rows;   // candidate rows; candidate c is row c * c_stride
  int64_t row_bytes;       // padded row pitch in bytes, multiple of 128
  int32_t n_q;             // queries
  int32_t n_c;             // candidates in candidate index space
  int32_t c_stride;        // row stride of the candidate sample (pass 1)
  int32_t n_ctiles;        // candidate tiles
  int32_t tiles_per_split; // candidate tiles per split
  int32_t q_offset;        // global index of query 0 (candidates are global)
  // grid: n_qblocks x n_splits workgroups, laid out by tile_coords()
  int32_t n_qblocks;       // query blocks (entries of block_list when given)
  int32_t n_splits;        // candidate splits
  int32_t group_q;         // query blocks per XCD group (they share candidate tiles)
  int32_t early_out;       // symmetric pass: 1 when most 32 x 32 blocks are expected to hold no key
  const int32_t* block_list;  // optional list of query blocks (re-scoring)
  const uint32_t* row_pop;    // binary rows: popcount per row
  // pass 1
  float* gmax;   // [n_q][ng]
  int32_t ng;
  int32_t skip_self;  // 1: a query never lists itself (brute_force_knn, src/knn.rs:204-206)
  // pass 2
  const float* tau;          // [n_q] score bound (>= passes)
  const uint64_t* tau_key;   // [n_q] optional key bound (re-scoring rounds)
  uint64_t* bufs;            // [splits][n_q][caps]
  int32_t* counts;           // [splits][n_q]
  int32_t caps;
  int32_t ccap;              // symmetric pass: candidate-side slots per row
  // symmetric pass 2 (query and candidate rows are the same matrix, q_offset 0):
  // each workgroup scores query block work[i].x against candidate tiles
  // [work[i].y, work[i].z) (all >= the query block) as split work[i].w, and
  // also appends every pair's key to the candidate row's list in cbufs
  const int4* work;          // [n_work]
  int32_t n_work;
  int32_t pad3;
  uint64_t* cbufs;           // [n_q][ccap] candidate-side keys
  int32_t* ccounts;          // [n_q] candidate-side key counts (may exceed ccap: overflow)
};

struct SelectArgs {
  const uint64_t* bufs;
  const int32_t* counts;
  int32_t n_q;
  int32_t n_splits;
  int32_t caps;
  int32_t K;
  const int32_t* query_list;  // optional
  int32_t n_list;
  int32_t q_offset;
  uint64_t* out_keys;     // [n_q][K]
  int32_t* overflow;      // [n_q] re-score flag: a buffer overflowed, or fewer than min_count keys
  uint64_t* tau_key_out;  // [n_q]
  // symmetric pass: the candidate-side list of each row, read as one more split
  const uint64_t* cbufs;  // [n_q][ccap] or null
  const int32_t* ccounts;
  int32_t ccap;
  // a query with fewer appended keys than this (over all segments, before the
  // caps) is flagged for re-scoring: its bound was an estimate that missed
  // (0: every bound of the pass was a guaranteed lower bound)
  int32_t min_count;
};

// One request of a batched exact search (the persistent batch kernels of
// knn_exact.hip and prep.hip).  Every pointer is into the batch workspace or the
// caller's device buffers.  The table of requests and the work-item lists live
// in device memory, so one captured launch sequence serves any batch.
struct BatchReq {
  TileArgs p1;            // threshold pass (has_p1)
  TileArgs p2;            // scoring pass (per-split)
  SelectArgs sel;         // K best keys per query (sel.out_keys)
  const uint8_t* x_in;    // caller rows on the device, ld_in elements apart
  int64_t ld_in;
  float* xf32;            // f32: prepared rows, f32_ld apart (the scored rows in the FP32 form)
  int64_t f32_ld;
  uint8_t* rows;          // BF16 / int8 scored rows, row_bytes apart (FP32 form: == xf32)
  int64_t row_bytes;
  float* norms;           // f32: row norms
  int32_t* flag;          // f32: prepare_float's renormalise-all flag
  float* gmax;            // threshold pass: [n][ng] group maxima
  float* tau;             // threshold pass: [n] score bounds
  uint64_t* rkeys;        // BF16 form: [n][k] FP32-re-scored keys
  int32_t* out_idx;       // [n][k]
  float* out_dist;        // [n][k]
  int32_t* ovf;           // set to 1 when any query's buffers overflowed
  int32_t n, d, K, k;
  int32_t ng, has_p1, dtype, bf16;
};
// Work-item lists of one batch: int4 items, one list per pass, at offsets into
// the item array.  rows items: (request, first row) of 64 rows; tile items:
// (request, query block, split); query items: (request, first query) of 4.
struct BatchHeader {
  int32_t n_reqs, n_rows, n_p1, n_p2, n_q;
  int32_t off_rows, off_p1, off_p2, off_q;
  int32_t pad[3];
};

// ---- launchers (knn_exact.hip) ----
int tile_shape(int kind);         // MFMA tile shape in force on this thread (-1 for binary rows)
void set_tile_shape(int shape);   // this thread's shape for the following launches (-1: default)
int tile_rows_c(int kind);        // candidate rows per tile
int tile_rows_q(int kind);        // query rows per tile
int groups_per_tile(int kind);    // pass-1 group maxima per candidate tile
size_t tile_lds_bytes(int kind, int pass);
size_t tile_lds_bytes_shape(int shape, int pass);  // LDS of the MFMA tile kernels of one shape
hipError_t launch_pass1(int kind, const TileArgs& a, int n_blocks, int n_splits, hipStream_t s);
hipError_t launch_pass2(int kind, const TileArgs& a, int n_blocks, int n_splits, hipStream_t s);
// out[q] = the rank-th largest group maximum of query q, for q < n_q or for
// the n_q queries of `list`
hipError_t launch_tau_select(const float* gmax, int n_q, int ng, int rank, float* out, hipStream_t s,
                             const int32_t* list = nullptr);
hipError_t launch_select(const SelectArgs& a, hipStream_t s);
// symmetric pass 2 over a.work (square tiles only: tile_rows_c == tile_rows_q)
hipError_t launch_pass2_sym(int kind, const TileArgs& a, hipStream_t s);
// zero ccounts of the queries in list[0..n)
hipError_t launch_zero_list(int32_t* ccounts, const int32_t* list, int n, hipStream_t s);
// order: optional permutation of the n_q queries to process them in (null: index order)
hipError_t launch_rescore(const float* X, int64_t ld, int d, int n_q, int q_offset,
                          const uint64_t* cand, int Kc, int k, uint64_t* out, hipStream_t s,
                          const int32_t* order = nullptr);
hipError_t launch_finalize(const uint64_t* keys, int n_q, int q_offset, int k, int dtype,
                           int32_t* idx, float* dist, hipStream_t s);
// The batched exact search of element kind `kind` (BF16, FP32 or I8; 128 x 128
// tiles; K <= 64) over the requests in `reqs`, every stage one persistent
// launch over its work items: clear flags, prepare rows, threshold pass,
// thresholds, scoring pass, selection, FP32 re-score (BF16 requests), output.
// Every kernel argument is fixed (the table, items and header live at fixed
// device addresses; `hdr` there, `h` its host copy sets the grids), so the
// sequence can be captured once and replayed for any batch, with
// batch_set_grids fitting the captured grids to each batch's item counts.
// Before any launch the dynamic LDS limits must have been raised
// (batch_prepare_attributes, outside any capture).
hipError_t batch_prepare_attributes(int kind);
hipError_t launch_batch(int kind, int n_cu, BatchReq* reqs, const int4* items, const BatchHeader* hdr,
                        const BatchHeader& h, hipStream_t s);
hipError_t batch_set_grids(hipGraph_t graph, hipGraphExec_t exec, int kind, int n_cu, const BatchHeader& h);

// ---- launchers (prep.hip) ----
// prepare_float (src/knn.rs:78-113) on the device, fused with padding and the
// optional BF16 copy: norms/flag are scratch ([n] floats, one int).
hipError_t launch_prepare_float(const float* X, int64_t ld, int64_t n, int d, float* out,
                                int64_t out_ld, uint16_t* out_bf16, int64_t bf_ld, float* norms,
                                int32_t* flag, hipStream_t s);
// batched input preparation (BatchReq): prepare_float for f32 requests (norms,
// renormalise-all flag, prepared FP32 and BF16 rows), padding for int8 ones
hipError_t launch_batch_prepare(unsigned norm_grid, unsigned write_grid, BatchReq* reqs, const int4* items,
                                const BatchHeader* hdr, hipStream_t s);
const void* batch_norm_kernel();   // the two preparation kernels (graph node identification)
const void* batch_write_kernel();
hipError_t launch_pad_rows(const uint8_t* X, int64_t ld_bytes, int64_t n, int row_bytes,
                           uint8_t* out, int64_t out_ld, hipStream_t s);
hipError_t launch_row_popcount(const uint8_t* X, int64_t ld, int64_t n, int row_bytes,
                               uint32_t* pop, hipStream_t s);

// ---- launchers (fuzzy.hip): locality order, the fuzzy graph and find_duplicates ----
// A permutation of the n rows ordered by their projection on a fixed +-1
// vector (rows of one cluster come out together).  keys / keys_out: n uint32,
// order_in: n int32 scratch, order: n int32 result; temp == nullptr queries
// temp_bytes.
hipError_t locality_order(void* temp, size_t& temp_bytes, const float* X, int64_t ld, int64_t n, int d,
                          uint32_t* keys, uint32_t* keys_out, int32_t* order_in, int32_t* order, hipStream_t s);
size_t fuzzy_partial_doubles();  // scratch doubles of launch_fuzzy_sigma's mean
hipError_t launch_fuzzy_sigma(const float* dist, int64_t n, int k, float n_neighbors, double* partial,
                              float* sigma, float* rho, hipStream_t s);
// 2 n k (key, value) slots; keys row * n + col, empty slots all ones
hipError_t launch_fuzzy_emit(const int32_t* inds, const float* dist, int64_t n, int k, const float* sigma,
                             const float* rho, int symmetrize, uint64_t* keys, float* vals, int32_t* bad,
                             hipStream_t s);
unsigned fuzzy_key_bits(int64_t n);  // key bits the sorts cover
hipError_t fuzzy_sort_pairs(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_out,
                            const float* vals_in, float* vals_out, int64_t m, unsigned bits, hipStream_t s);
hipError_t launch_fuzzy_csr(const uint64_t* sorted_keys, int64_t m, int64_t n, int64_t* indptr, hipStream_t s);
hipError_t launch_fuzzy_cols(const uint64_t* sorted_keys, int64_t nnz, int64_t n, int32_t* cols, hipStream_t s);
// n k keys min * n + max of duplicate pairs, empty slots all ones; dmax: one zeroed uint32
hipError_t launch_dup_emit(const int32_t* inds, const float* dist, int64_t n, int k, uint32_t* dmax,
                           uint64_t* keys, int32_t* bad, hipStream_t s);
// m distinct keys a * n + b -> pairs (a << 32 | b)
hipError_t launch_dup_pairs(const uint64_t* keys, int64_t m, int64_t n, uint64_t* pairs, hipStream_t s);
hipError_t dup_sort_unique(void* temp, size_t& temp_bytes, const uint64_t* keys_in, uint64_t* keys_sorted,
                           uint64_t* keys_unique, int64_t* n_unique, int64_t m, unsigned bits, hipStream_t s);

}  // namespace evoc
