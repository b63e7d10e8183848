// Host-side internals shared by the translation units of libnsp_amd.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/nsp_amd.h"

namespace nsp {

// Row bins.  Symbolic bins are chosen by the intermediate-product count of a
// row (an upper bound of its nnz), numeric bins by the exact nnz.  Bins
// 0..5 are handled by sub-wave or wave-per-row kernels with LDS tables,
// 6..7 by block-per-row kernels, 8 by the column-window kernels.
constexpr int kBins = 9;
constexpr int kHeavyBin = 8;

// Upper bound (inclusive) of each symbolic bin, on intermediate products.
constexpr int64_t kSymBound[kBins] = {32, 64, 128, 256, 512, 1024, 2048, 16384, INT64_MAX};
// Upper bound (inclusive) of each numeric bin, on row nnz.
constexpr int64_t kNumBound[kBins] = {16, 32, 64, 128, 256, 512, 2048, 4096, INT64_MAX};

// Stride (in 32-bit words) between the work-queue heads in Info.
constexpr int kQueueStride = 32;

// Per-call counters living in device memory; mirrored to pinned host memory.
struct Info {
  int64_t count_sym[16];
  int64_t cursor_sym[16];
  int64_t count_num[16];
  int64_t cursor_num[16];
  int64_t nprod;
  int64_t nnz;
  int64_t max_intprod;
  int64_t max_nz;
  uint32_t tile_ticket;
  uint32_t error;
  int64_t cap;  // capacity of C's column/value buffers (numeric rows beyond it are skipped)
  int64_t nlong;  // rows binned by the long-row pass
  int64_t smp_max;    // one-call product: most distinct columns of a sampled row
  int64_t spec_over;  // one-call product: a row had more columns than its slot
  // heads of the numeric row-bin work queues (one per XCD), each on a
  // 128-byte line of its own: draws on one queue do not queue behind draws
  // on another
  alignas(128) uint32_t queue_num[8 * kQueueStride];
};

// A batch of independent products stored as concatenated rows: matrix b owns
// rows [a_off[b], a_off[b+1]) of A and [b_off[b], b_off[b+1]) of B (device
// arrays of count+1 entries); columns are local to each matrix.
struct Batch {
  const int64_t *a_off = nullptr;
  const int64_t *b_off = nullptr;
  int64_t count = 0;
};

struct Workspace {
  void *base = nullptr;
  size_t bytes = 0;
};

struct Handle {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  nsp_stats_t stats{};

  // tuning knobs
  int grid_cap = 32;  // blocks per CU of the grid-stride launches: 8 rounds of 4 resident blocks, so that
                      // the dispatcher evens out the tail (27-point share: numeric 0.85 -> 0.81 ms against 8)
  int sym_small_rows_per_wave = 4;
  int sym_rows_per_wave = 2;
  int num_small_rows_per_wave = 4;
  int num_rows_per_wave = 2;
  int heavy_blocks = 0;  // 0: one block per CU
  // the backward may reuse ab_b0/ab_len when they were computed for the same
  // A and B buffers (the caller promises fixed patterns)
  int reuse_extents = 0;
  const void *ext_key[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t ext_rows = -1, ext_nnz = -1;
  int64_t ext_nprod = -1;  // nprod of the cached extents (known once read back)
  int num_table_factor = 2;
  int sym_table_factor = 2;  // symbolic slots per product bound in bins 0..3 (1 or 2)
  int narrow_mode = 1;       // 1: column bitmaps instead of hash tables when B has at most 4096 columns
  int fused_mode = 1;        // one-call products: 1 numeric pass into per-row slots + compaction when a sample allows
  int queue_mode = 1;        // 1: narrow numeric bins 0..5 in one launch fed by row-bin work queues
  int queue_chunk = 16;      // tiles per work-queue ticket
  int queue_blocks = 4;      // blocks per CU of the work-queue launch
  int fp32_accumulate = 0;   // 1: FP32 values summed in FP32 where the kernel offers it (narrow products)
  int64_t coo_chunk = int64_t(1) << 21;  // records per streamed chunk of the COO reader
  bool timing = false;
  bool timing_pending = false;
  hipEvent_t events[8] = {};

  // reusable device workspace (ws: product; aux: compaction and gathers)
  Workspace ws;
  Workspace aux;
  Workspace aux2;  // transpose
  Workspace bwd;   // backward: flags, row lists, product offsets, contributions
  Workspace segw;  // segment sort: tile lists, merge buffer
  Workspace rsort; // row sort: extents, keys, values of the rows being sorted
  Workspace coo;   // COO conversion: row bitmap, unsorted-row list, staging
  Workspace spec;  // one-call product: C rows in fixed-size slots before compaction
  void *pin = nullptr;  // pinned host staging of the COO reader
  size_t pin_bytes = 0;
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  bool poff_valid = false;  // bwd holds the product offsets of the cached extents
  const void *wpos_key = nullptr;  // A^T permutation whose product positions bwd holds
  // longest row of the C pattern the backward last ran on (its buffers and sizes)
  const void *bwd_c_key[2] = {nullptr, nullptr};
  int64_t bwd_c_rows = -1, bwd_c_nnz = -1, bwd_cmax = 0;
  int bwd_contrib = 1;  // dB: 1 contributions, 0 search in C's rows, 2 by C's longest row
  Info *d_info = nullptr;
  Info *h_info = nullptr;  // pinned

  // state carried from symbolic to numeric
  bool have_plan = false;
  // one-call product: C's rows wait in `spec` (slots of fused_ns entries)
  // for the compaction into the caller's buffers
  bool fused_plan = false;
  bool timing_fused = false;
  int64_t fused_ns = 0;
  const void *fused_key[2] = {nullptr, nullptr};
  // slot bin of the last one-call product that took the slot pass, with its
  // buffers and sizes (-1: none)
  int fc_bin = -1, fc_last = -1;
  const void *fc_key[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t fc_dims[4] = {0, 0, 0, 0};
  int64_t plan_rows = -1, plan_nnz = -1;
  const void *plan_a = nullptr, *plan_b = nullptr;
  int plan_sub = 16;
  int64_t num_count[kBins] = {0};
  int64_t *a_rpt64 = nullptr, *c_rpt64 = nullptr, *intprod = nullptr;
  int64_t *ab_b0 = nullptr;   // per A entry: start of the B row it selects
  int32_t *ab_len = nullptr;  // per A entry: length of that B row
  int32_t *row_nz = nullptr, *perm_sym = nullptr, *perm_num = nullptr;
  uint8_t *cls = nullptr;
  uint64_t *tile_status = nullptr;
  unsigned long long *ext_partial = nullptr;  // k_extents' per-block product sums (light row info)
  int64_t ntiles = 0;
};

inline Handle *handle_of(nsp_handle_t hs);

int set_error(Handle *h, int status, const std::string &msg);
int collect_timing(Handle *h);

#define NSP_HIP(h, call)                                                              \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return ::nsp::set_error((h), NSP_ERR_HIP,                                        \
                              std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define NSP_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ != NSP_OK) return s_; \
  } while (0)

}  // namespace nsp

// The opaque handle of the C ABI.
struct nsp_handle_s {
  nsp::Handle h;
};

inline nsp::Handle *nsp::handle_of(nsp_handle_t hs) { return &hs->h; }
