// Column compaction ("shrink") and B-row gather of one rank's local share, on
// one device.
//
// Reference path replaced: get_shrinked_col / shrink_col / get_shrinked_matrix
// (src/spspmpi.cu:106-182), getmct_4shrink with its compact -> thrust::sort ->
// unique -> three-branch concatenation (src/getmct.cu:135-220), the binary
// search remap apply_mask_permut_GPU_noSideEffects(_glob) (src/spspmpi.cu:
// 58-104), and the completedP assembly from exchanged halo rows
// (src/spspmpi.cu:584-906).
//
// The reference's permutation is sorted(unique(cols U [0..l])) for every input
// it accepts, so it is built here without a sort: flag every column value in
// the bounded range [lo, hi) U [0, l], exclusive-scan the flags, and the scan
// is both the permutation's inverse (new column = scan[col - base]) and its
// scatter target.  Integer-only, HBM-bound.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_utils.hpp"
#include "nsp_internal.hpp"
#include "scan.hpp"
#include "spgemm_internal.hpp"

namespace nsp {

__global__ void k_zero_u64(uint64_t *p, int64_t n, uint32_t *ticket) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = t; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
  if (t == 0) {
    ticket[0] = 0;  // scan ticket
    ticket[1] = 0;  // error flag
  }
}

__global__ void k_shrink_init(int64_t range, int64_t base, int64_t l, uint8_t *__restrict__ flags) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < range;
       x += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = x + base;
    flags[x] = (c >= 0 && c <= l) ? 1 : 0;
  }
}

__global__ void k_shrink_mark(int64_t nnz, const int32_t *__restrict__ col, int64_t base, int64_t range,
                              uint8_t *__restrict__ flags, uint32_t *__restrict__ bad) {
#pragma unroll 4
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = (int64_t)col[e] - base;
    if (x >= 0 && x < range) flags[x] = 1;
    else *bad = 1;
  }
}

// A column outside the flagged range (already reported by k_shrink_mark)
// maps to -1 instead of indexing past the position array.
__global__ void k_shrink_remap(int64_t nnz, const int32_t *__restrict__ col, int64_t base, int64_t range,
                               const int32_t *__restrict__ pos, int32_t *__restrict__ new_col) {
#pragma unroll 4
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = (int64_t)col[e] - base;
    new_col[e] = (x >= 0 && x < range) ? pos[x] : -1;
  }
}

__global__ void k_shrink_perm(int64_t range, int64_t base, const uint8_t *__restrict__ flags,
                              const int32_t *__restrict__ pos, int32_t *__restrict__ perm, int64_t cap) {
  for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < range;
       x += (int64_t)gridDim.x * blockDim.x)
    if (flags[x] && pos[x] < cap) perm[pos[x]] = (int32_t)(x + base);
}

template <typename RowT>
__global__ void k_gather_len(int64_t nr, const int32_t *__restrict__ rows, int64_t row_offset, int64_t MB,
                             const RowT *__restrict__ brpt, int64_t *__restrict__ len, uint32_t *bad) {
#pragma unroll 4
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = (int64_t)rows[i] + row_offset;
    if (g < 0 || g >= MB) {
      len[i] = 0;
      *bad = 1;
    } else {
      len[i] = (int64_t)brpt[g + 1] - (int64_t)brpt[g];
    }
  }
}

// A block copies kGatherRows gathered rows at a time as one run of output
// entries: the rows' source and destination starts are staged in LDS, then
// consecutive threads copy consecutive output entries (coalesced reads and
// writes across row boundaries; a stencil share's rows hold only ~27
// entries), each finding its row by a binary search over the staged
// destination starts.  Entries are taken only inside their source row, so a
// destination row pointer that disagrees with B's row lengths cannot read
// outside B; rows outside B copy nothing.
constexpr int kGatherRows = 64;

template <typename RowT, typename V>
__global__ __launch_bounds__(256) void k_gather_fill(int64_t nr, const int32_t *__restrict__ rows,
                                                     int64_t row_offset, int64_t MB,
                                                     const RowT *__restrict__ brpt,
                                                     const int32_t *__restrict__ bcol,
                                                     const V *__restrict__ bval, int64_t col_shift,
                                                     const RowT *__restrict__ orpt,
                                                     int32_t *__restrict__ ocol, V *__restrict__ oval) {
  __shared__ int64_t s_dst[kGatherRows + 1];  // destination start of each row, then the run's end
  __shared__ int64_t s_src[kGatherRows];      // source start of each row (-1: row outside B)
  __shared__ int64_t s_len[kGatherRows];      // source length of each row
  for (int64_t i0 = blockIdx.x * (int64_t)kGatherRows; i0 < nr; i0 += (int64_t)gridDim.x * kGatherRows) {
    const int nrow = (int)(nr - i0 < kGatherRows ? nr - i0 : kGatherRows);
    if ((int)threadIdx.x <= nrow) {
      const int64_t i = i0 + threadIdx.x;
      s_dst[threadIdx.x] = (int64_t)orpt[i];  // i <= nr: the row pointer has nr + 1 entries
      if ((int)threadIdx.x < nrow) {
        const int64_t g = (int64_t)rows[i] + row_offset;
        const bool in = g >= 0 && g < MB;
        s_src[threadIdx.x] = in ? (int64_t)brpt[g] : -1;
        s_len[threadIdx.x] = in ? (int64_t)brpt[g + 1] - (int64_t)brpt[g] : 0;
      }
    }
    __syncthreads();
    const int64_t D0 = s_dst[0], D1 = s_dst[nrow];
#pragma unroll 4
    for (int64_t e = D0 + threadIdx.x; e < D1; e += blockDim.x) {
      int lo = 0, hi = nrow;  // s_dst[lo] <= e < s_dst[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_dst[mid] <= e) lo = mid;
        else hi = mid;
      }
      const int64_t k = e - s_dst[lo];
      if (s_src[lo] >= 0 && k >= 0 && k < s_len[lo]) {
        ocol[e] = (int32_t)((int64_t)bcol[s_src[lo] + k] - col_shift);
        oval[e] = bval[s_src[lo] + k];
      }
    }
    __syncthreads();  // the staged starts are rewritten by the next run
  }
}

// ---------------------------------------------------------------------------
static int aux_reserve(Handle *h, size_t bytes) {
  if (bytes <= h->aux.bytes) return NSP_OK;
  if (h->aux.base) NSP_HIP(h, hipFree(h->aux.base));
  h->aux.base = nullptr;
  h->aux.bytes = 0;
  const size_t want = bytes + bytes / 4;
  if (hipMalloc(&h->aux.base, want) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, "auxiliary workspace allocation failed");
  }
  h->aux.bytes = want;
  return NSP_OK;
}

static int grid1d(const Handle *h, int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)h->num_cu * 16));
}

// Scratch layout for one scan of n items: status words, ticket, error flag.
struct ScanScratch {
  uint64_t *status;
  uint32_t *ticket;
  uint32_t *bad;
  int64_t ntiles;
};

static size_t scan_scratch_bytes(int64_t n) { return (size_t)(n / kScanTile + 2) * 8 + 256; }

static ScanScratch carve_scan(char *p, int64_t n) {
  ScanScratch s;
  s.ntiles = n / kScanTile + 2;
  s.status = (uint64_t *)p;
  s.ticket = (uint32_t *)(p + s.ntiles * 8);
  s.bad = s.ticket + 1;
  return s;
}

int shrink_cols(Handle *h, int64_t nnz, const int32_t *col, int64_t l, int64_t lo, int64_t hi,
                int32_t *new_col, int32_t *perm, int64_t perm_cap, int64_t *shrinked_m) {
  if (nnz < 0 || l < -1 || hi < lo || !shrinked_m || (nnz > 0 && (!col || !new_col)) || !perm)
    return set_error(h, NSP_ERR_INVALID, "nsp_shrink_cols: bad arguments");
  const int64_t base = std::min<int64_t>(lo, 0);
  const int64_t top = std::max<int64_t>(hi, l + 1);
  const int64_t range = top - base;
  if (range >= INT32_MAX) return set_error(h, NSP_ERR_UNSUPPORTED, "column range >= 2^31");
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t bytes = al(range) + al((range + 1) * 4) + al(scan_scratch_bytes(range));
  NSP_TRY(aux_reserve(h, bytes));
  char *p = (char *)h->aux.base;
  uint8_t *flags = (uint8_t *)p;
  int32_t *pos = (int32_t *)(p + al(range));
  ScanScratch sc = carve_scan(p + al(range) + al((range + 1) * 4), range);
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid1d(h, sc.ntiles)), dim3(256), 0, st, sc.status, sc.ntiles, sc.ticket);
  hipLaunchKernelGGL(k_shrink_init, dim3(grid1d(h, range)), dim3(256), 0, st, range, base, l, flags);
  if (nnz > 0)
    hipLaunchKernelGGL(k_shrink_mark, dim3(grid1d(h, nnz)), dim3(256), 0, st, nnz, col, base, range, flags,
                       sc.bad);
  hipLaunchKernelGGL((k_exscan<uint8_t, int32_t>), dim3((unsigned)(range / kScanTile + 1)), dim3(kScanBS), 0,
                     st, range, flags, pos, sc.status, sc.ticket);
  if (nnz > 0)
    hipLaunchKernelGGL(k_shrink_remap, dim3(grid1d(h, nnz)), dim3(256), 0, st, nnz, col, base, range, pos,
                       new_col);
  hipLaunchKernelGGL(k_shrink_perm, dim3(grid1d(h, range)), dim3(256), 0, st, range, base, flags, pos, perm,
                     perm_cap);
  NSP_HIP(h, hipGetLastError());
  int32_t total = 0;
  uint32_t bad = 0;
  NSP_HIP(h, hipMemcpyAsync(&total, pos + range, 4, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipMemcpyAsync(&bad, sc.bad, 4, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipStreamSynchronize(st));
  if (bad) return set_error(h, NSP_ERR_INVALID, "nsp_shrink_cols: a column lies outside [lo, hi)");
  *shrinked_m = total;
  if (total > perm_cap) return set_error(h, NSP_ERR_INVALID, "nsp_shrink_cols: perm capacity too small");
  return NSP_OK;
}

template <typename RowT>
static int gather_count_impl(Handle *h, const nsp_csr_t *B, int64_t nr, const int32_t *rows, int64_t row_offset,
                             void *out_rpt, int64_t *out_nnz) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  NSP_TRY(aux_reserve(h, al(nr * 8 + 8) + al(scan_scratch_bytes(nr))));
  char *p = (char *)h->aux.base;
  int64_t *len = (int64_t *)p;
  ScanScratch sc = carve_scan(p + al(nr * 8 + 8), nr);
  hipStream_t st = h->stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid1d(h, sc.ntiles)), dim3(256), 0, st, sc.status, sc.ntiles, sc.ticket);
  if (nr > 0)
    hipLaunchKernelGGL(k_gather_len<RowT>, dim3(grid1d(h, nr)), dim3(256), 0, st, nr, rows, row_offset, B->rows,
                       (const RowT *)B->row_ptr, len, sc.bad);
  hipLaunchKernelGGL((k_exscan<int64_t, RowT>), dim3((unsigned)(nr / kScanTile + 1)), dim3(kScanBS), 0, st, nr,
                     len, (RowT *)out_rpt, sc.status, sc.ticket);
  NSP_HIP(h, hipGetLastError());
  RowT total = 0;
  uint32_t bad = 0;
  NSP_HIP(h, hipMemcpyAsync(&total, (RowT *)out_rpt + nr, sizeof(RowT), hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipMemcpyAsync(&bad, sc.bad, 4, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipStreamSynchronize(st));
  if (bad) return set_error(h, NSP_ERR_INVALID, "nsp_gather_rows_count: a row id lies outside B");
  *out_nnz = (int64_t)total;
  return NSP_OK;
}

int gather_rows_count(Handle *h, nsp_index_t it, const nsp_csr_t *B, int64_t nr, const int32_t *rows,
                      int64_t row_offset, void *out_rpt, int64_t *out_nnz) {
  if (!B || !B->row_ptr || nr < 0 || (nr > 0 && !rows) || !out_rpt || !out_nnz)
    return set_error(h, NSP_ERR_INVALID, "nsp_gather_rows_count: bad arguments");
  if (it == NSP_IDX32) return gather_count_impl<int32_t>(h, B, nr, rows, row_offset, out_rpt, out_nnz);
  return gather_count_impl<int64_t>(h, B, nr, rows, row_offset, out_rpt, out_nnz);
}

int gather_rows_fill(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *B, int64_t nr,
                     const int32_t *rows, int64_t row_offset, int64_t col_shift, nsp_csr_t *out) {
  if (!B || !out || !out->row_ptr || nr < 0 || (nr > 0 && !rows))
    return set_error(h, NSP_ERR_INVALID, "nsp_gather_rows_fill: bad arguments");
  if (nr == 0 || out->nnz == 0) return NSP_OK;
  const int g = (int)std::max<int64_t>(1, std::min<int64_t>((nr + kGatherRows - 1) / kGatherRows, (int64_t)h->num_cu * 8));
  hipStream_t st = h->stream;
#define GATHER(RT, VT)                                                                                \
  hipLaunchKernelGGL((k_gather_fill<RT, VT>), dim3(g), dim3(256), 0, st, nr, rows, row_offset, B->rows, \
                     (const RT *)B->row_ptr, B->col, (const VT *)B->val, col_shift,                  \
                     (const RT *)out->row_ptr, out->col, (VT *)out->val)
  if (it == NSP_IDX32 && vt == NSP_F64) GATHER(int32_t, double);
  else if (it == NSP_IDX32) GATHER(int32_t, float);
  else if (vt == NSP_F64) GATHER(int64_t, double);
  else GATHER(int64_t, float);
#undef GATHER
  NSP_HIP(h, hipGetLastError());
  return NSP_OK;
}

}  // namespace nsp
