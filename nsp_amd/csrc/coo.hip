// Binary COO records -> device CSR: the reference's matrix loader on the
// device.
//
// Reference path replaced: READ_BINARY_MATRIX_MPI's header read, row-range
// binary search and record unpacking (src/READ_MATRIX_MPI.cpp:83-130,
// 146-356), and load_MatrixMPI's irow2iat2, check_and_fix_order and
// apply_shift (src/distribute.cu:998-1069, 1071-1125).
//
// The file's records {int64 i, int64 j, double v} are sorted by row, so the
// row pointer needs no counting: record q starts row r when the record
// before it is in an earlier row, and then row_ptr[x] = q for every row x
// after that earlier row up to r (empty rows in between included).  Every
// record is converted independently (col = j - 1 - col_shift, val = v), so
// the file streams through the device in chunks; a chunk is told the last
// record of the chunk before it.  Rows whose columns go down somewhere are
// recorded once each (a bitmap test-and-set, then a list) and sorted
// afterwards with keys (column, position in the row), which reproduces the
// order the reference's stable bubble sort leaves.  One 24-byte read and a
// 12-byte write per record: HBM-bound.
#include <fcntl.h>
#include <stdio.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "nsp_internal.hpp"
#include "segsort.hpp"

namespace nsp {

constexpr int kCooBlock = 256;

// Control words of one conversion in device memory.
struct CooCtl {
  uint32_t err;          // bit 0: records out of row order or outside the rows; bit 1: column overflow;
                         // bit 2: an ASCII record line not ending in a newline
  uint32_t pad;
  unsigned long long nunsorted;  // rows listed for sorting
};

__global__ void k_coo_init(uint32_t *__restrict__ bits, int64_t nwords, CooCtl *ctl) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t i = t; i < nwords; i += (int64_t)gridDim.x * blockDim.x) bits[i] = 0;
  if (t == 0) {
    ctl->err = 0;
    ctl->pad = 0;
    ctl->nunsorted = 0;
  }
}

template <typename RowT>
__global__ void k_coo_fill_rpt(RowT *rpt, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    rpt[i] = 0;
}

// One chunk of m records (words = 3 per record), global record numbers
// q0 .. q0+m-1 of nrec.  (prev_i, prev_j) is record q0-1 when q0 > 0.
// Row r (0-based, local) holds records with i = row1 + r, r < nrows.
// Records are staged through LDS in tiles of kCooTile: with VEC (words
// 16-byte aligned) a full tile is three 16-byte loads per thread, all
// issued before the LDS stores.  Full tiles and the one partial tail tile
// are separate calls with the tile size known to each: a tile size selected
// with min(kCooTile, m - base) next to the full-tile branch was compiled by
// this hipcc into a select that read a stale scalar condition code.
constexpr int kCooTile = 2 * kCooBlock;

struct CooArgs {
  const uint64_t *words;
  int64_t q0, prev_i, prev_j, row1, nrows, nrec, col_shift;
  uint32_t *bits;
  int64_t *list;
  CooCtl *ctl;
};

template <typename RowT, typename V, bool VEC>
__device__ __forceinline__ void coo_tile(const CooArgs &a, RowT *__restrict__ rpt, int32_t *__restrict__ col,
                                         V *__restrict__ val, uint64_t *W, int64_t base, int nb) {
  if (VEC) {
    // nb == kCooTile: 3*kCooTile words = 3*kCooBlock 16-byte vectors
    const uint4 *src = reinterpret_cast<const uint4 *>(a.words + 3 * base);
    uint4 *dst = reinterpret_cast<uint4 *>(W);
    const uint4 x0 = src[threadIdx.x], x1 = src[threadIdx.x + kCooBlock], x2 = src[threadIdx.x + 2 * kCooBlock];
    dst[threadIdx.x] = x0;
    dst[threadIdx.x + kCooBlock] = x1;
    dst[threadIdx.x + 2 * kCooBlock] = x2;
  } else {
    for (int k = threadIdx.x; k < 3 * nb; k += kCooBlock) W[k] = a.words[3 * base + k];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += kCooBlock) {
    const int64_t i = (int64_t)W[3 * t], j = (int64_t)W[3 * t + 1];
    const double v = __longlong_as_double((long long)W[3 * t + 2]);
    int64_t ip, jp;
    if (t > 0) {
      ip = (int64_t)W[3 * t - 3];
      jp = (int64_t)W[3 * t - 2];
    } else if (base > 0) {
      ip = (int64_t)a.words[3 * base - 3];
      jp = (int64_t)a.words[3 * base - 2];
    } else {
      ip = a.prev_i;
      jp = a.prev_j;
    }
    const int64_t q = a.q0 + base + t;
    const int64_t r = i - a.row1;
    uint32_t err = 0;
    if (r < 0 || r >= a.nrows) {
      err |= 1;
    } else {
      if (q == 0 || ip != i) {
        if (q > 0 && ip > i) {
          err |= 1;  // rows out of order
        } else {
          // rows after the previous record's row up to this one start here
          const int64_t x0 = (q == 0) ? 0 : std::max<int64_t>(0, ip - a.row1 + 1);
          for (int64_t x = x0; x <= r; x++) rpt[x] = (RowT)q;
        }
      } else if (j < jp) {
        const uint32_t bit = 1u << (r & 31);
        if (!(atomicOr(&a.bits[r >> 5], bit) & bit)) a.list[atomicAdd(&a.ctl->nunsorted, 1ull)] = r;
      }
      if (q == a.nrec - 1)
        for (int64_t x = r + 1; x <= a.nrows; x++) rpt[x] = (RowT)a.nrec;
    }
    const int64_t c = j - 1 - a.col_shift;
    if (c < INT32_MIN || c > INT32_MAX) err |= 2;
    col[q] = (int32_t)c;
    val[q] = (V)v;
    if (err) atomicOr(&a.ctl->err, err);
  }
  __syncthreads();
}

template <typename RowT, typename V, bool VEC>
__global__ __launch_bounds__(kCooBlock) void k_coo_chunk(CooArgs a, int64_t m, RowT *__restrict__ rpt,
                                                         int32_t *__restrict__ col, V *__restrict__ val) {
  __shared__ __align__(16) uint64_t W[3 * kCooTile];
  const int64_t nfull = m / kCooTile;
  for (int64_t tile = blockIdx.x; tile < nfull; tile += gridDim.x)
    coo_tile<RowT, V, VEC>(a, rpt, col, val, W, tile * kCooTile, kCooTile);
  const int tail = (int)(m - nfull * kCooTile);
  if (tail > 0 && blockIdx.x == (unsigned)(nfull % gridDim.x))
    coo_tile<RowT, V, false>(a, rpt, col, val, W, nfull * kCooTile, tail);
}

static int ws_reserve(Handle *h, Workspace &w, size_t bytes, const char *what) {
  if (bytes <= w.bytes) return NSP_OK;
  if (w.base) NSP_HIP(h, hipFree(w.base));
  w.base = nullptr;
  w.bytes = 0;
  if (hipMalloc(&w.base, bytes + bytes / 4) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, std::string(what) + " workspace allocation failed");
  }
  w.bytes = bytes + bytes / 4;
  return NSP_OK;
}

static int grid_of(const Handle *h, int64_t n, int per = kCooBlock) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + per - 1) / per, (int64_t)h->num_cu * 8));
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// State of one conversion: the rows bitmap, the unsorted-row list, the
// control words and (for the file reader) the device staging buffers.
struct CooRun {
  uint32_t *bits = nullptr;
  int64_t *list = nullptr;
  CooCtl *ctl = nullptr;
  uint64_t *stage[2] = {nullptr, nullptr};
  char *text[2] = {nullptr, nullptr};  // ASCII reader: record lines of a chunk
};

static int coo_begin(Handle *h, int64_t nrows, int64_t stage_recs, CooRun &run, size_t text_bytes = 0) {
  const int64_t nwords = (nrows + 31) / 32;
  const size_t b_bits = al256((size_t)std::max<int64_t>(nwords, 1) * 4);
  const size_t b_list = al256((size_t)std::max<int64_t>(nrows, 1) * 8);
  const size_t b_ctl = 256;
  const size_t b_stage = al256((size_t)stage_recs * 24);
  const size_t b_text = text_bytes ? al256(text_bytes + 16) : 0;  // + the parser's word-rounded tail
  NSP_TRY(ws_reserve(h, h->coo, b_bits + b_list + b_ctl + 2 * b_stage + 2 * b_text, "COO"));
  char *p = (char *)h->coo.base;
  run.bits = (uint32_t *)p;
  run.list = (int64_t *)(p + b_bits);
  run.ctl = (CooCtl *)(p + b_bits + b_list);
  run.stage[0] = (uint64_t *)(p + b_bits + b_list + b_ctl);
  run.stage[1] = (uint64_t *)(p + b_bits + b_list + b_ctl + b_stage);
  if (text_bytes) {
    run.text[0] = p + b_bits + b_list + b_ctl + 2 * b_stage;
    run.text[1] = run.text[0] + b_text;
  }
  hipLaunchKernelGGL(k_coo_init, dim3(grid_of(h, nwords)), dim3(256), 0, h->stream, run.bits, nwords, run.ctl);
  NSP_HIP(h, hipGetLastError());
  return NSP_OK;
}

// Device timing of the conversion (events 6 and 7 of the handle).
static void coo_mark(Handle *h, int i) {
  if (!h->timing) return;
  if (!h->events[6 + i]) (void)hipEventCreate(&h->events[6 + i]);
  (void)hipEventRecord(h->events[6 + i], h->stream);
}

static void coo_collect(Handle *h) {
  if (!h->timing || !h->events[6] || !h->events[7]) return;
  float t = 0;
  if (hipEventSynchronize(h->events[7]) == hipSuccess && hipEventElapsedTime(&t, h->events[6], h->events[7]) == hipSuccess)
    h->stats.t_coo_ms = t;
}

template <typename RowT, typename V>
static int coo_launch_chunk(Handle *h, const CooRun &run, const uint64_t *words, int64_t m, int64_t q0,
                            int64_t prev_i, int64_t prev_j, int64_t row1, int64_t nrows, int64_t nrec,
                            int64_t col_shift, nsp_csr_t *out) {
  if (m <= 0) return NSP_OK;
  const int g = grid_of(h, m, kCooTile);
  const CooArgs a{words, q0, prev_i, prev_j, row1, nrows, nrec, col_shift, run.bits, run.list, run.ctl};
  if (((uintptr_t)words & 15) == 0)
    hipLaunchKernelGGL((k_coo_chunk<RowT, V, true>), dim3(g), dim3(kCooBlock), 0, h->stream, a, m,
                       (RowT *)out->row_ptr, out->col, (V *)out->val);
  else
    hipLaunchKernelGGL((k_coo_chunk<RowT, V, false>), dim3(g), dim3(kCooBlock), 0, h->stream, a, m,
                       (RowT *)out->row_ptr, out->col, (V *)out->val);
  NSP_HIP(h, hipGetLastError());
  return NSP_OK;
}

// Reads the control words back, reports errors and sorts the listed rows.
template <typename RowT, typename V>
static int coo_finish(Handle *h, const CooRun &run, nsp_csr_t *out) {
  hipStream_t st = h->stream;
  CooCtl ctl{};
  NSP_HIP(h, hipMemcpyAsync(&ctl, run.ctl, sizeof ctl, hipMemcpyDeviceToHost, st));
  NSP_HIP(h, hipStreamSynchronize(st));
  if (ctl.err & 1)
    return set_error(h, NSP_ERR_INVALID, "COO records out of row order or outside the rank's rows");
  if (ctl.err & 4)
    return set_error(h, NSP_ERR_INVALID, "ASCII COO: a record line does not end in a newline (lines of unequal length)");
  if (ctl.err & 2) return set_error(h, NSP_ERR_OVERFLOW, "COO column does not fit int32 after the shift");
  const int64_t n = (int64_t)ctl.nunsorted;
  if (n == 0) return NSP_OK;
  return sort_listed_rows<RowT, V>(h, run.list, n, (const RowT *)out->row_ptr, out->col, (V *)out->val);
}

static int coo_check_args(Handle *h, nsp_index_t it, int64_t nrec, int64_t nrows, const nsp_csr_t *out) {
  if (!out || nrec < 0 || nrows < 0 || !out->row_ptr || (nrec > 0 && (!out->col || !out->val)))
    return set_error(h, NSP_ERR_INVALID, "COO: bad arguments");
  if (it == NSP_IDX32 && nrec > INT32_MAX) return set_error(h, NSP_ERR_OVERFLOW, "COO: nnz does not fit int32 row pointers");
  if (nrec > 0 && nrows == 0) return set_error(h, NSP_ERR_INVALID, "COO: records but no rows");
  return NSP_OK;
}

template <typename RowT, typename V>
static int coo_to_csr_impl(Handle *h, const void *records, int64_t nrec, int64_t first_row, int64_t col_shift,
                           nsp_csr_t *out) {
  const int64_t nrows = out->rows;
  if (nrec == 0) {
    hipLaunchKernelGGL((k_coo_fill_rpt<RowT>), dim3(grid_of(h, nrows + 1)), dim3(256), 0, h->stream,
                       (RowT *)out->row_ptr, nrows + 1);
    NSP_HIP(h, hipGetLastError());
    NSP_HIP(h, hipStreamSynchronize(h->stream));
    return NSP_OK;
  }
  CooRun run;
  NSP_TRY(coo_begin(h, nrows, 0, run));
  coo_mark(h, 0);
  NSP_TRY((coo_launch_chunk<RowT, V>(h, run, (const uint64_t *)records, nrec, 0, 0, 0, first_row + 1, nrows, nrec,
                                     col_shift, out)));
  coo_mark(h, 1);
  NSP_TRY((coo_finish<RowT, V>(h, run, out)));
  coo_collect(h);
  return NSP_OK;
}

int coo_to_csr(Handle *h, nsp_index_t it, nsp_value_t vt, const void *records, int64_t nrec, int64_t first_row,
               int64_t col_shift, nsp_csr_t *out) {
  NSP_TRY(coo_check_args(h, it, nrec, out ? out->rows : -1, out));
  if (nrec > 0 && !records) return set_error(h, NSP_ERR_INVALID, "COO: NULL records");
  out->nnz = nrec;
  if (it == NSP_IDX32 && vt == NSP_F64) return coo_to_csr_impl<int32_t, double>(h, records, nrec, first_row, col_shift, out);
  if (it == NSP_IDX32) return coo_to_csr_impl<int32_t, float>(h, records, nrec, first_row, col_shift, out);
  if (vt == NSP_F64) return coo_to_csr_impl<int64_t, double>(h, records, nrec, first_row, col_shift, out);
  return coo_to_csr_impl<int64_t, float>(h, records, nrec, first_row, col_shift, out);
}

// ---------------------------------------------------------------------------
// host side: header, rank ranges, streamed reads

struct Rec {
  int64_t i, j;
  double v;
};

static bool read_at(int fd, void *dst, size_t bytes, int64_t off) {
  char *p = (char *)dst;
  while (bytes > 0) {
    const ssize_t r = pread(fd, p, bytes, (off_t)off);
    if (r <= 0) return false;
    p += r;
    bytes -= (size_t)r;
    off += r;
  }
  return true;
}

// first record whose 1-based row id exceeds x (binsearch_iat,
// src/READ_MATRIX_MPI.cpp:107-130: records with row <= x come before it)
static int64_t upper_row(int fd, int64_t hdr, int64_t nt, int64_t x, bool &ok) {
  int64_t lo = 0, hi = nt;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    int64_t row;
    if (!read_at(fd, &row, 8, hdr + mid * 24)) {
      ok = false;
      return 0;
    }
    if (x < row) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

static int share_of(int fd, int64_t hdr, int64_t nr, int64_t nt, int nprocs, int rank, int64_t &b, int64_t &e,
                    int64_t &first_row, int64_t &nrows) {
  const int64_t per = nr / nprocs;
  const int64_t lo = per * rank, hi = (rank == nprocs - 1) ? nr : per * (rank + 1);
  bool ok = true;
  b = upper_row(fd, hdr, nt, lo, ok);
  e = upper_row(fd, hdr, nt, hi, ok);
  if (!ok) return NSP_ERR_INVALID;
  first_row = lo;
  nrows = 0;
  if (e > b) {
    int64_t r0, r1;
    if (!read_at(fd, &r0, 8, hdr + b * 24) || !read_at(fd, &r1, 8, hdr + (e - 1) * 24)) return NSP_ERR_INVALID;
    first_row = r0 - 1;
    nrows = r1 - r0 + 1;
  }
  return NSP_OK;
}

int coo_share(const char *path, int header_type, int nprocs, int rank, nsp_coo_share_t *out) {
  if (!path || !out || nprocs < 1 || rank < 0 || rank >= nprocs) return NSP_ERR_INVALID;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return NSP_ERR_INVALID;
  int64_t hd[3] = {0, 0, 0};
  const int nh = header_type ? 3 : 2;
  struct stat sb;
  int rc = NSP_OK;
  if (!read_at(fd, hd, 8 * nh, 0) || fstat(fd, &sb) != 0) rc = NSP_ERR_INVALID;
  nsp_coo_share_t s{};
  if (rc == NSP_OK) {
    s.nr = hd[0];
    s.nc = header_type ? hd[1] : hd[0];
    s.nt = header_type ? hd[2] : hd[1];
    s.header_bytes = 8 * nh;
    if (s.nr < 0 || s.nt < 0 || s.header_bytes + s.nt * 24 > (int64_t)sb.st_size) rc = NSP_ERR_INVALID;
  }
  if (rc == NSP_OK)
    rc = share_of(fd, s.header_bytes, s.nr, s.nt, nprocs, rank, s.rec_begin, s.rec_end, s.first_row, s.nrows);
  // col_shift = the rows spanned by the ranks before (MPI_Exscan of rowxproc)
  for (int r = 0; r < rank && rc == NSP_OK; r++) {
    int64_t b, e, f, n;
    rc = share_of(fd, s.header_bytes, s.nr, s.nt, nprocs, r, b, e, f, n);
    s.col_shift += n;
  }
  close(fd);
  if (rc == NSP_OK) *out = s;
  return rc;
}

static int pin_reserve(Handle *h, size_t bytes) {
  if (bytes <= h->pin_bytes) return NSP_OK;
  if (h->pin) NSP_HIP(h, hipHostFree(h->pin));
  h->pin = nullptr;
  h->pin_bytes = 0;
  if (hipHostMalloc(&h->pin, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(h, NSP_ERR_OOM, "pinned staging allocation failed");
  }
  h->pin_bytes = bytes;
  return NSP_OK;
}

// Parallel pread of [off, off+bytes) into dst.
static bool read_parallel(int fd, char *dst, size_t bytes, int64_t off, int nthreads) {
  if (nthreads <= 1 || bytes < (size_t(1) << 22)) return read_at(fd, dst, bytes, off);
  std::vector<std::thread> th;
  std::vector<char> ok((size_t)nthreads, 1);
  const size_t part = (bytes / nthreads + 4095) & ~(size_t)4095;
  for (int t = 0; t < nthreads; t++) {
    const size_t s = std::min(bytes, part * t), e = std::min(bytes, part * (t + 1));
    if (s >= e) break;
    th.emplace_back([&, t, s, e] { ok[t] = read_at(fd, dst + s, e - s, off + (int64_t)s); });
  }
  for (auto &x : th) x.join();
  for (char c : ok)
    if (!c) return false;
  return true;
}

template <typename RowT, typename V>
static int coo_read_impl(Handle *h, int fd, const nsp_coo_share_t *s, nsp_csr_t *out) {
  const int64_t nrec = s->rec_end - s->rec_begin, nrows = s->nrows;
  hipStream_t st = h->stream;
  if (nrec == 0) {
    hipLaunchKernelGGL((k_coo_fill_rpt<RowT>), dim3(grid_of(h, nrows + 1)), dim3(256), 0, st, (RowT *)out->row_ptr,
                       nrows + 1);
    NSP_HIP(h, hipGetLastError());
    NSP_HIP(h, hipStreamSynchronize(st));
    return NSP_OK;
  }
  // records per streamed chunk (default 2^21 records = 48 MiB per staging buffer)
  const int64_t chunk = std::min(h->coo_chunk, nrec);
  CooRun run;
  NSP_TRY(coo_begin(h, nrows, chunk, run));
  NSP_TRY(pin_reserve(h, (size_t)chunk * 24 * 2));
  for (int k = 0; k < 2; k++)
    if (!h->pin_ev[k]) NSP_HIP(h, hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming));
  const int nthreads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  int64_t prev_i = 0, prev_j = 0;
  int k = 0;
  for (int64_t q0 = 0; q0 < nrec; q0 += chunk, k++) {
    const int64_t m = std::min(chunk, nrec - q0);
    char *hbuf = (char *)h->pin + (size_t)(k & 1) * chunk * 24;
    // the copy out of this pinned buffer two chunks ago must be done
    if (k >= 2) NSP_HIP(h, hipEventSynchronize(h->pin_ev[k & 1]));
    if (!read_parallel(fd, hbuf, (size_t)m * 24, s->header_bytes + (s->rec_begin + q0) * 24, nthreads))
      return set_error(h, NSP_ERR_INVALID, "COO: short read");
    NSP_HIP(h, hipMemcpyAsync(run.stage[k & 1], hbuf, (size_t)m * 24, hipMemcpyHostToDevice, st));
    NSP_HIP(h, hipEventRecord(h->pin_ev[k & 1], st));
    if (k == 0) coo_mark(h, 0);
    NSP_TRY((coo_launch_chunk<RowT, V>(h, run, run.stage[k & 1], m, q0, prev_i, prev_j, s->first_row + 1, nrows,
                                       nrec, s->col_shift, out)));
    const Rec *last = (const Rec *)(hbuf + (size_t)(m - 1) * 24);
    prev_i = last->i;
    prev_j = last->j;
  }
  coo_mark(h, 1);
  NSP_TRY((coo_finish<RowT, V>(h, run, out)));
  coo_collect(h);
  return NSP_OK;
}

int coo_read(Handle *h, nsp_index_t it, nsp_value_t vt, const char *path, int header_type,
             const nsp_coo_share_t *s, nsp_csr_t *out) {
  if (!path || !s) return set_error(h, NSP_ERR_INVALID, "COO: bad arguments");
  const int64_t nrec = s->rec_end - s->rec_begin;
  if (out) out->rows = s->nrows;
  NSP_TRY(coo_check_args(h, it, nrec, s->nrows, out));
  if (s->header_bytes != (header_type ? 24 : 16) || s->rec_begin < 0 || nrec < 0)
    return set_error(h, NSP_ERR_INVALID, "COO: share does not match the header type");
  out->cols = s->nc;
  out->nnz = nrec;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return set_error(h, NSP_ERR_INVALID, std::string("COO: cannot open ") + path);
  int rc;
  if (it == NSP_IDX32 && vt == NSP_F64) rc = coo_read_impl<int32_t, double>(h, fd, s, out);
  else if (it == NSP_IDX32) rc = coo_read_impl<int32_t, float>(h, fd, s, out);
  else if (vt == NSP_F64) rc = coo_read_impl<int64_t, double>(h, fd, s, out);
  else rc = coo_read_impl<int64_t, float>(h, fd, s, out);
  if (rc != NSP_OK) (void)hipStreamSynchronize(h->stream);  // no copy may still read the pinned buffers
  close(fd);
  return rc;
}

// ---------------------------------------------------------------------------
// ASCII COO files: fixed-length record lines "i j v" parsed on the device,
// one lane per line, into the 24-byte records the chunk conversion above
// takes.  The reference reads them with atoi/atol and its own fast_atof
// (src/READ_MATRIX_MPI.cpp:21-77, 464-477); the parsers below follow the
// same arithmetic, so the values are bit-identical.

// atol over [p, end): leading white space, an optional sign, decimal digits.
__host__ __device__ inline int64_t text_atol(const char *p, const char *end) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\v' || *p == '\f' || *p == '\r')) p++;
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) neg = (*p++ == '-');
  uint64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') v = v * 10u + (uint64_t)(*p++ - '0');
  return neg ? -(int64_t)v : (int64_t)v;
}

// The reference's decimal parser, on [p, end): blanks (space, tab), an
// optional sign, integer digits accumulated as value*10 + d, fraction
// digits added one by one as d / 10^k (10^k built by repeated *10), then
// an optional exponent (clamped to 308) applied as a product of 1e50, 1e8
// and 10 factors, divided by for a negative exponent.  Every operation is a
// separately rounded FP64 operation (no fused multiply-add), as in the
// reference's host build.
__host__ __device__ inline double text_fast_atof(const char *p, const char *end) {
#pragma clang fp contract(off)
  while (p < end && (*p == ' ' || *p == '\t')) p++;
  double sign = 1.0;
  if (p < end && *p == '-') {
    sign = -1.0;
    p++;
  } else if (p < end && *p == '+') {
    p++;
  }
  double value = 0.0;
  for (; p < end && *p >= '0' && *p <= '9'; p++) value = value * 10.0 + (double)(*p - '0');
  if (p < end && *p == '.') {
    double pow10 = 10.0;
    for (p++; p < end && *p >= '0' && *p <= '9'; p++) {
      value += (double)(*p - '0') / pow10;
      pow10 *= 10.0;
    }
  }
  bool neg_exp = false;
  double scale = 1.0;
  if (p < end && (*p == 'e' || *p == 'E')) {
    p++;
    if (p < end && *p == '-') {
      neg_exp = true;
      p++;
    } else if (p < end && *p == '+') {
      p++;
    }
    uint32_t e = 0;
    for (; p < end && *p >= '0' && *p <= '9'; p++) e = e * 10u + (uint32_t)(*p - '0');
    if (e > 308) e = 308;
    for (; e >= 50; e -= 50) scale *= 1e50;
    for (; e >= 8; e -= 8) scale *= 1e8;
    for (; e > 0; e--) scale *= 10.0;
  }
  return sign * (neg_exp ? value / scale : value * scale);
}

// Lines per LDS tile of the parser: at most 16 KB of text, at most one line
// per thread.
constexpr int kAsciiTileBytes = 16384;

// m record lines of L bytes each at `text` (the buffer holds whole 4-byte
// words past the last line) -> 3 words per record {i, j, v}, 1-based as in
// the file.  A line whose last byte is not a newline sets error bit 2.
__global__ __launch_bounds__(kCooBlock) void k_ascii_parse(const char *__restrict__ text, int64_t m, int L,
                                                           int off_col, int off_val, uint64_t *__restrict__ words,
                                                           CooCtl *ctl) {
  __shared__ uint32_t tile[kAsciiTileBytes / 4 + 2];
  const int T = kAsciiTileBytes / L < kCooBlock ? kAsciiTileBytes / L : kCooBlock;  // lines per tile (>= 16)
  const int64_t ntiles = (m + T - 1) / T;
  const uint32_t *tw = reinterpret_cast<const uint32_t *>(text);
  for (int64_t tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
    const int64_t l0 = tl * T;
    const int nl = (int)(m - l0 < T ? m - l0 : T);
    const int64_t b0 = l0 * L, b1 = b0 + (int64_t)nl * L;
    const int64_t w0 = b0 >> 2, w1 = (b1 + 3) >> 2;  // words [w0, w1) cover the tile's bytes
    __syncthreads();  // the previous tile's lines are parsed
    for (int64_t w = w0 + threadIdx.x; w < w1; w += kCooBlock) tile[w - w0] = tw[w];
    __syncthreads();
    if (threadIdx.x < nl) {
      const char *line = reinterpret_cast<const char *>(tile) + (b0 - 4 * w0) + (int64_t)threadIdx.x * L;
      const char *end = line + L;
      const int64_t i = text_atol(line, end);
      const int64_t j = text_atol(line + off_col, end);
      const double v = text_fast_atof(line + off_val, end);
      if (line[L - 1] != '\n') atomicOr(&ctl->err, 4u);
      const int64_t q = l0 + threadIdx.x;
      words[3 * q] = (uint64_t)i;
      words[3 * q + 1] = (uint64_t)j;
      words[3 * q + 2] = (uint64_t)__double_as_longlong(v);
    }
  }
}

// Header and first record line: body offset, line length and the two field
// offsets, found as the reference finds them (src/READ_MATRIX_MPI.cpp:
// 381-429), except that the body starts after the '%' lines too.
static int ascii_layout(FILE *f, nsp_ascii_share_t &s) {
  char *line = nullptr;
  size_t cap = 0;
  int64_t off = 0;
  ssize_t n;
  int rc = NSP_ERR_INVALID;
  while ((n = getline(&line, &cap, f)) > 0 && line[0] == '%') off += n;
  if (n > 0) {
    long a = 0, b = 0, c = 0;
    const int k = sscanf(line, "%ld %ld %ld", &a, &b, &c);
    bool ok = true;
    if (k == 3) {
      s.nr = a;
      s.nc = b;
      s.nt = c;
    } else if (sscanf(line, "%ld %ld", &a, &b) == 2) {
      s.nr = s.nc = a;
      s.nt = b;
    } else {
      ok = false;
    }
    off += n;
    s.body_offset = off;
    if (ok && s.nr >= 0 && s.nt >= 0) {
      if (s.nt == 0) {
        s.line_bytes = 1;
        s.col_offset = s.val_offset = 0;
        rc = NSP_OK;
      } else if ((n = getline(&line, &cap, f)) > 1) {
        // first space after the row field, then after the column field
        ssize_t i = 0;
        while (i < n && line[i] == ' ') i++;
        while (i < n && line[i] != ' ' && line[i] != '\n') i++;
        s.col_offset = (int32_t)i;
        while (i < n && line[i] == ' ') i++;
        while (i < n && line[i] != ' ' && line[i] != '\n') i++;
        s.val_offset = (int32_t)i;
        s.line_bytes = line[n - 1] == '\n' ? n : n + 1;
        rc = NSP_OK;
      }
    }
  }
  free(line);
  return rc;
}

// Row field of record q (the reference's binsearch_iat_str reads it with atol).
static bool ascii_row(int fd, const nsp_ascii_share_t &s, int64_t q, int64_t &row) {
  char buf[64];
  const size_t want = (size_t)std::min<int64_t>(s.line_bytes, (int64_t)sizeof buf);
  const ssize_t got = pread(fd, buf, want, (off_t)(s.body_offset + q * s.line_bytes));
  if (got <= 0) return false;
  row = text_atol(buf, buf + got);
  return true;
}

// First record whose row exceeds x (src/READ_MATRIX_MPI.cpp:83-104).
static int64_t ascii_upper_row(int fd, const nsp_ascii_share_t &s, int64_t x, bool &ok) {
  int64_t lo = 0, hi = s.nt;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    int64_t row;
    if (!ascii_row(fd, s, mid, row)) {
      ok = false;
      return 0;
    }
    if (x < row) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

static int ascii_share_of(int fd, const nsp_ascii_share_t &s, int nprocs, int rank, int64_t &b, int64_t &e,
                          int64_t &first_row, int64_t &nrows) {
  const int64_t per = s.nr / nprocs;
  const int64_t lo = per * rank, hi = (rank == nprocs - 1) ? s.nr : per * (rank + 1);
  bool ok = true;
  b = ascii_upper_row(fd, s, lo, ok);
  e = ascii_upper_row(fd, s, hi, ok);
  if (!ok) return NSP_ERR_INVALID;
  first_row = lo;
  nrows = 0;
  if (e > b) {
    int64_t r0, r1;
    if (!ascii_row(fd, s, b, r0) || !ascii_row(fd, s, e - 1, r1)) return NSP_ERR_INVALID;
    first_row = r0 - 1;
    nrows = r1 - r0 + 1;
  }
  return NSP_OK;
}

int ascii_share(const char *path, int nprocs, int rank, nsp_ascii_share_t *out) {
  if (!path || !out || nprocs < 1 || rank < 0 || rank >= nprocs) return NSP_ERR_INVALID;
  FILE *f = fopen(path, "rb");
  if (!f) return NSP_ERR_INVALID;
  nsp_ascii_share_t s{};
  int rc = ascii_layout(f, s);
  struct stat sb;
  // nt lines of line_bytes each; the last one may end at the end of the file
  if (rc == NSP_OK && (fstat(fileno(f), &sb) != 0 ||
                       s.body_offset + s.nt * s.line_bytes - (s.nt > 0 ? 1 : 0) > (int64_t)sb.st_size))
    rc = NSP_ERR_INVALID;
  const int fd = fileno(f);
  if (rc == NSP_OK) rc = ascii_share_of(fd, s, nprocs, rank, s.rec_begin, s.rec_end, s.first_row, s.nrows);
  for (int r = 0; r < rank && rc == NSP_OK; r++) {  // col_shift: MPI_Exscan of rowxproc
    int64_t b, e, fr, n;
    rc = ascii_share_of(fd, s, nprocs, r, b, e, fr, n);
    s.col_shift += n;
  }
  fclose(f);
  if (rc == NSP_OK) *out = s;
  return rc;
}

template <typename RowT, typename V>
static int ascii_read_impl(Handle *h, int fd, int64_t file_bytes, const nsp_ascii_share_t *s, nsp_csr_t *out) {
  const int64_t nrec = s->rec_end - s->rec_begin, nrows = s->nrows, L = s->line_bytes;
  hipStream_t st = h->stream;
  if (nrec == 0) {
    hipLaunchKernelGGL((k_coo_fill_rpt<RowT>), dim3(grid_of(h, nrows + 1)), dim3(256), 0, st, (RowT *)out->row_ptr,
                       nrows + 1);
    NSP_HIP(h, hipGetLastError());
    NSP_HIP(h, hipStreamSynchronize(st));
    return NSP_OK;
  }
  // lines per streamed chunk: the records' chunk, held to 256 MiB of text
  const int64_t chunk = std::max<int64_t>(1, std::min({h->coo_chunk, nrec, (int64_t(1) << 28) / L}));
  CooRun run;
  NSP_TRY(coo_begin(h, nrows, chunk, run, (size_t)(chunk * L)));
  NSP_TRY(pin_reserve(h, (size_t)(chunk * L) * 2));
  for (int k = 0; k < 2; k++)
    if (!h->pin_ev[k]) NSP_HIP(h, hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming));
  const int nthreads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  const int T = std::min<int>(kAsciiTileBytes / (int)L, kCooBlock);
  int64_t prev_i = 0, prev_j = 0;
  int k = 0;
  for (int64_t q0 = 0; q0 < nrec; q0 += chunk, k++) {
    const int64_t m = std::min(chunk, nrec - q0);
    char *hbuf = (char *)h->pin + (size_t)(k & 1) * (size_t)(chunk * L);
    if (k >= 2) NSP_HIP(h, hipEventSynchronize(h->pin_ev[k & 1]));
    const int64_t off = s->body_offset + (s->rec_begin + q0) * L;
    // the file's last line may lack its newline: read what is there and
    // supply it
    const int64_t avail = std::min<int64_t>(m * L, file_bytes - off);
    if (avail < m * L - 1 || !read_parallel(fd, hbuf, (size_t)avail, off, nthreads))
      return set_error(h, NSP_ERR_INVALID, "ASCII COO: short read");
    if (avail < m * L) hbuf[m * L - 1] = '\n';
    NSP_HIP(h, hipMemcpyAsync(run.text[k & 1], hbuf, (size_t)(m * L), hipMemcpyHostToDevice, st));
    NSP_HIP(h, hipEventRecord(h->pin_ev[k & 1], st));
    if (k == 0) coo_mark(h, 0);
    const int g = grid_of(h, m, T);
    hipLaunchKernelGGL(k_ascii_parse, dim3(g), dim3(kCooBlock), 0, st, run.text[k & 1], m, (int)L, s->col_offset,
                       s->val_offset, run.stage[0], run.ctl);
    NSP_HIP(h, hipGetLastError());
    NSP_TRY((coo_launch_chunk<RowT, V>(h, run, run.stage[0], m, q0, prev_i, prev_j, s->first_row + 1, nrows, nrec,
                                       s->col_shift, out)));
    const char *last = hbuf + (m - 1) * L;
    prev_i = text_atol(last, last + L);
    prev_j = text_atol(last + s->col_offset, last + L);
  }
  coo_mark(h, 1);
  NSP_TRY((coo_finish<RowT, V>(h, run, out)));
  coo_collect(h);
  return NSP_OK;
}

int ascii_read(Handle *h, nsp_index_t it, nsp_value_t vt, const char *path, const nsp_ascii_share_t *s,
               nsp_csr_t *out) {
  if (!path || !s) return set_error(h, NSP_ERR_INVALID, "ASCII COO: bad arguments");
  const int64_t nrec = s->rec_end - s->rec_begin;
  if (out) out->rows = s->nrows;
  NSP_TRY(coo_check_args(h, it, nrec, s->nrows, out));
  if (s->rec_begin < 0 || nrec < 0 || s->line_bytes < 2 || s->line_bytes > kAsciiTileBytes / 16 ||
      s->col_offset < 0 || s->val_offset < s->col_offset || s->val_offset >= s->line_bytes)
    return set_error(h, NSP_ERR_INVALID, "ASCII COO: bad share (record lines of 2 to 1024 bytes)");
  out->cols = s->nc;
  out->nnz = nrec;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return set_error(h, NSP_ERR_INVALID, std::string("ASCII COO: cannot open ") + path);
  struct stat sb;
  if (fstat(fd, &sb) != 0) {
    close(fd);
    return set_error(h, NSP_ERR_INVALID, "ASCII COO: cannot stat the file");
  }
  int rc;
  if (it == NSP_IDX32 && vt == NSP_F64) rc = ascii_read_impl<int32_t, double>(h, fd, (int64_t)sb.st_size, s, out);
  else if (it == NSP_IDX32) rc = ascii_read_impl<int32_t, float>(h, fd, (int64_t)sb.st_size, s, out);
  else if (vt == NSP_F64) rc = ascii_read_impl<int64_t, double>(h, fd, (int64_t)sb.st_size, s, out);
  else rc = ascii_read_impl<int64_t, float>(h, fd, (int64_t)sb.st_size, s, out);
  if (rc != NSP_OK) (void)hipStreamSynchronize(h->stream);  // no copy may still read the pinned buffers
  close(fd);
  return rc;
}

}  // namespace nsp

extern "C" {

int nsp_coo_share(const char *path, int header_type, int nprocs, int rank, nsp_coo_share_t *out) {
  return nsp::coo_share(path, header_type, nprocs, rank, out);
}

int nsp_coo_to_csr(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const void *records, int64_t nrec,
                   int64_t first_row, int64_t col_shift, nsp_csr_t *out) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::coo_to_csr(nsp::handle_of(hs), it, vt, records, nrec, first_row, col_shift, out);
}

int nsp_coo_read(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const char *path, int header_type,
                 const nsp_coo_share_t *share, nsp_csr_t *out) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::coo_read(nsp::handle_of(hs), it, vt, path, header_type, share, out);
}

int nsp_ascii_share(const char *path, int nprocs, int rank, nsp_ascii_share_t *out) {
  return nsp::ascii_share(path, nprocs, rank, out);
}

int nsp_ascii_read(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const char *path, const nsp_ascii_share_t *share,
                   nsp_csr_t *out) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::ascii_read(nsp::handle_of(hs), it, vt, path, share, out);
}

}  // extern "C"
