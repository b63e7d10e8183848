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
  uint32_t err;          // bit 0: records out of row order or outside the rows; bit 1: column overflow
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
};

static int coo_begin(Handle *h, int64_t nrows, int64_t stage_recs, CooRun &run) {
  const int64_t nwords = (nrows + 31) / 32;
  const size_t b_bits = al256((size_t)std::max<int64_t>(nwords, 1) * 4);
  const size_t b_list = al256((size_t)std::max<int64_t>(nrows, 1) * 8);
  const size_t b_ctl = 256;
  const size_t b_stage = al256((size_t)stage_recs * 24);
  NSP_TRY(ws_reserve(h, h->coo, b_bits + b_list + b_ctl + 2 * b_stage, "COO"));
  char *p = (char *)h->coo.base;
  run.bits = (uint32_t *)p;
  run.list = (int64_t *)(p + b_bits);
  run.ctl = (CooCtl *)(p + b_bits + b_list);
  run.stage[0] = (uint64_t *)(p + b_bits + b_list + b_ctl);
  run.stage[1] = (uint64_t *)(p + b_bits + b_list + b_ctl + b_stage);
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

}  // extern "C"
