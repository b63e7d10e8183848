/*
 * nsp_amd.h -- C ABI of the MI355X (gfx950) local sparse product C = A*B.
 *
 * Every entry point takes plain pointers and sizes; device pointers are HIP
 * device memory on the handle's device, and work is ordered on the handle's
 * stream.  Each declaration names the reference interface it replaces
 * (paths relative to the reference tree).
 *
 * Conventions
 *   - CSR, 0-based, rows of B sorted ascending by column and duplicate-free
 *     (the reference's precondition: nsp/src/nsp.cu:654,1805,2105 and the
 *     loader's check_and_fix_order, src/distribute.cu:1050-1069).
 *   - row pointers int32 (NSP_IDX32, the reference's `int *rpt`,
 *     nsp/inc/nsp.h:57-70) or int64 (NSP_IDX64, for nnz >= 2^31); one width
 *     for A, B and C in a call.  Columns are always int32.
 *   - values FP64 (the reference's `real`) or FP32; FP32 products are formed
 *     in FP32 and summed in FP64 before the result is rounded to FP32.
 *   - C rows come out sorted ascending, structural zeros kept, as the
 *     reference produces them (nsp/src/nsp.cu:24-90, 1389-1399).
 *   - Functions return an nsp_status_t; they never exit the process (the
 *     reference exits through checkCudaErrors, nsp/inc/helper_cuda.h).
 */
#ifndef NSP_AMD_H
#define NSP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  NSP_OK = 0,
  NSP_ERR_INVALID = 1,      /* bad argument or malformed CSR */
  NSP_ERR_OOM = 2,          /* device allocation failed */
  NSP_ERR_OVERFLOW = 3,     /* result does not fit the chosen index width */
  NSP_ERR_HIP = 4,          /* HIP runtime error */
  NSP_ERR_STATE = 5,        /* call sequence error (numeric without symbolic) */
  NSP_ERR_UNSUPPORTED = 6   /* a size this build does not handle */
} nsp_status_t;

typedef enum { NSP_IDX32 = 0, NSP_IDX64 = 1 } nsp_index_t;
typedef enum { NSP_F64 = 0, NSP_F32 = 1 } nsp_value_t;

/* A CSR matrix.  row_ptr is int32_t* or int64_t* per nsp_index_t; val is
 * double* or float* per nsp_value_t.  Mirrors sfCSR's device fields
 * (d_rpt, d_col, d_val, M, N, nnz; nsp/inc/nsp.h:57-70). */
typedef struct {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  void *row_ptr;
  int32_t *col;
  void *val;
} nsp_csr_t;

typedef struct nsp_handle_s *nsp_handle_t;

/* Statistics of the last product on a handle. */
typedef struct {
  int64_t nprod;            /* sum_i sum_{k in A_i} nnz(B_k) */
  int64_t nnz;              /* nnz(C) */
  int64_t max_intprod;      /* largest intermediate-product count of a row
                             * (-1 after a one-call product that reused the
                             * previous call's slot size: it sums no rows) */
  int64_t max_row_nnz;      /* largest nnz of a C row */
  int64_t sym_bin[16];      /* rows per symbolic bin */
  int64_t num_bin[16];      /* rows per numeric bin */
  int32_t n_sym_bins, n_num_bins;
  int32_t direct_mode;      /* bit 0: symbolic pass walked all rows in natural
                             * order with one kernel, bit 1: numeric pass did */
  int32_t host_syncs;       /* blocking host synchronisations in the call */
  int32_t sym_launches;     /* symbolic hash kernels launched */
  int32_t num_launches;     /* numeric hash kernels launched */
  /* device time of each phase (HIP events on the handle's stream), filled
   * when the "timing" parameter is 1; milliseconds */
  float t_bin_ms, t_sym_ms, t_scan_ms, t_num_ms;
  /* device time of the record conversion of the last nsp_coo_to_csr (the
   * conversion kernel alone) or nsp_coo_read (first chunk's conversion to
   * the last one's end, copies included), when "timing" is 1 */
  float t_coo_ms;
  /* 1 when the last one-call product (nsp_spgemm, nsp_spgemm_into) ran its
   * numeric pass into per-row slots (no symbolic pass); t_num_ms is then
   * that pass, t_sym_ms 0, and t_compact_ms the copy of the slots into C */
  int32_t fused;
  float t_compact_ms;
} nsp_stats_t;

/* ---- handle ------------------------------------------------------------ */
/* Replaces the per-call state of nsp_init_bin/nsp_release_bin
 * (nsp/src/nsp.cu:248-298: 11 streams and 5 allocations per call): a handle
 * owns a reusable device workspace and one stream, bound to the device that
 * is current when it is created. */
int nsp_create(nsp_handle_t *h);
int nsp_destroy(nsp_handle_t h);
int nsp_set_stream(nsp_handle_t h, void *hip_stream);
const char *nsp_status_string(int status);
const char *nsp_last_error(nsp_handle_t h);
int nsp_get_stats(nsp_handle_t h, nsp_stats_t *out);
/* Tuning knobs (nsp_amd.autotune picks them per workload):
 *   "grid_cap"     blocks per CU of the grid-stride launches (1..64, default 32)
 *   "sym_small_rows_per_wave"  rows per wavefront in the symbolic bins of
 *                  up to 128 products (1, 2, 4, 8)
 *   "sym_rows_per_wave"  rows per wavefront in the symbolic bins of
 *                  129..1024 products (1, 2, 4)
 *   "num_small_rows_per_wave"  rows per wavefront in the numeric bins of up
 *                  to 64 entries (1, 2, 4, 8)
 *   "num_rows_per_wave"  rows per wavefront in the numeric bins of 65..256
 *                  entries (1, 2, 4)
 *   "num_table_factor"   numeric hash slots per row entry (2, 4)
 *   "narrow_mode"  1 (default): when B has at most 4096 columns, rows of up
 *                  to 1024 products / 512 entries keep their columns in an
 *                  LDS bitmap (no hash table, no sort); 0: hash tables
 *   "fused_mode"   1 (default): one-call products (nsp_spgemm, nsp_spgemm_into)
 *                  skip the symbolic pass when a sample of 1024 rows fits a
 *                  wave-level numeric bin (at most 512 columns): the numeric
 *                  pass writes each row into a slot of that bin's size and
 *                  counts it, a scan gives C's row pointer, and a copy packs
 *                  the slots into C (a row that outgrows its slot sends the
 *                  product back to two passes); 0: always two passes
 *   "queue_mode"   1 (default): with the column bitmaps, the numeric rows of
 *                  up to 512 entries run as one persistent launch whose
 *                  blocks draw tiles of rows from row-bin work queues (one
 *                  per XCD, longest rows first, idle blocks take from the
 *                  other XCDs' queues); 0: one launch per bin
 *   "queue_chunk"  tiles of rows per work-queue ticket (1..64, default 16)
 *   "queue_blocks" blocks per CU of the work-queue launch (1..16, default 4)
 *   "sym_table_factor"   symbolic hash slots per intermediate product in the
 *                  bins of up to 256 products (1, 2; default 2: rows whose
 *                  products are mostly distinct keep the table half empty)
 *   "heavy_blocks" blocks in flight for rows past 16384 products / 4096
 *                  entries (0 = one per CU)
 *   "reuse_extents" 1: nsp_spgemm_backward reuses the per-entry B row
 *                  extents, product offsets and longest C row of the
 *                  previous call on this handle when A's column, A's and B's
 *                  row-pointer and C's buffers are the same (the caller
 *                  promises the patterns did not change); such calls do not
 *                  synchronise the host
 *   "bwd_contrib"  how nsp_spgemm_backward forms dB when A^T's permutation
 *                  is given: 1 from per-product contributions written by the
 *                  dA pass (default), 0 by binary search in C's rows, 2
 *                  the contributions only when C has rows past 32 entries
 *   "timing"       1 records HIP events around each phase (see nsp_stats_t)
 *   "coo_chunk_records"  records per streamed chunk of nsp_coo_read
 *                  (1..2^26, default 2^21: two 48 MiB pinned buffers) */
int nsp_set_param(nsp_handle_t h, const char *name, int64_t value);

/* ---- local product ----------------------------------------------------- */
/* Symbolic phase of C = A*B: row binning by intermediate products, the hash
 * count of distinct columns per C row, and the exclusive scan into
 * c_row_ptr (rows+1 entries, caller-allocated device memory).  *c_nnz gets
 * nnz(C).  Replaces nsp_set_max_bin + nsp_set_row_nnz
 * (nsp/src/nsp.cu:417-471, 921-1130). */
int nsp_spgemm_symbolic(nsp_handle_t h, nsp_index_t it, nsp_value_t vt,
                        const nsp_csr_t *A, const nsp_csr_t *B,
                        void *c_row_ptr, int64_t *c_nnz);

/* Numeric phase: fills C->col and C->val (caller-allocated, C->nnz entries)
 * for the A, B and C->row_ptr of the preceding nsp_spgemm_symbolic on this
 * handle.  Replaces nsp_set_min_bin + nsp_calculate_value_col_bin
 * (nsp/src/nsp.cu:1241-1300, 2178-2440). */
int nsp_spgemm_numeric(nsp_handle_t h, nsp_index_t it, nsp_value_t vt,
                       const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C);

/* Allocation callbacks for the one-shot product; NULL selects hipMalloc /
 * hipFree. */
typedef void *(*nsp_alloc_fn)(size_t bytes, void *user);
typedef void (*nsp_free_fn)(void *ptr, void *user);

/* One-shot product with callee-allocated C (C->row_ptr, C->col, C->val set
 * on return; host-synchronous like the reference).  Replaces
 * void nsp_spgemm_kernel_hash(sfCSR *a, sfCSR *b, sfCSR *c)
 * (nsp/inc/nsp.h:105, nsp/src/nsp.cu:2448-2488). */
int nsp_spgemm(nsp_handle_t h, nsp_index_t it, nsp_value_t vt,
               const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C,
               nsp_alloc_fn alloc, nsp_free_fn dealloc, void *user);

/* One-call product into caller buffers: C->row_ptr (A->rows+1 entries) and
 * C->col / C->val with a capacity of C->nnz entries; *c_nnz gets nnz(C).
 * Same result as nsp_spgemm (and as nsp_spgemm_symbolic + _numeric), in
 * one pass over the products when the "fused_mode" sample allows.  C larger
 * than the capacity returns NSP_ERR_OVERFLOW with *c_nnz set and the row
 * pointer filled (C->col / C->val then hold an unspecified prefix; call
 * again with room).
 * Replaces nsp_spgemm_kernel_hash's computation (nsp/src/nsp.cu:2448-2488)
 * for callers that own C's buffers. */
int nsp_spgemm_into(nsp_handle_t h, nsp_index_t it, nsp_value_t vt,
                    const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C,
                    int64_t *c_nnz);

/* Bytes of the device array that nsp_spgemm_noalloc uses as its scratch
 * (one copy per device, held until the process ends). */
#define NSP_NOALLOC_SCRATCH_BYTES (64u << 20)

/* C = A*B with no allocation: no hipMalloc, hipFree or allocating library
 * call.  No counterpart in the reference; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The caller passes C->row_ptr
 * (A->rows+1 entries) and C->col / C->val with a capacity of C->nnz
 * entries; the per-call counters and the workspace live in the library's
 * static device array of NSP_NOALLOC_SCRATCH_BYTES (A needs about
 * 37 bytes per row + 12 bytes per entry of it; larger A is rejected with
 * NSP_ERR_UNSUPPORTED before any work).  The kernels are those of
 * nsp_spgemm, launched for every bin with the row counts read on the device,
 * so the call issues no host synchronisation between them.  With nnz
 * non-NULL it waits for the stream and returns nnz(C); rows past the
 * capacity are skipped and NSP_ERR_OVERFLOW returned with *nnz = the size C
 * needs.  With nnz NULL it only enqueues the work (nnz(C) is then
 * C->row_ptr[A->rows]).  Every call on one device shares that device's
 * array: calls that overlap on one device (other streams or threads) race. */
int nsp_spgemm_noalloc(nsp_handle_t h, nsp_index_t it, nsp_value_t vt,
                       const nsp_csr_t *A, const nsp_csr_t *B, nsp_csr_t *C, int64_t *nnz);

/* Frees a C returned by nsp_spgemm with the default allocator
 * (CSRm::free, src/CSR.cu:572-621). */
int nsp_csr_free(nsp_csr_t *C);

/* Intermediate-product count of every row of A*B into intprod (int64,
 * A->rows entries, device).  Returns the total in *nprod.  Replaces
 * nsp_set_intprod_num (nsp/src/nsp.cu:306-320). */
int nsp_intprod(nsp_handle_t h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B,
                int64_t *intprod, int64_t *nprod);

/* Device-side check of a CSR's preconditions: row_ptr monotone from 0 to nnz,
 * columns in [0, ncols), rows strictly increasing.  *bad_row gets the first
 * offending row or -1.  (The reference's CSRm::checkColumnsOrder,
 * src/CSR.cu:1485-1513.) */
int nsp_check_csr(nsp_handle_t h, nsp_index_t it, const nsp_csr_t *X, int64_t *bad_row);

/* ---- batched products and graph capture --------------------------------- */
/* Symbolic phase of `count` independent products C_b = A_b * B_b stored as
 * concatenated CSR: A holds the rows of A_0, A_1, ... one after another
 * (A->row_ptr runs over all of them), with columns local to each matrix;
 * a_row_off[b] is A_b's first row (device int64, count+1 entries, ending at
 * A->rows); B alike.  C comes out concatenated the same way (rows of C_b at
 * a_row_off[b]..).  Follow with nsp_spgemm_numeric on the same handle.  No
 * reference counterpart: the reference can only loop nsp_spgemm_kernel_hash
 * (nsp/src/nsp.cu:2448) over the matrices. */
int nsp_spgemm_batched_symbolic(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                                const int64_t *a_row_off, const nsp_csr_t *B, const int64_t *b_row_off,
                                int64_t count, void *c_row_ptr, int64_t *c_nnz);

typedef struct nsp_graph_s *nsp_graph_t;

/* Records the whole product -- binning, symbolic pass, scan, numeric pass --
 * into one hipGraph with no host synchronisation inside, for repeated
 * launches over the same A and B buffers (values may change between
 * launches).  a_row_off == NULL records a single product; otherwise a batch
 * as in nsp_spgemm_batched_symbolic.  C->row_ptr (A->rows+1 entries) and
 * C->col / C->val (capacity C->nnz entries) are written by every launch;
 * rows that would pass the capacity are skipped and reported by
 * nsp_graph_result.  The graph launches on the handle's stream at creation.
 * No reference counterpart (the reference synchronises the host 5-9 times
 * per product, nsp/src/nsp.cu:417-471, 921-1130, 2178-2440). */
int nsp_graph_create(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                     const int64_t *a_row_off, const nsp_csr_t *B, const int64_t *b_row_off,
                     int64_t count, nsp_csr_t *C, nsp_graph_t *g);
int nsp_graph_launch(nsp_graph_t g);                  /* asynchronous */
int nsp_graph_result(nsp_graph_t g, int64_t *nnz);    /* synchronises; nnz(C) and status */
const char *nsp_graph_last_error(nsp_graph_t g);
int nsp_graph_destroy(nsp_graph_t g);

/* ---- backward pair on fixed patterns ------------------------------------ */
/* Transpose of X into XT (XT->rows = X->cols, row pointer, columns and values
 * caller-allocated; XT rows come out sorted by X's row index).  Replaces the
 * reference's cuSPARSE csr2csc transpose CSRm::T (src/CSR.cu:739-797); cached
 * by the caller as A^T for nsp_spgemm_backward.  Rows of XT up to 4096
 * entries are sorted in LDS, longer ones by an LDS-tile + merge-path sort
 * (one extra host synchronisation when there are any). */
int nsp_transpose(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *X, nsp_csr_t *XT);

/* Gradients of C = A*B for a loss L(C) on fixed patterns, given dC = dL/dC on
 * C's pattern (C->row_ptr, C->col from the forward; dC aligned with C->col):
 *   dA = (dC B^T) sampled on pat(A)   (dA aligned with A->col)
 *   dB = (A^T dC) sampled on pat(B)   (dB aligned with B->col)
 * AT is A's transpose from nsp_transpose.  at_perm (optional, index type
 * `it`, AT->nnz entries) maps each A^T entry to its A entry; when it is given
 * A's values are read as A->val[at_perm[t]] and AT->val may be NULL, so a
 * cached A^T pattern serves any values of A.  Deterministic.  C rows of any
 * length: up to 1024 entries through an LDS table of the row, longer ones by
 * binary search in the sorted C row.  No host synchronisation when the
 * handle's cached B row extents apply.  No reference counterpart (the
 * reference has no backward). */
int nsp_spgemm_backward(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                        const nsp_csr_t *B, const nsp_csr_t *AT, const void *at_perm,
                        const nsp_csr_t *C, const void *dC, void *dA, void *dB);

/* ---- chained products ----------------------------------------------------- */
/* Galerkin (coarse-operator) product C = R * (A * P), with R = P^T built on
 * the device (nsp_transpose) when R is NULL; callee-allocated C as in
 * nsp_spgemm (the intermediate A*P and P^T are freed before returning).
 * No reference counterpart: the reference's driver computes A*P only and
 * names R(AP) as the product it serves (main.cu:138); src/setting.h:63
 * (GALERKIN_PRODUCT_TYPE) is the switch of the AMG code it comes from. */
int nsp_galerkin(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *R, const nsp_csr_t *A,
                 const nsp_csr_t *P, nsp_csr_t *C, nsp_alloc_fn alloc, nsp_free_fn dealloc, void *user);

/* ---- one rank's local share on one device ------------------------------- */
/* Column compaction ("shrink"): perm = sorted unique(col U [0..l]),
 * new_col[e] = position of col[e] in perm, *shrinked_m = |perm|.  Every
 * col[e] must lie in [lo, hi) (columns are in the rank frame and may be
 * negative); perm needs at most max(hi, l+1) - min(lo, 0) entries
 * (perm_cap).  Replaces get_shrinked_col / shrink_col (src/spspmpi.cu:106-
 * 162), getmct_4shrink (src/getmct.cu:135-220) and
 * apply_mask_permut_GPU_noSideEffects (src/spspmpi.cu:58-104). */
int nsp_shrink_cols(nsp_handle_t h, int64_t nnz, const int32_t *col, int64_t l, int64_t lo, int64_t hi,
                    int32_t *new_col, int32_t *perm, int64_t perm_cap, int64_t *shrinked_m);

/* Row pointer (nr+1 entries, width it) of the CSR made of B rows
 * rows[i] + row_offset, i < nr (device list); *out_nnz gets its nnz.  With
 * nsp_gather_rows_fill this assembles completedP on one device, replacing the
 * halo-row exchange and _completedP_rows2 (src/spspmpi.cu:37-55, 584-906). */
int nsp_gather_rows_count(nsp_handle_t h, nsp_index_t it, const nsp_csr_t *B, int64_t nr, const int32_t *rows,
                          int64_t row_offset, void *out_row_ptr, int64_t *out_nnz);

/* Fills out->col (B columns minus col_shift: the column re-shift of
 * src/spspmpi.cu:822-828) and out->val for the rows counted above;
 * out->row_ptr is the row pointer from nsp_gather_rows_count. */
int nsp_gather_rows_fill(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *B, int64_t nr,
                         const int32_t *rows, int64_t row_offset, int64_t col_shift, nsp_csr_t *out);

/* ---- binary COO files (the reference's on-disk matrix format) ----------- */
/* File layout: an int64 header {nr, nc, nt} (header_type 1) or {nr, nt}
 * (header_type 0, nc = nr), then nt records {int64 i, int64 j, double v},
 * 1-based, sorted by row (src/READ_MATRIX_MPI.cpp:146-356 with Ti = long,
 * the instantiation main.cu:164-165 uses).  Rank `rank` of `nprocs` owns
 * rows [nr/nprocs*rank, nr/nprocs*(rank+1)), the last rank up to nr. */
typedef struct {
  int64_t nr, nc, nt;          /* the header */
  int64_t header_bytes;        /* 24 or 16 */
  int64_t rec_begin, rec_end;  /* the rank's records [rec_begin, rec_end) */
  int64_t first_row;           /* 0-based row of the rank's first record */
  int64_t nrows;               /* rows spanned, first to last present row (rowxproc) */
  int64_t col_shift;           /* sum of nrows over the ranks before this one */
} nsp_coo_share_t;

/* Host only (no device work): reads the header and locates the rank's
 * records by binary search on the row id, and the col_shift that the
 * reference derives with MPI_Exscan over every rank's rowxproc.  Replaces the
 * header read and binsearch_iat of READ_BINARY_MATRIX_MPI
 * (src/READ_MATRIX_MPI.cpp:83-130, 230-270) and the rowxproc / col_shift
 * computation of load_MatrixMPI (src/distribute.cu:1094-1095).  A rank with
 * no records gets nrows 0 (the reference reads irow[-1] there). */
int nsp_coo_share(const char *path, int header_type, int nprocs, int rank, nsp_coo_share_t *out);

/* CSR of `nrec` records already in device memory (24-byte records as in the
 * file), rows first_row .. first_row+out->rows-1 (0-based): out->row_ptr
 * (out->rows+1 entries, width it), out->col = j-1-col_shift (int32),
 * out->val = v (FP64, or rounded to FP32 with NSP_F32), all caller-allocated
 * device memory of out->nnz = nrec entries.  Rows whose columns are out of
 * order are sorted stably, as check_and_fix_order's bubble sort does.
 * Records out of row order or outside the rows return NSP_ERR_INVALID; a
 * column that does not fit int32 returns NSP_ERR_OVERFLOW.  Replaces the
 * record unpacking of READ_BINARY_MATRIX_MPI (src/READ_MATRIX_MPI.cpp:
 * 329-341), irow2iat2, check_and_fix_order and apply_shift
 * (src/distribute.cu:998-1069, 1097-1105).  Synchronises the host once. */
int nsp_coo_to_csr(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const void *records, int64_t nrec,
                   int64_t first_row, int64_t col_shift, nsp_csr_t *out);

/* The rank's share of a binary COO file straight into a device CSR:
 * records stream from the file through pinned staging buffers and are
 * converted on the device chunk by chunk, reads overlapping copies and
 * conversion.  out->row_ptr (share->nrows+1), out->col and out->val
 * (share->rec_end - share->rec_begin entries) are caller-allocated device
 * memory; out->rows/cols/nnz are set.  Errors as nsp_coo_to_csr.  Replaces
 * load_MatrixMPI (src/distribute.cu:1071-1125) for one rank. */
int nsp_coo_read(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const char *path, int header_type,
                 const nsp_coo_share_t *share, nsp_csr_t *out);

/* ---- ASCII COO files (the reference's text format) --------------------- */
/* File layout (src/READ_MATRIX_MPI.cpp:364-578, read when load_MatrixMPI's
 * `ascii` is set, src/distribute.cu:1084-1086): lines starting with '%',
 * then a header line "nr nc nt" (or "nr nt", nc = nr), then nt record lines
 * "i j v" (1-based, sorted by row) that all have the length of the first
 * one, so record q starts at body_offset + q*line_bytes.  The column field
 * is read from col_offset and the value from val_offset, found in the first
 * record line as the reference finds them (the first space after the row
 * field, the first space after the column field).  Rows and columns are
 * read as atol reads them; values with the reference's own decimal parser
 * (fast_atof, src/READ_MATRIX_MPI.cpp:21-77: digits accumulated in FP64 as
 * value*10 + d, fraction digits added as d/10^k, the exponent applied as
 * factors 1e50, 1e8 and 10), so every value is bit-identical to the one
 * the reference reads.  Same rank split and col_shift as the binary files. */
typedef struct {
  int64_t nr, nc, nt;            /* the header */
  int64_t body_offset;           /* first byte of record 0 */
  int64_t line_bytes;            /* bytes per record line, newline included */
  int32_t col_offset, val_offset;
  int64_t rec_begin, rec_end;    /* the rank's records [rec_begin, rec_end) */
  int64_t first_row;             /* 0-based row of the rank's first record */
  int64_t nrows;                 /* rows spanned, first to last present row (rowxproc) */
  int64_t col_shift;             /* sum of nrows over the ranks before this one */
} nsp_ascii_share_t;

/* Host only: reads the header and the first record line, checks that the
 * file holds nt lines of that length, and locates the rank's records by
 * binary search on the row field (the reference's binsearch_iat_str,
 * src/READ_MATRIX_MPI.cpp:83-104, and its split, 440-455).  The body
 * starts after the '%' lines and the header line (the reference counts only
 * the header line's length, so it reads a file with '%' lines from the
 * wrong offset; files without them read the same).  NSP_ERR_INVALID for a
 * missing header, a missing or empty first record line, or a file too short
 * for nt lines. */
int nsp_ascii_share(const char *path, int nprocs, int rank, nsp_ascii_share_t *out);

/* The rank's block as a device CSR, exactly as nsp_coo_read builds it from
 * the same records (row pointer, columns in the rank frame, values FP64 or
 * rounded to FP32, unsorted rows sorted stably): the record lines are read
 * in chunks of "coo_chunk_records" lines through pinned buffers and parsed
 * on the device, one lane per line.  A record line whose last byte is not a
 * newline (lines of unequal length, which the reference would read
 * misaligned) returns NSP_ERR_INVALID, as do records out of row order. */
int nsp_ascii_read(nsp_handle_t h, nsp_index_t it, nsp_value_t vt, const char *path,
                   const nsp_ascii_share_t *share, nsp_csr_t *out);

#ifdef __cplusplus
}
#endif
#endif /* NSP_AMD_H */
