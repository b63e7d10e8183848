// Entry points of the local-product driver (spgemm.hip) used by the C ABI.
#pragma once
#include "nsp_internal.hpp"

namespace nsp {
int spgemm_symbolic(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    void *c_row_ptr, int64_t *c_nnz);
int spgemm_numeric(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                   nsp_csr_t *C);
// One-call product: C's row pointer and nnz (rows kept in the handle), then
// C's columns and values; the pair is called back to back with the same A, B.
// With `into` (C's buffers, capacity into->nnz) the rows may be copied
// there already: *filled says so, and spgemm_fused_fill is then not called.
int spgemm_fused_count(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                       void *c_row_ptr, int64_t *c_nnz, nsp_csr_t *into = nullptr, bool *filled = nullptr);
int spgemm_fused_fill(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                      nsp_csr_t *C);
int spgemm_intprod(Handle *h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, int64_t *intprod,
                   int64_t *nprod);
int check_csr(Handle *h, nsp_index_t it, const nsp_csr_t *X, int64_t *bad_row);
// Fills the handle's per-A-entry B row extents (ab_b0, ab_len) for A*B; ends
// any symbolic plan held by the handle.
int spgemm_batched_symbolic(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const int64_t *a_off,
                            const nsp_csr_t *B, const int64_t *b_off, int64_t count, void *c_row_ptr,
                            int64_t *c_nnz);
// Product with the workspace in the library's static device array (no
// allocation); noalloc_setup looks the array's device address up once per
// device.
int noalloc_setup(Handle *h);
int spgemm_noalloc(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                   nsp_csr_t *C, int64_t *nnz);
struct GraphPlan;
int graph_create(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const int64_t *a_off,
                 const nsp_csr_t *B, const int64_t *b_off, int64_t count, nsp_csr_t *C, GraphPlan **out);
int graph_launch(GraphPlan *gp);
int graph_result(GraphPlan *gp, int64_t *nnz);
const char *graph_last_error(GraphPlan *gp);
int graph_destroy(GraphPlan *gp);
/* B row extents of every A entry (ab_b0, ab_len) and nprod; *recomputed
 * tells whether the cached ones were replaced. */
int spgemm_row_extents(Handle *h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, bool *recomputed = nullptr);
int transpose(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *X, nsp_csr_t *XT);
int spgemm_backward(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    const nsp_csr_t *AT, const void *at_perm, const nsp_csr_t *C, const void *dc, void *da,
                    void *db);
int shrink_cols(Handle *h, int64_t nnz, const int32_t *col, int64_t l, int64_t lo, int64_t hi,
                int32_t *new_col, int32_t *perm, int64_t perm_cap, int64_t *shrinked_m);
int gather_rows_count(Handle *h, nsp_index_t it, const nsp_csr_t *B, int64_t nr, const int32_t *rows,
                      int64_t row_offset, void *out_rpt, int64_t *out_nnz);
int gather_rows_fill(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *B, int64_t nr,
                     const int32_t *rows, int64_t row_offset, int64_t col_shift, nsp_csr_t *out);
}  // namespace nsp
