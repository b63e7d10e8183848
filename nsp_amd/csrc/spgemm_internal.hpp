// Entry points of the local-product driver (spgemm.hip) used by the C ABI.
#pragma once
#include "nsp_internal.hpp"

namespace nsp {
int spgemm_symbolic(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    void *c_row_ptr, int64_t *c_nnz);
int spgemm_numeric(Handle *h, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                   nsp_csr_t *C);
int spgemm_intprod(Handle *h, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, int64_t *intprod,
                   int64_t *nprod);
int check_csr(Handle *h, nsp_index_t it, const nsp_csr_t *X, int64_t *bad_row);
}  // namespace nsp
