// extern "C" entry points of libnsp_amd.so (declared in include/nsp_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>

#include "device_utils.hpp"
#include "nsp_internal.hpp"
#include "spgemm_internal.hpp"

namespace nsp {

int set_error(Handle *h, int status, const std::string &msg) {
  if (h) h->last_error = msg;
  return status;
}

// CSR precondition check: row_ptr[0] == 0, monotone, row_ptr[rows] == nnz,
// columns in [0, cols), strictly increasing within a row.
template <typename RowT>
__global__ void k_check_csr(int64_t rows, int64_t cols, int64_t nnz, const RowT *__restrict__ rpt,
                            const int32_t *__restrict__ col, unsigned long long *bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows; i += stride) {
    const int64_t p0 = (int64_t)rpt[i], p1 = (int64_t)rpt[i + 1];
    bool ok = p0 <= p1 && p0 >= 0 && p1 <= nnz;
    if (i == 0 && p0 != 0) ok = false;
    if (i == rows - 1 && p1 != nnz) ok = false;
    if (ok) {
      int32_t prev = -1;
      for (int64_t p = p0; p < p1; p++) {
        const int32_t c = col[p];
        if (c <= prev || (int64_t)c >= cols) { ok = false; break; }
        prev = c;
      }
    }
    if (!ok) atomicMin(bad, (unsigned long long)i);
  }
}

int check_csr(Handle *h, nsp_index_t it, const nsp_csr_t *X, int64_t *bad_row) {
  if (!X || !bad_row) return set_error(h, NSP_ERR_INVALID, "NULL argument");
  *bad_row = -1;
  if (X->rows <= 0) return NSP_OK;
  unsigned long long *d_bad = nullptr;
  NSP_HIP(h, hipMallocAsync((void **)&d_bad, sizeof(unsigned long long), h->stream));
  const unsigned long long init = ~0ull;
  NSP_HIP(h, hipMemcpyAsync(d_bad, &init, sizeof init, hipMemcpyHostToDevice, h->stream));
  const int g = (int)std::min<int64_t>((X->rows + 255) / 256, (int64_t)h->num_cu * 8);
  if (it == NSP_IDX32)
    hipLaunchKernelGGL(k_check_csr<int32_t>, dim3(g), dim3(256), 0, h->stream, X->rows, X->cols, X->nnz,
                       (const int32_t *)X->row_ptr, X->col, d_bad);
  else
    hipLaunchKernelGGL(k_check_csr<int64_t>, dim3(g), dim3(256), 0, h->stream, X->rows, X->cols, X->nnz,
                       (const int64_t *)X->row_ptr, X->col, d_bad);
  unsigned long long hb = 0;
  NSP_HIP(h, hipMemcpyAsync(&hb, d_bad, sizeof hb, hipMemcpyDeviceToHost, h->stream));
  NSP_HIP(h, hipFreeAsync(d_bad, h->stream));
  NSP_HIP(h, hipStreamSynchronize(h->stream));
  *bad_row = (hb == ~0ull) ? -1 : (int64_t)hb;
  return NSP_OK;
}

}  // namespace nsp

using nsp::Handle;

extern "C" {

int nsp_create(nsp_handle_t *out) {
  if (!out) return NSP_ERR_INVALID;
  *out = nullptr;
  nsp_handle_s *hs = new (std::nothrow) nsp_handle_s();
  if (!hs) return NSP_ERR_OOM;
  Handle *h = &hs->h;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete hs;
    return NSP_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, h->device) == hipSuccess && prop.multiProcessorCount > 0)
    h->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete hs;
    return NSP_ERR_HIP;
  }
  h->own_stream = true;
  if (hipMalloc((void **)&h->d_info, sizeof(nsp::Info)) != hipSuccess ||
      hipHostMalloc((void **)&h->h_info, sizeof(nsp::Info), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    if (h->d_info) (void)hipFree(h->d_info);
    (void)hipStreamDestroy(h->stream);
    delete hs;
    return NSP_ERR_OOM;
  }
  std::memset(h->h_info, 0, sizeof(nsp::Info));
  // device address of the no-allocation product's scratch array on this
  // device (a failure here is reported again by nsp_spgemm_noalloc)
  (void)nsp::noalloc_setup(h);
  h->last_error.clear();
  *out = hs;
  return NSP_OK;
}

int nsp_destroy(nsp_handle_t hs) {
  if (!hs) return NSP_OK;
  Handle *h = &hs->h;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->ws.base) (void)hipFree(h->ws.base);
  if (h->aux.base) (void)hipFree(h->aux.base);
  if (h->aux2.base) (void)hipFree(h->aux2.base);
  if (h->bwd.base) (void)hipFree(h->bwd.base);
  if (h->segw.base) (void)hipFree(h->segw.base);
  if (h->coo.base) (void)hipFree(h->coo.base);
  if (h->rsort.base) (void)hipFree(h->rsort.base);
  if (h->spec.base) (void)hipFree(h->spec.base);
  if (h->pin) (void)hipHostFree(h->pin);
  for (hipEvent_t e : h->pin_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->events)
    if (e) (void)hipEventDestroy(e);
  if (h->d_info) (void)hipFree(h->d_info);
  if (h->h_info) (void)hipHostFree(h->h_info);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete hs;
  return NSP_OK;
}

int nsp_set_stream(nsp_handle_t hs, void *stream) {
  if (!hs) return NSP_ERR_INVALID;
  Handle *h = &hs->h;
  if (h->own_stream && h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  h->own_stream = false;
  h->stream = (hipStream_t)stream;
  return NSP_OK;
}

const char *nsp_status_string(int s) {
  switch (s) {
    case NSP_OK: return "NSP_OK";
    case NSP_ERR_INVALID: return "NSP_ERR_INVALID";
    case NSP_ERR_OOM: return "NSP_ERR_OOM";
    case NSP_ERR_OVERFLOW: return "NSP_ERR_OVERFLOW";
    case NSP_ERR_HIP: return "NSP_ERR_HIP";
    case NSP_ERR_STATE: return "NSP_ERR_STATE";
    case NSP_ERR_UNSUPPORTED: return "NSP_ERR_UNSUPPORTED";
    default: return "NSP_ERR_UNKNOWN";
  }
}

const char *nsp_last_error(nsp_handle_t hs) { return hs ? hs->h.last_error.c_str() : ""; }

int nsp_get_stats(nsp_handle_t hs, nsp_stats_t *out) {
  if (!hs || !out) return NSP_ERR_INVALID;
  NSP_TRY(nsp::collect_timing(&hs->h));
  *out = hs->h.stats;
  return NSP_OK;
}

int nsp_set_param(nsp_handle_t hs, const char *name, int64_t value) {
  if (!hs || !name) return NSP_ERR_INVALID;
  Handle *h = &hs->h;
  auto rpw = [&](int *dst, const char *what, bool small) {
    if (value != 1 && value != 2 && value != 4 && !(small && value == 8))
      return nsp::set_error(h, NSP_ERR_INVALID,
                            std::string(what) + (small ? " must be 1, 2, 4 or 8" : " must be 1, 2 or 4"));
    *dst = (int)value;
    return (int)NSP_OK;
  };
  if (!std::strcmp(name, "sym_small_rows_per_wave")) {
    return rpw(&h->sym_small_rows_per_wave, name, true);
  } else if (!std::strcmp(name, "sym_rows_per_wave")) {
    return rpw(&h->sym_rows_per_wave, name, false);
  } else if (!std::strcmp(name, "num_small_rows_per_wave")) {
    return rpw(&h->num_small_rows_per_wave, name, true);
  } else if (!std::strcmp(name, "num_rows_per_wave")) {
    return rpw(&h->num_rows_per_wave, name, false);
  } else if (!std::strcmp(name, "bwd_contrib")) {
    if (value < 0 || value > 2) return nsp::set_error(h, NSP_ERR_INVALID, "bwd_contrib must be 0, 1 or 2");
    h->bwd_contrib = (int)value;
  } else if (!std::strcmp(name, "fp32_accumulate")) {
    if (value != 0 && value != 1) return nsp::set_error(h, NSP_ERR_INVALID, "fp32_accumulate must be 0 or 1");
    h->fp32_accumulate = (int)value;
  } else if (!std::strcmp(name, "narrow_mode")) {
    if (value != 0 && value != 1) return nsp::set_error(h, NSP_ERR_INVALID, "narrow_mode must be 0 or 1");
    h->narrow_mode = (int)value;
  } else if (!std::strcmp(name, "fused_mode")) {
    if (value < 0 || value > 1) return nsp::set_error(h, NSP_ERR_INVALID, "fused_mode must be 0 or 1");
    h->fused_mode = (int)value;
  } else if (!std::strcmp(name, "sym_table_factor")) {
    if (value != 1 && value != 2) return nsp::set_error(h, NSP_ERR_INVALID, "sym_table_factor must be 1 or 2");
    h->sym_table_factor = (int)value;
  } else if (!std::strcmp(name, "num_table_factor")) {
    if (value != 2 && value != 4) return nsp::set_error(h, NSP_ERR_INVALID, "num_table_factor must be 2 or 4");
    h->num_table_factor = (int)value;
  } else if (!std::strcmp(name, "timing")) {
    h->timing = value != 0;
  } else if (!std::strcmp(name, "reuse_extents")) {
    h->reuse_extents = value != 0;
  } else if (!std::strcmp(name, "heavy_blocks")) {
    if (value < 0 || value > 65536) return nsp::set_error(h, NSP_ERR_INVALID, "heavy_blocks must be in [0, 65536]");
    h->heavy_blocks = (int)value;
  } else if (!std::strcmp(name, "coo_chunk_records")) {
    if (value < 1 || value > (int64_t(1) << 26))
      return nsp::set_error(h, NSP_ERR_INVALID, "coo_chunk_records must be in [1, 2^26]");
    h->coo_chunk = value;
  } else if (!std::strcmp(name, "queue_mode")) {
    if (value != 0 && value != 1) return nsp::set_error(h, NSP_ERR_INVALID, "queue_mode must be 0 or 1");
    h->queue_mode = (int)value;
  } else if (!std::strcmp(name, "queue_chunk")) {
    if (value < 1 || value > 64) return nsp::set_error(h, NSP_ERR_INVALID, "queue_chunk must be in [1, 64]");
    h->queue_chunk = (int)value;
  } else if (!std::strcmp(name, "queue_blocks")) {
    if (value < 1 || value > 16) return nsp::set_error(h, NSP_ERR_INVALID, "queue_blocks must be in [1, 16]");
    h->queue_blocks = (int)value;
  } else if (!std::strcmp(name, "grid_cap")) {
    if (value < 1 || value > 64) return nsp::set_error(h, NSP_ERR_INVALID, "grid_cap must be in [1, 64]");
    h->grid_cap = (int)value;
  } else {
    return nsp::set_error(h, NSP_ERR_INVALID, std::string("unknown parameter ") + name);
  }
  return NSP_OK;
}

int nsp_spgemm_symbolic(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                        const nsp_csr_t *B, void *c_row_ptr, int64_t *c_nnz) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_symbolic(&hs->h, it, vt, A, B, c_row_ptr, c_nnz);
}

int nsp_spgemm_numeric(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                       const nsp_csr_t *B, nsp_csr_t *C) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_numeric(&hs->h, it, vt, A, B, C);
}

static void *default_alloc(size_t bytes, void *) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
static void default_free(void *p, void *) {
  if (p) (void)hipFree(p);
}

int nsp_spgemm_noalloc(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                       nsp_csr_t *C, int64_t *nnz) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_noalloc(&hs->h, it, vt, A, B, C, nnz);
}

int nsp_spgemm_into(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                    nsp_csr_t *C, int64_t *c_nnz) {
  if (!hs || !C || !c_nnz) return NSP_ERR_INVALID;
  Handle *h = &hs->h;
  if (!C->row_ptr) return nsp::set_error(h, NSP_ERR_INVALID, "C.row_ptr is NULL");
  if (C->nnz < 0) return nsp::set_error(h, NSP_ERR_INVALID, "C.nnz (the capacity) is negative");
  const int64_t cap = C->nnz;
  if (cap > 0 && (!C->col || !C->val)) return nsp::set_error(h, NSP_ERR_INVALID, "C.col/C.val is NULL");
  int64_t nnz = 0;
  bool filled = false;
  nsp_csr_t into = *C;
  int s = nsp::spgemm_fused_count(h, it, vt, A, B, C->row_ptr, &nnz, &into, &filled);
  if (s != NSP_OK) return s;
  *c_nnz = nnz;
  if (nnz > cap) {
    h->fused_plan = false;
    h->have_plan = false;
    return nsp::set_error(h, NSP_ERR_OVERFLOW, "C needs " + std::to_string(nnz) + " entries, more than its capacity " +
                                                   std::to_string(cap) + " (C.row_ptr is filled)");
  }
  if (filled) return NSP_OK;
  nsp_csr_t c = *C;
  c.rows = A->rows;
  c.cols = B->cols;
  c.nnz = nnz;
  return nsp::spgemm_fused_fill(h, it, vt, A, B, &c);
}

int nsp_spgemm(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
               nsp_csr_t *C, nsp_alloc_fn alloc, nsp_free_fn dealloc, void *user) {
  if (!hs || !C) return NSP_ERR_INVALID;
  Handle *h = &hs->h;
  if (!alloc) alloc = default_alloc;
  if (!dealloc) dealloc = default_free;
  if (!A) return nsp::set_error(h, NSP_ERR_INVALID, "A is NULL");
  const size_t isz = (it == NSP_IDX64) ? 8 : 4, vsz = (vt == NSP_F32) ? 4 : 8;
  std::memset(C, 0, sizeof *C);
  void *rpt = alloc((size_t)(A->rows + 1) * isz, user);
  if (!rpt) return nsp::set_error(h, NSP_ERR_OOM, "C.row_ptr allocation failed");
  int64_t nnz = 0;
  int s = nsp::spgemm_fused_count(h, it, vt, A, B, rpt, &nnz);
  if (s != NSP_OK) {
    dealloc(rpt, user);
    return s;
  }
  void *col = alloc((size_t)nnz * 4, user);
  void *val = alloc((size_t)nnz * vsz, user);
  if (!col || !val) {
    dealloc(rpt, user);
    dealloc(col, user);
    dealloc(val, user);
    return nsp::set_error(h, NSP_ERR_OOM, "C.col/C.val allocation failed");
  }
  C->rows = A->rows;
  C->cols = B->cols;
  C->nnz = nnz;
  C->row_ptr = rpt;
  C->col = (int32_t *)col;
  C->val = val;
  s = nsp::spgemm_fused_fill(h, it, vt, A, B, C);
  if (s == NSP_OK && hipStreamSynchronize(h->stream) != hipSuccess)
    s = nsp::set_error(h, NSP_ERR_HIP, hipGetErrorString(hipGetLastError()));
  if (s != NSP_OK) {
    dealloc(rpt, user);
    dealloc(col, user);
    dealloc(val, user);
    std::memset(C, 0, sizeof *C);
  }
  return s;
}

int nsp_galerkin(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *R, const nsp_csr_t *A,
                 const nsp_csr_t *P, nsp_csr_t *C, nsp_alloc_fn alloc, nsp_free_fn dealloc, void *user) {
  if (!hs || !C) return NSP_ERR_INVALID;
  Handle *h = &hs->h;
  if (!A || !P) return nsp::set_error(h, NSP_ERR_INVALID, "nsp_galerkin: A or P is NULL");
  if (A->cols != P->rows) return nsp::set_error(h, NSP_ERR_INVALID, "nsp_galerkin: A.cols != P.rows");
  if (R ? R->cols != A->rows : P->rows != A->rows)
    return nsp::set_error(h, NSP_ERR_INVALID, "nsp_galerkin: R.cols (P.rows for R = P^T) != A.rows");
  std::memset(C, 0, sizeof *C);
  nsp_csr_t AP{};
  int s = nsp_spgemm(hs, it, vt, A, P, &AP, nullptr, nullptr, nullptr);
  if (s != NSP_OK) return s;
  nsp_csr_t PT{};
  if (!R) {
    const size_t isz = (it == NSP_IDX64) ? 8 : 4, vsz = (vt == NSP_F32) ? 4 : 8;
    PT.rows = P->cols;
    PT.cols = P->rows;
    PT.nnz = P->nnz;
    if (hipMalloc(&PT.row_ptr, (size_t)(PT.rows + 1) * isz) != hipSuccess ||
        hipMalloc((void **)&PT.col, std::max<size_t>((size_t)PT.nnz * 4, 4)) != hipSuccess ||
        hipMalloc(&PT.val, std::max<size_t>((size_t)PT.nnz * vsz, 8)) != hipSuccess) {
      (void)hipGetLastError();
      s = nsp::set_error(h, NSP_ERR_OOM, "nsp_galerkin: P^T allocation failed");
    } else {
      s = nsp_transpose(hs, it, vt, P, &PT);
    }
    R = &PT;
  }
  if (s == NSP_OK) s = nsp_spgemm(hs, it, vt, R, &AP, C, alloc, dealloc, user);
  (void)nsp_csr_free(&AP);
  if (PT.row_ptr || PT.col || PT.val) (void)nsp_csr_free(&PT);
  return s;
}

int nsp_csr_free(nsp_csr_t *C) {
  if (!C) return NSP_ERR_INVALID;
  if (C->row_ptr) (void)hipFree(C->row_ptr);
  if (C->col) (void)hipFree(C->col);
  if (C->val) (void)hipFree(C->val);
  std::memset(C, 0, sizeof *C);
  return NSP_OK;
}

int nsp_intprod(nsp_handle_t hs, nsp_index_t it, const nsp_csr_t *A, const nsp_csr_t *B, int64_t *intprod,
                int64_t *nprod) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_intprod(&hs->h, it, A, B, intprod, nprod);
}

int nsp_check_csr(nsp_handle_t hs, nsp_index_t it, const nsp_csr_t *X, int64_t *bad_row) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::check_csr(&hs->h, it, X, bad_row);
}

int nsp_shrink_cols(nsp_handle_t hs, int64_t nnz, const int32_t *col, int64_t l, int64_t lo, int64_t hi,
                    int32_t *new_col, int32_t *perm, int64_t perm_cap, int64_t *shrinked_m) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::shrink_cols(&hs->h, nnz, col, l, lo, hi, new_col, perm, perm_cap, shrinked_m);
}

int nsp_gather_rows_count(nsp_handle_t hs, nsp_index_t it, const nsp_csr_t *B, int64_t nr, const int32_t *rows,
                          int64_t row_offset, void *out_row_ptr, int64_t *out_nnz) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::gather_rows_count(&hs->h, it, B, nr, rows, row_offset, out_row_ptr, out_nnz);
}

int nsp_spgemm_batched_symbolic(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A,
                                const int64_t *a_row_off, const nsp_csr_t *B, const int64_t *b_row_off,
                                int64_t count, void *c_row_ptr, int64_t *c_nnz) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_batched_symbolic(&hs->h, it, vt, A, a_row_off, B, b_row_off, count, c_row_ptr, c_nnz);
}

struct nsp_graph_s {
  nsp::GraphPlan *gp;
};

int nsp_graph_create(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const int64_t *a_row_off,
                     const nsp_csr_t *B, const int64_t *b_row_off, int64_t count, nsp_csr_t *C, nsp_graph_t *g) {
  if (!hs || !g) return NSP_ERR_INVALID;
  *g = nullptr;
  nsp::GraphPlan *gp = nullptr;
  int s = nsp::graph_create(&hs->h, it, vt, A, a_row_off, B, b_row_off, count, C, &gp);
  if (s != NSP_OK) return s;
  *g = new (std::nothrow) nsp_graph_s{gp};
  if (!*g) {
    nsp::graph_destroy(gp);
    return NSP_ERR_OOM;
  }
  return NSP_OK;
}

int nsp_graph_launch(nsp_graph_t g) { return g ? nsp::graph_launch(g->gp) : NSP_ERR_INVALID; }

int nsp_graph_result(nsp_graph_t g, int64_t *nnz) { return g ? nsp::graph_result(g->gp, nnz) : NSP_ERR_INVALID; }

const char *nsp_graph_last_error(nsp_graph_t g) { return g ? nsp::graph_last_error(g->gp) : ""; }

int nsp_graph_destroy(nsp_graph_t g) {
  if (!g) return NSP_OK;
  int s = nsp::graph_destroy(g->gp);
  delete g;
  return s;
}

int nsp_transpose(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *X, nsp_csr_t *XT) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::transpose(&hs->h, it, vt, X, XT);
}

int nsp_spgemm_backward(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *A, const nsp_csr_t *B,
                        const nsp_csr_t *AT, const void *at_perm, const nsp_csr_t *C, const void *dC, void *dA,
                        void *dB) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::spgemm_backward(&hs->h, it, vt, A, B, AT, at_perm, C, dC, dA, dB);
}

int nsp_gather_rows_fill(nsp_handle_t hs, nsp_index_t it, nsp_value_t vt, const nsp_csr_t *B, int64_t nr,
                         const int32_t *rows, int64_t row_offset, int64_t col_shift, nsp_csr_t *out) {
  if (!hs) return NSP_ERR_INVALID;
  return nsp::gather_rows_fill(&hs->h, it, vt, B, nr, rows, row_offset, col_shift, out);
}

}  // extern "C"
