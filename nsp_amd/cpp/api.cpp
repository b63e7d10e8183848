// C++ host API (include/nsp_amd.hpp) over the C ABI of libnsp_amd.so.
// Host code only: device memory through the HIP runtime, compute through the
// library's entry points.
#include "nsp_amd.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

namespace nspamd {

namespace {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw Error(NSP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// One library handle per host thread and device, bound to the legacy (null)
// stream: the reference's calls are host-synchronous on the legacy stream,
// so work the caller queued there before a call is ordered before it.
struct ThreadHandles {
  std::map<int, nsp_handle_t> by_device;
  ~ThreadHandles() {
    for (auto &kv : by_device) nsp_destroy(kv.second);
  }
};

nsp_handle_t handle() {
  thread_local ThreadHandles th;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  auto it = th.by_device.find(dev);
  if (it != th.by_device.end()) return it->second;
  nsp_handle_t h = nullptr;
  const int s = nsp_create(&h);
  if (s != NSP_OK) throw Error(s, "nsp_create failed");
  (void)nsp_set_stream(h, nullptr);
  th.by_device[dev] = h;
  return h;
}

void check(int status, nsp_handle_t h, const char *what) {
  if (status != NSP_OK)
    throw Error(status, std::string(what) + ": " + nsp_status_string(status) + " (" +
                            (h ? nsp_last_error(h) : "") + ")");
}

template <typename T>
T *dalloc(size_t n) {
  void *p = nullptr;
  hip_check(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
  return static_cast<T *>(p);
}

nsp_csr_t view(const CSR *A, bool compacted) {
  nsp_csr_t v;
  v.rows = (int64_t)A->n;
  v.cols = compacted ? (int64_t)A->shrinked_m : (int64_t)A->m;
  v.nnz = (int64_t)A->nnz;
  v.row_ptr = A->row;
  v.col = compacted ? A->shrinked_col : A->col;
  v.val = A->val;
  return v;
}

}  // namespace

// ---------------------------------------------------------------------------
void nsp_spgemm_kernel_hash(sfCSR *a, sfCSR *b, sfCSR *c) {
  if (!a || !b || !c) throw Error(NSP_ERR_INVALID, "nsp_spgemm_kernel_hash: NULL matrix");
  nsp_handle_t h = handle();
  // the reference reads only M, N and the device arrays of a and b; nnz
  // comes from the row pointer (a->nnz is not trusted, nsp.cu:2448-2488)
  int a_nnz = 0, b_nnz = 0;
  hip_check(hipMemcpy(&a_nnz, a->d_rpt + a->M, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  hip_check(hipMemcpy(&b_nnz, b->d_rpt + b->M, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  nsp_csr_t A{a->M, b->M, a_nnz, a->d_rpt, a->d_col, a->d_val};
  nsp_csr_t B{b->M, b->N, b_nnz, b->d_rpt, b->d_col, b->d_val};
  nsp_csr_t C{};
  check(nsp_spgemm(h, NSP_IDX32, NSP_F64, &A, &B, &C, nullptr, nullptr, nullptr), h, "nsp_spgemm_kernel_hash");
  c->M = a->M;
  c->N = b->N;
  c->nnz = (int)C.nnz;
  c->d_rpt = (int *)C.row_ptr;
  c->d_col = (int *)C.col;
  c->d_val = (double *)C.val;
}

void sfCSR_free_device(sfCSR *c) {
  if (!c) return;
  (void)hipFree(c->d_rpt);
  (void)hipFree(c->d_col);
  (void)hipFree(c->d_val);
  c->d_rpt = c->d_col = nullptr;
  c->d_val = nullptr;
}

// ---------------------------------------------------------------------------
namespace CSRm {

CSR *init(stype n, gstype m, stype nnz, bool allocate_mem, bool on_the_device, bool is_symmetric, gstype full_n,
          gstype row_shift) {
  CSR *A = new CSR();
  A->n = n;
  A->m = m;
  A->nnz = nnz;
  A->full_n = full_n;
  A->full_m = m;
  A->row_shift = row_shift;
  A->on_the_device = on_the_device;
  A->is_symmetric = is_symmetric;
  A->shrinked_m = m;
  if (allocate_mem) {
    if (on_the_device) {
      A->row = dalloc<int>((size_t)n + 1);
      A->col = dalloc<int>((size_t)nnz);
      A->val = dalloc<double>((size_t)nnz);
    } else {
      A->row = new int[(size_t)n + 1];
      A->col = new int[std::max<size_t>((size_t)nnz, 1)];
      A->val = new double[std::max<size_t>((size_t)nnz, 1)];
    }
  }
  return A;
}

void free(CSR *A) {
  if (!A) return;
  if (A->on_the_device) {
    (void)hipFree(A->row);
    (void)hipFree(A->col);
    (void)hipFree(A->val);
  } else {
    delete[] A->row;
    delete[] A->col;
    delete[] A->val;
  }
  (void)hipFree(A->shrinked_col);
  (void)hipFree(A->shrinking_permut);
  delete A;
}

void freeStruct(CSR *A) { delete A; }

CSR *copyToDevice(const CSR *A) {
  CSR *D = init(A->n, A->m, A->nnz, true, true, A->is_symmetric, A->full_n, A->row_shift);
  D->full_m = A->full_m;
  hip_check(hipMemcpy(D->row, A->row, ((size_t)A->n + 1) * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  hip_check(hipMemcpy(D->col, A->col, (size_t)A->nnz * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
  hip_check(hipMemcpy(D->val, A->val, (size_t)A->nnz * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
  return D;
}

CSR *copyToHost(const CSR *A) {
  CSR *H = init(A->n, A->m, A->nnz, true, false, A->is_symmetric, A->full_n, A->row_shift);
  H->full_m = A->full_m;
  hip_check(hipMemcpy(H->row, A->row, ((size_t)A->n + 1) * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  hip_check(hipMemcpy(H->col, A->col, (size_t)A->nnz * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  hip_check(hipMemcpy(H->val, A->val, (size_t)A->nnz * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
  return H;
}

double norm_fro(const CSR *A) {
  std::vector<double> v((size_t)A->nnz);
  if (A->on_the_device)
    hip_check(hipMemcpy(v.data(), A->val, v.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
  else
    std::memcpy(v.data(), A->val, v.size() * sizeof(double));
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

void printInfo(const CSR *A, FILE *fp) {
  // src/CSR.cu:631-643: every size with %d, i.e. the low 32 bits of the
  // reference's unsigned long fields
  fprintf(fp, "Device?: %d\n", A->on_the_device ? 1 : 0);
  fprintf(fp, "nnz: %d\n", (int)A->nnz);
  fprintf(fp, "n: %d\n", (int)A->n);
  fprintf(fp, "m: %d\n", (int)A->m);
  fprintf(fp, "full_n: %d\n", (int)A->full_n);
  fprintf(fp, "row_shift: %d\n", (int)A->row_shift);
  fprintf(fp, A->is_symmetric ? "SYMMETRIC\n" : "GENERAL\n");
  fprintf(fp, "\n");
}

}  // namespace CSRm

// ---------------------------------------------------------------------------
bool shrink_col(CSR *A, CSR *P) {
  if (A->shrinked_flag) return false;
  if (!A->on_the_device) throw Error(NSP_ERR_INVALID, "shrink_col: A must be on the device");
  nsp_handle_t h = handle();
  const int64_t l = (int64_t)(P ? P->n : A->n) - 1;
  const int64_t shift = (int64_t)A->row_shift;
  const int64_t lo = A->col_hi > A->col_lo ? A->col_lo : -shift;
  const int64_t hi = A->col_hi > A->col_lo ? A->col_hi : (int64_t)A->m - shift;
  const int64_t cap = std::max(hi, l + 1) - std::min<int64_t>(lo, 0);
  int *perm = dalloc<int>((size_t)cap);
  int *new_col = dalloc<int>((size_t)A->nnz);
  int64_t m = 0;
  const int s = nsp_shrink_cols(h, (int64_t)A->nnz, A->col, l, lo, hi, new_col, perm, cap, &m);
  if (s != NSP_OK) {
    (void)hipFree(perm);
    (void)hipFree(new_col);
    check(s, h, "shrink_col");
  }
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  A->shrinked_col = new_col;
  A->shrinked_m = (stype)m;
  A->shrinking_permut = perm;
  A->shrinking_permut_len = (stype)m;
  A->shrinked_flag = true;
  return true;
}

CSR *get_shrinked_matrix(CSR *A, CSR *P) {
  shrink_col(A, P);
  CSR *S = CSRm::init(A->n, A->shrinked_m, A->nnz, false, A->on_the_device, A->is_symmetric, A->full_n, A->row_shift);
  S->row = A->row;
  S->col = A->shrinked_col;
  S->val = A->val;
  return S;
}

// ---------------------------------------------------------------------------
CSR *nsparseMGPU_version0(CSR *A, CSR *P, csrlocinfo * /*Plocal*/, double &loc_time) {
  if (!A || !P || !A->on_the_device || !P->on_the_device)
    throw Error(NSP_ERR_INVALID, "nsparseMGPU_version0: A and P must be on the device");
  if (A->m != P->n) throw Error(NSP_ERR_INVALID, "nsparseMGPU_version0: A->m != P->n");
  nsp_handle_t h = handle();
  nsp_csr_t a = view(A, false), b = view(P, false), c{};
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  const auto t0 = std::chrono::steady_clock::now();
  check(nsp_spgemm(h, NSP_IDX32, NSP_F64, &a, &b, &c, nullptr, nullptr, nullptr), h, "nsparseMGPU_version0");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  loc_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  CSR *C = CSRm::init(A->n, P->m, (stype)c.nnz, false, true, false, A->full_n, A->row_shift);
  C->row = (int *)c.row_ptr;
  C->col = (int *)c.col;
  C->val = (double *)c.val;
  return C;
}

CSR *nsparseMGPU_version0(CSR *A, CSR *P) {
  double t = 0.0;
  return nsparseMGPU_version0(A, P, nullptr, t);
}

CSR *nsparseMGPU_version1(CSR *Alocal, CSR *Pfull, double &loc_time, int p_local_n, CSR **A_out, CSR **P_out) {
  if (!Alocal || !Pfull || !Alocal->on_the_device || !Pfull->on_the_device)
    throw Error(NSP_ERR_INVALID, "nsparseMGPU_version1: matrices must be on the device");
  nsp_handle_t h = handle();
  const int64_t r0 = (int64_t)Alocal->row_shift;
  CSR Ploc;
  Ploc.n = p_local_n < 0 ? Alocal->n : (stype)p_local_n;
  // A's compaction: completedP rows = shrinking_permut(A) + r0
  if (Alocal->col_hi <= Alocal->col_lo) {
    Alocal->col_lo = -r0;
    Alocal->col_hi = (int64_t)Pfull->n - r0;
  }
  CSR *A_ = get_shrinked_matrix(Alocal, &Ploc);
  const int64_t nrows = (int64_t)Alocal->shrinking_permut_len;
  // completedP: gather of P rows perm + r0, columns re-shifted to the rank frame
  nsp_csr_t pf = view(Pfull, false);
  CSR *cP = CSRm::init((stype)nrows, Pfull->m, 0, false, true, false, Pfull->n, (gstype)r0);
  cP->row = dalloc<int>((size_t)nrows + 1);
  int64_t nnz = 0;
  check(nsp_gather_rows_count(h, NSP_IDX32, &pf, nrows, Alocal->shrinking_permut, r0, cP->row, &nnz), h,
        "completedP rows");
  cP->nnz = (stype)nnz;
  cP->col = dalloc<int>((size_t)nnz);
  cP->val = dalloc<double>((size_t)nnz);
  nsp_csr_t out = view(cP, false);
  check(nsp_gather_rows_fill(h, NSP_IDX32, NSP_F64, &pf, nrows, Alocal->shrinking_permut, r0, r0, &out), h,
        "completedP fill");
  cP->col_lo = -r0;
  cP->col_hi = (int64_t)Pfull->m - r0;
  CSR *P_ = get_shrinked_matrix(cP, nullptr);
  CSR *C = nsparseMGPU_version0(A_, P_, nullptr, loc_time);
  if (A_out) {
    *A_out = A_;
  } else {
    CSRm::freeStruct(A_);  // a view of Alocal's arrays
  }
  if (P_out) {
    // P_ owns completedP's arrays from here on
    P_->row = cP->row; P_->val = cP->val; P_->col = cP->shrinked_col;
    P_->shrinking_permut = cP->shrinking_permut;
    P_->shrinking_permut_len = cP->shrinking_permut_len;
    cP->row = nullptr; cP->val = nullptr; cP->shrinked_col = nullptr; cP->shrinking_permut = nullptr;
    *P_out = P_;
  } else {
    CSRm::freeStruct(P_);  // a view of completedP's arrays
  }
  CSRm::free(cP);
  return C;
}

CSR *nsparseMGPU_version1(handles *hs, CSR *Alocal, CSR *Plocal, double &loc_time) {
  if (!hs || !Alocal || !Plocal) throw Error(NSP_ERR_INVALID, "nsparseMGPU_version1: NULL argument");
  if (!hs->Pfull) {
    // one rank: Plocal holds every row of P (spspmpi.cu:385)
    csrlocinfo info{0, (int)Plocal->n, Plocal->row, Plocal->col, Plocal->val};
    return nsparseMGPU_version0(Alocal, Plocal, &info, loc_time);
  }
  if (Plocal->row_shift != Alocal->row_shift)
    throw Error(NSP_ERR_INVALID, "nsparseMGPU_version1: Alocal and Plocal start at different rows");
  if (Plocal->n > (stype)INT32_MAX) throw Error(NSP_ERR_OVERFLOW, "nsparseMGPU_version1: Plocal->n >= 2^31");
  return nsparseMGPU_version1(Alocal, const_cast<CSR *>(hs->Pfull), loc_time, (int)Plocal->n);
}

CSR *nsparseMGPU_version1(handles *hs, CSR *Alocal, CSR *Plocal) {
  double t = 0.0;
  return nsparseMGPU_version1(hs, Alocal, Plocal, t);
}

namespace Handles {
handles *init() {
  handles *h = new handles();
  hip_check(hipStreamCreate(&h->stream1), "hipStreamCreate");
  hip_check(hipStreamCreate(&h->stream2), "hipStreamCreate");
  return h;
}

void free(handles *h) {
  if (!h) return;
  if (h->stream1) (void)hipStreamDestroy(h->stream1);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  delete h;
}
}  // namespace Handles

// ---------------------------------------------------------------------------
CSR *galerkin_product(CSR *R, CSR *A, CSR *P) {
  if (!A || !P || !A->on_the_device || !P->on_the_device || (R && !R->on_the_device))
    throw Error(NSP_ERR_INVALID, "galerkin_product: device matrices expected");
  nsp_handle_t h = handle();
  const nsp_csr_t a = view(A, false), p = view(P, false);
  nsp_csr_t r{}, c{};
  if (R) r = view(R, false);
  check(nsp_galerkin(h, NSP_IDX32, NSP_F64, R ? &r : nullptr, &a, &p, &c, nullptr, nullptr, nullptr), h,
        "galerkin_product");
  if (c.nnz > INT32_MAX) {
    nsp_csr_free(&c);
    throw Error(NSP_ERR_OVERFLOW, "galerkin_product: nnz does not fit CSR's int row pointers");
  }
  CSR *C = CSRm::init((stype)c.rows, (gstype)c.cols, (stype)c.nnz, false, true, false, (gstype)c.rows, 0);
  C->row = (int *)c.row_ptr;
  C->col = c.col;
  C->val = (double *)c.val;
  return C;
}

// ---------------------------------------------------------------------------
namespace {

struct CooRec {
  int64_t i, j;
  double v;
};

nsp_coo_share_t coo_share(const char *path, int nprocs, int rank, int header_type, const char *who) {
  if (nprocs < 1 || rank < 0 || rank >= nprocs) throw Error(NSP_ERR_INVALID, std::string(who) + ": bad rank");
  nsp_coo_share_t s;
  const int st = nsp_coo_share(path, header_type, nprocs, rank, &s);
  if (st != NSP_OK) throw Error(st, std::string(who) + ": cannot read the header or rows of " + path);
  if (s.rec_end - s.rec_begin > INT32_MAX || s.nrows > INT32_MAX)
    throw Error(NSP_ERR_OVERFLOW, std::string(who) + ": the rank's share does not fit int32 CSR");
  return s;
}

}  // namespace

CSR *read_binary_matrix(const char *path, int nprocs, int rank, int header_type, int64_t *nr_out, int64_t *nc_out,
                        int64_t *nt_out) {
  const nsp_coo_share_t s = coo_share(path, nprocs, rank, header_type, "read_binary_matrix");
  const int64_t nrec = s.rec_end - s.rec_begin;
  std::vector<CooRec> rec((size_t)nrec);
  FILE *f = fopen(path, "rb");
  if (!f) throw Error(NSP_ERR_INVALID, std::string("read_binary_matrix: cannot open ") + path);
  const bool ok = rec.empty() || (fseeko(f, (off_t)(s.header_bytes + s.rec_begin * (int64_t)sizeof(CooRec)),
                                         SEEK_SET) == 0 &&
                                  fread(rec.data(), sizeof(CooRec), rec.size(), f) == rec.size());
  fclose(f);
  if (!ok) throw Error(NSP_ERR_INVALID, "read_binary_matrix: truncated records");
  // rows first to last present row (distribute.cu:1094), columns shifted by
  // the rows of the ranks before (distribute.cu:1095, 1103)
  const int n = (int)s.nrows, nnz = (int)nrec;
  CSR *A = CSRm::init((stype)n, (gstype)s.nc, (stype)nnz, true, false, false, (gstype)s.nr, (gstype)s.col_shift);
  std::fill(A->row, A->row + n + 1, 0);
  for (const CooRec &r : rec) {
    const int64_t x = r.i - 1 - s.first_row;
    if (x < 0 || x >= n) {
      CSRm::free(A);
      throw Error(NSP_ERR_INVALID, "read_binary_matrix: records out of row order");
    }
    A->row[x + 1]++;
  }
  for (int i = 0; i < n; i++) A->row[i + 1] += A->row[i];
  for (int q = 0; q < nnz; q++) {
    A->col[q] = (int)(rec[q].j - 1 - s.col_shift);
    A->val[q] = rec[q].v;
  }
  // rows in column order (check_and_fix_order): a stable sort per row
  std::vector<std::pair<int, double>> tmp;
  for (int i = 0; i < n; i++) {
    const int p0 = A->row[i], p1 = A->row[i + 1];
    bool sorted = true;
    for (int p = p0 + 1; p < p1 && sorted; p++) sorted = A->col[p - 1] <= A->col[p];
    if (sorted) continue;
    tmp.clear();
    for (int p = p0; p < p1; p++) tmp.emplace_back(A->col[p], A->val[p]);
    std::stable_sort(tmp.begin(), tmp.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (int p = p0; p < p1; p++) {
      A->col[p] = tmp[p - p0].first;
      A->val[p] = tmp[p - p0].second;
    }
  }
  if (nr_out) *nr_out = s.nr;
  if (nc_out) *nc_out = s.nc;
  if (nt_out) *nt_out = s.nt;
  return A;
}

CSR *load_MatrixMPI(const char *path, int nprocs, int rank, int header_type, bool ascii) {
  if (ascii) {
    nsp_ascii_share_t s{};
    const int rc = nsp_ascii_share(path, nprocs, rank, &s);
    if (rc != NSP_OK) throw Error(rc, std::string("load_MatrixMPI: cannot read the ASCII file ") + path);
    const int n = (int)s.nrows, nnz = (int)(s.rec_end - s.rec_begin);
    CSR *A = CSRm::init((stype)n, (gstype)s.nc, (stype)nnz, true, true, false, (gstype)s.nr, (gstype)s.col_shift);
    nsp_handle_t h = handle();
    nsp_csr_t out{n, s.nc, nnz, A->row, A->col, A->val};
    const int st = nsp_ascii_read(h, NSP_IDX32, NSP_F64, path, &s, &out);
    if (st != NSP_OK) {
      CSRm::free(A);
      check(st, h, "load_MatrixMPI");
    }
    return A;
  }
  const nsp_coo_share_t s = coo_share(path, nprocs, rank, header_type, "load_MatrixMPI");
  const int n = (int)s.nrows, nnz = (int)(s.rec_end - s.rec_begin);
  CSR *A = CSRm::init((stype)n, (gstype)s.nc, (stype)nnz, true, true, false, (gstype)s.nr, (gstype)s.col_shift);
  nsp_handle_t h = handle();
  nsp_csr_t out{n, s.nc, nnz, A->row, A->col, A->val};
  const int st = nsp_coo_read(h, NSP_IDX32, NSP_F64, path, header_type, &s, &out);
  if (st != NSP_OK) {
    CSRm::free(A);
    check(st, h, "load_MatrixMPI");
  }
  return A;
}

void CSRMatrixPrintMM(CSR *A_, const char *file_name) {
  CSR *A = A_->on_the_device ? CSRm::copyToHost(A_) : A_;
  FILE *fp = fopen(file_name, "w");
  if (!fp) {
    if (A != A_) CSRm::free(A);
    throw Error(NSP_ERR_INVALID, std::string("CSRMatrixPrintMM: cannot open ") + file_name);
  }
  // the reference's formats (matrixIO.cu:328-336), 1-based indices
  fprintf(fp, "%s\n", "%%MatrixMarket matrix coordinate real general");
  fprintf(fp, "%d  %d %d \n", (int)A->n, (int)A->m, (int)A->nnz);
  for (int i = 0; i < (int)A->n; i++)
    for (int j = A->row[i]; j < A->row[i + 1]; j++) fprintf(fp, "%d   %d  %lg\n", i + 1, A->col[j] + 1, A->val[j]);
  const bool ok = !ferror(fp);
  fclose(fp);
  if (A != A_) CSRm::free(A);
  if (!ok) throw Error(NSP_ERR_INVALID, std::string("CSRMatrixPrintMM: write failed on ") + file_name);
}

void append_outtime(const char *path, uint64_t tot_n, int64_t m, uint64_t tot_nnz, double loc_time, double tot_time,
                    double l2n) {
  FILE *fid = fopen(path, "a");
  if (!fid) throw Error(NSP_ERR_INVALID, std::string("append_outtime: cannot open ") + path);
  // main.cu:243-247
  fprintf(fid, "%15lu %15lu %15lu ", (unsigned long)tot_n, (unsigned long)m, (unsigned long)tot_nnz);
  fprintf(fid, "%10f %10f %50f\n", loc_time, tot_time, l2n);
  fclose(fid);
}

void write_binary_matrix(const char *path, const CSR *A, int64_t col_shift) {
  if (A->on_the_device) throw Error(NSP_ERR_INVALID, "write_binary_matrix: host matrix expected");
  FILE *f = fopen(path, "wb");
  if (!f) throw Error(NSP_ERR_INVALID, std::string("write_binary_matrix: cannot open ") + path);
  const int64_t h[3] = {(int64_t)A->n, (int64_t)A->m, (int64_t)A->nnz};
  bool ok = fwrite(h, 8, 3, f) == 3;
  std::vector<CooRec> buf;
  buf.reserve(1 << 16);
  for (int i = 0; i < (int)A->n && ok; i++) {
    for (int p = A->row[i]; p < A->row[i + 1]; p++) {
      buf.push_back(CooRec{(int64_t)i + 1, (int64_t)A->col[p] + col_shift + 1, A->val[p]});
      if (buf.size() == buf.capacity()) {
        ok = fwrite(buf.data(), sizeof(CooRec), buf.size(), f) == buf.size();
        buf.clear();
      }
    }
  }
  if (ok && !buf.empty()) ok = fwrite(buf.data(), sizeof(CooRec), buf.size(), f) == buf.size();
  fclose(f);
  if (!ok) throw Error(NSP_ERR_INVALID, std::string("write_binary_matrix: write failed on ") + path);
}

}  // namespace nspamd
