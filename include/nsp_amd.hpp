/* C++ host API of nsp_amd: the reference's types and entry points for the
 * local sparse product, on one device, over the C ABI of nsp_amd.h.
 *
 * Names, fields and argument meaning follow the reference (amavliut/nsp):
 *   sfCSR, nsp_spgemm_kernel_hash     nsp/inc/nsp.h:57-70, 105
 *   csrlocinfo                        src/csrseg.h:3-9
 *   stype, gstype, gsstype, itype     src/setting.h:21-28
 *   CSR (CSRtype<>)                   src/CSR.h:67-104
 *   handles, Handles::init/free       src/handles.h:5-16
 *   CSRm::init/free/copyToDevice/...  src/CSR.cu:28, 572-621, 656-700
 *   shrink_col, get_shrinked_matrix   src/spspmpi.cu:106-182
 *   nsparseMGPU_version0              src/spspmpi.cu:266-313
 *   nsparseMGPU_version1              src/spspmpi.cu:371-975, without MPI:
 *                                     the halo rows of P are gathered from the
 *                                     full P on the same device
 *   read_binary_matrix                src/READ_MATRIX_MPI.cpp:146-356 with the
 *                                     row split and column shift of
 *                                     src/distribute.cu:1071-1125 (host CSR)
 *   load_MatrixMPI                    src/distribute.cu:1071-1125 (device CSR)
 *   write_binary_matrix               main.cu:468-512
 *   CSRMatrixPrintMM                  src/matrixIO.cu:299-339
 *   CSRm::printInfo                   src/CSR.cu:631-643
 *   append_outtime                    main.cu:236-247 (the outtime.txt line)
 *   galerkin_product                  R(AP), named at main.cu:138
 * Errors: where the reference exits (checkCudaErrors, helper_cuda.h:714) these
 * functions throw nspamd::Error carrying the C ABI status.  Calls are
 * host-synchronous on return and ordered on the legacy (null) stream, as in
 * the reference; each host thread uses its own library handle on the current
 * device.  Code written against the reference compiles against this header
 * with `using namespace nspamd;` (CSR, CSRm::, Handles::, stype ... resolve
 * to the declarations below).
 *
 * Link with -lnsp_amd_cpp -lnsp_amd (and the HIP runtime). */
#ifndef NSP_AMD_HPP
#define NSP_AMD_HPP

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include <hip/hip_runtime.h>

#include "nsp_amd.h"

namespace nspamd {

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

/* nsp/inc/nsp.h:57-70.  The product reads M, N, d_rpt, d_col, d_val of a and
 * b and sets M, N, nnz and the device arrays of c. */
typedef struct {
  int *rpt;
  int *col;
  double *val;
  int *d_rpt;
  int *d_col;
  double *d_val;
  unsigned int M;
  unsigned int N;
  int nnz;
  int nnz_max;
  char *matrix_name;
} sfCSR;

/* C = A*B (nsp/src/nsp.cu:2448-2488): sorted rows, structural zeros kept;
 * c->d_rpt, c->d_col, c->d_val are allocated here with hipMalloc (the caller
 * frees them, e.g. with sfCSR_free_device). */
void nsp_spgemm_kernel_hash(sfCSR *a, sfCSR *b, sfCSR *c);
void sfCSR_free_device(sfCSR *c);

/* src/csrseg.h:3-9 (the rank's slice of P; kept for the signature of
 * nsparseMGPU_version0, which does not read it). */
typedef struct {
  int fr;
  int lr;
  int *row;
  int *col;
  double *val;
} csrlocinfo;

/* The reference's index and size types (src/setting.h:21-28): matrix sizes
 * and shifts are unsigned long, column and row-pointer entries int. */
typedef double vtype;
typedef int itype;
typedef unsigned long stype;
typedef unsigned long gstype;
typedef long int gsstype;

/* CSRtype<> of src/CSR.h:67-104, the fields with the reference's names,
 * types and order (the cuSPARSE descriptor, halo_info and rows_to_get of the
 * MPI exchange are not part of a one-device product and are left out).
 * Arrays are device pointers when on_the_device, host pointers otherwise.
 * Columns of a rank's block are in the rank frame (global column -
 * row_shift). */
struct CSR {
  stype nnz = 0;               // entries
  stype n = 0;                 // rows
  gstype m = 0;                // columns
  stype shrinked_m = 0;        // columns after compaction
  gstype full_n = 0;           // rows of the whole matrix
  gstype full_m = 0;           // columns of the whole matrix
  bool on_the_device = false;
  bool is_symmetric = false;
  bool shrinked_flag = false;
  bool custom_alloced = false;
  gsstype col_shifted = 0;     // set by the caller (main.cu:208-209), not read here
  gstype shrinked_firstrow = 0;
  gstype shrinked_lastrow = 0;
  vtype *val = nullptr;
  itype *col = nullptr;
  itype *row = nullptr;        // n+1 row pointers
  itype *shrinked_col = nullptr;  // compacted columns (device)
  gstype row_shift = 0;        // first global row of the block
  int *bitcol = nullptr;
  int bitcolsize = 0;
  int post_local = 0;
  /* Not in the reference's struct: the compaction map of shrink_col,
   * shrinking_permut[c'] = rank-frame column of compacted column c' (the
   * reference keeps it in a file-static buffer, src/spspmpi.cu:136-139),
   * and the bounds [col_lo, col_hi) of the rank-frame columns shrink_col
   * flags (default [-row_shift, m - row_shift)). */
  itype *shrinking_permut = nullptr;
  stype shrinking_permut_len = 0;
  int64_t col_lo = 0, col_hi = -1;
};

namespace CSRm {
/* src/CSR.cu:28 (declared at src/CSR.h:114), same parameters: allocate_mem
 * allocates n+1 row pointers and nnz columns and values on the device or the
 * host; is_symmetric is recorded (printInfo prints it). */
CSR *init(stype n, gstype m, stype nnz, bool allocate_mem, bool on_the_device, bool is_symmetric, gstype full_n,
          gstype row_shift = 0);
void free(CSR *A);
/* frees only the struct, not its arrays (src/CSR.cu:624): for views such as
 * get_shrinked_matrix's result */
void freeStruct(CSR *A);
CSR *copyToDevice(const CSR *A);
CSR *copyToHost(const CSR *A);
/* ||A||_F of a host or device matrix (main.cu:236-241 prints it per rank) */
double norm_fro(const CSR *A);
/* The reference's summary of a matrix (src/CSR.cu:631-643), same lines and
 * formats, SYMMETRIC or GENERAL from is_symmetric. */
void printInfo(const CSR *A, FILE *fp = stdout);
}  // namespace CSRm

/* src/handles.h:5-16: the per-process library state the reference passes to
 * nsparseMGPU_version1 (its streams and cuSPARSE/cuBLAS handles).  Here it
 * holds two streams at the default priority, and optionally the full P on
 * the device: with one device standing in for every rank, the halo rows a
 * rank would receive over MPI are gathered from it.  Pfull is not owned. */
struct handles {
  hipStream_t stream1 = nullptr;
  hipStream_t stream2 = nullptr;
  const CSR *Pfull = nullptr;
};
namespace Handles {
handles *init();
void free(handles *h);
}  // namespace Handles

/* Compacts A's columns once (src/spspmpi.cu:141-162): perm = sorted
 * unique(cols U [0..l]) with l = P->n - 1 when P is given, else A->n - 1;
 * A->shrinked_col, A->shrinked_m and A->shrinking_permut are set.  Returns
 * false when A was already compacted. */
bool shrink_col(CSR *A, CSR *P);
/* A view of A with compacted columns (src/spspmpi.cu:166-182); shares A's
 * row and val arrays. */
CSR *get_shrinked_matrix(CSR *A, CSR *P);

/* C = A*P on the device (src/spspmpi.cu:266-313): C->n = A->n,
 * C->m = P->m, C->full_n = A->full_n, C->row_shift = A->row_shift.
 * loc_time gets the product's time in seconds (host-synchronous around the
 * call, as the reference's timer). */
CSR *nsparseMGPU_version0(CSR *A, CSR *P, csrlocinfo *Plocal, double &loc_time);
CSR *nsparseMGPU_version0(CSR *A, CSR *P);

/* One rank's local product on one device (src/spspmpi.cu:371-975): Alocal is
 * the rank's row block (device, rank-frame columns, row_shift = first row);
 * Pfull holds every row of P (device, global columns).  completedP = P rows
 * shrinking_permut(A) + row_shift in ascending order, columns re-shifted to
 * the rank frame; then A_ and P_ are compacted and C = A_ * P_ (C's columns
 * are P_'s compacted columns).  p_local_n: rows of P the rank owns (default
 * Alocal->n).  A_out / P_out, when given, receive A_ and P_: A_ is a view of
 * Alocal's arrays (release it with CSRm::freeStruct), P_ owns its arrays
 * (CSRm::free). */
CSR *nsparseMGPU_version1(CSR *Alocal, CSR *Pfull, double &loc_time, int p_local_n = -1, CSR **A_out = nullptr,
                          CSR **P_out = nullptr);
/* src/spspmpi.cu:371 (called at main.cu:223), same parameters.  Without
 * h->Pfull, Plocal holds every row of P (one rank) and the product is
 * nsparseMGPU_version0(Alocal, Plocal), as the reference does when nprocs
 * is 1 (spspmpi.cu:385).  With h->Pfull, Plocal is the rank's own block of
 * P (its rows give the local row count, spspmpi.cu:400-404) and the rank's
 * share is the overload above with Pfull = h->Pfull. */
CSR *nsparseMGPU_version1(handles *h, CSR *Alocal, CSR *Plocal, double &loc_time);
CSR *nsparseMGPU_version1(handles *h, CSR *Alocal, CSR *Plocal);

/* Galerkin product C = R * (A * P) on the device, R = P^T when R is NULL
 * (nsp_galerkin): the coarse operator the reference's driver names as its
 * purpose (main.cu:138, "il prodotto e' R(AP)").  C is a device CSR
 * (CSRm::free). */
CSR *galerkin_product(CSR *R, CSR *A, CSR *P);

/* Rank `rank` of `nprocs` of a binary COO matrix file (header {nr, nc, nt}
 * when header_type = 1, {nr, nt} when 0, then nt records {int64 i, int64 j,
 * double v}, 1-based, sorted by row): host CSR whose n rows run from the
 * first to the last row present in the rank's range, columns shifted by
 * row_shift = the rows of the ranks before (the reference's MPI_Exscan of
 * rowxproc; the block's first row unless empty rows sit at partition ends),
 * rows sorted by column.  *nr_out, *nc_out and *nt_out get the file's sizes. */
CSR *read_binary_matrix(const char *path, int nprocs, int rank, int header_type = 1, int64_t *nr_out = nullptr,
                        int64_t *nc_out = nullptr, int64_t *nt_out = nullptr);
/* The same block as a device CSR built on the device (nsp_coo_read: records
 * stream from the file through pinned buffers and are converted there), as
 * the reference's load_MatrixMPI returns it (src/distribute.cu:1071-1125,
 * row partition).  With `ascii` the file is the reference's text format
 * (READ_ASCII_MATRIX_MPI, src/READ_MATRIX_MPI.cpp:364-578, taken when
 * load_MatrixMPI's `ascii` is set, distribute.cu:1084-1086), parsed on the
 * device by nsp_ascii_read; header_type is then unused (the text header
 * carries its own field count). */
CSR *load_MatrixMPI(const char *path, int nprocs, int rank, int header_type = 1, bool ascii = false);
/* Writes a host CSR as binary COO (header_type 1); global column =
 * col + col_shift.  The layout the reader takes: the reference's
 * write_binary_matrix template (main.cu:468-512) with otype = long. */
void write_binary_matrix(const char *path, const CSR *A, int64_t col_shift = 0);
/* MatrixMarket coordinate file of A (host or device) with the reference's
 * header and line formats (src/matrixIO.cu:299-339). */
void CSRMatrixPrintMM(CSR *A, const char *file_name);
/* Appends the reference's timing line (main.cu:243-247: rows, columns and
 * nnz of the product, local and total time, ||C||_F) to path. */
void append_outtime(const char *path, uint64_t tot_n, int64_t m, uint64_t tot_nnz, double loc_time, double tot_time,
                    double l2n);

}  // namespace nspamd

#endif /* NSP_AMD_HPP */
