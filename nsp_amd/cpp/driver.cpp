// nsp_amd_run: command-line driver of the C++ API, following the flow of the
// reference's main.cu (read -> split -> local product -> report nnz and
// ||C||_F, main.cu:223-241) on one device.
//
//   nsp_amd_run product A.bin B.bin [C.bin]
//       C = A*B through nsp_spgemm_kernel_hash (sfCSR interface)
//   nsp_amd_run local A.bin P.bin nprocs rank [C.bin|-] [outtime.txt]
//       rank `rank` of `nprocs`: its row block of A times the P rows it needs
//       (nsparseMGPU_version1 without MPI: P is read whole onto the device);
//       prints main.cu's TOTAL_TIME line and appends its outtime.txt line;
//       NSP_AMD_PRINTINFO=1 prints both matrices' summaries first
//       (CSRm::printInfo, main.cu's printinfo option)
//   nsp_amd_run split A.bin nprocs rank out.bin
//       the reader/splitter alone (host only): the rank's block, columns in
//       the rank frame, written back as binary COO
//   nsp_amd_run load A.bin nprocs rank out.bin
//       the same block built on the device (load_MatrixMPI), copied back
//       and written
//   nsp_amd_run ascii A.txt nprocs rank out.bin
//       the same from an ASCII COO file (load_MatrixMPI with ascii set, as
//       main.cu does for its ascii argument)
//   nsp_amd_run galerkin A.bin P.bin [C.bin]
//       the coarse operator P^T (A P) (galerkin_product)
//   nsp_amd_run mm A.bin out.mtx
//       A as a MatrixMarket file (CSRMatrixPrintMM; host only)
//
// Matrices are binary COO files (header {nr, nc, nt}, then {i, j, v}
// records, 1-based, sorted by row).  Prints one JSON line; C.bin, when given,
// receives C in the same format (columns as the product returns them).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nsp_amd.hpp"

using namespace nspamd;

static void write_device_csr(const char *path, CSR *C) {
  CSR *H = CSRm::copyToHost(C);
  write_binary_matrix(path, H, 0);
  CSRm::free(H);
}

static int run_product(int argc, char **argv) {
  if (argc < 4) return 2;
  CSR *A = read_binary_matrix(argv[2], 1, 0);
  CSR *B = read_binary_matrix(argv[3], 1, 0);
  CSR *Ad = CSRm::copyToDevice(A), *Bd = CSRm::copyToDevice(B);
  sfCSR a{}, b{}, c{};
  a.M = (unsigned)Ad->n;
  a.N = (unsigned)Ad->m;
  a.nnz = (int)Ad->nnz;
  a.d_rpt = Ad->row;
  a.d_col = Ad->col;
  a.d_val = Ad->val;
  b.M = (unsigned)Bd->n;
  b.N = (unsigned)Bd->m;
  b.nnz = (int)Bd->nnz;
  b.d_rpt = Bd->row;
  b.d_col = Bd->col;
  b.d_val = Bd->val;
  const auto t0 = std::chrono::steady_clock::now();
  nsp_spgemm_kernel_hash(&a, &b, &c);
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  CSR Cv;
  Cv.n = c.M;
  Cv.m = c.N;
  Cv.nnz = (stype)c.nnz;
  Cv.row = c.d_rpt;
  Cv.col = c.d_col;
  Cv.val = c.d_val;
  Cv.on_the_device = true;
  const double l2 = CSRm::norm_fro(&Cv);
  if (argc > 4) write_device_csr(argv[4], &Cv);
  printf("{\"mode\": \"product\", \"rows\": %u, \"cols\": %u, \"nnz\": %d, \"l2norm\": %.17g, \"time_s\": %.6f}\n",
         c.M, c.N, c.nnz, l2, t);
  sfCSR_free_device(&c);
  CSRm::free(Ad);
  CSRm::free(Bd);
  CSRm::free(A);
  CSRm::free(B);
  return 0;
}

static int run_local(int argc, char **argv) {
  if (argc < 6) return 2;
  const int nprocs = atoi(argv[4]), rank = atoi(argv[5]);
  // main.cu:164-165: both matrices through load_MatrixMPI (device CSR); P
  // whole, since its halo rows are gathered on this device
  const auto t0 = std::chrono::steady_clock::now();
  CSR *Ad = load_MatrixMPI(argv[2], nprocs, rank);
  CSR *Pd = load_MatrixMPI(argv[3], 1, 0);
  const double read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (getenv("NSP_AMD_PRINTINFO")) {  // main.cu's printinfo option
    CSRm::printInfo(Ad);
    CSRm::printInfo(Pd);
  }
  nsp_coo_share_t pshare;
  if (nsp_coo_share(argv[3], 1, nprocs, rank, &pshare) != NSP_OK) throw Error(NSP_ERR_INVALID, "P share");
  double loc_time = 0.0;
  CSR *A_ = nullptr, *P_ = nullptr;
  CSR *C = nsparseMGPU_version1(Ad, Pd, loc_time, (int)pshare.nrows, &A_, &P_);
  const double l2 = CSRm::norm_fro(C);
  if (argc > 6 && strcmp(argv[6], "-") != 0) write_device_csr(argv[6], C);
  // main.cu:241-247: the summary line and, when asked, the outtime.txt line
  // (one rank: loc_time is also the total)
  printf("TOTAL_TIME: %f\t\tlocal time: %f\t\tnnz = %lu\t\tl2norm: %f\n", loc_time, loc_time, (unsigned long)C->nnz, l2);
  if (argc > 7) append_outtime(argv[7], (uint64_t)C->n, (int64_t)C->m, (uint64_t)C->nnz, loc_time, loc_time, l2);
  printf("{\"mode\": \"local\", \"nprocs\": %d, \"rank\": %d, \"row_shift\": %lu, \"rows\": %lu, "
         "\"shrinked_m_A\": %lu, \"completedP_rows\": %lu, \"shrinked_m_P\": %lu, \"nnz\": %lu, "
         "\"l2norm\": %.17g, \"loc_time\": %.6f, \"read_s\": %.6f}\n",
         nprocs, rank, Ad->row_shift, C->n, (unsigned long)A_->m, P_->n, (unsigned long)P_->m, C->nnz, l2, loc_time,
         read_s);
  CSRm::free(C);
  CSRm::free(P_);
  CSRm::freeStruct(A_);
  CSRm::free(Ad);
  CSRm::free(Pd);
  return 0;
}

static int run_mm(int argc, char **argv) {
  if (argc < 4) return 2;
  CSR *A = read_binary_matrix(argv[2], 1, 0);
  CSRMatrixPrintMM(A, argv[3]);
  printf("{\"mode\": \"mm\", \"rows\": %lu, \"nnz\": %lu}\n", A->n, A->nnz);
  CSRm::free(A);
  return 0;
}

static int run_galerkin(int argc, char **argv) {
  if (argc < 4) return 2;
  CSR *A = load_MatrixMPI(argv[2], 1, 0);
  CSR *P = load_MatrixMPI(argv[3], 1, 0);
  const auto t0 = std::chrono::steady_clock::now();
  CSR *C = galerkin_product(nullptr, A, P);
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const double l2 = CSRm::norm_fro(C);
  if (argc > 4) write_device_csr(argv[4], C);
  printf("{\"mode\": \"galerkin\", \"rows\": %lu, \"cols\": %lu, \"nnz\": %lu, \"l2norm\": %.17g, "
         "\"time_s\": %.6f}\n",
         C->n, C->m, C->nnz, l2, t);
  CSRm::free(C);
  CSRm::free(A);
  CSRm::free(P);
  return 0;
}

static int run_split(int argc, char **argv, bool device, bool ascii = false) {
  if (argc < 6) return 2;
  const int nprocs = atoi(argv[3]), rank = atoi(argv[4]);
  int64_t nr = 0, nc = 0, nt = 0;
  const auto t0 = std::chrono::steady_clock::now();
  CSR *A = nullptr;
  if (device) {
    CSR *Ad = load_MatrixMPI(argv[2], nprocs, rank, 1, ascii);
    const double td = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    A = CSRm::copyToHost(Ad);
    CSRm::free(Ad);
    if (ascii) {
      nsp_ascii_share_t s;
      if (nsp_ascii_share(argv[2], nprocs, rank, &s) != NSP_OK) throw Error(NSP_ERR_INVALID, "share");
      nr = s.nr;
      nc = s.nc;
      nt = s.nt;
    } else {
      nsp_coo_share_t s;
      if (nsp_coo_share(argv[2], 1, nprocs, rank, &s) != NSP_OK) throw Error(NSP_ERR_INVALID, "share");
      nr = s.nr;
      nc = s.nc;
      nt = s.nt;
    }
    write_binary_matrix(argv[5], A, 0);
    printf("{\"mode\": \"%s\", \"nr\": %lld, \"nc\": %lld, \"nt\": %lld, \"rows\": %lu, \"nnz\": %lu, "
           "\"row_shift\": %lu, \"read_s\": %.6f}\n",
           ascii ? "ascii" : "load", (long long)nr, (long long)nc, (long long)nt, A->n, A->nnz, A->row_shift, td);
    CSRm::free(A);
    return 0;
  }
  A = read_binary_matrix(argv[2], nprocs, rank, 1, &nr, &nc, &nt);
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  write_binary_matrix(argv[5], A, 0);
  printf("{\"mode\": \"split\", \"nr\": %lld, \"nc\": %lld, \"nt\": %lld, \"rows\": %lu, \"nnz\": %lu, "
         "\"row_shift\": %lu, \"read_s\": %.6f}\n",
         (long long)nr, (long long)nc, (long long)nt, A->n, A->nnz, A->row_shift, t);
  CSRm::free(A);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr,
            "usage: %s product A.bin B.bin [C.bin] | local A.bin P.bin nprocs rank [C.bin] | "
            "split|load A.bin nprocs rank out.bin | ascii A.txt nprocs rank out.bin | galerkin A.bin P.bin [C.bin] | "
            "mm A.bin out.mtx\n",
            argv[0]);
    return 2;
  }
  try {
    int rc = 2;
    if (!strcmp(argv[1], "product")) rc = run_product(argc, argv);
    else if (!strcmp(argv[1], "local")) rc = run_local(argc, argv);
    else if (!strcmp(argv[1], "split")) rc = run_split(argc, argv, false);
    else if (!strcmp(argv[1], "load")) rc = run_split(argc, argv, true);
    else if (!strcmp(argv[1], "ascii")) rc = run_split(argc, argv, true, true);
    else if (!strcmp(argv[1], "galerkin")) rc = run_galerkin(argc, argv);
    else if (!strcmp(argv[1], "mm")) rc = run_mm(argc, argv);
    if (rc == 2) fprintf(stderr, "bad arguments\n");
    return rc;
  } catch (const Error &e) {
    fprintf(stderr, "nsp_amd_run: %s (status %d)\n", e.what(), e.status);
    return 1;
  }
}
