// The reference driver's call sequence (main.cu:158-250) written in the
// reference's own spelling against include/nsp_amd.hpp, for one rank of an
// nprocs-way row split on one device:
//
//   main_sequence A.bin P.bin nprocs rank C.bin [printinfo]
//
// The matrices are loaded as load_MatrixMPI builds them
// (src/distribute.cu:1071-1125: the rank's block read from the binary COO
// file, CSRm::init(rowxproc, nc, ntrank, true, false, false, nr, col_shift),
// arrays filled, copyToDevice, free of the host copy), the column shifts set
// as main.cu:208-209 sets them, the local product taken with
// nsparseMGPU_version1(h, Alocal, Plocal, loc_time) (main.cu:223), C copied
// to the host, and nnz and n read as the unsigned long fields MPI_Reduce
// reads (main.cu:236-237).  With one device standing in for every rank, the
// halo rows of P come from the full P registered in the handles (h->Pfull);
// with nprocs = 1 nothing is registered and Plocal is all of P.  C is written
// to C.bin (binary COO, columns as the product returns them) and a JSON line
// with the sizes and ||C||_F ends the output.  Built with -Wall -Werror.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "nsp_amd.hpp"

using namespace nspamd;

static_assert(std::is_same<decltype(CSR::nnz), unsigned long>::value, "CSR::nnz is stype (unsigned long)");
static_assert(std::is_same<decltype(CSR::n), unsigned long>::value, "CSR::n is stype (unsigned long)");
static_assert(std::is_same<decltype(CSR::m), unsigned long>::value, "CSR::m is gstype (unsigned long)");
static_assert(std::is_same<decltype(CSR::col_shifted), long>::value, "CSR::col_shifted is gsstype (long)");

// distribute.cu:1071-1125, with the reader of this library standing in for
// READ_BINARY_MATRIX_MPI + irow2iat2 + check_and_fix_order + apply_shift.
static CSR *load_rank(const char *path, int nprocs, int rank) {
  int64_t nr = 0, nc = 0, nt = 0;
  CSR *R = read_binary_matrix(path, nprocs, rank, 1, &nr, &nc, &nt);
  const long int rowxproc = (long int)R->n, ntrank = (long int)R->nnz;
  const long int col_shift = (long int)R->row_shift;
  CSR *Alocal = CSRm::init((stype)rowxproc, (gstype)nc, (stype)ntrank, true, false, false, nr, (long int)col_shift);
  for (unsigned long j = 0; j <= (unsigned long)rowxproc; j++) Alocal->row[j] = (itype)R->row[j];
  for (unsigned long k = 0; k < (unsigned long)ntrank; k++) {
    Alocal->col[k] = (itype)R->col[k];
    Alocal->val[k] = R->val[k];
  }
  CSRm::free(R);
  CSR *d_Alocal = CSRm::copyToDevice(Alocal);
  CSRm::free(Alocal);
  return d_Alocal;
}

int main(int argc, char *argv[]) {
  if (argc < 6) {
    fprintf(stderr, "usage: %s A.bin P.bin nprocs rank C.bin [printinfo]\n", argv[0]);
    return 2;
  }
  const int nprocs = atoi(argv[3]), myid = atoi(argv[4]);
  const int printinfo = argc > 6 ? atoi(argv[6]) : 0;
  try {
    handles *h = Handles::init();

    CSR *Alocal = load_rank(argv[1], nprocs, myid);
    CSR *Plocal = load_rank(argv[2], nprocs, myid);
    CSR *Pfull = nullptr;
    if (nprocs > 1) {
      Pfull = load_rank(argv[2], 1, 0);
      h->Pfull = Pfull;
    }

    Alocal->col_shifted = -Alocal->row_shift;
    Plocal->col_shifted = -Plocal->row_shift;

    if (printinfo) {
      CSRm::printInfo(Alocal);
      CSRm::printInfo(Plocal);
    }

    double TOT_TIMEM, loc_time;
    stype tot_nnz, tot_n;
    double l2n_loc = 0., l2n = 0.;

    (void)hipDeviceSynchronize();
    const auto t0 = std::chrono::steady_clock::now();
    CSR *APlocal = nsparseMGPU_version1(h, Alocal, Plocal, loc_time);
    (void)hipDeviceSynchronize();
    TOT_TIMEM = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (APlocal->on_the_device) {
      CSR *h_APlocal = CSRm::copyToHost(APlocal);
      CSRm::free(APlocal);
      APlocal = h_APlocal;
    }

    // MPI_Reduce(&(APlocal->nnz), &tot_nnz, 1, MPI_UNSIGNED_LONG, ...) over
    // one process: the 8 bytes of each field as MPI reads them
    std::memcpy(&tot_nnz, &(APlocal->nnz), sizeof(unsigned long));
    std::memcpy(&tot_n, &(APlocal->n), sizeof(unsigned long));
    for (long int i = 0; i < (long int)APlocal->nnz; i++) l2n_loc += APlocal->val[i] * APlocal->val[i];
    l2n = sqrt(l2n_loc);
    printf("TOTAL_TIME: %f\t\tlocal time: %f\t\tnnz = %lu\t\tl2norm: %f\n", TOT_TIMEM, loc_time, tot_nnz, l2n);

    write_binary_matrix(argv[5], APlocal, 0);
    printf("{\"nprocs\": %d, \"rank\": %d, \"tot_n\": %lu, \"m\": %lu, \"tot_nnz\": %lu, \"row_shift\": %lu, "
           "\"full_n\": %lu, \"l2norm\": %.17g}\n",
           nprocs, myid, tot_n, (gstype)APlocal->m, tot_nnz, APlocal->row_shift, APlocal->full_n, l2n);

    CSRm::free(APlocal);
    CSRm::free(Alocal);
    CSRm::free(Plocal);
    if (Pfull) CSRm::free(Pfull);
    Handles::free(h);
  } catch (const Error &e) {
    fprintf(stderr, "main_sequence: %s (status %d)\n", e.what(), e.status);
    return 1;
  }
  return 0;
}
