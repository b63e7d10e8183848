/*
 * oracle.c -- CPU restatement of nsp's local sparse product C = A*B and of the
 * steps around it.  TEST INFRASTRUCTURE ONLY: the HIP library never links,
 * loads or calls this file; it is the checker that the tests, the smoke test
 * and the CPU-baseline leg of the benchmark compare the HIP path against.
 *
 * Semantics restated (citations are to the reference tree):
 *   - intermediate-product count per row        nsp/src/nsp.cu:306-320
 *   - C structure = union of the B rows selected by A's row, structural zeros
 *     kept (the hash insert never tests the value)  nsp/src/nsp.cu:24-90
 *   - C rows sorted ascending by column           nsp/src/nsp.cu:1389-1399,
 *                                                 nsp_calc_val_sort_rows.cu:510-527
 *   - row_ptr = exclusive scan of row nnz          nsp/src/nsp.cu:937,1124-1126
 *   - column compaction ("shrink")                 src/getmct.cu:135-220,
 *                                                 src/spspmpi.cu:58-182
 *   - completedP row order and column re-shift     src/spspmpi.cu:584-906
 *   - binary COO reader / row splitter / CSR build src/READ_MATRIX_MPI.cpp:146-356,
 *                                                 src/distribute.cu:998-1125
 * Values are accumulated in FP64 with Neumaier-compensated sums, in the order
 * the A row and B rows list them; the reference's order is nondeterministic
 * (LDS atomics, nsp/src/nsp.cu:89), so values are compared with a tolerance
 * relative to sum_k |a_ik b_kj|, which this file also returns.
 *
 * Index conventions: row pointers int64, column indices int32, values double.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* intprod[i] = sum_{k in A_i} nnz(B_k)   (nsp/src/nsp.cu:306-320). Returns the total. */
int64_t orc_intprod(int64_t M, const int64_t *arpt, const int32_t *acol,
                    const int64_t *brpt, int64_t *intprod) {
    int64_t total = 0;
    #pragma omp parallel for reduction(+:total) schedule(static)
    for (int64_t i = 0; i < M; i++) {
        int64_t s = 0;
        for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
            int32_t k = acol[p];
            s += brpt[k + 1] - brpt[k];
        }
        intprod[i] = s;
        total += s;
    }
    return total;
}

/* out[0] = 0, out[i+1] = out[i] + in[i]   (C.rpt from row nnz, nsp/src/nsp.cu:1124). */
int64_t orc_exscan(int64_t n, const int64_t *in, int64_t *out) {
    int64_t s = 0;
    out[0] = 0;
    for (int64_t i = 0; i < n; i++) { s += in[i]; out[i + 1] = s; }
    return s;
}

/* Symbolic pass: number of distinct columns in each C row (union of B rows).
 * Marker array indexed by column, one per thread. */
int64_t orc_symbolic(int64_t M, const int64_t *arpt, const int32_t *acol,
                     const int64_t *brpt, const int32_t *bcol, int64_t ncolsB,
                     int64_t *row_nz) {
    int64_t total = 0;
    #pragma omp parallel reduction(+:total)
    {
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncolsB > 0 ? ncolsB : 1));
        for (int64_t j = 0; j < ncolsB; j++) mark[j] = -1;
        #pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < M; i++) {
            int64_t nz = 0;
            for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
                int32_t k = acol[p];
                for (int64_t q = brpt[k]; q < brpt[k + 1]; q++) {
                    int32_t j = bcol[q];
                    if (mark[j] != i) { mark[j] = i; nz++; }
                }
            }
            row_nz[i] = nz;
            total += nz;
        }
        free(mark);
    }
    return total;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* One C row by Gustavson with a dense accumulator; writes sorted columns,
 * compensated values and sum|a*b| starting at out offset. Returns nnz. */
static int64_t gustavson_row(int64_t i, const int64_t *arpt, const int32_t *acol, const double *aval,
                             const int64_t *brpt, const int32_t *bcol, const double *bval,
                             int64_t *mark, double *sum, double *comp, double *asum, int32_t *list,
                             int32_t *ccol, double *cval, double *cabs) {
    int64_t nz = 0;
    for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
        int32_t k = acol[p];
        double a = aval[p];
        for (int64_t q = brpt[k]; q < brpt[k + 1]; q++) {
            int32_t j = bcol[q];
            double prod = a * bval[q];
            if (mark[j] != i) {
                mark[j] = i; sum[j] = prod; comp[j] = 0.0; asum[j] = fabs(prod);
                list[nz++] = j;
            } else {
                /* Neumaier summation */
                double s = sum[j], t = s + prod;
                if (fabs(s) >= fabs(prod)) comp[j] += (s - t) + prod;
                else comp[j] += (prod - t) + s;
                sum[j] = t;
                asum[j] += fabs(prod);
            }
        }
    }
    qsort(list, (size_t)nz, sizeof(int32_t), cmp_i32);
    for (int64_t e = 0; e < nz; e++) {
        int32_t j = list[e];
        if (ccol) ccol[e] = j;
        if (cval) cval[e] = sum[j] + comp[j];
        if (cabs) cabs[e] = asum[j];
    }
    return nz;
}

/* Numeric pass over all rows: crpt must already hold the exclusive scan of the
 * symbolic counts. ccol/cval/cabs may be NULL to skip an output. */
void orc_spgemm(int64_t M, const int64_t *arpt, const int32_t *acol, const double *aval,
                const int64_t *brpt, const int32_t *bcol, const double *bval, int64_t ncolsB,
                const int64_t *crpt, int32_t *ccol, double *cval, double *cabs) {
    #pragma omp parallel
    {
        size_t n = (size_t)(ncolsB > 0 ? ncolsB : 1);
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * n);
        double *sum = (double *)malloc(sizeof(double) * n);
        double *comp = (double *)malloc(sizeof(double) * n);
        double *asum = (double *)malloc(sizeof(double) * n);
        int32_t *list = (int32_t *)malloc(sizeof(int32_t) * n);
        for (size_t j = 0; j < n; j++) mark[j] = -1;
        #pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < M; i++) {
            int64_t o = crpt[i];
            gustavson_row(i, arpt, acol, aval, brpt, bcol, bval, mark, sum, comp, asum, list,
                          ccol ? ccol + o : NULL, cval ? cval + o : NULL, cabs ? cabs + o : NULL);
        }
        free(mark); free(sum); free(comp); free(asum); free(list);
    }
}

/* The same product restricted to a list of rows (for matrices whose C does not
 * fit host memory). Output rows are packed in list order; out_rpt has
 * nrows+1 entries and must be sized from orc_symbolic_rows. */
void orc_symbolic_rows(int64_t nrows, const int64_t *rows, const int64_t *arpt, const int32_t *acol,
                       const int64_t *brpt, const int32_t *bcol, int64_t ncolsB, int64_t *row_nz) {
    #pragma omp parallel
    {
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncolsB > 0 ? ncolsB : 1));
        for (int64_t j = 0; j < ncolsB; j++) mark[j] = -1;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrows; r++) {
            int64_t i = rows[r], nz = 0;
            for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
                int32_t k = acol[p];
                for (int64_t q = brpt[k]; q < brpt[k + 1]; q++) {
                    int32_t j = bcol[q];
                    if (mark[j] != r) { mark[j] = r; nz++; }
                }
            }
            row_nz[r] = nz;
        }
        free(mark);
    }
}

void orc_spgemm_rows(int64_t nrows, const int64_t *rows, const int64_t *arpt, const int32_t *acol,
                     const double *aval, const int64_t *brpt, const int32_t *bcol, const double *bval,
                     int64_t ncolsB, const int64_t *out_rpt, int32_t *ccol, double *cval, double *cabs) {
    #pragma omp parallel
    {
        size_t n = (size_t)(ncolsB > 0 ? ncolsB : 1);
        int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * n);
        double *sum = (double *)malloc(sizeof(double) * n);
        double *comp = (double *)malloc(sizeof(double) * n);
        double *asum = (double *)malloc(sizeof(double) * n);
        int32_t *list = (int32_t *)malloc(sizeof(int32_t) * n);
        for (size_t j = 0; j < n; j++) mark[j] = -1;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t r = 0; r < nrows; r++) {
            int64_t o = out_rpt[r];
            /* gustavson_row keys the marker by row id; row ids in a list are
             * unique, so the global row id is a valid marker. */
            gustavson_row(rows[r], arpt, acol, aval, brpt, bcol, bval, mark, sum, comp, asum, list,
                          ccol ? ccol + o : NULL, cval ? cval + o : NULL, cabs ? cabs + o : NULL);
        }
        free(mark); free(sum); free(comp); free(asum); free(list);
    }
}

/* ------------------------------------------------------------------------- */
/* Column compaction (shrink).
 * The reference builds shrinking_permut as
 *   first rank :  [0..l] ++ uv
 *   last rank  :  uv ++ [0..l]
 *   middle rank:  [uv < 0] ++ [0..l] ++ [uv > l]
 * where uv = sorted unique columns outside [0, l] (src/getmct.cu:135-220), and
 * maps every column to its position by binary search
 * (src/spspmpi.cu:58-84).  For inputs the reference accepts, every branch
 * equals sorted(unique(cols U [0..l])); this restatement computes that set by
 * sorting, independently of the branch form (the tests check both forms).
 * perm must hold (l+1)+nnz entries.  Returns shrinked_m = |perm|. */
int64_t orc_shrink(int64_t nnz, const int32_t *col, int64_t l, int32_t *perm, int32_t *newcol) {
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    int64_t nu = 0;
    for (int64_t e = 0; e < nnz; e++)
        if (col[e] < 0 || col[e] > l) tmp[nu++] = col[e];
    qsort(tmp, (size_t)nu, sizeof(int32_t), cmp_i32);
    int64_t u = 0;
    for (int64_t e = 0; e < nu; e++)
        if (u == 0 || tmp[e] != tmp[u - 1]) tmp[u++] = tmp[e];
    /* merge: negatives, then 0..l, then > l */
    int64_t m = 0, e = 0;
    while (e < u && tmp[e] < 0) perm[m++] = tmp[e++];
    for (int64_t c = 0; c <= l; c++) perm[m++] = (int32_t)c;
    while (e < u) perm[m++] = tmp[e++];
    free(tmp);
    for (int64_t q = 0; q < nnz; q++) {
        int32_t c = col[q];
        int64_t lo = 0, hi = m - 1, pos = -1;
        while (lo <= hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (perm[mid] == c) { pos = mid; break; }
            if (c < perm[mid]) hi = mid - 1; else lo = mid + 1;
        }
        newcol[q] = (int32_t)pos;
    }
    return m;
}

/* Rows of B needed by one rank, in completedP order.  A_block columns are in
 * the rank frame (global - r0), l = p_local_n - 1.  The halo rows are the
 * sorted unique columns outside [0, l]; completedP lists the rows below the
 * local block, then the local block, then the rows above, all ascending
 * (src/spspmpi.cu:395-906, _completedP_rows2 at 37-55).  That order is exactly
 * sorted(unique(cols U [0..l])) + r0 in global row ids, i.e. the A shrink
 * permutation shifted by r0, so this returns orc_shrink's perm shifted.
 * grows must hold (l+1)+nnz entries; returns completedP.n. */
int64_t orc_completed_rows(int64_t nnz, const int32_t *acol_rank, int64_t l, int64_t r0,
                           int64_t *grows, int32_t *newcol) {
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(l + 1 + nnz));
    int64_t m = orc_shrink(nnz, acol_rank, l, perm, newcol);
    for (int64_t q = 0; q < m; q++) grows[q] = (int64_t)perm[q] + r0;
    free(perm);
    return m;
}

/* Gather B rows grows[0..nr) into a new CSR, shifting columns by -shift
 * (completedP assembly and the column re-shift, src/spspmpi.cu:822-828,
 * 845-885).  out_rpt must have nr+1 entries; call with out_col == NULL first
 * to size the output (returns nnz). */
int64_t orc_gather_rows(int64_t nr, const int64_t *grows, const int64_t *brpt, const int32_t *bcol,
                        const double *bval, int64_t shift, int64_t *out_rpt, int32_t *out_col,
                        double *out_val) {
    int64_t s = 0;
    out_rpt[0] = 0;
    for (int64_t r = 0; r < nr; r++) {
        int64_t g = grows[r];
        int64_t len = brpt[g + 1] - brpt[g];
        if (out_col) {
            for (int64_t q = 0; q < len; q++) {
                out_col[s + q] = (int32_t)(bcol[brpt[g] + q] - shift);
                out_val[s + q] = bval[brpt[g] + q];
            }
        }
        s += len;
        out_rpt[r + 1] = s;
    }
    return s;
}

/* ------------------------------------------------------------------------- */
/* Transpose (CSR -> CSR of the transpose, rows of the result sorted by the
 * original row index).  The reference's only transpose is cuSPARSE csr2csc
 * (src/CSR.cu:739-797); the backward pair uses it as the cached A^T. */
void orc_transpose(int64_t M, int64_t N, const int64_t *rpt, const int32_t *col, const double *val,
                   int64_t *trpt, int32_t *tcol, double *tval) {
    int64_t nnz = rpt[M];
    int64_t *cnt = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    for (int64_t e = 0; e < nnz; e++) cnt[col[e] + 1]++;
    for (int64_t j = 0; j < N; j++) cnt[j + 1] += cnt[j];
    memcpy(trpt, cnt, sizeof(int64_t) * (size_t)(N + 1));
    for (int64_t i = 0; i < M; i++)
        for (int64_t p = rpt[i]; p < rpt[i + 1]; p++) {
            int64_t d = cnt[col[p]]++;
            tcol[d] = (int32_t)i;
            if (tval) tval[d] = val[p];
        }
    free(cnt);
}

/* Backward of C = A*B on fixed patterns, for a loss L(C) with dL/dC = dC given
 * on C's pattern (crpt/ccol sorted):
 *   dA = (dC * B^T) sampled on pat(A):  dA_ik = sum_{j in B_k} dC_ij B_kj
 *   dB = (A^T * dC) sampled on pat(B):  dB_kj = sum_{i : k in A_i} A_ik dC_ij
 * Every (i,j) reached by a product lies in pat(C), so both are sums over the
 * forward's intermediate products.  Compensated sums; dB is accumulated in a
 * fixed (row-major over A) order. */
static int64_t find_col(const int32_t *col, int64_t lo, int64_t hi, int32_t j) {
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (col[mid] < j) lo = mid + 1; else hi = mid;
    }
    return lo;
}

void orc_backward(int64_t M, const int64_t *arpt, const int32_t *acol, const double *aval,
                  int64_t MB, const int64_t *brpt, const int32_t *bcol, const double *bval,
                  const int64_t *crpt, const int32_t *ccol, const double *dc,
                  double *da, double *db) {
    int64_t nnzA = arpt[M], nnzB = brpt[MB];
    double *dbc = (double *)calloc((size_t)(nnzB > 0 ? nnzB : 1), sizeof(double));
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < M; i++) {
        for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
            int32_t k = acol[p];
            double s = 0.0, c = 0.0;
            for (int64_t q = brpt[k]; q < brpt[k + 1]; q++) {
                int64_t pos = find_col(ccol, crpt[i], crpt[i + 1], bcol[q]);
                double t = dc[pos] * bval[q], u = s + t;
                if (fabs(s) >= fabs(t)) c += (s - u) + t; else c += (t - u) + s;
                s = u;
            }
            da[p] = s + c;
        }
    }
    for (int64_t e = 0; e < nnzB; e++) db[e] = 0.0;
    for (int64_t i = 0; i < M; i++) {
        for (int64_t p = arpt[i]; p < arpt[i + 1]; p++) {
            int32_t k = acol[p];
            double a = aval[p];
            for (int64_t q = brpt[k]; q < brpt[k + 1]; q++) {
                int64_t pos = find_col(ccol, crpt[i], crpt[i + 1], bcol[q]);
                double t = a * dc[pos], s = db[q], u = s + t;
                if (fabs(s) >= fabs(t)) dbc[q] += (s - u) + t; else dbc[q] += (t - u) + s;
                db[q] = u;
            }
        }
    }
    for (int64_t e = 0; e < nnzB; e++) db[e] += dbc[e];
    (void)nnzA;
    free(dbc);
}

/* ------------------------------------------------------------------------- */
/* Binary COO reader and row splitter (src/READ_MATRIX_MPI.cpp:146-356 with
 * partition = 0, src/distribute.cu:1071-1125).  File layout: a header of
 * int64 {nr, nc, nt} (header_type 1) or {nr, nt} (header_type 0), then nt
 * records {int64 i, int64 j, double v} with 1-based i, j, sorted by row.
 * Rank `rank` of `nprocs` owns rows [nr/nprocs*rank, ...), the last rank the
 * remainder (READ_MATRIX_MPI.cpp:256-270); its records are located by binary
 * search on the row id (binsearch_iat, 83-130).  The rank's rows run from
 * its first to its last present row (rowxproc, distribute.cu:1094); columns
 * become 0-based and are shifted by col_shift = the sum of rowxproc over the
 * ranks before it (the MPI_Exscan at distribute.cu:1095, applied at 1103 and
 * stored as the CSR's row_shift at 1105), which equals the rank's first row
 * unless empty rows sit at partition ends.  A rank with no records spans no
 * rows (the reference reads irow[-1] there).  Rows with unsorted columns are
 * sorted (check_and_fix_order, 1050-1069).
 * Returns 0 on success; on success *out_* are malloc'd and owned by the
 * caller (free with orc_free). */
typedef struct { int64_t i, j; double v; } orc_coo_rec;

/* The rank's n records (1-based, sorted by row) as its CSR block: rows from
 * the first to the last present row (distribute.cu:1094-1096; `first` when
 * there are none), columns 0-based minus col_shift (1103), rows with
 * unsorted columns sorted stably (check_and_fix_order, 1050-1069). */
static void coo_records_csr(const orc_coo_rec *rec, int64_t n, int64_t first, int64_t col_shift,
                            int64_t *out_nloc, int64_t **out_rpt, int32_t **out_col, double **out_val) {
    int64_t r_first = n > 0 ? rec[0].i - 1 : first;
    int64_t r_last = n > 0 ? rec[n - 1].i - 1 : first - 1;
    int64_t nloc = r_last - r_first + 1;
    int64_t *rpt = (int64_t *)calloc((size_t)nloc + 1, sizeof(int64_t));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    double *val = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (int64_t q = 0; q < n; q++) rpt[rec[q].i - 1 - r_first + 1]++;
    for (int64_t r = 0; r < nloc; r++) rpt[r + 1] += rpt[r];
    for (int64_t q = 0; q < n; q++) {
        col[q] = (int32_t)(rec[q].j - 1 - col_shift);
        val[q] = rec[q].v;
    }
    for (int64_t r = 0; r < nloc; r++) {
        for (int64_t p = rpt[r] + 1; p < rpt[r + 1]; p++) {
            int32_t c = col[p]; double v = val[p];
            int64_t q = p - 1;
            while (q >= rpt[r] && col[q] > c) { col[q + 1] = col[q]; val[q + 1] = val[q]; q--; }
            col[q + 1] = c; val[q + 1] = v;
        }
    }
    *out_nloc = nloc; *out_rpt = rpt; *out_col = col; *out_val = val;
}

static int64_t coo_lower_bound(FILE *f, int64_t hdr, int64_t nt, int64_t row1) {
    /* first record whose 1-based row id is >= row1 */
    int64_t lo = 0, hi = nt;
    orc_coo_rec r;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        fseeko(f, (off_t)(hdr + mid * (int64_t)sizeof(orc_coo_rec)), SEEK_SET);
        if (fread(&r, sizeof r, 1, f) != 1) return -1;
        if (r.i < row1) lo = mid + 1; else hi = mid;
    }
    return lo;
}

int orc_read_binary_coo(const char *path, int header_type, int nprocs, int rank,
                        int64_t *out_nr, int64_t *out_nc, int64_t *out_nt,
                        int64_t *out_nloc, int64_t *out_row_shift, int64_t *out_nnz,
                        int64_t **out_rpt, int32_t **out_col, double **out_val) {
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    int64_t h[3];
    int64_t hdr = header_type ? 24 : 16;
    if (fread(h, 8, header_type ? 3 : 2, f) != (size_t)(header_type ? 3 : 2)) { fclose(f); return 2; }
    int64_t nr = h[0], nc = header_type ? h[1] : h[0], nt = header_type ? h[2] : h[1];
    int64_t per = nr / nprocs;
    int64_t first = per * rank;
    int64_t last = (rank == nprocs - 1) ? nr : per * (rank + 1);
    int64_t b = coo_lower_bound(f, hdr, nt, first + 1);
    int64_t e = coo_lower_bound(f, hdr, nt, last + 1);
    if (b < 0 || e < 0) { fclose(f); return 3; }
    /* col_shift: rows spanned by the ranks before this one */
    int64_t col_shift = 0;
    for (int r = 0; r < rank; r++) {
        int64_t rb = coo_lower_bound(f, hdr, nt, per * r + 1);
        int64_t re = coo_lower_bound(f, hdr, nt, per * (r + 1) + 1);
        if (rb < 0 || re < 0) { fclose(f); return 3; }
        if (re > rb) {
            orc_coo_rec x, y;
            fseeko(f, (off_t)(hdr + rb * (int64_t)sizeof(orc_coo_rec)), SEEK_SET);
            if (fread(&x, sizeof x, 1, f) != 1) { fclose(f); return 3; }
            fseeko(f, (off_t)(hdr + (re - 1) * (int64_t)sizeof(orc_coo_rec)), SEEK_SET);
            if (fread(&y, sizeof y, 1, f) != 1) { fclose(f); return 3; }
            col_shift += y.i - x.i + 1;
        }
    }
    int64_t n = e - b;
    orc_coo_rec *rec = (orc_coo_rec *)malloc(sizeof(orc_coo_rec) * (size_t)(n > 0 ? n : 1));
    fseeko(f, (off_t)(hdr + b * (int64_t)sizeof(orc_coo_rec)), SEEK_SET);
    if (n > 0 && fread(rec, sizeof(orc_coo_rec), (size_t)n, f) != (size_t)n) { free(rec); fclose(f); return 4; }
    fclose(f);
    coo_records_csr(rec, n, first, col_shift, out_nloc, out_rpt, out_col, out_val);
    free(rec);
    *out_nr = nr; *out_nc = nc; *out_nt = nt;
    *out_row_shift = col_shift; *out_nnz = n;
    return 0;
}

void orc_free(void *p) { free(p); }

/* ---- ASCII COO files (READ_ASCII_MATRIX_MPI, src/READ_MATRIX_MPI.cpp:364-578) ----
 * Text file: lines starting with '%', a header "nr nc nt" or "nr nt" (nc =
 * nr, 381-388), then nt record lines "i j v" of the first record line's
 * length (row_size, 414-429), 1-based and sorted by row.  The column field
 * is read at off_1 and the value at off_2, the first space after the first
 * and second fields of the first record line (419-424).  Rows are split as
 * for the binary files (440-455, binsearch_iat_str 83-104), rows and
 * columns read with atol (column - 1, 474-477), values with the
 * reference's fast_atof (21-77).  The body starts after the '%' lines and
 * the header (the reference adds only the header line's length, 412-413,
 * which is the same offset for a file without '%' lines). */

/* atol over [p, end) */
static int64_t orc_atol(const char *p, const char *end) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\v' || *p == '\f' || *p == '\r')) p++;
    int neg = 0;
    if (p < end && (*p == '-' || *p == '+')) neg = (*p++ == '-');
    uint64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') v = v * 10u + (uint64_t)(*p++ - '0');
    return neg ? -(int64_t)v : (int64_t)v;
}

/* fast_atof (READ_MATRIX_MPI.cpp:21-77): blanks, sign, integer digits as
 * value*10 + d, fraction digits as value += d / pow10 with pow10 *= 10,
 * exponent (clamped to 308) as a product of 1e50, 1e8 and 10 factors. */
double orc_fast_atof(const char *p, const char *end) {
    while (p < end && (*p == ' ' || *p == '\t')) p++;
    double sign = 1.0;
    if (p < end && *p == '-') { sign = -1.0; p++; }
    else if (p < end && *p == '+') p++;
    double value = 0.0;
    for (; p < end && *p >= '0' && *p <= '9'; p++) value = value * 10.0 + (double)(*p - '0');
    if (p < end && *p == '.') {
        double pow10 = 10.0;
        for (p++; p < end && *p >= '0' && *p <= '9'; p++) { value += (double)(*p - '0') / pow10; pow10 *= 10.0; }
    }
    int frac = 0;
    double scale = 1.0;
    if (p < end && (*p == 'e' || *p == 'E')) {
        p++;
        if (p < end && *p == '-') { frac = 1; p++; }
        else if (p < end && *p == '+') p++;
        unsigned int e = 0;
        for (; p < end && *p >= '0' && *p <= '9'; p++) e = e * 10u + (unsigned int)(*p - '0');
        if (e > 308) e = 308;
        while (e >= 50) { scale *= 1e50; e -= 50; }
        while (e >= 8) { scale *= 1e8; e -= 8; }
        while (e > 0) { scale *= 10.0; e -= 1; }
    }
    return sign * (frac ? (value / scale) : (value * scale));
}

static int64_t ascii_upper(const char *body, int64_t L, int64_t nt, int64_t x) {
    /* first record line whose row exceeds x */
    int64_t lo = 0, hi = nt;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (x < orc_atol(body + mid * L, body + (mid + 1) * L)) hi = mid; else lo = mid + 1;
    }
    return lo;
}

int orc_read_ascii_coo(const char *path, int nprocs, int rank,
                       int64_t *out_nr, int64_t *out_nc, int64_t *out_nt,
                       int64_t *out_nloc, int64_t *out_row_shift, int64_t *out_nnz,
                       int64_t **out_rpt, int32_t **out_col, double **out_val) {
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    fseeko(f, 0, SEEK_END);
    int64_t size = (int64_t)ftello(f);
    fseeko(f, 0, SEEK_SET);
    char *buf = (char *)malloc((size_t)size + 2);
    if (fread(buf, 1, (size_t)size, f) != (size_t)size) { free(buf); fclose(f); return 2; }
    fclose(f);
    buf[size] = '\n';  /* the last line may lack its newline */
    buf[size + 1] = 0;
    int64_t off = 0;
    while (off < size && buf[off] == '%') { while (off < size && buf[off] != '\n') off++; off++; }
    if (off >= size) { free(buf); return 2; }
    long a = 0, b = 0, c = 0;
    int64_t nr, nc, nt;
    char hdr[1025];
    int64_t hl = 0;
    while (off + hl < size && buf[off + hl] != '\n' && hl < 1024) { hdr[hl] = buf[off + hl]; hl++; }
    hdr[hl] = 0;
    if (sscanf(hdr, "%ld %ld %ld", &a, &b, &c) == 3) { nr = a; nc = b; nt = c; }
    else if (sscanf(hdr, "%ld %ld", &a, &b) == 2) { nr = nc = a; nt = b; }
    else { free(buf); return 2; }
    const char *body = buf + off + hl + 1;
    int64_t body_bytes = size - (off + hl + 1);
    int64_t L = 1, off1 = 0, off2 = 0;
    if (nt > 0) {
        int64_t i = 0;
        while (i < body_bytes && body[i] == ' ') i++;
        while (i < body_bytes && body[i] != ' ' && body[i] != '\n') i++;
        off1 = i;
        while (i < body_bytes && body[i] == ' ') i++;
        while (i < body_bytes && body[i] != ' ' && body[i] != '\n') i++;
        off2 = i;
        while (i < body_bytes && body[i] != '\n') i++;
        L = i + 1;
        if (L < 2 || nt * L - 1 > body_bytes) { free(buf); return 3; }
    }
    int64_t per = nr / nprocs;
    int64_t first = per * rank, last = (rank == nprocs - 1) ? nr : per * (rank + 1);
    int64_t bq = ascii_upper(body, L, nt, first), eq = ascii_upper(body, L, nt, last);
    int64_t col_shift = 0;
    for (int r = 0; r < rank; r++) {
        int64_t rb = ascii_upper(body, L, nt, per * r), re = ascii_upper(body, L, nt, per * (r + 1));
        if (re > rb)
            col_shift += orc_atol(body + (re - 1) * L, body + re * L) - orc_atol(body + rb * L, body + (rb + 1) * L) + 1;
    }
    int64_t n = eq - bq;
    orc_coo_rec *rec = (orc_coo_rec *)malloc(sizeof(orc_coo_rec) * (size_t)(n > 0 ? n : 1));
    for (int64_t q = 0; q < n; q++) {
        const char *line = body + (bq + q) * L, *end = line + L;
        if (line[L - 1] != '\n') { free(rec); free(buf); return 4; }  /* lines of unequal length */
        rec[q].i = orc_atol(line, end);
        rec[q].j = orc_atol(line + off1, end);
        rec[q].v = orc_fast_atof(line + off2, end);
    }
    free(buf);
    for (int64_t q = 1; q < n; q++)
        if (rec[q].i < rec[q - 1].i) { free(rec); return 5; }
    coo_records_csr(rec, n, first, col_shift, out_nloc, out_rpt, out_col, out_val);
    free(rec);
    *out_nr = nr; *out_nc = nc; *out_nt = nt;
    *out_row_shift = col_shift; *out_nnz = n;
    return 0;
}

/* Binary COO writer in the same layout (main.cu:468-512). */
int orc_write_binary_coo(const char *path, int header_type, int64_t nr, int64_t nc,
                         const int64_t *rpt, const int32_t *col, const double *val, int64_t col_shift) {
    FILE *f = fopen(path, "wb");
    if (!f) return 1;
    int64_t nt = rpt[nr];
    int64_t h[3] = {nr, header_type ? nc : nt, nt};
    if (fwrite(h, 8, header_type ? 3 : 2, f) != (size_t)(header_type ? 3 : 2)) { fclose(f); return 2; }
    for (int64_t i = 0; i < nr; i++)
        for (int64_t p = rpt[i]; p < rpt[i + 1]; p++) {
            orc_coo_rec r = {i + 1, (int64_t)col[p] + col_shift + 1, val[p]};
            if (fwrite(&r, sizeof r, 1, f) != 1) { fclose(f); return 3; }
        }
    fclose(f);
    return 0;
}
