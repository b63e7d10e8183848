/* Self-test of the CPU oracle (test infrastructure only), built with
 * -fsanitize=address,undefined by `make selftest` and run by
 * tests/test_oracle_sanitized.py: random small integer matrices (empty
 * shapes and rows included) through intprod, symbolic, exscan, spgemm,
 * transpose, shrink and the binary COO writer/reader, each checked against
 * dense brute force, so that any out-of-bounds access, leak or undefined
 * behaviour in the restatement is reported by the sanitizers.
 * Exit status 0 = all checks passed. */
#include "oracle.c"

#include <stdio.h>

static uint64_t rs = 88172645463325252ull;
static uint64_t rnd(void) {
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return rs;
}
static int64_t rint_(int64_t n) { return n > 0 ? (int64_t)(rnd() % (uint64_t)n) : 0; }

#define FAIL(...) do { fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); return 1; } while (0)

typedef struct {
    int64_t r, c;
    int64_t *d;                 /* dense r x c */
    int64_t *rpt; int32_t *col; double *val;
} mat;

static void make(mat *m, int64_t r, int64_t c, int pct) {
    m->r = r; m->c = c;
    m->d = (int64_t *)calloc((size_t)(r * c + 1), sizeof(int64_t));
    for (int64_t i = 0; i < r * c; i++)
        if (rint_(100) < pct) m->d[i] = (1 + rint_(3)) * (rint_(2) ? 1 : -1);
    if (r > 0 && rint_(3) == 0) for (int64_t j = 0; j < c; j++) m->d[rint_(r) * c + j] = 0;  /* an empty row */
    m->rpt = (int64_t *)calloc((size_t)r + 1, sizeof(int64_t));
    int64_t nz = 0;
    for (int64_t i = 0; i < r * c; i++) nz += m->d[i] != 0;
    m->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nz + 1));
    m->val = (double *)malloc(sizeof(double) * (size_t)(nz + 1));
    int64_t p = 0;
    for (int64_t i = 0; i < r; i++) {
        for (int64_t j = 0; j < c; j++)
            if (m->d[i * c + j]) { m->col[p] = (int32_t)j; m->val[p] = (double)m->d[i * c + j]; p++; }
        m->rpt[i + 1] = p;
    }
}
static void drop(mat *m) { free(m->d); free(m->rpt); free(m->col); free(m->val); }

static int check_product(const mat *A, const mat *B) {
    const int64_t M = A->r, K = A->c, N = B->c;
    int64_t *ip = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *nz = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t *crpt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(M + 1));
    int64_t total = orc_intprod(M, A->rpt, A->col, B->rpt, ip);
    int64_t t2 = 0;
    for (int64_t i = 0; i < M; i++) {
        int64_t s = 0;
        for (int64_t k = 0; k < K; k++) if (A->d[i * K + k]) s += B->rpt[k + 1] - B->rpt[k];
        if (ip[i] != s) FAIL("intprod row %lld: %lld != %lld", (long long)i, (long long)ip[i], (long long)s);
        t2 += s;
    }
    if (total != t2) FAIL("intprod total");
    orc_symbolic(M, A->rpt, A->col, B->rpt, B->col, N, nz);
    const int64_t nnz = orc_exscan(M, nz, crpt);
    int32_t *ccol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    double *cval = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    double *cabs = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    orc_spgemm(M, A->rpt, A->col, A->val, B->rpt, B->col, B->val, N, crpt, ccol, cval, cabs);
    int64_t p = 0;
    for (int64_t i = 0; i < M; i++) {
        if (crpt[i] != p) FAIL("row_ptr %lld", (long long)i);
        for (int64_t j = 0; j < N; j++) {
            int pat = 0;
            int64_t v = 0;
            for (int64_t k = 0; k < K; k++)
                if (A->d[i * K + k] && B->d[k * N + j]) { pat = 1; v += A->d[i * K + k] * B->d[k * N + j]; }
            if (!pat) continue;
            if (p >= nnz || ccol[p] != j || cval[p] != (double)v) FAIL("C(%lld,%lld)", (long long)i, (long long)j);
            p++;
        }
    }
    if (p != nnz || crpt[M] != nnz) FAIL("nnz");
    free(ip); free(nz); free(crpt); free(ccol); free(cval); free(cabs);
    return 0;
}

static int check_transpose(const mat *X) {
    const int64_t nnz = X->rpt[X->r];
    int64_t *trpt = (int64_t *)malloc(sizeof(int64_t) * (size_t)(X->c + 1));
    int32_t *tcol = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz + 1));
    double *tval = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    orc_transpose(X->r, X->c, X->rpt, X->col, X->val, trpt, tcol, tval);
    int64_t p = 0;
    for (int64_t j = 0; j < X->c; j++) {
        if (trpt[j] != p) FAIL("transpose row_ptr");
        for (int64_t i = 0; i < X->r; i++)
            if (X->d[i * X->c + j]) {
                if (tcol[p] != i || tval[p] != (double)X->d[i * X->c + j]) FAIL("transpose entry");
                p++;
            }
    }
    free(trpt); free(tcol); free(tval);
    return 0;
}

static int check_shrink(void) {
    const int64_t n = rint_(60), l = rint_(30);
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + l + 2));
    int32_t *nc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; i++) col[i] = (int32_t)(rint_(120) - 20);
    const int64_t m = orc_shrink(n, col, l, perm, nc);
    for (int64_t q = 1; q < m; q++) if (perm[q] <= perm[q - 1]) FAIL("shrink: perm not increasing");
    for (int64_t v = 0; v <= l; v++) {
        int found = 0;
        for (int64_t q = 0; q < m; q++) found |= perm[q] == v;
        if (!found) FAIL("shrink: %lld missing", (long long)v);
    }
    for (int64_t i = 0; i < n; i++) if (nc[i] < 0 || nc[i] >= m || perm[nc[i]] != col[i]) FAIL("shrink map");
    free(col); free(perm); free(nc);
    return 0;
}

static int check_coo(const mat *X, const char *path) {
    if (orc_write_binary_coo(path, 1, X->r, X->c, X->rpt, X->col, X->val, 0)) FAIL("write");
    int64_t nr, nc, nt, nloc, shift, nnz;
    int64_t *rpt = NULL; int32_t *col = NULL; double *val = NULL;
    if (orc_read_binary_coo(path, 1, 1, 0, &nr, &nc, &nt, &nloc, &shift, &nnz, &rpt, &col, &val)) FAIL("read");
    /* one rank: the local rows run from the first to the last row that
     * holds a record (empty rows at the ends are not part of the share, as
     * in the reference's reader) */
    int64_t r0 = 0;
    while (r0 < X->r && X->rpt[r0 + 1] == X->rpt[r0]) r0++;
    if (nr != X->r || nnz != X->rpt[X->r] || r0 + nloc > X->r) FAIL("coo sizes");
    for (int64_t i = 0; i <= nloc; i++) if (rpt[i] != X->rpt[r0 + i] - X->rpt[r0]) FAIL("coo row_ptr");
    for (int64_t i = r0 + nloc; i < X->r; i++) if (X->rpt[i + 1] != X->rpt[i]) FAIL("coo: rows past the share");
    for (int64_t q = 0; q < nnz; q++) if (col[q] != X->col[q] || val[q] != X->val[q]) FAIL("coo entry");
    orc_free(rpt); orc_free(col); orc_free(val);
    return 0;
}

int main(int argc, char **argv) {
    const char *tmp = argc > 1 ? argv[1] : "selftest.coo";
    for (int t = 0; t < 300; t++) {
        mat A, B;
        const int64_t m = rint_(24), k = rint_(24), n = rint_(24);
        make(&A, m, k, (int)rint_(60));
        make(&B, k, n, (int)rint_(60));
        if (check_product(&A, &B) || check_transpose(&A) || check_shrink() ||
            (m > 0 && check_coo(&A, tmp))) {
            fprintf(stderr, "trial %d (%lld x %lld x %lld)\n", t, (long long)m, (long long)k, (long long)n);
            return 1;
        }
        drop(&A); drop(&B);
    }
    remove(tmp);
    printf("oracle self-test: 300 trials passed\n");
    return 0;
}
