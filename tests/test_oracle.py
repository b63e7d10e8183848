"""Pins the CPU oracle against what the mathematics fixes.

The reference holds no tests, fixtures or golden vectors for the local product,
so each oracle function is checked against closed forms (stencil products),
dense brute force on small integer matrices (exact in float64), scipy on
all-ones patterns, an independent restatement of the three getmct_4shrink
branch forms (src/getmct.cu:135-220), and invariants.  The ``planted``
tests perturb oracle outputs on the CPU and confirm the checks catch it.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from nsp_amd import gen


def dense_product_pattern(A, B):
    pa = _pat(A)
    pb = _pat(B)
    return (pa @ pb) > 0


def _pat(X):
    d = np.zeros((X.nrows, X.ncols), np.int64)
    rows = np.repeat(np.arange(X.nrows), np.diff(X.rpt))
    d[rows, X.col] = 1
    return d


def oracle_dense(orc, A, B):
    crpt, ccol, cval, cabs = orc.spgemm(A, B)
    d = np.zeros((A.nrows, B.ncols))
    s = np.zeros((A.nrows, B.ncols), bool)
    rows = np.repeat(np.arange(A.nrows), np.diff(crpt))
    d[rows, ccol] = cval
    s[rows, ccol] = True
    return crpt, ccol, cval, cabs, d, s


def check_sorted_rows(rpt, col):
    for i in range(len(rpt) - 1):
        seg = col[rpt[i]:rpt[i + 1]]
        assert np.all(np.diff(seg) > 0), f"row {i} not strictly increasing"


# ----------------------------------------------------------------- closed forms
def laplacian_sq_closed_form(m):
    """(L^2) for the 5-point Laplacian: C_ii = 16 + deg(i), -8 at grid
    distance 1, 1 at distance 2 in a straight line, 2 on diagonals."""
    n = m * m
    C = {}
    for i in range(n):
        x, y = i % m, i // m
        deg = (x > 0) + (x < m - 1) + (y > 0) + (y < m - 1)
        C[(i, i)] = 16.0 + deg
        for dx, dy, v in [(1, 0, -8.0), (-1, 0, -8.0), (0, 1, -8.0), (0, -1, -8.0),
                          (2, 0, 1.0), (-2, 0, 1.0), (0, 2, 1.0), (0, -2, 1.0),
                          (1, 1, 2.0), (1, -1, 2.0), (-1, 1, 2.0), (-1, -1, 2.0)]:
            xx, yy = x + dx, y + dy
            if 0 <= xx < m and 0 <= yy < m:
                C[(i, yy * m + xx)] = v
    return C


@pytest.mark.parametrize("m", [3, 7, 16])
def test_laplacian_square_closed_form(orc, m):
    L = gen.laplacian_5pt(m)
    crpt, ccol, cval, _ = orc.spgemm(L, L)
    check_sorted_rows(crpt, ccol)
    ref = laplacian_sq_closed_form(m)
    got = {}
    for i in range(L.nrows):
        for p in range(crpt[i], crpt[i + 1]):
            got[(i, int(ccol[p]))] = float(cval[p])
    assert got == ref  # exact: small integers


def test_laplacian_256_counts(orc):
    """Config-1 sizes: n=65,536, nnz(A)=326,656, nprod=1,629,192, nnz(C)=846,852."""
    L = gen.laplacian_5pt(256)
    assert L.nnz == 326_656
    _, total = orc.intprod(L, L)
    assert total == 1_629_192
    nz = orc.symbolic(L, L)
    assert int(nz.sum()) == 846_852 and int(nz.max()) == 13


@pytest.mark.parametrize("m", [4, 6])
def test_stencil27_square_structure_closed_form(orc, m):
    A = gen.stencil_27pt(m, seed=3)
    crpt, ccol, cval, _ = orc.spgemm(A, A)
    n = m ** 3
    for i in range(n):
        x, y, z = i % m, (i // m) % m, i // (m * m)
        want = sorted((z + dz) * m * m + (y + dy) * m + (x + dx)
                      for dz, dy, dx in itertools.product(range(-2, 3), repeat=3)
                      if 0 <= x + dx < m and 0 <= y + dy < m and 0 <= z + dz < m)
        assert list(ccol[crpt[i]:crpt[i + 1]]) == want


def test_stencil27_counts_closed_form(orc):
    m = 10
    A = gen.stencil_27pt(m, seed=1)
    nz = orc.symbolic(A, A)
    i = np.arange(m ** 3)
    span5 = lambda c: np.minimum(c, 2) + np.minimum(m - 1 - c, 2) + 1
    want = span5(i % m) * span5((i // m) % m) * span5(i // (m * m))
    np.testing.assert_array_equal(nz, want)
    ip, _ = orc.intprod(A, A)
    cnt = gen.stencil_27pt_counts(m, 0, m ** 3)
    want_ip = np.array([cnt[A.col[A.rpt[r]:A.rpt[r + 1]]].sum() for r in range(m ** 3)])
    np.testing.assert_array_equal(ip, want_ip)


def test_stencil_block_equals_full_rows():
    full = gen.stencil_27pt(8, seed=11)
    blk = gen.stencil_27pt(8, seed=11, r0=100, r1=300)
    ref = full.row_slice(100, 300)
    np.testing.assert_array_equal(blk.rpt, ref.rpt)
    np.testing.assert_array_equal(blk.col, ref.col)
    np.testing.assert_array_equal(blk.val, ref.val)
    d = blk.col == np.repeat(np.arange(100, 300), np.diff(blk.rpt))
    assert np.all(blk.val[d] == 27.0) and np.all(np.abs(blk.val[~d]) < 1)


# ---------------------------------------------------------------- brute force
@pytest.mark.parametrize("seed", range(6))
def test_random_integer_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    m, k, n = rng.integers(1, 40, 3)
    A = gen.random_csr(int(m), int(k), 0.2, seed=seed, integer_values=True, empty_rows=0.2)
    B = gen.random_csr(int(k), int(n), 0.25, seed=seed + 100, integer_values=True, empty_rows=0.2)
    crpt, ccol, cval, cabs, d, s = oracle_dense(orc, A, B)
    check_sorted_rows(crpt, ccol)
    np.testing.assert_array_equal(s, dense_product_pattern(A, B))      # zeros kept
    np.testing.assert_array_equal(d, A.to_dense() @ B.to_dense())       # exact integers
    absd = np.abs(A.to_dense()) @ np.abs(B.to_dense())
    rows = np.repeat(np.arange(A.nrows), np.diff(crpt))
    np.testing.assert_array_equal(cabs, absd[rows, ccol])


def test_structural_zeros_kept(orc):
    # row 0 of A = [1, 1], B rows [1] and [-1] in the same column: product 0
    A = gen.HostCSR(np.array([0, 2]), np.array([0, 1], np.int32), np.array([1.0, 1.0]), 1, 2)
    B = gen.HostCSR(np.array([0, 1, 2]), np.array([0, 0], np.int32), np.array([1.0, -1.0]), 2, 1)
    crpt, ccol, cval, _ = orc.spgemm(A, B)
    assert list(crpt) == [0, 1] and list(ccol) == [0] and cval[0] == 0.0


def test_scipy_pattern_product(orc):
    A = gen.random_csr(600, 500, 0.01, seed=7, empty_rows=0.1)
    B = gen.random_csr(500, 700, 0.01, seed=8, empty_rows=0.1)
    crpt, ccol, _, _ = orc.spgemm(A, B)
    pa = sp.csr_matrix((np.ones(A.nnz), A.col, A.rpt), shape=(A.nrows, A.ncols))
    pb = sp.csr_matrix((np.ones(B.nnz), B.col, B.rpt), shape=(B.nrows, B.ncols))
    pc = (pa @ pb).tocsr()
    pc.sort_indices()
    np.testing.assert_array_equal(crpt, pc.indptr)
    np.testing.assert_array_equal(ccol, pc.indices)


def test_compensated_values_random(orc):
    A = gen.random_csr(300, 250, 0.05, seed=1)
    B = gen.random_csr(250, 280, 0.05, seed=2)
    _, _, _, _, d, _ = oracle_dense(orc, A, B)
    np.testing.assert_allclose(d, A.to_dense() @ B.to_dense(), rtol=0, atol=1e-13)


def test_exscan_and_rows_subset(orc):
    A = gen.random_csr(200, 150, 0.05, seed=4, empty_rows=0.3)
    B = gen.random_csr(150, 180, 0.05, seed=5)
    crpt, ccol, cval, cabs = orc.spgemm(A, B)
    np.testing.assert_array_equal(crpt, np.concatenate([[0], np.cumsum(orc.symbolic(A, B))]))
    rows = np.array([199, 3, 50, 0, 77], np.int64)
    rrpt, rcol, rval, _ = orc.spgemm_rows(A, B, rows)
    for t, i in enumerate(rows):
        np.testing.assert_array_equal(rcol[rrpt[t]:rrpt[t + 1]], ccol[crpt[i]:crpt[i + 1]])
        np.testing.assert_array_equal(rval[rrpt[t]:rrpt[t + 1]], cval[crpt[i]:crpt[i + 1]])


def test_empty_inputs(orc):
    A = gen.HostCSR(np.zeros(5, np.int64), np.zeros(0, np.int32), np.zeros(0), 4, 3)
    B = gen.random_csr(3, 6, 0.5, seed=0)
    crpt, ccol, cval, _ = orc.spgemm(A, B)
    assert list(crpt) == [0] * 5 and ccol.size == 0
    Z = gen.HostCSR(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0), 0, 3)
    crpt, _, _, _ = orc.spgemm(Z, B)
    assert list(crpt) == [0]


# ---------------------------------------------------------------- compaction
def shrink_branch_form(col, l, which):
    """getmct_4shrink's three branches (src/getmct.cu:135-220), restated
    literally: uv = sorted unique cols outside [0, l]; the first rank appends
    uv after [0..l], the last rank prepends it, a middle rank splits it."""
    col = np.asarray(col, np.int64)
    uv = np.unique(col[(col < 0) | (col > l)])
    loc = np.arange(l + 1)
    if which == "first":
        perm = np.concatenate([loc, uv])
    elif which == "last":
        perm = np.concatenate([uv, loc])
    else:
        perm = np.concatenate([uv[uv < 0], loc, uv[uv > l]])
    newcol = np.array([np.flatnonzero(perm == c)[0] for c in col], np.int64)
    return perm, newcol


@pytest.mark.parametrize("which,lo,hi", [("first", 0, 60), ("last", -40, 20), ("middle", -40, 60)])
def test_shrink_matches_branch_forms(orc, which, lo, hi):
    rng = np.random.default_rng(5)
    l = 19
    col = rng.integers(lo, hi, 300).astype(np.int32)
    perm, newcol = orc.shrink(col, l)
    bperm, bnew = shrink_branch_form(col, l, which)
    np.testing.assert_array_equal(perm, bperm)
    np.testing.assert_array_equal(newcol, bnew)
    np.testing.assert_array_equal(perm[newcol], col)
    assert np.all(np.diff(perm) > 0)


def test_local_share_stitches_to_full_product(orc):
    """Per-rank products, mapped back to global columns and stacked, equal the
    full product (the multi-rank analogue of the reference's per-rank run)."""
    A = gen.random_csr(90, 90, 0.06, seed=21, empty_rows=0.05)
    n = A.nrows
    nprocs = 3
    crpt, ccol, cval, _ = orc.spgemm(A, A)
    per = n // nprocs
    for r in range(nprocs):
        r0 = per * r
        r1 = n if r == nprocs - 1 else per * (r + 1)
        A_rank = A.row_slice(r0, r1, col_shift=r0)
        sh = orc.local_share(A_rank, A, r0)
        A_, P_ = sh["A_"], sh["P_"]
        assert A_.ncols == P_.nrows                       # Alocal_->m == completedP->n
        lrpt, lcol, lval, _ = orc.spgemm(A_, P_)
        gcol = sh["perm_P"][lcol].astype(np.int64) + r0    # back to global columns
        for i in range(r1 - r0):
            seg = slice(lrpt[i], lrpt[i + 1])
            ref = slice(crpt[r0 + i], crpt[r0 + i + 1])
            np.testing.assert_array_equal(gcol[seg], ccol[ref])
            np.testing.assert_allclose(lval[seg], cval[ref], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("r0,r1", [(0, 30), (30, 60), (60, 90)])
def test_completed_rows_and_gather_closed_form(orc, r0, r1):
    """completedP's rows for a first, middle and last block, against their
    definition (src/spspmpi.cu:37-55, 395-906): the global rows
    sorted(unique((cols + r0) U [r0, r0 + l])), which puts the halo rows below
    the block, then the block, then the halo rows above, and each A column
    re-expressed as its position in that list; each gathered row is B's row
    with columns moved to the rank frame (global - r0, spspmpi.cu:822-828)."""
    A = gen.random_csr(90, 90, 0.06, seed=33, empty_rows=0.05)
    A_rank = A.row_slice(r0, r1, col_shift=r0)
    l = (r1 - r0) - 1
    grows, newcol = orc.completed_rows(A_rank.col, l, r0)
    gcols = A_rank.col.astype(np.int64) + r0
    expect = np.unique(np.concatenate([gcols, np.arange(r0, r0 + l + 1, dtype=np.int64)]))
    np.testing.assert_array_equal(grows, expect)
    np.testing.assert_array_equal(grows[newcol], gcols)
    below, local, above = grows[grows < r0], grows[(grows >= r0) & (grows < r1)], grows[grows >= r1]
    np.testing.assert_array_equal(local, np.arange(r0, r1))
    np.testing.assert_array_equal(grows, np.concatenate([below, local, above]))
    prpt, pcol, pval = orc.gather_rows(A, grows, r0)
    assert prpt[0] == 0 and prpt.size == grows.size + 1
    for i, g in enumerate(grows):
        b, e = A.rpt[g], A.rpt[g + 1]
        np.testing.assert_array_equal(pcol[prpt[i]:prpt[i + 1]], A.col[b:e].astype(np.int64) - r0)
        np.testing.assert_array_equal(pval[prpt[i]:prpt[i + 1]], A.val[b:e])


def test_config2_share_is_identity_map(orc):
    """Rank 0 of 4 of the 27-pt stencil: both compaction maps are the
    identity (m=16 keeps it small; 128^3 gives shrinked_m 540,672/557,056)."""
    m = 16
    n = m ** 3
    A = gen.stencil_27pt(m, seed=0)
    A0 = A.row_slice(0, n // 4)
    sh = orc.local_share(A0, A, 0)
    assert sh["A_"].ncols == n // 4 + m * m
    np.testing.assert_array_equal(sh["A_"].col, A0.col)
    np.testing.assert_array_equal(sh["perm_P"], np.arange(sh["perm_P"].size))
    assert sh["perm_P"].size == n // 4 + 2 * m * m


# ---------------------------------------------------------------- transpose / backward
def test_transpose(orc):
    X = gen.random_csr(40, 30, 0.1, seed=9)
    trpt, tcol, tval = orc.transpose(X)
    T = gen.HostCSR(trpt, tcol, tval, 30, 40)
    np.testing.assert_array_equal(T.to_dense(), X.to_dense().T)
    check_sorted_rows(trpt, tcol)


@pytest.mark.parametrize("seed", range(3))
def test_backward_closed_form(orc, seed):
    A = gen.random_csr(25, 20, 0.2, seed=seed, integer_values=True)
    B = gen.random_csr(20, 22, 0.2, seed=seed + 50, integer_values=True)
    crpt, ccol, cval, _ = orc.spgemm(A, B)
    dc = np.random.default_rng(seed).integers(-3, 4, ccol.size).astype(np.float64)
    da, db = orc.backward(A, B, crpt, ccol, dc)
    dC = np.zeros((A.nrows, B.ncols))
    dC[np.repeat(np.arange(A.nrows), np.diff(crpt)), ccol] = dc
    dA_full = dC @ B.to_dense().T
    dB_full = A.to_dense().T @ dC
    ra = np.repeat(np.arange(A.nrows), np.diff(A.rpt))
    rb = np.repeat(np.arange(B.nrows), np.diff(B.rpt))
    np.testing.assert_array_equal(da, dA_full[ra, A.col])
    np.testing.assert_array_equal(db, dB_full[rb, B.col])


# ---------------------------------------------------------------- reader / splitter
def test_binary_coo_roundtrip_and_split(orc, tmp_path):
    A = gen.random_csr(50, 50, 0.1, seed=3)
    # every row non-empty: add the diagonal
    from nsp_amd.gen import _from_rows_cols
    rows = np.concatenate([np.repeat(np.arange(50), np.diff(A.rpt)), np.arange(50)])
    cols = np.concatenate([A.col, np.arange(50)])
    key, idx = np.unique(rows * 50 + cols, return_index=True)
    vals = np.concatenate([A.val, np.full(50, 9.0)])[idx]
    A = _from_rows_cols(50, 50, key // 50, key % 50, vals)
    p = str(tmp_path / "a.bin")
    orc.write_binary_coo(p, A)
    nr, nc, nt, shift, X = orc.read_binary_coo(p)
    assert (nr, nc, nt, shift) == (50, 50, A.nnz, 0)
    np.testing.assert_array_equal(X.rpt, A.rpt)
    np.testing.assert_array_equal(X.col, A.col)
    np.testing.assert_array_equal(X.val, A.val)
    # 3 ranks: rows split nr/nprocs with the remainder on the last rank,
    # columns shifted by the rank's first row
    blocks = [orc.read_binary_coo(p, nprocs=3, rank=r) for r in range(3)]
    assert [b[3] for b in blocks] == [0, 16, 32]
    assert [b[4].nrows for b in blocks] == [16, 16, 18]
    for _, _, _, shift, B in blocks:
        ref = A.row_slice(shift, shift + B.nrows, col_shift=shift)
        np.testing.assert_array_equal(B.rpt, ref.rpt)
        np.testing.assert_array_equal(B.col, ref.col)


def test_reader_drops_empty_rows_at_block_ends(orc, tmp_path):
    """Reference quirk: a rank's row range runs from its first to its last
    present row (src/distribute.cu:1096), so trailing empty rows are lost."""
    A = gen.HostCSR(np.array([0, 1, 2, 2, 2]), np.array([0, 1], np.int32), np.ones(2), 4, 4)
    p = str(tmp_path / "e.bin")
    orc.write_binary_coo(p, A)
    _, _, _, shift, X = orc.read_binary_coo(p)
    assert shift == 0 and X.nrows == 2


def test_reader_shift_is_exscan_of_spanned_rows(orc, tmp_path):
    """Reference rule (src/distribute.cu:1094-1105): a rank's columns and
    row_shift are the sum of rowxproc (first to last present row) over the
    ranks before it.  Row 3 of 8 is empty, so with 2 ranks rank 0 spans 3
    rows and rank 1's shift is 3, not its first row 4."""
    p = str(tmp_path / "x.bin")
    recs = np.array([(i + 1, i + 1, float(i)) for i in (0, 1, 2, 4, 5, 6, 7)],
                    dtype=[("i", "<i8"), ("j", "<i8"), ("v", "<f8")])
    with open(p, "wb") as f:
        f.write(np.array([8, 8, len(recs)], "<i8").tobytes())
        f.write(recs.tobytes())
    _, _, _, s0, X0 = orc.read_binary_coo(p, nprocs=2, rank=0)
    _, _, _, s1, X1 = orc.read_binary_coo(p, nprocs=2, rank=1)
    assert (s0, X0.nrows, list(X0.rpt)) == (0, 3, [0, 1, 2, 3])
    assert (s1, X1.nrows, list(X1.rpt)) == (3, 4, [0, 1, 2, 3, 4])
    assert list(X1.col) == [4 - 3, 5 - 3, 6 - 3, 7 - 3]   # global col (0-based) minus 3
    assert list(X1.val) == [4.0, 5.0, 6.0, 7.0]


def test_reader_sorts_unsorted_rows(orc, tmp_path):
    p = str(tmp_path / "u.bin")
    recs = np.array([(1, 3, 1.0), (1, 1, 2.0), (2, 2, 3.0)],
                    dtype=[("i", "<i8"), ("j", "<i8"), ("v", "<f8")])
    with open(p, "wb") as f:
        f.write(np.array([2, 3, 3], "<i8").tobytes())
        f.write(recs.tobytes())
    _, _, _, _, X = orc.read_binary_coo(p)
    assert list(X.col) == [0, 2, 1] and list(X.val) == [2.0, 1.0, 3.0]


# ---------------------------------------------------------------- ASCII reader
def _with_diagonal(A):
    from nsp_amd.gen import _from_rows_cols
    n = A.nrows
    rows = np.concatenate([np.repeat(np.arange(n), np.diff(A.rpt)), np.arange(n)])
    cols = np.concatenate([A.col, np.arange(n)])
    key, idx = np.unique(rows * A.ncols + cols, return_index=True)
    vals = np.concatenate([A.val, np.full(n, 9.0)])[idx]
    return _from_rows_cols(n, A.ncols, key // A.ncols, key % A.ncols, vals)


@pytest.mark.parametrize("nprocs", [1, 2, 3, 4])
def test_ascii_coo_matches_binary_reader(orc, tmp_path, nprocs):
    """Values that the reference's parser reads exactly (half-integers: every
    step, value*10 + d and d/10, is exact): the ASCII block of every rank
    equals the binary reader's block of the same matrix bit for bit (same
    split, shift, row order)."""
    from tests.helpers import write_ascii_coo
    A = _with_diagonal(gen.random_csr(70, 70, 0.1, seed=4))
    A = gen.HostCSR(A.rpt, A.col, np.round(A.val * 200) / 2, A.nrows, A.ncols)
    pa, pb = str(tmp_path / "a.txt"), str(tmp_path / "a.bin")
    write_ascii_coo(pa, A, value_fmt="{:.1f}")
    orc.write_binary_coo(pb, A)
    for r in range(nprocs):
        ta = orc.read_ascii_coo(pa, nprocs, r)
        tb = orc.read_binary_coo(pb, nprocs, r)
        assert ta[:4] == tb[:4]
        np.testing.assert_array_equal(ta[4].rpt, tb[4].rpt)
        np.testing.assert_array_equal(ta[4].col, tb[4].col)
        np.testing.assert_array_equal(ta[4].val, tb[4].val)


def test_ascii_header_forms_comments_and_last_line(orc, tmp_path):
    """'%' lines before the header are skipped, "nr nt" means nc = nr, a file
    whose last line has no newline still holds nt records, and rows out of
    column order are sorted stably (check_and_fix_order)."""
    from tests.helpers import write_ascii_coo
    recs = [(1, 3, 1.5), (1, 1, 2.0), (1, 2, -0.5), (3, 3, 4.0)]
    p = str(tmp_path / "h.txt")
    write_ascii_coo(p, records=recs, nr=3, nc=3, header_fields=2, comments=["%%MatrixMarket x", "% note"],
                    value_fmt="{:g}", final_newline=False)
    nr, nc, nt, shift, X = orc.read_ascii_coo(p)
    assert (nr, nc, nt, shift) == (3, 3, 4, 0)
    assert list(X.rpt) == [0, 3, 3, 4]
    assert list(X.col) == [0, 1, 2, 2] and list(X.val) == [2.0, -0.5, 1.5, 4.0]


def test_ascii_unequal_lines_rejected(orc, tmp_path):
    """The reference reads record q at q*row_size (READ_MATRIX_MPI.cpp:482,
    556); a file with a longer line is read misaligned there and rejected
    here."""
    p = str(tmp_path / "u.txt")
    with open(p, "w") as f:
        f.write("2 2 2\n 1 1 1.0\n 2 2 10.25\n")
    with pytest.raises(IOError):
        orc.read_ascii_coo(p)


@pytest.mark.parametrize("text", ["0", "7", "-12", "+33", "9007199254740991", "0.5", "-3.5", "2.5e1", "2.5E-1",
                                  "3e2", "-4.5e+3", "1e0", "  6.5", "\t-8", "1.0e22", "5e-1", "12.50"])
def test_fast_atof_exact_inputs(orc, text):
    """Inputs on which every step of the reference's parser is exact
    (integer digits, a fraction digit 5 in the first place or 0 after it,
    powers of ten up to 1e22, which are exact doubles) give the correctly
    rounded value, as strtod does."""
    assert orc.fast_atof(text) == float(text)


def test_fast_atof_within_a_few_ulps_of_strtod(orc):
    """On general inputs the parser's separately rounded steps stay within
    a few units in the last place of the correctly rounded value (a wrong
    digit weight, exponent sign or scale would miss by orders of
    magnitude)."""
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, 3000) * 10.0 ** rng.integers(-20, 21, 3000)
    for v in x:
        for text in (f"{v:.17g}", f"{v:.6e}", f"{v:.9f}"):
            ref = float(text)
            got = orc.fast_atof(text)
            assert abs(got - ref) <= 8 * np.spacing(abs(ref)) + 1e-300, (text, got, ref)


# ---------------------------------------------------------------- planted mistakes
def test_planted_dropped_term_is_caught(orc):
    A = gen.random_csr(30, 30, 0.2, seed=31, integer_values=True)
    crpt, ccol, cval, cabs, d, s = oracle_dense(orc, A, A)
    bad = d.copy()
    i = int(np.argmax(np.diff(A.rpt)))
    k = A.col[A.rpt[i]]                                   # drop the first product term of row i
    bad[i] -= A.val[A.rpt[i]] * A.to_dense()[k]
    assert not np.array_equal(bad, A.to_dense() @ A.to_dense())


def test_planted_transposed_operand_is_caught(orc):
    A = gen.random_csr(30, 30, 0.15, seed=41, integer_values=True)
    B = gen.random_csr(30, 30, 0.15, seed=42, integer_values=True)
    Bt_rpt, Bt_col, Bt_val = orc.transpose(B)
    Bt = gen.HostCSR(Bt_rpt, Bt_col, Bt_val, 30, 30)
    _, _, _, _, d_wrong, _ = oracle_dense(orc, A, Bt)
    assert not np.array_equal(d_wrong, A.to_dense() @ B.to_dense())


def test_planted_wrong_shrink_branch_is_caught(orc):
    col = np.array([-5, 3, 25, -1, 0], np.int32)
    perm, newcol = orc.shrink(col, 9)
    wrong_perm, _ = shrink_branch_form(col, 9, "first")   # first-rank form on mixed halos
    assert not np.array_equal(perm, wrong_perm)


def test_uneven_batch_generator():
    X, off = gen.uneven_batch_concat(20, seed=3, nmin=64, nmax=300)
    assert off[0] == 0 and off[-1] == X.nrows and len(off) == 21
    sizes = np.diff(off)
    assert sizes.min() >= 64 and sizes.max() <= 300
    for b in range(20):
        m = X.row_slice(int(off[b]), int(off[b + 1]))
        assert m.col.max() < sizes[b]
        for i in range(m.nrows):
            seg = m.col[m.rpt[i]:m.rpt[i + 1]]
            assert seg.size >= 1 and np.all(np.diff(seg) > 0)
    mats = gen.uneven_batch(20, seed=3, nmin=64, nmax=300)
    np.testing.assert_array_equal(mats[7].col, X.row_slice(int(off[7]), int(off[8])).col)


def test_autotune_signature_and_cache(tmp_path, monkeypatch):
    """Host side of the autotuner: bucketed signatures and the JSON cache."""
    from nsp_amd import autotune
    assert autotune._bucket([0, 1, 2, 3, 1024]) == [0, 1, 2, 2, 11]

    class M:
        value_type, index_type = 0, 0
    s1 = autotune.signature({"sym_bin": [0, 5, 900], "num_bin": [7, 0]}, M, False)
    s2 = autotune.signature({"sym_bin": [0, 6, 1000], "num_bin": [6, 0]}, M, False)
    s3 = autotune.signature({"sym_bin": [0, 6, 1000], "num_bin": [6, 0]}, M, True)
    assert s1 == s2 and s1 != s3
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "x" / "t.json"))
    autotune._save_cache(autotune.cache_path(), {s1: {"params": {"grid_cap": 4}}})
    assert autotune._load_cache(autotune.cache_path())[s1]["params"]["grid_cap"] == 4
    assert autotune._load_cache(str(tmp_path / "missing.json")) == {}
