"""GPU parity of the local product C = A*B against the CPU oracle.

Structure (row_ptr, col_idx) must match bit-exactly; values within 1e-12
(FP64) or 1e-5 (FP32 with FP32 accumulation) of the FP64 oracle, relative to
sum_k |a_ik b_kj| per entry.  Every call goes through the C ABI of
libnsp_amd.so.
"""
import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import check_against_oracle, to_dtype

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    import nsp_amd
    return nsp_amd


def run(nsp, A, B, dtype="f64", index="i32", handle=None):
    dA = nsp.DeviceCSR.from_host(A, dtype, index)
    dB = nsp.DeviceCSR.from_host(B, dtype, index)
    h = handle or nsp.Handle()
    C = h.spgemm(dA, dB)
    return C, h


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_laplacian_256_closed_form(nsp, orc, dtype):
    """Config 1: 5-point Laplacian on 256^2, C = A*A."""
    print("case: laplacian 256 A*A", dtype, flush=True)
    L = gen.laplacian_5pt(256)
    C, h = run(nsp, L, L, dtype)
    st = h.stats()
    assert st["nprod"] == 1_629_192 and st["nnz"] == 846_852
    comp, _ = check_against_oracle(L, L, C, dtype, orc)
    if dtype == "f64":
        assert comp == 0.0   # small integers: exact


@pytest.mark.parametrize("dtype,index", [("f64", "i32"), ("f64", "i64"), ("f32", "i32"), ("f32", "i64")])
def test_stencil27_small(nsp, orc, dtype, index):
    print("case: stencil27 m=20", dtype, index, flush=True)
    A = to_dtype(gen.stencil_27pt(20, seed=5), dtype)
    C, _ = run(nsp, A, A, dtype, index)
    check_against_oracle(A, A, C, dtype, orc)


def mixed_rows(nrows, ncols, lo, hi, seed):
    """Rows with log-uniform lengths in [lo, hi], columns uniform."""
    rng = np.random.default_rng(seed)
    lens = np.exp(rng.uniform(np.log(lo), np.log(hi + 1), nrows)).astype(np.int64)
    lens = np.clip(lens, 0, ncols)
    rows = np.repeat(np.arange(nrows), lens)
    cols = np.concatenate([rng.choice(ncols, int(k), replace=False) for k in lens]) \
        if lens.sum() else np.zeros(0, np.int64)
    vals = rng.uniform(-1, 1, rows.size)
    return gen._from_rows_cols(nrows, ncols, rows, cols, vals)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_all_bins(nsp, orc, dtype):
    """Row lengths spread so that every symbolic bin up to 16384 products and
    every numeric bin up to 4096 entries is populated; B has 4000 columns, so
    the wave-level bins run the column-bitmap kernels by default and the
    hash-table kernels with narrow_mode 0."""
    print("case: all bins", dtype, flush=True)
    A = to_dtype(mixed_rows(3000, 2500, 1, 150, seed=1), dtype)
    B = to_dtype(mixed_rows(2500, 4000, 1, 100, seed=2), dtype)
    ref = orc.spgemm(A, B)
    for narrow in (1, 0):
        h = nsp.Handle()
        h.set_param("narrow_mode", narrow)
        C, h = run(nsp, A, B, dtype, handle=h)
        st = h.stats()
        assert all(c > 0 for c in st["sym_bin"][:8]), st["sym_bin"]
        assert all(c > 0 for c in st["num_bin"][:8]), st["num_bin"]
        check_against_oracle(A, B, C, dtype, orc, ref=ref)


@pytest.mark.parametrize("rpw", [1, 2, 4, 8])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_rows_per_wave_and_table_variants(nsp, orc, rpw, dtype):
    """Every wave-level kernel shape (1, 2 or 4 rows per wavefront in each
    bin group, 8 in the small bins; column bitmaps, and hash tables of both
    symbolic and both numeric table factors) gives the same product."""
    print("case: rows per wave", rpw, dtype, flush=True)
    A = to_dtype(mixed_rows(1500, 1200, 1, 60, seed=3), dtype)
    B = to_dtype(mixed_rows(1200, 1600, 1, 40, seed=4), dtype)
    ref = orc.spgemm(A, B)
    # (narrow_mode, sym_table_factor, num_table_factor): the column-bitmap
    # kernels (B has 1600 columns), then the hash kernels with both symbolic
    # and both numeric table sizes
    for narrow, stf, ntf in ((1, 2, 2), (0, 1, 2), (0, 2, 4)):
        h = nsp.Handle()
        for name in ("sym_small_rows_per_wave", "sym_rows_per_wave", "num_small_rows_per_wave",
                     "num_rows_per_wave"):
            h.set_param(name, min(rpw, 4) if name in ("sym_rows_per_wave", "num_rows_per_wave") else rpw)
        h.set_param("narrow_mode", narrow)
        h.set_param("sym_table_factor", stf)
        h.set_param("num_table_factor", ntf)
        C, _ = run(nsp, A, B, dtype, handle=h)
        st = h.stats()
        assert all(c > 0 for c in st["sym_bin"][:6]), st["sym_bin"]
        assert all(c > 0 for c in st["num_bin"][:5]), st["num_bin"]
        check_against_oracle(A, B, C, dtype, orc, ref=ref)


def bounded_rows(nrows, ncols, maxlen, seed, empty=0.0):
    """Rows of uniform length in [0, maxlen], a fraction `empty` of them
    emptied, columns uniform."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, maxlen + 1, nrows)
    lens[rng.random(nrows) < empty] = 0
    rows = np.repeat(np.arange(nrows), lens)
    cols = np.concatenate([rng.choice(ncols, int(k), replace=False) for k in lens])
    vals = rng.uniform(-1, 1, rows.size)
    return gen._from_rows_cols(nrows, ncols, rows, cols, vals)


@pytest.mark.parametrize("maxlen", [16, 32, 64])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_short_b_rows_empty_last_row(nsp, orc, maxlen, dtype):
    """B rows of up to 16, 32 or 64 entries, a tenth of them empty, the last
    one empty too (its row pointer entry is nnz(B), so a product walk that
    loaded at an empty row's start would read past B): the hash kernels of
    every group size and both numeric table factors, rows in every
    wave-level bin, and a 27-point stencil (B rows of 27)."""
    print("case: short B rows", maxlen, dtype, flush=True)
    A = to_dtype(mixed_rows(2000, 1800, 1, 120, seed=11), dtype)
    B = to_dtype(bounded_rows(1800, 6000, maxlen, seed=12, empty=0.1), dtype)
    rp = B.rpt.copy()
    rp[-1] = rp[-2]
    B = gen.HostCSR(rp, B.col[:rp[-1]], B.val[:rp[-1]], B.nrows, B.ncols)
    assert (A.col == B.nrows - 1).any()
    ref = orc.spgemm(A, B)
    for rpw in (1, 2, 4):
        for ntf in (2, 4):
            h = nsp.Handle()
            for name in ("sym_small_rows_per_wave", "sym_rows_per_wave", "num_small_rows_per_wave",
                         "num_rows_per_wave"):
                h.set_param(name, rpw)
            h.set_param("narrow_mode", 0)
            h.set_param("num_table_factor", ntf)
            C, _ = run(nsp, A, B, dtype, handle=h)
            st = h.stats()
            assert sum(st["sym_bin"][:6]) > 0 and sum(st["num_bin"][:5]) > 0
            check_against_oracle(A, B, C, dtype, orc, ref=ref)
    S = to_dtype(gen.stencil_27pt(24, seed=13), dtype)
    C, _ = run(nsp, S, S, dtype)
    check_against_oracle(S, S, C, dtype, orc)


def test_structural_zeros_and_empty_rows(nsp, orc):
    print("case: zeros and empty rows", flush=True)
    A = gen.random_csr(400, 300, 0.02, seed=9, integer_values=True, empty_rows=0.3)
    B = gen.random_csr(300, 350, 0.02, seed=10, integer_values=True, empty_rows=0.3)
    C, _ = run(nsp, A, B)
    comp, _ = check_against_oracle(A, B, C, "f64", orc)
    assert comp == 0.0
    # cancellation: two products summing to zero stay as an explicit entry
    A2 = gen.HostCSR(np.array([0, 2]), np.array([0, 1], np.int32), np.array([1.0, 1.0]), 1, 2)
    B2 = gen.HostCSR(np.array([0, 1, 2]), np.array([3, 3], np.int32), np.array([2.0, -2.0]), 2, 5)
    C2, _ = run(nsp, A2, B2)
    assert C2.row_ptr.cpu().tolist() == [0, 1] and C2.col.cpu().tolist() == [3]
    assert C2.val.cpu().tolist() == [0.0]


def test_empty_inputs(nsp):
    print("case: empty inputs", flush=True)
    Z = gen.HostCSR(np.zeros(6, np.int64), np.zeros(0, np.int32), np.zeros(0), 5, 4)
    B = gen.random_csr(4, 7, 0.5, seed=0)
    C, _ = run(nsp, Z, B)
    assert C.row_ptr.cpu().tolist() == [0] * 6 and C.nnz == 0
    E = gen.HostCSR(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0), 0, 4)
    C, _ = run(nsp, E, B)
    assert C.row_ptr.cpu().tolist() == [0] and C.nnz == 0


def test_repeated_calls_reuse_handle(nsp, orc):
    print("case: handle reuse", flush=True)
    h = nsp.Handle()
    for seed in range(3):
        A = mixed_rows(500 + 300 * seed, 400, 1, 40, seed=seed)
        B = mixed_rows(400, 600, 1, 40, seed=seed + 7)
        C, _ = run(nsp, A, B, handle=h)
        check_against_oracle(A, B, C, "f64", orc)


def test_intprod_matches_oracle(nsp, orc):
    print("case: intprod", flush=True)
    A = mixed_rows(1000, 800, 1, 50, seed=12)
    B = gen.random_csr(800, 900, 0.03, seed=13, empty_rows=0.2)
    h = nsp.Handle()
    ip, total = h.intprod(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
    ref, rtotal = orc.intprod(A, B)
    np.testing.assert_array_equal(ip.cpu().numpy(), ref)
    assert total == rtotal


def test_invalid_arguments_raise(nsp):
    print("case: invalid args", flush=True)
    A = nsp.DeviceCSR.from_host(gen.random_csr(10, 8, 0.3, seed=1))
    B = nsp.DeviceCSR.from_host(gen.random_csr(9, 8, 0.3, seed=2))
    h = nsp.Handle()
    with pytest.raises(nsp.NspError) as e:
        h.spgemm(A, B)                      # A.cols != B.rows
    assert e.value.status == 1
    with pytest.raises(nsp.NspError):
        h.numeric(A, A, A.row_ptr, 0)       # numeric without a symbolic phase


def test_check_csr(nsp):
    print("case: check_csr", flush=True)
    X = gen.random_csr(50, 40, 0.2, seed=3)
    h = nsp.Handle()
    assert h.check_csr(nsp.DeviceCSR.from_host(X)) == -1
    bad = gen.HostCSR(X.rpt.copy(), X.col.copy(), X.val, 50, 40)
    r = int(np.argmax(np.diff(bad.rpt) >= 2))
    bad.col[bad.rpt[r]], bad.col[bad.rpt[r] + 1] = bad.col[bad.rpt[r] + 1], bad.col[bad.rpt[r]]
    assert h.check_csr(nsp.DeviceCSR.from_host(bad)) == r


# ---------------------------------------------------------------- local share
@pytest.mark.parametrize("nprocs,rank", [(4, 0), (3, 1), (3, 2)])
def test_local_share_matches_oracle(nsp, orc, nprocs, rank):
    """One rank's share: compaction maps, completedP, and the compacted product
    equal the oracle's restatement of nsparseMGPU_version1."""
    print("case: local share", nprocs, rank, flush=True)
    from nsp_amd import csr
    A = mixed_rows(600, 600, 1, 12, seed=40 + rank)
    r0, r1 = csr.rank_rows(A.nrows, nprocs, rank)
    A_rank = A.row_slice(r0, r1, col_shift=r0)
    ref = orc.local_share(A_rank, A, r0)
    dAr = nsp.DeviceCSR.from_host(A_rank)
    Al = csr.CSR.from_device(dAr, full_n=A.nrows, row_shift=r0)
    Al.m = A.ncols
    Pfull = nsp.DeviceCSR.from_host(A)
    C, A_, P_, _ = csr.nsparse_mgpu_version1_local(Al, Pfull)
    np.testing.assert_array_equal(A_.col.cpu().numpy(), ref["A_"].col)
    assert A_.m == ref["A_"].ncols
    np.testing.assert_array_equal(P_.row.cpu().numpy().astype(np.int64), ref["P_"].rpt)
    np.testing.assert_array_equal(P_.col.cpu().numpy(), ref["P_"].col)
    np.testing.assert_array_equal(P_.val.cpu().numpy(), ref["P_"].val)
    np.testing.assert_array_equal(C.extra["shrinking_permut_P"].cpu().numpy(), ref["perm_P"])
    assert P_.m == ref["P_"].ncols
    check_against_oracle(ref["A_"], ref["P_"], nsp.DeviceCSR(C.row, C.col, C.val, C.n, C.m), "f64", orc)


def test_stencil_rank0_share_identity(nsp, orc):
    """27-point stencil, rank 0 of 4: both compaction maps are the identity and
    the product equals the direct block product (m=32)."""
    print("case: stencil share identity", flush=True)
    from nsp_amd import csr
    m = 32
    n = m ** 3
    A = gen.stencil_27pt(m, seed=2)
    A0 = A.row_slice(0, n // 4)
    Al = csr.CSR.from_device(nsp.DeviceCSR.from_host(A0), full_n=n, row_shift=0)
    Al.m = n
    C, A_, P_, _ = csr.nsparse_mgpu_version1_local(Al, nsp.DeviceCSR.from_host(A))
    assert A_.m == n // 4 + m * m and P_.m == n // 4 + 2 * m * m
    np.testing.assert_array_equal(A_.col.cpu().numpy(), A0.col)
    np.testing.assert_array_equal(C.extra["shrinking_permut_P"].cpu().numpy(), np.arange(P_.m))
    ref = orc.spgemm(A0, A)
    crpt, ccol, cval, cabs = ref
    np.testing.assert_array_equal(C.row.cpu().numpy().astype(np.int64), crpt)
    np.testing.assert_array_equal(C.col.cpu().numpy(), ccol)


def test_shrink_cols_rejects_out_of_range_column(nsp, orc):
    """A column outside the promised [lo, hi) is reported as NSP_ERR_INVALID;
    the remap kernel maps it to -1 instead of reading past its position
    array, and every in-range column is still remapped (checked against the
    oracle's shrink map on the valid entries)."""
    print("case: shrink out-of-range column", flush=True)
    import torch
    rng = np.random.default_rng(7)
    col = rng.integers(-50, 400, 5000).astype(np.int32)
    col[1234] = 2_000_000_000          # far outside [lo, hi)
    col[4321] = -2_000_000_000
    dcol = torch.from_numpy(col).cuda()
    h = nsp.Handle()
    with pytest.raises(nsp.NspError) as ei:
        h.shrink_cols(dcol, 99, -50, 400)
    assert ei.value.status == nsp._lib.NSP_ERR_INVALID
    # the handle stays usable and the in-range map is right
    ok = np.delete(col, [1234, 4321])
    new_col, perm = h.shrink_cols(torch.from_numpy(ok).cuda(), 99, -50, 400)
    rperm, rnew = orc.shrink(ok, 99)
    np.testing.assert_array_equal(perm.cpu().numpy(), rperm)
    np.testing.assert_array_equal(new_col.cpu().numpy(), rnew)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_wide_columns_unpacked_keys(nsp, orc, dtype):
    """C rows of 257-512 entries (numeric tables of 1024 slots) over 5 M
    columns: past 2^22 columns the table slot no longer fits beside the
    column in a 32-bit sort key, so these rows take the path that looks
    their values up by column; rows of smaller bins still pack."""
    print("case: wide columns", dtype, flush=True)
    rng = np.random.default_rng(21)
    ncols = 5_000_000
    nA, nB = 600, 1200
    rows, cols = [], []
    for i in range(nA):
        k = 10 if i % 3 else int(rng.integers(1, 4))
        rows.append(np.full(k, i))
        cols.append(rng.choice(nB, k, replace=False))
    A = gen._from_rows_cols(nA, nB, np.concatenate(rows), np.concatenate(cols),
                            rng.uniform(-1, 1, sum(len(c) for c in cols)))
    rows, cols = [], []
    for i in range(nB):
        rows.append(np.full(40, i))
        cols.append(rng.choice(ncols, 40, replace=False))
    B = gen._from_rows_cols(nB, ncols, np.concatenate(rows), np.concatenate(cols), rng.uniform(-1, 1, 40 * nB))
    A, B = to_dtype(A, dtype), to_dtype(B, dtype)
    h = nsp.Handle()
    C = h.spgemm(nsp.DeviceCSR.from_host(A, dtype), nsp.DeviceCSR.from_host(B, dtype))
    st = h.stats()
    assert st["num_bin"][5] > 0, st["num_bin"]
    check_against_oracle(A, B, C, dtype, orc)


@pytest.mark.parametrize("rpw", [8, 4])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_small_output_rows_with_many_products(nsp, orc, rpw, dtype):
    """Rows of 100-900 products that collapse onto at most 48 columns (the
    small numeric bins) with 8 and 4 rows per wave (groups of 8 and 16 lanes): each group walks several
    mask words and super-windows of products, so every word's entry-start
    mask and count must be rebuilt for each super-window and row."""
    print("case: small-output rows, many products, rows per wave", rpw, dtype, flush=True)
    rng = np.random.default_rng(31)
    nA, nB = 2000, 300
    pool = rng.choice(5000, 48, replace=False)
    rows, cols = [], []
    for i in range(nB):
        rows.append(np.full(20, i))
        cols.append(rng.choice(pool, 20, replace=False))
    B = gen._from_rows_cols(nB, 5000, np.concatenate(rows), np.concatenate(cols), rng.uniform(-1, 1, 20 * nB))
    rows, cols = [], []
    for i in range(nA):
        k = int(rng.integers(5, 46))
        rows.append(np.full(k, i))
        cols.append(rng.choice(nB, k, replace=False))
    A = gen._from_rows_cols(nA, nB, np.concatenate(rows), np.concatenate(cols),
                            rng.uniform(-1, 1, sum(len(c) for c in cols)))
    A, B = to_dtype(A, dtype), to_dtype(B, dtype)
    h = nsp.Handle()
    h.set_param("num_small_rows_per_wave", rpw)
    C, _ = run(nsp, A, B, dtype, handle=h)
    st = h.stats()
    assert st["num_bin"][2] > 0 and sum(st["num_bin"][3:]) == 0, st["num_bin"]
    check_against_oracle(A, B, C, dtype, orc)


@pytest.mark.parametrize("long_len", [1500, 12000])
def test_intprod_blocks_with_long_rows(nsp, orc, long_len):
    """Intermediate-product counts where 256-row blocks hold a long row
    (more than 1024 entries, summed by the long-row kernel): with 1500
    entries the block's entry run (8.7 K) still fits the four LDS stages
    of 4096, with 12000 (19.2 K) it does not and the block reads its rows directly; the
    other blocks span one to four stages."""
    print("case: intprod with long rows", long_len, flush=True)
    rng = np.random.default_rng(41)
    nA, nB = 2048, 16000
    lens = rng.integers(0, 60, nA)
    lens[300] = long_len
    lens[1800] = 1025
    rows = np.repeat(np.arange(nA), lens)
    cols = np.concatenate([rng.choice(nB, int(k), replace=False) for k in lens])
    A = gen._from_rows_cols(nA, nB, rows, cols, rng.uniform(-1, 1, rows.size))
    B = gen.random_csr(nB, 500, 0.02, seed=42, empty_rows=0.1)
    h = nsp.Handle()
    ip, total = h.intprod(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
    ref, rtotal = orc.intprod(A, B)
    np.testing.assert_array_equal(ip.cpu().numpy(), ref)
    assert total == rtotal


def test_b_beyond_32bit_offsets(nsp):
    """B with 537.9 M > 2^29 entries (FP32): 2^20 - 64 rows of 513 entries,
    then 64 rows of 32.  A's 64 rows take 3 of those last rows each (96
    products, at most 96 entries: the wave-level kernels), whose entries sit
    past 2^29, so the group kernels address B with 64-bit offsets (32-bit
    byte offsets serve only B with at most 2^29 entries).  C is checked
    against a host accumulation of the same products (FP32 products summed in
    FP64 and rounded once, as the kernels do)."""
    import torch
    print("case: B beyond 32-bit offsets (FP32, 537.9 M entries)", flush=True)
    nB, per, short, ncols = 1 << 20, 513, 32, 1 << 21
    nlong = nB - 64
    r = np.arange(nlong, dtype=np.int64)
    blong = ((r * 7)[:, None] + np.arange(per, dtype=np.int64)[None, :]) % ncols
    blong.sort(axis=1)
    rs = np.arange(nlong, nB, dtype=np.int64)
    bshort = ((rs * 7)[:, None] + np.arange(short, dtype=np.int64)[None, :]) % ncols
    bshort.sort(axis=1)
    bcol = np.concatenate([blong.astype(np.int32).ravel(), bshort.astype(np.int32).ravel()])
    del blong
    bval = (1.0 + (bcol % 3)).astype(np.float32)
    brpt = np.concatenate([np.arange(nlong + 1, dtype=np.int64) * per,
                           nlong * per + np.arange(1, 65, dtype=np.int64) * short])
    assert brpt[-1] == bcol.size and bcol.size > (1 << 29) and brpt[nlong] > (1 << 29)
    rng = np.random.default_rng(51)
    nA, k = 64, 3
    acol = np.sort(np.stack([nlong + rng.choice(64, k, replace=False) for _ in range(nA)]), axis=1)
    aval = rng.uniform(-1, 1, (nA, k)).astype(np.float32)
    arpt = np.arange(nA + 1, dtype=np.int64) * k
    dev = torch.device("cuda")
    dA = nsp.DeviceCSR(torch.from_numpy(arpt.astype(np.int32)).to(dev),
                       torch.from_numpy(acol.astype(np.int32).ravel()).to(dev), torch.from_numpy(aval.ravel()).to(dev),
                       nA, nB)
    dB = nsp.DeviceCSR(torch.from_numpy(brpt.astype(np.int32)).to(dev), torch.from_numpy(bcol).to(dev),
                       torch.from_numpy(bval).to(dev), nB, ncols)
    h = nsp.Handle()
    C = h.spgemm(dA, dB)
    st = h.stats()
    assert sum(st["sym_bin"][6:]) == 0 and sum(st["num_bin"][6:]) == 0, (st["sym_bin"], st["num_bin"])
    crpt = C.row_ptr.cpu().numpy().astype(np.int64)
    ccol, cval = C.col.cpu().numpy(), C.val.cpu().numpy()
    del dA, dB, C
    torch.cuda.empty_cache()
    for i in range(nA):
        cols = np.concatenate([bcol[brpt[j]:brpt[j + 1]] for j in acol[i]])
        prods = np.concatenate([np.float64(np.float32(aval[i, q]) * bval[brpt[j]:brpt[j + 1]])
                                for q, j in enumerate(acol[i])])
        uc, inv = np.unique(cols, return_inverse=True)
        acc = np.zeros(uc.size)
        np.add.at(acc, inv, prods)
        np.testing.assert_array_equal(ccol[crpt[i]:crpt[i + 1]], uc)
        np.testing.assert_allclose(cval[crpt[i]:crpt[i + 1]], acc.astype(np.float32), rtol=1e-6, atol=1e-6)


def test_allocation_failure_returns_oom(nsp, orc):
    """nsp_spgemm with an allocator that fails on its second request (C's
    columns): NSP_ERR_OOM, C zeroed, every buffer it did get handed back to
    the free callback, and the handle still computes the next product."""
    import ctypes as C
    import torch
    from nsp_amd import _lib
    print("case: allocation failure", flush=True)
    A = gen.random_csr(200, 150, 0.05, seed=71)
    B = gen.random_csr(150, 120, 0.05, seed=72)
    dA, dB = nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B)
    keep, calls = {}, [0]

    def _alloc(nbytes, user):
        calls[0] += 1
        if calls[0] == 2:
            return None
        t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")
        keep[t.data_ptr()] = t
        return t.data_ptr()

    def _free(ptr, user):
        if ptr:
            keep.pop(ptr)

    alloc, free = _lib.ALLOC_FN(_alloc), _lib.FREE_FN(_free)
    h = nsp.Handle()
    a, b, c = dA.c_struct(), dB.c_struct(), _lib.nsp_csr_t()
    status = h.L.nsp_spgemm(h.h, dA.index_type, dA.value_type, C.byref(a), C.byref(b), C.byref(c),
                            C.cast(alloc, C.c_void_p), C.cast(free, C.c_void_p), None)
    assert status == _lib.NSP_ERR_OOM
    assert calls[0] == 3 and not keep, (calls, keep)
    assert not c.row_ptr and not c.col and not c.val and c.nnz == 0
    check_against_oracle(A, B, h.spgemm(dA, dB), "f64", orc)
