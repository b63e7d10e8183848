"""Device CSR from the reference's binary COO files (nsp_coo_read, streamed
from the file, and nsp_coo_to_csr, from records already on the device)
against the oracle's restatement of READ_BINARY_MATRIX_MPI + load_MatrixMPI
(src/READ_MATRIX_MPI.cpp:146-356, src/distribute.cu:998-1125): row pointer,
columns and values bit-exact (FP32 output = the FP64 value rounded once).

Cases: a stencil split over 4 ranks; a random file with empty rows inside
and at partition ends, rows whose columns are out of order (short, and
long enough for the merge passes of the row sort) and a repeated column
(the sort must be stable, as the reference's bubble sort is); both header
types; streamed chunks down to one record.  Also the error paths and the
CSR transpose whose long columns now go through the same merge sort."""
import os

import numpy as np
import pytest

from nsp_amd import gen

pytestmark = pytest.mark.gpu

REC = [("i", "<i8"), ("j", "<i8"), ("v", "<f8")]


def write_records(path, rec, nr, nc, header_type=1):
    with open(path, "wb") as f:
        f.write(np.array([nr, nc, len(rec)] if header_type else [nr, len(rec)], "<i8").tobytes())
        f.write(np.asarray(rec, dtype=REC).tobytes())


def records_of(X, row_order=None):
    rows = np.repeat(np.arange(X.nrows), np.diff(X.rpt)) + 1
    rec = np.zeros(X.nnz, dtype=REC)
    rec["i"], rec["j"], rec["v"] = rows, X.col.astype(np.int64) + 1, X.val
    return rec


def mixed_records(seed=5):
    """60 rows over 50 columns: empty rows 3, 14, 15, 29, 30, 44; every
    fifth row's columns shuffled; row 7 repeats its first column."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(60):
        if i in (3, 14, 15, 29, 30, 44):
            continue
        cols = list(rng.choice(50, int(rng.integers(1, 9)), replace=False) + 1)
        if i % 5 == 0:
            rng.shuffle(cols)
        else:
            cols.sort()
        if i == 7:
            cols = cols + [cols[0]]
        out += [(i + 1, int(c), float(rng.normal())) for c in cols]
    return np.array(out, dtype=REC), 60, 50


def long_rows_records(seed=9):
    """Rows of 5000, 9000 and 33000 shuffled columns (past the 4096-key LDS
    tile: one, two and four merge passes), repeated columns in the longest,
    and short sorted rows between them."""
    rng = np.random.default_rng(seed)
    nc, out = 40000, []
    for i, L in enumerate([3, 5000, 2, 9000, 7, 33000, 1]):
        cols = rng.choice(nc, L, replace=False) + 1
        if L > 100:
            rng.shuffle(cols)
            if L == 33000:
                cols[100:200] = cols[:100]   # repeats: stability decides their order
        else:
            cols.sort()
        out += [(i + 1, int(c), float(v)) for c, v in zip(cols, rng.normal(size=L))]
    return np.array(out, dtype=REC), 7, nc


@pytest.fixture(scope="module")
def h():
    import nsp_amd
    return nsp_amd.Handle()


def check(X, ref, dtype):
    np.testing.assert_array_equal(X.row_ptr.cpu().numpy().astype(np.int64), ref.rpt)
    np.testing.assert_array_equal(X.col.cpu().numpy(), ref.col)
    want = ref.val.astype(np.float32) if dtype == "f32" else ref.val
    np.testing.assert_array_equal(X.val.cpu().numpy(), want)


def case_file(tmp_path, case, header_type):
    if case == "stencil":
        X = gen.stencil_27pt(12, seed=3)
        rec, nr, nc = records_of(X), X.nrows, X.ncols
    elif case == "mixed":
        rec, nr, nc = mixed_records()
    else:
        rec, nr, nc = long_rows_records()
    p = str(tmp_path / f"{case}_{header_type}.bin")
    write_records(p, rec, nr, nc if header_type else nr, header_type)
    return p


@pytest.mark.parametrize("case", ["stencil", "mixed", "long_rows"])
@pytest.mark.parametrize("index,dtype", [("i32", "f64"), ("i64", "f32")])
def test_coo_read_matches_oracle(h, orc, tmp_path, case, index, dtype):
    for header_type in (1, 0):
        p = case_file(tmp_path, case, header_type)
        chunks = {"stencil": (1 << 21, 1000), "mixed": (1 << 21, 7, 1), "long_rows": (1 << 21, 4099)}[case]
        for nprocs in ((1, 4) if case == "stencil" else (1, 2, 3, 4)):
            for rank in range(nprocs):
                nr, nc, nt, shift, ref = orc.read_binary_coo(p, nprocs, rank, header_type)
                for chunk in chunks:
                    print(f"  {case} ht={header_type} {nprocs}:{rank} chunk={chunk} {index}/{dtype}", flush=True)
                    h.set_param("coo_chunk_records", chunk)
                    s, X = h.coo_read(p, nprocs, rank, header_type, index=index, dtype=dtype)
                    assert (s.nr, s.nc, s.nt, s.col_shift) == (nr, nc, nt, shift)
                    assert X.nrows == ref.nrows
                    check(X, ref, dtype)
    h.set_param("coo_chunk_records", 1 << 21)


def test_coo_to_csr_from_device_records(h, orc, tmp_path):
    import torch
    import nsp_amd
    for case in ("mixed", "long_rows"):
        p = case_file(tmp_path, case, 1)
        for nprocs, rank in ((1, 0), (3, 1), (3, 2)):
            s = nsp_amd.coo_share(p, nprocs, rank)
            raw = np.fromfile(p, dtype=np.uint8)[s.header_bytes + 24 * s.rec_begin: s.header_bytes + 24 * s.rec_end]
            _, _, _, _, ref = orc.read_binary_coo(p, nprocs, rank)
            # records 16-byte aligned (vector loads) and 8 bytes off (word loads)
            for pad in (0, 8):
                print(f"  {case} {nprocs}:{rank} pad={pad}", flush=True)
                buf = torch.zeros(raw.size + 16, dtype=torch.uint8, device="cuda")
                buf[pad:pad + raw.size] = torch.from_numpy(raw.copy()).cuda()
                rec = buf[pad:pad + raw.size]
                assert (rec.data_ptr() % 16 == 0) == (pad == 0)
                X = h.coo_to_csr(rec, s.first_row, s.nrows, s.col_shift, ncols=s.nc)
                check(X, ref, "f64")


def test_coo_errors_leave_handle_usable(h, orc, tmp_path):
    import nsp_amd
    from nsp_amd import NspError
    from nsp_amd._lib import NSP_ERR_INVALID, NSP_ERR_OVERFLOW
    bad = str(tmp_path / "order.bin")    # row 2 before row 1 inside one rank
    write_records(bad, [(2, 1, 1.0), (1, 1, 1.0), (2, 2, 1.0)], 2, 2)
    s = nsp_amd._lib.nsp_coo_share_t(2, 2, 3, 24, 0, 3, 0, 2, 0)  # the whole file as one share
    print("  rows out of order", flush=True)
    with pytest.raises(NspError) as e:
        h.coo_read(bad, share=s)
    assert e.value.status == NSP_ERR_INVALID
    over = str(tmp_path / "over.bin")
    write_records(over, [(1, 1, 1.0), (2, 1 << 40, 1.0)], 2, 1 << 41)
    print("  column overflow", flush=True)
    with pytest.raises(NspError) as e:
        h.coo_read(over)
    assert e.value.status == NSP_ERR_OVERFLOW
    good = case_file(tmp_path, "mixed", 1)
    _, X = h.coo_read(good, 2, 1)
    check(X, orc.read_binary_coo(good, 2, 1)[4], "f64")


def test_empty_share(h, tmp_path):
    p = str(tmp_path / "e.bin")
    write_records(p, [(i + 1, 1, 1.0) for i in (0, 1, 4, 5, 6, 7)], 8, 8)
    s, X = h.coo_read(p, 4, 1)
    assert s.nrows == 0 and X.nrows == 0 and X.nnz == 0
    assert X.row_ptr.cpu().tolist() == [0]


@pytest.mark.parametrize("index,dtype", [("i32", "f64"), ("i64", "f32")])
def test_transpose_long_columns(h, index, dtype):
    """Columns of 5000, 9000 and 20000 entries (X^T rows past the LDS sort)
    next to a random pattern: X^T equals the CPU transpose exactly."""
    import nsp_amd
    rng = np.random.default_rng(4)
    n, m = 30000, 64
    R = gen.random_csr(n, m, 0.02, seed=11)
    dense_cols = {5: 5000, 17: 9000, 40: 20000}
    rows, cols, vals = [np.repeat(np.arange(n), np.diff(R.rpt))], [R.col.astype(np.int64)], [R.val]
    for c, L in dense_cols.items():
        rr = rng.choice(n, L, replace=False)
        rows.append(rr)
        cols.append(np.full(L, c))
        vals.append(rng.normal(size=L))
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    key, idx = np.unique(r * m + c, return_index=True)
    X = gen._from_rows_cols(n, m, key // m, key % m, v[idx])
    Xd = nsp_amd.DeviceCSR.from_host(X, dtype=dtype, index=index)
    print(f"  transpose {index}/{dtype}, longest column {max(np.bincount(X.col))}", flush=True)
    XT = h.transpose(Xd)
    order = np.lexsort((np.repeat(np.arange(n), np.diff(X.rpt)), X.col))
    trpt = np.concatenate([[0], np.cumsum(np.bincount(X.col, minlength=m))])
    np.testing.assert_array_equal(XT.row_ptr.cpu().numpy().astype(np.int64), trpt)
    np.testing.assert_array_equal(XT.col.cpu().numpy(), np.repeat(np.arange(n), np.diff(X.rpt))[order])
    want = X.val[order].astype(np.float32) if dtype == "f32" else X.val[order]
    np.testing.assert_array_equal(XT.val.cpu().numpy(), want)


# ---------------------------------------------------------------- ASCII files
def ascii_case(tmp_path, case, value_fmt):
    """Text files of the reference's ASCII format (fixed-length record
    lines): the 12^3 stencil, the mixed records (empty rows, unsorted rows,
    a repeated column) and the long shuffled rows, values written with the
    given format (17 significant digits, 7 in exponent form, 3 decimals)."""
    from tests.helpers import write_ascii_coo
    if case == "stencil":
        X = gen.stencil_27pt(12, seed=3)
        rec, nr, nc = records_of(X), X.nrows, X.ncols
    elif case == "mixed":
        rec, nr, nc = mixed_records()
    else:
        rec, nr, nc = long_rows_records()
    p = str(tmp_path / f"{case}.txt")
    write_ascii_coo(p, records=[(int(i), int(j), float(v)) for i, j, v in rec], nr=nr, nc=nc, value_fmt=value_fmt,
                    comments=["% written by the tests"] if case == "mixed" else ())
    return p


@pytest.mark.parametrize("case", ["stencil", "mixed", "long_rows"])
@pytest.mark.parametrize("value_fmt", ["{:.17g}", "{:.6e}", "{:+.3f}"])
@pytest.mark.parametrize("index,dtype", [("i32", "f64"), ("i64", "f32")])
def test_ascii_read_matches_oracle(h, orc, tmp_path, case, value_fmt, index, dtype):
    """nsp_ascii_read against the oracle's restatement of
    READ_ASCII_MATRIX_MPI + load_MatrixMPI: row pointer and columns
    bit-exact, values bit-exact (the device parses with the reference's
    fast_atof steps), for 1-4 ranks and streamed chunks down to one line."""
    import nsp_amd
    p = ascii_case(tmp_path, case, value_fmt)
    chunks = {"stencil": (1 << 21, 1000), "mixed": (1 << 21, 7, 1), "long_rows": (1 << 21, 4099)}[case]
    for nprocs in ((1, 4) if case == "stencil" else (1, 2, 3, 4)):
        for rank in range(nprocs):
            nr, nc, nt, shift, ref = orc.read_ascii_coo(p, nprocs, rank)
            s = nsp_amd.ascii_share(p, nprocs, rank)
            assert (s.nr, s.nc, s.nt, s.col_shift) == (nr, nc, nt, shift)
            for chunk in chunks:
                print(f"  ascii {case} {value_fmt} {nprocs}:{rank} chunk={chunk} {index}/{dtype}", flush=True)
                h.set_param("coo_chunk_records", chunk)
                _, X = h.ascii_read(p, nprocs, rank, index=index, dtype=dtype, share=s)
                assert X.nrows == ref.nrows
                check(X, ref, dtype)
    h.set_param("coo_chunk_records", 1 << 21)


def test_ascii_header_forms_and_errors(h, orc, tmp_path):
    """"nr nt" headers, '%' lines and a last line without its newline read
    as the oracle reads them; lines of unequal length and rows out of order
    return NSP_ERR_INVALID and leave the handle usable."""
    import nsp_amd
    from nsp_amd import NspError
    from nsp_amd._lib import NSP_ERR_INVALID
    from tests.helpers import write_ascii_coo
    p = str(tmp_path / "h.txt")
    write_ascii_coo(p, records=[(1, 3, 1.5), (1, 1, 2.0), (1, 2, -0.5), (3, 3, 4.0)], nr=3, nc=3, header_fields=2,
                    comments=["%%MatrixMarket matrix coordinate real general", "%"], value_fmt="{:g}",
                    final_newline=False)
    _, _, _, _, ref = orc.read_ascii_coo(p)
    s, X = h.ascii_read(p)
    assert (s.nr, s.nc, s.nt) == (3, 3, 4)
    check(X, ref, "f64")
    bad = str(tmp_path / "unequal.txt")
    with open(bad, "w") as f:
        f.write("2 2 2\n 1 1 1.0\n 2 2 10.25\n 1 1 1.0\n")
    print("  unequal lines", flush=True)
    with pytest.raises(NspError) as e:
        h.ascii_read(bad)
    assert e.value.status == NSP_ERR_INVALID
    order = str(tmp_path / "order.txt")
    write_ascii_coo(order, records=[(2, 1, 1.0), (1, 1, 1.0), (2, 2, 1.0)], nr=2, nc=2)
    s = nsp_amd.ascii_share(order)
    s.rec_begin, s.rec_end, s.first_row, s.nrows = 0, 3, 0, 2   # the whole file as one share
    print("  rows out of order", flush=True)
    with pytest.raises(NspError) as e:
        h.ascii_read(order, share=s)
    assert e.value.status == NSP_ERR_INVALID
    good = ascii_case(tmp_path, "mixed", "{:.17g}")
    _, X = h.ascii_read(good, 2, 1)
    check(X, orc.read_ascii_coo(good, 2, 1)[4], "f64")
