"""nsp_coo_share (host only: header, rank range, col_shift) against the
oracle's reader and against hand-computed cases of the reference's rules
(src/READ_MATRIX_MPI.cpp:83-130, 256-270; src/distribute.cu:1094-1105).
No GPU: the function reads the file and does no device work."""
import os

import numpy as np
import pytest

from nsp_amd import gen

REC = [("i", "<i8"), ("j", "<i8"), ("v", "<f8")]


def write_records(path, recs, nr, nc, header_type=1):
    with open(path, "wb") as f:
        f.write(np.array([nr, nc, len(recs)] if header_type else [nr, len(recs)], "<i8").tobytes())
        f.write(np.array(recs, dtype=REC).tobytes())


def share(path, nprocs, rank, header_type=1):
    from nsp_amd import coo_share
    return coo_share(path, nprocs, rank, header_type)


@pytest.fixture(scope="module")
def lib_built():
    from nsp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libnsp_amd.so is not built (run python -m nsp_amd.build)")


def test_exscan_shift_with_empty_row_at_partition_end(lib_built, tmp_path):
    """Rows 0-2 and 4-7 hold entries, row 3 is empty.  With 2 ranks, rank 0
    owns rows [0, 4) but spans rows 0..2 (rowxproc 3), so rank 1's columns
    are shifted by 3, not by its first row 4 (MPI_Exscan of rowxproc)."""
    recs = [(i + 1, i + 1, 1.0) for i in (0, 1, 2, 4, 5, 6, 7)]
    p = str(tmp_path / "q.bin")
    write_records(p, recs, 8, 8)
    s0, s1 = share(p, 2, 0), share(p, 2, 1)
    assert (s0.rec_begin, s0.rec_end, s0.first_row, s0.nrows, s0.col_shift) == (0, 3, 0, 3, 0)
    assert (s1.rec_begin, s1.rec_end, s1.first_row, s1.nrows, s1.col_shift) == (3, 7, 4, 4, 3)


def test_rank_without_records(lib_built, tmp_path):
    """Rows 2 and 3 are empty: rank 1 of 4 (rows [2, 4)) has no records and
    spans no rows; rank 2's shift counts only the rows rank 0 spans."""
    recs = [(i + 1, 1, 1.0) for i in (0, 1, 4, 5, 6, 7)]
    p = str(tmp_path / "e.bin")
    write_records(p, recs, 8, 8)
    s = [share(p, 4, r) for r in range(4)]
    assert [x.nrows for x in s] == [2, 0, 2, 2]
    assert [x.col_shift for x in s] == [0, 2, 2, 4]
    assert (s[1].rec_begin, s[1].rec_end) == (2, 2)


@pytest.mark.parametrize("header_type", [1, 0])
def test_share_matches_oracle(lib_built, orc, tmp_path, header_type):
    X = gen.random_csr(97, 97, 0.05, seed=7, empty_rows=0.25)
    p = str(tmp_path / "r.bin")
    rows = np.repeat(np.arange(X.nrows), np.diff(X.rpt)) + 1
    write_records(p, list(zip(rows.tolist(), (X.col.astype(np.int64) + 1).tolist(), X.val.tolist())),
                  X.nrows, X.ncols, header_type)
    for nprocs in (1, 2, 3, 5, 8):
        for rank in range(nprocs):
            s = share(p, nprocs, rank, header_type)
            nr, nc, nt, shift, ref = orc.read_binary_coo(p, nprocs, rank, header_type)
            assert (s.nr, s.nc, s.nt) == (nr, nc, nt)
            assert s.header_bytes == (24 if header_type else 16)
            assert s.col_shift == shift and s.nrows == ref.nrows
            assert s.rec_end - s.rec_begin == ref.nnz


def test_share_rejects_bad_files(lib_built, tmp_path):
    from nsp_amd import NspError
    p = str(tmp_path / "t.bin")
    write_records(p, [(1, 1, 1.0), (2, 2, 2.0)], 2, 2)
    with open(p, "r+b") as f:  # header claims more records than the file holds
        f.write(np.array([2, 2, 5], "<i8").tobytes())
    with pytest.raises(NspError):
        share(p, 1, 0)
    with pytest.raises(NspError):
        share(str(tmp_path / "missing.bin"), 1, 0)
    write_records(p, [(1, 1, 1.0)], 1, 1)
    with pytest.raises(NspError):
        share(p, 2, 2)
