"""The C++ host API's binary COO reader/splitter (nsp_amd_run split) against
the oracle's restatement of READ_MATRIX_MPI.cpp / distribute.cu: same row
block, row shift, rank-frame columns and values for every rank.  Host only."""
import json
import os
import subprocess

import numpy as np
import pytest

from nsp_amd import gen

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nsp_amd", "nsp_amd_run")


def run_split(path, nprocs, rank, out):
    r = subprocess.run([DRIVER, "split", path, str(nprocs), str(rank), out], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="nsp_amd_run not built")
@pytest.mark.parametrize("case", ["stencil", "random_empty_rows", "unsorted_rows"])
def test_split_matches_oracle_reader(orc, tmp_path, case):
    if case == "stencil":
        X = gen.stencil_27pt(7, seed=1)
    else:
        X = gen.random_csr(300, 300, 0.03, seed=2, empty_rows=0.2)
    path = str(tmp_path / "a.bin")
    if case == "unsorted_rows":
        # reversed columns inside every row: the reader sorts them back
        with open(path, "wb") as f:
            np.array([X.nrows, X.ncols, X.nnz], np.int64).tofile(f)
            rec = np.zeros(X.nnz, dtype=[("i", "<i8"), ("j", "<i8"), ("v", "<f8")])
            q = 0
            for i in range(X.nrows):
                for p in range(X.rpt[i + 1] - 1, X.rpt[i] - 1, -1):
                    rec[q] = (i + 1, X.col[p] + 1, X.val[p])
                    q += 1
            rec.tofile(f)
    else:
        orc.write_binary_coo(path, X)
    for nprocs in (1, 3, 4):
        for rank in range(nprocs):
            out = str(tmp_path / f"b{nprocs}_{rank}.bin")
            meta = run_split(path, nprocs, rank, out)
            nr, nc, nt, shift, ref = orc.read_binary_coo(path, nprocs, rank)
            assert (meta["nr"], meta["nc"], meta["nt"], meta["row_shift"]) == (nr, nc, nt, shift)
            assert (meta["rows"], meta["nnz"]) == (ref.nrows, ref.nnz)
            raw = np.fromfile(out, dtype=np.int64, count=3)
            rec = np.fromfile(out, dtype=[("i", "<i8"), ("j", "<i8"), ("v", "<f8")], offset=24)
            assert tuple(raw) == (ref.nrows, nc, ref.nnz)
            rows = np.repeat(np.arange(ref.nrows), np.diff(ref.rpt)) + 1
            np.testing.assert_array_equal(rec["i"], rows)
            np.testing.assert_array_equal(rec["j"], ref.col.astype(np.int64) + 1)
            np.testing.assert_array_equal(rec["v"], ref.val)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="nsp_amd_run not built")
def test_matrix_market_writer_formats(orc, tmp_path):
    """CSRMatrixPrintMM writes the reference's header and line formats
    (src/matrixIO.cu:328-336: "%d  %d %d \\n", "%d   %d  %lg\\n", 1-based)."""
    X = gen.random_csr(40, 55, 0.08, seed=12, empty_rows=0.2)
    X.val[::3] *= 1e-7          # exponents in %g
    X.val[1::5] = np.round(X.val[1::5] * 100)   # integral values
    path, out = str(tmp_path / "a.bin"), str(tmp_path / "a.mtx")
    orc.write_binary_coo(path, X)
    r = subprocess.run([DRIVER, "mm", path, out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = ["%%MatrixMarket matrix coordinate real general\n", "%d  %d %d \n" % (X.nrows, X.ncols, X.nnz)]
    for i in range(X.nrows):
        for p in range(X.rpt[i], X.rpt[i + 1]):
            lines.append("%d   %d  %g\n" % (i + 1, X.col[p] + 1, X.val[p]))
    assert open(out).read() == "".join(lines)
