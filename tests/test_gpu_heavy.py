"""Heavy rows: rows of more than 16384 intermediate products or more than
4096 entries in C (LDS bitmap windows, one block per row), against the
oracle; R-MAT power-law products with sampled-row checks.

Tolerances as everywhere: structure bit-exact, values 1e-12 (FP64) / 1e-5
(FP32) relative to sum_k |a_ik b_kj|."""
import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import TOL, check_against_oracle, to_dtype

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available()
    import nsp_amd
    return nsp_amd


def rows_matrix(lens, ncols, seed, spread=None):
    """Rows with the given lengths, columns uniform in [0, ncols) (or in a
    band of width `spread` around a random centre)."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i, k in enumerate(lens):
        k = int(min(k, ncols))
        if spread is None:
            c = rng.choice(ncols, k, replace=False)
        else:
            c0 = int(rng.integers(0, max(1, ncols - spread)))
            c = c0 + rng.choice(min(spread, ncols - c0), k, replace=False)
        rows.append(np.full(k, i))
        cols.append(c)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    return gen._from_rows_cols(len(lens), ncols, rows, cols, rng.uniform(-1, 1, rows.size))


@pytest.mark.parametrize("dtype,index", [("f64", "i32"), ("f32", "i32"), ("f64", "i64")])
def test_heavy_rows_one_window(nsp, orc, dtype, index):
    """A few rows of 300-600 entries over B rows of ~80 entries (24K-48K
    products, C rows of 10K-30K entries) among ordinary rows."""
    print("case: heavy one window", dtype, index, flush=True)
    rng = np.random.default_rng(1)
    lens = rng.integers(1, 30, 400)
    lens[[3, 77, 200, 399]] = [300, 450, 600, 350]
    A = to_dtype(rows_matrix(lens, 5000, seed=2), dtype)
    B = to_dtype(rows_matrix(rng.integers(40, 120, 5000), 60000, seed=3), dtype)
    h = nsp.Handle()
    C = h.spgemm(nsp.DeviceCSR.from_host(A, dtype, index), nsp.DeviceCSR.from_host(B, dtype, index))
    st = h.stats()
    assert st["sym_bin"][8] >= 4 and st["num_bin"][8] >= 4, (st["sym_bin"], st["num_bin"])
    check_against_oracle(A, B, C, dtype, orc)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_heavy_rows_many_windows(nsp, orc, dtype):
    """B has 3.2M columns and its rows are spread over all of them, so heavy
    C rows span several 2^20-column windows (symbolic and numeric); one
    row's products all land in one narrow band."""
    print("case: heavy many windows", dtype, flush=True)
    rng = np.random.default_rng(4)
    ncols = 3_200_000
    lens = rng.integers(1, 10, 200)
    lens[[5, 50, 150]] = [400, 700, 500]
    A = to_dtype(rows_matrix(lens, 3000, seed=5), dtype)
    Bw = rows_matrix(rng.integers(30, 60, 3000), ncols, seed=6)
    # rows 0..99 of B confined to a 100-column band: products collide heavily
    Bb = rows_matrix(rng.integers(30, 60, 100), ncols, seed=7, spread=100)
    rpt = np.concatenate([Bb.rpt, Bb.rpt[-1] + Bw.rpt[101:] - Bw.rpt[100]])
    B = gen.HostCSR(rpt, np.concatenate([Bb.col, Bw.col[Bw.rpt[100]:]]),
                    np.concatenate([Bb.val, Bw.val[Bw.rpt[100]:]]), 3000, ncols)
    B = to_dtype(B, dtype)
    h = nsp.Handle()
    C = h.spgemm(nsp.DeviceCSR.from_host(A, dtype), nsp.DeviceCSR.from_host(B, dtype))
    st = h.stats()
    assert st["sym_bin"][8] >= 3 and st["num_bin"][8] >= 3, (st["sym_bin"], st["num_bin"])
    check_against_oracle(A, B, C, dtype, orc)


def test_heavy_rows_handle_reuse(nsp, orc):
    """Heavy-row products on one handle, with different column counts and
    heavy-row counts in between, all match the oracle."""
    print("case: heavy handle reuse", flush=True)
    rng = np.random.default_rng(12)
    lens = rng.integers(1, 20, 300)
    lens[[1, 2, 3, 200]] = [450, 380, 520, 300]
    A1 = rows_matrix(lens, 4000, seed=13)
    B1 = rows_matrix(rng.integers(40, 100, 4000), 70000, seed=14)
    lens2 = rng.integers(1, 10, 100)
    lens2[[9]] = [700]
    A2 = rows_matrix(lens2, 3000, seed=15)
    B2 = rows_matrix(rng.integers(30, 80, 3000), 120000, seed=16)
    h = nsp.Handle()
    for A, B in ((A1, B1), (A2, B2), (A1, B1), (A2, B2)):
        C = h.spgemm(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
        assert h.stats()["num_bin"][8] >= 1
        check_against_oracle(A, B, C, "f64", orc)


def test_heavy_rows_in_graph(nsp, orc):
    print("case: heavy graph", flush=True)
    import torch
    rng = np.random.default_rng(8)
    lens = rng.integers(1, 20, 300)
    lens[[10, 20]] = [500, 400]
    A = rows_matrix(lens, 4000, seed=9)
    B = rows_matrix(rng.integers(40, 100, 4000), 80000, seed=10)
    ref = orc.spgemm(A, B)
    h = nsp.Handle()
    g = h.graph(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B), capacity=int(ref[0][-1]))
    for _ in range(2):
        g.launch()
        check_against_oracle(A, B, g.result(), "f64", orc, ref=ref)
        torch.cuda.synchronize()


def test_block_bins_in_graph_more_rows_than_grid(nsp, orc):
    """The graph launches the block-per-row bins on the grid a CU holds at
    once (2048 symbolic blocks for bin 6, 512 for bin 7 on 256 CUs; 512 and
    256 numeric): here those bins hold more rows than that, so each block
    walks several rows, and the product must still match the oracle on
    every replay."""
    print("case: block bins in graph", flush=True)
    import torch
    rng = np.random.default_rng(31)
    lens = np.concatenate([rng.integers(16, 25, 2600), rng.integers(30, 46, 1200)])
    A = rows_matrix(lens, 4000, seed=32)
    B = rows_matrix(np.full(4000, 80), 400000, seed=33)
    ref = orc.spgemm(A, B)
    h = nsp.Handle()
    dA, dB = nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B)
    h.spgemm(dA, dB)
    st = h.stats()
    assert st["sym_bin"][6] > 2048 and st["sym_bin"][7] > 512, st["sym_bin"]
    assert st["num_bin"][6] > 512 and st["num_bin"][7] > 256, st["num_bin"]
    g = h.graph(dA, dB, capacity=int(ref[0][-1]))
    for _ in range(2):
        g.launch()
        check_against_oracle(A, B, g.result(), "f64", orc, ref=ref)
        (torch.randn(1000, device="cuda") * 2).sum().item()
        torch.cuda.synchronize()


def check_sampled_rows(nsp, orc, A, B, C, rows, dtype):
    """Rows `rows` of C (device) against the oracle's product of those rows."""
    import torch
    rpt = C.row_ptr.to(torch.int64)
    r = torch.from_numpy(np.asarray(rows, np.int64)).cuda()
    s, e = rpt[r].cpu().numpy(), rpt[r + 1].cpu().numpy()
    ref_rpt, ref_col, ref_val, ref_abs = orc.spgemm_rows(A, B, np.asarray(rows, np.int64), with_abs=True)
    worst = 0.0
    for q, i in enumerate(rows):
        col = C.col[int(s[q]):int(e[q])].cpu().numpy()
        val = C.val[int(s[q]):int(e[q])].cpu().numpy().astype(np.float64)
        a, b = int(ref_rpt[q]), int(ref_rpt[q + 1])
        np.testing.assert_array_equal(col, ref_col[a:b], err_msg=f"row {i}")
        den = np.where(ref_abs[a:b] > 0, ref_abs[a:b], 1.0)
        if b > a:
            worst = max(worst, float(np.max(np.abs(val - ref_val[a:b]) / den)))
    assert worst <= TOL[dtype], worst
    return worst


def test_rmat_scale16_sampled(nsp, orc):
    """R-MAT scale 16 A*A in FP64: every row checked against the oracle."""
    print("case: rmat 16", flush=True)
    A = gen.rmat(16, seed=12345)
    h = nsp.Handle()
    dA = nsp.DeviceCSR.from_host(A, "f64", "i64")
    C = h.spgemm(dA, dA)
    st = h.stats()
    assert st["sym_bin"][8] > 0
    check_against_oracle(A, A, C, "f64", orc)


def test_rmat_scale20_sampled(nsp, orc):
    """R-MAT scale 20 A*A in FP64 (nnz(C) ~ 9.7e9, int64 row pointers): the
    whole row pointer (1,048,577 entries) and nnz(C) bit-exact against the
    exclusive scan of the oracle's per-row symbolic counts, the heaviest rows
    and a seeded sample against the oracle's values, every row strictly
    increasing."""
    print("case: rmat 20", flush=True)
    import torch
    A = gen.rmat(20, seed=12345)
    ip, nprod = orc.intprod(A, A)
    h = nsp.Handle()
    dA = nsp.DeviceCSR.from_host(A, "f64", "i64")
    C = h.spgemm(dA, dA)
    st = h.stats()
    print(f"  nprod {st['nprod']} nnz {st['nnz']} heavy rows sym {st['sym_bin'][8]} num {st['num_bin'][8]}",
          flush=True)
    assert st["nprod"] == nprod and st["nnz"] > 2**31
    ref_rpt = orc.exscan(orc.symbolic(A, A))
    got_rpt = C.row_ptr.cpu().numpy()
    assert got_rpt.shape == ref_rpt.shape
    np.testing.assert_array_equal(got_rpt, ref_rpt, err_msg="row_ptr differs from the oracle's counts")
    assert C.nnz == int(ref_rpt[-1]) == st["nnz"]
    print(f"  row_ptr of all {A.nrows} rows bit-exact, nnz(C) {C.nnz}", flush=True)
    rng = np.random.default_rng(0)
    heavy = np.argsort(ip)[-6:]
    sample = np.unique(np.concatenate([heavy, rng.choice(A.nrows, 60, replace=False)]))
    worst = check_sampled_rows(nsp, orc, A, A, C, sample, "f64")
    print(f"  sampled rows {sample.size}, worst componentwise error {worst:.2e}", flush=True)
    # strictly increasing columns in every row, in chunks to bound memory
    rpt = C.row_ptr
    start = torch.zeros(C.nnz, dtype=torch.bool, device="cuda")
    start[rpt[:-1][rpt[:-1] < rpt[1:]]] = True
    chunk = 1 << 30
    for p in range(1, C.nnz, chunk):
        q = min(C.nnz, p + chunk)
        ok = (C.col[p:q] > C.col[p - 1:q - 1]) | start[p:q]
        assert bool(ok.all()), f"unsorted entries in [{p}, {q})"


def test_heavy_rows_fp32_int64_graph(nsp, orc):
    """Heavy rows with FP32 values and int64 row pointers, through the graph."""
    print("case: heavy f32 i64 graph", flush=True)
    import torch
    from tests.helpers import to_dtype
    rng = np.random.default_rng(41)
    lens = rng.integers(1, 12, 300)
    lens[[7, 150]] = [520, 610]
    A = to_dtype(rows_matrix(lens, 4000, seed=42), "f32")
    B = to_dtype(rows_matrix(rng.integers(40, 90, 4000), 90000, seed=43), "f32")
    ref = orc.spgemm(A, B)
    h = nsp.Handle()
    g = h.graph(nsp.DeviceCSR.from_host(A, "f32", "i64"), nsp.DeviceCSR.from_host(B, "f32", "i64"),
                capacity=int(ref[0][-1]))
    for _ in range(2):
        g.launch()
        C = g.result()
        assert C.row_ptr.dtype == torch.int64
        check_against_oracle(A, B, C, "f32", orc, ref=ref)
