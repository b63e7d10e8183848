"""Batched products (concatenated CSR) and the hipGraph-captured product,
against the oracle and against the plain per-matrix path.

Graph replays are checked one by one, with eager GPU work and host
synchronisations between them (a captured sequence must give the same result
on every replay, not only the first)."""
import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import check_against_oracle, to_dtype

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available()
    import nsp_amd
    return nsp_amd


def dev_off(off):
    import torch
    return torch.from_numpy(np.asarray(off, np.int64)).cuda()


def oracle_batch(orc, X, off):
    """Per-matrix oracle products stitched into the concatenated layout."""
    rpts, cols, vals, abss = [np.zeros(1, np.int64)], [], [], []
    base = 0
    for b in range(len(off) - 1):
        m = X.row_slice(int(off[b]), int(off[b + 1]))
        m.ncols = int(off[b + 1] - off[b])
        r, c, v, a = orc.spgemm(m, m)
        rpts.append(r[1:] + base)
        base += int(r[-1])
        cols.append(c)
        vals.append(v)
        abss.append(a)
    return (np.concatenate(rpts), np.concatenate(cols), np.concatenate(vals), np.concatenate(abss))


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batched_matches_oracle(nsp, orc, dtype):
    print("case: batched", dtype, flush=True)
    X, off = gen.uneven_batch_concat(48, seed=5, nmin=16, nmax=700)
    X = to_dtype(X, dtype)
    h = nsp.Handle()
    dX = nsp.DeviceCSR.from_host(X, dtype)
    d_off = dev_off(off)
    C = h.batched_spgemm(dX, d_off, dX, d_off)
    check_against_oracle(X, X, C, dtype, orc, ref=oracle_batch(orc, X, off))


def test_batched_equals_per_matrix_loop(nsp):
    print("case: batched vs loop", flush=True)
    X, off = gen.uneven_batch_concat(32, seed=9, nmin=64, nmax=400)
    X = to_dtype(X, "f32")
    h = nsp.Handle()
    dX = nsp.DeviceCSR.from_host(X, "f32")
    C = h.batched_spgemm(dX, dev_off(off), dX, dev_off(off))
    crpt = C.row_ptr.cpu().numpy()
    ccol = C.col.cpu().numpy()
    cval = C.val.cpu().numpy()
    for b in range(len(off) - 1):
        m = X.row_slice(int(off[b]), int(off[b + 1]))
        m.ncols = int(off[b + 1] - off[b])
        dm = nsp.DeviceCSR.from_host(m, "f32")
        Cb = h.spgemm(dm, dm)
        r0, r1 = int(off[b]), int(off[b + 1])
        s0, s1 = int(crpt[r0]), int(crpt[r1])
        np.testing.assert_array_equal(Cb.row_ptr.cpu().numpy(), crpt[r0:r1 + 1] - s0)
        np.testing.assert_array_equal(Cb.col.cpu().numpy(), ccol[s0:s1])
        np.testing.assert_allclose(Cb.val.cpu().numpy(), cval[s0:s1], rtol=1e-5, atol=1e-6)


def test_graph_single_product_replays(nsp, orc):
    print("case: graph single", flush=True)
    import torch
    A = gen.stencil_27pt(18, seed=4)
    h = nsp.Handle()
    dA = nsp.DeviceCSR.from_host(A)
    ref = orc.spgemm(A, A)
    g = h.graph(dA, dA, capacity=int(ref[0][-1]) + 100)
    for rep in range(3):
        g.launch()
        C = g.result()
        check_against_oracle(A, A, C, "f64", orc, ref=ref)
        # eager work between replays
        x = torch.randn(1000, device="cuda")
        (x * 2).sum().item()
        h.spgemm(dA, dA)
        torch.cuda.synchronize()


def test_graph_batched_replays_and_value_updates(nsp, orc):
    print("case: graph batched", flush=True)
    import torch
    X, off = gen.uneven_batch_concat(40, seed=12, nmin=32, nmax=300)
    X = to_dtype(X, "f32")
    h = nsp.Handle()
    dX = nsp.DeviceCSR.from_host(X, "f32")
    d_off = dev_off(off)
    ref = oracle_batch(orc, X, off)
    g = h.graph(dX, dX, capacity=int(ref[0][-1]), a_off=d_off, b_off=d_off)
    for rep in range(3):
        g.launch()
        C = g.result()
        check_against_oracle(X, X, C, "f32", orc, ref=ref)
        h.spgemm(nsp.DeviceCSR.from_host(gen.laplacian_5pt(20)), nsp.DeviceCSR.from_host(gen.laplacian_5pt(20)))
        torch.cuda.synchronize()
    # new values, same pattern: the same graph computes the new product
    X2 = gen.HostCSR(X.rpt, X.col, (-2.0 * X.val).astype(np.float32).astype(np.float64), X.nrows, X.ncols)
    dX.val.copy_(torch.from_numpy(X2.val).to(dX.val.dtype))
    g.launch()
    check_against_oracle(X2, X2, g.result(), "f32", orc, ref=oracle_batch(orc, X2, off))


def test_graph_capacity_overflow_reported(nsp):
    print("case: graph capacity", flush=True)
    A = gen.stencil_27pt(10, seed=1)
    h = nsp.Handle()
    dA = nsp.DeviceCSR.from_host(A)
    g = h.graph(dA, dA, capacity=1000)
    g.launch()
    with pytest.raises(nsp.NspError) as e:
        g.result()
    assert e.value.status == 3


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_batched_int64_rows_and_graph(nsp, orc, dtype):
    """int64 row pointers through the batched path and the graph."""
    print("case: batched int64", dtype, flush=True)
    import torch
    X, off = gen.uneven_batch_concat(20, seed=21, nmin=16, nmax=500)
    X = to_dtype(X, dtype)
    h = nsp.Handle()
    dX = nsp.DeviceCSR.from_host(X, dtype, "i64")
    d_off = dev_off(off)
    ref = oracle_batch(orc, X, off)
    C = h.batched_spgemm(dX, d_off, dX, d_off)
    assert C.row_ptr.dtype == torch.int64
    check_against_oracle(X, X, C, dtype, orc, ref=ref)
    g = h.graph(dX, dX, capacity=int(ref[0][-1]), a_off=d_off, b_off=d_off)
    for _ in range(2):
        g.launch()
        check_against_oracle(X, X, g.result(), dtype, orc, ref=ref)
        torch.cuda.synchronize()


def test_batched_with_empty_matrices_and_rows(nsp, orc):
    """Matrices with no rows, and rows with no entries, inside a batch."""
    print("case: batched empty members", flush=True)
    rng = np.random.default_rng(31)
    mats = []
    for b in range(12):
        n = 0 if b in (0, 5, 11) else int(rng.integers(8, 200))
        m = gen.random_csr(n, n, 0.05, seed=100 + b, empty_rows=0.3) if n else \
            gen.HostCSR(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0), 0, 0)
        mats.append(m)
    off = np.zeros(len(mats) + 1, np.int64)
    np.cumsum([m.nrows for m in mats], out=off[1:])
    rpt = [np.zeros(1, np.int64)]
    base = 0
    for m in mats:
        rpt.append(m.rpt[1:] + base)
        base += m.nnz
    X = gen.HostCSR(np.concatenate(rpt), np.concatenate([m.col for m in mats]).astype(np.int32),
                    np.concatenate([m.val for m in mats]), int(off[-1]), max(m.ncols for m in mats))
    h = nsp.Handle()
    dX = nsp.DeviceCSR.from_host(X)
    C = h.batched_spgemm(dX, dev_off(off), dX, dev_off(off))
    check_against_oracle(X, X, C, "f64", orc, ref=oracle_batch(orc, X, off))


def test_batched_full_size_4096(nsp, orc):
    """The batched workload at full size (4096 uneven FP32 products,
    8.58 M rows): one graph launch; 64 seeded matrices against the oracle;
    C x = A (A x) over every row of the block-diagonal batch (SpMV on the
    host in FP64, relative to |A| (|A| |x|), 1e-5 for FP32 values)."""
    import scipy.sparse as sp
    import torch
    from tests.helpers import TOL
    print("case: batched 4096 full size", flush=True)
    X, off = gen.uneven_batch_concat(4096, seed=0)
    X = to_dtype(X, "f32")
    d_off = dev_off(off)
    dX = nsp.DeviceCSR.from_host(X, "f32")
    h = nsp.Handle()
    _, nnz = h.batched_symbolic(dX, d_off, dX, d_off)
    g = h.graph(dX, dX, capacity=nnz, a_off=d_off, b_off=d_off)
    g.launch()
    C = g.result()
    rpt = C.row_ptr.cpu().numpy().astype(np.int64)
    col = C.col.cpu().numpy()
    val = C.val.cpu().numpy().astype(np.float64)
    # 64 seeded matrices against the oracle
    worst = 0.0
    for b in np.random.default_rng(3).choice(4096, 64, replace=False):
        r0, r1 = int(off[b]), int(off[b + 1])
        m = X.row_slice(r0, r1)
        m.ncols = r1 - r0
        crpt, ccol, cval, cabs = orc.spgemm(m, m)
        s0 = rpt[r0]
        np.testing.assert_array_equal(rpt[r0:r1 + 1] - s0, crpt)
        np.testing.assert_array_equal(col[s0:rpt[r1]], ccol)
        comp, _ = orc.componentwise_error(val[s0:rpt[r1]], cval, cabs)
        worst = max(worst, comp)
    assert worst <= TOL["f32"], worst
    # block-diagonal identity over all rows: global column = local + matrix offset
    mat_of_row = np.repeat(np.arange(4096), np.diff(off))
    shift_c = np.repeat(off[:-1][mat_of_row], np.diff(rpt))
    Cs = sp.csr_matrix((val, col.astype(np.int64) + shift_c, rpt), shape=(X.nrows, X.nrows))
    shift_a = np.repeat(off[:-1][mat_of_row], np.diff(X.rpt))
    As = sp.csr_matrix((X.val, X.col.astype(np.int64) + shift_a, X.rpt), shape=(X.nrows, X.nrows))
    x = np.random.default_rng(4).uniform(-1, 1, X.nrows)
    y1, y2 = Cs @ x, As @ (As @ x)
    scale = abs(As) @ (abs(As) @ np.abs(x))
    rel = float(np.max(np.abs(y1 - y2) / np.where(scale > 0, scale, 1.0)))
    print(f"  nnz {C.nnz}, 64 matrices err {worst:.2e}, C x vs A(A x) err {rel:.2e}", flush=True)
    assert rel <= TOL["f32"], rel
    del C, g
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_work_queue_launch_every_bin(nsp, orc, dtype):
    """The numeric bins 0..5 of a column-bitmap product in one launch fed by
    the row-bin work queues (queue_mode 1, the default) against the per-bin
    launches (queue_mode 0) and the oracle: rows in every numeric bin up to
    the block bins, several ticket sizes and grid sizes (1 block per CU: the
    blocks steal from the other XCDs' queues), eager and graph-captured."""
    print("case: work queue", dtype, flush=True)
    from tests.test_gpu_spgemm import mixed_rows
    A = to_dtype(mixed_rows(3000, 2500, 1, 70, seed=31), dtype)
    B = to_dtype(mixed_rows(2500, 4000, 1, 12, seed=32), dtype)
    ref = orc.spgemm(A, B)
    dA, dB = nsp.DeviceCSR.from_host(A, dtype), nsp.DeviceCSR.from_host(B, dtype)
    h0 = nsp.Handle()
    h0.set_param("queue_mode", 0)
    C0 = h0.spgemm(dA, dB)
    nb = h0.stats()["num_bin"]
    assert all(nb[b] > 0 for b in range(6)), nb
    check_against_oracle(A, B, C0, dtype, orc, ref=ref)
    for chunk, blocks in ((4, 4), (1, 1), (3, 2), (64, 4)):
        h = nsp.Handle()
        h.set_param("queue_chunk", chunk)
        h.set_param("queue_blocks", blocks)
        C = h.spgemm(dA, dB)
        assert h.stats()["num_launches"] == 1 + sum(nb[b] > 0 for b in range(6, 9))
        # bin 5 runs 64 lanes a row here and 32 per bin, so sums may differ in
        # their last bits: structure bit-exact, values against the oracle
        np.testing.assert_array_equal(C.row_ptr.cpu().numpy(), C0.row_ptr.cpu().numpy())
        np.testing.assert_array_equal(C.col.cpu().numpy(), C0.col.cpu().numpy())
        check_against_oracle(A, B, C, dtype, orc, ref=ref)
    g = nsp.Handle().graph(dA, dB, capacity=int(ref[0][-1]))
    for rep in range(3):
        g.launch()
        check_against_oracle(A, B, g.result(), dtype, orc, ref=ref)
