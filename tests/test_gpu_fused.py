"""GPU parity of the one-call product (nsp_spgemm_into, nsp_spgemm).

When a sample of rows fits a wave-level numeric bin, the call runs the
numeric pass once into per-row slots, scans the row lengths into C's row
pointer and copies the slots into C; otherwise, or when a row outgrows its
slot, it runs the two passes.  Either way C must be the oracle's: row
pointer and columns bit-exact, values within 1e-12 (FP64) / 1e-5 (FP32).
Every call goes through the C ABI of libnsp_amd.so.
"""
import ctypes as C

import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import check_against_oracle, to_dtype
from tests.test_gpu_spgemm import mixed_rows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    import nsp_amd
    return nsp_amd


@pytest.mark.parametrize("dtype,index", [("f64", "i32"), ("f64", "i64"), ("f32", "i32"), ("f32", "i64")])
def test_stencil_one_pass(nsp, orc, dtype, index):
    """27-point stencil (rows of up to 125 columns): one numeric pass into
    128-entry slots; the same C as the two-phase product."""
    print("case: fused stencil", dtype, index, flush=True)
    S = to_dtype(gen.stencil_27pt(24, seed=41), dtype)
    ref = orc.spgemm(S, S)
    h = nsp.Handle()
    dS = nsp.DeviceCSR.from_host(S, dtype, index)
    C = h.spgemm_into(dS, dS)
    st = h.stats()
    assert st["fused"] == 1 and st["sym_launches"] == 0, st
    check_against_oracle(S, S, C, dtype, orc, ref=ref)
    C2 = h.spgemm(dS, dS)
    assert h.stats()["fused"] == 0
    np.testing.assert_array_equal(C.row_ptr.cpu().numpy(), C2.row_ptr.cpu().numpy())
    np.testing.assert_array_equal(C.col.cpu().numpy(), C2.col.cpu().numpy())
    np.testing.assert_array_equal(C.val.cpu().numpy(), C2.val.cpu().numpy())


@pytest.mark.parametrize("rpw", [1, 2, 4])
def test_random_rows_each_bin_size(nsp, orc, rpw):
    """Random rows whose longest C rows fall in each wave-level bin, with 1,
    2 or 4 rows per wavefront: the slot size follows the sample."""
    print("case: fused random", rpw, flush=True)
    for seed, hi in ((51, 6), (52, 12), (53, 25), (54, 45)):
        A = mixed_rows(3000, 2500, 1, hi, seed=seed)
        B = mixed_rows(2500, 60_000, 1, 10, seed=seed + 100)
        ref = orc.spgemm(A, B)
        h = nsp.Handle()
        for name in ("num_small_rows_per_wave", "num_rows_per_wave"):
            h.set_param(name, rpw)
        C = h.spgemm_into(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
        st = h.stats()
        check_against_oracle(A, B, C, "f64", orc, ref=ref)
        if st["fused"]:
            assert st["num_launches"] == 1


def test_row_outgrowing_its_slot_falls_back(nsp, orc):
    """The sample takes one row in each block of eight (row 0 in the first
    block); row 1 (never sampled) has far more columns than the sampled rows
    and more products than the largest wave-level bin, so it outgrows the
    slot: the product is redone in two passes and still matches the oracle."""
    print("case: fused fallback", flush=True)
    rng = np.random.default_rng(61)
    n = 8192
    lens = np.full(n, 2)
    lens[1] = 60
    A = gen._from_rows_cols(n, 5000, np.repeat(np.arange(n), lens),
                            np.concatenate([rng.choice(5000, k, replace=False) for k in lens]),
                            rng.uniform(-1, 1, lens.sum()))
    B = mixed_rows(5000, 100_000, 8, 12, seed=62)
    ref = orc.spgemm(A, B)
    h = nsp.Handle()
    C = h.spgemm_into(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
    assert h.stats()["fused"] == 0
    check_against_oracle(A, B, C, "f64", orc, ref=ref)


def test_not_taken_for_bitmap_and_block_rows_and_when_off(nsp, orc):
    """B of at most 4096 columns (column-bitmap kernels), rows past 1024
    products, and fused_mode 0 take two passes; all match the oracle."""
    print("case: fused not taken", flush=True)
    A = mixed_rows(2000, 1500, 1, 30, seed=71)
    Bn = mixed_rows(1500, 3000, 1, 20, seed=72)        # narrow B
    Bh = mixed_rows(1500, 90_000, 40, 80, seed=73)     # rows of A past 1024 products
    Bw = mixed_rows(1500, 90_000, 1, 10, seed=74)
    for B, mode in ((Bn, 1), (Bh, 1), (Bw, 0)):
        ref = orc.spgemm(A, B)
        h = nsp.Handle()
        h.set_param("fused_mode", mode)
        C = h.spgemm_into(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
        assert h.stats()["fused"] == 0
        check_against_oracle(A, B, C, "f64", orc, ref=ref)


def test_capacity_too_small_reports_size(nsp, orc):
    """Through the C ABI: a capacity below nnz(C) returns NSP_ERR_OVERFLOW
    with nnz(C) and the row pointer filled; a second call with room fills C."""
    print("case: fused capacity", flush=True)
    import torch
    from nsp_amd import _lib
    S = gen.stencil_27pt(16, seed=81)
    ref = orc.spgemm(S, S)
    nnz_ref = int(ref[0][-1])
    h = nsp.Handle()
    dS = nsp.DeviceCSR.from_host(S)
    rpt = torch.empty(S.nrows + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(nnz_ref, dtype=torch.int32, device="cuda")
    val = torch.empty(nnz_ref, dtype=torch.float64, device="cuda")
    a = dS.c_struct()
    nnz = C.c_int64()
    c = _lib.nsp_csr_t(S.nrows, S.ncols, 1000, rpt.data_ptr(), col.data_ptr(), val.data_ptr())
    st = h.L.nsp_spgemm_into(h.h, 0, 0, C.byref(a), C.byref(a), C.byref(c), C.byref(nnz))
    assert st == _lib.NSP_ERR_OVERFLOW and nnz.value == nnz_ref
    np.testing.assert_array_equal(rpt.cpu().numpy().astype(np.int64), ref[0])
    c = _lib.nsp_csr_t(S.nrows, S.ncols, nnz_ref, rpt.data_ptr(), col.data_ptr(), val.data_ptr())
    assert h.L.nsp_spgemm_into(h.h, 0, 0, C.byref(a), C.byref(a), C.byref(c), C.byref(nnz)) == _lib.NSP_OK
    check_against_oracle(S, S, nsp.DeviceCSR(rpt, col, val, S.nrows, S.ncols), "f64", orc, ref=ref)


def test_repeated_calls_new_values_and_mixed_paths(nsp, orc):
    """One handle: one-call products with new values each time, a two-phase
    product and an out-of-slot product in between; every result is the
    oracle's for the values of that call."""
    print("case: fused repeated", flush=True)
    import torch
    S = gen.stencil_27pt(20, seed=91)
    A = mixed_rows(2000, 1500, 1, 20, seed=92)
    B = mixed_rows(1500, 50_000, 1, 10, seed=93)
    h = nsp.Handle()
    dS = nsp.DeviceCSR.from_host(S)
    out = None
    for rep in range(3):
        vals = np.random.default_rng(100 + rep).uniform(-1, 1, S.nnz)
        dS.val.copy_(torch.from_numpy(vals))
        S_rep = gen.HostCSR(S.rpt, S.col, vals, S.nrows, S.ncols)
        out = h.spgemm_into(dS, dS, out=out)
        assert h.stats()["fused"] == 1
        check_against_oracle(S_rep, S_rep, out, "f64", orc)
        C2 = h.spgemm(nsp.DeviceCSR.from_host(A), nsp.DeviceCSR.from_host(B))
        check_against_oracle(A, B, C2, "f64", orc)


def test_cached_slot_size_with_a_changed_pattern(nsp, orc):
    """A second one-call product on the same buffers and sizes reuses the
    first call's slot size without a sample; when the pattern changed in
    place so that a row no longer fits, the row is caught and the product
    redone in two passes."""
    print("case: fused cached slot size", flush=True)
    import torch
    rng = np.random.default_rng(95)
    n, m = 8192, 5000
    B = mixed_rows(m, 100_000, 8, 12, seed=96)

    def rows(lens):
        cols = np.concatenate([rng.choice(m, k, replace=False) for k in lens])
        return gen._from_rows_cols(n, m, np.repeat(np.arange(n), lens), cols, rng.uniform(-1, 1, lens.sum()))

    l1 = np.full(n, 3)
    l2 = l1.copy()
    l2[1] += 57
    l2[2:59] -= 1
    A1, A2 = rows(l1), rows(l2)
    assert A1.nnz == A2.nnz
    h = nsp.Handle()
    dA, dB = nsp.DeviceCSR.from_host(A1), nsp.DeviceCSR.from_host(B)
    C = h.spgemm_into(dA, dB)
    st1 = h.stats()
    assert st1["fused"] == 1
    check_against_oracle(A1, B, C, "f64", orc)
    C = h.spgemm_into(dA, dB)  # same pattern: cached slot size, light row info
    st2 = h.stats()
    assert st2["fused"] == 1 and st2["host_syncs"] == 1
    assert st2["nprod"] == st1["nprod"] == int(orc.intprod(A1, B)[1]) and st2["max_intprod"] == -1
    check_against_oracle(A1, B, C, "f64", orc)
    d2 = nsp.DeviceCSR.from_host(A2)
    dA.row_ptr.copy_(d2.row_ptr)
    dA.col.copy_(d2.col)
    dA.val.copy_(d2.val)
    torch.cuda.synchronize()
    C = h.spgemm_into(dA, dB)
    assert h.stats()["fused"] == 0
    check_against_oracle(A2, B, C, "f64", orc)


def test_spgemm_once_allocates_through_torch(nsp, orc):
    """Handle.spgemm_once: nsp_spgemm with C allocated through torch's
    allocator, the one-call path underneath."""
    print("case: spgemm_once", flush=True)
    S = gen.stencil_27pt(18, seed=97)
    h = nsp.Handle()
    C = h.spgemm_once(nsp.DeviceCSR.from_host(S), nsp.DeviceCSR.from_host(S))
    assert h.stats()["fused"] == 1
    check_against_oracle(S, S, C, "f64", orc)


def test_grid_aligned_rows_take_one_pass(nsp, orc):
    """Stencil 32^3 times the aggregation prolongator: 32768 rows, so 1024
    evenly spaced rows would all lie on the x = 0 face, whose C rows are
    shorter than the interior's.  The sample and the product-count bound
    still size the slot for every row: one pass, the oracle's C.  (P has
    4096 columns, so the column-bitmap kernels are switched off.)"""
    print("case: fused grid-aligned", flush=True)
    S = gen.stencil_27pt(32, seed=98)
    P = gen.aggregation(S.nrows, seed=99)
    ref = orc.spgemm(S, P)
    h = nsp.Handle()
    h.set_param("narrow_mode", 0)
    C = h.spgemm_into(nsp.DeviceCSR.from_host(S), nsp.DeviceCSR.from_host(P))
    st = h.stats()
    assert st["fused"] == 1 and st["sym_launches"] == 0, st
    check_against_oracle(S, P, C, "f64", orc, ref=ref)
    # P^T (A P) through nsp_galerkin: both products in one pass
    Cg = h.galerkin(nsp.DeviceCSR.from_host(S), nsp.DeviceCSR.from_host(P))
    assert h.stats()["fused"] == 1
    trpt, tcol, tval = orc.transpose(P)
    PT = gen.HostCSR(trpt, tcol, tval, P.ncols, P.nrows)
    AP = gen.HostCSR(ref[0], ref[1], ref[2], S.nrows, P.ncols)
    check_against_oracle(PT, AP, Cg, "f64", orc)
