"""GPU backward pair on fixed patterns (config 4: 27-point stencil on 64^3 times
the aggregation prolongator), against the CPU oracle's gradient and central
finite differences evaluated on the affected rows only.

Tolerances: FP64 1e-12 and FP32 1e-5, componentwise relative to the sum of
absolute terms (|B_kj||dC_ij| for dA, |A_ik||dC_ij| for dB); FD 1e-6
relative (h = 1e-6, the loss is quadratic so central differences are exact up
to rounding)."""
import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import to_dtype

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 1e-5}


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available()
    import nsp_amd
    return nsp_amd


def abs_csr(X):
    return gen.HostCSR(X.rpt, X.col, np.abs(X.val), X.nrows, X.ncols)


def test_transpose_matches_oracle(nsp, orc):
    print("case: transpose", flush=True)
    h = nsp.Handle()
    for X in (gen.random_csr(300, 500, 0.02, seed=1, empty_rows=0.2), gen.stencil_27pt(10, seed=2)):
        XT = h.transpose(nsp.DeviceCSR.from_host(X))
        trpt, tcol, tval = orc.transpose(X)
        np.testing.assert_array_equal(XT.row_ptr.cpu().numpy(), trpt)
        np.testing.assert_array_equal(XT.col.cpu().numpy(), tcol)
        np.testing.assert_array_equal(XT.val.cpu().numpy(), tval)


def c4_problem(m, dtype, seed=0):
    A = to_dtype(gen.stencil_27pt(m, seed=seed), dtype)
    P = to_dtype(gen.aggregation(A.nrows, seed=seed + 1), dtype)
    return A, P


@pytest.mark.parametrize("dtype,m", [("f64", 64), ("f32", 64), ("f64", 12)])
def test_backward_matches_oracle(nsp, orc, dtype, m):
    """dA, dB from nsp_spgemm_backward against the oracle's closed forms."""
    print("case: backward vs oracle", dtype, m, flush=True)
    import torch
    A, P = c4_problem(m, dtype)
    h = nsp.Handle()
    dA_ = nsp.DeviceCSR.from_host(A, dtype)
    dP_ = nsp.DeviceCSR.from_host(P, dtype)
    C = h.spgemm(dA_, dP_)
    AT = h.transpose(dA_)
    rng = np.random.default_rng(7)
    dc = rng.standard_normal(C.nnz)
    if dtype == "f32":
        dc = dc.astype(np.float32).astype(np.float64)
    tdc = torch.from_numpy(dc).to(device="cuda", dtype=C.val.dtype)
    da, db = h.backward(dA_, dP_, AT, C, tdc)
    crpt = C.row_ptr.cpu().numpy().astype(np.int64)
    ccol = C.col.cpu().numpy()
    ref_da, ref_db = orc.backward(A, P, crpt, ccol, dc)
    abs_da, abs_db = orc.backward(abs_csr(A), abs_csr(P), crpt, ccol, np.abs(dc))
    err_a = np.max(np.abs(da.cpu().numpy() - ref_da) / np.where(abs_da > 0, abs_da, 1))
    err_b = np.max(np.abs(db.cpu().numpy() - ref_db) / np.where(abs_db > 0, abs_db, 1))
    print(f"  componentwise error dA {err_a:.2e} dB {err_b:.2e}", flush=True)
    assert err_a <= TOL[dtype] and err_b <= TOL[dtype]


def test_autograd_finite_differences(nsp, orc):
    """Config 4 in FP64: loss = 0.5 ||C - T||^2 on C's pattern; the autograd
    gradient against central differences (h = 1e-6) on 64 seeded coordinates
    (32 of A, 32 of B), each FD taken over the C rows the coordinate touches."""
    print("case: autograd FD", flush=True)
    import torch
    from nsp_amd.autograd import SpGEMMFixedPattern
    m = 64
    A, P = c4_problem(m, "f64", seed=3)
    dAp = nsp.DeviceCSR.from_host(A)
    dPp = nsp.DeviceCSR.from_host(P)
    op = SpGEMMFixedPattern(dAp, dPp)
    rng = np.random.default_rng(11)
    T = torch.from_numpy(rng.standard_normal(op.C.nnz)).cuda()
    a = dAp.val.clone().requires_grad_(True)
    b = dPp.val.clone().requires_grad_(True)
    c = op(a, b)
    loss = 0.5 * ((c - T) ** 2).sum()
    loss.backward()
    ga, gb = a.grad.cpu().numpy(), b.grad.cpu().numpy()

    crpt = op.C.row_ptr.cpu().numpy().astype(np.int64)
    Tn = T.cpu().numpy()
    h = nsp.Handle()
    hstep = 1e-6

    def rows_loss(rows, aval, bval):
        """0.5 sum (C_ij - T_ij)^2 over C rows `rows`, C recomputed on the GPU
        from those rows of A only."""
        sub = gen.HostCSR(np.concatenate([[0], np.cumsum(np.diff(A.rpt)[rows])]),
                          np.concatenate([A.col[A.rpt[r]:A.rpt[r + 1]] for r in rows]),
                          np.concatenate([aval[A.rpt[r]:A.rpt[r + 1]] for r in rows]),
                          len(rows), A.ncols)
        Pm = gen.HostCSR(P.rpt, P.col, bval, P.nrows, P.ncols)
        Cs = h.spgemm(nsp.DeviceCSR.from_host(sub), nsp.DeviceCSR.from_host(Pm))
        cv = Cs.val.cpu().numpy()
        tv = np.concatenate([Tn[crpt[r]:crpt[r + 1]] for r in rows])
        return 0.5 * float(np.sum((cv - tv) ** 2))

    rows_of_a = np.repeat(np.arange(A.nrows), np.diff(A.rpt))
    coords_a = rng.choice(A.nnz, 32, replace=False)
    coords_b = rng.choice(P.nnz, 32, replace=False)
    AT = orc.transpose(A)
    worst = 0.0
    for e in coords_a:
        rows = [int(rows_of_a[e])]
        ap, am = A.val.copy(), A.val.copy()
        ap[e] += hstep
        am[e] -= hstep
        fd = (rows_loss(rows, ap, P.val) - rows_loss(rows, am, P.val)) / (2 * hstep)
        worst = max(worst, abs(fd - ga[e]) / max(1.0, abs(ga[e])))
    for e in coords_b:
        k = int(np.searchsorted(P.rpt, e, side="right") - 1)
        rows = [int(r) for r in AT[1][AT[0][k]:AT[0][k + 1]]]
        bp, bm = P.val.copy(), P.val.copy()
        bp[e] += hstep
        bm[e] -= hstep
        fd = (rows_loss(rows, A.val, bp) - rows_loss(rows, A.val, bm)) / (2 * hstep)
        worst = max(worst, abs(fd - gb[e]) / max(1.0, abs(gb[e])))
    print(f"  worst FD relative error {worst:.2e}", flush=True)
    assert worst <= 1e-6


def _mixed(nrows, ncols, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lens = np.clip(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), nrows)).astype(np.int64), 0, ncols)
    rows = np.repeat(np.arange(nrows), lens)
    cols = np.concatenate([rng.choice(ncols, int(k), replace=False) for k in lens])
    return gen._from_rows_cols(nrows, ncols, rows, cols, rng.uniform(-1, 1, rows.size))


@pytest.mark.parametrize("case", ["stencil", "mixed", "rmat"])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("use_perm", [False, True])
def test_backward_general_patterns(nsp, orc, case, dtype, use_perm):
    """A*A on the 27-point stencil (C rows up to 125 entries), a product of
    mixed row lengths (C rows past 128, the second table tier) and R-MAT
    scale 14 A*A (6244 C rows past 1024 entries, up to 10,358: the search
    tier), with A's values given as A^T values or through the A^T -> A
    permutation."""
    print("case: backward general", case, dtype, use_perm, flush=True)
    import torch
    if case == "stencil":
        A = B = to_dtype(gen.stencil_27pt(10, seed=4), dtype)
    elif case == "mixed":
        A = to_dtype(_mixed(600, 500, 1, 40, seed=5), dtype)
        B = to_dtype(_mixed(500, 700, 1, 30, seed=6), dtype)
    else:
        A = B = to_dtype(gen.rmat(14, seed=3), dtype)
    h = nsp.Handle()
    dA_ = nsp.DeviceCSR.from_host(A, dtype)
    dB_ = nsp.DeviceCSR.from_host(B, dtype)
    C = h.spgemm(dA_, dB_)
    crpt = C.row_ptr.cpu().numpy().astype(np.int64)
    if case == "mixed":
        assert np.diff(crpt).max() > 128
    if case == "rmat":
        assert (np.diff(crpt) > 1024).sum() > 1000
    if use_perm:
        idx = torch.arange(A.nnz, dtype=torch.float64, device="cuda")
        ATi = h.transpose(nsp.DeviceCSR(dA_.row_ptr, dA_.col, idx, A.nrows, A.ncols))
        perm = ATi.val.to(dA_.row_ptr.dtype)
        AT = nsp.DeviceCSR(ATi.row_ptr, ATi.col, None, ATi.nrows, ATi.ncols)
    else:
        perm, AT = None, h.transpose(dA_)
    rng = np.random.default_rng(8)
    dc = rng.standard_normal(C.nnz)
    if dtype == "f32":
        dc = dc.astype(np.float32).astype(np.float64)
    ccol = C.col.cpu().numpy()
    ref_da, ref_db = orc.backward(A, B, crpt, ccol, dc)
    abs_da, abs_db = orc.backward(abs_csr(A), abs_csr(B), crpt, ccol, np.abs(dc))
    # dB by per-product contributions (1), by search in C's rows (0), and
    # the default choice by C's longest row (2, last: kept for below)
    for mode in ((1, 0, 2) if use_perm else (2,)):
        h.set_param("bwd_contrib", mode)
        da, db = h.backward(dA_, dB_, AT, C, torch.from_numpy(dc).to(device="cuda", dtype=C.val.dtype),
                            at_perm=perm)
        err_a = np.max(np.abs(da.cpu().numpy() - ref_da) / np.where(abs_da > 0, abs_da, 1))
        err_b = np.max(np.abs(db.cpu().numpy() - ref_db) / np.where(abs_db > 0, abs_db, 1))
        print(f"  bwd_contrib {mode}: componentwise error dA {err_a:.2e} dB {err_b:.2e}", flush=True)
        assert err_a <= TOL[dtype] and err_b <= TOL[dtype]
    # deterministic: a second call gives the same bits, also when it reuses
    # the first call's B row extents
    h.set_param("reuse_extents", 1)
    da2, db2 = h.backward(dA_, dB_, AT, C, torch.from_numpy(dc).to(device="cuda", dtype=C.val.dtype),
                          at_perm=perm)
    assert torch.equal(da, da2) and torch.equal(db, db2)
