"""Galerkin product C = R (A P) (nsp_galerkin) against two oracle products:
AP* = A P and C* = R AP* in FP64 (Neumaier sums), R = P^T from the oracle's
transpose when R is not given.  Structure bit-exact; values within
TOL[dtype] componentwise relative to the triple-product magnitude
(|R| |A| |P|)_ij, which bounds the rounding of both products.

Cases: the 27-point stencil with the aggregation prolongator (the AMG
coarse operator), a smoothed prolongator (several entries per row), an
explicit R that is not P^T, FP32, int64 row pointers, and the C++ API."""
import numpy as np
import pytest

from nsp_amd import gen
from nsp_amd.gen import HostCSR
from tests.helpers import TOL, to_dtype

pytestmark = pytest.mark.gpu


def host_abs(X):
    return HostCSR(X.rpt, X.col, np.abs(X.val), X.nrows, X.ncols)


def oracle_galerkin(orc, R, A, P):
    ap_rpt, ap_col, ap_val, _ = orc.spgemm(A, P, with_abs=False)
    AP = HostCSR(ap_rpt, ap_col, ap_val, A.nrows, P.ncols)
    crpt, ccol, cval, _ = orc.spgemm(R, AP, with_abs=False)
    # magnitude bound (|R| |A| |P|)_ij on the same structure
    m_rpt, m_col, m_val, _ = orc.spgemm(host_abs(A), host_abs(P), with_abs=False)
    M = HostCSR(m_rpt, m_col, m_val, A.nrows, P.ncols)
    _, _, cabs, _ = orc.spgemm(host_abs(R), M, with_abs=False)
    return crpt, ccol, cval, cabs


def transpose_host(orc, X):
    trpt, tcol, tval = orc.transpose(X)
    return HostCSR(trpt, tcol, tval, X.ncols, X.nrows)


def smoothed(orc, A, P0, omega=0.5):
    """(I - omega D^-1 A) P0, computed on the CPU: several entries per row."""
    d = np.zeros(A.nrows)
    rows = np.repeat(np.arange(A.nrows), np.diff(A.rpt))
    on = A.col == rows
    d[rows[on]] = A.val[on]
    S_val = -omega * A.val / d[rows]
    S_val[on] += 1.0
    S = HostCSR(A.rpt, A.col, S_val, A.nrows, A.ncols)
    rpt, col, val, _ = orc.spgemm(S, P0, with_abs=False)
    return HostCSR(rpt, col, val, A.nrows, P0.ncols)


def check(C, ref, dtype, orc):
    crpt, ccol, cval, cabs = ref
    np.testing.assert_array_equal(C.row_ptr.cpu().numpy().astype(np.int64), crpt)
    np.testing.assert_array_equal(C.col.cpu().numpy(), ccol)
    comp, _ = orc.componentwise_error(C.val.cpu().numpy().astype(np.float64), cval, cabs)
    assert comp <= TOL[dtype], f"componentwise error {comp:.3e} > {TOL[dtype]:.0e}"


@pytest.fixture(scope="module")
def h():
    import nsp_amd
    return nsp_amd.Handle()


@pytest.mark.parametrize("prolongator", ["aggregation", "smoothed"])
@pytest.mark.parametrize("index,dtype", [("i32", "f64"), ("i64", "f64"), ("i32", "f32")])
def test_galerkin_pt_a_p(h, orc, prolongator, index, dtype):
    import nsp_amd
    A = to_dtype(gen.stencil_27pt(16, seed=4), dtype)
    P = gen.aggregation(A.nrows, seed=2)
    if prolongator == "smoothed":
        P = smoothed(orc, A, P)
    P = to_dtype(P, dtype)
    print(f"  {prolongator} {index}/{dtype}: A {A.nrows}x{A.ncols} nnz {A.nnz}, P nnz {P.nnz}", flush=True)
    dA = nsp_amd.DeviceCSR.from_host(A, dtype, index)
    dP = nsp_amd.DeviceCSR.from_host(P, dtype, index)
    C = h.galerkin(dA, dP)
    assert (C.nrows, C.ncols) == (P.ncols, P.ncols)
    check(C, oracle_galerkin(orc, transpose_host(orc, P), A, P), dtype, orc)


def test_galerkin_explicit_r(h, orc):
    import nsp_amd
    A = gen.random_csr(300, 300, 0.02, seed=8)
    P = gen.random_csr(300, 40, 0.05, seed=9, empty_rows=0.1)
    R = gen.random_csr(25, 300, 0.04, seed=10)
    C = h.galerkin(*(nsp_amd.DeviceCSR.from_host(X) for X in (A, P)), R=nsp_amd.DeviceCSR.from_host(R))
    assert (C.nrows, C.ncols) == (25, 40)
    check(C, oracle_galerkin(orc, R, A, P), "f64", orc)


def test_galerkin_shape_errors(h):
    import nsp_amd
    from nsp_amd import NspError
    A = nsp_amd.DeviceCSR.from_host(gen.random_csr(30, 30, 0.1, seed=1))
    P = nsp_amd.DeviceCSR.from_host(gen.random_csr(20, 5, 0.2, seed=2))
    with pytest.raises(NspError):
        h.galerkin(A, P)


def test_galerkin_plan_relaunch_with_new_values(orc):
    """GalerkinPlan (two recorded graphs, P^T values by gather): launches
    after the values of A and P change in place match the oracle's products
    of the new values; the patterns stay."""
    import torch
    import nsp_amd
    from nsp_amd.galerkin import GalerkinPlan
    A = gen.stencil_27pt(12, seed=5)
    P = smoothed(orc, A, gen.aggregation(A.nrows, seed=6))
    dA, dP = nsp_amd.DeviceCSR.from_host(A), nsp_amd.DeviceCSR.from_host(P)
    plan = GalerkinPlan(dA, dP)
    rng = np.random.default_rng(7)
    for step in range(3):
        print(f"  plan launch {step}", flush=True)
        plan.launch()
        C = plan.result()
        check(C, oracle_galerkin(orc, transpose_host(orc, P), A, P), "f64", orc)
        # new values, same patterns
        A = HostCSR(A.rpt, A.col, A.val * rng.uniform(0.5, 1.5, A.nnz), A.nrows, A.ncols)
        P = HostCSR(P.rpt, P.col, P.val * rng.uniform(0.5, 1.5, P.nnz), P.nrows, P.ncols)
        dA.val.copy_(torch.from_numpy(A.val))
        dP.val.copy_(torch.from_numpy(P.val))
