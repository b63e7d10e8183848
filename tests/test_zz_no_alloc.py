"""nsp_spgemm_noalloc: the product with its workspace in the library's static
device array, called through the C ABI, against nsp_spgemm (the one-call
product that allocates) on the same inputs: row pointer and columns
bit-exact, values within the parity tolerance (FP64 1e-12, FP32 1e-5,
componentwise relative to sum_k |a_ik b_kj|, the oracle's magnitudes).  Two
calls in a row on one handle, the second with a smaller input, so the second
runs on the array the first left behind; then a capacity that is too small
(NSP_ERR_OVERFLOW and the needed nnz) and an A too large for the array
(NSP_ERR_UNSUPPORTED before any work)."""
import ctypes as C

import numpy as np
import pytest

from nsp_amd import gen
from tests.helpers import TOL, to_dtype

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nsp():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    import nsp_amd
    return nsp_amd


def call_spgemm(nsp, h, dA, dB):
    """nsp_spgemm with the default (hipMalloc) allocator; host copies of C."""
    L = h.L
    a, b, c = dA.c_struct(), dB.c_struct(), nsp._lib.nsp_csr_t()
    h._check(L.nsp_spgemm(h.h, dA.index_type, dA.value_type, C.byref(a), C.byref(b), C.byref(c), None, None, None))
    import torch
    torch.cuda.synchronize()
    it = np.int64 if dA.index_type == nsp._lib.NSP_IDX64 else np.int32
    vt = np.float32 if dA.value_type == nsp._lib.NSP_F32 else np.float64
    rows, nnz = int(c.rows), int(c.nnz)
    rpt = np.empty(rows + 1, it)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, vt)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    assert hip.hipMemcpy(rpt.ctypes.data, c.row_ptr, rpt.nbytes, 2) == 0
    if nnz:
        assert hip.hipMemcpy(col.ctypes.data, c.col, col.nbytes, 2) == 0
        assert hip.hipMemcpy(val.ctypes.data, c.val, val.nbytes, 2) == 0
    L.nsp_csr_free(C.byref(c))
    return rpt.astype(np.int64), col, val.astype(np.float64)


def call_noalloc(nsp, h, dA, dB, capacity):
    import torch
    L = h.L
    itype = torch.int64 if dA.index_type == nsp._lib.NSP_IDX64 else torch.int32
    out = nsp.DeviceCSR(torch.empty(dA.nrows + 1, dtype=itype, device="cuda"),
                        torch.empty(max(capacity, 1), dtype=torch.int32, device="cuda"),
                        torch.empty(max(capacity, 1), dtype=dA.val.dtype, device="cuda"), dA.nrows, dB.ncols)
    a, b = dA.c_struct(), dB.c_struct()
    c = nsp._lib.nsp_csr_t(dA.nrows, dB.ncols, capacity, out.row_ptr.data_ptr(), out.col.data_ptr(),
                           out.val.data_ptr())
    nnz = C.c_int64(-1)
    st = L.nsp_spgemm_noalloc(h.h, dA.index_type, dA.value_type, C.byref(a), C.byref(b), C.byref(c), C.byref(nnz))
    return st, int(nnz.value), out


def check_pair(nsp, orc, A, B, dtype, index, h):
    dA = nsp.DeviceCSR.from_host(A, dtype, index)
    dB = nsp.DeviceCSR.from_host(B, dtype, index)
    rpt, col, val = call_spgemm(nsp, h, dA, dB)
    st, nnz, out = call_noalloc(nsp, h, dA, dB, capacity=int(rpt[-1]) + 17)
    assert st == nsp._lib.NSP_OK, nsp._lib.load().nsp_last_error(h.h).decode()
    assert nnz == int(rpt[-1])
    np.testing.assert_array_equal(out.row_ptr.cpu().numpy().astype(np.int64), rpt)
    np.testing.assert_array_equal(out.col[:nnz].cpu().numpy(), col)
    _, _, ref_val, ref_abs = orc.spgemm(A, B)
    den = np.where(ref_abs > 0, ref_abs, 1.0)
    got = out.val[:nnz].cpu().numpy().astype(np.float64)
    err_pair = float(np.max(np.abs(got - val) / den)) if nnz else 0.0
    err_orc = float(np.max(np.abs(got - ref_val) / den)) if nnz else 0.0
    print(f"  nnz {nnz}: noalloc vs nsp_spgemm {err_pair:.2e}, vs oracle {err_orc:.2e}", flush=True)
    assert err_pair <= TOL[dtype] and err_orc <= TOL[dtype]
    return dA, dB, nnz


def test_noalloc_matches_spgemm_two_calls(nsp, orc):
    print("case: noalloc two calls (large then small)", flush=True)
    h = nsp.Handle()
    # first call: 27-point stencil 24^3 times a random matrix (several bins,
    # heavy-free), FP64 with int32 row pointers
    A1 = gen.stencil_27pt(24, seed=3)
    B1 = gen.random_csr(A1.ncols, 5000, 0.004, seed=4)
    check_pair(nsp, orc, A1, B1, "f64", "i32", h)
    # second call on the array the first left: a smaller product with mixed
    # row lengths, FP32 values and int64 row pointers
    A2 = to_dtype(gen.random_csr(700, 900, 0.03, seed=5, empty_rows=0.2), "f32")
    B2 = to_dtype(gen.random_csr(900, 800, 0.05, seed=6), "f32")
    check_pair(nsp, orc, A2, B2, "f32", "i64", h)


def test_noalloc_capacity_and_size_errors(nsp, orc):
    print("case: noalloc capacity and size errors", flush=True)
    import torch
    h = nsp.Handle()
    A = gen.stencil_27pt(10, seed=7)
    dA = nsp.DeviceCSR.from_host(A)
    ref_rpt = orc.spgemm(A, A)[0]
    st, nnz, _ = call_noalloc(nsp, h, dA, dA, capacity=int(ref_rpt[-1]) // 2)
    assert st == nsp._lib.NSP_ERR_OVERFLOW and nnz == int(ref_rpt[-1])
    # the handle and the array stay usable
    check_pair(nsp, orc, A, A, "f64", "i32", h)
    # an A whose workspace passes the array: rejected before any work
    rows, per = 1 << 20, 8
    big = nsp.DeviceCSR(torch.arange(0, rows * per + 1, per, dtype=torch.int32, device="cuda"),
                        torch.zeros(rows * per, dtype=torch.int32, device="cuda"),
                        torch.ones(rows * per, dtype=torch.float64, device="cuda"), rows, 1)
    one = nsp.DeviceCSR(torch.tensor([0, 1], dtype=torch.int32, device="cuda"),
                        torch.zeros(1, dtype=torch.int32, device="cuda"),
                        torch.ones(1, dtype=torch.float64, device="cuda"), 1, 1)
    st, _, _ = call_noalloc(nsp, h, big, one, capacity=rows)
    assert st == nsp._lib.NSP_ERR_UNSUPPORTED
    assert "scratch" in nsp._lib.load().nsp_last_error(h.h).decode()
