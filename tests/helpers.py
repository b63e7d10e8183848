"""Shared checks for the GPU parity tests: structure bit-exact against the
oracle, values within a tolerance relative to sum_k |a_ik b_kj|."""
from __future__ import annotations

import numpy as np

# Tolerances (componentwise, relative to sum_k |a_ik b_kj|):
#   FP64 values with FP64 accumulation: 1e-12
#   FP32 values with FP32 accumulation: 1e-5
TOL = {"f64": 1e-12, "f32": 1e-5}


def to_dtype(h, dtype):
    """Host CSR with values rounded to the GPU value type (the oracle then
    sees exactly the inputs the GPU sees)."""
    from nsp_amd.gen import HostCSR
    if dtype == "f32":
        return HostCSR(h.rpt, h.col, h.val.astype(np.float32).astype(np.float64), h.nrows, h.ncols)
    return h


def check_against_oracle(A, B, C, dtype="f64", oracle=None, ref=None):
    """A, B: HostCSR inputs (already rounded to dtype); C: DeviceCSR result.
    Returns (componentwise error, Frobenius relative error)."""
    if oracle is None:
        from oracle import oracle
    if ref is None:
        ref = oracle.spgemm(A, B)
    crpt, ccol, cval, cabs = ref
    got_rpt = C.row_ptr.cpu().numpy().astype(np.int64)
    got_col = C.col.cpu().numpy()
    got_val = C.val.cpu().numpy().astype(np.float64)
    assert got_rpt.shape == crpt.shape
    np.testing.assert_array_equal(got_rpt, crpt, err_msg="row_ptr differs from the oracle")
    np.testing.assert_array_equal(got_col, ccol, err_msg="col_idx differs from the oracle")
    comp, fro = oracle.componentwise_error(got_val, cval, cabs)
    assert comp <= TOL[dtype], f"componentwise error {comp:.3e} > {TOL[dtype]:.0e} ({dtype})"
    return comp, fro
