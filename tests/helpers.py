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


def write_ascii_coo(path, X=None, records=None, nr=None, nc=None, header_fields=3, comments=(),
                    value_fmt="{:.17g}", final_newline=True, col_shift=0):
    """ASCII COO file of the reference's text format: '%' comment lines, a
    header "nr nc nt" (or "nr nt"), then fixed-length record lines "i j v"
    (1-based; right-aligned row and column fields, the value left-aligned
    and padded), so every line has the first line's field offsets.  From a
    HostCSR X (global column = col + col_shift) or a list of (i, j, v)
    records with 1-based indices (v a float, or a string written as is)."""
    if records is None:
        rows = np.repeat(np.arange(X.nrows), np.diff(X.rpt))
        records = [(int(i) + 1, int(j) + col_shift + 1, float(v)) for i, j, v in zip(rows, X.col, X.val)]
        nr, nc = X.nrows, X.ncols + col_shift
    texts = [v if isinstance(v, str) else value_fmt.format(v) for _, _, v in records]
    w1 = len(str(max([nr] + [i for i, _, _ in records]))) + 1
    w2 = len(str(max([nc] + [j for _, j, _ in records]))) + 1
    w3 = max([len(t) for t in texts] + [1])
    lines = [f"{i:>{w1}d} {j:>{w2}d} {t:<{w3}s}\n" for (i, j, _), t in zip(records, texts)]
    head = f"{nr} {nc} {len(records)}\n" if header_fields == 3 else f"{nr} {len(records)}\n"
    body = "".join(lines)
    if not final_newline and body:
        body = body[:-1]
    with open(path, "w") as f:
        f.write("".join(c + "\n" for c in comments) + head + body)
    return records
