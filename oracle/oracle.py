"""ctypes wrapper of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY.  Only the tests, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` import this module; the HIP library and
the ``nsp_amd`` package never do.  Every function here restates the
reference's algorithm for the local product (citations in oracle.c).

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  intprod, exscan          closed forms (stencils) + brute force on small inputs
  symbolic / spgemm        closed forms for 5-point and 27-point stencils,
                           dense brute force on small random integer matrices
                           (exact in float64), scipy on all-ones patterns
  shrink / completed_rows  the three branch forms of getmct_4shrink restated
                           separately in the test + searchsorted closed form
  transpose                numpy dense transpose on small inputs
  backward                 dense closed form dA = (dC B^T)*pat(A),
                           dB = (A^T dC)*pat(B) on small integer inputs
  read/write binary COO    round trip + split/stitch invariants
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        I64, VP = C.c_int64, C.c_void_p
        L.orc_intprod.restype = I64
        L.orc_intprod.argtypes = [I64, _i64p, _i32p, _i64p, _i64p]
        L.orc_exscan.restype = I64
        L.orc_exscan.argtypes = [I64, _i64p, _i64p]
        L.orc_symbolic.restype = I64
        L.orc_symbolic.argtypes = [I64, _i64p, _i32p, _i64p, _i32p, I64, _i64p]
        L.orc_spgemm.restype = None
        L.orc_spgemm.argtypes = [I64, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p, I64, _i64p,
                                 VP, VP, VP]
        L.orc_symbolic_rows.restype = None
        L.orc_symbolic_rows.argtypes = [I64, _i64p, _i64p, _i32p, _i64p, _i32p, I64, _i64p]
        L.orc_spgemm_rows.restype = None
        L.orc_spgemm_rows.argtypes = [I64, _i64p, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p, I64,
                                      _i64p, VP, VP, VP]
        L.orc_shrink.restype = I64
        L.orc_shrink.argtypes = [I64, _i32p, I64, _i32p, _i32p]
        L.orc_completed_rows.restype = I64
        L.orc_completed_rows.argtypes = [I64, _i32p, I64, I64, _i64p, _i32p]
        L.orc_gather_rows.restype = I64
        L.orc_gather_rows.argtypes = [I64, _i64p, _i64p, _i32p, _f64p, I64, _i64p, VP, VP]
        L.orc_transpose.restype = None
        L.orc_transpose.argtypes = [I64, I64, _i64p, _i32p, _f64p, _i64p, _i32p, VP]
        L.orc_backward.restype = None
        L.orc_backward.argtypes = [I64, _i64p, _i32p, _f64p, I64, _i64p, _i32p, _f64p,
                                   _i64p, _i32p, _f64p, _f64p, _f64p]
        L.orc_read_binary_coo.restype = C.c_int
        L.orc_read_binary_coo.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int] + \
            [C.POINTER(C.c_int64)] * 6 + [C.POINTER(VP)] * 3
        L.orc_read_ascii_coo.restype = C.c_int
        L.orc_read_ascii_coo.argtypes = [C.c_char_p, C.c_int, C.c_int] + [C.POINTER(C.c_int64)] * 6 + \
            [C.POINTER(C.c_void_p)] * 3
        L.orc_fast_atof.restype = C.c_double
        L.orc_fast_atof.argtypes = [C.c_char_p, C.c_char_p]
        L.orc_write_binary_coo.restype = C.c_int
        L.orc_write_binary_coo.argtypes = [C.c_char_p, C.c_int, I64, I64, _i64p, _i32p, _f64p, I64]
        L.orc_free.argtypes = [VP]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _LIB = L
    return _LIB


def _csr(x):
    return (np.ascontiguousarray(x.rpt, np.int64), np.ascontiguousarray(x.col, np.int32),
            np.ascontiguousarray(x.val, np.float64))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def num_threads() -> int:
    return lib().orc_num_threads()


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


def intprod(A, B):
    arpt, acol, _ = _csr(A)
    brpt = np.ascontiguousarray(B.rpt, np.int64)
    out = np.empty(A.nrows, np.int64)
    total = lib().orc_intprod(A.nrows, arpt, acol, brpt, out)
    return out, int(total)


def exscan(x):
    x = np.ascontiguousarray(x, np.int64)
    out = np.empty(x.size + 1, np.int64)
    lib().orc_exscan(x.size, x, out)
    return out


def symbolic(A, B):
    arpt, acol, _ = _csr(A)
    brpt, bcol, _ = _csr(B)
    nz = np.empty(A.nrows, np.int64)
    lib().orc_symbolic(A.nrows, arpt, acol, brpt, bcol, B.ncols, nz)
    return nz


def spgemm(A, B, with_abs=True):
    """C = A*B -> (rpt int64, col int32, val float64, abs float64 or None)."""
    arpt, acol, aval = _csr(A)
    brpt, bcol, bval = _csr(B)
    crpt = exscan(symbolic(A, B))
    nnz = int(crpt[-1])
    ccol = np.empty(nnz, np.int32)
    cval = np.empty(nnz, np.float64)
    cabs = np.empty(nnz, np.float64) if with_abs else None
    lib().orc_spgemm(A.nrows, arpt, acol, aval, brpt, bcol, bval, B.ncols, crpt,
                     _ptr(ccol), _ptr(cval), _ptr(cabs))
    return crpt, ccol, cval, cabs


def spgemm_rows(A, B, rows, with_abs=True):
    """Product restricted to `rows` (packed in that order)."""
    arpt, acol, aval = _csr(A)
    brpt, bcol, bval = _csr(B)
    rows = np.ascontiguousarray(rows, np.int64)
    nz = np.empty(rows.size, np.int64)
    lib().orc_symbolic_rows(rows.size, rows, arpt, acol, brpt, bcol, B.ncols, nz)
    rpt = exscan(nz)
    nnz = int(rpt[-1])
    ccol = np.empty(nnz, np.int32)
    cval = np.empty(nnz, np.float64)
    cabs = np.empty(nnz, np.float64) if with_abs else None
    lib().orc_spgemm_rows(rows.size, rows, arpt, acol, aval, brpt, bcol, bval, B.ncols, rpt,
                          _ptr(ccol), _ptr(cval), _ptr(cabs))
    return rpt, ccol, cval, cabs


def shrink(col, l):
    """(perm, newcol): perm = sorted unique(col U [0..l]), newcol = index in perm."""
    col = np.ascontiguousarray(col, np.int32)
    perm = np.empty(int(l) + 1 + col.size, np.int32)
    newcol = np.empty(col.size, np.int32)
    m = lib().orc_shrink(col.size, col, int(l), perm, newcol)
    return perm[:m].copy(), newcol


def completed_rows(acol_rank, l, r0):
    acol_rank = np.ascontiguousarray(acol_rank, np.int32)
    grows = np.empty(int(l) + 1 + acol_rank.size, np.int64)
    newcol = np.empty(acol_rank.size, np.int32)
    m = lib().orc_completed_rows(acol_rank.size, acol_rank, int(l), int(r0), grows, newcol)
    return grows[:m].copy(), newcol


def gather_rows(B, grows, shift):
    brpt, bcol, bval = _csr(B)
    grows = np.ascontiguousarray(grows, np.int64)
    rpt = np.empty(grows.size + 1, np.int64)
    nnz = lib().orc_gather_rows(grows.size, grows, brpt, bcol, bval, int(shift), rpt, None, None)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib().orc_gather_rows(grows.size, grows, brpt, bcol, bval, int(shift), rpt, _ptr(col), _ptr(val))
    return rpt, col, val


def local_share(A_rank, B_full, r0, p_local_n=None):
    """One rank's local product on one device, as nsparseMGPU_version1 builds
    it (src/spspmpi.cu:371-975) without MPI.

    A_rank: the rank's row block of A, columns in the rank frame (global - r0)
    B_full: all rows of B, global columns
    Returns dict with A_ (compacted), P_ (completedP, compacted), their
    shrinked_m, the completedP global row list and perm_P (rank frame)."""
    from nsp_amd.gen import HostCSR
    if p_local_n is None:
        p_local_n = A_rank.nrows
    l = p_local_n - 1
    grows, a_newcol = completed_rows(A_rank.col, l, r0)
    A_ = HostCSR(A_rank.rpt.copy(), a_newcol, A_rank.val.copy(), A_rank.nrows, grows.size)
    prpt, pcol, pval = gather_rows(B_full, grows, r0)
    perm_p, p_newcol = shrink(pcol, grows.size - 1)
    P_ = HostCSR(prpt, p_newcol, pval, grows.size, perm_p.size)
    return {"A_": A_, "P_": P_, "grows": grows, "perm_P": perm_p, "completedP_col": pcol}


def transpose(X):
    rpt, col, val = _csr(X)
    trpt = np.empty(X.ncols + 1, np.int64)
    tcol = np.empty(X.nnz, np.int32)
    tval = np.empty(X.nnz, np.float64)
    lib().orc_transpose(X.nrows, X.ncols, rpt, col, val, trpt, tcol, _ptr(tval))
    return trpt, tcol, tval


def backward(A, B, crpt, ccol, dc):
    arpt, acol, aval = _csr(A)
    brpt, bcol, bval = _csr(B)
    da = np.empty(A.nnz, np.float64)
    db = np.empty(B.nnz, np.float64)
    lib().orc_backward(A.nrows, arpt, acol, aval, B.nrows, brpt, bcol, bval,
                       np.ascontiguousarray(crpt, np.int64), np.ascontiguousarray(ccol, np.int32),
                       np.ascontiguousarray(dc, np.float64), da, db)
    return da, db


def read_binary_coo(path, nprocs=1, rank=0, header_type=1):
    """Rank `rank`'s CSR block: (nr, nc, nt, row_shift, HostCSR) with columns
    in the rank frame, as load_MatrixMPI returns it."""
    from nsp_amd.gen import HostCSR
    vals = [C.c_int64() for _ in range(6)]
    ptrs = [C.c_void_p() for _ in range(3)]
    rc = lib().orc_read_binary_coo(path.encode(), header_type, nprocs, rank,
                                   *[C.byref(v) for v in vals], *[C.byref(p) for p in ptrs])
    if rc != 0:
        raise IOError(f"orc_read_binary_coo failed ({rc}) on {path}")
    nr, nc, nt, nloc, shift, nnz = (v.value for v in vals)
    rpt = np.ctypeslib.as_array(C.cast(ptrs[0], C.POINTER(C.c_int64)), (nloc + 1,)).copy()
    col = np.ctypeslib.as_array(C.cast(ptrs[1], C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(C.cast(ptrs[2], C.POINTER(C.c_double)), (max(nnz, 1),))[:nnz].copy()
    for p in ptrs:
        lib().orc_free(p)
    return nr, nc, nt, shift, HostCSR(rpt, col, val, nloc, nc)


def read_ascii_coo(path, nprocs=1, rank=0):
    """Rank `rank`'s CSR block of an ASCII COO file: (nr, nc, nt, row_shift,
    HostCSR), as load_MatrixMPI returns it with `ascii` set."""
    from nsp_amd.gen import HostCSR
    vals = [C.c_int64() for _ in range(6)]
    ptrs = [C.c_void_p() for _ in range(3)]
    rc = lib().orc_read_ascii_coo(path.encode(), nprocs, rank, *[C.byref(v) for v in vals],
                                  *[C.byref(p) for p in ptrs])
    if rc != 0:
        raise IOError(f"orc_read_ascii_coo failed ({rc}) on {path}")
    nr, nc, nt, nloc, shift, nnz = (v.value for v in vals)
    rpt = np.ctypeslib.as_array(C.cast(ptrs[0], C.POINTER(C.c_int64)), (nloc + 1,)).copy()
    col = np.ctypeslib.as_array(C.cast(ptrs[1], C.POINTER(C.c_int32)), (max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(C.cast(ptrs[2], C.POINTER(C.c_double)), (max(nnz, 1),))[:nnz].copy()
    for p in ptrs:
        lib().orc_free(p)
    return nr, nc, nt, shift, HostCSR(rpt, col, val, nloc, nc)


def fast_atof(text: str) -> float:
    """The reference's decimal parser on one field."""
    b = text.encode()
    buf = C.create_string_buffer(b, len(b))
    end = C.cast(C.addressof(buf) + len(b), C.c_char_p)
    return lib().orc_fast_atof(buf, end)


def write_binary_coo(path, X, header_type=1, col_shift=0):
    rpt, col, val = _csr(X)
    rc = lib().orc_write_binary_coo(path.encode(), header_type, X.nrows, X.ncols, rpt, col, val,
                                    int(col_shift))
    if rc != 0:
        raise IOError(f"orc_write_binary_coo failed ({rc}) on {path}")


def componentwise_error(val, ref, ref_abs):
    """max_ij |c - c*| / sum_k |a_ik b_kj| (0/0 counts as 0) and
    ||C - C*||_F / ||C*||_F."""
    val = np.asarray(val, np.float64)
    diff = np.abs(val - ref)
    denom = np.where(ref_abs > 0, ref_abs, 1.0)
    comp = float((diff / denom).max()) if diff.size else 0.0
    nref = float(np.linalg.norm(ref))
    fro = float(np.linalg.norm(val - ref) / nref) if nref > 0 else float(np.linalg.norm(val - ref))
    return comp, fro
