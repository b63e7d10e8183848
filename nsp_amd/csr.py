"""The reference's CSR type and per-rank entry points, on one device.

``CSR`` keeps the field names of the reference's ``CSRtype`` (src/CSR.h:
67-104) that the local product reads or writes; device arrays are torch
tensors.  The entry points mirror src/spspmpi.cu:

  get_shrinked_matrix(A, P)            spspmpi.cu:166-182 (shrink_col 141-162)
  nsparse_mgpu_version0(A, P)          spspmpi.cu:266-313
  nsparse_mgpu_version1_local(A, Pfull) spspmpi.cu:371-975 without MPI: the
      halo rows come from the full B on the same device (a gather) instead of
      MPI_Alltoallv, then A and completedP are compacted and multiplied.

Row partition of a rank (src/READ_MATRIX_MPI.cpp:256-270): nr/nprocs rows
per rank, the remainder to the last rank; a rank's rows run from its first
to its last present row and its columns are shifted by the rows of the ranks
before it (src/distribute.cu:1094-1105).

  load_MatrixMPI(path, nprocs, rank)   distribute.cu:1071-1125 (binary or ASCII files,
      row partition): the rank's block built on the device.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from . import DeviceCSR, Handle, default_handle


@dataclass
class CSR:
    nnz: int
    n: int                     # rows
    m: int                     # columns
    full_n: int                # rows of the whole matrix
    row_shift: int             # first global row of this block
    row: "object" = None       # row pointer (device)
    col: "object" = None       # columns (device; rank frame for a block)
    val: "object" = None       # values (device)
    shrinked_m: int = 0
    shrinked_col: "object" = None
    shrinked_flag: bool = False
    on_the_device: bool = True
    extra: dict = field(default_factory=dict)

    def device_csr(self, compacted: bool = False) -> DeviceCSR:
        if compacted:
            return DeviceCSR(self.row, self.shrinked_col, self.val, self.n, self.shrinked_m)
        return DeviceCSR(self.row, self.col, self.val, self.n, self.m)

    @classmethod
    def from_device(cls, X: DeviceCSR, full_n: int | None = None, row_shift: int = 0) -> "CSR":
        return cls(X.nnz, X.nrows, X.ncols, X.nrows if full_n is None else full_n, row_shift,
                   X.row_ptr, X.col, X.val)


def load_MatrixMPI(path: str, nprocs: int = 1, rank: int = 0, header_type: int = 1,
                   h: Handle | None = None, ascii: bool = False) -> CSR:
    """Rank `rank`'s block of a binary COO file (or, with `ascii`, of an
    ASCII one: distribute.cu:1084-1086) as a device CSR (int32 row pointer,
    FP64), row_shift = the rows of the ranks before it."""
    h = h or default_handle()
    if ascii:
        s, X = h.ascii_read(path, nprocs, rank)
    else:
        s, X = h.coo_read(path, nprocs, rank, header_type)
    return CSR.from_device(X, full_n=int(s.nr), row_shift=int(s.col_shift))


def rank_rows(nr: int, nprocs: int, rank: int) -> tuple[int, int]:
    """[first, last) rows of a rank under the reference's row split."""
    per = nr // nprocs
    first = per * rank
    last = nr if rank == nprocs - 1 else per * (rank + 1)
    return first, last


def shrink_col(A: CSR, P: CSR | None = None, h: Handle | None = None) -> bool:
    """Compacts A's columns once (spspmpi.cu:141-162): l = P.n - 1 when P is
    given (the rank's local row count of P), else A.n - 1."""
    if A.shrinked_flag:
        return False
    h = h or default_handle()
    l = (P.n if P is not None else A.n) - 1
    lo, hi = -A.row_shift, A.m - A.row_shift
    if "col_bounds" in A.extra:
        lo, hi = A.extra["col_bounds"]
    new_col, perm = h.shrink_cols(A.col, l, lo, hi)
    A.shrinked_col = new_col
    A.shrinked_m = int(perm.numel())
    A.shrinked_flag = True
    A.extra["shrinking_permut"] = perm
    return True


def get_shrinked_matrix(A: CSR, P: CSR | None = None, h: Handle | None = None) -> CSR:
    """A with compacted columns (spspmpi.cu:166-182)."""
    shrink_col(A, P, h)
    return CSR(A.nnz, A.n, A.shrinked_m, A.full_n, A.row_shift, A.row, A.shrinked_col, A.val)


def nsparse_mgpu_version0(A: CSR, P: CSR, h: Handle | None = None):
    """C = A * P on the device; returns (C, loc_time in seconds, the
    kernel-side time of the product as the reference's loc_time measures it,
    spspmpi.cu:289-305)."""
    import torch
    h = h or default_handle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Cd = h.spgemm(A.device_csr(), P.device_csr())
    torch.cuda.synchronize()
    loc_time = time.perf_counter() - t0
    C = CSR(Cd.nnz, Cd.nrows, Cd.ncols, A.full_n, A.row_shift, Cd.row_ptr, Cd.col, Cd.val)
    return C, loc_time


def nsparse_mgpu_version1_local(Alocal: CSR, Pfull: DeviceCSR, p_local_n: int | None = None,
                                h: Handle | None = None):
    """One rank's local product on one device (spspmpi.cu:371-975, no MPI).

    Alocal: the rank's row block of A with columns in the rank frame
            (global - row_shift), as load_MatrixMPI returns it.
    Pfull:  every row of P (global columns) on the same device -- the rows the
            reference fetches from the other ranks are gathered from it.
    p_local_n: rows of P the rank owns (defaults to Alocal.n, the reference's
            equal split of A and P).
    Returns (C, A_, P_, loc_time): C's columns are compacted P_ columns, as
    the reference returns them."""
    h = h or default_handle()
    r0 = Alocal.row_shift
    if p_local_n is None:
        p_local_n = Alocal.n
    Ploc = CSR(0, p_local_n, Pfull.ncols, Pfull.nrows, r0)
    # halo discovery + A's shrink: completedP rows = shrinking_permut + r0
    Alocal.extra.setdefault("col_bounds", (-r0, Pfull.nrows - r0))
    A_ = get_shrinked_matrix(Alocal, Ploc, h)
    perm_a = Alocal.extra["shrinking_permut"]
    # completedP: rows [pre-halo | local | post-halo] ascending = perm_a + r0,
    # columns re-shifted to the rank frame
    Pd = h.gather_rows(Pfull, perm_a, row_offset=r0, col_shift=r0)
    completedP = CSR(Pd.nnz, Pd.nrows, Pfull.ncols, Pfull.nrows, r0, Pd.row_ptr, Pd.col, Pd.val)
    completedP.extra["col_bounds"] = (-r0, Pfull.ncols - r0)
    assert A_.m == completedP.n
    P_ = get_shrinked_matrix(completedP, None, h)
    C, loc_time = nsparse_mgpu_version0(A_, P_, h)
    C.extra["shrinking_permut_P"] = completedP.extra["shrinking_permut"]
    return C, A_, P_, loc_time
