"""torch.autograd wrapper of the local product on fixed sparsity patterns.

    op = SpGEMMFixedPattern(A_pattern, B_pattern)   # DeviceCSR patterns
    c_val = op(a_val, b_val)                        # values of C = A*B on op.C
    loss = 0.5 * ((c_val - t) ** 2).sum()
    loss.backward()                                 # a_val.grad, b_val.grad

The forward is the library's symbolic + numeric product; C's pattern is
computed once.  The backward is nsp_spgemm_backward:
    dA = (dC B^T) on pat(A),   dB = (A^T dC) on pat(B)
with A^T's pattern cached: the transpose is taken once of A's pattern with
values 0..nnz-1, which gives the permutation that maps A's values onto A^T's
(each backward gathers a_val through it).  No reference counterpart (the
reference has no backward).
"""
from __future__ import annotations

import torch

from . import DeviceCSR, Handle, default_handle


class _SpGEMMFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_val, b_val, op):
        A = DeviceCSR(op.A.row_ptr, op.A.col, a_val, op.A.nrows, op.A.ncols)
        B = DeviceCSR(op.B.row_ptr, op.B.col, b_val, op.B.nrows, op.B.ncols)
        C = op.h.spgemm(A, B)
        ctx.op = op
        ctx.save_for_backward(a_val, b_val)
        return C.val

    @staticmethod
    def backward(ctx, dc):
        op = ctx.op
        a_val, b_val = ctx.saved_tensors
        A = DeviceCSR(op.A.row_ptr, op.A.col, a_val, op.A.nrows, op.A.ncols)
        B = DeviceCSR(op.B.row_ptr, op.B.col, b_val, op.B.nrows, op.B.ncols)
        AT = DeviceCSR(op.AT.row_ptr, op.AT.col, a_val[op.at_perm], op.AT.nrows, op.AT.ncols)
        da, db = op.h.backward(A, B, AT, op.C, dc.contiguous())
        return da, db, None


class SpGEMMFixedPattern:
    """C = A*B as a differentiable function of A's and B's values."""

    def __init__(self, A: DeviceCSR, B: DeviceCSR, handle: Handle | None = None):
        self.h = handle or default_handle()
        self.A, self.B = A, B
        C = self.h.spgemm(A, B)
        self.C = C
        # A^T pattern + value permutation, computed once
        idx = torch.arange(A.nnz, dtype=torch.float64, device=A.col.device)
        AT = self.h.transpose(DeviceCSR(A.row_ptr, A.col, idx, A.nrows, A.ncols))
        self.at_perm = AT.val.to(torch.int64)
        self.AT = DeviceCSR(AT.row_ptr, AT.col, AT.val.to(A.val.dtype), AT.nrows, AT.ncols)

    def __call__(self, a_val, b_val):
        return _SpGEMMFn.apply(a_val, b_val, self)
