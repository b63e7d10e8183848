"""torch.autograd wrapper of the local product on fixed sparsity patterns.

    op = SpGEMMFixedPattern(A_pattern, B_pattern)   # DeviceCSR patterns
    c_val = op(a_val, b_val)                        # values of C = A*B on op.C
    loss = 0.5 * ((c_val - t) ** 2).sum()
    loss.backward()                                 # a_val.grad, b_val.grad

C's pattern is computed once (symbolic phase + scan) at construction; each
forward is the numeric phase alone on that plan, so a forward costs one
hash-accumulate pass and no host synchronisation.  The backward is
nsp_spgemm_backward:
    dA = (dC B^T) on pat(A),   dB = (A^T dC) on pat(B)
with A^T's pattern cached: the transpose is taken once of A's pattern with
values 0..nnz-1, which gives the permutation from A^T's entries to A's; the
backward reads a_val through it.  The backward runs on a second
handle: its row-extent pass would otherwise replace the forward's plan.
No reference counterpart (the reference has no backward).
"""
from __future__ import annotations

import torch

from . import DeviceCSR, Handle


class _SpGEMMFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_val, b_val, op):
        a_val, b_val = a_val.contiguous(), b_val.contiguous()
        ctx.op = op
        ctx.save_for_backward(a_val, b_val)
        return op.forward_values(a_val, b_val)

    @staticmethod
    def backward(ctx, dc):
        op = ctx.op
        a_val, b_val = ctx.saved_tensors
        da, db = op.backward_values(a_val, b_val, dc.contiguous())
        return da, db, None


class SpGEMMFixedPattern:
    """C = A*B as a differentiable function of A's and B's values."""

    def __init__(self, A: DeviceCSR, B: DeviceCSR, handle: Handle | None = None):
        self.h = handle or Handle()
        self.hb = Handle(self.h.stream)
        self.hb.set_param("reuse_extents", 1)  # A's and B's patterns are fixed
        self.A, self.B = A, B
        rpt, nnz = self.h.symbolic(A, B)
        self.C = self.h.numeric(A, B, rpt, nnz)
        # A^T pattern + value permutation, computed once
        idx = torch.arange(A.nnz, dtype=torch.float64, device=A.col.device)
        AT = self.hb.transpose(DeviceCSR(A.row_ptr, A.col, idx, A.nrows, A.ncols))
        self.at_perm = AT.val.to(A.row_ptr.dtype)
        self.AT = DeviceCSR(AT.row_ptr, AT.col, None, AT.nrows, AT.ncols)

    def forward_values(self, a_val, b_val):
        """Values of C on the cached pattern (numeric phase only)."""
        A = DeviceCSR(self.A.row_ptr, self.A.col, a_val, self.A.nrows, self.A.ncols)
        B = DeviceCSR(self.B.row_ptr, self.B.col, b_val, self.B.nrows, self.B.ncols)
        out = DeviceCSR(self.C.row_ptr, self.C.col, torch.empty_like(self.C.val), self.C.nrows,
                        self.C.ncols)
        return self.h.numeric(A, B, self.C.row_ptr, self.C.nnz, out=out).val

    def backward_values(self, a_val, b_val, dc):
        A = DeviceCSR(self.A.row_ptr, self.A.col, a_val, self.A.nrows, self.A.ncols)
        B = DeviceCSR(self.B.row_ptr, self.B.col, b_val, self.B.nrows, self.B.ncols)
        return self.hb.backward(A, B, self.AT, self.C, dc, at_perm=self.at_perm)

    def __call__(self, a_val, b_val):
        return _SpGEMMFn.apply(a_val, b_val, self)
