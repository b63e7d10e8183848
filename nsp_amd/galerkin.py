"""Repeated Galerkin products C = P^T (A P) on fixed patterns.

The reference's driver names R(AP) as the product it serves (main.cu:138);
in an AMG setup the coarse operator is rebuilt whenever the values of A (or
P) change while their patterns stay.  ``GalerkinPlan`` records A*P and
P^T*(AP) as two hipGraphs once (``nsp_graph_create``; the second reads the
first's C buffers) and caches P^T's pattern with the permutation from P^T's
entries to P's, so a rebuild is one value gather and two graph launches on
the handle's stream with no host synchronisation; ``result`` synchronises.
``nsp_galerkin`` (``Handle.galerkin``) is the one-shot form.
"""
from __future__ import annotations

from . import DeviceCSR, Handle, _torch


class GalerkinPlan:
    def __init__(self, A: DeviceCSR, P: DeviceCSR, handle: Handle | None = None):
        torch = _torch()
        if A.ncols != P.nrows or A.nrows != P.nrows:
            raise ValueError("GalerkinPlan: A must be square with as many rows as P")
        self.h = handle or Handle()
        self.A, self.P = A, P
        dev = P.col.device
        # P^T's pattern and, through values that number P's entries, the
        # permutation from P^T's entries to P's (exact in FP64 below 2^53)
        idx = torch.arange(P.nnz, dtype=torch.float64, device=dev)
        PTi = self.h.transpose(DeviceCSR(P.row_ptr, P.col, idx, P.nrows, P.ncols))
        self.pt_perm = PTi.val.to(torch.int64)
        self.PT = DeviceCSR(PTi.row_ptr, PTi.col, torch.empty(P.nnz, dtype=P.val.dtype, device=dev),
                            P.ncols, P.nrows)
        # sizes of A*P and of the coarse operator from one eager product each
        AP = self.h.spgemm(A, P)
        torch.index_select(P.val, 0, self.pt_perm, out=self.PT.val)
        C = self.h.spgemm(self.PT, AP)
        # a recorded graph launches every bin's kernels whether or not the
        # pattern puts rows in them; the plan's graphs are recorded with 8
        # blocks per CU so that the empty launches stay short (a handle on
        # the same stream with the caller's parameters)
        gh = Handle(stream=self.h.stream)
        for k, v in self.h.params.items():
            gh.set_param(k, v)
        gh.set_param("grid_cap", min(8, self.h.params.get("grid_cap", 8)))
        self.g_ap = gh.graph(A, P, capacity=AP.nnz)
        self.AP = DeviceCSR(self.g_ap.C.row_ptr, self.g_ap.C.col[:AP.nnz], self.g_ap.C.val[:AP.nnz],
                            A.nrows, P.ncols)
        self.g_c = gh.graph(self.PT, self.AP, capacity=C.nnz)
        self._gh = gh
        self.nnz_ap, self.nnz = AP.nnz, C.nnz

    def launch(self):
        """Rebuilds the coarse operator from the current values of A.val and
        P.val (update them in place between launches); asynchronous."""
        torch = _torch()
        torch.index_select(self.P.val, 0, self.pt_perm, out=self.PT.val)
        self.g_ap.launch()
        self.g_c.launch()

    def result(self) -> DeviceCSR:
        """Synchronises; C of the last launch (row pointer, columns, values)."""
        ap = self.g_ap.result()
        if ap.nnz != self.nnz_ap:
            raise RuntimeError("GalerkinPlan: the pattern of A*P changed")
        return self.g_c.result()
