"""nsp_amd -- MI355X (gfx950) local sparse product C = A*B.

Python host side over the C ABI of libnsp_amd.so.  torch supplies device
memory and the stream; the compute is the library's HIP kernels.

    import nsp_amd
    A = nsp_amd.DeviceCSR.from_host(host_csr)           # FP64, int32 row_ptr
    C = nsp_amd.spgemm(A, A)                            # sorted rows, zeros kept

The reference's entry points this mirrors: nsp_spgemm_kernel_hash(sfCSR*,
sfCSR*, sfCSR*) (nsp/src/nsp.cu:2448) -> ``spgemm``; the per-rank wrapper
nsparseMGPU_version0/1 (src/spspmpi.cu:266, 371) -> ``nsparse_mgpu_version0``
and ``nsparse_mgpu_version1_local`` in ``nsp_amd.csr``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import NSP_F32, NSP_F64, NSP_IDX32, NSP_IDX64, NspError


def _torch():
    import torch
    return torch


@dataclass
class DeviceCSR:
    """CSR on the GPU: row_ptr int32/int64, col int32, val float64/float32."""
    row_ptr: "object"
    col: "object"
    val: "object"
    nrows: int
    ncols: int

    @property
    def nnz(self) -> int:
        return int(self.col.numel())

    @property
    def index_type(self) -> int:
        return NSP_IDX64 if self.row_ptr.dtype == _torch().int64 else NSP_IDX32

    @property
    def value_type(self) -> int:
        return NSP_F32 if self.val.dtype == _torch().float32 else NSP_F64

    @classmethod
    def from_host(cls, h, dtype: str = "f64", index: str = "i32", device: str = "cuda"):
        torch = _torch()
        vt = torch.float32 if dtype == "f32" else torch.float64
        it = torch.int64 if index == "i64" else torch.int32
        return cls(torch.from_numpy(np.asarray(h.rpt)).to(device=device, dtype=it),
                   torch.from_numpy(np.asarray(h.col, np.int32)).to(device),
                   torch.from_numpy(np.asarray(h.val, np.float64)).to(device=device, dtype=vt),
                   int(h.nrows), int(h.ncols))

    def to_host(self):
        from .gen import HostCSR
        return HostCSR(self.row_ptr.cpu().numpy().astype(np.int64), self.col.cpu().numpy(),
                       self.val.cpu().numpy().astype(np.float64), self.nrows, self.ncols)

    def c_struct(self) -> _lib.nsp_csr_t:
        return _lib.nsp_csr_t(self.nrows, self.ncols, self.nnz, self.row_ptr.data_ptr(),
                              self.col.data_ptr() if self.nnz else None,
                              self.val.data_ptr() if self.nnz and self.val is not None else None)


class Handle:
    """Owns an nsp handle (workspace + stream binding).  By default work is
    ordered on torch's current stream of the current device."""

    def __init__(self, stream=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("nsp_amd needs a GPU (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.h = C.c_void_p()
        self._check(self.L.nsp_create(C.byref(self.h)))
        s = stream if stream is not None else torch.cuda.current_stream()
        self.stream = s
        self._check(self.L.nsp_set_stream(self.h, C.c_void_p(s.cuda_stream)))
        self.params = {}  # parameters set on this handle (name -> value)

    def _check(self, status: int):
        if status != _lib.NSP_OK:
            msg = self.L.nsp_last_error(self.h).decode() if getattr(self, "h", None) else ""
            raise NspError(status, msg)

    def close(self):
        if getattr(self, "h", None):
            self.L.nsp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_param(self, name: str, value: int):
        self._check(self.L.nsp_set_param(self.h, name.encode(), int(value)))
        self.params[name] = int(value)

    def stats(self) -> dict:
        st = _lib.nsp_stats_t()
        self._check(self.L.nsp_get_stats(self.h, C.byref(st)))
        return {"nprod": st.nprod, "nnz": st.nnz, "max_intprod": st.max_intprod,
                "max_row_nnz": st.max_row_nnz, "sym_bin": list(st.sym_bin[:st.n_sym_bins]),
                "num_bin": list(st.num_bin[:st.n_num_bins]), "direct_mode": st.direct_mode,
                "host_syncs": st.host_syncs, "sym_launches": st.sym_launches,
                "num_launches": st.num_launches, "t_bin_ms": st.t_bin_ms,
                "t_sym_ms": st.t_sym_ms, "t_scan_ms": st.t_scan_ms, "t_num_ms": st.t_num_ms,
                "t_coo_ms": st.t_coo_ms, "fused": st.fused, "t_compact_ms": st.t_compact_ms}

    # -- local product ---------------------------------------------------
    def symbolic(self, A: DeviceCSR, B: DeviceCSR):
        """Returns (C row_ptr tensor, nnz)."""
        torch = _torch()
        _same_types(A, B)
        rpt = torch.empty(A.nrows + 1, dtype=A.row_ptr.dtype, device=A.row_ptr.device)
        nnz = C.c_int64()
        a, b = A.c_struct(), B.c_struct()
        self._check(self.L.nsp_spgemm_symbolic(self.h, A.index_type, A.value_type, C.byref(a),
                                               C.byref(b), C.c_void_p(rpt.data_ptr()),
                                               C.byref(nnz)))
        return rpt, int(nnz.value)

    def numeric(self, A: DeviceCSR, B: DeviceCSR, rpt, nnz: int, out=None) -> DeviceCSR:
        torch = _torch()
        if out is None:
            col = torch.empty(nnz, dtype=torch.int32, device=A.col.device)
            val = torch.empty(nnz, dtype=A.val.dtype, device=A.val.device)
            out = DeviceCSR(rpt, col, val, A.nrows, B.ncols)
        a, b, c = A.c_struct(), B.c_struct(), out.c_struct()
        self._check(self.L.nsp_spgemm_numeric(self.h, A.index_type, A.value_type, C.byref(a),
                                              C.byref(b), C.byref(c)))
        return out

    def spgemm(self, A: DeviceCSR, B: DeviceCSR) -> DeviceCSR:
        rpt, nnz = self.symbolic(A, B)
        return self.numeric(A, B, rpt, nnz)

    def spgemm_once(self, A: DeviceCSR, B: DeviceCSR) -> DeviceCSR:
        """One-call product with C allocated by the library through torch's
        allocator (nsp_spgemm, the reference's nsp_spgemm_kernel_hash): the
        symbolic pass is skipped when the row sample allows."""
        torch = _torch()
        _same_types(A, B)
        dev = A.col.device
        keep = {}

        def _alloc(nbytes, user):
            t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
            keep[t.data_ptr()] = t
            return t.data_ptr()

        def _free(ptr, user):
            keep.pop(ptr, None)

        alloc, free = _lib.ALLOC_FN(_alloc), _lib.FREE_FN(_free)
        a, b = A.c_struct(), B.c_struct()
        c = _lib.nsp_csr_t()
        self._check(self.L.nsp_spgemm(self.h, A.index_type, A.value_type, C.byref(a), C.byref(b), C.byref(c),
                                      C.cast(alloc, C.c_void_p), C.cast(free, C.c_void_p), None))
        nnz, rows = int(c.nnz), int(c.rows)
        rpt = keep[c.row_ptr].view(A.row_ptr.dtype)[:rows + 1]
        col = keep[c.col].view(torch.int32)[:nnz] if c.col else torch.empty(0, dtype=torch.int32, device=dev)
        val = keep[c.val].view(A.val.dtype)[:nnz] if c.val else torch.empty(0, dtype=A.val.dtype, device=dev)
        return DeviceCSR(rpt, col, val, rows, int(c.cols))

    def spgemm_into(self, A: DeviceCSR, B: DeviceCSR, out: DeviceCSR | None = None,
                    capacity: int | None = None) -> DeviceCSR:
        """One-call product (nsp_spgemm_into) into `out`'s buffers (row
        pointer of A.nrows+1 entries, col/val of their length as the
        capacity), or into new buffers of `capacity` entries; a C larger
        than the capacity is counted, the buffers are regrown and the call
        repeated.  Returns C as a DeviceCSR viewing the first nnz(C) entries."""
        torch = _torch()
        _same_types(A, B)
        dev = A.col.device
        if out is None:
            cap = int(capacity or 0)
            out = DeviceCSR(torch.empty(A.nrows + 1, dtype=A.row_ptr.dtype, device=dev),
                            torch.empty(cap, dtype=torch.int32, device=dev),
                            torch.empty(cap, dtype=A.val.dtype, device=dev), A.nrows, B.ncols)
        for _ in range(2):
            c = _lib.nsp_csr_t(A.nrows, B.ncols, out.col.numel(), out.row_ptr.data_ptr(),
                               out.col.data_ptr() if out.col.numel() else None,
                               out.val.data_ptr() if out.val.numel() else None)
            nnz = C.c_int64()
            a, b = A.c_struct(), B.c_struct()
            st = self.L.nsp_spgemm_into(self.h, A.index_type, A.value_type, C.byref(a), C.byref(b),
                                        C.byref(c), C.byref(nnz))
            if st == _lib.NSP_ERR_OVERFLOW and nnz.value > out.col.numel():  # regrow and repeat
                out = DeviceCSR(out.row_ptr, torch.empty(nnz.value, dtype=torch.int32, device=dev),
                                torch.empty(nnz.value, dtype=A.val.dtype, device=dev), A.nrows, B.ncols)
                continue
            self._check(st)
            n = int(nnz.value)
            return DeviceCSR(out.row_ptr, out.col[:n], out.val[:n], A.nrows, B.ncols)
        raise NspError(_lib.NSP_ERR_OVERFLOW, "C grew between two counts")

    def intprod(self, A: DeviceCSR, B: DeviceCSR):
        torch = _torch()
        out = torch.empty(max(A.nrows, 1), dtype=torch.int64, device=A.col.device)
        nprod = C.c_int64()
        a, b = A.c_struct(), B.c_struct()
        self._check(self.L.nsp_intprod(self.h, A.index_type, C.byref(a), C.byref(b),
                                       C.c_void_p(out.data_ptr()), C.byref(nprod)))
        return out[:A.nrows], int(nprod.value)

    def check_csr(self, X: DeviceCSR) -> int:
        bad = C.c_int64()
        x = X.c_struct()
        self._check(self.L.nsp_check_csr(self.h, X.index_type, C.byref(x), C.byref(bad)))
        return int(bad.value)

    # -- batched products and graphs ------------------------------------------
    def batched_symbolic(self, A: DeviceCSR, a_off, B: DeviceCSR, b_off):
        """Symbolic phase of the batch (concatenated rows, int64 device row
        offsets); returns (C row_ptr, nnz); follow with numeric()."""
        torch = _torch()
        _same_types(A, B)
        rpt = torch.empty(A.nrows + 1, dtype=A.row_ptr.dtype, device=A.row_ptr.device)
        nnz = C.c_int64()
        a, b = A.c_struct(), B.c_struct()
        self._check(self.L.nsp_spgemm_batched_symbolic(
            self.h, A.index_type, A.value_type, C.byref(a), C.c_void_p(a_off.data_ptr()), C.byref(b),
            C.c_void_p(b_off.data_ptr()), a_off.numel() - 1, C.c_void_p(rpt.data_ptr()), C.byref(nnz)))
        return rpt, int(nnz.value)

    def batched_spgemm(self, A: DeviceCSR, a_off, B: DeviceCSR, b_off) -> DeviceCSR:
        rpt, nnz = self.batched_symbolic(A, a_off, B, b_off)
        return self.numeric(A, B, rpt, nnz)

    def graph(self, A: DeviceCSR, B: DeviceCSR, capacity: int, a_off=None, b_off=None) -> "Graph":
        return Graph(self, A, B, capacity, a_off, b_off)

    # -- backward pair -----------------------------------------------------
    def transpose(self, X: DeviceCSR) -> DeviceCSR:
        torch = _torch()
        XT = DeviceCSR(torch.empty(X.ncols + 1, dtype=X.row_ptr.dtype, device=X.col.device),
                       torch.empty(X.nnz, dtype=torch.int32, device=X.col.device),
                       torch.empty(X.nnz, dtype=X.val.dtype, device=X.col.device), X.ncols, X.nrows)
        x, xt = X.c_struct(), XT.c_struct()
        self._check(self.L.nsp_transpose(self.h, X.index_type, X.value_type, C.byref(x), C.byref(xt)))
        return XT

    def backward(self, A: DeviceCSR, B: DeviceCSR, AT: DeviceCSR, Cpat: DeviceCSR, dC, at_perm=None):
        """(dA, dB) aligned with A.col and B.col for dC aligned with Cpat.col.
        With at_perm (A^T entry -> A entry, A's index dtype) A's values come
        from A.val through it and AT.val is not read."""
        torch = _torch()
        dA = torch.empty(A.nnz, dtype=A.val.dtype, device=A.col.device)
        dB = torch.empty(B.nnz, dtype=B.val.dtype, device=B.col.device)
        a, b, at, c = A.c_struct(), B.c_struct(), AT.c_struct(), Cpat.c_struct()
        if at_perm is not None:
            if at_perm.dtype != A.row_ptr.dtype or at_perm.numel() != A.nnz:
                raise ValueError("at_perm must have A.nnz entries of A's index dtype")
        perm_p = C.c_void_p(at_perm.data_ptr() if at_perm is not None and at_perm.numel() else None)
        self._check(self.L.nsp_spgemm_backward(
            self.h, A.index_type, A.value_type, C.byref(a), C.byref(b), C.byref(at), perm_p, C.byref(c),
            C.c_void_p(dC.data_ptr() if dC.numel() else None),
            C.c_void_p(dA.data_ptr() if dA.numel() else None),
            C.c_void_p(dB.data_ptr() if dB.numel() else None)))
        return dA, dB

    # -- one rank's local share -------------------------------------------
    def shrink_cols(self, col, l: int, lo: int, hi: int):
        """(new_col, perm): perm = sorted unique(col U [0..l]) (int32 device
        tensors); every column must lie in [lo, hi)."""
        torch = _torch()
        cap = max(hi, l + 1) - min(lo, 0)
        new_col = torch.empty_like(col)
        perm = torch.empty(max(cap, 1), dtype=torch.int32, device=col.device)
        m = C.c_int64()
        self._check(self.L.nsp_shrink_cols(self.h, col.numel(), C.c_void_p(col.data_ptr()), int(l),
                                           int(lo), int(hi), C.c_void_p(new_col.data_ptr()),
                                           C.c_void_p(perm.data_ptr()), cap, C.byref(m)))
        return new_col, perm[:m.value]

    def gather_rows(self, B: DeviceCSR, rows, row_offset: int = 0, col_shift: int = 0,
                    ncols: int | None = None) -> DeviceCSR:
        """CSR of B rows rows[i] + row_offset, columns minus col_shift."""
        torch = _torch()
        nr = rows.numel()
        rpt = torch.empty(nr + 1, dtype=B.row_ptr.dtype, device=B.col.device)
        nnz = C.c_int64()
        b = B.c_struct()
        self._check(self.L.nsp_gather_rows_count(self.h, B.index_type, C.byref(b), nr,
                                                 C.c_void_p(rows.data_ptr() if nr else None),
                                                 int(row_offset), C.c_void_p(rpt.data_ptr()),
                                                 C.byref(nnz)))
        out = DeviceCSR(rpt, torch.empty(nnz.value, dtype=torch.int32, device=B.col.device),
                        torch.empty(nnz.value, dtype=B.val.dtype, device=B.col.device), nr,
                        B.ncols if ncols is None else ncols)
        o = out.c_struct()
        self._check(self.L.nsp_gather_rows_fill(self.h, B.index_type, B.value_type, C.byref(b), nr,
                                                C.c_void_p(rows.data_ptr() if nr else None),
                                                int(row_offset), int(col_shift), C.byref(o)))
        return out

    # -- chained products ---------------------------------------------------
    def galerkin(self, A: DeviceCSR, P: DeviceCSR, R: DeviceCSR | None = None) -> DeviceCSR:
        """C = R (A P), R = P^T (built on the device) when R is None
        (nsp_galerkin); C's arrays live in torch's allocator."""
        torch = _torch()
        dev = A.col.device
        keep = {}

        def _alloc(nbytes, user):
            t = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
            keep[t.data_ptr()] = t
            return t.data_ptr()

        def _free(ptr, user):
            keep.pop(ptr, None)

        alloc, free = _lib.ALLOC_FN(_alloc), _lib.FREE_FN(_free)
        a, p = A.c_struct(), P.c_struct()
        r = R.c_struct() if R is not None else None
        c = _lib.nsp_csr_t()
        self._check(self.L.nsp_galerkin(self.h, A.index_type, A.value_type, C.byref(r) if r is not None else None,
                                        C.byref(a), C.byref(p), C.byref(c), C.cast(alloc, C.c_void_p),
                                        C.cast(free, C.c_void_p), None))
        nnz, rows = int(c.nnz), int(c.rows)
        rpt = keep[c.row_ptr].view(A.row_ptr.dtype)[:rows + 1]
        col = keep[c.col].view(torch.int32)[:nnz] if c.col else torch.empty(0, dtype=torch.int32, device=dev)
        val = keep[c.val].view(A.val.dtype)[:nnz] if c.val else torch.empty(0, dtype=A.val.dtype, device=dev)
        return DeviceCSR(rpt, col, val, rows, int(c.cols))

    # -- binary COO files ---------------------------------------------------
    def coo_to_csr(self, records, first_row: int, nrows: int, col_shift: int = 0, ncols: int = 0,
                   index: str = "i32", dtype: str = "f64") -> DeviceCSR:
        """CSR of 24-byte {int64 i, int64 j, double v} records (a uint8 or
        int64 device tensor holding them), rows first_row .. first_row +
        nrows - 1, columns j - 1 - col_shift (nsp_coo_to_csr)."""
        torch = _torch()
        nrec = records.numel() * records.element_size() // 24
        dev = records.device
        out = DeviceCSR(torch.empty(nrows + 1, dtype=torch.int64 if index == "i64" else torch.int32, device=dev),
                        torch.empty(nrec, dtype=torch.int32, device=dev),
                        torch.empty(nrec, dtype=torch.float32 if dtype == "f32" else torch.float64, device=dev),
                        int(nrows), int(ncols))
        o = out.c_struct()
        self._check(self.L.nsp_coo_to_csr(self.h, out.index_type, out.value_type,
                                          C.c_void_p(records.data_ptr() if nrec else None), nrec,
                                          int(first_row), int(col_shift), C.byref(o)))
        return out

    def coo_read(self, path: str, nprocs: int = 1, rank: int = 0, header_type: int = 1,
                 index: str = "i32", dtype: str = "f64", share=None):
        """(share, DeviceCSR): rank `rank` of `nprocs` of a binary COO file
        built on the device (nsp_coo_read); share as coo_share returns it."""
        torch = _torch()
        s = share if share is not None else coo_share(path, nprocs, rank, header_type)
        nrec = s.rec_end - s.rec_begin
        dev = torch.device("cuda", torch.cuda.current_device())
        out = DeviceCSR(torch.empty(s.nrows + 1, dtype=torch.int64 if index == "i64" else torch.int32,
                                    device=dev),
                        torch.empty(nrec, dtype=torch.int32, device=dev),
                        torch.empty(nrec, dtype=torch.float32 if dtype == "f32" else torch.float64, device=dev),
                        int(s.nrows), int(s.nc))
        o = out.c_struct()
        self._check(self.L.nsp_coo_read(self.h, out.index_type, out.value_type, path.encode(), int(header_type),
                                        C.byref(s), C.byref(o)))
        return s, out

    def ascii_read(self, path: str, nprocs: int = 1, rank: int = 0, index: str = "i32", dtype: str = "f64",
                   share=None):
        """(share, DeviceCSR): rank `rank` of `nprocs` of an ASCII COO file
        parsed and built on the device (nsp_ascii_read); share as
        ascii_share returns it."""
        torch = _torch()
        s = share if share is not None else ascii_share(path, nprocs, rank)
        nrec = s.rec_end - s.rec_begin
        dev = torch.device("cuda", torch.cuda.current_device())
        out = DeviceCSR(torch.empty(s.nrows + 1, dtype=torch.int64 if index == "i64" else torch.int32,
                                    device=dev),
                        torch.empty(nrec, dtype=torch.int32, device=dev),
                        torch.empty(nrec, dtype=torch.float32 if dtype == "f32" else torch.float64, device=dev),
                        int(s.nrows), int(s.nc))
        o = out.c_struct()
        self._check(self.L.nsp_ascii_read(self.h, out.index_type, out.value_type, path.encode(), C.byref(s),
                                          C.byref(o)))
        return s, out


def ascii_share(path: str, nprocs: int = 1, rank: int = 0) -> "_lib.nsp_ascii_share_t":
    """Header, line layout and rank range of an ASCII COO file
    (nsp_ascii_share; host only, no GPU needed)."""
    s = _lib.nsp_ascii_share_t()
    st = _lib.load().nsp_ascii_share(path.encode(), int(nprocs), int(rank), C.byref(s))
    if st != _lib.NSP_OK:
        raise NspError(st, f"cannot read the header, record lines or rows of {path}")
    return s


def coo_share(path: str, nprocs: int = 1, rank: int = 0, header_type: int = 1) -> "_lib.nsp_coo_share_t":
    """Header and rank range of a binary COO file (nsp_coo_share; host
    only, no GPU needed)."""
    s = _lib.nsp_coo_share_t()
    st = _lib.load().nsp_coo_share(path.encode(), int(header_type), int(nprocs), int(rank), C.byref(s))
    if st != _lib.NSP_OK:
        raise NspError(st, f"cannot read the header or rows of {path}")
    return s


class Graph:
    """A product (or batch of products) recorded into one hipGraph; launch()
    re-runs binning, symbolic pass, scan and numeric pass with no host
    synchronisation; result() synchronises and returns C (row_ptr, col and
    val views of nnz entries)."""

    def __init__(self, handle: Handle, A: DeviceCSR, B: DeviceCSR, capacity: int, a_off=None, b_off=None):
        torch = _torch()
        _same_types(A, B)
        self.handle, self.L = handle, handle.L
        self.keep = (A, B, a_off, b_off)
        dev = A.col.device
        self.C = DeviceCSR(torch.empty(A.nrows + 1, dtype=A.row_ptr.dtype, device=dev),
                           torch.empty(max(capacity, 1), dtype=torch.int32, device=dev),
                           torch.empty(max(capacity, 1), dtype=A.val.dtype, device=dev), A.nrows, B.ncols)
        c = _lib.nsp_csr_t(A.nrows, B.ncols, capacity, self.C.row_ptr.data_ptr(), self.C.col.data_ptr(),
                           self.C.val.data_ptr())
        a, b = A.c_struct(), B.c_struct()
        self.g = C.c_void_p()
        count = a_off.numel() - 1 if a_off is not None else 0
        handle._check(self.L.nsp_graph_create(
            handle.h, A.index_type, A.value_type, C.byref(a),
            C.c_void_p(a_off.data_ptr() if a_off is not None else None), C.byref(b),
            C.c_void_p(b_off.data_ptr() if b_off is not None else None), count, C.byref(c),
            C.byref(self.g)))

    def _check(self, s):
        if s != _lib.NSP_OK:
            raise NspError(s, self.L.nsp_graph_last_error(self.g).decode())

    def launch(self):
        self._check(self.L.nsp_graph_launch(self.g))

    def result(self) -> DeviceCSR:
        nnz = C.c_int64()
        self._check(self.L.nsp_graph_result(self.g, C.byref(nnz)))
        n = int(nnz.value)
        return DeviceCSR(self.C.row_ptr, self.C.col[:n], self.C.val[:n], self.C.nrows, self.C.ncols)

    def close(self):
        if getattr(self, "g", None):
            self.L.nsp_graph_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _same_types(A: DeviceCSR, B: DeviceCSR):
    if A.row_ptr.dtype != B.row_ptr.dtype:
        raise ValueError("A and B must use the same row-pointer width")
    if A.val.dtype != B.val.dtype:
        raise ValueError("A and B must use the same value type")


_default = {}


def default_handle() -> Handle:
    torch = _torch()
    key = (torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream)
    h = _default.get(key)
    if h is None:
        h = _default[key] = Handle()
    return h


def spgemm(A: DeviceCSR, B: DeviceCSR, handle: Handle | None = None) -> DeviceCSR:
    """C = A*B on the GPU (sorted rows, structural zeros kept)."""
    return (handle or default_handle()).spgemm(A, B)
