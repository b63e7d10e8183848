"""ctypes binding of libnsp_amd.so (the C ABI in include/nsp_amd.h).

Loading never falls back to anything: if the library is missing the import
of the compute functions raises, naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsp_amd.so")

NSP_OK, NSP_ERR_INVALID, NSP_ERR_OOM, NSP_ERR_OVERFLOW, NSP_ERR_HIP, NSP_ERR_STATE, \
    NSP_ERR_UNSUPPORTED = range(7)
NSP_IDX32, NSP_IDX64 = 0, 1
NSP_F64, NSP_F32 = 0, 1


class nsp_csr_t(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


class nsp_stats_t(C.Structure):
    _fields_ = [("nprod", C.c_int64), ("nnz", C.c_int64), ("max_intprod", C.c_int64),
                ("max_row_nnz", C.c_int64), ("sym_bin", C.c_int64 * 16),
                ("num_bin", C.c_int64 * 16), ("n_sym_bins", C.c_int32),
                ("n_num_bins", C.c_int32), ("direct_mode", C.c_int32), ("host_syncs", C.c_int32),
                ("sym_launches", C.c_int32), ("num_launches", C.c_int32),
                ("t_bin_ms", C.c_float), ("t_sym_ms", C.c_float), ("t_scan_ms", C.c_float),
                ("t_num_ms", C.c_float), ("t_coo_ms", C.c_float),
                ("fused", C.c_int32), ("t_compact_ms", C.c_float)]


class nsp_coo_share_t(C.Structure):
    _fields_ = [("nr", C.c_int64), ("nc", C.c_int64), ("nt", C.c_int64), ("header_bytes", C.c_int64),
                ("rec_begin", C.c_int64), ("rec_end", C.c_int64), ("first_row", C.c_int64),
                ("nrows", C.c_int64), ("col_shift", C.c_int64)]


class nsp_ascii_share_t(C.Structure):
    _fields_ = [("nr", C.c_int64), ("nc", C.c_int64), ("nt", C.c_int64), ("body_offset", C.c_int64),
                ("line_bytes", C.c_int64), ("col_offset", C.c_int32), ("val_offset", C.c_int32),
                ("rec_begin", C.c_int64), ("rec_end", C.c_int64), ("first_row", C.c_int64),
                ("nrows", C.c_int64), ("col_shift", C.c_int64)]


class NspError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{status_string(status)}: {msg}")
        self.status = status


_lib = None

# (name, restype, argtypes) of every symbol include/nsp_amd.h declares
_H, _I, _VP, _I64, _P64 = C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)
_CSRP = C.POINTER(nsp_csr_t)
SIGNATURES = [
    ("nsp_create", _I, [C.POINTER(_H)]),
    ("nsp_destroy", _I, [_H]),
    ("nsp_set_stream", _I, [_H, _VP]),
    ("nsp_status_string", C.c_char_p, [_I]),
    ("nsp_last_error", C.c_char_p, [_H]),
    ("nsp_get_stats", _I, [_H, C.POINTER(nsp_stats_t)]),
    ("nsp_set_param", _I, [_H, C.c_char_p, _I64]),
    ("nsp_spgemm_symbolic", _I, [_H, _I, _I, _CSRP, _CSRP, _VP, _P64]),
    ("nsp_spgemm_numeric", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP]),
    ("nsp_spgemm", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP, _VP, _VP, _VP]),
    ("nsp_spgemm_noalloc", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP, _P64]),
    ("nsp_spgemm_into", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP, _P64]),
    ("nsp_csr_free", _I, [_CSRP]),
    ("nsp_intprod", _I, [_H, _I, _CSRP, _CSRP, _VP, _P64]),
    ("nsp_check_csr", _I, [_H, _I, _CSRP, _P64]),
    ("nsp_shrink_cols", _I, [_H, _I64, _VP, _I64, _I64, _I64, _VP, _VP, _I64, _P64]),
    ("nsp_gather_rows_count", _I, [_H, _I, _CSRP, _I64, _VP, _I64, _VP, _P64]),
    ("nsp_gather_rows_fill", _I, [_H, _I, _I, _CSRP, _I64, _VP, _I64, _I64, _CSRP]),
    ("nsp_spgemm_batched_symbolic", _I, [_H, _I, _I, _CSRP, _VP, _CSRP, _VP, _I64, _VP, _P64]),
    ("nsp_graph_create", _I, [_H, _I, _I, _CSRP, _VP, _CSRP, _VP, _I64, _CSRP, C.POINTER(_H)]),
    ("nsp_graph_launch", _I, [_H]),
    ("nsp_graph_result", _I, [_H, _P64]),
    ("nsp_graph_last_error", C.c_char_p, [_H]),
    ("nsp_graph_destroy", _I, [_H]),
    ("nsp_transpose", _I, [_H, _I, _I, _CSRP, _CSRP]),
    ("nsp_spgemm_backward", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP, _VP, _CSRP, _VP, _VP, _VP]),
    ("nsp_galerkin", _I, [_H, _I, _I, _CSRP, _CSRP, _CSRP, _CSRP, _VP, _VP, _VP]),
    ("nsp_coo_share", _I, [C.c_char_p, _I, _I, _I, C.POINTER(nsp_coo_share_t)]),
    ("nsp_coo_to_csr", _I, [_H, _I, _I, _VP, _I64, _I64, _I64, _CSRP]),
    ("nsp_coo_read", _I, [_H, _I, _I, C.c_char_p, _I, C.POINTER(nsp_coo_share_t), _CSRP]),
    ("nsp_ascii_share", _I, [C.c_char_p, _I, _I, C.POINTER(nsp_ascii_share_t)]),
    ("nsp_ascii_read", _I, [_H, _I, _I, C.c_char_p, C.POINTER(nsp_ascii_share_t), _CSRP]),
]


# allocation callbacks of the callee-allocating entry points
ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_size_t, C.c_void_p)
FREE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nsp_amd.build` "
                              "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(s: int) -> str:
    names = ["NSP_OK", "NSP_ERR_INVALID", "NSP_ERR_OOM", "NSP_ERR_OVERFLOW", "NSP_ERR_HIP",
             "NSP_ERR_STATE", "NSP_ERR_UNSUPPORTED"]
    return names[s] if 0 <= s < len(names) else f"status {s}"
