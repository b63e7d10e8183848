"""The C-ABI library loads and exports every symbol include/nsp_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nsp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "nsp_spgemm_symbolic" in names and "nsp_spgemm_numeric" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from nsp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libnsp_amd.so is not built (run python -m nsp_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound, "python binding misses a declared symbol"


def test_status_strings_without_gpu():
    from nsp_amd import _lib
    L = _lib.load()
    assert L.nsp_status_string(0) == b"NSP_OK"
    assert L.nsp_status_string(6) == b"NSP_ERR_UNSUPPORTED"
