"""The C-ABI library loads and exports every function include/szx_amd.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

import szx_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "szx_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_header_declares_the_reference_entry_points():
    names = declared_functions()
    for ref_name in ("SZx_fast_compress_args", "SZx_fast_decompress",
                     "SZx_fast_compress_args_unpredictable_blocked",
                     "SZx_fast_decompress_args_unpredictable_blocked",
                     "SZx_computeDataLength"):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(szx_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_lists_every_export():
    assert sorted(szx_amd.EXPORTS) == declared_functions()


def test_compute_data_length_without_gpu():
    L = szx_amd.lib()
    assert L.SZx_computeDataLength(0, 0, 100, 500, 500) == 25_000_000
    assert L.SZx_computeDataLength(0, 0, 0, 0, 7) == 7
    assert L.SZx_computeDataLength(0, 0, 0, 0, 0) == 0
