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
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)   # preprocessor lines
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


def test_batch_workspace_validates_fields_without_gpu():
    """szx_amd_batch_workspace is host arithmetic: it rejects bad fields with
    0 and sizes good ones (every field's single-field workspace and record
    index, plus the tables) -- no device call is made."""
    L = szx_amd.lib()
    F = szx_amd.Field

    def ws(fields, bs=128):
        arr = (F * len(fields))(*fields)
        return L.szx_amd_batch_workspace(arr, len(fields), bs)

    good = [F(16, 1000, szx_amd.SZx_FLOAT, szx_amd.REL, 0.0, 1e-3, 16, 0),
            F(16, 70000, szx_amd.SZx_DOUBLE, szx_amd.ABS, 1e-4, 0.0, 16, 0)]
    total = ws(good)
    assert total > 0
    single = sum(L.szx_amd_compress_workspace(f.n, f.dataType, 128) +
                 8 * L.szx_amd_index_words(f.n, f.dataType, 128) for f in good)
    assert total >= single
    assert ws(good, bs=100) == 0
    assert ws([F(16, 0, szx_amd.SZx_FLOAT, szx_amd.ABS, 1.0, 0.0, 16, 0)]) == 0
    assert ws([F(16, 10, 5, szx_amd.ABS, 1.0, 0.0, 16, 0)]) == 0
    assert ws([F(16, 10, szx_amd.SZx_FLOAT, szx_amd.FXR, 1.0, 0.0, 16, 0)]) == 0
    assert L.szx_amd_batch_workspace(None, 0, 128) == 0


def test_block_size_range_without_gpu():
    """Every blockSize 1..256 sizes its buffers (the powers of two 4..256 use
    the tiled layout and index, the others the per-block ones); anything else
    is refused with 0.  Host-side size queries only."""
    L = szx_amd.lib()
    F, D = szx_amd.SZx_FLOAT, szx_amd.SZx_DOUBLE
    n = 100_003
    for bs in (1, 2, 3, 5, 7, 100, 128, 255, 256):
        for dt in (F, D):
            assert L.szx_amd_compress_capacity(n, dt, bs) > 0
            assert L.szx_amd_compress_workspace(n, dt, bs) > 0
            assert L.szx_amd_decompress_workspace(n, dt, bs, 1 << 20) > 0
    nb = -(-n // 100)
    assert L.szx_amd_index_words(n, F, 100) == 2 * (nb + 1)
    nb128, tiles = -(-n // 128), -(-n // 4096)
    assert L.szx_amd_index_words(n, F, 128) == 2 * (tiles + 1) + (nb128 + 3) // 4
    for bs in (0, 257, -1):
        assert L.szx_amd_compress_capacity(n, F, bs) == 0
        assert L.szx_amd_compress_workspace(n, F, bs) == 0
        assert L.szx_amd_decompress_workspace(n, F, bs, 1 << 20) == 0
        assert L.szx_amd_index_words(n, F, bs) == 0
    # the fixed-rate solver takes the tiled block sizes only
    assert L.szx_amd_fixed_rate_workspace(n, F, 100) == 0
    assert L.szx_amd_fixed_rate_workspace(n, F, 128) > 0


def test_library_exports_nothing_undeclared():
    """Every dynamic symbol the library defines is declared in the header
    (the C++ internals stay local)."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", szx_amd.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    defined = sorted({ln.split()[-1] for ln in out.splitlines() if ln.strip()})
    extra = [s for s in defined if s not in declared_functions()]
    assert not extra, extra


def _dim_grid():
    import itertools
    for r in itertools.product((0, 1, 2, 3), repeat=5):
        yield r   # (r5, r4, r3, r2, r1), zeros anywhere


def test_dimension_helpers_match_the_compiled_reference():
    """SZx_computeDimension / SZx_filterDimension / SZx_computeDataLength
    against the reference's own functions (szx.cc:41-229), every dims tuple
    over {0,1,2,3}^5 (no GPU call)."""
    import pytest
    import oracle as O
    if not O.ref_available():
        pytest.skip("compiled reference not built (oracle/Makefile ref)")
    R = O.ref()
    L = szx_amd.lib()
    sz = C.c_size_t
    for f in (R.SZx_computeDimension, R.SZx_filterDimension, R.SZx_computeDataLength):
        f.argtypes = [sz] * 5 + ([C.POINTER(sz)] if f is R.SZx_filterDimension else [])
    R.SZx_computeDataLength.restype = sz
    for r in _dim_grid():
        assert L.SZx_computeDimension(*r) == R.SZx_computeDimension(*r), r
        assert L.SZx_computeDataLength(*r) == R.SZx_computeDataLength(*r), r
        a, b = (sz * 5)(), (sz * 5)()
        assert L.SZx_filterDimension(*r, a) == R.SZx_filterDimension(*r, b), r
        assert list(a) == list(b), r


def test_c_caller_builds_against_both_headers_and_runs_host_only_modes():
    """The C caller compiled against include/szx_amd.h (and against the
    reference's szx.h, where it was built) links to libszx_amd.so; its
    host-only modes (unions, dims) run without a GPU and agree."""
    import subprocess
    import pytest
    here = os.path.join(os.path.dirname(__file__), "c_abi")
    exe = os.path.join(here, "szx_caller")
    assert os.path.exists(exe), "tests/c_abi/szx_caller is not built (make -C tests/c_abi)"
    assert subprocess.run([exe, "union"], capture_output=True).returncode == 0
    dims = subprocess.run([exe, "dims"], check=True, capture_output=True, text=True).stdout
    ref_exe = os.path.join(here, "szx_caller_refhdr")
    if not os.path.exists(ref_exe):
        pytest.skip("caller against the reference's szx.h not built")
    assert subprocess.run([ref_exe, "dims"], check=True, capture_output=True, text=True).stdout == dims
