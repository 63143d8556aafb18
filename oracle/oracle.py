"""ctypes front-end to the CPU checker.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package ``szx_amd``.

Two libraries are wrapped:

* ``liboracle_szx.so`` -- this repo's C restatement of the fastMode-2 codec
  (``szx_oracle.c``) and of the fixed-rate search (``szx_oracle_fxr.c``).
* ``_ref/libszx_ref.so`` -- the reference's own CPU codec compiled from its
  sources by ``make ref``; optional (absent where the sources are absent).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_szx.so")
REF_SO = os.path.join(HERE, "_ref", "libszx_ref.so")

_u8p = C.POINTER(C.c_uint8)
_sz = C.c_size_t


def build(ref: bool = True) -> None:
    """Compile the restatement (and the reference when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_szx.so"], check=True)
    if ref and os.path.isdir("/root/reference/szx/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        L.szxo_compress_bound.restype = _sz
        L.szxo_compress_bound.argtypes = [_sz, C.c_int, C.c_int]
        for nm in ("szxo_compress_f32", "szxo_compress_f64"):
            f = getattr(L, nm)
            f.restype = _sz
            f.argtypes = [vp, _sz, C.c_float, C.c_int, vp]
        L.szxo_decompress_f32.restype = C.c_int
        L.szxo_decompress_f32.argtypes = [vp, _sz, _sz, vp]
        L.szxo_decompress_f64.restype = C.c_int
        L.szxo_decompress_f64.argtypes = [vp, _sz, _sz, vp, C.c_int]
        for nm in ("szxo_block_stats_f32", "szxo_block_stats_f64"):
            f = getattr(L, nm)
            f.restype = _sz
            f.argtypes = [vp, _sz, C.c_float, C.c_int, vp, vp, vp]
        L.szxo_req_length.restype = C.c_int
        L.szxo_req_length.argtypes = [C.c_int, C.c_float, C.c_float, C.POINTER(C.c_int)]
        L.szxo_record_size.restype = _sz
        L.szxo_record_size.argtypes = [vp, _sz, C.c_int]
        L.szxo_record_offsets.restype = C.c_long
        L.szxo_record_offsets.argtypes = [vp, _sz, _sz, C.c_int, vp]
        for nm, rt in (("szxo_value_range_f32", C.c_float), ("szxo_value_range_f64", C.c_double)):
            f = getattr(L, nm)
            f.restype = rt
            f.argtypes = [vp, _sz]
        for nm in ("szxo_rel_eb_f32", "szxo_rel_eb_f64"):
            f = getattr(L, nm)
            f.restype = C.c_float
            f.argtypes = [vp, _sz, C.c_float]
        L.szxo_compress_ra_bound.restype = _sz
        L.szxo_compress_ra_bound.argtypes = [_sz, C.c_int, C.c_int]
        L.szxo_compress_ra.restype = _sz
        L.szxo_compress_ra.argtypes = [vp, _sz, C.c_float, C.c_int, C.c_int, C.c_int, vp]
        L.szxo_decompress_ra.restype = C.c_int
        L.szxo_decompress_ra.argtypes = [vp, _sz, _sz, C.c_int, C.c_int, vp, C.c_int]
        L.szxo_compress_nb_bound.restype = _sz
        L.szxo_compress_nb_bound.argtypes = [_sz, C.c_int]
        L.szxo_compress_nb.restype = _sz
        L.szxo_compress_nb.argtypes = [vp, _sz, C.c_float, C.c_int, vp]
        L.szxo_decompress_nb.restype = C.c_int
        L.szxo_decompress_nb.argtypes = [vp, _sz, _sz, C.c_int, vp, C.c_int]
        L.szxo_cuszx_bound.restype = _sz
        L.szxo_cuszx_bound.argtypes = [_sz, C.c_int]
        L.szxo_cuszx_compress.restype = C.c_long
        L.szxo_cuszx_compress.argtypes = [vp, _sz, C.c_float, C.c_int, vp]
        L.szxo_cuszx_decompress.restype = C.c_int
        L.szxo_cuszx_decompress.argtypes = [vp, _sz, _sz, vp]
        L.szxo_fxr_error_bound.restype = C.c_float
        L.szxo_fxr_error_bound.argtypes = [vp, C.c_int, _sz, C.c_float, C.c_float, C.c_int,
                                           C.POINTER(C.c_int)]
        L.szxo_fxr_estimate_cr.restype = C.c_float
        L.szxo_fxr_estimate_cr.argtypes = [C.c_float, vp, C.c_int, vp, vp, C.c_int, C.c_int, _sz,
                                           C.POINTER(_sz), C.POINTER(_sz)]
        L.szxo_fxr_sample.restype = C.c_float
        L.szxo_fxr_sample.argtypes = [vp, C.c_int, _sz, C.c_int, C.c_int, vp, vp, vp]
        L.szxo_fxr_search.restype = C.c_float
        L.szxo_fxr_search.argtypes = [C.c_float, C.c_float, C.c_int, C.c_float, C.c_int, _sz, vp,
                                      C.c_int, vp, vp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _width(x: np.ndarray) -> int:
    if x.dtype == np.float32:
        return 4
    if x.dtype == np.float64:
        return 8
    raise TypeError(f"unsupported dtype {x.dtype}")


def compress_bound(n: int, bs: int, w: int) -> int:
    return lib().szxo_compress_bound(n, bs, w)


def compress(x: np.ndarray, eb: float, bs: int = 128) -> bytes:
    """fastMode-2 stream of x at absolute bound eb (eb is rounded to float)."""
    x = np.ascontiguousarray(x).ravel()
    w = _width(x)
    out = np.empty(compress_bound(x.size, bs, w), np.uint8)
    f = lib().szxo_compress_f32 if w == 4 else lib().szxo_compress_f64
    n = f(_ptr(x), x.size, float(np.float32(eb)), bs, _ptr(out))
    return out[:n].tobytes()


def decompress(stream: bytes, n: int, dtype, ref_quirk: bool = True) -> np.ndarray:
    """Decode a fastMode-2 stream.  ref_quirk (fp64) mirrors the reference
    decoder bit for bit, including its reqBytes=5 byte-7 read."""
    dtype = np.dtype(dtype)
    s = np.frombuffer(stream, np.uint8)
    out = np.empty(n, dtype)
    if dtype == np.float32:
        rc = lib().szxo_decompress_f32(_ptr(s), s.size, n, _ptr(out))
    else:
        rc = lib().szxo_decompress_f64(_ptr(s), s.size, n, _ptr(out), int(ref_quirk))
    if rc != 0:
        raise ValueError(f"malformed stream (oracle rc={rc})")
    return out


def compress_ra(x: np.ndarray, eb: float, bs: int = 128, fast_mode: int = 3) -> bytes:
    """Random-access container (fastMode 3: big-endian medians; fastMode 4:
    native medians and, for fp64, float-radius block statistics)."""
    x = np.ascontiguousarray(x).ravel()
    w = _width(x)
    out = np.empty(lib().szxo_compress_ra_bound(x.size, bs, w), np.uint8)
    n = lib().szxo_compress_ra(_ptr(x), x.size, float(np.float32(eb)), bs, w, fast_mode, _ptr(out))
    return out[:n].tobytes()


def decompress_ra(stream: bytes, n: int, dtype, fast_mode: int = 3,
                  ref_quirk: bool = True) -> np.ndarray:
    dtype = np.dtype(dtype)
    s = np.frombuffer(stream, np.uint8)
    out = np.empty(n, dtype)
    rc = lib().szxo_decompress_ra(_ptr(s), s.size, n, dtype.itemsize, fast_mode, _ptr(out),
                                  int(ref_quirk))
    if rc != 0:
        raise ValueError(f"malformed stream (oracle rc={rc})")
    return out


def compress_nb(x: np.ndarray, eb: float) -> bytes:
    """Non-blocked stream (fastMode 1): one record against the global median."""
    x = np.ascontiguousarray(x).ravel()
    w = _width(x)
    out = np.empty(lib().szxo_compress_nb_bound(x.size, w), np.uint8)
    n = lib().szxo_compress_nb(_ptr(x), x.size, float(np.float32(eb)), w, _ptr(out))
    return out[:n].tobytes()


def decompress_nb(stream: bytes, n: int, dtype, ref_quirk: bool = True) -> np.ndarray:
    dtype = np.dtype(dtype)
    s = np.frombuffer(stream, np.uint8)
    out = np.empty(n, dtype)
    rc = lib().szxo_decompress_nb(_ptr(s), s.size, n, dtype.itemsize, _ptr(out), int(ref_quirk))
    if rc != 0:
        raise ValueError(f"malformed stream (oracle rc={rc})")
    return out


def compress_cuszx(x: np.ndarray, eb: float, bs: int = 128) -> bytes:
    """The stream the reference's CUDA codec (cuSZx) writes for fp32 x."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.empty(max(lib().szxo_cuszx_bound(x.size, bs), 1), np.uint8)
    n = lib().szxo_cuszx_compress(_ptr(x), x.size, float(np.float32(eb)), bs, _ptr(out))
    if n < 0:
        raise ValueError(f"cuSZx stream undefined for this input (oracle rc={n})")
    return out[:n].tobytes()


def decompress_cuszx(stream: bytes, n: int) -> np.ndarray:
    """What the reference's cuSZx decoder reads from a stream (fp32)."""
    s = np.frombuffer(stream, np.uint8)
    out = np.empty(n, np.float32)
    rc = lib().szxo_cuszx_decompress(_ptr(s), s.size, n, _ptr(out))
    if rc != 0:
        raise ValueError(f"malformed cuSZx stream (oracle rc={rc})")
    return out


def block_stats(x: np.ndarray, eb: float, bs: int = 128):
    x = np.ascontiguousarray(x).ravel()
    nb = -(-x.size // bs)
    st = np.empty(nb, np.uint8)
    med = np.empty(nb, np.float32)
    rad = np.empty(nb, np.float32)
    f = lib().szxo_block_stats_f32 if x.dtype == np.float32 else lib().szxo_block_stats_f64
    nc = f(_ptr(x), x.size, float(np.float32(eb)), bs, _ptr(st), _ptr(med), _ptr(rad))
    return nc, st, med, rad


def req_length(w: int, eb: float, radius: float):
    cl = C.c_int(0)
    ln = lib().szxo_req_length(w, float(np.float32(eb)), float(np.float32(radius)), C.byref(cl))
    return ln, bool(cl.value)


def record_offsets(stream: bytes, n: int, w: int) -> np.ndarray:
    s = np.frombuffer(stream, np.uint8)
    offs = np.empty(max(n, 1), np.uint64)
    k = lib().szxo_record_offsets(_ptr(s), s.size, n, w, _ptr(offs))
    if k < 0:
        raise ValueError(f"malformed stream (oracle rc={k})")
    return offs[:k].copy()


def rel_eb(x: np.ndarray, rel: float) -> float:
    x = np.ascontiguousarray(x).ravel()
    f = lib().szxo_rel_eb_f32 if x.dtype == np.float32 else lib().szxo_rel_eb_f64
    return f(_ptr(x), x.size, float(np.float32(rel)))


def value_range(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x).ravel()
    f = lib().szxo_value_range_f32 if x.dtype == np.float32 else lib().szxo_value_range_f64
    return f(_ptr(x), x.size)


def fxr_error_bound(x: np.ndarray, target_cr: float, tol: float, bs: int = 128):
    """Reference fixed-rate bound choice; returns (eb, model evaluations)."""
    x = np.ascontiguousarray(x).ravel()
    ne = C.c_int(0)
    eb = lib().szxo_fxr_error_bound(_ptr(x), _width(x), x.size, float(target_cr), float(tol), bs,
                                     C.byref(ne))
    return eb, ne.value


def fxr_search_from(x: np.ndarray, target_cr: float, tol: float, init: float, bs: int = 128,
                    rate: int = 10):
    """The fixed-rate search from an explicit start bound; (eb, evaluations)."""
    x = np.ascontiguousarray(x).ravel()
    w = _width(x)
    ns = x.size // (bs * rate)
    buf = np.zeros((ns + 1) * bs, x.dtype)
    med = np.zeros(ns + 1, np.float32)
    rad = np.zeros(ns + 1, np.float32)
    L = lib()
    L.szxo_fxr_sample(_ptr(x), w, x.size, rate, bs, _ptr(buf), _ptr(med), _ptr(rad))
    ne = C.c_int(0)
    eb = L.szxo_fxr_search(float(target_cr), float(tol), rate, float(init), bs, x.size, _ptr(buf), w,
                           _ptr(med), _ptr(rad), C.byref(ne))
    return eb, ne.value


# ---------------------------------------------------------------------------
# the compiled reference (optional)
# ---------------------------------------------------------------------------

_ref = None
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        vp = C.c_void_p
        R.SZx_fast_compress_args.restype = vp
        R.SZx_fast_compress_args.argtypes = [C.c_int, C.c_int, vp, C.POINTER(_sz), C.c_int, C.c_float,
                                             C.c_float, C.c_float, C.c_float, _sz, _sz, _sz, _sz, _sz]
        R.SZx_fast_decompress.restype = vp
        R.SZx_fast_decompress.argtypes = [C.c_int, C.c_int, vp, _sz, _sz, _sz, _sz, _sz, _sz]
        R.SZx_fast_compress_args_unpredictable_blocked.restype = vp
        R.SZx_fast_compress_args_unpredictable_blocked.argtypes = [vp, C.c_int, C.POINTER(_sz),
                                                                   C.c_float, _sz, C.c_int]
        f = getattr(R, "_ZN3szx42estimateErrorBoundbasedonCR_buffered_floatEffifimPfS0_S0_")
        f.restype = C.c_float
        f.argtypes = [C.c_float, C.c_float, C.c_int, C.c_float, C.c_int, _sz, vp, vp, vp]
        f = getattr(R, "_ZN3szx43estimateErrorBoundbasedonCR_buffered_doubleEffifimPdPfS1_")
        f.restype = C.c_float
        f.argtypes = [C.c_float, C.c_float, C.c_int, C.c_float, C.c_int, _sz, vp, vp, vp]
        _ref = R
    return _ref


def _take(ptr: int, nbytes: int) -> bytes:
    out = C.string_at(ptr, nbytes)
    _libc.free(ptr)
    return out


def ref_compress(x: np.ndarray, eb: float, bs: int = 128) -> bytes:
    """The reference's blocked compressor (any blockSize)."""
    x = np.ascontiguousarray(x).ravel()
    dt = 0 if x.dtype == np.float32 else 1
    osz = _sz(0)
    p = ref().SZx_fast_compress_args_unpredictable_blocked(_ptr(x), dt, C.byref(osz),
                                                           float(np.float32(eb)), x.size, bs)
    return _take(p, osz.value)


def ref_compress_args(x: np.ndarray, mode: int, abs_eb: float = 0.0, rel: float = 0.0,
                      fast_mode: int = 2, cr: float = 0.0, tol: float = 0.0) -> bytes:
    """SZx_fast_compress_args with 1-D dims (fixed blockSize 128).  For FXR the
    reference writes into a buffer of N*W/cr*(1+2 tol) bytes: callers must
    know the stream fits (it overflows the heap otherwise)."""
    x = np.ascontiguousarray(x).ravel()
    dt = 0 if x.dtype == np.float32 else 1
    osz = _sz(0)
    p = ref().SZx_fast_compress_args(fast_mode, dt, _ptr(x), C.byref(osz), mode, float(abs_eb),
                                     float(rel), float(cr), float(tol), 0, 0, 0, 0, x.size)
    return _take(p, osz.value)


def ref_decompress(stream: bytes, n: int, dtype, fast_mode: int = 2) -> np.ndarray:
    dtype = np.dtype(dtype)
    dt = 0 if dtype == np.float32 else 1
    buf = C.create_string_buffer(stream, len(stream))
    p = ref().SZx_fast_decompress(fast_mode, dt, buf, len(stream), 0, 0, 0, 0, n)
    raw = _take(p, n * dtype.itemsize)
    return np.frombuffer(raw, dtype).copy()
