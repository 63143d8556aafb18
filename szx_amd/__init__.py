"""szx_amd -- SZx's blockwise fast codec (fastMode 2) on MI355X.

The product is ``libszx_amd.so`` (HIP kernels for gfx950 behind a C ABI,
declared in ``include/szx_amd.h``).  This package is a thin ctypes binding to
it, used by the tests and the benchmark:

* the reference's drop-in entry points (host arrays in, bytes out), and
* the device-resident entry points, with torch providing device memory and
  streams.

There is no fallback: if the library is missing, importing the binding's
functions raises, and every codec call runs on the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libszx_amd.so")

# szx_defines.h
ABS, REL, FXR = 0, 1, 2
SZx_FLOAT, SZx_DOUBLE = 0, 1
SZx_NO_BLOCK_FAST_CMPR = 1
SZx_WITH_BLOCK_FAST_CMPR = 2
SZx_RANDOMACCESS_FAST_CMPR = 3
SZx_OPENMP_FAST_CMPR = 4
SZx_SCES, SZx_NSCS, SZx_FERR, SZx_TERR, SZx_DERR, SZx_MERR, SZx_BERR = 0, -1, -2, -3, -4, -5, -6
DECODE_REFERENCE = 0
DECODE_EXACT = 1

# every symbol include/szx_amd.h declares
EXPORTS = (
    "SZx_computeDataLength", "SZx_fast_compress_args", "SZx_fast_decompress",
    "SZx_fast_compress_args_unpredictable_blocked",
    "SZx_fast_decompress_args_unpredictable_blocked", "SZx_computeValueRange",
    "SZx_estimateErrorBoundbasedonCR", "SZ_fast_compress_args", "SZ_fast_decompress",
    "szx_amd_last_error", "szx_amd_compress_bound", "szx_amd_compress_capacity",
    "szx_amd_compress_workspace", "szx_amd_compress_dev", "szx_amd_decompress_workspace",
    "szx_amd_decompress_dev", "szx_amd_compress_ra_dev", "szx_amd_decompress_ra_dev",
    "szx_amd_nb_capacity", "szx_amd_nb_workspace", "szx_amd_compress_nb_dev", "szx_amd_decompress_nb_dev",
    "szx_amd_index_words", "szx_amd_timing_begin",
    "szx_amd_timing_end", "szx_amd_copy_dev", "szx_amd_fxr_workspace", "szx_amd_fxr_bound_dev",
    "szx_amd_fixed_rate_workspace", "szx_amd_fixed_rate_dev", "szx_amd_rate_stats_bytes",
    "szx_amd_rate_stats_workspace", "szx_amd_rate_stats_dev", "szx_amd_rate_inputs_dev",
    "szx_amd_mlp_forward_dev", "szx_amd_mlp_grad_workspace", "szx_amd_mlp_grad_dev",
    "szx_amd_mlp_adam_dev", "szx_amd_rate_octave_workspace", "szx_amd_rate_octave_dev",
    "szx_amd_rate_octave_span_dev",
    "szx_amd_batch_workspace", "szx_amd_batch_create", "szx_amd_batch_destroy",
    "szx_amd_batch_compress", "szx_amd_batch_results_dev", "szx_amd_batch_status_dev",
    "szx_amd_batch_index_dev", "szx_amd_batch_results", "szx_amd_batch_decompress",
    "szx_amd_batch_set_shape", "szx_amd_batch_tune",
    "cuSZx_fast_decompress_args_unpredictable_blocked_float", "szx_amd_cuszx_workspace",
    "szx_amd_decompress_cuszx_dev", "cuSZx_fast_compress_args_unpredictable_blocked_float",
    "szx_amd_cuszx_capacity", "szx_amd_cuszx_compress_workspace", "szx_amd_compress_cuszx_dev",
    "SZx_computeDimension", "SZx_filterDimension",
    "SZx_fast_compress_args_unpredictable_blocked_args",
    "SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp",
    "SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp",
    "SZx_fast_decompress_pred", "computeRadiusBuffer", "estimateCRbasedonErrorBound_buffered",
    "szx_amd_compress_dev_noalloc",
)
BATCH_FUSED, BATCH_PER_FIELD = 0, 1
OCTAVE_BINS = 512
RATE_BISECT, RATE_LEARNED = 0, 1
RATE_MODEL_PARAMS = 5825
RATE_MODEL_INPUTS = 24
RATE_CHUNK = 524288
NOALLOC_SCRATCH_BYTES = 16 << 20
RATE_MODEL_PATH = os.path.join(HERE, "rate_model_fp32.bin")


class Field(C.Structure):
    """szx_amd_field: one field of a batched launch (device pointers)."""
    _fields_ = [("d_in", C.c_void_p), ("n", C.c_size_t), ("dataType", C.c_int),
                ("errBoundMode", C.c_int), ("absErrBound", C.c_float),
                ("relBoundRatio", C.c_float), ("d_out", C.c_void_p), ("d_dec", C.c_void_p)]


class RateInfo(C.Structure):
    _fields_ = [("errorBound", C.c_float), ("bitsPerValue", C.c_double),
                ("codecPasses", C.c_int), ("featurePasses", C.c_int), ("hit", C.c_int)]

_lib = None
_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


class SZxError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libszx_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m szx_amd.build`")
    L = C.CDLL(LIB_PATH)
    sz, vp, f32, i32 = C.c_size_t, C.c_void_p, C.c_float, C.c_int
    L.szx_amd_last_error.restype = C.c_char_p
    L.SZx_computeDataLength.restype = sz
    L.SZx_computeDataLength.argtypes = [sz] * 5
    L.SZx_fast_compress_args.restype = vp
    L.SZx_fast_compress_args.argtypes = [i32, i32, vp, C.POINTER(sz), i32, f32, f32, f32, f32,
                                         sz, sz, sz, sz, sz]
    L.SZ_fast_compress_args.restype = vp
    L.SZ_fast_compress_args.argtypes = L.SZx_fast_compress_args.argtypes
    L.SZx_fast_decompress.restype = vp
    L.SZx_fast_decompress.argtypes = [i32, i32, vp, sz, sz, sz, sz, sz, sz]
    L.SZ_fast_decompress.restype = vp
    L.SZ_fast_decompress.argtypes = L.SZx_fast_decompress.argtypes
    L.SZx_fast_compress_args_unpredictable_blocked.restype = vp
    L.SZx_fast_compress_args_unpredictable_blocked.argtypes = [vp, i32, C.POINTER(sz), f32, sz, i32]
    L.SZx_fast_decompress_args_unpredictable_blocked.restype = None
    L.SZx_fast_decompress_args_unpredictable_blocked.argtypes = [vp, i32, sz, vp]
    L.SZx_computeValueRange.restype = C.c_double
    L.SZx_computeValueRange.argtypes = [vp, i32, sz, C.POINTER(f32), C.POINTER(f32)]
    L.SZx_estimateErrorBoundbasedonCR.restype = f32
    L.SZx_estimateErrorBoundbasedonCR.argtypes = [i32, f32, f32, vp, f32, i32, sz]
    L.SZx_computeDimension.restype = i32
    L.SZx_computeDimension.argtypes = [sz] * 5
    L.SZx_filterDimension.restype = i32
    L.SZx_filterDimension.argtypes = [sz] * 5 + [C.POINTER(sz)]
    L.SZx_fast_compress_args_unpredictable_blocked_args.restype = sz
    L.SZx_fast_compress_args_unpredictable_blocked_args.argtypes = [vp, i32, vp, f32, sz, i32]
    L.SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp.restype = vp
    L.SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp.argtypes = [vp, i32, C.POINTER(sz), f32,
                                                                                   sz, i32]
    L.SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp.restype = vp
    L.SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp.argtypes = [i32, sz, vp]
    L.SZx_fast_decompress_pred.restype = vp
    L.SZx_fast_decompress_pred.argtypes = [i32, vp, vp, sz, sz, sz, sz, sz, sz]
    L.computeRadiusBuffer.restype = f32
    L.computeRadiusBuffer.argtypes = [vp, i32, sz, i32, i32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p)]
    L.estimateCRbasedonErrorBound_buffered.restype = f32
    L.estimateCRbasedonErrorBound_buffered.argtypes = [i32, f32, vp, vp, vp, i32, i32, sz, C.POINTER(sz),
                                                       C.POINTER(sz)]
    L.szx_amd_compress_bound.restype = sz
    L.szx_amd_compress_bound.argtypes = [sz, i32, i32]
    L.szx_amd_compress_capacity.restype = sz
    L.szx_amd_compress_capacity.argtypes = [sz, i32, i32]
    L.szx_amd_compress_workspace.restype = sz
    L.szx_amd_compress_workspace.argtypes = [sz, i32, i32]
    L.szx_amd_compress_dev.restype = i32
    L.szx_amd_compress_dev.argtypes = [vp, sz, i32, i32, f32, f32, i32, vp, sz, vp, sz, vp, vp,
                                       vp]
    L.szx_amd_compress_dev_noalloc.restype = i32
    L.szx_amd_compress_dev_noalloc.argtypes = [vp, sz, i32, i32, f32, f32, i32, vp, sz, vp, vp, vp]
    L.szx_amd_index_words.restype = sz
    L.szx_amd_index_words.argtypes = [sz, i32, i32]
    L.szx_amd_decompress_workspace.restype = sz
    L.szx_amd_decompress_workspace.argtypes = [sz, i32, i32, sz]
    L.szx_amd_decompress_dev.restype = i32
    L.szx_amd_decompress_dev.argtypes = [vp, sz, sz, i32, i32, vp, vp, sz, vp, i32, vp, vp]
    L.szx_amd_nb_capacity.restype = sz
    L.szx_amd_nb_capacity.argtypes = [sz, i32]
    L.szx_amd_nb_workspace.restype = sz
    L.szx_amd_nb_workspace.argtypes = [sz, i32]
    L.szx_amd_compress_nb_dev.restype = i32
    L.szx_amd_compress_nb_dev.argtypes = [vp, sz, i32, i32, f32, f32, vp, sz, vp, sz, vp, vp]
    L.szx_amd_decompress_nb_dev.restype = i32
    L.szx_amd_decompress_nb_dev.argtypes = [vp, sz, sz, i32, vp, vp, sz, i32, vp, vp]
    L.cuSZx_fast_decompress_args_unpredictable_blocked_float.restype = None
    L.cuSZx_fast_decompress_args_unpredictable_blocked_float.argtypes = [C.POINTER(C.c_void_p), sz, vp]
    L.cuSZx_fast_compress_args_unpredictable_blocked_float.restype = vp
    L.cuSZx_fast_compress_args_unpredictable_blocked_float.argtypes = [vp, C.POINTER(sz), f32, sz, i32]
    L.szx_amd_cuszx_capacity.restype = sz
    L.szx_amd_cuszx_capacity.argtypes = [sz, i32]
    L.szx_amd_cuszx_compress_workspace.restype = sz
    L.szx_amd_cuszx_compress_workspace.argtypes = [sz, i32]
    L.szx_amd_compress_cuszx_dev.restype = i32
    L.szx_amd_compress_cuszx_dev.argtypes = [vp, sz, f32, i32, vp, sz, vp, sz, vp, vp]
    L.szx_amd_cuszx_workspace.restype = sz
    L.szx_amd_cuszx_workspace.argtypes = [sz]
    L.szx_amd_decompress_cuszx_dev.restype = i32
    L.szx_amd_decompress_cuszx_dev.argtypes = [vp, sz, sz, vp, vp, sz, vp, vp]
    L.szx_amd_compress_ra_dev.restype = i32
    L.szx_amd_compress_ra_dev.argtypes = [i32] + list(L.szx_amd_compress_dev.argtypes)
    L.szx_amd_decompress_ra_dev.restype = i32
    L.szx_amd_decompress_ra_dev.argtypes = [i32] + list(L.szx_amd_decompress_dev.argtypes)
    L.szx_amd_timing_begin.restype = None
    L.szx_amd_timing_begin.argtypes = []
    L.szx_amd_timing_end.restype = i32
    L.szx_amd_timing_end.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.szx_amd_copy_dev.restype = i32
    L.szx_amd_copy_dev.argtypes = [vp, vp, sz, vp]
    L.szx_amd_fxr_workspace.restype = sz
    L.szx_amd_fxr_workspace.argtypes = [sz, i32, i32]
    L.szx_amd_fxr_bound_dev.restype = i32
    L.szx_amd_fxr_bound_dev.argtypes = [vp, sz, i32, i32, f32, f32, f32, vp, sz, C.POINTER(f32),
                                        C.POINTER(i32), vp]
    L.szx_amd_fixed_rate_workspace.restype = sz
    L.szx_amd_fixed_rate_workspace.argtypes = [sz, i32, i32]
    L.szx_amd_fixed_rate_dev.restype = i32
    L.szx_amd_fixed_rate_dev.argtypes = [vp, sz, i32, i32, f32, f32, i32, vp, vp, sz, vp, sz, vp,
                                         vp, C.POINTER(RateInfo), vp]
    L.szx_amd_rate_octave_workspace.restype = sz
    L.szx_amd_rate_octave_workspace.argtypes = []
    L.szx_amd_rate_octave_dev.restype = i32
    L.szx_amd_rate_octave_dev.argtypes = [vp, sz, i32, i32, i32, vp, vp, sz, vp]
    L.szx_amd_rate_octave_span_dev.restype = i32
    L.szx_amd_rate_octave_span_dev.argtypes = [vp, sz, i32, i32, i32, i32, vp, vp, sz, vp]
    L.szx_amd_rate_stats_bytes.restype = sz
    L.szx_amd_rate_stats_bytes.argtypes = [sz]
    L.szx_amd_rate_stats_workspace.restype = sz
    L.szx_amd_rate_stats_workspace.argtypes = [sz]
    L.szx_amd_rate_stats_dev.restype = i32
    L.szx_amd_rate_stats_dev.argtypes = [vp, sz, i32, i32, vp, vp, sz, vp]
    L.szx_amd_rate_inputs_dev.restype = i32
    L.szx_amd_rate_inputs_dev.argtypes = [vp, sz, vp, i32, vp, vp]
    L.szx_amd_mlp_forward_dev.restype = i32
    L.szx_amd_mlp_forward_dev.argtypes = [vp, vp, sz, vp, vp]
    L.szx_amd_mlp_grad_workspace.restype = sz
    L.szx_amd_mlp_grad_workspace.argtypes = []
    L.szx_amd_mlp_grad_dev.restype = i32
    L.szx_amd_mlp_grad_dev.argtypes = [vp, vp, vp, sz, vp, vp, vp, sz, vp]
    L.szx_amd_mlp_adam_dev.restype = i32
    L.szx_amd_mlp_adam_dev.argtypes = [vp, vp, vp, vp, f32, f32, f32, f32, i32, vp]
    fp = C.POINTER(Field)
    L.szx_amd_batch_workspace.restype = sz
    L.szx_amd_batch_workspace.argtypes = [fp, i32, i32]
    L.szx_amd_batch_create.restype = vp
    L.szx_amd_batch_create.argtypes = [fp, i32, i32, vp, sz, vp]
    L.szx_amd_batch_destroy.restype = None
    L.szx_amd_batch_destroy.argtypes = [vp]
    L.szx_amd_batch_compress.restype = i32
    L.szx_amd_batch_compress.argtypes = [vp, i32, vp]
    L.szx_amd_batch_results_dev.restype = vp
    L.szx_amd_batch_results_dev.argtypes = [vp]
    L.szx_amd_batch_status_dev.restype = vp
    L.szx_amd_batch_status_dev.argtypes = [vp]
    L.szx_amd_batch_index_dev.restype = vp
    L.szx_amd_batch_index_dev.argtypes = [vp, i32]
    L.szx_amd_batch_results.restype = i32
    L.szx_amd_batch_results.argtypes = [vp, vp, vp]
    L.szx_amd_batch_decompress.restype = i32
    L.szx_amd_batch_decompress.argtypes = [vp, i32, i32, vp]
    L.szx_amd_batch_set_shape.restype = i32
    L.szx_amd_batch_set_shape.argtypes = [vp, i32, i32]
    L.szx_amd_batch_tune.restype = i32
    L.szx_amd_batch_tune.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(f32), vp]
    _lib = L
    return L


KERNEL_KINDS = ("compress", "decode", "index", "copy")


def timing_begin() -> None:
    """Start bracketing this thread's main kernel launches with HIP events."""
    lib().szx_amd_timing_begin()


def timing_end() -> dict:
    """{kind: (summed ms, launches)} for the launches since timing_begin()."""
    ms = (C.c_double * len(KERNEL_KINDS))()
    cnt = (C.c_int * len(KERNEL_KINDS))()
    if lib().szx_amd_timing_end(ms, cnt) != 0:
        raise SZxError(last_error())
    return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNEL_KINDS)}


def copy_dev(src, dst, stream=None) -> None:
    """Copy device tensor src into dst (same byte size, a multiple of 16) with
    the library's plain copy kernel: the copy rate that bounds the codec."""
    torch = _torch()
    if not isinstance(src, torch.Tensor) or not isinstance(dst, torch.Tensor):
        raise SZxError("copy_dev takes two torch tensors")
    nb = src.numel() * src.element_size()
    if not src.is_contiguous() or not dst.is_contiguous() or dst.numel() * dst.element_size() != nb:
        raise SZxError("copy_dev needs contiguous tensors of the same byte size")
    if not src.is_cuda or not dst.is_cuda:
        raise SZxError("copy_dev takes device tensors")
    _check(lib().szx_amd_copy_dev(src.data_ptr(), dst.data_ptr(), nb, _stream_handle(stream)))


def last_error() -> str:
    return lib().szx_amd_last_error().decode()


def _dtype_code(dtype) -> int:
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return SZx_FLOAT
    if dtype == np.float64:
        return SZx_DOUBLE
    raise TypeError(f"unsupported dtype {dtype}")


def _take(ptr, nbytes: int) -> bytes:
    out = C.string_at(ptr, nbytes)
    _libc.free(ptr)
    return out


# ---------------------------------------------------------------------------
# drop-in host API (szx.h)
# ---------------------------------------------------------------------------

def _dims(shape):
    dims = list(shape)[::-1][:5]           # r1 is the fastest dimension
    dims += [0] * (5 - len(dims))
    return dims[4], dims[3], dims[2], dims[1], dims[0]


def SZx_fast_compress_args(fastMode: int, data: np.ndarray, errBoundMode: int,
                           absErrBound: float = 0.0, relBoundRatio: float = 0.0,
                           compressionRatio: float = 0.0, tolerance: float = 0.0) -> bytes:
    """szx.h:77 with the dims taken from data.shape (C order, last axis = r1)."""
    x = np.ascontiguousarray(data)
    osz = C.c_size_t(0)
    p = lib().SZx_fast_compress_args(fastMode, _dtype_code(x.dtype), x.ctypes.data, C.byref(osz),
                                     errBoundMode, absErrBound, relBoundRatio, compressionRatio,
                                     tolerance, *_dims(x.shape))
    if not p:
        raise SZxError(last_error())
    return _take(p, osz.value)


def SZx_fast_decompress(fastMode: int, dtype, stream: bytes, shape) -> np.ndarray:
    """szx.h:80: decode into an array of `shape`."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape))
    buf = C.create_string_buffer(bytes(stream), len(stream))
    p = lib().SZx_fast_decompress(fastMode, _dtype_code(dtype), buf, len(stream), *_dims(shape))
    if not p:
        raise SZxError(last_error())
    return np.frombuffer(_take(p, n * dtype.itemsize), dtype).reshape(shape).copy()


def compress_blocked(data: np.ndarray, absErrBound: float, blockSize: int = 128) -> bytes:
    """szx.h:89 SZx_fast_compress_args_unpredictable_blocked."""
    x = np.ascontiguousarray(data).ravel()
    osz = C.c_size_t(0)
    p = lib().SZx_fast_compress_args_unpredictable_blocked(x.ctypes.data, _dtype_code(x.dtype),
                                                           C.byref(osz), absErrBound, x.size,
                                                           blockSize)
    if not p:
        raise SZxError(last_error())
    return _take(p, osz.value)


def decompress_blocked(stream: bytes, n: int, dtype) -> np.ndarray:
    """szx.h:98 SZx_fast_decompress_args_unpredictable_blocked (length-free)."""
    dtype = np.dtype(dtype)
    buf = C.create_string_buffer(bytes(stream), len(stream))
    out = C.c_void_p(0)
    lib().SZx_fast_decompress_args_unpredictable_blocked(C.byref(out), _dtype_code(dtype), n, buf)
    if not out.value:
        raise SZxError(last_error())
    return np.frombuffer(_take(out.value, n * dtype.itemsize), dtype).copy()


def cuszx_compress(data: np.ndarray, absErrBound: float, blockSize: int = 128) -> bytes:
    """cuszx_entry.h:10 cuSZx_fast_compress_args_unpredictable_blocked_float:
    the stream the reference's CUDA codec writes (fp32)."""
    x = np.ascontiguousarray(data, np.float32).ravel()
    osz = C.c_size_t(0)
    p = lib().cuSZx_fast_compress_args_unpredictable_blocked_float(x.ctypes.data, C.byref(osz), absErrBound,
                                                                   x.size, blockSize)
    if not p:
        raise SZxError(last_error())
    return _take(p, osz.value)


def cuszx_decompress(stream: bytes, n: int) -> np.ndarray:
    """cuszx_entry.h:11 cuSZx_fast_decompress_args_unpredictable_blocked_float:
    read a stream the reference's CUDA codec wrote (fp32, length-free)."""
    buf = C.create_string_buffer(bytes(stream), len(stream))
    out = C.c_void_p(0)
    lib().cuSZx_fast_decompress_args_unpredictable_blocked_float(C.byref(out), n, buf)
    if not out.value:
        raise SZxError(last_error())
    return np.frombuffer(_take(out.value, n * 4), np.float32).copy()


def SZx_computeValueRange(data: np.ndarray):
    """szx.h:81: (range, radius, median) of the array."""
    x = np.ascontiguousarray(data).ravel()
    rad, med = C.c_float(0), C.c_float(0)
    r = lib().SZx_computeValueRange(x.ctypes.data, _dtype_code(x.dtype), x.size, C.byref(rad),
                                    C.byref(med))
    if last_error():
        raise SZxError(last_error())
    return r, rad.value, med.value


def SZx_estimateErrorBoundbasedonCR(data: np.ndarray, targetCompressionRatio: float,
                                    tolerance: float, initErrorBound: float,
                                    blockSize: int = 128) -> float:
    """szx.h:87: the fixed-rate bound search from initErrorBound."""
    x = np.ascontiguousarray(data).ravel()
    eb = lib().SZx_estimateErrorBoundbasedonCR(_dtype_code(x.dtype), targetCompressionRatio,
                                               tolerance, x.ctypes.data, initErrorBound,
                                               blockSize, x.size)
    if last_error():
        raise SZxError(last_error())
    return eb


# ---------------------------------------------------------------------------
# device-resident API (torch tensors as device memory)
# ---------------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _torch_code(t) -> int:
    """SZx dataType of a torch tensor (float32 / float64 only)."""
    torch = _torch()
    if t.dtype == torch.float32:
        return SZx_FLOAT
    if t.dtype == torch.float64:
        return SZx_DOUBLE
    raise SZxError(f"unsupported tensor dtype {t.dtype} (float32 and float64 are)")


def _dev_tensor(t, name: str, n: int, dtype=None):
    """Check, before any launch, that t is a contiguous device tensor of at
    least n elements (of `dtype`, a numpy float dtype, when given): the
    kernels read and write n elements from its data pointer."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise SZxError(f"{name} must be a torch tensor")
    if not t.is_contiguous():
        raise SZxError(f"{name} must be contiguous")
    if dtype is not None and _torch_code(t) != _dtype_code(dtype):
        raise SZxError(f"{name} is {t.dtype}, expected {np.dtype(dtype)}")
    if t.numel() < n:
        raise SZxError(f"{name} holds {t.numel()} elements, fewer than {n}")
    if not t.is_cuda:
        raise SZxError(f"{name} must be a device tensor")


def _dev_bytes(t, name: str, nbytes: int):
    """Check that t is a contiguous uint8 device tensor of at least nbytes."""
    torch = _torch()
    if isinstance(t, torch.Tensor) and t.dtype != torch.uint8:
        raise SZxError(f"{name} must be uint8, not {t.dtype}")
    _dev_tensor(t, name, nbytes)


def _stream_handle(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class DeviceCompressor:
    """Holds the output buffer and workspace for repeated compresses of one
    shape/dtype/blockSize, so the timed path allocates nothing.  fastMode 2
    writes the blocked stream, 3 / 4 the random-access container."""

    def __init__(self, n: int, dtype, blockSize: int = 128, device="cuda",
                 fastMode: int = SZx_WITH_BLOCK_FAST_CMPR):
        torch = _torch()
        self.n, self.dtype, self.bs, self.fast_mode = n, np.dtype(dtype), blockSize, fastMode
        code = _dtype_code(self.dtype)
        L = lib()
        self.capacity = L.szx_amd_compress_capacity(n, code, blockSize)
        self.wsbytes = L.szx_amd_compress_workspace(n, code, blockSize)
        if not self.capacity or not self.wsbytes:
            raise SZxError(f"unsupported parameters (blockSize={blockSize})")
        self.out = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.result = torch.zeros(4, dtype=torch.int64, device=device)
        self.index = torch.zeros(L.szx_amd_index_words(n, code, blockSize), dtype=torch.int64,
                                 device=device)

    def __call__(self, x, errBoundMode: int = ABS, absErrBound: float = 0.0,
                 relBoundRatio: float = 0.0, stream=None) -> None:
        """Enqueue one compress of tensor x (asynchronous)."""
        _dev_tensor(x, "x", self.n, self.dtype)
        args = (x.data_ptr(), self.n, _dtype_code(self.dtype), errBoundMode, absErrBound,
                relBoundRatio, self.bs, self.out.data_ptr(), self.capacity, self.ws.data_ptr(),
                self.wsbytes, self.result.data_ptr(), self.index.data_ptr(), _stream_handle(stream))
        if self.fast_mode == SZx_WITH_BLOCK_FAST_CMPR:
            rc = lib().szx_amd_compress_dev(*args)
        else:
            rc = lib().szx_amd_compress_ra_dev(self.fast_mode, *args)
        if rc != 0:
            raise SZxError(last_error())

    def stream_view(self):
        """(device uint8 tensor view of the stream, offset, size, eb) -- syncs."""
        off, size, ebbits, err = (int(v) for v in self.result.cpu().tolist())
        if err:
            raise SZxError("compress failed: a bounded wait between workgroups ran out")
        eb = float(np.array([ebbits], np.uint32).view(np.float32)[0])
        return self.out[off:off + size], off, size, eb


class DeviceDecompressor:
    def __init__(self, n: int, dtype, blockSize: int, stream_capacity: int, device="cuda",
                 fastMode: int = SZx_WITH_BLOCK_FAST_CMPR):
        torch = _torch()
        self.n, self.dtype, self.bs, self.fast_mode = n, np.dtype(dtype), blockSize, fastMode
        code = _dtype_code(self.dtype)
        self.wsbytes = lib().szx_amd_decompress_workspace(n, code, blockSize, stream_capacity)
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        tdt = torch.float32 if self.dtype == np.float32 else torch.float64
        self.out = torch.empty(n, dtype=tdt, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)

    def __call__(self, stream_tensor, nbytes: int | None = None, flags: int = DECODE_REFERENCE,
                 index=None, stream=None) -> None:
        """Enqueue one decode; `index` is the compressor's record index
        (DeviceCompressor.index) or None to rebuild it from the stream."""
        nb = stream_tensor.numel() if nbytes is None else nbytes
        if index is not None:
            if index.element_size() != 8:
                raise SZxError("index must hold 64-bit words")
            _dev_tensor(index, "index", lib().szx_amd_index_words(self.n, _dtype_code(self.dtype), self.bs))
        _dev_bytes(stream_tensor, "stream_tensor", nb)
        args = (stream_tensor.data_ptr(), nb, self.n, _dtype_code(self.dtype), self.bs,
                self.out.data_ptr(), self.ws.data_ptr(), self.wsbytes,
                None if index is None else index.data_ptr(), flags, self.status.data_ptr(),
                _stream_handle(stream))
        if self.fast_mode == SZx_WITH_BLOCK_FAST_CMPR:
            rc = lib().szx_amd_decompress_dev(*args)
        else:
            rc = lib().szx_amd_decompress_ra_dev(self.fast_mode, *args)
        if rc != 0:
            raise SZxError(last_error())

    def check(self) -> int:
        return int(self.status.item())


class DeviceNonBlocked:
    """fastMode 1 on device tensors: one record over the whole array
    (szx_amd_compress_nb_dev / szx_amd_decompress_nb_dev)."""

    def __init__(self, n: int, dtype, device="cuda"):
        torch = _torch()
        self.n, self.dtype = n, np.dtype(dtype)
        code = _dtype_code(self.dtype)
        L = lib()
        self.capacity = L.szx_amd_nb_capacity(n, code)
        self.wsbytes = L.szx_amd_nb_workspace(n, code)
        if not self.capacity or not self.wsbytes:
            raise SZxError("unsupported parameters")
        self.out = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.result = torch.zeros(4, dtype=torch.int64, device=device)
        tdt = torch.float32 if self.dtype == np.float32 else torch.float64
        self.dec = torch.empty(n, dtype=tdt, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)

    def compress(self, x, errBoundMode: int = ABS, absErrBound: float = 0.0, relBoundRatio: float = 0.0,
                 stream=None) -> None:
        _dev_tensor(x, "x", self.n, self.dtype)
        rc = lib().szx_amd_compress_nb_dev(x.data_ptr(), self.n, _dtype_code(self.dtype), errBoundMode,
                                           absErrBound, relBoundRatio, self.out.data_ptr(), self.capacity,
                                           self.ws.data_ptr(), self.wsbytes, self.result.data_ptr(),
                                           _stream_handle(stream))
        if rc != 0:
            raise SZxError(last_error())

    def stream_view(self):
        """(device uint8 view of the stream, size, eb) -- syncs."""
        off, size, ebbits, _ = (int(v) for v in self.result.cpu().tolist())
        eb = float(np.array([ebbits], np.uint32).view(np.float32)[0])
        return self.out[off:off + size], size, eb

    def decompress(self, stream_tensor, nbytes: int, flags: int = DECODE_REFERENCE, stream=None) -> None:
        _dev_bytes(stream_tensor, "stream_tensor", nbytes)
        rc = lib().szx_amd_decompress_nb_dev(stream_tensor.data_ptr(), nbytes, self.n, _dtype_code(self.dtype),
                                             self.dec.data_ptr(), self.ws.data_ptr(), self.wsbytes, flags,
                                             self.status.data_ptr(), _stream_handle(stream))
        if rc != 0:
            raise SZxError(last_error())

    def check(self) -> int:
        return int(self.status.item())


class DeviceCuszxWriter:
    """Writer of cuSZx streams on device tensors (szx_amd_compress_cuszx_dev);
    the stream starts at out[0]."""

    def __init__(self, n: int, blockSize: int = 128, device="cuda"):
        torch = _torch()
        L = lib()
        self.n, self.bs = n, blockSize
        self.capacity = L.szx_amd_cuszx_capacity(n, blockSize)
        self.wsbytes = L.szx_amd_cuszx_compress_workspace(n, blockSize)
        if not self.capacity or not self.wsbytes:
            raise SZxError("unsupported parameters")
        self.out = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.result = torch.zeros(4, dtype=torch.int64, device=device)

    def compress(self, x, absErrBound: float, stream=None) -> None:
        _dev_tensor(x, "x", self.n, np.float32)
        rc = lib().szx_amd_compress_cuszx_dev(x.data_ptr(), self.n, absErrBound, self.bs, self.out.data_ptr(),
                                              self.capacity, self.ws.data_ptr(), self.wsbytes,
                                              self.result.data_ptr(), _stream_handle(stream))
        if rc != 0:
            raise SZxError(last_error())

    def stream_view(self):
        """(device uint8 view of the stream, size, refused) -- syncs."""
        _, size, _, bad = (int(v) for v in self.result.cpu().tolist())
        return self.out[:size], size, bool(bad)


class DeviceCuszxReader:
    """Reader of cuSZx (the reference's CUDA codec) streams on device tensors
    (szx_amd_decompress_cuszx_dev)."""

    def __init__(self, n: int, device="cuda"):
        torch = _torch()
        self.n = n
        self.wsbytes = lib().szx_amd_cuszx_workspace(n)
        if not self.wsbytes:
            raise SZxError("unsupported parameters")
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.dec = torch.empty(n, dtype=torch.float32, device=device)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)

    def decompress(self, stream_tensor, nbytes: int, out=None, stream=None) -> None:
        dst = self.dec if out is None else out
        _dev_bytes(stream_tensor, "stream_tensor", nbytes)
        _dev_tensor(dst, "out", self.n, np.float32)
        rc = lib().szx_amd_decompress_cuszx_dev(stream_tensor.data_ptr(), nbytes, self.n, dst.data_ptr(),
                                                self.ws.data_ptr(), self.wsbytes, self.status.data_ptr(),
                                                _stream_handle(stream))
        if rc != 0:
            raise SZxError(last_error())

    def check(self) -> int:
        return int(self.status.item())


class BatchCodec:
    """Many device fields compressed / decompressed by one launch sequence
    each way (szx_amd_batch_*).  fields: sequence of (tensor, errBoundMode,
    bound), where bound is the ABS bound or the REL ratio.  Output buffers,
    decode buffers and the workspace are allocated here once; compress() and
    decompress() only enqueue kernels (capturable in a CUDA/HIP graph)."""

    def __init__(self, fields, blockSize: int = 128, decode: bool = True, device="cuda"):
        torch = _torch()
        L = lib()
        self.bs = blockSize
        self.inputs = [f[0] for f in fields]
        nf = len(fields)
        self.nf = nf
        arr = (Field * nf)()
        self.outs, self.decs = [], []
        for i, (x, mode, bound) in enumerate(fields):
            _dev_tensor(x, f"field {i}", x.numel())
            code = _torch_code(x)
            cap = L.szx_amd_compress_capacity(x.numel(), code, blockSize)
            if not cap:
                raise SZxError(f"unsupported parameters (blockSize={blockSize})")
            out = torch.empty(cap, dtype=torch.uint8, device=device)
            dec = torch.empty_like(x) if decode else None
            self.outs.append(out)
            self.decs.append(dec)
            arr[i] = Field(x.data_ptr(), x.numel(), code, mode,
                           bound if mode == ABS else 0.0, bound if mode == REL else 0.0,
                           out.data_ptr(), dec.data_ptr() if dec is not None else None)
        self._fields = arr
        self.wsbytes = L.szx_amd_batch_workspace(arr, nf, blockSize)
        if not self.wsbytes:
            raise SZxError(last_error())
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.handle = L.szx_amd_batch_create(arr, nf, blockSize, self.ws.data_ptr(), self.wsbytes,
                                             _stream_handle(None))
        if not self.handle:
            raise SZxError(last_error())

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().szx_amd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def compress(self, mode: int = BATCH_FUSED, stream=None) -> None:
        _check(lib().szx_amd_batch_compress(self.handle, mode, _stream_handle(stream)))

    def results(self, stream=None) -> np.ndarray:
        """(nf, 4) result words -- stream offset, size, bound bits, error; syncs."""
        out = np.zeros((self.nf, 4), np.uint64)
        _check(lib().szx_amd_batch_results(self.handle, out.ctypes.data, _stream_handle(stream)))
        return out

    def decompress(self, flags: int = DECODE_REFERENCE, mode: int = BATCH_FUSED,
                   stream=None) -> None:
        _check(lib().szx_amd_batch_decompress(self.handle, flags, mode, _stream_handle(stream)))

    def status(self):
        """Device int32 tensor view of the nf decode statuses."""
        torch = _torch()
        st = torch.empty(self.nf, dtype=torch.int32, device=self.ws.device)
        # copy out of the workspace (the plan's status array lives there)
        base = lib().szx_amd_batch_status_dev(self.handle) - self.ws.data_ptr()
        st.copy_(self.ws[base:base + 4 * self.nf].view(torch.int32))
        return st

    def stream(self, i: int, res=None):
        """Device uint8 view of field i's stream (res: results() rows)."""
        r = self.results() if res is None else res
        off, size = int(r[i, 0]), int(r[i, 1])
        return self.outs[i][off:off + size]

    def set_shape(self, valuesPerLane: int, tilesPerGroup: int) -> None:
        _check(lib().szx_amd_batch_set_shape(self.handle, valuesPerLane, tilesPerGroup))

    def tune(self, reps: int = 3, stream=None):
        """Time every tile shape on this batch and keep the fastest:
        returns (valuesPerLane, tilesPerGroup, ms per compress)."""
        lv, ipt, ms = C.c_int(0), C.c_int(0), C.c_float(0.0)
        _check(lib().szx_amd_batch_tune(self.handle, reps, C.byref(lv), C.byref(ipt), C.byref(ms),
                                        _stream_handle(stream)))
        return lv.value, ipt.value, ms.value


class DeviceFixedRateBound:
    """The reference's fixed-rate bound choice for device data (synchronous)."""

    def __init__(self, n: int, dtype, blockSize: int = 128, device="cuda"):
        torch = _torch()
        self.n, self.dtype, self.bs = n, np.dtype(dtype), blockSize
        self.wsbytes = lib().szx_amd_fxr_workspace(n, _dtype_code(self.dtype), blockSize)
        if not self.wsbytes:
            raise SZxError(f"unsupported parameters (blockSize={blockSize})")
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)

    def __call__(self, x, targetCompressionRatio: float, tolerance: float,
                 initErrorBound: float = 0.0, stream=None):
        """(bound, model evaluations); initErrorBound <= 0 starts at 1% of the
        sampled range, as the reference's fixed-rate compressor does."""
        eb, ne = C.c_float(0), C.c_int(0)
        _dev_tensor(x, "x", self.n, self.dtype)
        rc = lib().szx_amd_fxr_bound_dev(x.data_ptr(), self.n, _dtype_code(self.dtype), self.bs,
                                         targetCompressionRatio, tolerance, initErrorBound,
                                         self.ws.data_ptr(), self.wsbytes, C.byref(eb),
                                         C.byref(ne), _stream_handle(stream))
        if rc != 0:
            raise SZxError(last_error())
        return eb.value, ne.value


def _check(rc: int) -> None:
    if rc != 0:
        raise SZxError(last_error())


class RateModel:
    """The fixed-rate rate model on the device: forward, gradient and Adam
    kernels of libszx_amd.so over an FP32 parameter tensor."""

    def __init__(self, params=None, device="cuda"):
        torch = _torch()
        if params is None:
            params = np.fromfile(RATE_MODEL_PATH, dtype="<f4")
        p = np.asarray(params, np.float32).ravel()
        if p.size != RATE_MODEL_PARAMS:
            raise ValueError(f"expected {RATE_MODEL_PARAMS} parameters, got {p.size}")
        self.params = torch.from_numpy(p.copy()).to(device)
        self.device = device
        self._ws = None

    def forward(self, X, stream=None):
        torch = _torch()
        y = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        _check(lib().szx_amd_mlp_forward_dev(self.params.data_ptr(), X.data_ptr(), X.shape[0],
                                             y.data_ptr(), _stream_handle(stream)))
        return y

    def grad(self, X, T, stream=None):
        """(gradient tensor, loss tensor) of mean((model(X) - T)^2)."""
        torch = _torch()
        if self._ws is None:
            self._ws = torch.empty(lib().szx_amd_mlp_grad_workspace(), dtype=torch.uint8,
                                   device=X.device)
        g = torch.empty(RATE_MODEL_PARAMS, dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        _check(lib().szx_amd_mlp_grad_dev(self.params.data_ptr(), X.data_ptr(), T.data_ptr(),
                                          X.shape[0], g.data_ptr(), loss.data_ptr(),
                                          self._ws.data_ptr(), self._ws.numel(),
                                          _stream_handle(stream)))
        return g, loss

    def adam(self, grad, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        _check(lib().szx_amd_mlp_adam_dev(self.params.data_ptr(), grad.data_ptr(), m.data_ptr(),
                                          v.data_ptr(), lr, beta1, beta2, eps, step,
                                          _stream_handle(stream)))


def rate_stats(x, blockSize: int = 128, stream=None):
    """Per-chunk statistics of device array x (raw bytes tensor)."""
    torch = _torch()
    n = x.numel()
    code = _torch_code(x)
    _dev_tensor(x, "x", n)
    st = torch.empty(lib().szx_amd_rate_stats_bytes(n), dtype=torch.uint8, device=x.device)
    ws = torch.empty(lib().szx_amd_rate_stats_workspace(n), dtype=torch.uint8, device=x.device)
    _check(lib().szx_amd_rate_stats_dev(x.data_ptr(), n, code, blockSize, st.data_ptr(),
                                        ws.data_ptr(), ws.numel(), _stream_handle(stream)))
    return st


def rate_inputs(stats, n: int, ebs, stream=None):
    """Model inputs [K * nchunks, 24] for the bounds in ebs (device float32)."""
    torch = _torch()
    nch = (n + RATE_CHUNK - 1) // RATE_CHUNK
    X = torch.empty(ebs.numel() * nch, RATE_MODEL_INPUTS, dtype=torch.float32, device=ebs.device)
    _check(lib().szx_amd_rate_inputs_dev(stats.data_ptr(), n, ebs.data_ptr(), ebs.numel(),
                                         X.data_ptr(), _stream_handle(stream)))
    return X


def rate_octave(x, reqExpo: int, blockSize: int = 128, stream=None, with_end: bool = False):
    """Exact stream sizes (bytes) for the OCTAVE_BINS bounds of one octave:
    (numpy sizes, numpy float32 bounds).  with_end adds the size the octave
    tends to as the bound approaches 2^(reqExpo+1) from below (every block in
    the octave's radius bins constant), the last size before the rate jumps
    at the octave edge."""
    torch = _torch()
    n = x.numel()
    code = _torch_code(x)
    _dev_tensor(x, "x", n)
    acc = torch.zeros(1 + 2 * OCTAVE_BINS, dtype=torch.int64, device=x.device)
    ws = torch.empty(lib().szx_amd_rate_octave_workspace(), dtype=torch.uint8, device=x.device)
    _check(lib().szx_amd_rate_octave_dev(x.data_ptr(), n, code, blockSize, reqExpo,
                                         acc.data_ptr(), ws.data_ptr(), ws.numel(),
                                         _stream_handle(stream)))
    a = acc.cpu().numpy().astype(np.uint64)
    nb = (n + blockSize - 1) // blockSize
    cnt, rec = a[1::2], a[2::2]
    const_before = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rec_from = np.cumsum(rec[::-1])[::-1]
    sizes = 13 + (nb + 7) // 8 + a[0] + 4 * const_before + rec_from
    ebs = np.ldexp(np.float32(1) + np.arange(OCTAVE_BINS, dtype=np.float32) / np.float32(OCTAVE_BINS),
                   reqExpo).astype(np.float32)
    if with_end:
        end = int(13 + (nb + 7) // 8 + a[0] + 4 * int(cnt.sum()))
        return sizes.astype(np.int64), ebs, end
    return sizes.astype(np.int64), ebs


def rate_octave_span(x, reqExpo: int, nOctaves: int = 2, blockSize: int = 128, stream=None):
    """Exact stream sizes (bytes) for the OCTAVE_BINS bounds of each of the
    nOctaves octaves reqExpo, reqExpo - 1, ... sized in one pass, lowest
    bound first, and for the bounds just outside them: (numpy sizes, numpy
    float32 bounds, (size at 2^(reqExpo+1), its bound), (size at the float
    just below the lowest octave, its bound))."""
    torch = _torch()
    n = x.numel()
    code = _torch_code(x)
    _dev_tensor(x, "x", n)
    width = 1 + 2 * OCTAVE_BINS
    acc = torch.zeros(nOctaves * width + 4, dtype=torch.int64, device=x.device)
    ws = torch.empty(lib().szx_amd_rate_octave_workspace(), dtype=torch.uint8, device=x.device)
    _check(lib().szx_amd_rate_octave_span_dev(x.data_ptr(), n, code, blockSize, reqExpo, nOctaves,
                                              acc.data_ptr(), ws.data_ptr(), ws.numel(),
                                              _stream_handle(stream)))
    a = acc.cpu().numpy().astype(np.uint64)
    nb = (n + blockSize - 1) // blockSize
    head = 13 + (nb + 7) // 8
    sizes, ebs = [], []
    for j in reversed(range(nOctaves)):
        o = a[j * width:(j + 1) * width]
        cnt, rec = o[1::2], o[2::2]
        const_before = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        rec_from = np.cumsum(rec[::-1])[::-1]
        sizes.append(head + o[0] + 4 * const_before + rec_from)
        ebs.append(np.ldexp(np.float32(1) + np.arange(OCTAVE_BINS, dtype=np.float32) / np.float32(OCTAVE_BINS),
                            reqExpo - j).astype(np.float32))
    g = a[nOctaves * width:]
    up = (int(head + 4 * g[0] + g[1]), np.float32(np.ldexp(np.float32(1), reqExpo + 1)))
    dn = (int(head + 4 * g[2] + g[3]),
          np.nextafter(np.float32(np.ldexp(np.float32(1), reqExpo - nOctaves + 1)), np.float32(0)))
    return np.concatenate(sizes).astype(np.int64), np.concatenate(ebs), up, dn


class FixedRateCompressor:
    """Fixed-rate compress of one device field to a target bits/value."""

    def __init__(self, n: int, dtype, blockSize: int = 128, model: RateModel | None = None,
                 device="cuda"):
        torch = _torch()
        self.n, self.dtype, self.bs = n, np.dtype(dtype), blockSize
        code = _dtype_code(self.dtype)
        L = lib()
        self.capacity = L.szx_amd_compress_capacity(n, code, blockSize)
        self.wsbytes = L.szx_amd_fixed_rate_workspace(n, code, blockSize)
        if not self.capacity or not self.wsbytes:
            raise SZxError(f"unsupported parameters (blockSize={blockSize})")
        self.out = torch.empty(self.capacity, dtype=torch.uint8, device=device)
        self.ws = torch.empty(self.wsbytes, dtype=torch.uint8, device=device)
        self.result = torch.zeros(4, dtype=torch.int64, device=device)
        self.model = model

    def __call__(self, x, targetBitsPerValue: float, tolerance: float = 0.02,
                 method: int = RATE_LEARNED, stream=None) -> RateInfo:
        info = RateInfo()
        _dev_tensor(x, "x", self.n, self.dtype)
        params = self.model.params.data_ptr() if (method == RATE_LEARNED and self.model) else None
        _check(lib().szx_amd_fixed_rate_dev(x.data_ptr(), self.n, _dtype_code(self.dtype), self.bs,
                                            targetBitsPerValue, tolerance, method, params,
                                            self.out.data_ptr(), self.capacity, self.ws.data_ptr(),
                                            self.wsbytes, self.result.data_ptr(), None,
                                            C.byref(info), _stream_handle(stream)))
        return info

    def stream_view(self):
        off, size, _, err = (int(v) for v in self.result.cpu().tolist())
        if err:
            raise SZxError("compress failed")
        return self.out[off:off + size]
