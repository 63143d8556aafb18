"""The Python wrappers of the device entry points refuse, before any launch,
tensors the kernels would read or write out of bounds: host tensors, strided
views, a wrong element type, too few elements, a stream shorter than the byte
count given, an index shorter than szx_amd_index_words.  No GPU is needed:
the checks run before the library is called."""
import numpy as np
import pytest
import torch

import szx_amd as Z


def bare(cls, **attrs):
    """An instance with its attributes set by hand (no device buffers)."""
    o = object.__new__(cls)
    for k, v in attrs.items():
        setattr(o, k, v)
    return o


def test_torch_dtype_mapping_is_strict():
    assert Z._torch_code(torch.zeros(2, dtype=torch.float32)) == Z.SZx_FLOAT
    assert Z._torch_code(torch.zeros(2, dtype=torch.float64)) == Z.SZx_DOUBLE
    for dt in (torch.float16, torch.bfloat16, torch.int32, torch.int64, torch.uint8):
        with pytest.raises(Z.SZxError, match="unsupported tensor dtype"):
            Z._torch_code(torch.zeros(2, dtype=dt))


@pytest.mark.parametrize("x, msg", [
    (np.zeros(100, np.float32), "torch tensor"),
    (torch.zeros(200, dtype=torch.float32)[::2], "contiguous"),
    (torch.zeros(100, dtype=torch.float64), "expected float32"),
    (torch.zeros(99, dtype=torch.float32), "fewer than 100"),
    (torch.zeros(100, dtype=torch.int32), "unsupported tensor dtype"),
    (torch.zeros(100, dtype=torch.float32), "device tensor"),
])
def test_compressor_refuses_bad_input(x, msg):
    comp = bare(Z.DeviceCompressor, n=100, dtype=np.dtype(np.float32), bs=128,
                fast_mode=Z.SZx_WITH_BLOCK_FAST_CMPR)
    with pytest.raises(Z.SZxError, match=msg):
        comp(x, Z.REL, relBoundRatio=1e-3)


def test_decompressor_refuses_bad_stream_and_index():
    dec = bare(Z.DeviceDecompressor, n=1000, dtype=np.dtype(np.float64), bs=128,
               fast_mode=Z.SZx_WITH_BLOCK_FAST_CMPR)
    with pytest.raises(Z.SZxError, match="fewer than 500"):
        dec(torch.zeros(400, dtype=torch.uint8), 500)
    with pytest.raises(Z.SZxError, match="must be uint8"):
        dec(torch.zeros(600, dtype=torch.float32), 500)
    words = Z.lib().szx_amd_index_words(1000, Z.SZx_DOUBLE, 128)
    with pytest.raises(Z.SZxError, match="index holds"):
        dec(torch.zeros(600, dtype=torch.uint8), 500, index=torch.zeros(words - 1, dtype=torch.int64))
    with pytest.raises(Z.SZxError, match="64-bit words"):
        dec(torch.zeros(600, dtype=torch.uint8), 500, index=torch.zeros(2 * words, dtype=torch.int32))
    with pytest.raises(Z.SZxError, match="device tensor"):   # host tensors, everything else right
        dec(torch.zeros(600, dtype=torch.uint8), 500)


def test_other_wrappers_check_their_inputs():
    nb = bare(Z.DeviceNonBlocked, n=50, dtype=np.dtype(np.float32))
    with pytest.raises(Z.SZxError, match="expected float32"):
        nb.compress(torch.zeros(50, dtype=torch.float64))
    with pytest.raises(Z.SZxError, match="fewer than 80"):
        nb.decompress(torch.zeros(70, dtype=torch.uint8), 80)
    with pytest.raises(Z.SZxError, match="contiguous"):
        Z.rate_stats(torch.zeros(64, dtype=torch.float32)[::2])
    with pytest.raises(Z.SZxError, match="contiguous"):
        Z.BatchCodec([(torch.zeros(64, dtype=torch.float32)[::2], Z.ABS, 1e-3)])


@pytest.mark.parametrize("n, reqexpo, noct", [
    (0, 0, 1),        # no values
    (100, 0, 3),      # one or two octaves only
    (100, 0, 0),
    (100, 127, 1),    # the bound above 2^(reqExpo+1) must be a float
    (100, -149, 1),   # the float below the lowest octave must exist
    (100, -148, 2),
])
def test_octave_span_refuses_bad_arguments(n, reqexpo, noct):
    """The octave-span entry point rejects its arguments before any launch
    (a dummy, never dereferenced device address stands in for the field)."""
    L = Z.lib()
    ws = L.szx_amd_rate_octave_workspace()
    rc = L.szx_amd_rate_octave_span_dev(16, n, Z.SZx_FLOAT, 128, reqexpo, noct, 16, 16, ws, None)
    assert rc != 0
    assert "invalid arguments" in Z.last_error()
