"""The oracle's restatement of the reference's CUDA codec stream (cuSZx,
cuszx_entry.cu / cuszx_float.cu / cuszxd_float.cu), pinned without a CUDA
device: hand-built streams whose values follow from the format alone, the
stream size formula, the fastMode-2 oracle (itself pinned by the compiled
reference's golden vectors) on the same blocks, and the error bound.
Also checks the host-side validation of the C ABI's cuSZx reader, which runs
before any device call."""
import struct

import numpy as np
import pytest

import oracle as O


def f32(v):
    return struct.pack("<f", v)


def header(bs, nconst, states, tail=()):
    """[1,1,1,0,bs], BE u64 constant count, MSB-first bitmap, raw tail."""
    bits = bytearray((len(states) + 7) // 8)
    for i, s in enumerate(states):
        if s:
            bits[i // 8] |= 0x80 >> (i % 8)
    return bytes([1, 1, 1, 0, bs]) + struct.pack(">Q", nconst) + bytes(bits) + b"".join(f32(v) for v in tail)


def record(median, req_length, codes, mids):
    return f32(median) + bytes([req_length]) + bytes(codes) + bytes(mids)


# One block of 32 values, two significant bytes (reqLength 16, no shift),
# median 1.0.  Words (value - median as float bits), by hand:
#   v0 = 0x3F800000 ( 1.0)   lead 0 vs the zero word: bytes 0x80, 0x3F
#   v1 = 0x3F800000 ( 1.0)   lead 2: no bytes
#   v2 = 0x3F900000 ( 1.125) lead 1: byte 0x90
#   v3 = 0xBF900000 (-1.125) lead 0: bytes 0x90, 0xBF
#   v4..v31 = v3             lead 2
# Codes: byte 0 = 0<<6 | 2<<4 | 1<<2 | 0 = 0x24, bytes 1..7 = 0xAA.
KAT_MIDS = [0x80, 0x3F, 0x90, 0x90, 0xBF]
KAT_REC = record(1.0, 16, [0x24] + [0xAA] * 7, KAT_MIDS)
KAT_VALUES = [2.0, 2.0, 2.125] + [-0.125] * 29


def kat_stream():
    """Block 0 constant (median 4.5), block 1 the record above, 3 raw values."""
    return (header(32, 1, [0, 1], tail=(7.0, -2.5, 0.125)) + f32(4.5) + struct.pack("<h", len(KAT_MIDS))
            + KAT_REC)


def test_known_answer_stream():
    y = O.decompress_cuszx(kat_stream(), 67)
    want = np.array([4.5] * 32 + KAT_VALUES + [7.0, -2.5, 0.125], np.float32)
    assert np.array_equal(y.view(np.uint32), want.view(np.uint32))


def test_one_significant_byte_reads_its_own_top_byte():
    """cuszxd_float.cu:235-239 marks every value as the source of its top
    byte when one byte is significant, so a lead-1 value decodes a zero byte:
    v0 = 0x40000000 + 1.0 = 3.0, every later value (lead 1) = 0 + 1.0."""
    s = header(32, 0, [1]) + struct.pack("<h", 1) + record(1.0, 8, [0x15] + [0x55] * 7, [0x40])
    y = O.decompress_cuszx(s, 32)
    assert y[0] == 3.0 and np.all(y[1:] == 1.0)


def test_negative_mid_byte_position_is_an_error():
    """Lead 3 with two significant bytes counts -1 byte; twenty of them put
    the next value's bytes before its slot (the reference reads outside its
    shared buffer there)."""
    codes = [0xFF] * 5 + [0x00] * 3   # values 0..19 lead 3, then lead 0
    s = header(32, 0, [1]) + struct.pack("<h", 4) + record(1.0, 16, codes, [1, 2, 3, 4])
    with pytest.raises(ValueError, match="rc=-7"):
        O.decompress_cuszx(s, 32)


def test_record_longer_than_its_block_is_refused():
    """cuszx_entry.cu:165-169 stops when count + 5 + bs/4 > 4 bs."""
    mids = bytes(4 * 32)
    s = header(32, 0, [1]) + struct.pack("<h", len(mids)) + record(1.0, 32, [0] * 8, mids)
    with pytest.raises(ValueError, match="rc=-6"):
        O.decompress_cuszx(s, 32)


def test_block_sizes_outside_the_kernel_are_refused():
    for bs in (0, 16, 100, 255):
        with pytest.raises(ValueError, match="rc=-2"):
            O.decompress_cuszx(header(bs, 0, []), 8)


def test_equal_neighbours_at_four_bytes_spill_into_the_previous_code():
    """cuszx_float.cu:163 gives lead 4 for an equal word at 4 significant
    bytes, and the shift-OR packing (:38-51) puts 4 << 4 into value 4's field.
    Signs alternate so every other lead is 0; x5 = x4."""
    x = np.array([(-1) ** i * (0.5 + 0.01 * i) for i in range(32)], np.float32)
    x[5] = x[4]
    x[6:] = -x[6:]   # keep alternating after the repeat
    s = O.compress_cuszx(x, 1e-6, 32)
    q_rec = 13 + 1 + 2   # header, bitmap, one count
    assert 25 <= s[q_rec + 4] <= 32   # four significant bytes
    codes = s[q_rec + 5:q_rec + 13]
    assert codes[1] == 0x40 and all(c == 0 for i, c in enumerate(codes) if i != 1)


def test_stream_size_formula():
    r = np.random.default_rng(7)
    for bs in (32, 64, 96, 128, 224):
        n = bs * 37 + 11
        x = np.cumsum(r.standard_normal(n)).astype(np.float32)
        x[bs * 3:bs * 6] = 2.0   # constant blocks
        s = O.compress_cuszx(x, 1e-3, bs)
        nb = n // bs
        nconst = struct.unpack(">Q", s[5:13])[0]
        nc = nb - nconst
        q_off = 13 + (nb + 7) // 8 + 4 * (n % bs) + 4 * nconst
        counts = struct.unpack(f"<{nc}h", s[q_off:q_off + 2 * nc])
        assert len(s) == q_off + 2 * nc + sum(c + 5 + bs // 4 for c in counts)
        assert nconst >= 3


@pytest.mark.parametrize("bs", [32, 64, 128])
@pytest.mark.parametrize("eb", [1e-2, 1e-4, 1e-6])
def test_same_values_as_the_blocked_stream(bs, eb):
    """Where reqLength stays within 32 bits and no word repeats at 4 bytes, a
    full block holds the same mid bits in both formats: the cuSZx decode equals
    the fastMode-2 oracle's decode bit for bit, and the raw tail is exact."""
    r = np.random.default_rng(bs + int(-np.log10(eb)))
    n = bs * 40 + 5
    x = (np.cumsum(r.standard_normal(n)) * 0.05).astype(np.float32)
    full = n // bs * bs
    y = O.decompress_cuszx(O.compress_cuszx(x, eb, bs), n)
    y2 = O.decompress(O.compress(x[:full], eb, bs), full, np.float32)
    assert np.array_equal(y[:full].view(np.uint32), y2.view(np.uint32))
    assert np.array_equal(y[full:], x[full:])
    assert np.abs(y.astype(np.float64) - x).max() <= eb


def test_host_reader_validates_before_the_device():
    import szx_amd as Z
    Z.lib()
    with pytest.raises(Z.SZxError, match="multiple of 32"):
        Z.cuszx_decompress(header(100, 0, []), 400)
    mids = bytes(4 * 32)
    s = header(32, 0, [1]) + struct.pack("<h", len(mids)) + record(1.0, 32, [0] * 8, mids)
    with pytest.raises(Z.SZxError, match="larger than its block"):
        Z.cuszx_decompress(s, 32)
    for bs in (0, 16, 100, 256):
        with pytest.raises(Z.SZxError, match="multiple of 32"):
            Z.cuszx_compress(np.zeros(512, np.float32), 1e-3, bs)


def test_device_entries_validate_arguments_without_gpu():
    """Argument checks of the device-resident cuSZx entries run before any
    launch (NULL buffers, short workspace or capacity, block size)."""
    import ctypes as C
    import szx_amd as Z
    L = Z.lib()
    n = 1 << 16
    assert L.szx_amd_cuszx_workspace(0) == 0
    assert L.szx_amd_cuszx_capacity(n, 100) == 0 and L.szx_amd_cuszx_compress_workspace(n, 16) == 0
    cap, ws = L.szx_amd_cuszx_capacity(n, 128), L.szx_amd_cuszx_compress_workspace(n, 128)
    assert cap > 13 + n // 128 // 8 and ws > 0
    fake = C.c_void_p(0x1000)   # never dereferenced: every call below fails its checks first
    assert L.szx_amd_decompress_cuszx_dev(None, 100, n, fake, fake, L.szx_amd_cuszx_workspace(n), fake, None) != 0
    assert L.szx_amd_decompress_cuszx_dev(fake, 100, n, fake, fake, 16, fake, None) != 0
    assert "workspace" in Z.last_error()
    assert L.szx_amd_compress_cuszx_dev(fake, n, 1e-3, 100, fake, cap, fake, ws, fake, None) != 0
    assert "multiple of 32" in Z.last_error()
    assert L.szx_amd_compress_cuszx_dev(fake, n, 1e-3, 128, fake, cap - 1, fake, ws, fake, None) != 0
    assert "too small" in Z.last_error()
