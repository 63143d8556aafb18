"""GPU parity of libszx_amd.so against the oracle and the golden vectors.

Every call goes through the C ABI (the reference's drop-in entry points or
the device-resident ones).  The bar is bit-exact: the stream must equal the
reference's byte for byte, and the decode must equal the reference decoder's
bit for bit (the fp64 exact mode is checked against the oracle's exact mode).
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
import szx_amd as Z

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "f*.npz")))   # fastMode 2
GOLDEN_RA = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ra*.npz")))  # fastMode 3/4
POW2 = (4, 8, 16, 32, 64, 128, 256)
# block sizes outside the tiled kernels (per-block kernels)
GENERIC = (1, 2, 3, 5, 7, 12, 33, 100, 200, 255)


def say(*a):
    print(*a, flush=True)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits_equal(a, b):
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(
        a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_stream_and_decode(path, device):
    g = load(path)
    x, bs, eb = g["x"], int(g["bs"]), float(g["eb"])
    name = os.path.basename(path)
    say("case", name, x.dtype, x.size, bs, eb)
    if "sin_rel" in name:   # produced by SZx_fast_compress_args(REL)
        s = Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, x, Z.REL,
                                     relBoundRatio=float(g["rel"]))
    else:
        s = Z.compress_blocked(x, eb, bs)
    assert s == g["stream"].tobytes()
    if bs == 128:
        d = Z.SZx_fast_decompress(Z.SZx_WITH_BLOCK_FAST_CMPR, x.dtype, s, x.shape)
    else:
        d = Z.decompress_blocked(s, x.size, x.dtype)
    want = g["decoded"] if g["decoded"].size else O.decompress(s, x.size, x.dtype)
    assert bits_equal(d, want)


def _field(r, n, dtype, kind):
    t = np.arange(n)
    if kind == 0:
        x = np.sin(t * r.uniform(1e-4, 0.2)) * 10 ** r.uniform(-4, 4)
    elif kind == 1:
        x = np.cumsum(r.standard_normal(n)) * 10 ** r.uniform(-6, 2)
    elif kind == 2:
        x = np.where(r.random(n) < 0.5, 0.75, np.sin(t * 0.01))
    elif kind == 3:
        x = r.standard_normal(n) * 10.0 ** r.uniform(-300 if dtype == np.float64 else -40, 0)
    else:
        x = np.where(r.random(n) < 0.3, -0.0, r.standard_normal(n) * 1e-3)
    x = x + r.standard_normal(n) * 10 ** r.uniform(-9, -3)
    return x.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_parity_host_api(dtype, device):
    r = np.random.default_rng(100 if dtype == np.float32 else 101)
    for trial in range(60):
        n = int(r.choice([1, 2, 3, 5, 63, 64, 127, 128, 129, 1000, 4095, 4096, 4097, 50000,
                          100003]))
        bs = int(r.choice(POW2))
        x = _field(r, n, dtype, trial % 5)
        eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 10.0 ** r.uniform(-12, -1)))
        say("trial", trial, dtype.__name__, n, bs, eb)
        s = Z.compress_blocked(x, eb, bs)
        assert s == O.compress(x, eb, bs), (trial, n, bs, eb)
        d = Z.decompress_blocked(s, n, dtype)
        assert bits_equal(d, O.decompress(s, n, dtype, ref_quirk=True)), trial


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("rel", [1e-2, 1e-3, 1e-4, 1e-6])
def test_rel_mode(dtype, rel, device):
    r = np.random.default_rng(7)
    x = _field(r, 70001, dtype, 1)
    say("rel", dtype.__name__, rel)
    s = Z.SZx_fast_compress_args(Z.SZx_WITH_BLOCK_FAST_CMPR, x, Z.REL, relBoundRatio=rel)
    assert s == O.compress(x, O.rel_eb(x, rel), 128)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_device_api_and_exact_decode(dtype, device):
    import torch
    r = np.random.default_rng(5)
    n = 300_007
    x = _field(r, n, dtype, 1)
    eb = float(np.float32(np.ptp(x) * 1e-8))
    say("device api", dtype.__name__, n, eb)
    comp = Z.DeviceCompressor(n, dtype, 128)
    xt = torch.from_numpy(x).to(device)
    comp(xt, Z.ABS, eb)
    view, off, size, eb_used = comp.stream_view()
    s = view.cpu().numpy().tobytes()
    assert np.float32(eb_used) == np.float32(eb)
    assert s == O.compress(x, eb, 128)
    dec = Z.DeviceDecompressor(n, dtype, 128, size)
    ref_dec = O.decompress(s, n, dtype, ref_quirk=True)
    dec(view, size, Z.DECODE_REFERENCE)                      # index rebuilt from the stream
    assert dec.check() == 0
    assert bits_equal(dec.out.cpu().numpy(), ref_dec)
    dec.out.zero_()
    dec(view, size, Z.DECODE_REFERENCE, index=comp.index)    # the compressor's index
    assert dec.check() == 0
    assert bits_equal(dec.out.cpu().numpy(), ref_dec)
    dec(view, size, Z.DECODE_EXACT)
    assert dec.check() == 0
    exact = dec.out.cpu().numpy()
    assert bits_equal(exact, O.decompress(s, n, dtype, ref_quirk=False))
    if dtype == np.float32:
        assert np.max(np.abs(exact.astype(np.float64) - x)) <= eb


def test_repeated_calls_reuse_buffers(device):
    import torch
    r = np.random.default_rng(9)
    n = 1 << 20
    comp = Z.DeviceCompressor(n, np.float32, 128)
    for rep in range(3):
        x = _field(r, n, np.float32, rep % 3)
        eb = float(np.float32(np.ptp(x) * 1e-4))
        say("repeat", rep)
        comp(torch.from_numpy(x).to(device), Z.ABS, eb)
        view, off, size, _ = comp.stream_view()
        assert view.cpu().numpy().tobytes() == O.compress(x, eb, 128)


def test_malformed_streams_are_rejected(device):
    x = np.sin(np.arange(10000) * 0.01).astype(np.float32)
    s = Z.compress_blocked(x, 1e-4, 128)
    say("truncated")
    with pytest.raises(Z.SZxError):
        Z.SZx_fast_decompress(Z.SZx_WITH_BLOCK_FAST_CMPR, np.float32, s[: len(s) // 2], x.shape)
    bad = bytearray(s)
    bad[5:13] = (10 ** 9).to_bytes(8, "big")        # impossible constant-block count
    with pytest.raises(Z.SZxError):
        Z.SZx_fast_decompress(Z.SZx_WITH_BLOCK_FAST_CMPR, np.float32, bytes(bad), x.shape)
    # a decode after a rejected stream still works
    d = Z.SZx_fast_decompress(Z.SZx_WITH_BLOCK_FAST_CMPR, np.float32, s, x.shape)
    assert bits_equal(d, O.decompress(s, x.size, np.float32))


@pytest.mark.parametrize("bs", POW2)
def test_index_path_all_block_sizes(bs, device):
    import torch
    r = np.random.default_rng(bs)
    for dtype in (np.float32, np.float64):
        n = int(r.integers(20000, 90000))
        x = _field(r, n, dtype, int(r.integers(0, 5)))
        eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 10.0 ** r.uniform(-9, -2)))
        say("index path", bs, dtype.__name__, n, eb)
        comp = Z.DeviceCompressor(n, dtype, bs)
        comp(torch.from_numpy(x).to(device), Z.ABS, eb)
        view, off, size, _ = comp.stream_view()
        s = view.cpu().numpy().tobytes()
        assert s == O.compress(x, eb, bs)
        dec = Z.DeviceDecompressor(n, dtype, bs, size)
        dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
        assert dec.check() == 0
        assert bits_equal(dec.out.cpu().numpy(), O.decompress(s, n, dtype, ref_quirk=True))


def test_corrupt_index_is_rejected(device):
    import torch
    x = np.sin(np.arange(200000) * 0.001).astype(np.float32)
    comp = Z.DeviceCompressor(x.size, np.float32, 128)
    comp(torch.from_numpy(x).to(device), Z.ABS, 1e-5)
    view, off, size, _ = comp.stream_view()
    bad = comp.index.clone()
    bad[4] += 3                      # shift one tile's start off its first record
    dec = Z.DeviceDecompressor(x.size, np.float32, 128, size)
    say("corrupt index")
    dec(view, size, Z.DECODE_REFERENCE, index=bad)
    assert dec.check() != 0


@pytest.mark.parametrize("bs", POW2)
def test_index_free_decode(bs, device):
    """Decode without the compressor's index: the record index is rebuilt in
    parallel from the stream (chunk exit maps, chain marking, scans)."""
    import torch
    r = np.random.default_rng(1000 + bs)
    for dtype in (np.float32, np.float64):
        for trial in range(3):
            n = int(r.choice([1, 5, 4097, 30011, 200003, 1 << 20]))
            x = _field(r, n, dtype, int(r.integers(0, 5)))
            if trial == 1:
                x[: n // 2] = 0.5            # long constant run: few records
            eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 10.0 ** r.uniform(-9, -2)))
            s = O.compress(x, eb, bs)
            say("index-free", bs, dtype.__name__, n, trial, len(s))
            want = O.decompress(s, n, dtype, ref_quirk=True)
            st = torch.from_numpy(np.frombuffer(s, np.uint8).copy()).to(device)
            dec = Z.DeviceDecompressor(n, dtype, bs, len(s) + 64)
            dec(st, len(s), Z.DECODE_REFERENCE)
            assert dec.check() == 0
            assert bits_equal(dec.out.cpu().numpy(), want)
            # a buffer longer than the stream (trailing bytes) decodes the same
            padded = torch.cat([st, torch.full((64,), 0xAB, dtype=torch.uint8, device=device)])
            dec(padded, len(s) + 64, Z.DECODE_REFERENCE)
            assert dec.check() == 0
            assert bits_equal(dec.out.cpu().numpy(), want)


@pytest.mark.parametrize("bs", [8, 128, 256])
def test_index_free_decode_in_order_repair(bs, device, monkeypatch):
    """The index builder's in-order repair (normally only for chunks whose
    speculative chain never met the true one) on every chunk: speculation
    switched off, each chunk's entry comes from the repair walking the chunk
    before it.  Same bytes as the oracle's decoder."""
    import torch
    monkeypatch.setenv("SZX_AMD_INDEX_NO_SPECULATION", "1")
    r = np.random.default_rng(2000 + bs)
    for dtype in (np.float32, np.float64):
        n = 300007
        x = _field(r, n, dtype, int(r.integers(0, 5)))
        x[n // 3: n // 2] = 0.25
        eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 1e-5))
        s = O.compress(x, eb, bs)
        say("in-order repair", bs, dtype.__name__, n, len(s))
        want = O.decompress(s, n, dtype, ref_quirk=True)
        st = torch.from_numpy(np.frombuffer(s, np.uint8).copy()).to(device)
        dec = Z.DeviceDecompressor(n, dtype, bs, len(s))
        dec(st, len(s), Z.DECODE_REFERENCE)
        assert dec.check() == 0
        assert bits_equal(dec.out.cpu().numpy(), want)


def test_index_free_rejects_damage(device):
    import torch
    x = np.sin(np.arange(300000) * 0.001).astype(np.float32)
    s = bytearray(O.compress(x, 1e-6, 128))
    dec = Z.DeviceDecompressor(x.size, np.float32, 128, len(s))
    say("truncated payload")
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    dec(st[: len(s) - 100], len(s) - 100, Z.DECODE_REFERENCE)
    assert dec.check() != 0
    say("impossible reqLength")
    q0 = 13 + (((x.size + 127) // 128) + 7) // 8 + 4 * int.from_bytes(s[5:13], "big")
    s[q0] = 0                                # first record's reqLength byte
    st = torch.from_numpy(np.frombuffer(bytes(s), np.uint8).copy()).to(device)
    dec(st, len(s), Z.DECODE_REFERENCE)
    assert dec.check() != 0


@pytest.mark.parametrize("bs", GENERIC)
def test_generic_block_sizes(bs, device):
    """Block sizes the tiled kernels do not take: stream bytes, compressor
    index, index-free decode and the fp64 exact mode against the oracle."""
    import torch
    r = np.random.default_rng(7000 + bs)
    for dtype in (np.float32, np.float64):
        for n in (1, bs + 1, int(r.integers(3000, 70000))):
            x = _field(r, n, dtype, int(r.integers(0, 5)))
            if n > 2 * bs:
                x[: n // 3] = 0.25            # constant blocks
            eb = float(np.float32((float(np.ptp(x)) + 1e-30) * 10.0 ** r.uniform(-9, -2)))
            say("generic", bs, dtype.__name__, n, eb)
            want_s = O.compress(x, eb, bs)
            comp = Z.DeviceCompressor(n, dtype, bs)
            comp(torch.from_numpy(x).to(device), Z.ABS, eb)
            view, off, size, _ = comp.stream_view()
            s = view.cpu().numpy().tobytes()
            assert s == want_s
            want = O.decompress(s, n, dtype, ref_quirk=True)
            dec = Z.DeviceDecompressor(n, dtype, bs, size + 64)
            dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
            assert dec.check() == 0
            assert bits_equal(dec.out.cpu().numpy(), want)
            st = torch.from_numpy(np.frombuffer(s, np.uint8).copy()).to(device)
            dec(st, size, Z.DECODE_REFERENCE)                 # index rebuilt from the stream
            assert dec.check() == 0
            assert bits_equal(dec.out.cpu().numpy(), want)
            if dtype == np.float64:
                dec(view, size, Z.DECODE_EXACT, index=comp.index)
                assert dec.check() == 0
                assert bits_equal(dec.out.cpu().numpy(), O.decompress(s, n, dtype, ref_quirk=False))
            # host API
            assert Z.compress_blocked(x, eb, bs) == want_s
            assert bits_equal(Z.decompress_blocked(want_s, n, dtype), want)


def _tile_at_stage_cap(extra):
    """An fp32 field whose first 16384-value tile (block size 128: one
    compressor tile, 48 KiB stage) has records of 49152 + extra bytes in all.
    117 blocks alternate +-1.5 (eb 2^-14: reqLength 24, 3 bytes a value, lead
    code 0: 421-byte records), 11 blocks are constant, and block 0 trades
    bytes for lead codes: a repeated value (code 3) saves 3 bytes, a value of
    the same sign and exponent as the one before (code 1) saves 1."""
    def block(codes):
        v, prev = [], None
        for c in codes:
            if prev is None:
                x = 1.5
            elif c == 3:
                x = prev
            elif c == 1:
                x = np.copysign(1.25 if abs(prev) == 1.5 else 1.5, prev)
            else:
                x = -np.copysign(1.5, prev)
            v.append(x)
            prev = x
        return np.array(v, np.float32)

    saved = 117 * 421 - (49152 + extra)
    n3, n1 = saved // 3, saved % 3
    codes = [0] * 128
    for j in range(n3):
        codes[1 + j] = 3
    for j in range(n1):
        codes[n3 + 4 + 2 * j] = 1
    alt = block([0] * 128)
    tile = np.concatenate([block(codes)] + [alt] * 116 + [np.zeros(128 * 11, np.float32)])
    rest = np.sin(np.arange(2 * 16384 + 777) * 0.01).astype(np.float32)
    return np.concatenate([tile, rest])


@pytest.mark.parametrize("extra", [0, 1])
def test_stage_cap_edge(extra, device):
    """A tile whose record bytes are exactly the stage capacity is staged;
    one byte more and it is written straight to the payload.  Both streams
    match the oracle's, and the tile sizes are checked on the oracle's stream
    first."""
    import torch
    x = _tile_at_stage_cap(extra)
    eb = 2.0 ** -14
    want = O.compress(x, eb, 128)
    offs = O.record_offsets(want, x.size, 4)
    assert int(offs[117]) - int(offs[0]) == 49152 + extra
    say("stage cap edge", extra)
    comp = Z.DeviceCompressor(x.size, np.float32, 128)
    comp(torch.from_numpy(x).to(device), Z.ABS, eb)
    view, off, size, _ = comp.stream_view()
    assert view.cpu().numpy().tobytes() == want
    dec = Z.DeviceDecompressor(x.size, np.float32, 128, size)
    dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
    assert dec.check() == 0
    assert bits_equal(dec.out.cpu().numpy(), O.decompress(want, x.size, np.float32))


def test_generic_rel_and_damage(device):
    import torch
    x = (np.sin(np.arange(100003) * 0.003) * 3).astype(np.float32)
    bs = 100
    comp = Z.DeviceCompressor(x.size, np.float32, bs)
    comp(torch.from_numpy(x).to(device), Z.REL, 0.0, 1e-4)
    view, off, size, eb = comp.stream_view()
    s = view.cpu().numpy().tobytes()
    say("generic REL", eb)
    assert eb == O.rel_eb(x, 1e-4)
    assert s == O.compress(x, eb, bs)
    dec = Z.DeviceDecompressor(x.size, np.float32, bs, size)
    say("generic truncated")
    dec(view[: size - 50], size - 50, Z.DECODE_REFERENCE)
    assert dec.check() != 0
    say("generic damaged index")
    bad = comp.index.clone()
    bad[2 * 50] += 1
    dec(view, size, Z.DECODE_REFERENCE, index=bad)
    assert dec.check() != 0
    dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
    assert dec.check() == 0


@pytest.mark.parametrize("bs", [4, 8, 16, 32, 64, 128, 256])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_incompressible_tiles(dtype, bs, device):
    """Tiles whose records outgrow the compressor's LDS stage are written
    straight to the payload; mixed with ordinary tiles in one field.  The
    stage holds 48 KiB for the 64 KiB tiles of block sizes 16..128 with fewer
    than 512 blocks per tile, 44 KiB for those with 512 or more (fp32 block
    sizes 16 and 32, fp64 16), and 20 KiB for the smaller tiles of block sizes
    4, 8 and 256."""
    import torch
    r = np.random.default_rng(77)
    n = 300_001
    x = r.standard_normal(n).astype(dtype)
    x[100_000:200_000] = np.sin(np.arange(100_000) * 1e-3)     # compressible middle
    for eb in (1e-7, 1e-3):
        say("incompressible", dtype.__name__, bs, eb)
        comp = Z.DeviceCompressor(n, dtype, bs)
        comp(torch.from_numpy(x).to(device), Z.ABS, eb)
        view, off, size, _ = comp.stream_view()
        s = view.cpu().numpy().tobytes()
        assert s == O.compress(x, eb, bs)
        dec = Z.DeviceDecompressor(n, dtype, bs, size)
        dec(view, size, Z.DECODE_REFERENCE, index=comp.index)
        assert dec.check() == 0
        assert bits_equal(dec.out.cpu().numpy(), O.decompress(s, n, dtype, ref_quirk=True))


def test_copy_kernel(device):
    """The measurement helper copies exactly (sizes that are not multiples of
    the 4 KiB a workgroup moves included) and refuses what it cannot copy."""
    import torch
    for n in (4, 1024, 1024 + 4, (1 << 20) + 12):
        x = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device=device)
        y = torch.zeros_like(x)
        Z.copy_dev(x, y)
        torch.cuda.synchronize()
        assert torch.equal(x, y), n
    with pytest.raises(Z.SZxError):
        Z.copy_dev(torch.zeros(6, dtype=torch.int8, device=device), torch.zeros(6, dtype=torch.int8, device=device))
    with pytest.raises(Z.SZxError):
        Z.copy_dev(torch.zeros(16, device=device), torch.zeros(8, device=device))
